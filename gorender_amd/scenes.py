"""Scene data in the reference's data model and seeded synthetic generators.

Meshes follow mesh.go:8-26 (Vec4 vertices/normals, per-face index triples,
per-corner UVs, per-face texture), textures texture.go:19-26, objects
mesh.go:71-79 with the matrices renderer.go:255-264 derives from them.

Randomness: SplitMix64 with seed 0x6F72656E + config index; floats are
(x >> 40) * 2**-24.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib

SEED_BASE = 0x6F72656E
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


# ---------------------------------------------------------------------------
# PRNG
# ---------------------------------------------------------------------------

class SplitMix64:
    """SplitMix64 stream; draws are vectorised with numpy uint64 wrap-around."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * _GOLDEN
            self.state = self.state + np.uint64(n) * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))

    def uniform(self, n: int) -> np.ndarray:
        """float64 values in [0, 1) with 24 random bits (exact in float32)."""
        return (self.next_u64(n) >> np.uint64(40)).astype(np.float64) * 2.0 ** -24

    def randint(self, n: int, lo: int, hi: int) -> np.ndarray:
        """Integers in [lo, hi] (inclusive), by multiply-shift of 24-bit draws."""
        u = (self.next_u64(n) >> np.uint64(40)).astype(np.int64)
        return lo + (u * (hi - lo + 1)) // (1 << 24)


# ---------------------------------------------------------------------------
# Data model
# ---------------------------------------------------------------------------

@dataclass
class MeshData:
    vertices: np.ndarray                       # (V,4) float32
    vertex_indices: np.ndarray                 # (F,3) int32
    vertex_normals: Optional[np.ndarray] = None  # (N,4) float32
    normal_indices: Optional[np.ndarray] = None  # (F,3) int32
    uvs: Optional[np.ndarray] = None             # (F,3,2) float32
    face_textures: Optional[np.ndarray] = None   # (F,) int32, -1 = nil
    face_normals: Optional[np.ndarray] = None    # (F,4) float32; None = computed as NewMesh does
    name: str = ""

    @property
    def num_faces(self) -> int:
        return int(self.vertex_indices.shape[0])

    @property
    def num_vertices(self) -> int:
        return int(self.vertices.shape[0])

    def compulsory_bytes(self, textured: bool) -> int:
        """Inputs read once: V*(12 + 12*hasNormals + 8*textured) + F*12."""
        v = self.num_vertices
        has_n = self.vertex_normals is not None and len(self.vertex_normals) > 0
        return v * (12 + 12 * has_n + 8 * textured) + self.num_faces * 12


@dataclass
class TextureData:
    type: int
    width: int = 0
    height: int = 0
    scale: float = 1.0
    color: tuple = (0, 0, 0, 0)
    rgba: Optional[np.ndarray] = None  # (H,W,4) uint8


@dataclass
class ObjectData:
    mesh: int
    mvp: np.ndarray    # (4,4) float32
    world: np.ndarray  # (4,4) float32


@dataclass
class FrameData:
    width: int
    height: int
    flags: int = _lib.GR_DEFAULT_FLAGS
    ref_tiles: int = 16
    screen: Optional[np.ndarray] = None
    light: Optional[np.ndarray] = None
    ambient: float = 0.5
    diffuse: float = 0.5
    znear: float = 0.0
    zfar: float = 50.0
    face_color: tuple = (200, 200, 200, 255)
    clear_color: tuple = (50, 50, 50, 255)
    dot_color: tuple = (100, 100, 100, 255)
    dot_step: int = 10

    def __post_init__(self):
        if self.screen is None:
            self.screen = screen_matrix(self.width, self.height)
        if self.light is None:
            self.light = vec3_normalize(np.array([-1, 1, 1], np.float32))


@dataclass
class SceneData:
    frame: FrameData
    meshes: List[MeshData]
    textures: List[TextureData] = field(default_factory=list)
    objects: List[ObjectData] = field(default_factory=list)
    name: str = ""

    @property
    def num_faces(self) -> int:
        return sum(self.meshes[o.mesh].num_faces for o in self.objects)

    def compulsory_bytes(self) -> int:
        """Algorithmic bytes of one frame: inputs read once, outputs written once."""
        show_tex = bool(self.frame.flags & _lib.GR_SHOW_TEXTURES)
        b = 0
        used = set()
        seen_meshes = set()
        for o in self.objects:
            m = self.meshes[o.mesh]
            tex = []
            if show_tex and m.face_textures is not None:
                tex = [int(t) for t in np.unique(m.face_textures) if t >= 0]
            textured = any(self.textures[t].type != _lib.GR_TEXTURE_SOLID for t in tex)
            b += m.compulsory_bytes(textured)
            seen_meshes.add(o.mesh)
            used.update(t for t in tex if self.textures[t].type != _lib.GR_TEXTURE_SOLID)
        for t in used:
            b += self.textures[t].width * self.textures[t].height * 4
        return b + self.frame.width * self.frame.height * 8


# ---------------------------------------------------------------------------
# Host math (the product library's float32 reproduction of matrix.go)
# ---------------------------------------------------------------------------

def _f32p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def screen_matrix(w: int, h: int) -> np.ndarray:
    out = np.zeros((4, 4), np.float32)
    _lib.load().gr_matrix_screen(w, h, _f32p(out))
    return out


def perspective_matrix(fov: float, aspect: float, znear: float, zfar: float) -> np.ndarray:
    out = np.zeros((4, 4), np.float32)
    _lib.load().gr_matrix_perspective(float(np.float32(fov)), float(np.float32(aspect)),
                                      float(np.float32(znear)), float(np.float32(zfar)),
                                      _f32p(out))
    return out


def view_matrix(eye, direction, up) -> np.ndarray:
    out = np.zeros((4, 4), np.float32)
    e = np.asarray(eye, np.float32)
    d = np.asarray(direction, np.float32)
    u = np.asarray(up, np.float32)
    _lib.load().gr_matrix_view(_f32p(e), _f32p(d), _f32p(u), _f32p(out))
    return out


def world_matrix(scale=(1, 1, 1), rotation=(0, 0, 0), translation=(0, 0, 0)) -> np.ndarray:
    out = np.zeros((4, 4), np.float32)
    s = np.asarray(scale, np.float32)
    r = np.asarray(rotation, np.float32)
    t = np.asarray(translation, np.float32)
    _lib.load().gr_matrix_world(_f32p(s), _f32p(r), _f32p(t), _f32p(out))
    return out


def mvp_matrix(p: np.ndarray, v: np.ndarray, w: np.ndarray) -> np.ndarray:
    out = np.zeros((4, 4), np.float32)
    p = np.ascontiguousarray(p, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    _lib.load().gr_matrix_mvp(_f32p(p), _f32p(v), _f32p(w), _f32p(out))
    return out


def vec3_normalize(v) -> np.ndarray:
    out = np.zeros(3, np.float32)
    a = np.asarray(v, np.float32)
    _lib.load().gr_vec3_normalize(_f32p(a), _f32p(out))
    return out


def float32_const(x: float) -> float:
    return float(np.float32(x))


FOV_45 = float32_const(math.pi / 4)  # float32(45 * (math.Pi / 180)), renderer.go:118


def camera_object(frame: FrameData, mesh: int, eye, direction, up=(0, 1, 0), fov=FOV_45,
                  world: Optional[np.ndarray] = None) -> ObjectData:
    """Matrices of projectObject (renderer.go:255-262) for one object."""
    aspect = np.float32(frame.width) / np.float32(frame.height)  # renderer.go:115
    p = perspective_matrix(fov, float(aspect), frame.znear, frame.zfar)
    v = view_matrix(eye, direction, up)
    w = world if world is not None else world_matrix()
    return ObjectData(mesh, mvp_matrix(p, v, w), w)


# ---------------------------------------------------------------------------
# Geometry generators
# ---------------------------------------------------------------------------

def icosphere(level: int) -> MeshData:
    """Icosahedron subdivided `level` times, projected to r=1, CCW outward.

    Normals are the positions; UVs are zero; W = 1 (as obj.go:42-58 stores them).
    """
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}
        nf = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = (verts[a] + verts[b]) / 2.0
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    p = np.array(verts, np.float64).astype(np.float32)
    v4 = np.concatenate([p, np.ones((len(p), 1), np.float32)], axis=1)
    f = np.array(faces, np.int32)
    _orient_outward(v4, f)
    return MeshData(vertices=v4, vertex_indices=f, vertex_normals=v4.copy(),
                    normal_indices=f.copy(), uvs=np.zeros((len(f), 3, 2), np.float32),
                    name=f"icosphere{level}")


def _orient_outward(v4: np.ndarray, f: np.ndarray) -> None:
    """Make every face wind counter-clockwise seen from outside (in place)."""
    p = v4[:, :3].astype(np.float64)
    a, b, c = p[f[:, 0]], p[f[:, 1]], p[f[:, 2]]
    n = np.cross(b - a, c - a)
    flip = np.einsum("ij,ij->i", n, (a + b + c)) < 0
    f[flip] = f[flip][:, [0, 2, 1]]


def latlong_sphere(nlat: int, nlon: int, rng: Optional[SplitMix64] = None,
                   noise: float = 0.05) -> MeshData:
    """nlat x nlon quads -> 2*nlat*nlon triangles over (nlat+1) x (nlon+1) vertices.

    Radius 1 + U[-noise, noise] per surface point (seam column and pole rows
    share their radius so the surface is closed), normals = normalize(position),
    UV = (j/nlon, i/nlat), CCW outward.
    """
    i = np.arange(nlat + 1, dtype=np.float64)[:, None]
    j = np.arange(nlon + 1, dtype=np.float64)[None, :]
    if rng is not None and noise > 0:
        u = rng.uniform((nlat + 1) * nlon).reshape(nlat + 1, nlon)
        r = 1.0 + (u - 0.5) * (2 * noise)
        r[0, :] = r[0, 0]
        r[-1, :] = r[-1, 0]
        r = np.concatenate([r, r[:, :1]], axis=1)
    else:
        r = np.ones((nlat + 1, nlon + 1))
    theta = math.pi * i / nlat
    phi = 2 * math.pi * j / nlon
    x = r * np.sin(theta) * np.cos(phi)
    y = r * np.cos(theta) * np.ones_like(phi)
    z = -r * np.sin(theta) * np.sin(phi)
    pos = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    nrm = pos / np.linalg.norm(pos, axis=1, keepdims=True)
    V = pos.shape[0]
    v4 = np.concatenate([pos.astype(np.float32), np.ones((V, 1), np.float32)], axis=1)
    n4 = np.concatenate([nrm.astype(np.float32), np.ones((V, 1), np.float32)], axis=1)
    ii, jj = np.meshgrid(np.arange(nlat), np.arange(nlon), indexing="ij")
    a = ii * (nlon + 1) + jj
    b = a + 1
    c = a + (nlon + 1)
    d = c + 1
    f = np.stack([np.stack([a, c, b], -1), np.stack([b, c, d], -1)], axis=2).reshape(-1, 3)
    f = f.astype(np.int32)
    _orient_outward(v4, f)
    uv_v = np.stack([(np.broadcast_to(j, (nlat + 1, nlon + 1)) / nlon),
                     (np.broadcast_to(i, (nlat + 1, nlon + 1)) / nlat)], -1)
    uv_v = uv_v.reshape(-1, 2).astype(np.float32)
    uvs = uv_v[f]
    return MeshData(vertices=v4, vertex_indices=f, vertex_normals=n4, normal_indices=f.copy(),
                    uvs=uvs, name=f"latlong{nlat}x{nlon}")


def cube() -> MeshData:
    """Unit cube (side 2, centred) with per-face normals and UVs, like models/cube.obj."""
    p = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    normals = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
    f, ni, uvs = [], [], []
    quv = [(0, 0), (1, 0), (1, 1), (0, 1)]
    for q, (a, b, c, d) in enumerate(quads):
        f += [(a, b, c), (a, c, d)]
        ni += [(q, q, q), (q, q, q)]
        uvs += [(quv[0], quv[1], quv[2]), (quv[0], quv[2], quv[3])]
    v4 = np.concatenate([p, np.ones((8, 1), np.float32)], axis=1)
    f = np.array(f, np.int32)
    n4 = np.concatenate([np.array(normals, np.float32), np.ones((6, 1), np.float32)], axis=1)
    mesh = MeshData(vertices=v4, vertex_indices=f, vertex_normals=n4,
                    normal_indices=np.array(ni, np.int32), uvs=np.array(uvs, np.float32),
                    name="cube")
    # Keep normal indices aligned with any winding flip.
    p64 = v4[:, :3].astype(np.float64)
    a, b, c = p64[f[:, 0]], p64[f[:, 1]], p64[f[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0
    mesh.vertex_indices[flip] = mesh.vertex_indices[flip][:, [0, 2, 1]]
    mesh.uvs[flip] = mesh.uvs[flip][:, [0, 2, 1]]
    return mesh


def random_texture(rng: SplitMix64, w: int, h: int, scale: float = 1.0) -> TextureData:
    px = (rng.next_u64(w * h) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    rgba = px.view(np.uint8).reshape(h, w, 4).copy()
    kind = _lib.GR_TEXTURE_IMAGE_FAST if (w & (w - 1) == 0 and h & (h - 1) == 0) \
        else _lib.GR_TEXTURE_IMAGE
    return TextureData(type=kind, width=w, height=h, scale=scale, rgba=rgba)


def with_texture(mesh: MeshData, tex: int) -> MeshData:
    mesh.face_textures = np.full(mesh.num_faces, tex, np.int32)
    return mesh


# ---------------------------------------------------------------------------
# Benchmark configurations
# ---------------------------------------------------------------------------

def config1(variant: str = "ref") -> SceneData:
    """Icosphere L4 (5,120 tris), 640x360, camera (0,0,3) -> origin, fovY 60,
    near 0.1, far 100, Gouraud, culling, untextured. variant "ref" keeps the
    reference light/face colour, "cfg" uses light (0.577,0.577,-0.577) and
    albedo 0.8 (face colour 204)."""
    fr = FrameData(640, 360, znear=0.1, zfar=100.0)
    if variant == "cfg":
        fr.light = np.array([0.577, 0.577, -0.577], np.float32)
        fr.face_color = (204, 204, 204, 255)
    mesh = icosphere(4)
    obj = camera_object(fr, 0, (0, 0, 3), (0, 0, -1), fov=float32_const(math.pi / 3))
    return SceneData(fr, [mesh], [], [obj], name=f"config1-{variant}")


def config2(nlat: int = 250, nlon: int = 400, width: int = 1280, height: int = 720) -> SceneData:
    """Displaced lat-long sphere (250x400 quads = 200,000 tris) with a 16x16 random
    texture, 1280x720, camera at (0,0,1.2) looking at the origin, reference
    defaults fovY 45 / near 0 / far 50 (renderer.go:118-121)."""
    rng = SplitMix64(SEED_BASE + 2)
    mesh = latlong_sphere(nlat, nlon, rng)
    tex = random_texture(rng, 16, 16)
    with_texture(mesh, 0)
    fr = FrameData(width, height)
    obj = camera_object(fr, 0, (0, 0, 1.2), (0, 0, -1))
    return SceneData(fr, [mesh], [tex], [obj], name=f"config2-{width}x{height}")


def config3(total_faces: int = 1_000_000, width: int = 1920, height: int = 1080,
            instances: int = 64) -> SceneData:
    """64 instances of icospheres L2..L6 and cubes in a 20^3 box (levels raised
    in seeded order until the faces reach at most `total_faces`), random unit
    quaternion rotation, scale [0.5, 3], textures 16^2..256^2, half of the meshes
    without vertex normals (flat), camera at the origin looking down -z."""
    rng = SplitMix64(SEED_BASE + 3)
    fr = FrameData(width, height)
    base = {lvl: icosphere(lvl) for lvl in range(2, 7)}
    kinds = [int(v) for v in rng.randint(instances, 0, 5)]
    # Weighted to the requested total: icospheres drawn in seeded order are
    # raised one level (4x the faces) while the total stays within it.
    faces = [12 if kd == 5 else 20 * 4 ** (kd + 2) for kd in kinds]
    for i in [int(v) for v in rng.randint(4 * instances, 0, instances - 1)]:
        if kinds[i] < 4 and sum(faces) + 3 * faces[i] <= total_faces:
            kinds[i] += 1
            faces[i] *= 4
    meshes, textures, objects = [], [], []
    for k in range(instances):
        kind = kinds[k]
        src = cube() if kind == 5 else base[kind + 2]
        m = MeshData(src.vertices.copy(), src.vertex_indices.copy(),
                     src.vertex_normals.copy(), src.normal_indices.copy(), src.uvs.copy())
        if kind != 5:
            # spherical UVs so textures vary over the surface
            p = m.vertices[:, :3].astype(np.float64)
            u = (np.arctan2(p[:, 2], p[:, 0]) / (2 * np.pi) + 0.5).astype(np.float32)
            v = (np.arccos(np.clip(p[:, 1], -1, 1)) / np.pi).astype(np.float32)
            m.uvs = np.stack([u, v], -1)[m.vertex_indices]
        if k % 2 == 1:
            m.vertex_normals = None
            m.normal_indices = None
        side = 1 << int(rng.randint(1, 4, 8)[0])
        textures.append(random_texture(rng, side, side))
        with_texture(m, len(textures) - 1)
        meshes.append(m)
        pos = (rng.uniform(3) * 20.0 - 10.0).astype(np.float32)
        q = _shoemake(rng)
        s = np.float32(0.5 + rng.uniform(1)[0] * 2.5)
        w = _quat_world(q, s, pos)
        objects.append(camera_object(fr, k, (0, 0, 0), (0, 0, -1), world=w))
    return SceneData(fr, meshes, textures, objects, name="config3")


def _shoemake(rng: SplitMix64) -> np.ndarray:
    u1, u2, u3 = rng.uniform(3)
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    return np.array([a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
                     b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)])


def _quat_world(q, s, t) -> np.ndarray:
    x, y, z, w = q
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r * float(s)
    m[:3, 3] = t
    return m.astype(np.float32)


RESOLUTIONS = [(320, 240), (640, 360), (1280, 720), (1920, 1080)]


def config4_frames(n: int = 256, seed_offset: int = 4) -> List[SceneData]:
    """Batched form: n frames, faces log-uniform in [1k, 500k] as lat-long grids
    with a 250:400 aspect, resolution from RESOLUTIONS, textured with p=0.5."""
    rng = SplitMix64(SEED_BASE + seed_offset)
    out = []
    for k in range(n):
        F = math.exp(math.log(1e3) + rng.uniform(1)[0] * (math.log(5e5) - math.log(1e3)))
        nlat = max(2, int(round(math.sqrt(F / 2 * 250 / 400))))
        nlon = max(3, int(round(F / 2 / nlat)))
        w, h = RESOLUTIONS[int(rng.randint(1, 0, 3)[0])]
        textured = rng.uniform(1)[0] < 0.5
        mesh = latlong_sphere(nlat, nlon, rng)
        texs = []
        if textured:
            texs = [random_texture(rng, 16, 16)]
            with_texture(mesh, 0)
        fr = FrameData(w, h)
        obj = camera_object(fr, 0, (0, 0, 1.2), (0, 0, -1))
        out.append(SceneData(fr, [mesh], texs, [obj], name=f"config4-{k}"))
    return out


@dataclass
class ResidentSet:
    """Meshes and textures uploaded once and drawn by many requests."""
    meshes: List[MeshData]
    textures: List[TextureData]
    requests: List[SceneData]


SERVE_RESOLUTIONS = [(640, 360), (1280, 720), (1920, 1080)]


def config5(num_requests: int = 64, seed_offset: int = 5) -> ResidentSet:
    """Serving form: 32 resident meshes totalling ~8 M triangles (16 displaced
    lat-long spheres of 224k..673k triangles, 8 icospheres L6, 8 icospheres L5),
    textures of side 2^U{4..10}; each request draws 1-8 objects (random mesh,
    unit-quaternion rotation, translation U[-1.5, 1.5]^3, scale U[0.4, 1.2]) from
    a camera on a shell r in [2, 6] looking at the origin, at one of three
    resolutions. Every request lists all 32 meshes; objects refer to them by
    index."""
    rng = SplitMix64(SEED_BASE + seed_offset)
    ico = {5: icosphere(5), 6: icosphere(6)}
    meshes, textures = [], []
    for k in range(32):
        side = 1 << int(rng.randint(1, 4, 10)[0])
        textures.append(random_texture(rng, side, side))
        if k % 2 == 0:
            F = 449_000 * (0.5 + (k // 2) / 15.0)
            nlat = max(2, int(round(math.sqrt(F / 3.2))))
            nlon = max(3, int(round(F / 2 / nlat)))
            m = latlong_sphere(nlat, nlon, rng)
        else:
            src = ico[6 if k % 4 == 1 else 5]
            m = MeshData(src.vertices.copy(), src.vertex_indices.copy(),
                         src.vertex_normals.copy(), src.normal_indices.copy(), src.uvs.copy(),
                         name=src.name)
            p = m.vertices[:, :3].astype(np.float64)
            u = (np.arctan2(p[:, 2], p[:, 0]) / (2 * np.pi) + 0.5).astype(np.float32)
            v = (np.arccos(np.clip(p[:, 1], -1, 1)) / np.pi).astype(np.float32)
            m.uvs = np.stack([u, v], -1)[m.vertex_indices]
        with_texture(m, k)
        meshes.append(m)
    requests = []
    for r in range(num_requests):
        w, h = SERVE_RESOLUTIONS[int(rng.randint(1, 0, 2)[0])]
        fr = FrameData(w, h)
        radius = 2.0 + 4.0 * rng.uniform(1)[0]
        z = 2.0 * rng.uniform(1)[0] - 1.0
        phi = 2.0 * math.pi * rng.uniform(1)[0]
        rxy = math.sqrt(max(0.0, 1.0 - z * z))
        eye = np.array([rxy * math.cos(phi), z, rxy * math.sin(phi)]) * radius
        direction = -eye / np.linalg.norm(eye)
        up = (0, 1, 0) if abs(direction[1]) < 0.99 else (1, 0, 0)
        nobj = int(rng.randint(1, 1, 8)[0])
        objects = []
        for _ in range(nobj):
            mi = int(rng.randint(1, 0, 31)[0])
            q = _shoemake(rng)
            t = (rng.uniform(3) * 3.0 - 1.5).astype(np.float32)
            sc = np.float32(0.4 + 0.8 * rng.uniform(1)[0])
            objects.append(camera_object(fr, mi, tuple(eye.astype(np.float32)),
                                         tuple(direction.astype(np.float32)), up,
                                         world=_quat_world(q, sc, t)))
        requests.append(SceneData(fr, meshes, textures, objects, name=f"config5-{r}"))
    return ResidentSet(meshes, textures, requests)


def subscene(s: SceneData) -> SceneData:
    """The same draw with only the meshes its objects use (renumbered); textures
    are kept whole so face texture ids stay valid."""
    used = sorted({o.mesh for o in s.objects})
    remap = {m: i for i, m in enumerate(used)}
    objs = [ObjectData(remap[o.mesh], o.mvp, o.world) for o in s.objects]
    return SceneData(s.frame, [s.meshes[m] for m in used], s.textures, objs, name=s.name)

