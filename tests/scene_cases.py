"""Small seeded scenes that exercise every branch of the render path.

Each builder returns a gorender_amd.scenes.SceneData; the GPU parity tests draw
them through the C ABI and compare with the CPU oracle.
"""
import functools
import math

import numpy as np

from gorender_amd import _lib
from gorender_amd import scenes as S

ALL = (_lib.GR_SHOW_FACES | _lib.GR_BACKFACE_CULLING | _lib.GR_LIGHTING
       | _lib.GR_FRUSTUM_CLIPPING | _lib.GR_SHOW_TEXTURES)


def sphere(flags=ALL, w=160, h=96, eye=(0, 0, 3.0), nlat=12, nlon=20, textured=True,
           tex_size=(8, 8), seed=1, fov=S.FOV_45, znear=0.0, zfar=50.0, noise=0.05,
           scale=1.0, drop_normals=False):
    rng = S.SplitMix64(S.SEED_BASE + 100 + seed)
    mesh = S.latlong_sphere(nlat, nlon, rng, noise=noise)
    if drop_normals:
        mesh.vertex_normals = None
        mesh.normal_indices = None
    texs = []
    if textured:
        texs = [S.random_texture(rng, tex_size[0], tex_size[1], scale=scale)]
        S.with_texture(mesh, 0)
    fr = S.FrameData(w, h, flags=flags, znear=znear, zfar=zfar)
    d = -np.asarray(eye, np.float64)
    obj = S.camera_object(fr, 0, eye, tuple(d / np.linalg.norm(d)), fov=fov)
    return S.SceneData(fr, [mesh], texs, [obj], name=f"sphere-{seed}")


def overlapping_objects(w=128, h=96, seed=2):
    """Three objects; the second and third are the same mesh with the same
    matrices (every fragment ties in depth), so the third must win everywhere."""
    rng = S.SplitMix64(S.SEED_BASE + 200 + seed)
    m = S.latlong_sphere(10, 16, rng)
    t0 = S.random_texture(rng, 4, 4)
    t1 = S.random_texture(rng, 8, 4)
    t2 = S.TextureData(type=_lib.GR_TEXTURE_SOLID, color=(10, 200, 30, 255))
    m0 = S.latlong_sphere(10, 16, rng)
    S.with_texture(m0, 0)
    m1 = S.MeshData(m.vertices.copy(), m.vertex_indices.copy(), m.vertex_normals.copy(),
                    m.normal_indices.copy(), m.uvs.copy())
    S.with_texture(m1, 1)
    m2 = S.MeshData(m.vertices.copy(), m.vertex_indices.copy(), m.vertex_normals.copy(),
                    m.normal_indices.copy(), m.uvs.copy())
    S.with_texture(m2, 2)
    fr = S.FrameData(w, h)
    objs = [
        S.camera_object(fr, 0, (0, 0, 4), (0, 0, -1),
                        world=S.world_matrix((1, 1, 1), (0.3, 0.2, 0), (0.5, 0, 0))),
        S.camera_object(fr, 1, (0, 0, 4), (0, 0, -1),
                        world=S.world_matrix((1.2, 1, 1), (0, 0.4, 0), (-0.3, 0.1, 0))),
        S.camera_object(fr, 2, (0, 0, 4), (0, 0, -1),
                        world=S.world_matrix((1.2, 1, 1), (0, 0.4, 0), (-0.3, 0.1, 0))),
    ]
    return S.SceneData(fr, [m0, m1, m2], [t0, t1, t2], objs, name="overlap")


def no(bit):
    """The default toggles without `bit`."""
    return ALL & ~bit


def with_flags(scene, flags):
    scene.frame.flags = flags
    return scene


def inside_sphere(flags=ALL, znear=0.0, w=128, h=96, seed=3):
    """Camera inside a sphere (faces wound inward) looking along -z: near-plane
    and side clipping of triangles that pass behind the camera."""
    rng = S.SplitMix64(S.SEED_BASE + 300 + seed)
    m = S.latlong_sphere(16, 24, rng, noise=0.1)
    m.vertex_indices = np.ascontiguousarray(m.vertex_indices[:, [0, 2, 1]])
    m.normal_indices = np.ascontiguousarray(m.normal_indices[:, [0, 2, 1]])
    m.uvs = np.ascontiguousarray(m.uvs[:, [0, 2, 1]])
    t = S.random_texture(rng, 16, 16)
    S.with_texture(m, 0)
    fr = S.FrameData(w, h, flags=flags, znear=znear)
    obj = S.camera_object(fr, 0, (0.1, 0.05, 0.6), (0.2, -0.1, -1.0))
    return S.SceneData(fr, [m], [t], [obj], name="inside")


def cube_scene(flags=ALL | _lib.GR_FLAT_SHADING, w=96, h=64, tex_kind="npot"):
    rng = S.SplitMix64(S.SEED_BASE + 400)
    m = S.cube()
    if tex_kind == "npot":
        t = S.random_texture(rng, 5, 3, scale=2.5)
    elif tex_kind == "solid":
        t = S.TextureData(type=_lib.GR_TEXTURE_SOLID, color=(250, 20, 99, 128))
    else:
        t = S.random_texture(rng, 4, 8, scale=0.75)
    S.with_texture(m, 0)
    fr = S.FrameData(w, h, flags=flags)
    obj = S.camera_object(fr, 0, (2.5, 1.7, 3.1), (-2.5, -1.7, -3.1),
                          world=S.world_matrix((0.9, 1.1, 1.0), (0.1, 0.7, -0.2), (0, 0, 0)))
    return S.SceneData(fr, [m], [t], [obj], name=f"cube-{tex_kind}")


def mixed_textures(w=120, h=80):
    rng = S.SplitMix64(S.SEED_BASE + 500)
    m = S.latlong_sphere(8, 12, rng)
    texs = [S.random_texture(rng, 8, 8), S.random_texture(rng, 3, 7, scale=1.5),
            S.TextureData(type=_lib.GR_TEXTURE_SOLID, color=(1, 2, 3, 4))]
    ft = (np.arange(m.num_faces) % 4).astype(np.int32) - 1  # -1 (nil), 0, 1, 2
    m.face_textures = ft
    fr = S.FrameData(w, h)
    obj = S.camera_object(fr, 0, (0, 0.3, 2.6), (0, -0.3, -2.6))
    return S.SceneData(fr, [m], texs, [obj], name="mixed-tex")


def clip_space_soup(n=300, seed=4, flags=ALL, w=96, h=72, exact=True, w_range=(-3.0, 0.5),
                    span=3.0):
    """Independent triangles given directly in clip space (identity matrices):
    vertices straddle every plane, W crosses zero, and a share of coordinates
    come from a small exact set so plane tests and depth values tie."""
    rng = S.SplitMix64(S.SEED_BASE + 600 + seed)
    V = 3 * n
    xyz = (rng.uniform(3 * V).reshape(V, 3) * 2 - 1) * span
    xyz[:, 2] = rng.uniform(V) * 6 - 1
    wv = w_range[0] + rng.uniform(V) * (w_range[1] - w_range[0])
    if exact:
        choices = np.array([-2.0, -1.0, -0.5, 0.0, 0.5, 1.0, 2.0])
        pick = rng.randint(4 * V, 0, len(choices) - 1).reshape(V, 4)
        mask = rng.uniform(4 * V).reshape(V, 4) < 0.15
        allv = np.concatenate([xyz, wv[:, None]], axis=1)
        allv[mask] = choices[pick][mask]
        xyz, wv = allv[:, :3], allv[:, 3]
    v4 = np.concatenate([xyz, wv[:, None]], axis=1).astype(np.float32)
    f = np.arange(V, dtype=np.int32).reshape(n, 3)
    nrm = np.concatenate([(rng.uniform(3 * V).reshape(V, 3) * 2 - 1),
                          np.ones((V, 1))], axis=1).astype(np.float32)
    uvs = (rng.uniform(6 * n).reshape(n, 3, 2) * 3 - 1).astype(np.float32)
    m = S.MeshData(v4, f, nrm, f.copy(), uvs)
    t = S.random_texture(rng, 8, 8)
    S.with_texture(m, 0)
    fr = S.FrameData(w, h, flags=flags)
    eye = np.eye(4, dtype=np.float32)
    return S.SceneData(fr, [m], [t], [S.ObjectData(0, eye, eye)], name=f"soup-{seed}")


def degenerate(w=64, h=48):
    """Zero-area, single-point and collinear triangles plus a normal one."""
    v = np.array([[0.1, 0.1, 0.5, -1], [0.1, 0.1, 0.5, -1], [0.1, 0.1, 0.5, -1],
                  [-0.5, -0.5, 0.5, -1], [0.5, 0.5, 0.5, -1], [0.0, 0.0, 0.5, -1],
                  [-0.9, 0.8, 0.3, -1], [0.7, 0.8, 0.3, -1], [0.0, -0.9, 0.3, -1],
                  [0.33, 0.33, 0.2, -1], [0.33, 0.34, 0.2, -1], [0.34, 0.33, 0.2, -1],
                  [-5, -5, -1, 1], [5, 5, 5, 1]],  # unused: widen the bounding box
                 np.float32)
    f = np.arange(12, dtype=np.int32).reshape(4, 3)
    m = S.MeshData(v, f)
    fr = S.FrameData(w, h, flags=ALL & ~_lib.GR_BACKFACE_CULLING)
    eye = np.eye(4, dtype=np.float32)
    return S.SceneData(fr, [m], [], [S.ObjectData(0, eye, eye)], name="degenerate")


# Hand-derived frames (16x16). mvp = diag(1, 1, 1, -1) puts model vertex
# (X, Y, Z, Wm) at clip (X, Y, Z, -Wm); with the 16x16 screen matrix the three
# model vertices (1,1,1), (1,0,1), (0,1,1) land on screen (0,0), (0,8), (8,0).
# Under the reference's edge functions and top-left rule that triangle covers
# exactly the 36 pixels with x >= 0, y >= 0, x + y <= 7, at clip W = -1 (depth
# (a + b) + g, within 2^-22 of 1). It is facing the camera in clip space, and
# its bounding box is inside the frustum, so it is never clipped. The light is
# (1, 0, 0), so a world-space normal (nx, 0, 0, nw) lights a face with
# 0.5 + 0.5 * nx / sqrt(nx^2 + nw^2) (4-D normalisation, vector.go:115-125).
HAND_MVP = np.diag([1.0, 1.0, 1.0, -1.0]).astype(np.float32)
HAND_TRI = np.array([[1, 1, 1, 1], [1, 0, 1, 1], [0, 1, 1, 1]], np.float32)
HAND_FLAGS = (_lib.GR_SHOW_FACES | _lib.GR_BACKFACE_CULLING | _lib.GR_LIGHTING
              | _lib.GR_FRUSTUM_CLIPPING)


def hand_frame(kind="flat-ab"):
    """Small frames whose every pixel is derived by hand (tests/test_oracle_pins_frame.py):

    flat-ab / flat-ba  two objects, the same triangle and matrices except the
                       world matrix: A (identity world, face normal (1,0,0,1):
                       colour uint8(200 * (0.5 + 0.5/sqrt 2)) = 170) and B
                       (world translated by (2,0,0), so the normal becomes
                       (3,0,0,1): uint8(200 * (0.5 + 1.5/sqrt 10)) = 194).
                       Every fragment ties in depth: the later object wins.
    gouraud            one object lit per corner through normal indices
                       (1, 1, 1) over normals that differ per vertex; world
                       translated by (2,0,0): every corner 194.
    unlit              lighting off, ambient 0.613: uint8(200 * 0.613) = 122.
    w-zero             the third vertex has clip W = -0: x/w, y/w are NaN or
                       infinite, the triangle is never binned (background only).
    z-inf              the third vertex has z/w = -inf with finite x/w, y/w: the
                       screen matrix's zero terms give 0 * inf = NaN, so the
                       triangle is never binned.
    z-big              as z-inf with z/w finite (-2e30): binned and drawn.
    """
    fr = S.FrameData(16, 16, flags=HAND_FLAGS, light=np.array([1, 0, 0], np.float32))
    eye = np.eye(4, dtype=np.float32)
    shifted = eye.copy()
    shifted[0, 3] = 2.0
    f = np.array([[0, 1, 2]], np.int32)
    fn = np.array([[1, 0, 0, 1]], np.float32)
    if kind in ("flat-ab", "flat-ba"):
        fr.flags |= _lib.GR_FLAT_SHADING
        m = S.MeshData(HAND_TRI.copy(), f, face_normals=fn)
        a, b = S.ObjectData(0, HAND_MVP, eye), S.ObjectData(0, HAND_MVP, shifted)
        objs = [a, b] if kind == "flat-ab" else [b, a]
        return S.SceneData(fr, [m], [], objs, name=f"hand-{kind}")
    if kind == "gouraud":
        vn = np.array([[7, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1]], np.float32)
        m = S.MeshData(HAND_TRI.copy(), f, vertex_normals=vn,
                       normal_indices=np.array([[1, 1, 1]], np.int32), face_normals=fn)
        return S.SceneData(fr, [m], [], [S.ObjectData(0, HAND_MVP, shifted)],
                           name="hand-gouraud")
    if kind == "unlit":
        fr.flags &= ~_lib.GR_LIGHTING
        fr.ambient = 0.613
        m = S.MeshData(HAND_TRI.copy(), f, face_normals=fn)
        return S.SceneData(fr, [m], [], [S.ObjectData(0, HAND_MVP, eye)], name="hand-unlit")
    v = HAND_TRI.copy()
    fr.flags &= ~(_lib.GR_BACKFACE_CULLING | _lib.GR_FRUSTUM_CLIPPING)
    if kind == "w-zero":
        v[2] = [0, 1, 1, 0]
    elif kind == "z-inf":
        v[2] = [0, 1, 3e38, 0.5]
    elif kind == "z-big":
        v[2] = [0, 1, 1e30, 0.5]
    else:
        raise ValueError(kind)
    m = S.MeshData(v, f, face_normals=fn)
    return S.SceneData(fr, [m], [], [S.ObjectData(0, HAND_MVP, eye)], name=f"hand-{kind}")


HAND_KINDS = ("flat-ab", "flat-ba", "gouraud", "unlit", "w-zero", "z-inf", "z-big")


def all_cases():
    """(name, builder) pairs of the small parity suite."""
    return [
        ("sphere-default", lambda: sphere()),
        ("sphere-near", lambda: sphere(eye=(0, 0, 1.15), nlat=24, nlon=40)),
        ("sphere-flat", lambda: sphere(flags=ALL | _lib.GR_FLAT_SHADING)),
        ("sphere-no-normals", lambda: sphere(drop_normals=True)),
        ("sphere-unlit", lambda: sphere(flags=no(_lib.GR_LIGHTING))),
        ("sphere-nocull", lambda: sphere(flags=no(_lib.GR_BACKFACE_CULLING))),
        ("sphere-noclip", lambda: sphere(flags=no(_lib.GR_FRUSTUM_CLIPPING), eye=(0, 0, 1.3))),
        ("sphere-notex", lambda: sphere(flags=no(_lib.GR_SHOW_TEXTURES))),
        ("sphere-nofaces", lambda: sphere(flags=no(_lib.GR_SHOW_FACES))),
        ("sphere-untextured", lambda: sphere(textured=False)),
        ("sphere-npot-scale", lambda: sphere(tex_size=(5, 7), scale=3.0)),
        ("sphere-odd-size", lambda: sphere(w=77, h=53)),
        ("sphere-fov60-near", lambda: sphere(fov=S.float32_const(math.pi / 3), znear=0.1,
                                             zfar=100.0)),
        ("sphere-far-clip", lambda: sphere(zfar=2.5, eye=(0, 0, 2.2))),
        ("overlap-ties", lambda: overlapping_objects()),
        ("inside-near0", lambda: inside_sphere()),
        ("inside-near01", lambda: inside_sphere(znear=0.1)),
        ("inside-nocull", lambda: inside_sphere(flags=no(_lib.GR_BACKFACE_CULLING))),
        ("cube-npot", lambda: cube_scene()),
        ("cube-solid", lambda: cube_scene(tex_kind="solid")),
        ("cube-gouraud", lambda: cube_scene(flags=ALL, tex_kind="pot")),
        ("mixed-textures", lambda: mixed_textures()),
        ("soup-clip", lambda: clip_space_soup()),
        ("soup-noclip", lambda: clip_space_soup(flags=no(_lib.GR_FRUSTUM_CLIPPING), seed=5)),
        ("soup-nocull", lambda: clip_space_soup(flags=no(_lib.GR_BACKFACE_CULLING), seed=6)),
        ("soup-wide", lambda: clip_space_soup(flags=no(_lib.GR_FRUSTUM_CLIPPING), seed=7,
                                              w_range=(-1e-4, 1e-4))),
        ("degenerate", lambda: degenerate()),
        ("overlay-edges", lambda: sphere(flags=ALL | _lib.GR_SHOW_EDGES)),
        ("overlay-vertices", lambda: sphere(flags=ALL | _lib.GR_SHOW_VERTICES, seed=3)),
        ("overlay-both-nofaces", lambda: sphere(
            flags=no(_lib.GR_SHOW_FACES) | _lib.GR_SHOW_EDGES | _lib.GR_SHOW_VERTICES, seed=4)),
        ("overlay-soup", lambda: clip_space_soup(
            flags=ALL | _lib.GR_SHOW_EDGES | _lib.GR_SHOW_VERTICES, seed=8)),
        ("overlay-ties", lambda: with_flags(overlapping_objects(), ALL | _lib.GR_SHOW_EDGES)),
        ("overlay-inside", lambda: inside_sphere(flags=ALL | _lib.GR_SHOW_EDGES
                                                 | _lib.GR_SHOW_VERTICES)),
    ] + [(f"hand-{k}", (lambda k=k: hand_frame(k))) for k in HAND_KINDS]


CONFIG4_FRAMES = 256   # the batch of config 4
CONFIG5_REQUESTS = 64  # the distinct requests the serving benchmark replays


@functools.lru_cache(maxsize=2)
def config4_all():
    """All frames of config 4 (the first n of them are config4_frames(n))."""
    return S.config4_frames(CONFIG4_FRAMES)


def config4_batch(n=16):
    return config4_all()[:n]


@functools.lru_cache(maxsize=1)
def config5_all():
    return S.config5(CONFIG5_REQUESTS)


def config5_set(n=8):
    rs = config5_all()
    return S.ResidentSet(rs.meshes, rs.textures, rs.requests[:n])


def golden_cases():
    """Every scene with a committed oracle digest under tests/golden: the small
    cases, configs 1-3, all 256 frames of config 4 and all 64 requests of
    config 5 (as the subscene of the meshes each one draws)."""
    cases = list(all_cases())
    cases += [("config1-ref", lambda: S.config1("ref")), ("config1-cfg", lambda: S.config1("cfg")),
              ("config2", S.config2), ("config3", S.config3)]
    cases += [(f"config4-{k}", (lambda k=k: config4_all()[k])) for k in range(CONFIG4_FRAMES)]
    cases += [(f"config5-{k}", (lambda k=k: S.subscene(config5_all().requests[k])))
              for k in range(CONFIG5_REQUESTS)]
    return cases
