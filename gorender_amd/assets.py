"""OBJ/MTL and PNG ingestion through the C ABI (gr_obj_open, gr_png_decode):
the reference's LoadObjFile (obj.go:196-309) and LoadTextureFile
(texture.go:91-103). Parsing runs on the host; `ObjFile.upload` puts the
meshes and textures into a context's resident store. `SceneFile` reads the
reference's JSON scene manifests (scene.go:56-133) on top of both."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List

import numpy as np

from . import _lib
from .scenes import (FrameData, MeshData, SceneData, TextureData, camera_object,
                     world_matrix)


def png_decode(path: str) -> np.ndarray:
    """(H, W, 4) uint8 as color.RGBAModel converts the decoded image."""
    lib = _lib.load()
    w, h = ctypes.c_int32(), ctypes.c_int32()
    px = ctypes.POINTER(ctypes.c_uint8)()
    _lib.check(lib.gr_png_decode(str(path).encode(), ctypes.byref(w), ctypes.byref(h),
                                 ctypes.byref(px)))
    try:
        n = w.value * h.value * 4
        out = np.ctypeslib.as_array(px, shape=(n,)).copy().reshape(h.value, w.value, 4)
    finally:
        lib.gr_host_free(px)
    return out


class ObjFile:
    """A parsed OBJ file: meshes (face_textures index `textures`) and textures."""

    def __init__(self, path: str, single_mesh: bool = False):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self.lib.gr_obj_open(str(path).encode(), int(single_mesh), ctypes.byref(h)))
        self.handle = h
        try:
            self.meshes = [self._mesh(i) for i in range(self.lib.gr_obj_num_meshes(h))]
            self.textures = [self._texture(i) for i in range(self.lib.gr_obj_num_textures(h))]
        except Exception:
            self.close()
            raise

    def _mesh(self, i: int) -> MeshData:
        d = _lib.gr_mesh_desc()
        _lib.check(self.lib.gr_obj_mesh(self.handle, i, ctypes.byref(d)))

        def arr(ptr, ctype, shape):
            n = int(np.prod(shape))
            if n == 0 or not ptr:
                return np.zeros(shape, np.float32 if ctype is ctypes.c_float else np.int32)
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)),
                                         shape=(n,)).copy().reshape(shape)

        V, N, F = d.num_vertices, d.num_vertex_normals, d.num_faces
        return MeshData(
            vertices=arr(d.vertices, ctypes.c_float, (V, 4)),
            vertex_indices=arr(d.vertex_indices, ctypes.c_int32, (F, 3)),
            vertex_normals=arr(d.vertex_normals, ctypes.c_float, (N, 4)) if N else None,
            normal_indices=arr(d.normal_indices, ctypes.c_int32, (F, 3)),
            uvs=arr(d.uvs, ctypes.c_float, (F, 3, 2)),
            face_textures=arr(d.face_textures, ctypes.c_int32, (F,)),
            name=f"obj-mesh-{i}")

    def _texture(self, i: int) -> TextureData:
        d = _lib.gr_texture_desc()
        px = ctypes.POINTER(ctypes.c_uint8)()
        _lib.check(self.lib.gr_obj_texture(self.handle, i, ctypes.byref(d), ctypes.byref(px)))
        rgba = None
        if px:
            rgba = np.ctypeslib.as_array(px, shape=(d.width * d.height * 4,)).copy()
            rgba = rgba.reshape(d.height, d.width, 4)
        return TextureData(type=d.type, width=d.width, height=d.height, scale=d.scale,
                           color=tuple(int(c) for c in d.color), rgba=rgba)

    def upload(self, ctx) -> List[int]:
        n = len(self.meshes)
        handles = (ctypes.c_int32 * max(1, n))()
        got = ctypes.c_int32()
        _lib.check(self.lib.gr_obj_upload(ctx.handle, self.handle, handles, n,
                                          ctypes.byref(got)))
        return [handles[i] for i in range(got.value)]

    def close(self) -> None:
        if self.handle:
            self.lib.gr_obj_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# JSON scenes: LoadSceneFile (scene.go:56-133)
# ---------------------------------------------------------------------------

class SceneFileError(ValueError):
    """A scene file the reference's LoadSceneFile rejects, with its message."""


def go_float32(text: str) -> np.float32:
    """strconv.ParseFloat(text, 32) as encoding/json applies it to a float32
    field: the decimal rounded once, to nearest-even float32. A finite value
    beyond float32's range is an error (json's UnmarshalTypeError)."""
    from fractions import Fraction
    x = Fraction(text)
    try:
        with np.errstate(over="ignore"):
            f = np.float32(float(x))  # float(Fraction) rounds once, to float64
    except OverflowError:  # beyond float64 too
        f = np.float32(np.inf)
    if not np.isfinite(f):
        raise SceneFileError(f"json: cannot unmarshal number {text} into a float32")
    # float64 then float32 can round twice; pick the nearest float32 exactly.
    with np.errstate(over="ignore"):
        down = np.nextafter(f, np.float32(-np.inf))
        up = np.nextafter(f, np.float32(np.inf))
    best = None
    for c in (down, f, up):
        if not np.isfinite(c):
            continue
        key = (abs(Fraction(float(c)) - x), int(np.array(c).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, c)
    if best[1] == 0:  # a zero keeps the decimal's sign ("-0", "-1e-50")
        return np.float32(-0.0) if text.lstrip().startswith("-") else np.float32(0.0)
    return np.float32(best[1])


class _Obj(list):
    """A JSON object as its (key, value) pairs, in order, duplicates kept."""


def _json_fields(pairs, fields: dict, what: str) -> dict:
    """Go's struct decoding: each key sets the field whose name matches it
    exactly or, failing that, ignoring case; later keys overwrite earlier
    ones; unknown keys are ignored; null leaves the field as it is."""
    if not isinstance(pairs, _Obj):
        raise SceneFileError(f"json: cannot unmarshal a non-object into {what}")
    out = {}
    for key, value in pairs:
        name = key if key in fields else next((f for f in fields if f.lower() == key.lower()),
                                              None)
        if name is None or value is None:
            continue
        out[name] = fields[name](value)
    return out


def _json_string(v):
    if not isinstance(v, str):
        raise SceneFileError("json: cannot unmarshal a non-string into a string field")
    return v


def _json_f32(v):
    if not isinstance(v, _Number):
        raise SceneFileError("json: cannot unmarshal a non-number into a float32 field")
    return go_float32(v.text)


def _json_vec3(v):
    # [3]float32: missing elements stay zero, extra elements are dropped.
    if type(v) is not list:
        raise SceneFileError("json: cannot unmarshal a non-array into [3]float32")
    out = np.zeros(3, np.float32)
    for i, e in enumerate(v[:3]):
        if e is not None:
            out[i] = _json_f32(e)
    return out


def _json_constant(name):
    # NaN and Infinity are not JSON: Go's decoder stops at their first letter.
    raise SceneFileError(f"invalid character '{name[0]}' looking for beginning of value")


class _Number:
    """A JSON number kept as its text, so float32 fields round it once."""

    def __init__(self, text: str):
        self.text = text


def _json_array(v, item, what):
    # A null element decodes to the zero struct.
    if type(v) is not list:
        raise SceneFileError(f"json: cannot unmarshal a non-array into {what}")
    return [item(e) if e is not None else {} for e in v]


@dataclass
class SceneObject:
    """One object of a scene file (scene.go:113-128, mesh.go:81-89): its mesh
    (index into SceneFile.meshes) and its transform, rotation in radians."""
    mesh: int
    scale: np.ndarray
    rotation: np.ndarray
    translation: np.ndarray

    def world(self) -> np.ndarray:
        return world_matrix(self.scale, self.rotation, self.translation)


class SceneFile:
    """A JSON scene manifest loaded as LoadSceneFile (scene.go:56-133) does:

    - every mesh entry loads the first mesh of its OBJ file (LoadMeshFile with
      singleMesh, mesh.go:91-103); all its faces take the entry's texture file
      (with `textureScale` when non-zero), or the solid 200,200,200,255
      default texture when the entry names none (the OBJ's own materials are
      dropped);
    - a later mesh entry with the same id replaces an earlier one;
    - objects name a mesh id; rotation is given in degrees and converted with
      Vec3.ToRadians (vector.go:82-85) in float32; a zero scale only warns.

    Paths are joined to the manifest's directory as Go's path.Join does.
    `meshes`, `textures` and `objects` feed SceneData through `scene()`.
    """

    def __init__(self, path: str):
        import json
        import posixpath
        path = str(path)
        try:
            with open(path, "rb") as f:
                raw = f.read()
        except OSError as e:
            raise SceneFileError(f"failed to open level file: {e}") from e
        try:
            # json.Decoder.Decode reads the first value and ignores the rest;
            # a top-level null leaves the scene empty.
            text = raw.decode("utf-8", errors="replace").lstrip(" \t\r\n")
            doc, _ = json.JSONDecoder(object_pairs_hook=_Obj, parse_float=_Number,
                                      parse_int=_Number,
                                      parse_constant=_json_constant).raw_decode(text)
            if doc is None:
                doc = _Obj()
            mesh_fields = {"id": _json_string, "objFile": _json_string,
                           "texture": _json_string, "textureScale": _json_f32}
            obj_fields = {"meshID": _json_string, "position": _json_vec3,
                          "rotation": _json_vec3, "scale": _json_vec3}
            top = _json_fields(doc, {
                "name": _json_string,
                "meshes": lambda v: _json_array(
                    v, lambda e: _json_fields(e, mesh_fields, "SceneMeshData"), "[]SceneMeshData"),
                "objects": lambda v: _json_array(
                    v, lambda e: _json_fields(e, obj_fields, "SceneObjectData"),
                    "[]SceneObjectData"),
            }, "SceneData")
        except SceneFileError as e:
            raise SceneFileError(f"failed to read scene manifest: {e}") from e
        except ValueError as e:
            raise SceneFileError(f"failed to read scene manifest: {e}") from e

        root = posixpath.dirname(path) or "."

        def join(name: str) -> str:
            return posixpath.normpath(root + "/" + name) if name else posixpath.normpath(root)

        self.name = top.get("name", "")
        self.meshes: List[MeshData] = []
        self.textures: List[TextureData] = []
        self.objects: List[SceneObject] = []
        self.warnings: List[str] = []
        default_tex = None
        by_id = {}
        for md in top.get("meshes", []):
            mid = md.get("id", "")
            obj_path = join(md.get("objFile", ""))
            ext = posixpath.splitext(obj_path)[1]
            if ext != ".obj":
                raise SceneFileError(
                    f"failed to load mesh '{mid}': unsupported mesh format: {ext}")
            try:
                of = ObjFile(obj_path, single_mesh=True)
                mesh = of.meshes[0]
                of.close()
            except Exception as e:
                raise SceneFileError(f"failed to load mesh '{mid}': {e}") from e
            if md.get("texture", ""):
                try:
                    rgba = png_decode(join(md["texture"]))
                except Exception as e:
                    raise SceneFileError(f"failed to load texture {mid}: {e}") from e
                h, w = rgba.shape[:2]
                pow2 = w != 0 and w & (w - 1) == 0 and h != 0 and h & (h - 1) == 0
                scale = md.get("textureScale", np.float32(0))
                self.textures.append(TextureData(
                    type=_lib.GR_TEXTURE_IMAGE_FAST if pow2 else _lib.GR_TEXTURE_IMAGE,
                    width=w, height=h, scale=float(scale) if scale != 0 else 1.0, rgba=rgba))
                tex = len(self.textures) - 1
            else:
                if default_tex is None:
                    self.textures.append(TextureData(type=_lib.GR_TEXTURE_SOLID,
                                                     color=(200, 200, 200, 255)))
                    default_tex = len(self.textures) - 1
                tex = default_tex
            mesh.face_textures = np.full(mesh.num_faces, tex, np.int32)
            mesh.name = mid
            self.meshes.append(mesh)
            by_id[mid] = len(self.meshes) - 1
        deg = np.float32(np.float32(np.pi) / np.float32(180))  # pi32 / 180
        for od in top.get("objects", []):
            mid = od.get("meshID", "")
            if mid not in by_id:
                raise SceneFileError(f"mesh id not found: {mid}")
            scale = od.get("scale", np.zeros(3, np.float32))
            if not scale.any():
                self.warnings.append(f"[WARN] object scale is zero: {mid}")
            rot = od.get("rotation", np.zeros(3, np.float32))
            self.objects.append(SceneObject(
                mesh=by_id[mid], scale=scale, rotation=(rot * deg).astype(np.float32),
                translation=od.get("position", np.zeros(3, np.float32))))

    # Scene.NumObjects / NumVertices / NumTriangles (scene.go:36-54)
    def num_objects(self) -> int:
        return len(self.objects)

    def num_vertices(self) -> int:
        return sum(self.meshes[o.mesh].vertices.shape[0] for o in self.objects)

    def num_triangles(self) -> int:
        return sum(self.meshes[o.mesh].num_faces for o in self.objects)

    def scene(self, frame: FrameData, eye, direction, up=(0, 1, 0), fov=None) -> SceneData:
        """The scene seen from a camera: every object's matrices as
        projectObject builds them (renderer.go:255-262)."""
        kw = {} if fov is None else {"fov": fov}
        objs = [camera_object(frame, o.mesh, eye, direction, up=up, world=o.world(), **kw)
                for o in self.objects]
        return SceneData(frame, self.meshes, self.textures, objs, name=self.name)
