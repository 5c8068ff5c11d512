"""OBJ/MTL and PNG ingestion through the C ABI (gr_obj_open, gr_png_decode):
the reference's LoadObjFile (obj.go:196-309) and LoadTextureFile
(texture.go:91-103). Parsing runs on the host; `ObjFile.upload` puts the
meshes and textures into a context's resident store."""
from __future__ import annotations

import ctypes
from typing import List

import numpy as np

from . import _lib
from .scenes import MeshData, TextureData


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
