"""ctypes binding of the CPU oracle (oracle/build/librefcpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, Structure, c_float, c_int32, c_int64, c_uint8, c_uint32, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("REFCPU_LIB") or os.path.join(HERE, "build", "librefcpu.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libmatrix_amd64.so")


class rc_vec2(Structure):
    _fields_ = [("x", c_float), ("y", c_float)]


class rc_vec3(Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float)]


class rc_vec4(Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float), ("w", c_float)]


class rc_uv(Structure):
    _fields_ = [("u", c_float), ("v", c_float)]


class rc_matrix(Structure):
    _fields_ = [("m", c_float * 16)]


class rc_texture(Structure):
    _fields_ = [("type", c_int32), ("width", c_int32), ("height", c_int32),
                ("width_f", c_float), ("height_f", c_float), ("scale", c_float),
                ("color", c_uint8 * 4), ("pixels", c_void_p)]


class rc_mesh(Structure):
    _fields_ = [("num_vertices", c_int32), ("vertices", c_void_p),
                ("num_vertex_normals", c_int32), ("vertex_normals", c_void_p),
                ("num_faces", c_int32), ("vertex_indices", c_void_p),
                ("normal_indices", c_void_p), ("uvs", c_void_p), ("face_textures", c_void_p),
                ("face_normals", c_void_p), ("bbox", rc_vec4 * 8)]


class rc_object(Structure):
    _fields_ = [("mesh", c_int32), ("mvp", rc_matrix), ("world", rc_matrix)]


class rc_frame(Structure):
    _fields_ = [("width", c_int32), ("height", c_int32), ("flags", c_uint32),
                ("ref_tiles", c_int32), ("screen", rc_matrix), ("light", rc_vec3),
                ("ambient", c_float), ("diffuse", c_float), ("znear", c_float),
                ("zfar", c_float), ("face_color", c_uint8 * 4), ("clear_color", c_uint8 * 4),
                ("dot_color", c_uint8 * 4), ("dot_step", c_int32)]


class rc_stats(Structure):
    _fields_ = [("tpf", c_int64), ("triangles", c_int64), ("fragments_covered", c_int64),
                ("fragments_passed", c_int64), ("faces_front", c_int64),
                ("clip_overflow", c_int64), ("pixels_written", c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = c_void_p
    sig = {
        "rc_go_int": (c_int64, [c_float]),
        "rc_go_uint8": (c_uint8, [c_float]),
        "rc_go_min": (c_float, [c_float, c_float]),
        "rc_go_max": (c_float, [c_float, c_float]),
        "rc_sqrt32": (c_float, [c_float]),
        "rc_transform_batch": (None, [P, P, c_int64]),
        "rc_transform_batch_pairwise": (None, [P, P, c_int64]),
        "rc_matrix_multiply": (rc_matrix, [P, P]),
        "rc_world_matrix": (rc_matrix, [rc_vec3, rc_vec3, rc_vec3]),
        "rc_view_matrix": (rc_matrix, [rc_vec3, rc_vec3, rc_vec3]),
        "rc_perspective_matrix": (rc_matrix, [c_float, c_float, c_float, c_float]),
        "rc_screen_matrix": (rc_matrix, [c_int32, c_int32]),
        "rc_rotation_matrix": (rc_matrix, [c_float, c_float, c_float]),
        "rc_vec3_normalize": (rc_vec3, [rc_vec3]),
        "rc_vec4_normalize": (rc_vec4, [rc_vec4]),
        "rc_face_normals": (None, [P, P, c_int32, P]),
        "rc_bounding_box": (None, [P, c_int32, P]),
        "rc_box_visibility": (c_int32, [c_float, c_float, P]),
        "rc_clip_triangle": (c_int32, [c_float, c_float, P, P, P, P, P, P]),
        "rc_facing_camera": (c_int32, [P]),
        "rc_sample": (None, [POINTER(rc_texture), c_float, c_float, P]),
        "rc_color_intensity": (None, [P, c_float, P]),
        "rc_triangle": (c_int64, [c_int32, c_int32, P, P, c_int64, c_int64, c_float, c_float,
                                  c_float, c_int64, c_int64, c_float, c_float, c_float,
                                  c_int64, c_int64, c_float, c_float, c_float, c_int64,
                                  c_int64, c_int64, c_int64, c_float, c_float, c_float, P, P,
                                  POINTER(c_int64)]),
        "rc_clear": (None, [c_int32, c_int32, P, P, P]),
        "rc_dot_grid": (None, [c_int32, c_int32, P, P, c_int32]),
        "rc_pixel": (None, [c_int32, c_int32, P, c_int64, c_int64, P]),
        "rc_go_abs": (c_int64, [c_int64]),
        "rc_rect": (None, [c_int32, c_int32, P, c_int64, c_int64, c_int64, c_int64, P]),
        "rc_line": (None, [c_int32, c_int32, P, c_int64, c_int64, c_int64, c_int64, P]),
        "rc_tile_bounds": (None, [c_int32, c_int32, c_int32, c_int32, POINTER(rc_vec2),
                                  POINTER(rc_vec2)]),
        "rc_fog": (None, [c_int32, c_int32, P, P, ctypes.c_float, ctypes.c_float, P]),
        "rc_light_intensity": (c_float, [rc_vec4, rc_vec3, c_float, c_float]),
        "rc_project_point": (rc_vec4, [rc_vec4, P]),
        "rc_tile_refs": (c_int64, [P, c_int32, c_int32, c_int32]),
        "rc_draw": (c_int32, [POINTER(rc_frame), P, c_int32, P, c_int32, P, c_int32, P, P,
                              POINTER(rc_stats)]),
        "rc_draw_parallel": (c_int32, [POINTER(rc_frame), P, c_int32, P, c_int32, P, c_int32,
                                       P, P, POINTER(rc_stats), c_int32, c_int32, c_int32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def matrix(a) -> rc_matrix:
    m = rc_matrix()
    for i, v in enumerate(np.asarray(a, np.float32).reshape(16)):
        m.m[i] = float(v)
    return m


def matrix_np(m: rc_matrix) -> np.ndarray:
    return np.array(list(m.m), np.float32).reshape(4, 4)


def face_normals(vertices: np.ndarray, vi: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32)
    f = np.ascontiguousarray(vi, np.int32)
    out = np.zeros((len(f), 4), np.float32)
    load().rc_face_normals(_p(v), _p(f), len(f), _p(out))
    return out


def bounding_box(vertices: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32)
    out = np.zeros((8, 4), np.float32)
    load().rc_bounding_box(_p(v), len(v), _p(out))
    return out


class PreparedScene:
    """Scene converted to the oracle's structs (keeps the numpy arrays alive)."""

    def __init__(self, scene):
        lib = load()
        self.scene = scene
        self.keep = []
        fr = scene.frame
        f = rc_frame()
        f.width, f.height, f.flags, f.ref_tiles = fr.width, fr.height, fr.flags, fr.ref_tiles
        f.screen = matrix(fr.screen)
        lt = np.asarray(fr.light, np.float32)
        f.light = rc_vec3(*[float(x) for x in lt])
        f.ambient, f.diffuse = float(np.float32(fr.ambient)), float(np.float32(fr.diffuse))
        f.znear, f.zfar = float(np.float32(fr.znear)), float(np.float32(fr.zfar))
        for i in range(4):
            f.face_color[i] = fr.face_color[i]
            f.clear_color[i] = fr.clear_color[i]
            f.dot_color[i] = fr.dot_color[i]
        f.dot_step = fr.dot_step
        self.frame = f

        self.meshes = (rc_mesh * max(1, len(scene.meshes)))()
        for i, m in enumerate(scene.meshes):
            F = m.num_faces
            v = np.ascontiguousarray(m.vertices, np.float32)
            vi = np.ascontiguousarray(m.vertex_indices, np.int32)
            vn = None if m.vertex_normals is None else np.ascontiguousarray(m.vertex_normals,
                                                                            np.float32)
            ni = (np.zeros((F, 3), np.int32) if m.normal_indices is None
                  else np.ascontiguousarray(m.normal_indices, np.int32))
            uv = (np.zeros((F, 3, 2), np.float32) if m.uvs is None
                  else np.ascontiguousarray(m.uvs, np.float32))
            ft = (np.full(F, -1, np.int32) if m.face_textures is None
                  else np.ascontiguousarray(m.face_textures, np.int32))
            fn = (face_normals(v, vi) if m.face_normals is None
                  else np.ascontiguousarray(m.face_normals, np.float32))
            self.keep += [v, vi, vn, ni, uv, ft, fn]
            rm = self.meshes[i]
            rm.num_vertices, rm.vertices = len(v), _p(v)
            rm.num_vertex_normals = 0 if vn is None else len(vn)
            rm.vertex_normals = _p(vn)
            rm.num_faces, rm.vertex_indices, rm.normal_indices = F, _p(vi), _p(ni)
            rm.uvs, rm.face_textures, rm.face_normals = _p(uv), _p(ft), _p(fn)
            lib.rc_bounding_box(_p(v), len(v), ctypes.cast(rm.bbox, c_void_p))

        self.textures = (rc_texture * max(1, len(scene.textures)))()
        for i, t in enumerate(scene.textures):
            rt = self.textures[i]
            rt.type, rt.width, rt.height = t.type, t.width, t.height
            rt.width_f, rt.height_f = float(np.float32(t.width)), float(np.float32(t.height))
            rt.scale = float(np.float32(t.scale))
            for k in range(4):
                rt.color[k] = t.color[k]
            if t.rgba is not None:
                px = np.ascontiguousarray(t.rgba, np.uint8)
                self.keep.append(px)
                rt.pixels = _p(px)

        self.objects = (rc_object * max(1, len(scene.objects)))()
        for i, o in enumerate(scene.objects):
            self.objects[i].mesh = o.mesh
            self.objects[i].mvp = matrix(o.mvp)
            self.objects[i].world = matrix(o.world)


def draw(scene, prepared: PreparedScene = None, parallel: int = 0, tiles=(4, 4)):
    """Render with the oracle. parallel=0: sequential rc_draw; >0 threads:
    rc_draw_parallel. Returns (rgba, depth, stats)."""
    lib = load()
    ps = prepared or PreparedScene(scene)
    fr = scene.frame
    rgba = np.empty((fr.height, fr.width, 4), np.uint8)
    depth = np.empty((fr.height, fr.width), np.float32)
    st = rc_stats()
    if parallel:
        rc = lib.rc_draw_parallel(ctypes.byref(ps.frame), ps.meshes, len(scene.meshes),
                                  ps.textures, len(scene.textures), ps.objects,
                                  len(scene.objects), _p(rgba), _p(depth), ctypes.byref(st),
                                  parallel, tiles[0], tiles[1])
    else:
        rc = lib.rc_draw(ctypes.byref(ps.frame), ps.meshes, len(scene.meshes), ps.textures,
                         len(scene.textures), ps.objects, len(scene.objects), _p(rgba),
                         _p(depth), ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle rc_draw failed with code {rc}")
    return rgba, depth, st.as_dict()


def fog(rgba: np.ndarray, depth: np.ndarray, fog_start: float, fog_end: float, color) -> np.ndarray:
    """FrameBuffer.Fog (rasterizer.go:185-207) on a copy of an (H, W, 4) frame."""
    lib = load()
    out = np.ascontiguousarray(rgba, np.uint8).copy()
    d = np.ascontiguousarray(depth, np.float32)
    c = np.asarray(color, np.uint8)
    h, w = out.shape[:2]
    lib.rc_fog(w, h, _p(out), _p(d), float(np.float32(fog_start)), float(np.float32(fog_end)),
               _p(c))
    return out
