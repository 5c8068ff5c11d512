"""ctypes binding of the C ABI declared in include/gorender_hip.h.

The shared library is built in-tree (gorender_amd/libgorender_hip.so). There is
no fallback: importing the render path without the library raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_double, c_float, c_int32, c_int64,
                    c_uint8, c_uint32, c_void_p)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgorender_hip.so")
# Another build of the library can be named here (the sanitizer test loads the
# host-only ASan/UBSan build this way); symbols it lacks are left unbound.
ALT_LIB_PATH = os.environ.get("GORENDER_HIP_LIB")

GR_OK = 0
GR_ERR_INVALID_ARGUMENT = 1
GR_ERR_INVALID_HANDLE = 2
GR_ERR_OUT_OF_MEMORY = 3
GR_ERR_DEVICE = 4
GR_ERR_INDEX_OUT_OF_RANGE = 5
GR_ERR_UNSUPPORTED = 6
GR_ERR_CAPACITY = 7
GR_ERR_CLIP_OVERFLOW = 8
GR_ERR_TIMEOUT = 9
GR_ERR_BUSY = 10

GR_SHOW_FACES = 1 << 0
GR_BACKFACE_CULLING = 1 << 1
GR_LIGHTING = 1 << 2
GR_FLAT_SHADING = 1 << 3
GR_FRUSTUM_CLIPPING = 1 << 4
GR_SHOW_TEXTURES = 1 << 5
GR_SHOW_EDGES = 1 << 6
GR_SHOW_VERTICES = 1 << 7
GR_SHOW_TILE_BOUNDS = 1 << 8
GR_SHOW_CROSSHAIR = 1 << 9
GR_DEFAULT_FLAGS = (GR_SHOW_FACES | GR_BACKFACE_CULLING | GR_LIGHTING | GR_FRUSTUM_CLIPPING
                    | GR_SHOW_TEXTURES)

GR_TEXTURE_SOLID = 0
GR_TEXTURE_IMAGE = 1
GR_TEXTURE_IMAGE_FAST = 2


class gr_texture_desc(Structure):
    _fields_ = [("type", c_int32), ("width", c_int32), ("height", c_int32),
                ("scale", c_float), ("color", c_uint8 * 4)]


class gr_mesh_desc(Structure):
    _fields_ = [("num_vertices", c_int32), ("vertices", c_void_p),
                ("num_vertex_normals", c_int32), ("vertex_normals", c_void_p),
                ("num_faces", c_int32), ("vertex_indices", c_void_p),
                ("normal_indices", c_void_p), ("uvs", c_void_p),
                ("face_textures", c_void_p), ("face_normals", c_void_p)]


class gr_object(Structure):
    _fields_ = [("mesh", c_int32), ("reserved", c_int32), ("mvp", c_float * 16),
                ("world", c_float * 16)]


class gr_frame(Structure):
    _fields_ = [("width", c_int32), ("height", c_int32), ("flags", c_uint32),
                ("ref_tiles", c_int32), ("screen", c_float * 16), ("light", c_float * 3),
                ("ambient", c_float), ("diffuse", c_float), ("znear", c_float),
                ("zfar", c_float), ("face_color", c_uint8 * 4), ("clear_color", c_uint8 * 4),
                ("dot_color", c_uint8 * 4), ("dot_step", c_int32)]


class gr_stats(Structure):
    _fields_ = [("tpf", c_int64), ("triangles", c_int64), ("fragments_covered", c_int64),
                ("pixels_shaded", c_int64), ("faces_front", c_int64),
                ("clip_overflow", c_int64), ("refs", c_int64), ("tiles_split", c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "reserved"}


class gr_batch_frame(Structure):
    _fields_ = [("frame", gr_frame), ("first_object", c_int32), ("num_objects", c_int32),
                ("d_rgba", c_void_p), ("d_depth", c_void_p)]


class gr_config(Structure):
    _fields_ = [("tile_w", c_int32), ("tile_h", c_int32), ("refs_capacity", c_int64),
                ("raster_waves", c_int32), ("tile_split", c_int32), ("pipelines", c_int32),
                ("raster_grid", c_int32)]


class gr_tune_result(Structure):
    _fields_ = [("tile_w", c_int32), ("tile_h", c_int32), ("tile_split", c_int32),
                ("raster_waves", c_int32), ("ms_per_launch", c_double),
                ("checksum", ctypes.c_uint64)]


class gr_server_config(Structure):
    _fields_ = [("target_budget", c_int64), ("max_batch", c_int32), ("max_inflight", c_int32),
                ("colour_only", c_int32), ("max_queue", c_int32), ("deadline_ms", c_int32),
                ("reserved", c_int32)]


class gr_serve_result(Structure):
    _fields_ = [("status", c_int32), ("width", c_int32), ("height", c_int32),
                ("batch_size", c_int32), ("rgba", c_void_p), ("depth", c_void_p),
                ("stats", gr_stats), ("submit_ns", c_int64), ("start_ns", c_int64),
                ("done_ns", c_int64)]


# name -> (restype, argtypes); every symbol of include/gorender_hip.h
SIGNATURES = {
    "gr_create": (c_int32, [c_int32, POINTER(c_void_p)]),
    "gr_destroy": (None, [c_void_p]),
    "gr_last_error": (c_char_p, []),
    "gr_abi_version": (c_int32, []),
    "gr_set_config": (c_int32, [c_void_p, POINTER(gr_config)]),
    "gr_get_config": (c_int32, [c_void_p, POINTER(gr_config)]),
    "gr_set_stream": (c_int32, [c_void_p, c_void_p]),
    "gr_get_stream": (c_void_p, [c_void_p]),
    "gr_texture_create": (c_int32, [c_void_p, POINTER(gr_texture_desc), c_void_p,
                                    POINTER(c_int32)]),
    "gr_texture_destroy": (c_int32, [c_void_p, c_int32]),
    "gr_texture_kind_for_size": (c_int32, [c_int32, c_int32]),
    "gr_mesh_create": (c_int32, [c_void_p, POINTER(gr_mesh_desc), POINTER(c_int32)]),
    "gr_mesh_destroy": (c_int32, [c_void_p, c_int32]),
    "gr_mesh_device_bytes": (c_int64, [c_void_p, c_int32]),
    "gr_frame_defaults": (None, [POINTER(gr_frame), c_int32, c_int32]),
    "gr_matrix_multiply": (None, [c_void_p, c_void_p, c_void_p]),
    "gr_matrix_world": (None, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "gr_matrix_view": (None, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "gr_matrix_perspective": (None, [c_float, c_float, c_float, c_float, c_void_p]),
    "gr_matrix_screen": (None, [c_int32, c_int32, c_void_p]),
    "gr_matrix_mvp": (None, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "gr_vec3_normalize": (None, [c_void_p, c_void_p]),
    "gr_box_visibility": (c_int32, [c_void_p, c_void_p, c_float, c_float]),
    "gr_draw": (c_int32, [c_void_p, POINTER(gr_frame), c_void_p, c_int32, c_void_p, c_void_p,
                          POINTER(gr_stats)]),
    "gr_draw_device": (c_int32, [c_void_p, POINTER(gr_frame), c_void_p, c_int32, c_void_p,
                                 c_void_p]),
    "gr_draw_batch": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_int32]),
    "gr_draw_device_noalloc": (c_int32, [c_void_p, POINTER(gr_frame), c_void_p, c_int32,
                                         c_void_p, c_void_p]),
    "gr_batch_graph_create": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_int32,
                                        POINTER(c_void_p)]),
    "gr_batch_graph_update": (c_int32, [c_void_p, c_void_p, c_int32]),
    "gr_batch_graph_launch": (c_int32, [c_void_p]),
    "gr_batch_graph_destroy": (None, [c_void_p]),
    "gr_autotune": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_int32,
                              c_void_p, c_int32, POINTER(c_int32), POINTER(gr_config)]),
    "gr_present_draw": (c_int32, [c_void_p, POINTER(gr_frame), c_void_p, c_int32, c_int32]),
    "gr_present_wait": (c_int32, [c_void_p, POINTER(ctypes.POINTER(c_uint8)),
                                  POINTER(ctypes.POINTER(ctypes.c_float)), POINTER(gr_stats)]),
    "gr_synchronize": (c_int32, [c_void_p]),
    "gr_last_stats": (c_int32, [c_void_p, c_void_p, c_int32, POINTER(c_int32)]),
    "gr_device_alloc": (c_int32, [c_void_p, ctypes.c_size_t, POINTER(c_void_p)]),
    "gr_device_free": (c_int32, [c_void_p, c_void_p]),
    "gr_device_fill": (c_int32, [c_void_p, c_void_p, c_int32, ctypes.c_size_t]),
    "gr_device_copy": (c_int32, [c_void_p, c_void_p, c_void_p, ctypes.c_size_t]),
    "gr_fog": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_float, c_float,
                         POINTER(c_uint8)]),
    "gr_obj_open": (c_int32, [ctypes.c_char_p, c_int32, POINTER(c_void_p)]),
    "gr_obj_close": (None, [c_void_p]),
    "gr_obj_num_meshes": (c_int32, [c_void_p]),
    "gr_obj_num_textures": (c_int32, [c_void_p]),
    "gr_obj_mesh": (c_int32, [c_void_p, c_int32, POINTER(gr_mesh_desc)]),
    "gr_obj_texture": (c_int32, [c_void_p, c_int32, POINTER(gr_texture_desc),
                                 POINTER(POINTER(ctypes.c_uint8))]),
    "gr_obj_upload": (c_int32, [c_void_p, c_void_p, POINTER(c_int32), c_int32,
                                POINTER(c_int32)]),
    "gr_png_decode": (c_int32, [ctypes.c_char_p, POINTER(c_int32), POINTER(c_int32),
                                POINTER(POINTER(ctypes.c_uint8))]),
    "gr_host_free": (None, [c_void_p]),
    "gr_texture_load": (c_int32, [c_void_p, ctypes.c_char_p, POINTER(c_int32)]),
    "gr_check_exact_division": (c_int32, [c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                          POINTER(ctypes.c_uint64), POINTER(ctypes.c_float)]),
    "gr_device_checksum": (c_int32, [c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "gr_copy_to_host": (c_int32, [c_void_p, c_void_p, c_void_p, ctypes.c_size_t]),
    "gr_set_kernel_timing": (c_int32, [c_void_p, c_int32]),
    "gr_kernel_times": (c_int32, [c_void_p, POINTER(c_char_p), POINTER(c_double),
                                  POINTER(c_int64), c_int32]),
    "gr_reset_kernel_times": (None, [c_void_p]),
    "gr_server_create": (c_int32, [c_void_p, POINTER(gr_server_config), POINTER(c_void_p)]),
    "gr_server_destroy": (None, [c_void_p]),
    "gr_server_submit": (c_int32, [c_void_p, POINTER(gr_frame), c_void_p, c_int32,
                                   POINTER(c_int64)]),
    "gr_server_wait": (c_int32, [c_void_p, c_int64, c_int32, POINTER(gr_serve_result)]),
    "gr_server_release": (c_int32, [c_void_p, c_int64]),
    "gr_server_replay": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
                                   c_void_p, c_int32, c_void_p, c_void_p]),
}

_lib = None


class GorenderError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gorender_hip status {status}: {msg}")
        self.status = status


def load() -> ctypes.CDLL:
    """Load the in-tree HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = ALT_LIB_PATH or LIB_PATH
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: run __graft_entry__.build() first")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        if ALT_LIB_PATH and not hasattr(lib, name):
            continue
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != GR_OK:
        msg = load().gr_last_error()
        raise GorenderError(status, msg.decode() if msg else "")
