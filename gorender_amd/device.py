"""Python handle on a gorender_hip context (one GPU, one HIP stream).

Thin layer over the C ABI: uploads meshes/textures to HBM once, then draws
scenes synchronously into numpy buffers or asynchronously into device memory.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .scenes import FrameData, MeshData, SceneData, TextureData


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def frame_struct(fr: FrameData) -> _lib.gr_frame:
    f = _lib.gr_frame()
    f.width, f.height, f.flags, f.ref_tiles = fr.width, fr.height, fr.flags, fr.ref_tiles
    for i, v in enumerate(np.asarray(fr.screen, np.float32).reshape(16)):
        f.screen[i] = float(v)
    for i, v in enumerate(np.asarray(fr.light, np.float32).reshape(3)):
        f.light[i] = float(v)
    f.ambient, f.diffuse = float(np.float32(fr.ambient)), float(np.float32(fr.diffuse))
    f.znear, f.zfar = float(np.float32(fr.znear)), float(np.float32(fr.zfar))
    for i in range(4):
        f.face_color[i] = fr.face_color[i]
        f.clear_color[i] = fr.clear_color[i]
        f.dot_color[i] = fr.dot_color[i]
    f.dot_step = fr.dot_step
    return f


class Context:
    """A gorender_hip context on one GPU."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self.lib.gr_create(device, ctypes.byref(h)))
        self.handle = h
        self._mesh_cache: Dict[int, int] = {}
        self._tex_cache: Dict[int, int] = {}

    def close(self) -> None:
        if self.handle:
            self.lib.gr_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- configuration -------------------------------------------------
    def set_config(self, tile_w: int = 0, tile_h: int = 0, refs_capacity: int = 0,
                   tile_split: int = 0, pipelines: int = 0, raster_grid: int = 0) -> None:
        """Zero keeps the current value (gr_set_config)."""
        c = _lib.gr_config(tile_w, tile_h, refs_capacity, 0, tile_split, pipelines, raster_grid)
        _lib.check(self.lib.gr_set_config(self.handle, ctypes.byref(c)))

    def config(self) -> _lib.gr_config:
        c = _lib.gr_config()
        _lib.check(self.lib.gr_get_config(self.handle, ctypes.byref(c)))
        return c

    def set_stream(self, stream_ptr: Optional[int]) -> None:
        _lib.check(self.lib.gr_set_stream(self.handle, stream_ptr))

    def stream(self) -> int:
        return self.lib.gr_get_stream(self.handle) or 0

    # ---- assets ----------------------------------------------------------
    def upload_texture(self, t: TextureData) -> int:
        d = _lib.gr_texture_desc(t.type, t.width, t.height, float(np.float32(t.scale)))
        for i in range(4):
            d.color[i] = t.color[i]
        rgba = None
        if t.rgba is not None:
            rgba = np.ascontiguousarray(t.rgba, np.uint8)
        h = ctypes.c_int32()
        _lib.check(self.lib.gr_texture_create(self.handle, ctypes.byref(d), _ptr(rgba),
                                              ctypes.byref(h)))
        return h.value

    def upload_mesh(self, m: MeshData, tex_handles: Sequence[int] = ()) -> int:
        v = np.ascontiguousarray(m.vertices, np.float32)
        vi = np.ascontiguousarray(m.vertex_indices, np.int32)
        vn = None if m.vertex_normals is None else np.ascontiguousarray(m.vertex_normals,
                                                                        np.float32)
        ni = None if m.normal_indices is None else np.ascontiguousarray(m.normal_indices,
                                                                        np.int32)
        uv = None if m.uvs is None else np.ascontiguousarray(m.uvs, np.float32)
        ft = None
        if m.face_textures is not None:
            lut = np.asarray(list(tex_handles) + [-1], np.int32)
            ft = np.ascontiguousarray(np.where(m.face_textures >= 0,
                                               lut[np.clip(m.face_textures, 0, None)], -1),
                                      np.int32)
        fn = None if m.face_normals is None else np.ascontiguousarray(m.face_normals, np.float32)
        d = _lib.gr_mesh_desc(len(v), _ptr(v), 0 if vn is None else len(vn), _ptr(vn), len(vi),
                              _ptr(vi), _ptr(ni), _ptr(uv), _ptr(ft), _ptr(fn))
        h = ctypes.c_int32()
        _lib.check(self.lib.gr_mesh_create(self.handle, ctypes.byref(d), ctypes.byref(h)))
        return h.value

    def destroy_mesh(self, h: int) -> None:
        _lib.check(self.lib.gr_mesh_destroy(self.handle, h))

    def destroy_texture(self, h: int) -> None:
        _lib.check(self.lib.gr_texture_destroy(self.handle, h))

    def mesh_bytes(self, h: int) -> int:
        return int(self.lib.gr_mesh_device_bytes(self.handle, h))

    def upload_scene(self, s: SceneData) -> Tuple[List[int], List[int]]:
        th = [self.upload_texture(t) for t in s.textures]
        mh = [self.upload_mesh(m, th) for m in s.meshes]
        return mh, th

    # ---- drawing -----------------------------------------------------------
    @staticmethod
    def objects_struct(s: SceneData, mesh_handles: Sequence[int]):
        arr = (_lib.gr_object * max(1, len(s.objects)))()
        for i, o in enumerate(s.objects):
            arr[i].mesh = mesh_handles[o.mesh]
            for k, val in enumerate(np.asarray(o.mvp, np.float32).reshape(16)):
                arr[i].mvp[k] = float(val)
            for k, val in enumerate(np.asarray(o.world, np.float32).reshape(16)):
                arr[i].world[k] = float(val)
        return arr

    def draw(self, s: SceneData, mesh_handles: Optional[Sequence[int]] = None):
        """Synchronous Draw; returns (rgba (H,W,4) uint8, depth (H,W) float32, stats)."""
        if mesh_handles is None:
            mesh_handles, _ = self.upload_scene(s)
        fr = frame_struct(s.frame)
        objs = self.objects_struct(s, mesh_handles)
        rgba = np.empty((s.frame.height, s.frame.width, 4), np.uint8)
        depth = np.empty((s.frame.height, s.frame.width), np.float32)
        st = _lib.gr_stats()
        _lib.check(self.lib.gr_draw(self.handle, ctypes.byref(fr), objs, len(s.objects),
                                    _ptr(rgba), _ptr(depth), ctypes.byref(st)))
        return rgba, depth, st.as_dict()

    def draw_device(self, s: SceneData, mesh_handles: Sequence[int], d_rgba: int,
                    d_depth: int) -> None:
        fr = frame_struct(s.frame)
        objs = self.objects_struct(s, mesh_handles)
        _lib.check(self.lib.gr_draw_device(self.handle, ctypes.byref(fr), objs,
                                           len(s.objects), d_rgba, d_depth))

    def draw_device_noalloc(self, s: SceneData, mesh_handles: Sequence[int], d_rgba: int,
                            d_depth: int) -> None:
        """gr_draw_device_noalloc: the same draw with every per-draw buffer in the
        library's fixed device array (no allocation)."""
        fr = frame_struct(s.frame)
        objs = self.objects_struct(s, mesh_handles)
        _lib.check(self.lib.gr_draw_device_noalloc(self.handle, ctypes.byref(fr), objs,
                                                   len(s.objects), d_rgba, d_depth))

    def present_draw(self, s: SceneData, mesh_handles: Sequence[int],
                     with_depth: bool = True) -> None:
        """gr_present_draw: queue a frame whose pixels land in pinned host memory."""
        fr = frame_struct(s.frame)
        objs = self.objects_struct(s, mesh_handles)
        _lib.check(self.lib.gr_present_draw(self.handle, ctypes.byref(fr), objs, len(s.objects),
                                            1 if with_depth else 0))

    def present_wait(self, width: int, height: int):
        """gr_present_wait: (rgba, depth or None, stats) of the oldest queued
        frame, copied out of the context's pinned buffers."""
        rp = ctypes.POINTER(ctypes.c_uint8)()
        dp = ctypes.POINTER(ctypes.c_float)()
        st = _lib.gr_stats()
        _lib.check(self.lib.gr_present_wait(self.handle, ctypes.byref(rp), ctypes.byref(dp),
                                            ctypes.byref(st)))
        rgba = np.ctypeslib.as_array(rp, (height, width, 4)).copy()
        depth = np.ctypeslib.as_array(dp, (height, width)).copy() if dp else None
        return rgba, depth, st.as_dict()

    def prepare_batch(self, scenes: Sequence[SceneData], mesh_handles: Sequence[Sequence[int]],
                      outputs: Sequence[Tuple[int, int]]):
        """Build the ctypes arguments of a gr_draw_batch call once (for loops
        that resubmit the same batch)."""
        nf = len(scenes)
        frames = (_lib.gr_batch_frame * nf)()
        total = sum(len(s.objects) for s in scenes)
        objs = (_lib.gr_object * max(1, total))()
        k = 0
        for f, s in enumerate(scenes):
            frames[f].frame = frame_struct(s.frame)
            frames[f].first_object = k
            frames[f].num_objects = len(s.objects)
            frames[f].d_rgba, frames[f].d_depth = outputs[f]
            part = self.objects_struct(s, mesh_handles[f])
            for i in range(len(s.objects)):
                objs[k + i] = part[i]
            k += len(s.objects)
        return (frames, nf, objs, total)

    def submit_batch(self, prepared) -> None:
        frames, nf, objs, total = prepared
        _lib.check(self.lib.gr_draw_batch(self.handle, frames, nf, objs, total))

    def draw_batch(self, scenes: Sequence[SceneData], mesh_handles: Sequence[Sequence[int]],
                   outputs: Sequence[Tuple[int, int]]) -> None:
        """One asynchronous batch: frame k draws scenes[k] into outputs[k]."""
        self.submit_batch(self.prepare_batch(scenes, mesh_handles, outputs))

    def batch_graph(self, prepared) -> "BatchGraph":
        """gr_batch_graph_create over a prepared batch (see prepare_batch)."""
        return BatchGraph(self, prepared)

    def autotune(self, prepared, reps: int = 3, max_results: int = 64):
        """gr_autotune over a prepared batch: returns (best config as a dict,
        every candidate tried as dicts); the context keeps the best config."""
        frames, nf, objs, total = prepared
        res = (_lib.gr_tune_result * max_results)()
        n = ctypes.c_int32()
        best = _lib.gr_config()
        _lib.check(self.lib.gr_autotune(self.handle, frames, nf, objs, total, reps, res,
                                        max_results, ctypes.byref(n), ctypes.byref(best)))
        tried = [{"tile": [res[i].tile_w, res[i].tile_h], "tile_split": res[i].tile_split,
                  "raster_waves": res[i].raster_waves,
                  "ms_per_launch": round(res[i].ms_per_launch, 4),
                  "checksum": f"{res[i].checksum:016x}"} for i in range(n.value)]
        return ({"tile": [best.tile_w, best.tile_h], "tile_split": best.tile_split,
                 "raster_waves": best.raster_waves, "refs_capacity": best.refs_capacity},
                tried)

    def synchronize(self) -> None:
        _lib.check(self.lib.gr_synchronize(self.handle))

    def last_stats(self, max_frames: int = 1024) -> List[dict]:
        arr = (_lib.gr_stats * max_frames)()
        n = ctypes.c_int32()
        _lib.check(self.lib.gr_last_stats(self.handle, arr, max_frames, ctypes.byref(n)))
        return [arr[i].as_dict() for i in range(n.value)]

    # ---- device memory -----------------------------------------------------
    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def frame_buffers(self, width: int, height: int) -> Tuple["DeviceBuffer", "DeviceBuffer"]:
        """Device colour (H*W*4 bytes) and depth (H*W floats) buffers."""
        return self.alloc(width * height * 4), self.alloc(width * height * 4)

    def fog(self, rgba: "DeviceBuffer", depth: "DeviceBuffer", width: int, height: int,
            fog_start: float, fog_end: float, color) -> None:
        """FrameBuffer.Fog (rasterizer.go:185-207) over a frame in device
        memory, on the context stream."""
        c = (ctypes.c_uint8 * 4)(*[int(v) for v in color])
        _lib.check(self.lib.gr_fog(self.handle, rgba.ptr, depth.ptr, width, height,
                                   float(fog_start), float(fog_end), c))

    # ---- measurement -------------------------------------------------------
    def set_kernel_timing(self, on: bool) -> None:
        _lib.check(self.lib.gr_set_kernel_timing(self.handle, 1 if on else 0))

    def kernel_times(self) -> Dict[str, Tuple[float, int]]:
        n = 32
        names = (ctypes.c_char_p * n)()
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int64 * n)()
        k = self.lib.gr_kernel_times(self.handle, names, ms, cnt, n)
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k)}

    def reset_kernel_times(self) -> None:
        self.lib.gr_reset_kernel_times(self.handle)


class DeviceBuffer:
    """Device memory owned by a context (for the asynchronous draw forms)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _lib.check(ctx.lib.gr_device_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def fill(self, value: int) -> None:
        _lib.check(self.ctx.lib.gr_device_fill(self.ctx.handle, self.ptr, value, self.nbytes))

    def copy_from(self, src: "DeviceBuffer", nbytes: int | None = None) -> None:
        n = min(self.nbytes, src.nbytes) if nbytes is None else int(nbytes)
        assert n <= self.nbytes and n <= src.nbytes
        _lib.check(self.ctx.lib.gr_device_copy(self.ctx.handle, self.ptr, src.ptr, n))

    def checksum_into(self, digest: "DeviceBuffer", nbytes: int | None = None) -> None:
        """Adds the digest of the first nbytes (all by default) to the 8-byte
        device buffer `digest` (gr_device_checksum, on the context stream)."""
        n = self.nbytes if nbytes is None else int(nbytes)
        _lib.check(self.ctx.lib.gr_device_checksum(self.ctx.handle, self.ptr, n, digest.ptr))

    def to_numpy(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _lib.check(self.ctx.lib.gr_copy_to_host(self.ctx.handle, out.ctypes.data_as(
            ctypes.c_void_p), self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr and self.ctx.handle:
            _lib.check(self.ctx.lib.gr_device_free(self.ctx.handle, self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BatchGraph:
    """One batch's launches captured in HIP graphs (gr_batch_graph_*)."""

    def __init__(self, ctx: Context, prepared):
        self.ctx = ctx
        frames, nf, objs, total = prepared
        self._prepared = prepared
        h = ctypes.c_void_p()
        _lib.check(ctx.lib.gr_batch_graph_create(ctx.handle, frames, nf, objs, total,
                                                 ctypes.byref(h)))
        self.handle = h.value

    def launch(self) -> None:
        _lib.check(self.ctx.lib.gr_batch_graph_launch(self.handle))

    def update(self, objs, total: int) -> None:
        _lib.check(self.ctx.lib.gr_batch_graph_update(self.handle, objs, total))

    def destroy(self) -> None:
        if self.handle and self.ctx.handle:
            self.ctx.lib.gr_batch_graph_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
