"""Python handle on the serving loop (gr_server_* of the C ABI) and the
open-loop load generator used to measure it.

A Server runs on one Context whose meshes and textures are already resident;
requests (a frame and its objects, SceneData) are submitted from any thread
and come back as colour/depth in pinned host memory with their timings.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .device import Context, frame_struct
from .scenes import SceneData


@dataclass
class ServeResult:
    rgba: np.ndarray       # (H, W, 4) uint8, copied out of the server's pinned buffer
    depth: Optional[np.ndarray]  # (H, W) float32; None when the server reads colour only
    stats: dict
    submit_ns: int
    start_ns: int
    done_ns: int
    batch_size: int

    @property
    def latency_ms(self) -> float:
        return (self.done_ns - self.submit_ns) / 1e6


class PreparedRequest:
    """ctypes arguments of one request, built ahead of a timed run."""

    def __init__(self, scene: SceneData, mesh_handles: Sequence[int]):
        self.scene = scene
        self.frame = frame_struct(scene.frame)
        self.objects = Context.objects_struct(scene, mesh_handles)
        self.num_objects = len(scene.objects)


class Server:
    def __init__(self, ctx: Context, target_budget: int = 1 << 30, max_batch: int = 16,
                 max_inflight: int = 3, colour_only: bool = False, max_queue: int = 0,
                 deadline_ms: int = 0):
        self.ctx = ctx
        self.lib = ctx.lib
        cfg = _lib.gr_server_config(target_budget, max_batch, max_inflight, int(colour_only),
                                    max_queue, deadline_ms, 0)
        h = ctypes.c_void_p()
        _lib.check(self.lib.gr_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            self.lib.gr_server_destroy(self.handle)
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

    def submit(self, req: PreparedRequest) -> int:
        t = ctypes.c_int64()
        _lib.check(self.lib.gr_server_submit(self.handle, ctypes.byref(req.frame), req.objects,
                                             req.num_objects, ctypes.byref(t)))
        return t.value

    def wait_raw(self, ticket: int, timeout_ms: int = -1) -> _lib.gr_serve_result:
        res = _lib.gr_serve_result()
        _lib.check(self.lib.gr_server_wait(self.handle, ticket, timeout_ms, ctypes.byref(res)))
        return res

    def release(self, ticket: int) -> None:
        _lib.check(self.lib.gr_server_release(self.handle, ticket))

    def wait(self, ticket: int, timeout_ms: int = -1, copy: bool = True) -> ServeResult:
        """Wait for a request and release it; copy=False skips the pixel copies."""
        r = self.wait_raw(ticket, timeout_ms)
        h, w = r.height, r.width
        rgba = depth = None
        if copy:
            rgba = np.ctypeslib.as_array(ctypes.cast(r.rgba, ctypes.POINTER(ctypes.c_uint8)),
                                         shape=(h, w, 4)).copy()
            if r.depth:
                depth = np.ctypeslib.as_array(
                    ctypes.cast(r.depth, ctypes.POINTER(ctypes.c_float)), shape=(h, w)).copy()
        out = ServeResult(rgba, depth, r.stats.as_dict(), r.submit_ns, r.start_ns, r.done_ns,
                          r.batch_size)
        self.release(ticket)
        return out


def poisson_arrivals(n: int, qps: float, seed: int) -> np.ndarray:
    """Arrival offsets in seconds of n requests of a Poisson stream at qps."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / qps, n))


def replay(server: Server, requests: Sequence[PreparedRequest], arrivals: np.ndarray):
    """Open-loop load (gr_server_replay): request i % len(requests) is
    submitted at arrivals[i] seconds; returns (latency_ms, batch_size) arrays,
    latency from submission to colour and depth in host memory (-1e-6 ms for a
    request a full queue refused, -2e-6 ms for one shed past the deadline)."""
    d = len(requests)
    frames = (_lib.gr_frame * d)()
    total = sum(r.num_objects for r in requests)
    objs = (_lib.gr_object * max(1, total))()
    first = np.zeros(d, np.int32)
    count = np.zeros(d, np.int32)
    k = 0
    for i, r in enumerate(requests):
        frames[i] = r.frame
        first[i], count[i] = k, r.num_objects
        for j in range(r.num_objects):
            objs[k + j] = r.objects[j]
        k += r.num_objects
    arr = np.ascontiguousarray(arrivals, np.float64)
    n = len(arr)
    lat = np.zeros(n, np.int64)
    bs = np.zeros(n, np.int32)
    _lib.check(server.lib.gr_server_replay(
        server.handle, frames, objs, first.ctypes.data, count.ctypes.data, d,
        arr.ctypes.data, n, lat.ctypes.data, bs.ctypes.data))
    return lat / 1e6, bs
