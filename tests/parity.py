"""Comparison helpers shared by the GPU parity tests and the golden fixtures."""
import hashlib

import numpy as np


def assert_same(scene, got, want, check_stats=True):
    rgba, depth, st = got
    rrgba, rdepth, rst = want
    dbits, rbits = depth.view(np.uint32), rdepth.view(np.uint32)
    bad_c = np.any(rgba != rrgba, axis=-1)
    bad_d = dbits != rbits
    if bad_c.any() or bad_d.any():
        ys, xs = np.where(bad_c | bad_d)
        y, x = int(ys[0]), int(xs[0])
        raise AssertionError(
            f"{scene.name}: {int(bad_c.sum())} colour / {int(bad_d.sum())} depth pixels differ; "
            f"first at (x={x}, y={y}): gpu {rgba[y, x]} {depth[y, x]!r} "
            f"oracle {rrgba[y, x]} {rdepth[y, x]!r}")
    if check_stats:
        for k in ("tpf", "triangles", "fragments_covered", "faces_front", "clip_overflow"):
            assert st[k] == rst[k], f"{scene.name}: stat {k} gpu {st[k]} oracle {rst[k]}"
        assert st["pixels_shaded"] == rst["pixels_written"], scene.name


def digest(rgba, depth, stats):
    """SHA-256 of the colour and depth buffers plus the counters that both the
    oracle and the GPU report."""
    return {"rgba": hashlib.sha256(np.ascontiguousarray(rgba).tobytes()).hexdigest(),
            "depth": hashlib.sha256(np.ascontiguousarray(depth).view(np.uint32).tobytes())
            .hexdigest(),
            "tpf": int(stats["tpf"]), "triangles": int(stats["triangles"]),
            "fragments_covered": int(stats["fragments_covered"]),
            "faces_front": int(stats["faces_front"]),
            "pixels_shaded": int(stats.get("pixels_shaded", stats.get("pixels_written", -1)))}
