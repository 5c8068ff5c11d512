"""The draw that allocates nothing (gr_draw_device_noalloc) against the draw it
is a variant of (gr_draw_device), through the C ABI: the same colour and depth
buffers bit for bit and the same counters, on two calls in a row, the second
with a smaller frame and fewer faces (the fixed device array keeps the first
call's contents). Sizes beyond the array are refused before any launch."""
import numpy as np
import pytest

from gorender_amd import _lib
from gorender_amd import scenes as S
from oracle import refcpu

import scene_cases as C
from parity import assert_same

pytestmark = pytest.mark.gpu


def _draw(ctx, scene, noalloc):
    fr = scene.frame
    mh, th = ctx.upload_scene(scene)
    rgba_b, depth_b = ctx.frame_buffers(fr.width, fr.height)
    try:
        rgba_b.fill(0x5A)
        depth_b.fill(0x5A)
        if noalloc:
            ctx.draw_device_noalloc(scene, mh, rgba_b.ptr, depth_b.ptr)
        else:
            ctx.draw_device(scene, mh, rgba_b.ptr, depth_b.ptr)
        ctx.synchronize()
        st = ctx.last_stats()
        assert len(st) == 1
        rgba = rgba_b.to_numpy(np.uint8, (fr.height, fr.width, 4))
        depth = depth_b.to_numpy(np.float32, (fr.height, fr.width))
    finally:
        rgba_b.free()
        depth_b.free()
        for h in mh:
            ctx.destroy_mesh(h)
        for h in th:
            ctx.destroy_texture(h)
    return rgba, depth, st[0]


def _same(scene, a, b):
    rgba, depth, st = a
    rgba2, depth2, st2 = b
    assert np.array_equal(rgba, rgba2), scene.name
    assert np.array_equal(depth.view(np.uint32), depth2.view(np.uint32)), scene.name
    for k in ("tpf", "triangles", "fragments_covered", "pixels_shaded", "faces_front",
              "clip_overflow", "refs", "tiles_split"):
        assert st[k] == st2[k], f"{scene.name}: {k} {st[k]} != {st2[k]}"


def test_noalloc_equals_draw_device_twice(ctx):
    # First call: config 1 (icosphere L4, 5,120 faces, 640x360); second call:
    # a smaller textured frame (160x96, 480 faces) over the array the first
    # call left behind.
    for scene in (S.config1("ref"), C.sphere()):
        want = _draw(ctx, scene, noalloc=False)
        got = _draw(ctx, scene, noalloc=True)
        _same(scene, got, want)
        # and the oracle, as every other draw is checked
        ro = refcpu.draw(scene)
        assert_same(scene, (got[0], got[1], got[2]), ro)


def test_noalloc_refuses_large_frames(ctx):
    scene = S.config2()  # 200,000 faces: records alone exceed the array
    fr = scene.frame
    mh, th = ctx.upload_scene(scene)
    rgba_b, depth_b = ctx.frame_buffers(fr.width, fr.height)
    try:
        with pytest.raises(_lib.GorenderError) as e:
            ctx.draw_device_noalloc(scene, mh, rgba_b.ptr, depth_b.ptr)
        assert e.value.status == _lib.GR_ERR_INVALID_ARGUMENT
        assert "no-allocation" in str(e.value)
        ctx.synchronize()
    finally:
        rgba_b.free()
        depth_b.free()
        for h in mh:
            ctx.destroy_mesh(h)
        for h in th:
            ctx.destroy_texture(h)
