"""The draw that allocates nothing (gr_draw_device_noalloc) against the draw it
is a variant of (gr_draw_device), through the C ABI: the same colour and depth
buffers bit for bit and the same counters, on two calls in a row, the second
with a smaller frame and fewer faces (the fixed device array keeps the first
call's contents). Frames whose fixed parts exceed the array are refused
before any launch; one that needs more triangle records or tile-list entries
than the array holds is reported by gr_synchronize."""
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


def test_noalloc_frame_past_per_block_records(ctx):
    # 50,000 faces: records reserved at 1792 per 256-face block would take
    # 67 MB; 320 per block plus the overflow records fit the array.
    scene = S.config2(nlat=125, nlon=200)
    want = _draw(ctx, scene, noalloc=False)
    got = _draw(ctx, scene, noalloc=True)
    _same(scene, got, want)


def test_noalloc_reports_capacity(ctx):
    # Sixteen copies of a 480-face sphere filling a 4096 x 4096 frame in 8 x 8
    # tiles: millions of tile-list entries, more than the array holds. The
    # draw is refused at synchronize, as a capacity error of that array, and
    # the context draws normally afterwards.
    mesh = S.latlong_sphere(12, 20, S.SplitMix64(7))
    fr = S.FrameData(4096, 4096)
    obj = S.camera_object(fr, 0, (0, 0, 2.5), (0, 0, -1))
    scene = S.SceneData(fr, [mesh], [], [obj] * 16, name="sixteen-spheres")
    old = ctx.config()
    ctx.set_config(tile_w=8, tile_h=8)
    mh, th = ctx.upload_scene(scene)
    rgba_b, depth_b = ctx.frame_buffers(fr.width, fr.height)
    try:
        ctx.draw_device_noalloc(scene, mh, rgba_b.ptr, depth_b.ptr)
        with pytest.raises(_lib.GorenderError) as e:
            ctx.synchronize()
        assert e.value.status == _lib.GR_ERR_CAPACITY
        assert "no-allocation" in str(e.value)
        ctx.synchronize()
    finally:
        ctx.set_config(tile_w=old.tile_w, tile_h=old.tile_h)
        rgba_b.free()
        depth_b.free()
        for h in mh:
            ctx.destroy_mesh(h)
        for h in th:
            ctx.destroy_texture(h)
    scene = C.sphere()
    _same(scene, _draw(ctx, scene, noalloc=True), _draw(ctx, scene, noalloc=False))


def test_noalloc_refuses_large_frames(ctx):
    # 16384 x 16384 pixels in 8 x 8 tiles: the tile counters, offsets and
    # cursors alone (4M tiles) exceed the array; refused before any launch.
    scene = C.sphere(w=16384, h=16384)
    fr = scene.frame
    old = ctx.config()
    ctx.set_config(tile_w=8, tile_h=8)
    mh, th = ctx.upload_scene(scene)
    rgba_b, depth_b = ctx.frame_buffers(fr.width, fr.height)
    try:
        with pytest.raises(_lib.GorenderError) as e:
            ctx.draw_device_noalloc(scene, mh, rgba_b.ptr, depth_b.ptr)
        assert e.value.status == _lib.GR_ERR_INVALID_ARGUMENT
        assert "no-allocation" in str(e.value)
        ctx.synchronize()
    finally:
        ctx.set_config(tile_w=old.tile_w, tile_h=old.tile_h)
        rgba_b.free()
        depth_b.free()
        for h in mh:
            ctx.destroy_mesh(h)
        for h in th:
            ctx.destroy_texture(h)
