"""The present path (gr_present_draw / gr_present_wait): the window loop's
double-buffered Pixels / Pixels2 (main.go:201-227, rasterizer.go:32-34) as two
frames in flight whose pixels land in pinned host memory. Every frame equals
the oracle's bit for bit, with or without its depth copied out; a frame whose
tile lists overflow is drawn again inside gr_present_wait."""
import numpy as np
import pytest

from gorender_amd import _lib
from gorender_amd import scenes as S
from oracle import refcpu

import scene_cases as C
from parity import assert_same

pytestmark = pytest.mark.gpu


def _check(scene, got, want):
    rgba, depth, st = got
    if depth is None:
        assert np.array_equal(rgba, want[0]), scene.name
        assert st["tpf"] == want[2]["tpf"], scene.name
    else:
        assert_same(scene, got, want)


def test_present_two_frames_in_flight(ctx):
    # 77x53: the pixel copies take the copy kernel's unaligned path.
    scenes = [S.config1("ref"), C.sphere(), S.config1("cfg"), C.cube_scene(),
              C.sphere(w=200, h=120, eye=(0, 0, 1.3)), C.sphere(w=77, h=53)]
    handles = [ctx.upload_scene(s)[0] for s in scenes]
    wants = [refcpu.draw(s) for s in scenes]
    with pytest.raises(_lib.GorenderError):
        ctx.present_wait(1, 1)  # nothing queued
    ctx.present_draw(scenes[0], handles[0])
    for k in range(1, len(scenes) + 1):
        if k < len(scenes):
            ctx.present_draw(scenes[k], handles[k], with_depth=(k % 2 == 0))
            if k == 1:
                with pytest.raises(_lib.GorenderError) as e:
                    ctx.present_draw(scenes[0], handles[0])  # a third frame in flight
                assert e.value.status == _lib.GR_ERR_INVALID_ARGUMENT
        fr = scenes[k - 1].frame
        _check(scenes[k - 1], ctx.present_wait(fr.width, fr.height), wants[k - 1])


def test_present_overflow_redraws(ctx):
    scene = C.sphere(w=320, h=240, eye=(0, 0, 1.3), nlat=24, nlon=40)
    want = refcpu.draw(scene)
    mh, _ = ctx.upload_scene(scene)
    old = ctx.config()
    try:
        ctx.set_config(tile_w=8, tile_h=8, refs_capacity=1024)
        for _ in range(2):
            ctx.present_draw(scene, mh)
            _check(scene, ctx.present_wait(320, 240), want)
    finally:
        ctx.set_config(tile_w=old.tile_w, tile_h=old.tile_h, refs_capacity=old.refs_capacity)
