"""The batched form captured in HIP graphs (gr_batch_graph_*) and the autotuner
(gr_autotune), through the C ABI.

Bar: every replay of a graph gives the colour, depth and counters of the eager
gr_draw_batch of the same frames, bit for bit (and the committed oracle digests
of config 4); graphs are replayed several times with eager work and
synchronisations in between; the autotuner's candidates all produce one image.
"""
import json
import os

import numpy as np
import pytest

from gorender_amd import _lib
from gorender_amd import scenes as S
from oracle import refcpu
from parity import assert_same, digest
from scene_cases import config4_batch, cube_scene, sphere

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_digests.json")))


def _outputs(ctx, scenes):
    outs = [ctx.frame_buffers(s.frame.width, s.frame.height) for s in scenes]
    return outs, [(o[0].ptr, o[1].ptr) for o in outs]


def _junk(outs):
    for rg, dp in outs:
        rg.fill(0xA5)
        dp.fill(0x5A)


def _read(scenes, outs, stats):
    res = []
    for k, s in enumerate(scenes):
        h, w = s.frame.height, s.frame.width
        res.append((outs[k][0].to_numpy(np.uint8, (h, w, 4)),
                    outs[k][1].to_numpy(np.float32, (h, w)), stats[k]))
    return res


def _free(outs):
    for rg, dp in outs:
        rg.free()
        dp.free()


def _same(a, b, name):
    assert np.array_equal(a[0], b[0]), name
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)), name
    assert a[2] == b[2], name


def test_graph_replays_equal_eager_and_digests(ctx):
    frames = list(config4_batch())
    handles = [ctx.upload_scene(s)[0] for s in frames]
    outs, ptrs = _outputs(ctx, frames)
    try:
        prepared = ctx.prepare_batch(frames, handles, ptrs)
        _junk(outs)
        ctx.submit_batch(prepared)
        try:
            ctx.synchronize()
        except _lib.GorenderError as e:
            assert e.status == _lib.GR_ERR_CAPACITY
            ctx.submit_batch(prepared)
            ctx.synchronize()
        eager = _read(frames, outs, ctx.last_stats())
        for k, got in enumerate(eager):
            assert digest(*got) == GOLDEN[f"config4-{k}"], k
        g = ctx.batch_graph(prepared)
        try:
            for rep in range(3):
                _junk(outs)
                g.launch()
                ctx.synchronize()
                res = _read(frames, outs, ctx.last_stats())
                for k in range(len(frames)):
                    _same(res[k], eager[k], f"replay {rep} frame {k}")
                # eager work of the program between replays
                ctx.submit_batch(ctx.prepare_batch(frames[:2], handles[:2], ptrs[:2]))
                ctx.synchronize()
        finally:
            g.destroy()
    finally:
        _free(outs)


def test_two_graphs_alternate_on_two_pipes(ctx):
    """Two graphs replayed back to back without waiting: the second one's
    geometry overlaps the first one's raster; frames stay exact."""
    frames = list(config4_batch())[:8]
    handles = [ctx.upload_scene(s)[0] for s in frames]
    outs, ptrs = _outputs(ctx, frames)
    try:
        a = ctx.prepare_batch(frames[:4], handles[:4], ptrs[:4])
        b = ctx.prepare_batch(frames[4:], handles[4:], ptrs[4:])
        ga, gb = ctx.batch_graph(a), ctx.batch_graph(b)
        try:
            for rep in range(2):
                _junk(outs)
                for _ in range(3):
                    ga.launch()
                    gb.launch()
                ctx.synchronize()
                stats_b = ctx.last_stats()
                res = _read(frames[4:], outs[4:], stats_b)
                for k, got in enumerate(res):
                    assert digest(*got) == GOLDEN[f"config4-{k + 4}"], (rep, k)
                for k in range(4):
                    h, w = frames[k].frame.height, frames[k].frame.width
                    rg = outs[k][0].to_numpy(np.uint8, (h, w, 4))
                    dp = outs[k][1].to_numpy(np.float32, (h, w))
                    want = GOLDEN[f"config4-{k}"]
                    got = digest(rg, dp, {"tpf": want["tpf"], "triangles": want["triangles"],
                                          "fragments_covered": want["fragments_covered"],
                                          "faces_front": want["faces_front"],
                                          "pixels_shaded": want["pixels_shaded"]})
                    assert got == want, (rep, k)
        finally:
            ga.destroy()
            gb.destroy()
    finally:
        _free(outs)


def test_graph_update_and_regrow(ctx):
    """New matrices through gr_batch_graph_update; a replay whose tile lists
    overflow (the graph was sized for a far, small sphere) reports
    GR_ERR_CAPACITY and the next launch regrows them."""
    far = sphere(w=320, h=240, eye=(0, 0, 9.0), nlat=24, nlon=40)
    near = sphere(w=320, h=240, eye=(0, 0, 1.3), nlat=24, nlon=40)
    old = ctx.config()
    mh, th = ctx.upload_scene(far)
    outs, ptrs = _outputs(ctx, [far])
    try:
        ctx.set_config(tile_w=8, tile_h=8, refs_capacity=1024)
        g = ctx.batch_graph(ctx.prepare_batch([far], [mh], ptrs))
        try:
            _junk(outs)
            g.launch()
            ctx.synchronize()
            assert_same(far, _read([far], outs, ctx.last_stats())[0], refcpu.draw(far))
            g.update(ctx.objects_struct(near, mh), len(near.objects))
            _junk(outs)
            g.launch()
            with pytest.raises(_lib.GorenderError) as e:
                ctx.synchronize()
            assert e.value.status == _lib.GR_ERR_CAPACITY
            for _ in range(2):
                _junk(outs)
                g.launch()
                ctx.synchronize()
                assert_same(near, _read([near], outs, ctx.last_stats())[0], refcpu.draw(near))
        finally:
            g.destroy()
    finally:
        ctx.set_config(tile_w=old.tile_w, tile_h=old.tile_h, refs_capacity=old.refs_capacity)
        _free(outs)


def test_autotune_keeps_the_image(ctx):
    scenes = [sphere(w=200, h=120, eye=(0, 0, 1.4)), cube_scene(), S.config1("ref")]
    handles = [ctx.upload_scene(s)[0] for s in scenes]
    outs, ptrs = _outputs(ctx, scenes)
    old = ctx.config()
    try:
        prepared = ctx.prepare_batch(scenes, handles, ptrs)
        best, tried = ctx.autotune(prepared, reps=1)
        assert len(tried) == 64
        assert len({t["checksum"] for t in tried}) == 1
        assert best["tile"][0] in (8, 16, 32, 64) and best["tile_split"] in (256, 512, 1024, 4096)
        cfg = ctx.config()
        assert [cfg.tile_w, cfg.tile_h] == best["tile"]
        _junk(outs)
        ctx.submit_batch(prepared)
        ctx.synchronize()
        for s, got in zip(scenes, _read(scenes, outs, ctx.last_stats())):
            assert_same(s, got, refcpu.draw(s))
    finally:
        ctx.set_config(tile_w=old.tile_w, tile_h=old.tile_h, tile_split=old.tile_split,
                       refs_capacity=old.refs_capacity)
        _free(outs)
