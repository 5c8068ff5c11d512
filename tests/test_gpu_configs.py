"""GPU parity on the benchmark configurations, bit for bit against the oracle's
committed digests (tests/golden) and, on a mismatch, against the live oracle
for a pixel-level report.

config 3: 64 instances (~780k triangles, 1920x1080, mixed flat/Gouraud and
texture sizes); config 4: 16 frames of 1k-500k triangles at four resolutions
in one batched call; config 5: 32 resident meshes (~8 M triangles) drawn by 8
requests of 1-8 objects, singly and as one batch.
"""
import json
import os

import numpy as np
import pytest

from gorender_amd import _lib
from gorender_amd import scenes as S
from oracle import refcpu
from parity import assert_same, digest
from scene_cases import config4_all, config4_batch, config5_all, config5_set

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_digests.json")))


def check(name, scene, got, oracle_scene=None):
    rgba, depth, st = got
    if digest(rgba, depth, st) != GOLDEN[name]:
        assert_same(scene, got, refcpu.draw(oracle_scene or scene))
        raise AssertionError(f"{name}: digest differs from the committed oracle digest")


def draw_batch(ctx, scenes, handles):
    outs = []
    for s in scenes:
        rg, dp = ctx.frame_buffers(s.frame.width, s.frame.height)
        rg.fill(0xA5)
        dp.fill(0x5A)
        outs.append((rg, dp))
    ptrs = [(o[0].ptr, o[1].ptr) for o in outs]
    ctx.draw_batch(scenes, handles, ptrs)
    try:
        ctx.synchronize()
    except _lib.GorenderError as e:
        # The asynchronous form reports a tile-list overflow after the fact and
        # has grown its capacity: the caller draws the batch again.
        assert e.status == _lib.GR_ERR_CAPACITY
        ctx.draw_batch(scenes, handles, ptrs)
        ctx.synchronize()
    stats = ctx.last_stats()
    res = []
    for k, s in enumerate(scenes):
        h, w = s.frame.height, s.frame.width
        res.append((outs[k][0].to_numpy(np.uint8, (h, w, 4)),
                    outs[k][1].to_numpy(np.float32, (h, w)), stats[k]))
        outs[k][0].free()
        outs[k][1].free()
    return res


def test_config3(ctx):
    scene = S.config3()
    check("config3", scene, ctx.draw(scene))


def test_config4_batched(ctx):
    frames = config4_batch()
    handles = [ctx.upload_scene(s)[0] for s in frames]
    for k, got in enumerate(draw_batch(ctx, frames, handles)):
        check(f"config4-{k}", frames[k], got)


def test_config5_resident(ctx):
    rs = config5_set()
    th = [ctx.upload_texture(t) for t in rs.textures]
    mh = [ctx.upload_mesh(m, th) for m in rs.meshes]
    for k, req in enumerate(rs.requests):
        check(f"config5-{k}", req, ctx.draw(req, mh), S.subscene(req))
    for k, got in enumerate(draw_batch(ctx, rs.requests, [mh] * len(rs.requests))):
        check(f"config5-{k}", rs.requests[k], got, S.subscene(rs.requests[k]))


def test_config4_256_frames_one_batch(ctx):
    """Config 4 at its stated size: all 256 frames (1k-500k triangles, four
    resolutions, half textured) in ONE gr_draw_batch, each frame against the
    oracle's committed digest."""
    frames = config4_all()
    assert len(frames) == 256
    handles = [ctx.upload_scene(s)[0] for s in frames]
    print(f"config4: {sum(s.num_faces for s in frames)} triangles in one batch", flush=True)
    for k, got in enumerate(draw_batch(ctx, frames, handles)):
        check(f"config4-{k}", frames[k], got)


def test_config5_all_requests(ctx):
    """All 64 requests of config 5 over the 32 resident meshes, as four batches
    of 16, against the oracle's digests of each request drawn alone."""
    rs = config5_all()
    assert len(rs.requests) == 64
    th = [ctx.upload_texture(t) for t in rs.textures]
    mh = [ctx.upload_mesh(m, th) for m in rs.meshes]
    for b in range(4):
        reqs = rs.requests[16 * b:16 * (b + 1)]
        for k, got in enumerate(draw_batch(ctx, reqs, [mh] * len(reqs))):
            check(f"config5-{16 * b + k}", reqs[k], got, S.subscene(reqs[k]))


@pytest.mark.parametrize("tile,split", [((32, 32), 64), ((16, 16), 64), ((64, 64), 64)])
def test_config4_split_tiles(ctx, tile, split):
    """Tile lists longer than tile_split entries are resolved by several raster
    workgroups and merged: the frames stay bit-exact. At 64x64 tiles the slab
    budget runs out and the remaining long lists stay whole."""
    before = ctx.config()
    ctx.set_config(tile_w=tile[0], tile_h=tile[1], tile_split=split)
    try:
        frames = config4_batch()
        handles = [ctx.upload_scene(s)[0] for s in frames]
        res = draw_batch(ctx, frames, handles)
        split_tiles = sum(r[2]["tiles_split"] for r in res)
        print(f"tile {tile} split {split}: {split_tiles} tiles split", flush=True)
        assert split_tiles > 0
        for k, got in enumerate(res):
            check(f"config4-{k}", frames[k], got)
    finally:
        ctx.set_config(tile_w=before.tile_w, tile_h=before.tile_h,
                       tile_split=before.tile_split)


def test_config5_split_tiles(ctx):
    """The serving batch with every list longer than 64 entries split."""
    before = ctx.config()
    ctx.set_config(tile_split=64)
    try:
        rs = config5_set()
        th = [ctx.upload_texture(t) for t in rs.textures]
        mh = [ctx.upload_mesh(m, th) for m in rs.meshes]
        res = draw_batch(ctx, rs.requests, [mh] * len(rs.requests))
        assert sum(r[2]["tiles_split"] for r in res) > 0
        for k, got in enumerate(res):
            check(f"config5-{k}", rs.requests[k], got, S.subscene(rs.requests[k]))
    finally:
        ctx.set_config(tile_split=before.tile_split)


def test_persistent_raster_grid(ctx):
    """gr_config.raster_grid = 2 (the raster stage as a persistent grid striding
    over the tiles): the 256 config-4 frames in one batch and the config-5
    serving batch with split tiles give the committed oracle digests."""
    before = ctx.config()
    ctx.set_config(raster_grid=2)
    try:
        assert ctx.config().raster_grid == 2
        frames = config4_all()
        handles = [ctx.upload_scene(s)[0] for s in frames]
        for k, got in enumerate(draw_batch(ctx, frames, handles)):
            check(f"config4-{k}", frames[k], got)
        ctx.set_config(tile_split=64)
        rs = config5_set()
        th = [ctx.upload_texture(t) for t in rs.textures]
        mh = [ctx.upload_mesh(m, th) for m in rs.meshes]
        res = draw_batch(ctx, rs.requests, [mh] * len(rs.requests))
        assert sum(r[2]["tiles_split"] for r in res) > 0
        for k, got in enumerate(res):
            check(f"config5-{k}", rs.requests[k], got, S.subscene(rs.requests[k]))
    finally:
        ctx.set_config(tile_split=before.tile_split, raster_grid=before.raster_grid)


def test_config_validation(ctx):
    """gr_set_config refuses out-of-range fields and keeps the configuration."""
    before = ctx.config()
    for kw in ({"raster_grid": 3}, {"tile_w": 24}, {"tile_split": 32}, {"pipelines": 3}):
        with pytest.raises(_lib.GorenderError) as e:
            ctx.set_config(**kw)
        assert e.value.status == _lib.GR_ERR_INVALID_ARGUMENT, kw
        now = ctx.config()
        assert (now.tile_w, now.tile_h, now.tile_split, now.pipelines, now.raster_grid) == (
            before.tile_w, before.tile_h, before.tile_split, before.pipelines,
            before.raster_grid), kw
    assert before.raster_grid == 1
