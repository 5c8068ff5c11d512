"""Seeded random parity sweep: scene kind, camera, render toggles, clip
planes, frame size, raster tile shape and tile splitting are drawn from a
seeded generator; every frame drawn on the GPU, alone and inside a batch,
must equal the oracle's (colour, depth and counters bit for bit).

The fixed cases of test_gpu_parity pin each toggle and edge case on its own;
this sweep covers their combinations."""
import numpy as np
import pytest

from gorender_amd import _lib
from oracle import refcpu
from parity import assert_same
import scene_cases as C

pytestmark = pytest.mark.gpu

FLAG_BITS = [_lib.GR_SHOW_FACES, _lib.GR_BACKFACE_CULLING, _lib.GR_LIGHTING,
             _lib.GR_FLAT_SHADING, _lib.GR_FRUSTUM_CLIPPING, _lib.GR_SHOW_TEXTURES,
             _lib.GR_SHOW_EDGES, _lib.GR_SHOW_VERTICES, _lib.GR_SHOW_TILE_BOUNDS,
             _lib.GR_SHOW_CROSSHAIR]
FLAG_P = [0.85, 0.7, 0.7, 0.3, 0.85, 0.8, 0.15, 0.15, 0.15, 0.15]
TILES = [(8, 8), (16, 16), (32, 16), (16, 32), (32, 32), (64, 16), (64, 64), (8, 64)]
N_SINGLE = 32
N_BATCH = 8


def random_flags(rng):
    return sum(b for b, p in zip(FLAG_BITS, FLAG_P) if rng.random() < p)


def random_scene(rng, k):
    kind = int(rng.integers(0, 4))
    flags = random_flags(rng)
    w, h = int(rng.integers(9, 200)), int(rng.integers(7, 140))
    if kind == 0:
        # A displaced sphere seen from a random direction (kept off the up
        # axis, where the view matrix degenerates) and distance.
        d = rng.normal(size=3)
        d[1] *= 0.5
        d /= np.linalg.norm(d)
        eye = tuple(float(v) for v in d * rng.uniform(1.05, 4.0))
        tw = int(2 ** rng.integers(0, 5)) if rng.random() < 0.5 else int(rng.integers(1, 9))
        return C.sphere(flags=flags, w=w, h=h, eye=eye, nlat=int(rng.integers(3, 30)),
                        nlon=int(rng.integers(3, 40)), seed=100 + k,
                        textured=bool(rng.random() < 0.8), tex_size=(tw, int(rng.integers(1, 9))),
                        znear=float(rng.choice([0.0, 0.05, 0.5])),
                        zfar=float(rng.choice([2.0, 3.5, 50.0])),
                        scale=float(rng.choice([1.0, 0.5, 3.0])),
                        drop_normals=bool(rng.random() < 0.2))
    if kind == 1:
        # Clip-space triangle soup: W crosses zero, coordinates from wide to
        # beyond the int32 fast path.
        wr = [(-3.0, 0.5), (-1e-4, 1e-4), (-2.0, -0.1)][int(rng.integers(0, 3))]
        if not flags & _lib.GR_FRUSTUM_CLIPPING:
            # Unclipped overlay lines between far-off vertices: the reference's
            # line walk (and so the oracle) steps through every point of them,
            # minutes of CPU per frame.
            flags &= ~(_lib.GR_SHOW_EDGES | _lib.GR_SHOW_VERTICES)
        return C.clip_space_soup(n=int(rng.integers(1, 400)), seed=1000 + k, flags=flags, w=w,
                                 h=h, w_range=wr, span=float(rng.choice([1.0, 3.0, 50.0])),
                                 exact=bool(rng.random() < 0.7))
    if kind == 2:
        return C.with_flags(C.overlapping_objects(w=w, h=h, seed=2000 + k), flags)
    return C.inside_sphere(flags=flags, znear=float(rng.choice([0.0, 0.1])), w=w, h=h,
                           seed=3000 + k)


def random_config(ctx, rng):
    tw, th = TILES[int(rng.integers(0, len(TILES)))]
    split = int(rng.choice([64, 512, 1 << 24]))
    ctx.set_config(tile_w=tw, tile_h=th, tile_split=split)
    return tw, th, split


@pytest.fixture
def restore_config(ctx):
    old = ctx.config()
    yield
    ctx.set_config(tile_w=old.tile_w, tile_h=old.tile_h, tile_split=old.tile_split)


@pytest.mark.parametrize("k", range(N_SINGLE))
def test_random_single(ctx, restore_config, k):
    rng = np.random.default_rng(7100 + k)
    scene = random_scene(rng, k)
    cfg = random_config(ctx, rng)
    print(f"case {k}: {scene.name} {scene.frame.width}x{scene.frame.height} "
          f"flags {scene.frame.flags:#x} tile/split {cfg}", flush=True)
    assert_same(scene, ctx.draw(scene), refcpu.draw(scene))


@pytest.mark.parametrize("k", range(N_BATCH))
def test_random_batch(ctx, restore_config, k):
    """Three to five random frames in one gr_draw_batch."""
    rng = np.random.default_rng(7300 + k)
    frames = [random_scene(rng, 100 + 8 * k + i) for i in range(int(rng.integers(3, 6)))]
    cfg = random_config(ctx, rng)
    print(f"batch {k}: {[s.name for s in frames]} tile/split {cfg}", flush=True)
    handles = [ctx.upload_scene(s)[0] for s in frames]
    outs = []
    for s in frames:
        rg, dp = ctx.frame_buffers(s.frame.width, s.frame.height)
        rg.fill(0xA5)  # every pixel must be written by the draw
        dp.fill(0x5A)
        outs.append((rg, dp))
    ctx.draw_batch(frames, handles, [(o[0].ptr, o[1].ptr) for o in outs])
    ctx.synchronize()
    stats = ctx.last_stats()
    for i, s in enumerate(frames):
        h, w = s.frame.height, s.frame.width
        got = (outs[i][0].to_numpy(np.uint8, (h, w, 4)),
               outs[i][1].to_numpy(np.float32, (h, w)), stats[i])
        assert_same(s, got, refcpu.draw(s))
    for rg, dp in outs:
        rg.free()
        dp.free()
