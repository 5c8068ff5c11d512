"""GPU parity: the HIP render path against the CPU oracle, through the C ABI.

Bar: colour and depth buffers bit-identical to the oracle (which restates the
reference's sequential draw, renderer.go:466-474), and identical counters
(TPF renderer.go:436-441, binned triangles, covered fragments, front faces).
"""
import numpy as np
import pytest

from gorender_amd import _lib
from gorender_amd import scenes as S
from oracle import refcpu

from parity import assert_same
from scene_cases import all_cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,builder", all_cases(), ids=[c[0] for c in all_cases()])
def test_case(ctx, name, builder):
    scene = builder()
    want = refcpu.draw(scene)
    got = ctx.draw(scene)
    assert_same(scene, got, want)


@pytest.mark.parametrize("tw,th", [(8, 8), (16, 16), (32, 16), (64, 64), (8, 64), (64, 8)])
def test_tile_shapes(ctx, tw, th):
    """The image does not depend on the raster tile shape."""
    import scene_cases as C
    scene = C.sphere(w=150, h=110, eye=(0, 0, 1.25), nlat=30, nlon=50)
    want = refcpu.draw(scene)
    old = ctx.config()
    try:
        ctx.set_config(tile_w=tw, tile_h=th)
        got = ctx.draw(scene)
    finally:
        ctx.set_config(tile_w=old.tile_w, tile_h=old.tile_h)
    assert_same(scene, got, want)


def test_config1_both_variants(ctx):
    for variant in ("ref", "cfg"):
        scene = S.config1(variant)
        assert_same(scene, ctx.draw(scene), refcpu.draw(scene))


def test_config2_full(ctx):
    """Headline frame: 200,000 triangles, 1280x720, textured, side-plane clipping."""
    scene = S.config2()
    assert_same(scene, ctx.draw(scene), refcpu.draw(scene))


def test_refs_capacity_retry(ctx):
    """A tiny tile-list capacity is grown and the frame redrawn (same output)."""
    import scene_cases as C
    scene = C.sphere(w=200, h=120, eye=(0, 0, 1.3), nlat=20, nlon=30)
    want = refcpu.draw(scene)
    old = ctx.config()
    try:
        ctx.set_config(refs_capacity=1024)
        got = ctx.draw(scene)
    finally:
        ctx.set_config(refs_capacity=old.refs_capacity)
    assert_same(scene, got, want)


def test_repeat_is_deterministic(ctx):
    scene = S.config1()
    mh, _ = ctx.upload_scene(scene)
    a = ctx.draw(scene, mh)
    b = ctx.draw(scene, mh)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint32),
                                                         b[1].view(np.uint32))


def test_ref_tiles_one(ctx):
    """TPF on the sequential renderer's single tile equals the binned count."""
    scene = S.config1()
    scene.frame.ref_tiles = 1
    got = ctx.draw(scene)
    want = refcpu.draw(scene)
    assert_same(scene, got, want)
    assert got[2]["tpf"] == got[2]["triangles"]


def test_batch_equals_single(ctx):
    """gr_draw_batch of several uneven frames == each frame drawn alone."""
    import scene_cases as C
    frames = [C.sphere(seed=11, w=64, h=48), C.overlapping_objects(w=100, h=70),
              C.cube_scene(w=33, h=65), C.clip_space_soup(n=200, seed=9)]
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
    for k, s in enumerate(frames):
        want = refcpu.draw(s)
        h, w = s.frame.height, s.frame.width
        got = (outs[k][0].to_numpy(np.uint8, (h, w, 4)),
               outs[k][1].to_numpy(np.float32, (h, w)), stats[k])
        assert_same(s, got, want)


def test_errors(ctx):
    """Reference panics become status codes (renderer.go:318-330; mesh.go:29)."""
    bad = S.MeshData(np.zeros((3, 4), np.float32), np.array([[0, 1, 3]], np.int32))
    with pytest.raises(_lib.GorenderError) as e:
        ctx.upload_mesh(bad)
    assert e.value.status == _lib.GR_ERR_INDEX_OUT_OF_RANGE
    empty = S.MeshData(np.zeros((0, 4), np.float32), np.zeros((0, 3), np.int32))
    with pytest.raises(_lib.GorenderError):
        ctx.upload_mesh(empty)
    # Gouraud shading with a normal index out of range (obj.go:77-89 v//vn faces)
    v = np.array([[0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1]], np.float32)
    m = S.MeshData(v, np.array([[0, 1, 2]], np.int32), v.copy(),
                   np.array([[0, 1, -1]], np.int32))
    h = ctx.upload_mesh(m)
    fr = S.FrameData(32, 32)
    scene = S.SceneData(fr, [m], [], [S.camera_object(fr, 0, (0, 0, 3), (0, 0, -1))])
    with pytest.raises(_lib.GorenderError) as e:
        ctx.draw(scene, [h])
    assert e.value.status == _lib.GR_ERR_INDEX_OUT_OF_RANGE
    # ... but flat shading never reads the normal indices, as in the reference
    scene.frame.flags |= _lib.GR_FLAT_SHADING
    ctx.draw(scene, [h])
    # Overlay lines of unclipped geometry far off screen would walk billions of
    # pixels in the reference; the GPU path reports them as unsupported and the
    # context stays usable.
    import scene_cases as C
    wide = C.clip_space_soup(flags=C.no(_lib.GR_FRUSTUM_CLIPPING) | _lib.GR_SHOW_EDGES, seed=7,
                             w_range=(-1e-4, 1e-4))
    with pytest.raises(_lib.GorenderError) as e:
        ctx.draw(wide)
    assert e.value.status == _lib.GR_ERR_UNSUPPORTED
    ctx.draw(scene, [h])
    with pytest.raises(_lib.GorenderError) as e:
        ctx.draw(scene, [h + 1000])
    assert e.value.status == _lib.GR_ERR_INVALID_HANDLE


@pytest.mark.parametrize("flags,ref_tiles,size", [
    (_lib.GR_SHOW_CROSSHAIR, 16, (160, 96)),
    (_lib.GR_SHOW_TILE_BOUNDS, 16, (77, 53)),
    (_lib.GR_SHOW_TILE_BOUNDS | _lib.GR_SHOW_CROSSHAIR, 1, (64, 48)),
    (_lib.GR_SHOW_TILE_BOUNDS | _lib.GR_SHOW_CROSSHAIR | _lib.GR_SHOW_EDGES, 16, (120, 90)),
    (_lib.GR_SHOW_TILE_BOUNDS | _lib.GR_SHOW_CROSSHAIR, 16, (5, 7)),
])
def test_decorations(ctx, flags, ref_tiles, size):
    """Tile borders and crosshair (renderer.go:422-434,476-480;
    rasterizer.go:209-217) over the drawn frame, as the oracle draws them."""
    import scene_cases as C
    scene = C.sphere(w=size[0], h=size[1], eye=(0, 0, 2.4), flags=C.ALL | flags)
    scene.frame.ref_tiles = ref_tiles
    assert_same(scene, ctx.draw(scene), refcpu.draw(scene))


def test_decorations_batched(ctx):
    """Decorated and plain frames in one batch into shared buffers: each frame
    keeps only its own decorations, and a following draw into the same
    buffers is not painted over."""
    import scene_cases as C
    frames = [C.sphere(w=96, h=64, seed=s, flags=C.ALL | f) for s, f in
              ((21, _lib.GR_SHOW_CROSSHAIR), (22, 0),
               (23, _lib.GR_SHOW_TILE_BOUNDS | _lib.GR_SHOW_CROSSHAIR))]
    handles = [ctx.upload_scene(s)[0] for s in frames]
    outs = [ctx.frame_buffers(96, 64) for _ in frames]
    ctx.draw_batch(frames, handles, [(o[0].ptr, o[1].ptr) for o in outs])
    # The undecorated frame again into the first frame's buffers, right after.
    ctx.draw_batch(frames[1:2], handles[1:2], [(outs[0][0].ptr, outs[0][1].ptr)])
    ctx.synchronize()
    for k, src in ((0, 1), (1, 1), (2, 2)):
        want = refcpu.draw(frames[src])
        got = outs[k][0].to_numpy(np.uint8, (64, 96, 4))
        assert np.array_equal(got, want[0]), f"buffers {k} differ from frame {src}"
    for rg, dp in outs:
        rg.free()
        dp.free()


def test_fog(ctx):
    """gr_fog (FrameBuffer.Fog, rasterizer.go:185-207) over a GPU frame equals
    the oracle's fog over the oracle's frame; the depths span keep, blend and
    fog colour (the background's -1 included)."""
    import scene_cases as C
    scene = C.sphere(w=120, h=90, eye=(0, 0, 2.4))
    want_rgba, want_depth, _ = refcpu.draw(scene)
    d = want_depth[want_depth > 0]
    start, end = float(np.quantile(d, 0.7)), float(np.quantile(d, 0.2))
    want = refcpu.fog(want_rgba, want_depth, start, end, (100, 120, 140, 255))
    h, w = 90, 120
    handles = ctx.upload_scene(scene)[0]
    rg, dp = ctx.frame_buffers(w, h)
    ctx.draw_device(scene, handles, rg.ptr, dp.ptr)
    ctx.fog(rg, dp, w, h, start, end, (100, 120, 140, 255))
    ctx.synchronize()
    got = rg.to_numpy(np.uint8, (h, w, 4))
    assert np.array_equal(got, want)
    assert (got != want_rgba).any(axis=-1).sum() > 1000
    rg.free()
    dp.free()


def test_subpixel_triangles(ctx):
    """A dense sphere far from the camera: most triangles fall between pixel
    centres, and those whose box is at most 2x2 pixels with no covered pixel
    stay out of the tile lists. The image and the counters still equal the
    oracle's, and the lists hold fewer entries than there are triangles."""
    import scene_cases as C
    scene = C.sphere(w=160, h=96, eye=(0, 0, 14.0), nlat=100, nlon=160)
    got = ctx.draw(scene)
    assert_same(scene, got, refcpu.draw(scene))
    st = got[2]
    assert st["refs"] < st["triangles"], st


def _straddling_fans(n=512, seed=9):
    """n triangles given in clip space (identity matrices), each crossing the
    x = -w, x = w and y = w planes: every face reaches the clipper and reserves
    the largest fan, 7 records, so each 256-face block needs 1,792 records and
    takes all but its first 320 from the overflow pools."""
    rng = S.SplitMix64(S.SEED_BASE + 700 + seed)
    # Inside the clip volume is w < 0, |x|, |y| <= -w and znear <= z <= zfar.
    base = np.array([[-2.0, -0.2, 0.5, -1.0], [0.0, 2.0, 0.5, -1.0], [2.0, -0.1, 0.5, -1.0]])
    jit = (rng.uniform(n * 12).reshape(n, 3, 4) - 0.5) * np.array([0.4, 0.4, 0.2, 0.0])
    v4 = (base[None] + jit).reshape(3 * n, 4).astype(np.float32)
    f = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    nrm = np.tile(np.array([[0.0, 0.0, 1.0, 1.0]], np.float32), (3 * n, 1))
    uvs = (rng.uniform(6 * n).reshape(n, 3, 2)).astype(np.float32)
    m = S.MeshData(v4, f, nrm, f.copy(), uvs)
    t = S.random_texture(rng, 8, 8)
    S.with_texture(m, 0)
    fr = S.FrameData(96, 72, flags=_lib.GR_DEFAULT_FLAGS & ~_lib.GR_BACKFACE_CULLING)
    eye = np.eye(4, dtype=np.float32)
    return S.SceneData(fr, [m], [t], [S.ObjectData(0, eye, eye)], name="straddling-fans")


def test_overflow_records(ctx):
    """Faces that all reach the clipper: most of their records come from the
    overflow pools. Equal to the oracle at the default capacity, and with a
    capacity so small that the pools run out (the batch reports it, the
    capacity grows and the frame is drawn again)."""
    scene = _straddling_fans()
    want = refcpu.draw(scene)
    assert want[2]["triangles"] > 0
    assert_same(scene, ctx.draw(scene), want)
    old = ctx.config()
    try:
        ctx.set_config(refs_capacity=1024)
        got = ctx.draw(scene)
        assert ctx.config().refs_capacity > 1024
    finally:
        ctx.set_config(refs_capacity=old.refs_capacity)
    assert_same(scene, got, want)
