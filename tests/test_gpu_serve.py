"""Serving loop (gr_server_*): every request's frame and counters equal the
oracle's for that request drawn alone (committed digests), whatever batch it
shares, under tight target budgets, and after a tile-list overflow makes the
server grow its capacity and draw a batch again."""
import ctypes
import json
import os

import numpy as np
import pytest

from gorender_amd import _lib
from gorender_amd import scenes as S
from gorender_amd.serve import PreparedRequest, Server, poisson_arrivals, replay
from parity import digest
from gorender_amd import scenes as S
from scene_cases import config5_all

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_digests.json")))


@pytest.fixture(scope="module")
def resident_all(ctx):
    """The 32 resident meshes and all 64 requests the serving benchmark replays."""
    rs = config5_all()
    th = [ctx.upload_texture(t) for t in rs.textures]
    mh = [ctx.upload_mesh(m, th) for m in rs.meshes]
    return rs, [PreparedRequest(r, mh) for r in rs.requests]


@pytest.fixture(scope="module")
def resident(resident_all):
    """The first 8 requests over the same resident meshes."""
    rs, reqs = resident_all
    return S.ResidentSet(rs.meshes, rs.textures, rs.requests[:8]), reqs[:8]


def check_result(k, res):
    assert res.rgba is not None
    got = digest(res.rgba, res.depth, res.stats)
    assert got == GOLDEN[f"config5-{k}"], f"config5-{k}: served frame differs from the oracle"
    assert res.done_ns >= res.start_ns >= res.submit_ns


@pytest.mark.parametrize("budget_mib,max_batch,max_inflight", [(1024, 8, 2), (40, 3, 2),
                                                                (17, 1, 1)])
def test_mix_equals_alone(ctx, resident, budget_mib, max_batch, max_inflight):
    """Each of the 8 requests is submitted twice in a shuffled order; all are
    in flight together and batched as the budget allows (17 MiB holds a single
    1920x1080 request)."""
    rs, reqs = resident
    order = list(range(len(reqs))) * 2
    np.random.default_rng(budget_mib).shuffle(order)
    print(f"budget {budget_mib} MiB, max_batch {max_batch}, max_inflight {max_inflight}",
          flush=True)
    with Server(ctx, budget_mib << 20, max_batch, max_inflight) as srv:
        tickets = [srv.submit(reqs[k]) for k in order]
        results = [srv.wait(t) for t in tickets]
    sizes = [r.batch_size for r in results]
    assert max(sizes) <= max_batch
    if max_batch > 1:
        assert max(sizes) > 1, "requests in flight together were never batched"
    for k, res in zip(order, results):
        check_result(k, res)


def test_all_benchmark_requests(ctx, resident_all):
    """The 64 distinct requests of the serving benchmark, served at its
    settings (2 GiB budget, batches of up to 16, 3 launches in flight), each
    equal to the oracle's digest of that request drawn alone."""
    rs, reqs = resident_all
    assert len(reqs) == 64
    with Server(ctx, 2 << 30, 16, 3) as srv:
        tickets = [srv.submit(r) for r in reqs]
        results = [srv.wait(t) for t in tickets]
    assert max(r.batch_size for r in results) > 1
    for k, res in enumerate(results):
        check_result(k, res)


def test_overflow_regrows_and_redraws(ctx, resident):
    """A tile-list capacity far below what the requests need: the server grows
    it and draws the overflowing batches again."""
    rs, reqs = resident
    ctx.set_config(refs_capacity=1 << 16)
    with Server(ctx, 1 << 30, 8, 2) as srv:
        tickets = [srv.submit(r) for r in reqs]
        results = [srv.wait(t) for t in tickets]
    assert ctx.config().refs_capacity > 1 << 16
    for k, res in enumerate(results):
        check_result(k, res)


def test_rejects_bad_requests(ctx, resident):
    rs, reqs = resident
    budget = 4 << 20  # holds 640x360 (1.8 MB), not 1280x720 (7.4 MB)
    with Server(ctx, budget, 4, 2) as srv:
        big = next(r for r in reqs if r.scene.frame.width * r.scene.frame.height * 8 > budget)
        with pytest.raises(_lib.GorenderError) as e:
            srv.submit(big)
        assert e.value.status == _lib.GR_ERR_INVALID_ARGUMENT
        small = next(s for s in rs.requests if s.frame.width * s.frame.height * 8 <= budget)
        bad = PreparedRequest(small, [10_000] * len(rs.meshes))
        with pytest.raises(_lib.GorenderError) as e:
            srv.submit(bad)
        assert e.value.status == _lib.GR_ERR_INVALID_HANDLE
        with pytest.raises(_lib.GorenderError) as e:
            srv.wait_raw(12345)
        assert e.value.status == _lib.GR_ERR_INVALID_HANDLE


def test_open_loop_replay(ctx, resident):
    """Native open-loop driver: every request completes, latencies are
    positive and some requests share a launch at 2000 requests/s."""
    rs, reqs = resident
    with Server(ctx, 1 << 30, 16, 2) as srv:
        lat, bs = replay(srv, reqs, poisson_arrivals(96, 2000.0, seed=7))
    assert len(lat) == 96 and (lat > 0).all()
    assert bs.min() >= 1 and bs.max() > 1
    print(f"replay p50 {np.percentile(lat, 50):.3f} ms p99 {np.percentile(lat, 99):.3f} ms "
          f"mean batch {bs.mean():.2f}", flush=True)


def test_bad_requests_refused_alone(ctx, resident):
    """Requests that a batch would refuse (a reference-tile count other than 1
    or 16; Gouraud lighting over normal indices out of range) are refused at
    submit, and a request whose overlay lines are too long fails alone inside
    its batch: the server keeps serving the others, each frame equal to the
    oracle's."""
    rs, reqs = resident
    v = np.array([[0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1]], np.float32)
    m = S.MeshData(v, np.array([[0, 1, 2]], np.int32), v.copy(),
                   np.array([[0, 1, -1]], np.int32))
    h = ctx.upload_mesh(m)  # before the server owns the context
    fr = S.FrameData(32, 32)
    lit = PreparedRequest(
        S.SceneData(fr, [m], [], [S.camera_object(fr, 0, (0, 0, 3), (0, 0, -1))]), [h])
    # Overlay lines between far-off unclipped vertices: this frame fails alone
    # (GR_ERR_UNSUPPORTED) inside a batch whose other frames complete.
    import scene_cases as C
    wide_scene = C.clip_space_soup(flags=C.no(_lib.GR_FRUSTUM_CLIPPING) | _lib.GR_SHOW_EDGES,
                                   seed=7, w_range=(-1e-4, 1e-4))
    wide = PreparedRequest(wide_scene, ctx.upload_scene(wide_scene)[0])
    ok = reqs[0]
    tiles = ok.frame.ref_tiles
    with Server(ctx, 1 << 30, 8, 2) as srv:
        ok.frame.ref_tiles = 3
        try:
            with pytest.raises(_lib.GorenderError) as e:
                srv.submit(ok)
        finally:
            ok.frame.ref_tiles = tiles
        assert e.value.status == _lib.GR_ERR_INVALID_ARGUMENT
        with pytest.raises(_lib.GorenderError) as e:
            srv.submit(lit)
        assert e.value.status == _lib.GR_ERR_INDEX_OUT_OF_RANGE
        tw = srv.submit(wide)
        tickets = [srv.submit(r) for r in reqs]
        with pytest.raises(_lib.GorenderError) as e:
            srv.wait_raw(tw)
        assert e.value.status == _lib.GR_ERR_UNSUPPORTED
        assert "65536" in str(e.value)
        srv.release(tw)
        results = [srv.wait(t) for t in tickets]
    assert max(r.batch_size for r in results) > 1
    for k, res in enumerate(results):
        check_result(k, res)


def test_colour_only(ctx, resident):
    """colour_only: the frames come back without depth (result depth NULL), the
    colour and counters of each still equal to the oracle's."""
    import hashlib
    rs, reqs = resident
    with Server(ctx, 1 << 30, 8, 2, colour_only=True) as srv:
        tickets = [srv.submit(r) for r in reqs]
        results = [srv.wait(t) for t in tickets]
    assert max(r.batch_size for r in results) > 1
    for k, res in enumerate(results):
        assert res.depth is None
        want = GOLDEN[f"config5-{k}"]
        assert hashlib.sha256(np.ascontiguousarray(res.rgba).tobytes()).hexdigest() == want["rgba"]
        assert res.stats["tpf"] == want["tpf"] and res.stats["triangles"] == want["triangles"]


def test_server_config_checked(ctx):
    h = ctypes.c_void_p()
    for bad in (_lib.gr_server_config(1 << 20, 4, 2, 2, 0, 0, 0),
                _lib.gr_server_config(1 << 20, 4, 2, 0, -1, 0, 0),
                _lib.gr_server_config(1 << 20, 4, 2, 0, 0, -5, 0)):
        assert ctx.lib.gr_server_create(ctx.handle, ctypes.byref(bad), ctypes.byref(h)) == \
            _lib.GR_ERR_INVALID_ARGUMENT


def test_bounded_queue_refuses_and_admitted_are_exact(ctx, resident):
    """max_queue: a burst beyond the queue bound is refused at submission
    (GR_ERR_BUSY); every admitted request is still exact."""
    rs, reqs = resident
    tickets, refused = [], 0
    with Server(ctx, 1 << 30, 8, 2, max_queue=8) as srv:
        for i in range(256):
            try:
                tickets.append((i % 8, srv.submit(reqs[i % 8])))
            except _lib.GorenderError as e:
                assert e.status == _lib.GR_ERR_BUSY
                refused += 1
        results = [(k, srv.wait(t)) for k, t in tickets]
    print(f"{len(results)} admitted, {refused} refused", flush=True)
    assert refused > 0 and results
    for k, res in results:
        check_result(k, res)


def test_deadline_sheds_late_requests(ctx, resident):
    """deadline_ms: requests that waited longer for admission are shed (their
    wait returns GR_ERR_BUSY); the others are exact."""
    rs, reqs = resident
    shed, done = 0, 0
    with Server(ctx, 1 << 30, 2, 1, deadline_ms=1) as srv:
        tickets = [(i % 8, srv.submit(reqs[i % 8])) for i in range(64)]
        for k, t in tickets:
            try:
                check_result(k, srv.wait(t))
                done += 1
            except _lib.GorenderError as e:
                assert e.status == _lib.GR_ERR_BUSY, e
                srv.release(t)
                shed += 1
    print(f"{done} served, {shed} shed", flush=True)
    assert shed > 0 and done > 0


@pytest.mark.parametrize("colour_only", [False, True])
def test_odd_frame_sizes(ctx, colour_only):
    """Frames whose pixel bytes are not a multiple of 16 (the readback copy
    kernel's unaligned path), batched together, equal the oracle bit for bit;
    colour_only returns no depth."""
    from oracle import refcpu
    from scene_cases import cube_scene, sphere
    scenes = [sphere(w=77, h=53), cube_scene(w=45, h=31), sphere(w=200, h=120, eye=(0, 0, 1.4))]
    reqs = [PreparedRequest(sc, ctx.upload_scene(sc)[0]) for sc in scenes]
    with Server(ctx, 64 << 20, 8, 3, colour_only=colour_only) as srv:
        tickets = [srv.submit(r) for r in reqs]
        results = [srv.wait(t) for t in tickets]
    for sc, res in zip(scenes, results):
        want = refcpu.draw(sc)
        h, w = sc.frame.height, sc.frame.width
        assert np.array_equal(res.rgba.reshape(h, w, 4), want[0]), sc.name
        if colour_only:
            assert res.depth is None
        else:
            assert np.array_equal(res.depth.reshape(h, w).view(np.uint32),
                                  want[1].view(np.uint32)), sc.name
        assert res.stats["tpf"] == want[2]["tpf"], sc.name

