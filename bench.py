"""Benchmark of the render path on one MI355X.

Workload: the headline frame (displaced lat-long sphere of 200,000 triangles,
16x16 RGBA8 texture, 1280x720, camera at r=1.2). Frame k rotates the object by
0.01*k rad about Y, as the reference's demo loop does (main.go:229-233), so
the frames of a launch are distinct work. A step is one batched launch of
--batch frames (512 by default, k = 0..511): the full path (vertex transform,
clip, bin, raster, shade, colour + depth written to HBM) over 512 distinct
frames, meshes and textures resident in HBM. --warmup steps run untimed
first, then exactly --steps steps are timed between two device
synchronisations.

Forms timed: "graph" (each batch's launches captured in HIP graphs and
replayed, two batches alternating on the context's two pipelines), "stream"
(the same batches submitted with gr_draw_batch) and "plain" (one
gr_draw_device call per frame, with the library's two pipelines and strictly
one after another). The value is the faster batched form. Beside it: the
graph form with the raster stage as a persistent grid, config 4
(256 mixed frames: autotuned configuration, one graph launch for all 256
against 256 plain launches, outputs compared on the device), configs 3 and 5,
the serving loop at fixed request rates, per-kernel times, the HBM roofline of
the dominant kernel and a CPU baseline.

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64,
                    help="timed steps; a step is one batched launch of --batch frames")
    ap.add_argument("--warmup", type=int, default=16, help="untimed steps before timing")
    ap.add_argument("--batch", type=int, default=512, help="frames per batched launch")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="bound on the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: no CPU baseline, no latency loops, no "
                         "other configs")
    ap.add_argument("--mode", choices=["all", "stream", "graph", "plain", "config5"],
                    default="all",
                    help="forms to run (rocprofv3 runs profile one form at a time); config5 "
                         "only launches the 8-request serving batch --steps times")
    ap.add_argument("--tile", default="32x32", help="raster tile WxH")
    ap.add_argument("--tile-split", type=int, default=0,
                    help="tile-list entries per raster workgroup before a tile is split "
                         "(0 keeps the library default)")
    ap.add_argument("--pipelines", type=int, default=2, choices=[1, 2],
                    help="context pipelines: 2 overlaps a draw's geometry kernels with the "
                         "previous draw's raster (the library default)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip configs 3-5 and the serving loop")
    ap.add_argument("--calib-copy-mib", type=int, default=1024,
                    help="MiB copied by k_copy16 to measure the device's copy bandwidth (also "
                         "the known-size kernel that HBM byte counters are calibrated against)")
    return ap.parse_args()


def profile_file(name):
    return os.path.join(ROOT, "profiles", name)


def measured_traffic(kernel, cfg, frames_per_launch):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by the ratio a known-size
    copy kernel reads under the same counters), when it was taken on this
    kernel, tile and batch size; else None."""
    path = profile_file("hbm_traffic_config2.txt")
    if not os.path.exists(path):
        return None
    want = f"{kernel} tile={cfg.tile_w}x{cfg.tile_h} frames_per_launch={frames_per_launch}"
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch ") and want in line:
            return int(line.rsplit("=", 1)[1])
    return None


# Wave64 VALU issue of the whole chip: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64
# VALU instruction per SIMD every 4 cycles.
VALU_PEAK_GINSTR_S = 256 * 4 * 2.4 / 4


def sq_counters(kernel, cfg, frames_per_launch):
    """SQ counters per launch of `kernel` from the committed rocprofv3 summary
    of the headline batch, when it was taken with this tile and batch size
    (its first line names the frames per launch)."""
    path = profile_file("sq_counters_config2.txt")
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        if f", {frames_per_launch} frames per launch" not in fh.readline():
            return {}
    name = f"{kernel}<{cfg.tile_w}, {cfg.tile_h}>" if kernel.startswith("k_raster") else kernel
    tile_ok, out = False, {}
    for line in open(path):
        if line.startswith("# raster tile "):
            tile_ok = line.split()[-1] == f"{cfg.tile_w}x{cfg.tile_h}"
        elif tile_ok and line.startswith(name + " "):
            for kv in line.split()[len(name.split()):]:
                if "=" in kv:
                    k, v = kv.split("=", 1)
                    try:
                        out[k] = float(v)
                    except ValueError:  # the [vgpr=.. scratch=..] annotation
                        pass
    return out


def capacity_safe(ctx, fn):
    """Runs fn() then waits; a tile-list overflow (the context grows its
    capacity) runs it again. Returns the passes run."""
    from gorender_amd import _lib
    for n in range(1, 9):
        fn()
        try:
            ctx.synchronize()
            return n
        except _lib.GorenderError as e:
            if e.status != _lib.GR_ERR_CAPACITY:
                raise
    raise RuntimeError("tile-list capacity did not settle")


def warm_until_stable(ctx, fn, min_passes=4):
    """Untimed passes until two in a row needed no regrowth and at least
    min_passes ran (with two pipelines both pipes then hold buffers of their
    final size, so no allocation falls inside a timed region)."""
    from gorender_amd import _lib
    n = clean = 0
    while (clean < 2 or n < min_passes) and n < 32:
        fn()
        n += 1
        try:
            ctx.synchronize()
            clean += 1
        except _lib.GorenderError as e:
            if e.status != _lib.GR_ERR_CAPACITY:
                raise
            clean = 0
    return n


def timed_passes(ctx, fn, min_s=0.25, min_passes=20):
    """Passes of fn until at least min_s seconds and min_passes passes have
    run; returns (passes, seconds)."""
    ctx.synchronize()
    t0 = time.perf_counter()
    n = 0
    while True:
        fn()
        n += 1
        if n >= min_passes and n % 4 == 0:
            ctx.synchronize()
            dt = time.perf_counter() - t0
            if dt >= min_s:
                return n, dt


def frames_digest(ctx, outs):
    """gr_device_checksum over every frame's colour and depth buffers."""
    import numpy as np
    d = ctx.alloc(8)
    try:
        d.fill(0)
        for rg, dp in outs:
            rg.checksum_into(d)
            dp.checksum_into(d)
        ctx.synchronize()
        return int(d.to_numpy(np.uint64, (1,))[0])
    finally:
        d.free()


def config4_leg(ctx, scenes):
    """Config 4: 256 independent frames (1k-500k triangles, four resolutions,
    half textured). The autotuner picks tile size, tile_split and tile-list
    capacity on the batch; then one graph launch draws all 256 frames, against
    256 plain gr_draw_device launches (the library's two pipelines, and
    strictly one after another) under the same configuration. The outputs of
    the forms are compared by device checksums."""
    frames = scenes.config4_frames(256)
    handles = [ctx.upload_scene(s)[0] for s in frames]
    outs = [ctx.frame_buffers(s.frame.width, s.frame.height) for s in frames]
    ptrs = [(o[0].ptr, o[1].ptr) for o in outs]
    tris = sum(s.num_faces for s in frames)
    saved = ctx.config()
    res = {"frames": len(frames), "triangles_total": tris}
    try:
        prepared = ctx.prepare_batch(frames, handles, ptrs)
        t0 = time.perf_counter()
        best, tried = ctx.autotune(prepared, reps=2)
        res["autotune"] = {"choice": best, "candidates": len(tried),
                           "seconds": round(time.perf_counter() - t0, 2),
                           "fastest_ms": min(t["ms_per_launch"] for t in tried),
                           "slowest_ms": max(t["ms_per_launch"] for t in tried),
                           "default_32x32_512_ms": next(
                               (t["ms_per_launch"] for t in tried
                                if t["tile"] == [32, 32] and t["tile_split"] == 512), None),
                           "one_image": len({t["checksum"] for t in tried}) == 1}
        g = ctx.batch_graph(prepared)
        try:
            warm_until_stable(ctx, g.launch, min_passes=2)
            n, dt = timed_passes(ctx, g.launch, min_s=0.3, min_passes=8)
            fps_b = n * len(frames) / dt
            dig_b = frames_digest(ctx, outs)
        finally:
            g.destroy()
        singles = [ctx.prepare_batch([s], [h], [p]) for s, h, p in zip(frames, handles, ptrs)]

        def plain():
            for pb in singles:
                ctx.submit_batch(pb)

        for o in outs:
            o[0].fill(0)
            o[1].fill(0)
        warm_until_stable(ctx, plain, min_passes=2)
        n, dt = timed_passes(ctx, plain, min_s=0.3, min_passes=4)
        fps_p = n * len(frames) / dt
        dig_p = frames_digest(ctx, outs)
        ctx.set_config(pipelines=1)
        warm_until_stable(ctx, plain, min_passes=2)
        n, dt = timed_passes(ctx, plain, min_s=0.3, min_passes=4)
        fps_p1 = n * len(frames) / dt
        best_plain = max(fps_p, fps_p1)
        res.update({
            "frames_per_s": {"batched_graph": round(fps_b, 1), "plain": round(fps_p, 1),
                             "plain_one_pipeline": round(fps_p1, 1)},
            "mtris_s": {"batched_graph": round(tris / len(frames) * fps_b / 1e6, 1),
                        "plain": round(tris / len(frames) * fps_p / 1e6, 1)},
            "batched_over_best_plain": round(fps_b / best_plain, 3),
            "bit_exact_batched_vs_plain": dig_b == dig_p,
            "digest": f"{dig_b:016x}"})
    finally:
        ctx.set_config(tile_w=saved.tile_w, tile_h=saved.tile_h, tile_split=saved.tile_split,
                       pipelines=saved.pipelines)
        for o in outs:
            o[0].free()
            o[1].free()
        for h in handles:
            for m in h:
                ctx.destroy_mesh(m)
    return res


def other_configs(ctx, scenes):
    """Throughput of configs 3 and 5 on the same context: config 3 (64
    instances, ~1M triangles, 1920x1080) drawn one frame per call; config 5 (8
    requests of 1-8 objects over 32 resident meshes, ~8M triangles) as one
    batch. Each leg warms until its buffers stop growing, then times passes for
    at least 0.25 s."""
    out = {}

    def run(name, frames, handles, batch):
        bufs = [ctx.frame_buffers(s.frame.width, s.frame.height) for s in frames]
        ptrs = [(b[0].ptr, b[1].ptr) for b in bufs]
        tris = sum(s.num_faces for s in frames)
        if batch:
            prepared = [ctx.prepare_batch(frames, handles, ptrs)]
        else:
            prepared = [ctx.prepare_batch([s], [h], [p]) for s, h, p in zip(frames, handles, ptrs)]

        def one_pass():
            for pb in prepared:
                ctx.submit_batch(pb)

        warm = warm_until_stable(ctx, one_pass)
        n, dt = timed_passes(ctx, one_pass)
        out[name] = {"mtris_s": round(tris * n / dt / 1e6, 1),
                     "frames_per_s": round(len(frames) * n / dt, 1),
                     "triangles_per_frame_avg": int(tris / len(frames)),
                     "form": "batched" if batch else "plain",
                     "warmup_passes": warm, "timed_passes": n, "timed_s": round(dt, 3)}
        for b in bufs:
            b[0].free()
            b[1].free()

    c3 = scenes.config3()
    th = [ctx.upload_texture(t) for t in c3.textures]
    mh = [ctx.upload_mesh(m, th) for m in c3.meshes]
    run("config3_1080p_64_instances", [c3], [mh], False)
    c5 = scenes.config5(64)
    th = [ctx.upload_texture(t) for t in c5.textures]
    mh = [ctx.upload_mesh(m, th) for m in c5.meshes]
    run("config5_8_requests_8M_resident", c5.requests[:8], [mh] * 8, True)
    out["config5_serving"] = serving(ctx, c5.requests, mh)
    return out


def serving(ctx, requests, mh):
    """Serving form on the config-5 resident set: 64 distinct requests (1-8
    objects, 640x360 to 1920x1080), latency from submission to colour and depth
    in host memory. 'serial' draws one request at a time (gr_draw: launch, wait,
    copy out); 'saturated' submits a burst of 512 at once; the qps lines are
    open-loop Poisson streams of the stated rate (gr_server_replay);
    'saturated_colour_only' is the burst with colour read back alone."""
    import numpy as np
    from gorender_amd.serve import PreparedRequest, Server, poisson_arrivals, replay

    def summary(lat, wall, n, bs=None):
        d = {"p50_ms": round(float(np.percentile(lat, 50)), 3),
             "p99_ms": round(float(np.percentile(lat, 99)), 3),
             "requests": int(n), "completed_per_s": round(n / wall, 1)}
        if bs is not None:
            d["mean_batch"] = round(float(np.mean(bs)), 2)
        return d

    for s in requests[:8]:  # warm-up: capacity and allocations
        ctx.draw(s, mh)
    lat = []
    t0 = time.perf_counter()
    for s in requests:
        t = time.perf_counter()
        ctx.draw(s, mh)
        lat.append((time.perf_counter() - t) * 1e3)
    res = {"serial": summary(lat, time.perf_counter() - t0, len(requests))}
    reqs = [PreparedRequest(s, mh) for s in requests]
    with Server(ctx, target_budget=2 << 30, max_batch=16, max_inflight=3) as srv:
        replay(srv, reqs, np.zeros(64))  # warm-up: target pools
        t = time.perf_counter()
        lat, bs = replay(srv, reqs, np.zeros(512))
        res["saturated"] = summary(lat, time.perf_counter() - t, 512, bs)
        for qps, n in ((100, 200), (1000, 1000), (10000, 2000)):
            t = time.perf_counter()
            lat, bs = replay(srv, reqs, poisson_arrivals(n, qps, seed=qps))
            res[f"qps_{qps}"] = summary(lat, time.perf_counter() - t, n, bs)
    # The 10,000 requests/s stream exceeds capacity: with a bounded queue (64
    # waiting) and a 20 ms admission deadline the latency of what is served
    # stays bounded; refused and shed requests are counted beside it.
    with Server(ctx, target_budget=2 << 30, max_batch=16, max_inflight=3, max_queue=64,
                deadline_ms=20) as srv:
        replay(srv, reqs, np.zeros(64))
        t = time.perf_counter()
        lat, bs = replay(srv, reqs, poisson_arrivals(2000, 10000, seed=10000))
        wall = time.perf_counter() - t
        code = np.round(lat * 1e6).astype(np.int64)
        ok = lat >= 0
        d = summary(lat[ok], wall, int(ok.sum()), bs[ok])
        d.update({"submitted": 2000, "refused_queue_full": int((code == -1).sum()),
                  "shed_past_deadline": int((code == -2).sum()), "max_queue": 64,
                  "deadline_ms": 20})
        res["qps_10000_bounded"] = d
    # A window's present needs colour only: the same burst without depth readback.
    with Server(ctx, target_budget=2 << 30, max_batch=16, max_inflight=3,
                colour_only=True) as srv:
        replay(srv, reqs, np.zeros(64))
        t = time.perf_counter()
        lat, bs = replay(srv, reqs, np.zeros(512))
        res["saturated_colour_only"] = summary(lat, time.perf_counter() - t, 512, bs)
    res["config"] = {"target_budget_bytes": 2 << 30, "max_batch": 16, "max_inflight": 3,
                     "distinct_requests": len(requests)}
    return res


def profile_config5(steps, warmup, tile_split, pipelines):
    """The config-5 batch (8 requests over the 32 resident meshes) launched
    `steps` times, for kernel traces of the split-tile path."""
    from gorender_amd import scenes
    from gorender_amd.device import Context
    ctx = Context(0)
    ctx.set_config(tile_split=tile_split, pipelines=pipelines)
    rs = scenes.config5(8)
    th = [ctx.upload_texture(t) for t in rs.textures]
    mh = [ctx.upload_mesh(m, th) for m in rs.meshes]
    bufs = [ctx.frame_buffers(s.frame.width, s.frame.height) for s in rs.requests]
    pb = ctx.prepare_batch(rs.requests, [mh] * 8, [(b[0].ptr, b[1].ptr) for b in bufs])
    warm_until_stable(ctx, lambda: ctx.submit_batch(pb), min_passes=max(2, warmup))
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.submit_batch(pb)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    st = ctx.last_stats()
    print(json.dumps({"mode": "config5", "launches": steps, "ms_per_launch": round(dt / steps * 1e3, 4),
                      "tiles_split": sum(x["tiles_split"] for x in st)}))


def present_rates(ctx, variants, mh, W, H, n=256):
    """Frames per second through gr_present_draw / gr_present_wait with two
    frames in flight, as a window loop runs them (main.go:201-227): frame k+1
    is queued before frame k is waited for; each wait leaves the frame's pixels
    in pinned host memory (read where they lie, no further copy). Colour only
    (what a window presents), and colour with depth."""
    import ctypes
    from gorender_amd import _lib
    rp = ctypes.POINTER(ctypes.c_uint8)()
    dp = ctypes.POINTER(ctypes.c_float)()
    st = _lib.gr_stats()

    def wait():
        _lib.check(ctx.lib.gr_present_wait(ctx.handle, ctypes.byref(rp), ctypes.byref(dp),
                                           ctypes.byref(st)))
        return rp[0]

    out = {}
    for name, with_depth in (("colour", False), ("colour_depth", True)):
        ctx.present_draw(variants[0], mh, with_depth=with_depth)  # warm-up: buffers
        wait()
        t0 = time.perf_counter()
        ctx.present_draw(variants[0], mh, with_depth=with_depth)
        for k in range(1, n):
            ctx.present_draw(variants[k % len(variants)], mh, with_depth=with_depth)
            wait()
        wait()
        out[name] = round(n / (time.perf_counter() - t0), 1)
    out["bytes_per_frame"] = {"colour": W * H * 4, "colour_depth": W * H * 8}
    return out


def pct(xs, p):
    xs = sorted(xs)
    k = min(len(xs) - 1, max(0, int(math.ceil(p / 100.0 * len(xs))) - 1))
    return xs[k]


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this render path runs on one GPU; --gpus must be 1 "
                                   "(sharding frames over several GPUs is a design note, "
                                   "not built)"}))
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0 or args.batch < 1:
        raise SystemExit("--steps >= 1, --warmup >= 0 and --batch >= 1")

    import numpy as np
    from gorender_amd import scenes
    from gorender_amd.device import Context

    if args.mode == "config5":
        profile_config5(args.steps, args.warmup, args.tile_split, args.pipelines)
        return

    scene = scenes.config2()
    fr = scene.frame
    W, H = fr.width, fr.height
    faces = scene.num_faces
    aspect = float(np.float32(W) / np.float32(H))
    P = scenes.perspective_matrix(scenes.FOV_45, aspect, fr.znear, fr.zfar)
    Vw = scenes.view_matrix((0, 0, 1.2), (0, 0, -1), (0, 1, 0))

    def frame_scene(k):
        w = scenes.world_matrix((1, 1, 1), (0, 0.01 * k, 0), (0, 0, 0))
        return scenes.SceneData(fr, scene.meshes, scene.textures,
                                [scenes.ObjectData(0, scenes.mvp_matrix(P, Vw, w), w)],
                                name=f"c2-{k}")

    ctx = Context(0)
    mh, _ = ctx.upload_scene(scene)
    B = args.batch
    n_variants = B  # the B distinct frames every launch draws
    variants = [frame_scene(k) for k in range(n_variants)]
    outs = [ctx.frame_buffers(W, H) for _ in range(B)]
    out_ptrs = [(o[0].ptr, o[1].ptr) for o in outs]
    # ctypes arguments are built before timing: the timed loops only submit.
    # Two copies of the batch, for the two graphs that alternate on the pipes.
    batches = [ctx.prepare_batch(variants, [mh] * B, out_ptrs) for _ in range(2)]
    singles = [ctx.prepare_batch([variants[k]], [mh], [out_ptrs[k]]) for k in range(B)]
    tw, th = [int(v) for v in args.tile.split("x")]
    ctx.set_config(tile_w=tw, tile_h=th, tile_split=args.tile_split, pipelines=args.pipelines)

    step_i = [0]

    def stream_step():
        ctx.submit_batch(batches[step_i[0] & 1])
        step_i[0] += 1

    def plain_step():  # B frames, one launch each
        for k in range(B):
            ctx.submit_batch(singles[k])
        step_i[0] += 1

    def run_form(step, warmup, steps):
        """Warm-up steps (capacity settles: a pass that overflowed is drawn
        again), then exactly `steps` steps between two synchronisations."""
        for _ in range(warmup):
            capacity_safe(ctx, step)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        ctx.synchronize()
        return time.perf_counter() - t0

    K, Wm = args.steps, args.warmup
    modes = {"all": ("stream", "graph", "plain"), "stream": ("stream",), "graph": ("graph",),
             "plain": ("plain",)}[args.mode]
    # The stream form first: the context's buffers are allocated once at their
    # largest size before the single-frame form runs.
    capacity_safe(ctx, stream_step)
    capacity_safe(ctx, stream_step)
    wall = {}
    if "stream" in modes:
        wall["stream"] = run_form(stream_step, Wm, K)
    graphs = []
    if "graph" in modes:
        graphs = [ctx.batch_graph(b) for b in batches]

        def graph_step():
            graphs[step_i[0] & 1].launch()
            step_i[0] += 1

        wall["graph"] = run_form(graph_step, Wm, K)
    stats = ctx.last_stats()[0]
    persistent = None
    if "graph" in modes and not args.profile:
        # The raster stage as a persistent grid striding over the tiles
        # (gr_config.raster_grid 2), same batches in the graph form: reported
        # beside the default one workgroup per tile, never as `value`.
        for g in graphs:
            g.destroy()
        ctx.set_config(raster_grid=2)
        graphs = [ctx.batch_graph(b) for b in batches]
        persistent = K * B / run_form(graph_step, Wm, K)
        for g in graphs:
            g.destroy()
        graphs = []
        ctx.set_config(raster_grid=1)
    if "plain" in modes:
        wall["plain"] = run_form(plain_step, min(Wm, 4), K)
        ctx.set_config(pipelines=1)
        wall["plain_one_pipeline"] = run_form(plain_step, min(Wm, 4), K)
        ctx.set_config(pipelines=args.pipelines)
    fps = {k: K * B / v for k, v in wall.items()}

    # Per-kernel durations: HIP events around every launch, one pipeline, so
    # that no two kernels overlap inside a measured duration.
    kt_batched, kt_plain = {}, {}
    ctx.set_config(pipelines=1)
    if "stream" in modes or "graph" in modes:
        ctx.reset_kernel_times()
        ctx.set_kernel_timing(True)
        for _ in range(4):
            stream_step()
        ctx.synchronize()
        kt_batched = ctx.kernel_times()
        ctx.set_kernel_timing(False)
    if "plain" in modes:
        ctx.reset_kernel_times()
        ctx.set_kernel_timing(True)
        plain_step()
        ctx.synchronize()
        kt_plain = ctx.kernel_times()
        ctx.set_kernel_timing(False)
    ctx.set_config(pipelines=args.pipelines)

    # The device's copy bandwidth: k_copy16 over calib_copy_mib MiB, timed by
    # HIP events (read + write bytes per second).
    copy_gbs = None
    if args.calib_copy_mib > 0:
        nb = args.calib_copy_mib << 20
        src, dst = ctx.alloc(nb), ctx.alloc(nb)
        src.fill(0x3C)
        dst.copy_from(src)
        ctx.synchronize()
        ctx.reset_kernel_times()
        ctx.set_kernel_timing(True)
        for _ in range(4):
            dst.copy_from(src)
        ctx.synchronize()
        ms, n = ctx.kernel_times().get("k_copy", (0.0, 0))
        ctx.set_kernel_timing(False)
        ctx.reset_kernel_times()
        if n:
            copy_gbs = 2 * nb / (ms / n / 1e3) / 1e9
        src.free()
        dst.free()

    c4 = None
    extra = None
    if not args.no_extra and not args.profile:
        for g in graphs:
            g.destroy()
        graphs = []
        c4 = config4_leg(ctx, scenes)
        extra = other_configs(ctx, scenes)

    lat_dev, lat_host = [], []
    if not args.profile:
        for k in range(100):
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.submit_batch(singles[k % n_variants])
            ctx.synchronize()
            lat_dev.append((time.perf_counter() - t0) * 1e3)
        ctx.draw(variants[0], mh)  # untimed: the pinned readback buffer is allocated once
        for k in range(50):
            t0 = time.perf_counter()
            ctx.draw(variants[k % n_variants], mh)
            lat_host.append((time.perf_counter() - t0) * 1e3)
    present = None if args.profile else present_rates(ctx, variants, mh, W, H)

    batched = [k for k in ("graph", "stream") if k in fps]
    main_form = max(batched, key=lambda k: fps[k]) if batched else "plain"
    main_fps = fps[main_form]
    shaded = stats["pixels_shaded"]
    covered = stats["fragments_covered"]
    B_frame = scene.compulsory_bytes()

    # Roofline of the dominant kernel in the batched form.
    kt_dom = kt_batched or kt_plain
    frames_per_launch = B if kt_batched else 1
    dom_name, (dom_ms, dom_launches) = max(kt_dom.items(), key=lambda kv: kv[1][0])
    avg_launch_s = dom_ms / dom_launches / 1e3
    alg_bytes = {
        "k_raster": W * H * 8 + 16 * 16 * 4,
        "k_face": faces * 12 + scene.meshes[0].num_vertices * (12 + 12 + 8),
    }
    dom_bytes = alg_bytes.get(dom_name, 0) * frames_per_launch
    achieved = dom_bytes / avg_launch_s / 1e9 if dom_bytes else 0.0
    sq = sq_counters(dom_name, ctx.config(), frames_per_launch)
    valu_issue = None
    if sq.get("SQ_INSTS_VALU"):
        ach = sq["SQ_INSTS_VALU"] / avg_launch_s / 1e9
        valu_issue = {"kernel": dom_name, "instr_per_launch": int(sq["SQ_INSTS_VALU"]),
                      "achieved_ginstr_s": round(ach, 1), "peak_ginstr_s": VALU_PEAK_GINSTR_S,
                      "frac": round(ach / VALU_PEAK_GINSTR_S, 4),
                      "source": "SQ_INSTS_VALU, profiles/sq_counters_config2.txt"}
        if sq.get("SQ_ACTIVE_INST_VALU2") is not None:
            # Two VALU instructions can issue in one quad-cycle: the share that
            # did, and the fraction of the launch a SIMD's VALU issue was busy
            # counting such a pair once (1,024 SIMDs at 2.4 GHz, 4 cycles each).
            pairs = sq["SQ_ACTIVE_INST_VALU2"]
            busy = (sq["SQ_INSTS_VALU"] - pairs) * 4 / (1024 * 2.4e9 * avg_launch_s)
            valu_issue["paired_frac"] = round(2 * pairs / sq["SQ_INSTS_VALU"], 4)
            valu_issue["issue_busy_frac"] = round(busy, 4)
        if sq.get("SQ_THREAD_CYCLES_VALU") is not None:
            valu_issue["active_lanes_per_instr"] = round(
                sq["SQ_THREAD_CYCLES_VALU"] / sq["SQ_INSTS_VALU"], 2)
    lds_conflict = (round(sq["SQ_LDS_BANK_CONFLICT"] / sq["SQ_ACTIVE_INST_LDS"], 4)
                    if sq.get("SQ_ACTIVE_INST_LDS") else None)
    occupancy = (round(sq["MEAN_WAVES_PER_SIMD"], 3) if sq.get("MEAN_WAVES_PER_SIMD") else None)

    cpu, cpu_forms = None, None
    if not args.no_cpu and not args.profile:
        from oracle import refcpu
        ps = refcpu.PreparedScene(variants[0])

        def time_cpu(threads, tiles, seconds):
            refcpu.draw(variants[0], ps, parallel=threads, tiles=(tiles, tiles))
            t0 = time.perf_counter()
            n = 0
            while time.perf_counter() - t0 < seconds or n < 3:
                refcpu.draw(variants[0], ps, parallel=threads, tiles=(tiles, tiles))
                n += 1
            dt = time.perf_counter() - t0
            return faces * n / dt / 1e6, n, dt

        # par16: the reference's parallel structure (object-parallel geometry,
        # 4x4 tile raster, 16 workers); seq: one thread, sequential draw order.
        threads = min(16, os.cpu_count() or 1)
        tiles = int(math.ceil(math.sqrt(threads)))
        v16, n16, dt16 = time_cpu(threads, tiles, 0.7 * args.cpu_seconds)
        v1, n1, dt1 = time_cpu(0, 1, 0.3 * args.cpu_seconds)
        cpu = {"value": round(v16, 3), "unit": "Mtris/s", "cores": threads,
               "kind": "port",
               "sample": f"{n16} frames of the same 200k-triangle 1280x720 frame, "
                         f"oracle/refcpu.c rc_draw_parallel ({threads} threads, "
                         f"{tiles}x{tiles} tiles), {dt16:.1f} s; a C restatement, not "
                         f"gorender's Go binary"}
        cpu_forms = {"par16": round(v16, 3), "seq": round(v1, 3), "seq_frames": n1,
                     "unit": "Mtris/s", "host_cpus": os.cpu_count()}

    out = {
        "metric": "Mtris/s + shaded Mpix/s @1280x720; p50/p99 frame latency, batched vs plain",
        "value": round(faces * main_fps / 1e6, 3),
        "unit": "Mtris/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": round(wall[main_form] / K * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 displaced sphere + 16x16 texture)",
        "config": {"workload": "headline frame: 250x400 lat-long displaced sphere, 200,000 "
                               "triangles, 16x16 RGBA8 texture, 1280x720, camera r=1.2",
                   "step": f"one batched launch of {B} distinct frames",
                   "form": main_form,
                   "frames_per_batch": B, "triangles_per_frame": faces,
                   "tile": [ctx.config().tile_w, ctx.config().tile_h],
                   "tile_split": ctx.config().tile_split,
                   "pipelines": ctx.config().pipelines},
        "shaded_mpix_s": round(shaded * main_fps / 1e6, 3),
        "covered_mfrag_s": round(covered * main_fps / 1e6, 3),
        "frames_per_s": {k: round(v, 2) for k, v in fps.items()},
        "batched_over_plain": (round(main_fps / max(fps["plain"], fps["plain_one_pipeline"]), 3)
                               if "plain" in fps else None),
        "persistent_raster_graph": ({"frames_per_s": round(persistent, 2),
                                     "vs_one_workgroup_per_tile": round(persistent / fps["graph"], 3)}
                                    if persistent else None),
        "mtris_s": {k: round(faces * v / 1e6, 3) for k, v in fps.items()},
        "latency_ms": ({"device_p50": round(pct(lat_dev, 50), 4),
                        "device_p99": round(pct(lat_dev, 99), 4),
                        "host_p50": round(pct(lat_host, 50), 4),
                        "host_p99": round(pct(lat_host, 99), 4)} if lat_dev else None),
        # Frames delivered into host memory by the present pair (two in
        # flight): the PCIe-inclusive rate of a window loop, never `value`.
        "present_frames_per_s": present,
        "frame_stats": stats,
        "compulsory_bytes_per_frame": B_frame,
        "frame_roofline": {"achieved_GBs": round(B_frame * main_fps / 1e9, 2),
                           "frac": round(B_frame * main_fps / 1e9 / HBM_PEAK_GBS, 5)},
        "kernel_ms_per_frame": {"batched": {k: round(v[0] / (v[1] * B), 5)
                                            for k, v in kt_batched.items()},
                                "plain": {k: round(v[0] / v[1], 5)
                                          for k, v in kt_plain.items()}},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic(dom_name, ctx.config(), frames_per_launch),
                     "avg_launch_us": round(avg_launch_s * 1e6, 3),
                     "algorithmic_bytes_per_launch": dom_bytes,
                     "measured_copy_GBs": round(copy_gbs, 1) if copy_gbs else None,
                     "lds_bank_conflicts_per_lds_instr": lds_conflict,
                     "mean_waves_per_simd": occupancy},
        # The same kernel against its compute ceiling: VALU instructions per
        # launch (counter profile) over the live launch time.
        "valu_issue": valu_issue,
        "cpu_baseline": cpu,
        "cpu_baseline_forms": cpu_forms,
        "config4_256_mixed_frames": c4,
        "other_configs": extra,
    }
    print(json.dumps(out))
    for g in graphs:
        g.destroy()
    ctx.close()


if __name__ == "__main__":
    main()
