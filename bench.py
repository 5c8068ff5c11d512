"""Benchmark of the render path on one MI355X.

Workload: the headline frame (displaced lat-long sphere of 200,000 triangles,
16x16 RGBA8 texture, 1280x720, camera at r=1.2). A step is one frame: the full
path (vertex transform, clip, bin, raster, shade, colour+depth written to HBM)
with meshes and textures resident in HBM. Frame k rotates the object by
0.01*k rad about Y, as the reference's demo loop does (main.go:229-233), so
every frame is distinct work.

Modes timed: "batched" (B frames per gr_draw_batch call, the reported value)
and "plain" (one gr_draw_device call per frame). Latency p50/p99 is measured
per frame from submit to output-ready in device memory, and separately to
output-ready in host memory (PCIe readback included).

Prints one JSON line (the contract of this repo's driver).
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
    ap.add_argument("--steps", type=int, default=4096)
    ap.add_argument("--warmup", type=int, default=1024)
    ap.add_argument("--batch", type=int, default=64, help="frames per batched launch")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="bound on the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: no CPU baseline, no latency loops")
    ap.add_argument("--mode", choices=["both", "batched", "plain", "config5"], default="both",
                    help="forms to run (rocprofv3 runs profile one form at a time); config5 "
                         "only launches the 8-request serving batch --steps times")
    ap.add_argument("--tile", default="32x32",
                    help="raster tile WxH, or 'auto' to time the candidates first and keep "
                         "the fastest")
    ap.add_argument("--tile-split", type=int, default=0,
                    help="tile-list entries per raster workgroup before a tile is split "
                         "(0 keeps the library default)")
    ap.add_argument("--pipelines", type=int, default=2, choices=[1, 2],
                    help="context pipelines: 2 overlaps a draw's geometry kernels with the "
                         "previous draw's raster (the library default)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the throughput lines of configs 3-5")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run k_copy over this many MiB (known-size kernel that the "
                         "HBM byte counters are calibrated against)")
    return ap.parse_args()


def measured_traffic(kernel, cfg, frames_per_launch):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by the ratio a known-size
    copy kernel reads under the same counters), when it was taken on this
    kernel, tile and batch size; else None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_config2.txt")
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


def measured_valu(kernel, cfg, frames_per_launch):
    """VALU wave-instructions per launch of `kernel` (SQ_INSTS_VALU) from the
    committed rocprofv3 counter summary of the headline batch, when it was
    taken with this tile and 64 frames per launch; else None."""
    path = os.path.join(ROOT, "profiles", "sq_counters_config2.txt")
    if frames_per_launch != 64 or not os.path.exists(path):
        return None
    name = f"{kernel}<{cfg.tile_w}, {cfg.tile_h}>" if kernel.startswith("k_raster") else kernel
    tile_ok = False
    for line in open(path):
        if line.startswith("# raster tile "):
            tile_ok = line.split()[-1] == f"{cfg.tile_w}x{cfg.tile_h}"
        elif tile_ok and line.startswith(name + " ") and "SQ_INSTS_VALU=" in line:
            return int(line.split("SQ_INSTS_VALU=")[1].split()[0])
    return None


def other_configs(ctx, scenes):
    """Throughput of configs 3-5 on the same context: config 3 (64 instances,
    ~780k triangles, 1920x1080) drawn one frame per call; config 4 (16 frames of
    1k-500k triangles at four resolutions) as one batch; config 5 (8 requests of
    1-8 objects over 32 resident meshes, ~8M triangles) as one batch."""
    import numpy as np
    out = {}

    def run(name, frames, handles, reps, batch):
        bufs = [ctx.frame_buffers(s.frame.width, s.frame.height) for s in frames]
        ptrs = [(b[0].ptr, b[1].ptr) for b in bufs]
        tris = sum(s.num_faces for s in frames)
        if batch:
            prepared = [ctx.prepare_batch(frames, handles, ptrs)]
        else:
            prepared = [ctx.prepare_batch([s], [h], [p]) for s, h, p in zip(frames, handles, ptrs)]
        for _ in range(2):  # warm-up (grows tile-list capacity if needed)
            for pb in prepared:
                ctx.submit_batch(pb)
            try:
                ctx.synchronize()
            except Exception:
                pass
        t0 = time.perf_counter()
        for _ in range(reps):
            for pb in prepared:
                ctx.submit_batch(pb)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        out[name] = {"mtris_s": round(tris * reps / dt / 1e6, 1),
                     "frames_per_s": round(len(frames) * reps / dt, 1),
                     "triangles_per_frame_avg": int(tris / len(frames)),
                     "form": "batched" if batch else "plain"}
        for b in bufs:
            b[0].free()
            b[1].free()

    c3 = scenes.config3()
    th = [ctx.upload_texture(t) for t in c3.textures]
    mh = [ctx.upload_mesh(m, th) for m in c3.meshes]
    run("config3_1080p_64_instances", [c3], [mh], 100, False)
    c4 = scenes.config4_frames(16)
    h4 = [ctx.upload_scene(s)[0] for s in c4]
    run("config4_16_mixed_frames", c4, h4, 40, True)
    c5 = scenes.config5(64)
    th = [ctx.upload_texture(t) for t in c5.textures]
    mh = [ctx.upload_mesh(m, th) for m in c5.meshes]
    run("config5_8_requests_8M_resident", c5.requests[:8], [mh] * 8, 20, True)
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
    with Server(ctx, target_budget=2 << 30, max_batch=16, max_inflight=2) as srv:
        replay(srv, reqs, np.zeros(64))  # warm-up: target pools
        t = time.perf_counter()
        lat, bs = replay(srv, reqs, np.zeros(512))
        res["saturated"] = summary(lat, time.perf_counter() - t, 512, bs)
        for qps, n in ((100, 200), (1000, 1000), (10000, 2000)):
            t = time.perf_counter()
            lat, bs = replay(srv, reqs, poisson_arrivals(n, qps, seed=qps))
            res[f"qps_{qps}"] = summary(lat, time.perf_counter() - t, n, bs)
    # A window's present needs colour only: the same burst without depth readback.
    with Server(ctx, target_budget=2 << 30, max_batch=16, max_inflight=2,
                colour_only=True) as srv:
        replay(srv, reqs, np.zeros(64))
        t = time.perf_counter()
        lat, bs = replay(srv, reqs, np.zeros(512))
        res["saturated_colour_only"] = summary(lat, time.perf_counter() - t, 512, bs)
    res["config"] = {"target_budget_bytes": 2 << 30, "max_batch": 16, "max_inflight": 2,
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
    for _ in range(max(2, warmup)):
        ctx.submit_batch(pb)
        try:
            ctx.synchronize()
        except Exception:
            pass  # the first launch grows the tile-list capacity
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.submit_batch(pb)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    st = ctx.last_stats()
    print(json.dumps({"mode": "config5", "launches": steps, "ms_per_launch": round(dt / steps * 1e3, 4),
                      "tiles_split": sum(x["tiles_split"] for x in st)}))


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
        s = scenes.SceneData(fr, scene.meshes, scene.textures,
                             [scenes.ObjectData(0, scenes.mvp_matrix(P, Vw, w), w)],
                             name=f"c2-{k}")
        return s

    ctx = Context(0)
    mh, _ = ctx.upload_scene(scene)

    n_variants = 64
    variants = [frame_scene(k) for k in range(n_variants)]
    B = args.batch
    outs = [ctx.frame_buffers(W, H) for _ in range(B)]
    out_ptrs = [(o[0].ptr, o[1].ptr) for o in outs]

    # ctypes arguments are built before timing: the timed loops only submit.
    batches = [ctx.prepare_batch([variants[(j * B + i) % n_variants] for i in range(B)],
                                 [mh] * B, out_ptrs) for j in range(n_variants)]
    singles = [ctx.prepare_batch([variants[k]], [mh], [out_ptrs[k % B]])
               for k in range(n_variants)]

    def run_batched(k0, nframes):
        assert nframes % B == 0
        for j in range(k0 // B, (k0 + nframes) // B):
            ctx.submit_batch(batches[j % n_variants])

    def run_plain(k0, nframes):
        for k in range(k0, k0 + nframes):
            ctx.submit_batch(singles[k % n_variants])

    def set_tile(tw, th):
        ctx.set_config(tile_w=tw, tile_h=th, tile_split=args.tile_split,
                       pipelines=args.pipelines)

    tile_sweep = {}
    if args.tile == "auto":
        # Short timed runs of the batched form per candidate; the fastest is kept.
        for tw, th in [(16, 16), (32, 16), (16, 32), (32, 32), (64, 16)]:
            set_tile(tw, th)
            run_batched(0, 4 * B)
            ctx.synchronize()
            t0 = time.perf_counter()
            run_batched(0, 16 * B)
            ctx.synchronize()
            tile_sweep[f"{tw}x{th}"] = round(16 * B / (time.perf_counter() - t0), 1)
        best = max(tile_sweep, key=tile_sweep.get)
        set_tile(*[int(v) for v in best.split("x")])
    else:
        set_tile(*[int(v) for v in args.tile.split("x")])

    do_b = args.mode in ("both", "batched")
    do_p = args.mode in ("both", "plain")
    steps = (max(args.steps, B) + B - 1) // B * B
    warm = (max(args.warmup, B) + B - 1) // B * B
    # warm-up (also grows tile-list capacity if needed); the batched form first,
    # so that the context's buffers are allocated once at their largest size
    def warm_up():
        if do_b:
            run_batched(0, warm)
        if do_p:
            run_plain(0, max(1, warm // 4))
        if do_b:
            run_batched(0, warm)
        ctx.synchronize()

    from gorender_amd import _lib
    try:
        warm_up()
    except _lib.GorenderError as e:
        if e.status != _lib.GR_ERR_CAPACITY:
            raise
        warm_up()  # the library has grown the tile-list capacity: draw again
    stats = ctx.last_stats()[0]

    def timed(fn, nframes):
        ctx.synchronize()
        t0 = time.perf_counter()
        fn(0, nframes)
        ctx.synchronize()
        return time.perf_counter() - t0

    plain_steps = max(B, steps // 2)
    wall_b = timed(run_batched, steps) if do_b else float("nan")
    wall_p = timed(run_plain, plain_steps) if do_p else float("nan")

    # per-kernel durations (HIP events around every launch, separate pass with
    # one pipeline, so that no two kernels overlap inside a measured duration)
    kt_batched, kt_plain = {}, {}
    ctx.set_config(pipelines=1)
    # plain form strictly one frame after another (one pipeline: no two
    # frames' kernels overlap)
    wall_p1 = float("nan")
    if do_p:
        run_plain(0, 64)
        wall_p1 = timed(run_plain, plain_steps)
    if do_b:
        ctx.reset_kernel_times()
        ctx.set_kernel_timing(True)
        run_batched(0, 4 * B)
        ctx.synchronize()
        kt_batched = ctx.kernel_times()
        ctx.set_kernel_timing(False)
    if do_p:
        ctx.reset_kernel_times()
        ctx.set_kernel_timing(True)
        run_plain(0, 64)
        ctx.synchronize()
        kt_plain = ctx.kernel_times()
        ctx.set_kernel_timing(False)
    ctx.set_config(pipelines=args.pipelines)

    if args.calib_copy_mib > 0:
        nb = args.calib_copy_mib << 20
        src, dst = ctx.alloc(nb), ctx.alloc(nb)
        src.fill(0x3C)
        for _ in range(4):
            dst.copy_from(src)
        ctx.synchronize()
        src.free()
        dst.free()

    extra = None
    if not args.no_extra and not args.profile:
        extra = other_configs(ctx, scenes)

    lat_dev, lat_host = [], []
    if not args.profile:
        for k in range(100):
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.submit_batch(singles[k % n_variants])
            ctx.synchronize()
            lat_dev.append((time.perf_counter() - t0) * 1e3)
        for k in range(50):
            t0 = time.perf_counter()
            ctx.draw(variants[k % n_variants], mh)
            lat_host.append((time.perf_counter() - t0) * 1e3)

    # correctness guard on the measured workload: a frame of the timed loop
    # against the oracle is done by tests; here record the counters only.
    frames_s_b = steps / wall_b
    frames_s_p = plain_steps / wall_p
    frames_s_p1 = plain_steps / wall_p1
    mtris_b = faces * frames_s_b / 1e6
    mtris_p = faces * frames_s_p / 1e6
    # the reported value is the batched form (the plain form when run alone)
    main_fps = frames_s_b if do_b else frames_s_p
    main_steps, main_wall = (steps, wall_b) if do_b else (plain_steps, wall_p)
    shaded = stats["pixels_shaded"]
    covered = stats["fragments_covered"]
    B_frame = scene.compulsory_bytes()

    # roofline of the dominant kernel in the batched form
    kt_dom = kt_batched or kt_plain
    frames_per_launch = B if kt_batched else 1
    dom = max(kt_dom.items(), key=lambda kv: kv[1][0])
    dom_name, (dom_ms, dom_launches) = dom[0], dom[1]
    avg_launch_s = dom_ms / dom_launches / 1e3
    alg_bytes = {
        "k_raster": W * H * 8 + 16 * 16 * 4,
        "k_face": faces * 12 + scene.meshes[0].num_vertices * (12 + 12 + 8),
        "k_vertex": scene.meshes[0].num_vertices * 24,
        "k_bin_fill": 0,
        "k_bin_scan": 0,
    }
    dom_bytes = alg_bytes.get(dom_name, 0) * frames_per_launch
    achieved = dom_bytes / avg_launch_s / 1e9 if dom_bytes else 0.0

    n_valu = measured_valu(dom_name, ctx.config(), frames_per_launch)
    valu_issue = None
    if n_valu:
        ach = n_valu / avg_launch_s / 1e9
        valu_issue = {"kernel": dom_name, "instr_per_launch": n_valu,
                      "achieved_ginstr_s": round(ach, 1), "peak_ginstr_s": VALU_PEAK_GINSTR_S,
                      "frac": round(ach / VALU_PEAK_GINSTR_S, 4),
                      "source": "SQ_INSTS_VALU, profiles/sq_counters_config2.txt"}

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
        "steps": main_steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_wall / main_steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 displaced sphere + 16x16 texture)",
        "config": {"workload": "headline frame: 250x400 lat-long displaced sphere, 200,000 "
                               "triangles, 16x16 RGBA8 texture, 1280x720, camera r=1.2",
                   "frames_per_batch": B, "triangles_per_frame": faces,
                   "tile": [ctx.config().tile_w, ctx.config().tile_h],
                   "tile_split": ctx.config().tile_split,
                   "pipelines": ctx.config().pipelines},
        "shaded_mpix_s": round(shaded * main_fps / 1e6, 3),
        "covered_mfrag_s": round(covered * main_fps / 1e6, 3),
        "frames_per_s": {"batched": round(frames_s_b, 2), "plain": round(frames_s_p, 2),
                         "plain_one_pipeline": round(frames_s_p1, 2),
                         "batched_over_plain": round(frames_s_b / frames_s_p, 3),
                         "batched_over_plain_one_pipeline": round(frames_s_b / frames_s_p1, 3)},
        "mtris_s": {"batched": round(mtris_b, 3), "plain": round(mtris_p, 3)},
        "latency_ms": ({"device_p50": round(pct(lat_dev, 50), 4),
                        "device_p99": round(pct(lat_dev, 99), 4),
                        "host_p50": round(pct(lat_host, 50), 4),
                        "host_p99": round(pct(lat_host, 99), 4)} if lat_dev else None),
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
                     "algorithmic_bytes_per_launch": dom_bytes},
        # The same kernel against its compute ceiling: VALU instructions per
        # launch (counter profile) over the live launch time.
        "valu_issue": valu_issue,
        "cpu_baseline": cpu,
        "cpu_baseline_forms": cpu_forms,
        "tile_sweep_fps": tile_sweep or None,
        "other_configs": extra,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
