"""Benchmark: closest-hit throughput of the traversal microbenchmark.

Workload (one GPU): 1,048,576 random triangles (centres U[0,1]^3, log-uniform
edges in [1e-3, 1e-1]) and 4,194,304 rays (half coherent primary rays of a
2048x2048 pinhole, half incoherent).  One step = one closest-hit query of all
rays, rays and outputs resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu]

Roofline: algorithmic bytes per ray = 24 (ray in) + 16 (prim, t, u, v out) =
40 B; achieved GB/s = 40 * rays / kernel time, against 8 TB/s.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_RAY_CLOSEST = 40
BYTES_PER_RAY_ANY = 25
HBM_PEAK_GBS = 8000.0


def _counters_from_profiles():
    """rocprofv3 PMC figures for the traversal kernel (committed summary of the
    FETCH_SIZE / WRITE_SIZE / TCC / SQ passes over this same workload)."""
    p = os.path.join(ROOT, "profiles", "trace_wide_microbench_traffic.json")
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        return json.load(f)


# Host threads for the CPU baselines: a one-GPU job is granted 16 host cores
# (a share of the node's), whatever the node's logical CPU count says, so
# the CPU legs use at most 16 threads; the node's count and CPU model are
# recorded beside each baseline.
CPU_THREADS = 16


def host_info():
    """CPU model and the logical CPUs the OS reports on the host running the bench."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "hardware_concurrency": os.cpu_count(),
            "threads_used": None, "thread_cap": CPU_THREADS,
            "thread_cap_reason": "the one-GPU job's share of the host is 16 cores"}


def _cpu_threads():
    import oracle
    return min(CPU_THREADS, oracle.num_threads())


def _host(threads):
    h = host_info()
    h["threads_used"] = threads
    return h


def cpu_baseline(scene, o, d, budget_s=12.0):
    """The oracle (reference algorithm restated in C++) on host cores over a
    bounded sample of the same rays: closest-hit rays/s."""
    import numpy as np

    import oracle
    threads = _cpu_threads()
    t0 = time.time()
    orc = oracle.OracleScene(scene)
    build_s = time.time() - t0
    rng = np.random.default_rng(1234)
    done, spent, chunk = 0, 0.0, 4096
    while spent < budget_s and done < len(o):
        sel = rng.choice(len(o), chunk, replace=False)
        t1 = time.perf_counter()
        orc.trace(o[sel], d[sel], threads=threads)
        spent += time.perf_counter() - t1
        done += chunk
        if spent < budget_s / 8:
            chunk *= 2
    return {"value": done / spent / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{done} rays drawn uniformly from the same 4,194,304-ray workload through the same "
                      f"1,048,576-triangle scene; reference median-split tree built in {build_s:.1f} s (untimed)",
            "host": _host(threads)}


_TERRAIN = {}


def _terrain(ctx):
    """The 1080p path-integration scene (2.9 M triangles), built once per run."""
    if "scene" not in _TERRAIN:
        from voidray_amd import workloads as W
        s = W.terrain_spheres()
        _TERRAIN["scene"], _TERRAIN["accel"] = s, s.build_acceleration(ctx)
    return _TERRAIN["scene"], _TERRAIN["accel"]


def _cache_leg(ctx, torch, stream, width=960, height=540, ref_spp=4096, train_steps=8, warm_frames=0,
               warm_size=(240, 135, 2), query_roughness=0.3, max_resolution=1024, lr_grid=1e-2, lr_mlp=1e-3,
               depths=(2, 1), train_every=1):
    """Learned radiance cache on the terrain scene (depth 16, RR from 4, env IS
    + MIS) at width x height, compared with the plain tracer at ref_spp by
    relMSE = mean((x - r)^2 / (r^2 + 0.01)) at equal device time.
    Plain: progressive 4-spp passes, image snapshot after each.  Cached:
    `warm_frames` warm-up frames that only train (the same camera at
    warm_size = (w, h, spp), every path a training path, the image discarded;
    counted in the cached run's time; none by default: the cache learns from
    the frames it renders), then 32 frames of 4 spp.  In each frame
    ~2^16 randomly chosen paths run to the end as training paths and record up
    to 4 vertices with their suffix radiance; the others stop at vertex
    `cache_depth` and add the cache's answer; 8 Adam steps on disjoint slices
    of the samples follow.  The same run with training off (untrained cache:
    truncation alone) is the control.  Every cache object first renders one
    untimed frame without training, which sizes its buffers as the reference
    render sized the plain tracer's.  Device time per pass from HIP events;
    snapshots are excluded."""
    import numpy as np

    from voidray_amd import _native as N
    from voidray_amd.device import RadianceCache, make_render_desc, scene_bounds
    from voidray_amd.scene import RenderSettings
    sp = stream.cuda_stream
    torch.cuda.set_stream(stream)  # snapshots (clone, cpu) must follow the renders on this stream
    scene, accel = _terrain(ctx)
    w, h = width, height
    ext = N.VR_RENDER_ENV_IS | N.VR_RENDER_RUSSIAN_ROULETTE | N.VR_RENDER_ALPHA_PER_CALL
    st = RenderSettings(total_samples=1, max_bounces=16, firefly_clamp=float("inf"))

    def ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record(stream)
        return e

    def relmse(x, r):
        x = x[..., :3].astype(np.float64)
        return float(np.mean((x - r) ** 2 / (r * r + 0.01)))

    img = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    e0 = ev()
    for k in range(0, ref_spp, 64):
        accel.render_device(img.data_ptr(), make_render_desc(w, h, 64, ref_spp, st, sample_offset=k, seed=4242,
                                                             flags=ext, rr_depth=4), stream=sp)
    e1 = ev()
    torch.cuda.synchronize()
    ref_ms = e0.elapsed_time(e1)
    ref = img[..., :3].double().cpu().numpy()

    def box8(x):
        hh, ww = x.shape[0] // 8 * 8, x.shape[1] // 8 * 8
        return x[:hh, :ww, :3].reshape(hh // 8, 8, ww // 8, 8, 3).mean((1, 3))

    ref8 = box8(ref)

    def curve(passes, run_pass, total):
        """Runs `passes` passes with an image snapshot after each; returns
        [(cumulative device ms, spp so far, relMSE of the snapshot rescaled to
        its spp, relMSE of its 8x8-pixel box means)]."""
        img.zero_()
        torch.cuda.synchronize()
        spans, snaps = [], []
        done = 0
        for k in range(passes):
            a = ev()
            done += run_pass(k)
            b = ev()
            spans.append((a, b))
            snaps.append((done, img.clone() if done else None))
        torch.cuda.synchronize()
        cum = np.cumsum([a.elapsed_time(b) for a, b in spans])
        out = []
        for k, (d, s) in enumerate(snaps):
            if d:
                x = s.cpu().numpy()[..., :3].astype(np.float64) * (total / d)
                out.append((float(cum[k]), d, relmse(x, ref), relmse(box8(x), ref8)))
        return out

    spf = 4

    def plain_pass(k):
        accel.render_device(img.data_ptr(), make_render_desc(w, h, spf, 32 * spf, st, sample_offset=k * spf, seed=7,
                                                             flags=ext, rr_depth=4), stream=sp)
        return spf

    plain_pts = curve(32, plain_pass, 32 * spf)
    plain_img = img[..., :3].double().cpu().numpy()
    out = {"scene": f"terrain {w}x{h}, {scene.num_triangles} triangles", "reference_spp": ref_spp,
           "reference_ms": round(ref_ms, 1),
           "plain": [{"ms": round(t, 2), "spp": s, "relmse": round(e, 5), "relmse_8x8": round(e8, 6)}
                     for t, s, e, e8 in plain_pts[::4] + plain_pts[-1:]]}
    lo, hi = scene_bounds(scene)
    warm, frames = warm_frames, 32
    ww, wh, wspp = warm_size
    scratch = torch.zeros_like(img)

    for depth in depths:
        for trained in (True, False):
            cache = RadianceCache(ctx, lo, hi, log2_table=19, base_resolution=16, max_resolution=max_resolution,
                                  lr_grid=lr_grid, lr_mlp=lr_mlp)
            for sw, sh, sspp in ((w, h, spf), (ww, wh, wspp)):  # untimed buffer sizing, no training
                cache.render_frame(accel, scratch.data_ptr(), make_render_desc(sw, sh, sspp, sspp, st, seed=3, flags=ext,
                                                                               rr_depth=4),
                                   train_paths=1 << 16, train_vertices=4, cache_depth=depth, train_steps=0, frame=0,
                                   stream=sp)
            torch.cuda.synchronize()

            def cached_pass(k, depth=depth, cache=cache, trained=trained):
                if k < warm:  # training-only frame: small image, every path trains, image discarded
                    cache.render_frame(accel, scratch.data_ptr(),
                                       make_render_desc(ww, wh, wspp, wspp, st, sample_offset=0, seed=11 + k, flags=ext,
                                                        rr_depth=4),
                                       train_paths=ww * wh * wspp, train_vertices=4, cache_depth=depth,
                                       train_steps=train_steps if trained else 0, frame=k, stream=sp,
                                       query_roughness=query_roughness)
                    return 0
                train = (k - warm) % train_every == 0  # frames between training frames only query
                cache.render_frame(accel, img.data_ptr(), make_render_desc(w, h, spf, frames * spf, st,
                                                                           sample_offset=(k - warm) * spf, seed=7,
                                                                           flags=ext, rr_depth=4),
                                   train_paths=(1 << 16) if train else 0, train_vertices=4, cache_depth=depth,
                                   train_steps=train_steps if trained and train else 0, frame=k, stream=sp,
                                   query_roughness=query_roughness)
                return spf

            pts = curve(warm + frames, cached_pass, frames * spf)
            stats = cache.stats()
            cache.close()
            final = img[..., :3].double().cpu().numpy()
            # plain relMSE at the cached run's final time: log-log interpolation of the plain curve
            pt = np.log([p[0] for p in plain_pts])
            t_end, _, e_end, e8_end = pts[-1]

            def plain_at(col):
                """Plain tracer's error (column col) at the cached run's final
                time: log-log interpolation, 1/t beyond its last point."""
                v = float(np.exp(np.interp(np.log(t_end), pt, np.log([p[col] for p in plain_pts]),
                                           left=np.nan, right=np.nan)))
                return v if np.isfinite(v) else plain_pts[-1][col] * plain_pts[-1][0] / t_end

            # time the cached run needs to reach the plain tracer's final error
            target_t = plain_pts[-1][0]
            reach = next((p[0] for p in pts if p[2] <= plain_pts[-1][2]), None)
            reach8 = next((p[0] for p in pts if p[3] <= plain_pts[-1][3]), None)
            out[f"cache_depth_{depth}_{'trained' if trained else 'untrained'}"] = {
                "points": [{"ms": round(t, 2), "spp": s, "relmse": round(e, 5), "relmse_8x8": round(e8, 6)}
                           for t, s, e, e8 in pts[::4] + pts[-1:]],
                "final": {"ms": round(t_end, 2), "relmse": round(e_end, 5),
                          "mean_over_reference": round(float(final.mean() / ref.mean()), 4),
                          "relmse_8x8": round(relmse(box8(final), ref8), 6)},
                "train_samples_per_frame": stats["train_samples"], "queries_per_frame": stats["queries"],
                "train_ms_per_frame": round(stats["train_ms"], 3), "render_ms_per_frame": round(stats["render_ms"], 3),
                "equal_time_relmse_ratio_plain_over_cache": round(plain_at(2) / e_end, 3),
                "time_to_reach_plain_final_relmse_ratio": (round(reach / target_t, 3) if reach else None),
                # the same on 8x8-pixel box means: low-frequency error, which a
                # biased (dark) cache cannot average away
                "equal_time_relmse_8x8_ratio_plain_over_cache": round(plain_at(3) / e8_end, 3),
                "time_to_reach_plain_final_relmse_8x8_ratio": (round(reach8 / target_t, 3) if reach8 else None)}
    out["plain_final"] = {"ms": round(plain_pts[-1][0], 2), "relmse": round(plain_pts[-1][2], 5),
                          "mean_over_reference": round(float(plain_img.mean() / ref.mean()), 4),
                          "relmse_8x8": round(relmse(box8(plain_img), ref8), 6)}
    out["query_roughness"] = query_roughness
    out["grid"] = {"log2_table": 19, "base_resolution": 16, "max_resolution": max_resolution, "lr_grid": lr_grid,
                   "lr_mlp": lr_mlp, "warm_frames": warm_frames, "train_every": train_every}
    out["roofline"] = _cache_roofline(ctx, torch, stream, lo, hi)
    return out


NRC_FLOP_INFER = 2 * (128 * 64 + 3 * 64 * 64 + 64 * 16)   # 43,008 per query: the MLP forward
NRC_FLOP_TRAIN = 3 * NRC_FLOP_INFER                          # forward, input gradients, weight gradients
MFMA_BF16_PEAK_TFLOPS = 2500.0                               # MI355X dense BF16


def _cache_roofline(ctx, torch, stream, lo, hi, n_infer=1 << 20, n_train=1 << 16, reps=10):
    """MFMA roofline of the learned cache's kernels, timed live with HIP events
    on the stream they run on: inference (one k_nrc_infer launch over n_infer
    queries) and one training step over n_train samples (k_nrc_train with its
    grid-gradient atomics, k_nrc_wgrad, the reduction and Adam).  Queries:
    positions uniform in the scene box, random unit directions and normals,
    roughness U(0, 1); training targets U(0, 1)."""
    from voidray_amd.device import RadianceCache
    sp = stream.cuda_stream
    g = torch.Generator(device="cuda").manual_seed(5)
    n = max(n_infer, n_train)
    lo_t = torch.tensor(lo, dtype=torch.float32, device="cuda")
    hi_t = torch.tensor(hi, dtype=torch.float32, device="cuda")
    pos = lo_t + (hi_t - lo_t) * torch.rand((n, 3), device="cuda", generator=g)
    dirs = torch.nn.functional.normalize(torch.randn((n, 3), device="cuda", generator=g), dim=1)
    nrm = torch.nn.functional.normalize(torch.randn((n, 3), device="cuda", generator=g), dim=1)
    rough = torch.rand(n, device="cuda", generator=g)
    target = torch.rand((n, 3), device="cuda", generator=g)
    out = torch.empty((n, 3), device="cuda")
    torch.cuda.synchronize()
    cache = RadianceCache(ctx, lo, hi, log2_table=19, base_resolution=16, max_resolution=2048)
    args = (pos.data_ptr(), dirs.data_ptr(), nrm.data_ptr(), rough.data_ptr())

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms_inf = timed(lambda: cache.infer(n_infer, *args, out.data_ptr(), stream=sp))
    ms_tr = timed(lambda: cache.train(n_train, *args, target.data_ptr(), stream=sp))
    cache.close()
    tf_inf = NRC_FLOP_INFER * n_infer / (ms_inf * 1e-3) / 1e12
    tf_tr = NRC_FLOP_TRAIN * n_train / (ms_tr * 1e-3) / 1e12
    return {"bound": "mfma", "kernel": "k_nrc_infer", "achieved": round(tf_inf, 2), "peak": MFMA_BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tf_inf / MFMA_BF16_PEAK_TFLOPS, 5), "traffic": None,
            "infer_ms": round(ms_inf, 4), "queries": n_infer, "flop_per_query": NRC_FLOP_INFER,
            "grid_gather_bytes_per_query": 512,
            "train_step": {"ms": round(ms_tr, 4), "samples": n_train, "flop_per_sample": NRC_FLOP_TRAIN,
                           "achieved_tflops": round(tf_tr, 2), "frac": round(tf_tr / MFMA_BF16_PEAK_TFLOPS, 5),
                           "kernels": "k_nrc_train + k_nrc_wgrad + k_nrc_reduce + k_adam + k_adam_grid4"}}


def _render_roofline(st, env_cdf=False):
    """HBM roofline of a render by its algorithmic bytes: per path segment 144 B
    (ray SoA 24 B written and read, hit 16 B written and read, path state 32 B
    read and written), 16 B per sample (RGBA accumulation), 32 B per textured
    hit, 48 B per environment miss and, with environment importance sampling,
    8 B per CDF lookup; counts from the run's own counters."""
    b = 144 * st["segments"] + 16 * st["samples"] + 32 * st["textured_hits"] + 48 * st["env_misses"]
    if env_cdf:
        b += 8 * st["env_is_samples"]
    gbs = b / (st["kernel_ms"] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 2), "peak": 8000.0, "unit": "GB/s", "frac": round(gbs / 8000.0, 5),
            "algorithmic_bytes": int(b), "textured_hits": st["textured_hits"], "env_misses": st["env_misses"],
            "env_is_samples": st["env_is_samples"]}


def _precision(img, ref, rows=None):
    """Max per-pixel relative error (floor |r| >= 1e-3) and relMSE of img against ref."""
    import numpy as np
    x = img[..., :3].astype(np.float64)
    r = ref[..., :3].astype(np.float64)
    if rows is not None:
        x, r = x[rows], r[rows]
    return {"max_rel_err": float(np.max(np.abs(x - r) / np.maximum(np.abs(r), 1e-3))),
            "relmse": float(np.mean((x - r) ** 2 / (r * r + 0.01))),
            "bit_identical_fraction": round(float(np.mean(x == r)), 4)}


def _render_legs(ctx, cpu=True, c3_spp=64):
    """Path-integration legs reported beside the headline metric:
    * cornell 512x512, 16 spp, depth 8, reference estimator (firefly clamp 3);
    * terrain+spheres 1920x1080, 64 spp, depth 16, RR from 4, env IS + MIS,
      and the same frame with the reference estimator (no RR, no env IS,
      firefly clamp 3).
    Device time of vr_render (HIP events), samples/s, segments (rays)/s, the
    render's HBM roofline by algorithmic bytes, and, with the CPU legs, the
    oracle timed on the host and the GPU image's error against it."""
    import numpy as np

    from voidray_amd import _native as N
    from voidray_amd.device import make_render_desc
    from voidray_amd import workloads as W
    from voidray_amd.scene import RenderSettings
    out = {}
    c1 = W.cornell_room()
    a1 = c1.build_acceleration(ctx)
    s1 = RenderSettings(total_samples=16, max_bounces=8)
    img = np.zeros((512, 512, 4), np.float32)
    a1.render(img, make_render_desc(512, 512, 1, 16, s1, seed=99))  # warm-up
    img[:] = 0
    d1 = make_render_desc(512, 512, 16, 16, s1, seed=1)
    st = a1.render(img, d1)
    leg = {"ms": round(st["kernel_ms"], 3), "samples_per_s": round(st["samples"] / (st["kernel_ms"] * 1e-3), 1),
           "mrays_per_s": round(st["segments"] / (st["kernel_ms"] * 1e3), 2), "segments": st["segments"],
           "exact_engine_rays": st["fallback_rays"], "roofline": _render_roofline(st)}
    if cpu:
        import oracle
        threads = _cpu_threads()
        ref, sec, _ = oracle.OracleScene(c1).render(d1, threads=threads)
        n = 512 * 512 * 16
        leg["cpu_baseline"] = {"samples_per_s": round(n / sec, 1), "cores": threads, "kind": "port",
                               "sample": "the full frame (512x512, 16 spp), reference estimator and traversal",
                               "host": _host(threads)}
        leg["gpu_over_cpu"] = round(leg["samples_per_s"] / (n / sec), 1)
        leg["precision_vs_f32_restatement"] = _precision(img, ref)
    out["cornell_512_16spp"] = leg
    t0 = time.time()
    c3, a3 = _terrain(ctx)
    build3 = time.time() - t0
    s3 = RenderSettings(total_samples=c3_spp, max_bounces=16, firefly_clamp=float("inf"))
    ext = N.VR_RENDER_ENV_IS | N.VR_RENDER_RUSSIAN_ROULETTE | N.VR_RENDER_ALPHA_PER_CALL
    img3 = np.zeros((1080, 1920, 4), np.float32)
    a3.render(img3, make_render_desc(1920, 1080, 1, c3_spp, s3, seed=99, flags=ext, rr_depth=4))  # warm-up
    img3[:] = 0
    st = a3.render(img3, make_render_desc(1920, 1080, c3_spp, c3_spp, s3, seed=1, flags=ext, rr_depth=4))
    leg = {"ms": round(st["kernel_ms"], 2), "samples_per_s": round(st["samples"] / (st["kernel_ms"] * 1e-3), 1),
           "mrays_per_s": round(st["segments"] / (st["kernel_ms"] * 1e3), 2), "segments": st["segments"],
           "exact_engine_rays": st["fallback_rays"], "triangles": c3.num_triangles, "scene_build_s": round(build3, 2),
           "roofline": _render_roofline(st, env_cdf=True)}
    if cpu:
        import oracle
        threads = _cpu_threads()
        orc = oracle.OracleScene(c3)
        # bounded sample: 1 spp on every 3rd row (360 rows x 1920 pixels), same estimator, reference traversal
        d1 = make_render_desc(1920, 1080, 1, c3_spp, s3, seed=5, flags=ext, rr_depth=4)
        ref, sec, ost = orc.render(d1, threads=threads, row_step=3)
        n = 1920 * 360
        leg["cpu_baseline"] = {"samples_per_s": round(n / sec, 1), "cores": threads, "kind": "port",
                               "sample": f"{n} samples (1 spp, every 3rd row), reference median-split traversal",
                               "host": _host(threads)}
        leg["gpu_over_cpu"] = round(leg["samples_per_s"] / (n / sec), 1)
        # precision: the same 1-spp frame on the GPU against the f32 restatement on
        # those rows, and against the FP64 oracle on every 30th row
        g1 = np.zeros((1080, 1920, 4), np.float32)
        a3.render(g1, d1)
        leg["precision_vs_f32_restatement"] = _precision(g1, ref, rows=slice(0, None, 3))
        ref64, _, _ = oracle.OracleScene(c3, "f64").render(d1, threads=threads, row_step=30, prune=True)
        leg["precision_vs_fp64_oracle"] = _precision(g1, ref64, rows=slice(0, None, 30))
        leg["precision_note"] = "1 spp, seed 5, same random streams; f32 rows 0,3,6,...; FP64 rows 0,30,60,..."
    out["terrain_1080p_64spp"] = leg
    # the same frame with the 200 spheres as instances of one unit sphere
    # (one shared object-space BVH): same random draws, same streams
    t0 = time.time()
    c3i = W.terrain_spheres(instanced=True)
    a3i = c3i.build_acceleration(ctx)
    build3i = time.time() - t0
    imi = np.zeros((1080, 1920, 4), np.float32)
    a3i.render(imi, make_render_desc(1920, 1080, 1, c3_spp, s3, seed=99, flags=ext, rr_depth=4))  # warm-up
    imi[:] = 0
    st = a3i.render(imi, make_render_desc(1920, 1080, c3_spp, c3_spp, s3, seed=1, flags=ext, rr_depth=4))
    out["terrain_1080p_64spp_instanced"] = {
        "ms": round(st["kernel_ms"], 2), "samples_per_s": round(st["samples"] / (st["kernel_ms"] * 1e-3), 1),
        "mrays_per_s": round(st["segments"] / (st["kernel_ms"] * 1e3), 2), "segments": st["segments"],
        "exact_engine_rays": st["fallback_rays"], "instances": len(c3i.instances), "scene_build_s": round(build3i, 2),
        "device_bytes": a3i.stats()["device_bytes"], "device_bytes_flattened": a3.stats()["device_bytes"],
        # the flattened spheres are generated in f64 and rounded, the instances
        # place one f32 unit sphere: the same picture, not the same bits
        "mean_over_flattened": round(float(imi[..., :3].mean() / img3[..., :3].mean()), 5),
        "roofline": _render_roofline(st, env_cdf=True)}
    a3i.release()
    # the same frame under the reference estimator (no RR, no env IS, clamp 3)
    sr = RenderSettings(total_samples=c3_spp, max_bounces=16)
    imr = np.zeros((1080, 1920, 4), np.float32)
    a3.render(imr, make_render_desc(1920, 1080, 1, c3_spp, sr, seed=98))  # warm-up
    imr[:] = 0
    st = a3.render(imr, make_render_desc(1920, 1080, c3_spp, c3_spp, sr, seed=1))
    out["terrain_1080p_64spp_reference_estimator"] = {
        "ms": round(st["kernel_ms"], 2), "samples_per_s": round(st["samples"] / (st["kernel_ms"] * 1e-3), 1),
        "mrays_per_s": round(st["segments"] / (st["kernel_ms"] * 1e3), 2), "segments": st["segments"],
        "exact_engine_rays": st["fallback_rays"], "roofline": _render_roofline(st)}
    return out


def _pct(x, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(x), q)), 3)


def _serving_cpu_baseline(pool, reqs, flags, row_step=64):
    """The 64 requests rendered one at a time by the oracle on the host, each
    on every `row_step`-th row at its full spp (the same estimator, the
    reference median-split traversal): samples/s over the rendered samples.
    Oracle scene builds are not timed."""
    import oracle
    from voidray_amd.device import make_render_desc
    from voidray_amd.scene import RenderSettings
    threads = _cpu_threads()
    t0 = time.time()
    orcs = [oracle.OracleScene(s) for s in pool]
    build_s = time.time() - t0
    spent, n = 0.0, 0
    for i, (si, w, h, spp, seed) in enumerate(reqs):
        st = RenderSettings(total_samples=spp, max_bounces=16, firefly_clamp=float("inf"))
        _, sec, _ = orcs[si].render(make_render_desc(w, h, spp, spp, st, seed=seed, request=i, flags=flags,
                                                     rr_depth=4), threads=threads, row_step=row_step)
        spent += sec
        n += w * len(range(0, h, row_step)) * spp
    return {"samples_per_s": round(n / spent, 1), "cores": threads, "kind": "port", "seconds": round(spent, 2),
            "sample": f"all 64 requests one at a time, every {row_step}th row at full spp ({n} samples); "
                      f"oracle scene builds ({build_s:.1f} s) not timed",
            "host": _host(threads)}


def _serving_leg(ctx, torch, stream, cpu=True):
    """64 render requests (resolutions 256^2..1920x1080, 1-32 spp, own seeds)
    over eight scenes of 10k-2M triangles; depth 16, RR from 4, env IS + MIS.
    Batched: one captured hipGraph launch with the autotuned chunk/block size
    (tuned on a proxy with spp / 8, the pick confirmed against the largest
    chunk on the real batch); completion time of each request from device
    time stamps.  Sequential: the same requests one render call each,
    shortest first, back to back on the same stream; completion times from HIP
    events.  Latency = completion time with all requests submitted at t = 0.
    The per-image checksums of both runs must agree (same random streams)."""
    from voidray_amd import _native as N
    from voidray_amd import workloads as W
    from voidray_amd.device import BatchRenderer, make_render_desc
    from voidray_amd.scene import RenderSettings
    sp = stream.cuda_stream
    t0 = time.time()
    pool = W.serving_scene_pool()
    accels = [s.build_acceleration(ctx) for s in pool]
    build_s = time.time() - t0
    reqs = W.serving_requests()
    flags = (N.VR_RENDER_ENV_IS | N.VR_RENDER_RUSSIAN_ROULETTE | N.VR_RENDER_ALPHA_PER_CALL
             | N.VR_RENDER_OVERWRITE)
    targets, batch, proxy = [], [], []
    for i, (si, w, h, spp, seed) in enumerate(reqs):
        st = RenderSettings(total_samples=spp, max_bounces=16, firefly_clamp=float("inf"))
        t = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
        targets.append(t)
        batch.append((accels[si], make_render_desc(w, h, spp, spp, st, seed=seed, request=i, flags=flags,
                                                   rr_depth=4), t.data_ptr()))
        ps = max(1, spp // 8)
        proxy.append((accels[si], make_render_desc(w, h, ps, ps, st, seed=seed, request=i, flags=flags,
                                                   rr_depth=4), t.data_ptr()))
    samples = sum(w * h * spp for _, w, h, spp, _ in reqs)
    torch.cuda.synchronize()
    br = BatchRenderer(ctx)
    t1 = time.time()
    (chunk, block), tune_ms = br.autotune(proxy, reps=1, shortest_first=True, stream=sp)
    # the proxy has 8x fewer paths, so it cannot see what the largest chunks
    # gain on the real batch: the proxy's pick and the same block at the
    # largest chunk are each run once on the real batch, and the faster one kept
    confirm = {}
    for cand in dict.fromkeys([(chunk, block), (1 << 27, block)]):
        br.submit(batch, tuning=cand, graph=True, shortest_first=True, stream=sp)
        br.stats()
        c0 = torch.cuda.Event(enable_timing=True)
        c1 = torch.cuda.Event(enable_timing=True)
        c0.record(stream)
        br.relaunch(stream=sp)
        c1.record(stream)
        torch.cuda.synchronize()
        confirm[cand] = c0.elapsed_time(c1)
    chunk, block = min(confirm, key=confirm.get)
    tune_s = time.time() - t1
    # batched: capture + first launch (warm-up), then one timed relaunch
    t1 = time.time()
    br.submit(batch, tuning=(chunk, block), graph=True, timestamps=True, shortest_first=True, stream=sp)
    br.stats()
    capture_first_s = time.time() - t1
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    br.relaunch(stream=sp)
    e1.record(stream)
    torch.cuda.synchronize()
    batch_ms = e0.elapsed_time(e1)
    lat_b = br.latencies(len(reqs))
    bst = br.stats()
    sums_b = torch.stack([t.double().sum() for t in targets]).cpu()
    nonfinite = int(sum((~torch.isfinite(t)).sum().item() for t in targets))
    # sequential: one request per call on the same stream, shortest first
    order = sorted(range(len(reqs)), key=lambda k: reqs[k][1] * reqs[k][2] * reqs[k][3])
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(len(reqs) + 1)]
    evs[0].record(stream)
    for k, i in enumerate(order):
        acc, desc, ptr = batch[i]
        acc.render_device(ptr, desc, stream=sp)
        evs[k + 1].record(stream)
    torch.cuda.synchronize()
    lat_s = [evs[0].elapsed_time(e) for e in evs[1:]]
    svc_s = [evs[k].elapsed_time(evs[k + 1]) for k in range(len(reqs))]
    seq_ms = lat_s[-1]
    sums_s = torch.stack([t.double().sum() for t in targets]).cpu()
    roof = _render_roofline(dict(bst, kernel_ms=batch_ms), env_cdf=True)
    cpu_leg = _serving_cpu_baseline(pool, reqs, flags & ~N.VR_RENDER_OVERWRITE) if cpu else None
    return {
        "requests": len(reqs), "samples": samples, "scenes_triangles": [s.num_triangles for s in pool],
        "scene_build_s": round(build_s, 2), "settings": "depth 16, RR from 4, env IS + MIS",
        "autotune": {"chunk_paths": chunk, "block": block, "proxy_ms": round(tune_ms, 2),
                     "seconds": round(tune_s, 1), "proxy": "same requests with spp // 8",
                     "confirmed_on_batch_ms": {f"{c}x{b}": round(v, 1) for (c, b), v in confirm.items()}},
        "batched": {"ms": round(batch_ms, 2), "samples_per_s": round(samples / (batch_ms * 1e-3), 1),
                    "mrays_per_s": round(bst["segments"] / (batch_ms * 1e3), 2),
                    "p50_ms": _pct(lat_b, 50), "p99_ms": _pct(lat_b, 99),
                    "capture_and_first_launch_s": round(capture_first_s, 2)},
        "sequential": {"ms": round(seq_ms, 2), "samples_per_s": round(samples / (seq_ms * 1e-3), 1),
                       "p50_ms": _pct(lat_s, 50), "p99_ms": _pct(lat_s, 99),
                       "service_p50_ms": _pct(svc_s, 50), "service_p99_ms": _pct(svc_s, 99)},
        "batched_over_sequential": round(seq_ms / batch_ms, 3),
        "roofline": roof,
        "cpu_baseline": cpu_leg,
        "gpu_over_cpu": (round(samples / (batch_ms * 1e-3) / cpu_leg["samples_per_s"], 1) if cpu_leg else None),
        "images_equal_checksums": bool(torch.equal(sums_b, sums_s)),
        "nonfinite_values": nonfinite,
    }


def _progress(msg):
    """One line on stderr per finished leg (the JSON result stays the one line on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1 << 20)
    ap.add_argument("--rays", type=int, default=1 << 22)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-render", action="store_true", help="skip the path-integration legs")
    ap.add_argument("--no-serving", action="store_true", help="skip the batched-serving leg")
    ap.add_argument("--only-serving", action="store_true", help="run the batched-serving leg alone (prints it)")
    ap.add_argument("--no-cache", action="store_true", help="skip the learned-radiance-cache leg")
    ap.add_argument("--only-cache", action="store_true", help="run the learned-radiance-cache leg alone (prints it)")
    ap.add_argument("--cache-query-roughness", type=float, default=0.3,
                    help="cached paths end at the first vertex from the query depth on whose roughness input is "
                         "at least this (0: at the query depth whatever the material)")
    ap.add_argument("--cache-train-steps", type=int, default=8,
                    help="Adam steps per cached frame (disjoint slices of the frame's 2^16 samples)")
    ap.add_argument("--cache-max-res", type=int, default=1024, help="finest hash-grid resolution")
    ap.add_argument("--cache-lr", type=float, nargs=2, default=(1e-2, 1e-3), metavar=("GRID", "MLP"),
                    help="Adam learning rates of the hash grid and the MLP")
    ap.add_argument("--cache-depths", default="2,1", help="query depths to run (comma-separated)")
    ap.add_argument("--cache-warm", type=int, default=0, help="training-only warm-up frames")
    ap.add_argument("--cache-train-every", type=int, default=1, help="train on every n-th rendered frame")
    ap.add_argument("--nrc-probe", action="store_true",
                    help="the learned cache's kernels alone (for profiler runs): inference over 2^20 queries and "
                         "training steps over 2^16 samples, 11 launches each, timed with HIP events; prints one line")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="counter runs: 3 closest-hit queries of the workload, then a 1 GiB device copy "
                         "(the calibration kernel for FETCH_SIZE / WRITE_SIZE); prints nothing else")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this implementation runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is not built)"}))
        sys.exit(2)

    import numpy as np
    import torch

    from voidray_amd import workloads as W
    from voidray_amd.device import Context

    dev = torch.device("cuda:0")
    ctx = Context(0)
    if args.only_cache:
        st = torch.cuda.Stream(device=dev)
        torch.cuda.set_stream(st)
        print(json.dumps({"radiance_cache": _cache_leg(ctx, torch, st, train_steps=args.cache_train_steps,
                                                       query_roughness=args.cache_query_roughness,
                                                       max_resolution=args.cache_max_res, lr_grid=args.cache_lr[0],
                                                       lr_mlp=args.cache_lr[1], warm_frames=args.cache_warm,
                                                       depths=tuple(int(x) for x in args.cache_depths.split(",")),
                                                       train_every=args.cache_train_every)}),
              flush=True)
        return
    if args.nrc_probe:
        st = torch.cuda.Stream(device=dev)
        torch.cuda.set_stream(st)
        lo, hi = (-10.0, -3.0, -10.0), (10.0, 6.0, 10.0)  # the terrain scene's box
        print(json.dumps({"radiance_cache_kernels": _cache_roofline(ctx, torch, st, lo, hi)}), flush=True)
        return
    if args.only_serving:
        st = torch.cuda.Stream(device=dev)
        torch.cuda.set_stream(st)
        print(json.dumps({"serving": _serving_leg(ctx, torch, st, cpu=not args.no_cpu)}), flush=True)
        return
    scene = W.random_triangle_scene(args.tris)
    o, d = W.microbench_rays(args.rays)
    t0 = time.time()
    acc = scene.build_acceleration(ctx)
    build_s = time.time() - t0
    n = len(o)
    to = torch.from_numpy(o).to(dev)
    td = torch.from_numpy(d).to(dev)
    prim = torch.empty(n, dtype=torch.int32, device=dev)
    tt = torch.empty(n, dtype=torch.float32, device=dev)
    tu = torch.empty(n, dtype=torch.float32, device=dev)
    tv = torch.empty(n, dtype=torch.float32, device=dev)
    thit = torch.empty(n, dtype=torch.uint8, device=dev)
    # a dedicated stream (default priority): kernels and timing events share it
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream

    def step(lo=0, cnt=n, any_hit=False, timing=False):
        if any_hit:
            acc.trace_device(to.data_ptr() + 12 * lo, td.data_ptr() + 12 * lo, cnt, any_hit=True,
                             d_hit=thit.data_ptr() + lo, stream=sp)
        else:
            acc.trace_device(to.data_ptr() + 12 * lo, td.data_ptr() + 12 * lo, cnt, d_prim=prim.data_ptr() + 4 * lo,
                             d_t=tt.data_ptr() + 4 * lo, d_u=tu.data_ptr() + 4 * lo, d_v=tv.data_ptr() + 4 * lo,
                             stream=sp, timing=timing)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        t1 = time.perf_counter()
        for _ in range(steps):
            fn()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        return ev0.elapsed_time(ev1) / steps, wall

    if args.pmc_probe:
        for _ in range(3):
            step()
        torch.cuda.synchronize()
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": "done", "copy_bytes": 4 << 28, "rays": n}), flush=True)
        return

    ms, wall = timed(step, args.steps, args.warmup)
    fallbacks = acc.last_fallbacks()
    # duration of the dominant kernel alone (HIP events the library records around it)
    kern = []
    for _ in range(args.steps):
        step(timing=True)
        kern.append(acc.last_wide_ms())
    kernel_ms = sum(kern) / len(kern)
    half = n // 2
    ms_coh, _ = timed(lambda: step(0, half), max(5, args.steps // 2), 1)
    ms_inc, _ = timed(lambda: step(half, n - half), max(5, args.steps // 2), 1)
    ms_any, _ = timed(lambda: step(any_hit=True), max(5, args.steps // 2), 1)
    hits = int((prim != -1).sum().item())

    mrays = n / (ms * 1e-3) / 1e6
    achieved = BYTES_PER_RAY_CLOSEST * n / (kernel_ms * 1e-3) / 1e9
    _progress(f"traversal microbench: {mrays:.1f} Mrays/s")
    pmc = _counters_from_profiles()
    cpu = None if args.no_cpu else cpu_baseline(scene, o, d)
    _progress("traversal CPU baseline done")
    renders = None if args.no_render else _render_legs(ctx, cpu=not args.no_cpu)
    _progress("render legs done")
    serving = None if args.no_serving else _serving_leg(ctx, torch, stream, cpu=not args.no_cpu)
    _progress("serving leg done")
    cache = None if args.no_cache else _cache_leg(ctx, torch, stream, train_steps=args.cache_train_steps,
                                                  query_roughness=args.cache_query_roughness,
                                                  max_resolution=args.cache_max_res, lr_grid=args.cache_lr[0],
                                                  lr_mlp=args.cache_lr[1], warm_frames=args.cache_warm,
                                                  depths=tuple(int(x) for x in args.cache_depths.split(",")),
                                                  train_every=args.cache_train_every)
    out = {
        "metric": "Mrays/s closest-hit (traversal microbench: 1M random tris, 4M rays)",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "traversal microbench", "triangles": args.tris, "rays": n,
                   "rays_mix": "1/2 coherent 2048^2 pinhole, 1/2 incoherent", "bvh": "SAH + spatial splits, 4-wide, 8-bit quantised",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": pmc.get("hbm_bytes_per_launch"),
                     "kernel": "k_trace_persist<false, false, false>", "kernel_ms": round(kernel_ms, 4),
                     "bytes_per_ray": BYTES_PER_RAY_CLOSEST,
                     # rocprofv3 PMC passes over the same workload (profiles/)
                     "traffic_bytes_per_ray": pmc.get("hbm_bytes_per_ray"),
                     "l2_hit_rate": pmc.get("l2_hit_rate"),
                     "occupancy_waves_per_simd": pmc.get("occupancy_waves_per_simd"),
                     "valu_busy": pmc.get("valu_busy"),
                     "valu_lane_utilization": pmc.get("valu_lane_utilization")},
        "cpu_baseline": cpu,
        "extra": {
            "coherent_mrays_s": round(half / (ms_coh * 1e-3) / 1e6, 2),
            "incoherent_mrays_s": round((n - half) / (ms_inc * 1e-3) / 1e6, 2),
            "any_hit_mrays_s": round(n / (ms_any * 1e-3) / 1e6, 2),
            "any_hit_roofline_frac": round(BYTES_PER_RAY_ANY * n / (ms_any * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "hit_fraction": round(hits / n, 4),
            "exact_engine_fallback_rays": fallbacks,
            "scene_build_s": round(build_s, 2),
            "scene_stats": acc.stats(),
            "wall_s": round(wall, 3),
            "gpu_over_cpu": (round(mrays / cpu["value"], 1) if cpu else None),
            "render": renders,
            "serving": serving,
            "radiance_cache": cache,
        },
    }
    _progress("radiance cache leg done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
