#!/usr/bin/env python3
"""LC2 evaluation throughput on one MI355X (3D affine, 7^3 patch).

Workload: MR phantom 256x256x192 float32 (1 mm, the moving high-resolution volume)
against a simulated ultrasound volume 192x192x160 uint8 (0.5 mm, the fixed image
that defines the virtual grid), patch radius 3. A step is one fused launch that
evaluates P candidate poses (the 2n+1 / finite-difference style candidate sets an
optimizer step produces). Inputs are resident in HBM before timing starts.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK = 157.3e12
FP64_PEAK = 78.6e12
# Algorithmic work per voxel-pose of the 3D path (US uint8 static grid, MR float32 warped):
# 1 B of ultrasound + the MR footprint inside the US field of view ((0.5/1)^3 * 4 B).
ALG_BYTES_PER_VOXEL = 1.0 + 0.5
ALG_F32_PER_VOXEL = 116.0   # warp 40, Sobel 27, products 5, window sums 42, reduction 2
ALG_F64_PER_VOXEL = 45.0    # centred moments, 2x2 solve, variance/measure, thresholds
ALG_F32_SWEEP = 89.0        # 2D frames from 3D: warp 40, Sobel 14, products 5, sums 28, reduction 2
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "k_fused3d_r3_affine3d_us192_traffic.json")
SIMDS, CLOCK_HZ = 256 * 4, 2.4e9  # MI355X: 256 CUs of 4 SIMDs, peak engine clock


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _f32_equiv(units, f32=ALG_F32_PER_VOXEL, f64=ALG_F64_PER_VOXEL):
    """Algorithmic flops of `units` voxel-poses in FP32 equivalents: an FP64 flop takes
    two FP32 issue slots (78.6 vs 157.3 TFLOP/s), so achieved/157.3 = t_VALU/t."""
    return units * (f32 + f64 * FP32_PEAK / FP64_PEAK)


def _kernel_time(p, mats, launches):
    """Fused-kernel time of `launches` launches of this batch: HIP events around the kernel
    on the engine's stream (timed launches are not graph-replayed, so the events bracket
    the kernel alone). Returns (total ms, launches)."""
    p.profile(True)
    for _ in range(launches):
        p.eval_batch(mats)
    out = p.profile_read()
    p.profile(False)
    return out


def cpu_model():
    """The host CPU's model name (as lscpu reports it) and the cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0)), os.cpu_count()


def _oracle_grids(O, w, ndim=3):
    f, m = w.fixed, w.moving
    fg, mg = O.geom(f.dims, f.spacing, f.origin), O.geom(m.dims, m.spacing, m.origin)
    vd = O.virtual_dims(fg, w.resampled_spacing, ndim)
    return fg, mg, O.geom(vd, w.resampled_spacing, f.origin), vd


def cpu_baseline(w, M, budget_s, checks=()):
    """The CPU leg. (1) The reference's algorithm (FP32 CGLS per voxel, restated in the
    oracle) on the host cores: full warps + Sobel, CGLS on a slab of output planes, scaled
    to the grid; with all OpenMP threads, then with one. (2) The FP64 oracle as the checker
    of the GPU figures in `checks`: each is (name, workload, kind, matrices, GPU LC2s,
    moving data as the GPU stored it); |dLC2| is reported against the FP64 oracle on the
    workload's own data and, for rounded storage, on the rounded data the GPU read."""
    from oracle import oracle as O
    f, m = w.fixed, w.moving
    fg, mg, vg, vd = _oracle_grids(O, w)
    r = w.patch_radius
    planes = vd[2] - 2 * r
    zc = vd[2] // 2
    all_threads = O.num_threads()

    def sample(threads, names, budget):
        O.set_num_threads(threads)
        t0 = time.perf_counter()
        U = O.warp_3d(f.data.astype(np.float32), fg, vg, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0])
        I = O.warp_3d(m.data, mg, vg, M)
        G = O.sobel(I, "ref")
        t_ws = time.perf_counter() - t0

        def timed(fn, k):
            t = time.perf_counter()
            fn(zc - k // 2, zc - k // 2 + k)
            return time.perf_counter() - t

        fns = {"cg_fp32": lambda a, b: O.lc2_cg(U, I, G, r, zlo=a, zhi=b),
               "exact_fp64": lambda a, b: O.lc2_exact(U, I, G, r, zlo=a, zhi=b)}
        out = {}
        for name in names:
            t1 = timed(fns[name], 1)
            k = int(max(1, min(planes, budget / max(t1, 1e-6))))
            tk = timed(fns[name], k) if k > 1 else t1
            t_eval = t_ws + tk * planes / k
            out[name] = (1.0 / t_eval, k, t_eval, t_ws)
            log(f"cpu {name}, {threads} thread(s): warps+Sobel {t_ws:.2f}s, {k} of {planes} "
                f"planes in {tk:.2f}s -> {t_eval:.2f} s/eval")
        return out

    multi = sample(all_threads, ("cg_fp32", "exact_fp64"), budget_s)
    one = sample(1, ("cg_fp32",), budget_s)
    O.set_num_threads(all_threads)
    model, usable, logical = cpu_model()
    v, k, t_eval, _ = multi["cg_fp32"]
    v1, k1, t1_eval, t1_ws = one["cg_fp32"]
    out = {"value": v, "unit": "LC2 evals/s", "cores": all_threads, "kind": "port",
           "sample": (f"reference algorithm (FP32 CGLS, <=20 it) restated in oracle/, "
                      f"{all_threads} OpenMP threads: full warps + Sobel, CGLS on {k} of {planes} "
                      f"output planes, scaled; {t_eval:.2f} s per evaluation"),
           "cpu_model": model, "cpus_usable": usable, "cpus_logical": logical,
           "one_thread": {"value": v1, "unit": "LC2 evals/s", "cores": 1,
                          "sample": f"warps + Sobel in full ({t1_ws:.2f} s), CGLS on {k1} of "
                                    f"{planes} output planes, scaled; {t1_eval:.1f} s per "
                                    f"evaluation"},
           "exact_fp64_evals_per_s": multi["exact_fp64"][0]}
    if checks:
        out["fp64_oracle_check"] = fp64_oracle_checks(O, checks)
    return out


def fp64_oracle_checks(O, checks):
    """|dLC2| of GPU results against the FP64 oracle (exact least squares, FP64 sums) on the
    same inputs. kind "3d": volumes on the ultrasound grid, MR/CT warped; kind "sweep":
    tracked frames, the MR warped onto each frame and the sums taken over all frames."""
    out = {}
    for name, w, kind, mats, gpu, stored in checks:
        t = time.perf_counter()
        f, m = w.fixed, w.moving
        mg = O.geom(m.dims, m.spacing, m.origin)
        sources = {"source": m.data.astype(np.float32)}
        if stored is not None:
            sources["stored"] = np.asarray(stored, np.float32)
        res, U = {}, None
        if kind == "3d":
            fg, _, vg, _ = _oracle_grids(O, w)
            U = O.warp_3d(f.data.astype(np.float32), fg, vg, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0])
        for src_name, src in sources.items():
            l64 = []
            for M in mats:
                if kind == "3d":
                    I = O.warp_3d(src, mg, vg, np.asarray(M).ravel()[:12])
                    l64.append(O.lc2_exact(U, I, O.sobel(I, "sep"), w.patch_radius).lc2)
                else:
                    X, Y = f.dims[0], f.dims[1]
                    vg = O.geom((X, Y, 1), (f.spacing[0], f.spacing[1], 1.0), (0, 0, 0))
                    sm = sw = 0.0
                    for k in range(M.shape[0]):
                        I = O.warp_3d(src, mg, vg, M[k].ravel())[0]
                        r = O.lc2_exact(f.data[k].astype(np.float32), I, O.sobel(I, "sep"),
                                        w.patch_radius)
                        sm += r.sum_measure
                        sw += r.sum_w
                    l64.append(sm / sw if sw > 0 else 0.0)
            res[src_name] = np.asarray(l64)
        g = np.asarray(gpu, np.float64)
        out[name] = {"poses": len(mats),
                     "lc2_gpu": [float(x) for x in g],
                     "lc2_fp64_oracle": [float(x) for x in res["source"]],
                     "max_abs_dlc2_vs_fp64": float(np.abs(g - res["source"]).max())}
        if "stored" in res:
            out[name]["max_abs_dlc2_vs_fp64_on_stored_values"] = float(
                np.abs(g - res["stored"]).max())
        log(f"FP64 oracle check {name}: |dLC2| {out[name]['max_abs_dlc2_vs_fp64']:.2e} "
            f"({time.perf_counter() - t:.1f}s)")
    return out


def batching_figures(eng, E, synth, w, pair, fv, mats, res, batched_f32, checks):
    """The batching and precision forms against the plain single-pose FP32 form (one pose
    per launch, launched eagerly), on the 3D affine workload and on a 256x256 2D affine
    case: batched launches with the MR stored f32 and f16, each launched eagerly and
    replayed from a captured HIP graph. The f16 results join `checks` for the FP64 oracle."""
    m = w.moving
    P = len(mats)
    out = {}

    def rate(fn, n, reps):
        fn()
        fn()   # with graphs on, the second call captures
        eng.sync()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        return n * reps / (time.perf_counter() - t)

    k = iter(range(10 ** 9))
    pair.use_graph(False)
    single = rate(lambda: pair.eval(mats[next(k) % P]), 1, 20)
    pair.use_graph(True)
    single_graph = rate(lambda: pair.eval_batch(mats[next(k) % P][None]), 1, 20)
    batched_graph = rate(lambda: pair.eval_batch(mats), P, 10)
    pair.use_graph(False)
    mv16 = eng.upload(m.data, m.spacing, m.origin, E.F16)
    p16 = eng.pair(fv, mv16, w.resampled_spacing, True, False, w.patch_radius, 3)
    r16 = p16.eval_batch(mats)
    batched16 = rate(lambda: p16.eval_batch(mats), P, 10)
    p16.close()
    mv16.close()
    checks.append(("3d_affine_mr_f16_storage", w, "3d", mats[:2], r16.lc2[:2],
                   m.data.astype(np.float16)))
    out["batching_3d"] = {
        "single_pose_fp32_evals_per_s": single,
        "single_pose_fp32_graph_evals_per_s": single_graph,
        "batched_fp32_evals_per_s": batched_f32,
        "batched_fp32_graph_evals_per_s": batched_graph,
        "batched_fp16_evals_per_s": batched16,
        "batched_fp32_over_single_fp32": batched_f32 / single,
        "batched_fp16_over_single_fp32": batched16 / single,
        "fp16_vs_fp32_max_abs_dlc2": float(np.abs(r16.lc2 - res.lc2).max()),
        "note": "one pose of the 192x192x160 grid already fills the GPU, so batching saves "
                "only per-launch overhead and the partial last wave; f16 storage does not "
                "speed up a kernel bound by vector ALU, and rounding the MR to f16 moves LC2 "
                "(FP64 oracle check)"}
    log(f"3D: single-pose fp32 {single:.0f} evals/s (graph {single_graph:.0f}), batched f32 "
        f"{batched_f32:.0f} (graph {batched_graph:.0f}), batched f16 {batched16:.0f} evals/s")
    # 2D 256x256 affine (ultrasound warped onto the static image): one pose is a few
    # hundred work-groups, far from filling the GPU, so the batch carries the throughput
    w0 = synth.config0(seed=1)
    f0, m0 = w0.fixed, w0.moving
    a = eng.upload(f0.data, f0.spacing, f0.origin)
    b = eng.upload(m0.data, m0.spacing, m0.origin)
    p2 = eng.pair(a, b, w0.resampled_spacing, w0.high_res_is_moving, w0.center_overlaid,
                  w0.patch_radius, 2)
    rng = np.random.default_rng(2)
    m2 = np.stack([p2.pose_to_matrix(synth.perturb(w0.true_pose, rng, 2)) for _ in range(64)])
    s2 = rate(lambda: p2.eval(m2[next(k) % 64]), 1, 200)
    b2 = rate(lambda: p2.eval_batch(m2), 64, 50)
    p2.use_graph(True)
    s2g = rate(lambda: p2.eval_batch(m2[next(k) % 64][None]), 1, 200)
    b2g = rate(lambda: p2.eval_batch(m2), 64, 50)
    out["batching_2d_256"] = {"single_pose_fp32_evals_per_s": s2,
                              "single_pose_fp32_graph_evals_per_s": s2g,
                              "batched_64_fp32_evals_per_s": b2,
                              "batched_64_fp32_graph_evals_per_s": b2g,
                              "batched_over_single": b2 / s2,
                              "config": "256x256 f32 fixed vs u8 moving affine, 7x7 patch, "
                                        "64 poses per launch"}
    log(f"2D 256^2: single-pose {s2:.0f} evals/s (graph {s2g:.0f}), batched 64 {b2:.0f} "
        f"(graph {b2g:.0f}) evals/s")
    for h in (p2, a, b):
        h.close()
    return out


def sweep_figures(eng, E, synth, steps=3, registration=True):
    """Tracked 2D sweep against an MR volume (2D-to-3D): 128 uint8 frames of 512x384
    (0.3 mm) against MR 320x320x256 stored f16, patch 7x7, 64 candidate poses per launch,
    one matrix per frame and pose. Returns whole-sweep evaluations/s and pixel-poses/s, and
    with `registration` the rigid sweep registration figures as well."""
    t0 = time.time()
    w = synth.config2()
    log(f"sweep workload synthesised in {time.time() - t0:.1f}s")
    f, m = w.fixed, w.moving
    fv = eng.upload(f.data, f.spacing, f.origin)
    mv = eng.upload(m.data, m.spacing, m.origin, E.F16)
    p = eng.pair(fv, mv, w.resampled_spacing, True, False, w.patch_radius, 2)
    p.set_per_frame(True)
    rng = np.random.default_rng(7)
    poses = [w.true_pose] + [w.true_pose + np.r_[rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 3)]
                             for _ in range(63)]
    mats = np.stack([synth.sweep_matrices(q, w.extra["frames"], w.extra["centre"]) for q in poses])
    P = len(mats)
    res = p.eval_batch(mats)
    eng.sync()
    t = time.perf_counter()
    for _ in range(steps):
        res = p.eval_batch(mats)
    dt = (time.perf_counter() - t) / steps
    kms, nl = _kernel_time(p, mats, steps)
    X, Y, F = p.virt_dims
    px = P * F * X * Y
    t_k = kms / max(nl, 1) / 1e3
    t_valu = px * (ALG_F32_SWEEP / FP32_PEAK + ALG_F64_PER_VOXEL / FP64_PEAK)
    out = {"sweep_evals_per_s": P / dt, "sweep_pixel_poses_per_s": px / dt,
           "sweep_kernel_ms_per_launch": t_k * 1e3,
           "sweep_roofline": {"bound": "valu", "achieved": _f32_equiv(px, ALG_F32_SWEEP) / t_k / 1e12,
                              "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s", "frac": t_valu / t_k},
           "sweep_config": f"{F} frames {X}x{Y} u8 (0.3 mm) vs MR 320x320x256 f16, 7x7 patch, "
                           f"{P} poses per launch",
           "sweep_lc2_at_true_pose": float(res.lc2[0])}
    check = ("sweep_mr_f16_storage", w, "sweep", mats[:1], res.lc2[:1],
             m.data.astype(np.float16))
    log(f"sweep: {out['sweep_evals_per_s']:.1f} whole-sweep evals/s "
        f"({out['sweep_pixel_poses_per_s'] / 1e9:.2f} G pixel-poses/s)")
    if registration:
        out.update(sweep_registration_figures(p, w, synth))
    p.close()
    fv.close()
    mv.close()
    return out, check


def sweep_registration_figures(p, w, synth, n_runs=3, seed=43):
    """Rigid registrations of the 128-frame tracked sweep (6 DoF shared by all frames,
    one matrix per frame) with the batched trust region (lc2reg_minimize_qmtr) on
    -log(LC2), from the generating pose moved by up to 3 mm / 3 deg per parameter. Every
    objective call evaluates its whole candidate set in one launch. Returns p50 / p99 wall
    time, evaluations and the largest error against the generating pose."""
    from lc2registration_amd import registration as REG
    rng = np.random.default_rng(seed)
    truth = np.asarray(w.true_pose, np.float64)[:6]
    frames, centre = w.extra["frames"], w.extra["centre"]

    def objective(xs):
        mats = np.stack([synth.sweep_matrices(q, frames, centre) for q in xs])
        l = p.eval_batch(mats).lc2
        return np.where(l > 0, -np.log(np.maximum(l, 1e-300)), np.inf)

    times, evals, errs = [], [], []
    for _ in range(n_runs):
        x0 = truth + rng.uniform(-3, 3, 6)
        t = time.perf_counter()
        x, _, nev = REG.minimize_qmtr(objective, x0, x0 - 10, x0 + 10, xtol=1e-3)
        times.append(time.perf_counter() - t)
        evals.append(nev)
        errs.append(np.abs(x - truth))
    ms = np.array(times) * 1e3
    e = np.array(errs)
    out = {"sweep_registration_p50_ms": float(np.percentile(ms, 50)),
           "sweep_registration_p99_ms": float(np.percentile(ms, 99)),
           "sweep_registration_evaluations": [int(v) for v in evals],
           "sweep_registration_max_error_mm": float(e[:, :3].max()),
           "sweep_registration_max_error_deg": float(e[:, 3:].max())}
    log(f"sweep registration p50 {out['sweep_registration_p50_ms']:.0f} ms, {evals} evaluations, "
        f"max error {out['sweep_registration_max_error_mm']:.3f} mm / "
        f"{out['sweep_registration_max_error_deg']:.3f} deg")
    return out


def registration_latency_figures(synth, n_runs=8):
    """Whole 2D registrations through the product path (lc2reg_register_yaml: YAML and raw
    files in, DIRECT stage + two BOBYQA-stage rounds on the GPU, result YAML and log out)
    on seeded 256x256 affine cases (7x7 patch). Returns p50 / p99 wall time and the
    largest parameter error against the generating pose."""
    import tempfile
    from lc2registration_amd import registration as REG
    times, errs = [], []
    with tempfile.TemporaryDirectory() as tmp:
        for seed in range(n_runs):
            w = synth.config0(seed=seed)
            f, m = w.fixed, w.moving
            fp, mp = os.path.join(tmp, f"f{seed}.raw"), os.path.join(tmp, f"m{seed}.raw")
            f.data.tofile(fp)
            m.data.tofile(mp)
            cfg = REG.write_parameters_yaml(
                os.path.join(tmp, f"p{seed}.yaml"), 2,
                (fp, "float32", f.dims, f.spacing, f.origin), (mp, "uint8", m.dims, m.spacing, m.origin),
                "Fixed", w.resampled_spacing, w.patch_radius, tmp, False, [0, 0, 0, 1, 1, 0],
                [-15, -15, -15, 0.9, 0.9, -0.1], [15, 15, 15, 1.1, 1.1, 0.1], direct_maxeval=150,
                direct_entries=(True, True, False, False),
                rounds=[((True, True, False, False), 1e-4), ((True, True, True, True), 1e-4)])
            t = time.perf_counter()
            x, lc2, nev, _ = REG.register_yaml("2D", cfg, write_images=False)
            times.append(time.perf_counter() - t)
            errs.append(float(np.abs(x - w.true_pose).max()))
    ms = np.array(times[1:] if len(times) > 1 else times) * 1e3   # first run pays set-up
    ok = sum(e < 0.5 for e in errs)
    out = {"registration_2d_p50_ms": float(np.percentile(ms, 50)),
           "registration_2d_p99_ms": float(np.percentile(ms, 99)),
           "registration_2d_converged": f"{ok}/{len(errs)} within 0.5 (px, deg, scale) of the "
                                        f"generating pose",
           "registration_2d_median_param_error": float(np.median(errs)),
           "registration_2d_config": "256x256 f32 fixed vs u8 moving affine, 7x7 patch, DIRECT 150 "
                                     "(translation, rotation) + BOBYQA-stage rounds (T+R, all), "
                                     "from identity"}
    log(f"2D registration latency p50 {out['registration_2d_p50_ms']:.1f} ms, "
        f"p99 {out['registration_2d_p99_ms']:.1f} ms, {ok}/{len(errs)} converged")
    return out


def _serve_closed_loop(S, eng, pair, mats, max_poses, slots, duration_s, clients, seed):
    """`clients` threads, each submitting 1-8 poses of `mats` and waiting, for duration_s."""
    import threading
    srv = S.Server(eng, max_poses=max_poses, window_us=0.0, slots=slots)
    lat, stop = [[] for _ in range(clients)], time.perf_counter() + duration_s

    def client(c):
        rng = np.random.default_rng(seed + c)
        while time.perf_counter() < stop:
            idx = rng.integers(0, len(mats), int(rng.integers(1, 9)))
            t = time.perf_counter()
            srv.wait(srv.submit(pair, mats[idx]))
            lat[c].append(time.perf_counter() - t)

    t0 = time.perf_counter()
    th = [threading.Thread(target=client, args=(c,)) for c in range(clients)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    dt = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    ms = np.array([x for l in lat for x in l]) * 1e3
    return {"poses_per_s": st["poses"] / dt, "p50_ms": float(np.percentile(ms, 50)),
            "p99_ms": float(np.percentile(ms, 99)),
            "poses_per_launch": st["poses"] / max(st["launches"], 1)}


def _serve_poisson(S, eng, classes, lam, max_poses, slots, duration_s, seed):
    """Open-loop Poisson arrivals at lam requests/s; each request draws its class by weight
    and 1-8 candidate poses of that class's pair. Returns per-class p50 / p99 latency (from
    arrival to result) and the served poses/s."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    n = max(int(lam * duration_s), 1)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, n))
    weights = np.array([c[3] for c in classes], np.float64)
    kind = rng.choice(len(classes), n, p=weights / weights.sum())
    srv = S.Server(eng, max_poses=max_poses, window_us=0.0, slots=slots)
    lat = [[] for _ in classes]

    def collect(ticket, c, t_arr):
        srv.wait(ticket)
        lat[c].append(time.perf_counter() - t_arr)

    with ThreadPoolExecutor(max_workers=256) as pool:   # one waiter per outstanding request
        t0 = time.perf_counter()
        for a, c in zip(arrivals, kind):
            now = time.perf_counter() - t0
            if a > now:
                time.sleep(a - now)
            _, pair, mats, _ = classes[c]
            idx = rng.integers(0, len(mats), int(rng.integers(1, 9)))
            t_arr = t0 + a
            pool.submit(collect, srv.submit(pair, mats[idx]), c, t_arr)
    dt = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    out = {"served_poses_per_s": st["poses"] / dt,
           "poses_per_launch": st["poses"] / max(st["launches"], 1),
           "max_launches_in_flight": st["max_in_flight"]}
    for (name, _, _, _), l in zip(classes, lat):
        ms = np.array(l) * 1e3
        out[name] = {"requests": len(l), "p50_ms": float(np.percentile(ms, 50)) if len(l) else None,
                     "p99_ms": float(np.percentile(ms, 99)) if len(l) else None}
    return out


def serving_figures(eng, E, synth, pair3d, mats3d, w_ct):
    """Serving loop (lc2serve.h) on the mixed serving workload: six resident pairs in three
    request classes, each request a candidate set of 1-8 poses of one pair:
      40% 2D affine, 256x256 or 512x512 (u8 ultrasound warped onto an f32 image);
      40% 2D-to-3D tracked sweeps of 32 or 64 frames 512x384 u8 against one shared
          MR 256x256x192 stored f16;
      20% 3D: US 192x192x160 u8 against MR 256x256x192 f32, or US 256^3 u8 against the
          int16 CT 512x512x400 (stored f16).
    Open-loop Poisson arrivals at 5, 20, 50, 100, 200 and 1000 requests/s through continuous
    batching (each pair's queued requests merged into one launch of up to 64 poses, up to
    4 launches of different pairs on the device at once while their work-groups fit it
    together), and at 200 and 1000 requests/s through a
    plain queue (one request per launch, one launch at a time); then 8 closed-loop clients
    on the 3D MR pair, in both modes."""
    from lc2registration_amd import serve as S
    t0 = time.time()
    w256, w512 = synth.config0(seed=21), synth.config0(seed=23, n=512)
    ws32 = synth.config2(seed=22, mr_dims=(256, 256, 192), n_frames=32)
    ws64 = synth.config2(seed=22, mr_dims=(256, 256, 192), n_frames=64)   # same MR
    log(f"serving workloads synthesised in {time.time() - t0:.1f}s")
    keep, rng = [], np.random.default_rng(5)

    def pair2d(w):
        f, m = w.fixed, w.moving
        fv, mv = eng.upload(f.data, f.spacing, f.origin), eng.upload(m.data, m.spacing, m.origin)
        p = eng.pair(fv, mv, w.resampled_spacing, w.high_res_is_moving, w.center_overlaid,
                     w.patch_radius, 2)
        keep.extend([p, fv, mv])
        return p, np.stack([p.pose_to_matrix(synth.perturb(w.true_pose, rng, 2)) for _ in range(64)])

    mr = eng.upload(ws32.moving.data, ws32.moving.spacing, ws32.moving.origin, E.F16)

    def sweep(w):
        f = w.fixed
        fv = eng.upload(f.data, f.spacing, f.origin)
        p = eng.pair(fv, mr, w.resampled_spacing, True, False, w.patch_radius, 2)
        p.set_per_frame(True)
        keep.extend([p, fv])
        return p, np.stack([synth.sweep_matrices(w.true_pose + np.r_[rng.uniform(-2, 2, 3),
                                                                   rng.uniform(-2, 2, 3)],
                                                 w.extra["frames"], w.extra["centre"])
                            for _ in range(64)])

    f, m = w_ct.fixed, w_ct.moving
    fct, mct = eng.upload(f.data, f.spacing, f.origin), eng.upload(m.data, m.spacing, m.origin)
    pct = eng.pair(fct, mct, w_ct.resampled_spacing, True, False, w_ct.patch_radius, 3)
    keep.extend([pct, fct, mct])
    mct_mats = np.stack([pct.pose_to_matrix(synth.perturb(w_ct.true_pose, rng, 3)) for _ in range(64)])
    classes = [("2d_affine_256", *pair2d(w256), 0.2), ("2d_affine_512", *pair2d(w512), 0.2),
               ("sweep_32_frames", *sweep(ws32), 0.2), ("sweep_64_frames", *sweep(ws64), 0.2),
               ("3d_mr_us192x192x160", pair3d, mats3d, 0.1), ("3d_ct_us256", pct, mct_mats, 0.1)]
    keep.append(mr)
    used, _ = eng.memory()
    for _, p, mm, _ in classes:            # first launch of each pair and batch size out of the way
        for k in range(1, 9):
            p.eval_batch(mm[:k])
    out = {"serve_resident_bytes": used, "serve_poisson": {}}
    for lam, mp, sl in ((5, 64, 4), (20, 64, 4), (50, 64, 4), (100, 64, 4), (200, 64, 4),
                        (200, 1, 1), (1000, 64, 4), (1000, 1, 1)):
        key = f"{lam}_per_s" + ("_continuous" if mp > 1 else "_queue_one_launch_at_a_time")
        # the same arrivals and requests for both modes at one rate
        r = _serve_poisson(S, eng, classes, lam, mp, sl, max(2.0, 60.0 / lam), seed=lam)
        out["serve_poisson"][key] = r
        log(f"serving Poisson {key}: " + ", ".join(
            f"{n} {r[n]['p50_ms']:.2f}/{r[n]['p99_ms']:.2f}" for n, *_ in classes
            if r[n]["requests"]) + f" ms (p50/p99); {r['poses_per_launch']:.1f} poses per "
            f"launch, up to {r['max_launches_in_flight']} in flight")
    for mode, mp, sl in (("continuous", 64, 4), ("queue_one_launch_at_a_time", 1, 1)):
        r = _serve_closed_loop(S, eng, pair3d, mats3d, mp, sl, 2.0, 8, 11)
        out[f"serve_closed_loop_3d_{mode}"] = r
        log(f"serving closed loop 3D ({mode}): {r['poses_per_s']:.0f} poses/s, p50 "
            f"{r['p50_ms']:.2f} ms, p99 {r['p99_ms']:.2f} ms, {r['poses_per_launch']:.1f} poses per launch")
    for h in keep:
        h.close()
    return out


def registration_serving_figures(eng, E, synth, w3d, w_ct, rates=(5, 20), duration_s=8.0):
    """Whole registrations as requests (lc2s_registrar over the serving loop): open-loop
    Poisson arrivals of registrations drawn from the serving mix -- 40% 2D affine 256x256 /
    512x512 (6 parameters, from the generating pose moved by up to 2 px / 2 deg / 0.02),
    40% rigid tracked-sweep registrations of 32 / 64 frames against one MR 256x256x192 f16,
    20% rigid 3D (US 192x192x160 vs MR 256x256x192 f32; US 256^3 vs CT 512x512x400 f16),
    the last two from the generating pose moved by up to 2 mm / 2 deg. Each is admitted
    (its pair built) when a session is free, runs BOBYQA (xtol_rel 1e-4) with every
    candidate set one serving request, and frees its pair. Returns per-class p50 / p99
    latency from arrival to result, evaluations, and registrations/s."""
    from concurrent.futures import ThreadPoolExecutor
    from lc2registration_amd import serve as S
    keep, classes = [], []
    rng = np.random.default_rng(71)
    for seed, n in ((41, 256), (43, 512)):
        w = synth.config0(seed=seed, n=n)
        f, m = w.fixed, w.moving
        fv, mv = eng.upload(f.data, f.spacing, f.origin), eng.upload(m.data, m.spacing, m.origin)
        keep += [fv, mv]
        half = np.r_[5, 5, 5, 0.08, 0.08, 0.05]

        def make(w=w, fv=fv, mv=mv, half=half):
            x0 = w.true_pose + np.r_[rng.uniform(-2, 2, 3), rng.uniform(-0.02, 0.02, 2), 0]
            return dict(fixed=fv, moving=mv, resampled_spacing=w.resampled_spacing,
                        high_res_is_moving=w.high_res_is_moving,
                        center_overlaid=w.center_overlaid, patch_radius=w.patch_radius, dim=2,
                        x0=x0, lower=x0 - half, upper=x0 + half), w.true_pose
        classes.append((f"2d_affine_{n}", make, 0.2))
    mr = None
    for frames in (32, 64):
        w = synth.config2(seed=22, mr_dims=(256, 256, 192), n_frames=frames)
        if mr is None:
            mr = eng.upload(w.moving.data, w.moving.spacing, w.moving.origin, E.F16)
            keep.append(mr)
        fv = eng.upload(w.fixed.data, w.fixed.spacing, w.fixed.origin)
        keep.append(fv)

        def make(w=w, fv=fv):
            truth = np.asarray(w.true_pose, np.float64)[:6]
            x0 = truth + rng.uniform(-2, 2, 6)
            return dict(fixed=fv, moving=mr, resampled_spacing=w.resampled_spacing,
                        high_res_is_moving=True, center_overlaid=False,
                        patch_radius=w.patch_radius, dim=2, x0=x0, lower=x0 - 4, upper=x0 + 4,
                        frames=w.extra["frames"], frames_centre=w.extra["centre"]), truth
        classes.append((f"sweep_{frames}_frames", make, 0.2))
    for name, w in (("3d_mr_us192x192x160", w3d), ("3d_ct_us256", w_ct)):
        f, m = w.fixed, w.moving
        fv, mv = eng.upload(f.data, f.spacing, f.origin), eng.upload(m.data, m.spacing, m.origin)
        keep += [fv, mv]

        def make(w=w, fv=fv, mv=mv):
            truth = np.asarray(w.true_pose, np.float64)[:6]
            x0 = truth + rng.uniform(-2, 2, 6)
            return dict(fixed=fv, moving=mv, resampled_spacing=w.resampled_spacing,
                        high_res_is_moving=True, center_overlaid=False,
                        patch_radius=w.patch_radius, dim=3, x0=x0, lower=x0 - 4,
                        upper=x0 + 4), truth
        classes.append((name, make, 0.1))
    out = {"config": "registrations drawn from the serving mix; BOBYQA, xtol_rel 1e-4; "
                     "8 sessions, 4 launch slots, 64 poses per merged launch"}
    weights = np.array([c[2] for c in classes])
    for lam in rates:
        srv = S.Server(eng, max_poses=64, slots=4)
        reg = S.Registrar(srv, eng, max_sessions=8)
        n = max(int(lam * duration_s), 1)
        arrivals = np.cumsum(rng.exponential(1.0 / lam, n))
        kind = rng.choice(len(classes), n, p=weights / weights.sum())
        lat = [[] for _ in classes]
        nev = [[] for _ in classes]
        err = [[] for _ in classes]

        def collect(ticket, c, t_arr, truth):
            x, _, ne, _, _ = reg.wait(ticket)
            lat[c].append(time.perf_counter() - t_arr)
            nev[c].append(ne)
            err[c].append(float(np.abs(x[:3] - truth[:3]).max()))

        with ThreadPoolExecutor(max_workers=64) as pool:
            t0 = time.perf_counter()
            for a, c in zip(arrivals, kind):
                now = time.perf_counter() - t0
                if a > now:
                    time.sleep(a - now)
                kw, truth = classes[c][1]()
                pool.submit(collect, reg.submit(**kw), c, t0 + a, truth)
        dt = time.perf_counter() - t0
        st = reg.stats()
        reg.close()
        srv.close()
        r = {"registrations_per_s": st["completed"] / dt, "completed": st["completed"],
             "failed": st["failed"], "max_running": st["max_running"]}
        for (name, _, _), l, ne, e in zip(classes, lat, nev, err):
            ms = np.array(l) * 1e3
            r[name] = {"registrations": len(l),
                       "p50_ms": float(np.percentile(ms, 50)) if len(l) else None,
                       "p99_ms": float(np.percentile(ms, 99)) if len(l) else None,
                       "median_evaluations": float(np.median(ne)) if ne else None,
                       "max_translation_error": max(e) if e else None}
        out[f"{lam}_per_s"] = r
        log(f"registration serving {lam}/s: " + ", ".join(
            f"{nm} {r[nm]['p50_ms']:.0f}/{r[nm]['p99_ms']:.0f} ms" for nm, _, _ in classes
            if r[nm]["registrations"]) + f"; up to {st['max_running']} running")
    for h in keep:
        h.close()
    return {"serve_registrations": out}


def ct_precision_figures(eng, E, synth, w, poses=32, steps=3):
    """CT precision workload: int16 CT 512x512x400 (0.8 mm, HU) against a 256^3 uint8
    ultrasound (0.6 mm), 7^3 patch, 32 poses per launch. The CT stored as half (exact for
    HU in [-2048, 2048], the engine's automatic choice) and as f32: evaluations/s, kernel
    time and whether LC2 is bit-identical."""
    f, m = w.fixed, w.moving
    fv = eng.upload(f.data, f.spacing, f.origin)
    out, res = {}, {}
    rng = np.random.default_rng(31)
    qs = [w.true_pose] + [synth.perturb(w.true_pose, rng, 3) for _ in range(poses - 1)]
    mats = None
    for name, storage, nbytes in (("f16", E.STORE_AUTO, 2), ("f32", E.F32, 4)):
        u0, _ = eng.memory()
        mv = eng.upload(m.data, m.spacing, m.origin, storage)
        u1, _ = eng.memory()
        p = eng.pair(fv, mv, w.resampled_spacing, True, False, w.patch_radius, 3)
        mats = np.stack([p.pose_to_matrix(q) for q in qs])
        res[name] = p.eval_batch(mats)
        eng.sync()
        t = time.perf_counter()
        for _ in range(steps):
            p.eval_batch(mats)
        dt = (time.perf_counter() - t) / steps
        kms, nl = _kernel_time(p, mats, steps)
        t_k = kms / max(nl, 1) / 1e3
        V = int(np.prod(p.virt_dims))
        # algorithmic bytes per pose: the u8 ultrasound once, and the CT region inside the
        # ultrasound field of view at the CT's own resolution and storage width
        fov = V * float(np.prod(np.asarray(w.resampled_spacing) / np.asarray(m.spacing)))
        alg = len(mats) * (V * 1.0 + fov * nbytes)
        t_valu = len(mats) * V * (ALG_F32_PER_VOXEL / FP32_PEAK + ALG_F64_PER_VOXEL / FP64_PEAK)
        out[f"ct_{name}_evals_per_s"] = len(mats) / dt
        out[f"ct_{name}_kernel_ms_per_launch"] = t_k * 1e3
        out[f"ct_{name}_volume_bytes"] = u1 - u0
        out[f"ct_{name}_roofline"] = {
            "bound": "valu" if t_valu >= alg / (HBM_PEAK_GBS * 1e9) else "hbm",
            "hbm": {"achieved": alg / t_k / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": alg / t_k / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": alg},
            "valu": {"achieved": _f32_equiv(len(mats) * V) / t_k / 1e12, "peak": FP32_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": t_valu / t_k}}
        p.close()
        mv.close()
    fv.close()
    out["ct_f16_lc2_bit_identical_to_f32"] = bool(np.array_equal(res["f16"].lc2, res["f32"].lc2))
    out["ct_f16_vs_f32_max_abs_dlc2"] = float(np.abs(res["f16"].lc2 - res["f32"].lc2).max())
    out["ct_lc2_at_true_pose"] = float(res["f16"].lc2[0])
    out["ct_config"] = (f"CT int16 512x512x400 (0.8 mm, HU {int(m.data.min())}..{int(m.data.max())}) "
                        f"vs US 256^3 u8 (0.6 mm), 7^3 patch, {poses} poses per launch")
    log(f"CT: f16 {out['ct_f16_evals_per_s']:.0f} evals/s, f32 {out['ct_f32_evals_per_s']:.0f} "
        f"evals/s, LC2 bit-identical: {out['ct_f16_lc2_bit_identical_to_f32']}")
    check = ("ct_f16_storage", w, "3d", mats[:2], res["f16"].lc2[:2], None)
    return out, check


def registration_3d_figures(w, n_runs=3, seed=41):
    """Whole 3D registrations through the product path (lc2reg_register_yaml) on the 3D
    affine workload (MR 256x256x192 f32 against US 192x192x160 u8, 7^3 patch): from the
    generating pose moved by up to 3 mm / 3 deg in each rigid parameter, with identity
    scale and shear, two BOBYQA-stage rounds (rigid, then all 12 parameters). Returns
    p50 / p99 wall time, evaluations, and the largest error against the generating pose."""
    import tempfile
    from lc2registration_amd import registration as REG
    rng = np.random.default_rng(seed)
    f, m = w.fixed, w.moving
    truth = np.asarray(w.true_pose, np.float64)
    times, evals, errs = [], [], []
    with tempfile.TemporaryDirectory() as tmp:
        fp, mp = os.path.join(tmp, "us.raw"), os.path.join(tmp, "mr.raw")
        f.data.tofile(fp)
        m.data.astype(np.float32).tofile(mp)
        for k in range(n_runs):
            init = np.r_[truth[:6] + rng.uniform(-3, 3, 6), 1, 1, 1, 0, 0, 0]
            cfg = REG.write_parameters_yaml(
                os.path.join(tmp, f"p{k}.yaml"), 3, (fp, "uint8", f.dims, f.spacing, f.origin),
                (mp, "float32", m.dims, m.spacing, m.origin), "Moving", w.resampled_spacing,
                w.patch_radius, tmp, False, init, [-10] * 6 + [0.9] * 3 + [-0.1] * 3,
                [10] * 6 + [1.1] * 3 + [0.1] * 3,
                rounds=[((True, True, False, False), 1e-4), ((True, True, True, True), 1e-4)])
            t = time.perf_counter()
            x, lc2, nev, _ = REG.register_yaml("3D", cfg, write_images=False)
            times.append(time.perf_counter() - t)
            evals.append(nev)
            errs.append(np.abs(x - truth))
    ms = np.array(times) * 1e3
    e = np.array(errs)
    out = {"registration_3d_p50_ms": float(np.percentile(ms, 50)),
           "registration_3d_p99_ms": float(np.percentile(ms, 99)),
           "registration_3d_evaluations": [int(v) for v in evals],
           "registration_3d_max_error_mm": float(e[:, :3].max()),
           "registration_3d_max_error_deg": float(e[:, 3:6].max()),
           "registration_3d_max_error_scale_shear": float(e[:, 6:].max()),
           "registration_3d_config": "3D affine workload, from the generating pose moved by "
                                     "<= 3 mm / 3 deg per rigid parameter, BOBYQA-stage rounds "
                                     f"(rigid, all 12), {n_runs} runs"}
    log(f"3D registration latency p50 {out['registration_3d_p50_ms']:.0f} ms, "
        f"{evals} evaluations, max error {out['registration_3d_max_error_mm']:.3f} mm / "
        f"{out['registration_3d_max_error_deg']:.3f} deg")
    return out


def calibrate_copy(nbytes):
    """A copy kernel of known size: under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE per
    dispatch, divided by nbytes, give the factor that turns those counters into bytes."""
    import torch
    a = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "dispatches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--poses", type=int, default=32, help="candidate poses per fused launch")
    ap.add_argument("--zchunk", type=int, default=0,
                    help="z planes per work-group (0: the engine's launch model)")
    ap.add_argument("--autotune", action="store_true",
                    help="time the z-chunk candidates first (cached per shape)")
    ap.add_argument("--mr-storage", choices=("f32", "f16"), default="f32",
                    help="storage of the MR volume (f16 rounds its values: LC2 changes)")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds per CPU sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip single-pose / fp16 figures")
    ap.add_argument("--graph", action="store_true",
                    help="replay each step from a captured HIP graph instead of launching it "
                         "eagerly (on this stack a replay costs 5-13 us more per batch)")
    ap.add_argument("--sweep-only", action="store_true",
                    help="only the tracked-sweep figure (2D kernel), timed over --steps launches "
                         "(for profiling that kernel)")
    ap.add_argument("--ct-only", action="store_true",
                    help="only the CT precision figures (f16 and f32 storage), timed over --steps "
                         "launches (for profiling that workload)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES three times (the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.sweep_only:
        from lc2registration_amd import engine as E
        from lc2registration_amd import synth
        print(json.dumps(sweep_figures(E.Engine(0), E, synth, steps=args.steps,
                                       registration=False)[0]), flush=True)
        return
    if args.ct_only:
        from lc2registration_amd import engine as E
        from lc2registration_amd import synth
        print(json.dumps(ct_precision_figures(E.Engine(0), E, synth, synth.config3(),
                                              steps=args.steps)[0]), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from lc2registration_amd import engine as E
    from lc2registration_amd import synth

    t0 = time.time()
    w = synth.config1()
    log(f"workload synthesised in {time.time() - t0:.1f}s")
    eng = E.Engine(0)
    f, m = w.fixed, w.moving
    fv = eng.upload(f.data, f.spacing, f.origin)
    mv = eng.upload(m.data, m.spacing, m.origin, E.F16 if args.mr_storage == "f16" else E.F32)
    pair = eng.pair(fv, mv, w.resampled_spacing, True, False, w.patch_radius, 3)
    V = int(np.prod(pair.virt_dims))
    rng = np.random.default_rng(123)
    poses = [w.true_pose] + [synth.perturb(w.true_pose, rng, 3) for _ in range(args.poses - 1)]
    mats = np.stack([pair.pose_to_matrix(p) for p in poses])
    P = len(mats)
    if args.zchunk:
        pair.set_tuning_n(P, args.zchunk)
    elif args.autotune:
        from lc2registration_amd import autotune
        zc = autotune.tune(pair, mats)
        log(f"autotuned z-chunk {zc}")

    # the product path: each step uploads the matrices, launches the fused kernel and the
    # fixed-order reduction, and downloads the results (eagerly, or replayed from a HIP
    # graph with --graph)
    pair.use_graph(args.graph)
    for _ in range(max(args.warmup, 2)):   # the second call captures the graph
        pair.eval_batch(mats)
    eng.sync()
    log(f"warmup done ({args.warmup} steps of {P} poses); timing {args.steps} steps")
    lat = []
    eng.sync()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        res = pair.eval_batch(mats)
        lat.append(time.perf_counter() - ts)
    eng.sync()
    elapsed = time.perf_counter() - t_start
    kernel_ms, launches = _kernel_time(pair, mats, args.steps)
    zchunk_batch = pair.get_tuning()
    avg_kernel_s = kernel_ms / max(launches, 1) / 1e3
    value = args.steps * P / elapsed
    log(f"{value:.1f} evals/s, kernel {avg_kernel_s * 1e3:.3f} ms per launch of {P} poses")

    alg_bytes = P * V * ALG_BYTES_PER_VOXEL
    achieved_gbs = alg_bytes / avg_kernel_s / 1e9
    t_valu = P * V * (ALG_F32_PER_VOXEL / FP32_PEAK + ALG_F64_PER_VOXEL / FP64_PEAK)
    traffic, valu_insts, tr = None, None, {}
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as fh:
            tr = json.load(fh)
        if tr.get("poses") == P and tr.get("mr_storage", "f32") == args.mr_storage:
            traffic = tr.get("hbm_bytes_per_launch")
            valu_insts = tr.get("sq_insts_valu_per_launch")
        else:
            tr = {}

    def ratio(num, den, complement=False):
        a, b = tr.get(num), tr.get(den)
        if not a or not b:
            return None
        return 1.0 - a / b if complement else a / b
    # The binding roof is the vector ALU: achieved = algorithmic FP32 + FP64 work in FP32
    # equivalents per second, so frac = t_VALU / t_kernel.
    roofline = {"bound": "valu", "achieved": _f32_equiv(P * V) / avg_kernel_s / 1e12,
                "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s", "frac": t_valu / avg_kernel_s,
                "traffic": traffic,
                "kernel": "k_fused3d<R=3, moving high-res>",
                "kernel_ms_per_launch": avg_kernel_s * 1e3, "launches": launches,
                "algorithmic_flops_per_launch": {"f32": P * V * ALG_F32_PER_VOXEL,
                                                 "f64": P * V * ALG_F64_PER_VOXEL},
                "valu_note": "algorithmic FP32 116 + FP64 45 flop per voxel-pose; t_VALU = "
                             "F32 / 157.3 T + F64 / 78.6 T",
                "hbm": {"achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved_gbs / HBM_PEAK_GBS,
                        "algorithmic_bytes_per_launch": alg_bytes, "traffic": traffic,
                        "traffic_gbs": traffic / avg_kernel_s / 1e9 if traffic else None},
                # VALU issue: SQ_INSTS_VALU of the profiled launch, 4 cycles per wave64
                # instruction on a 16-lane SIMD, over 1024 SIMDs at 2.4 GHz
                "valu_issue_frac": (valu_insts * 4 / (SIMDS * CLOCK_HZ) / avg_kernel_s
                                    if valu_insts else None),
                "valu_insts_per_voxel_pose": (valu_insts * 64 / (P * V) if valu_insts else None),
                # cache behaviour of the profiled launch (TCC / TCP / SQ counters)
                "l2_hit_rate": (tr["tcc_hit_per_launch"] / (tr["tcc_hit_per_launch"] +
                                                            tr["tcc_miss_per_launch"])
                                if tr.get("tcc_hit_per_launch") else None),
                "vl1d_hit_rate": ratio("tcp_tcc_read_req_per_launch",
                                       "tcp_total_cache_accesses_per_launch", complement=True),
                "lds_conflict_cycle_frac": ratio("sq_lds_bank_conflict_per_launch",
                                                 "sq_active_inst_lds_per_launch")}

    stored = m.data.astype(np.float16) if args.mr_storage == "f16" else None
    extra, checks = {}, [(f"3d_affine_mr_{args.mr_storage}", w, "3d", mats[:2], res.lc2[:2], stored)]
    if not args.no_extra:
        extra.update(batching_figures(eng, E, synth, w, pair, fv, mats, res, value, checks))
        # the mutual-information metric (patchSize 0, 256 x 256 bins) on the same pair
        pair.eval_mi(mats)
        t = time.perf_counter()
        nmi = None
        for _ in range(3):
            nmi = pair.eval_mi(mats)
        extra["mi_evals_per_s"] = 3 * P / (time.perf_counter() - t)
        extra["mi_at_true_pose"] = float(nmi[0])
        log(f"mutual information {extra['mi_evals_per_s']:.1f} evals/s")
        t0 = time.time()
        w_ct = synth.config3()
        log(f"CT workload synthesised in {time.time() - t0:.1f}s")
        ct, ck = ct_precision_figures(eng, E, synth, w_ct)
        extra.update(ct)
        checks.append(ck)
        extra.update(serving_figures(eng, E, synth, pair, mats, w_ct))
        extra.update(registration_serving_figures(eng, E, synth, w, w_ct))
        sw, ck = sweep_figures(eng, E, synth)
        extra.update(sw)
        checks.append(ck)
        extra.update(registration_latency_figures(synth))
        extra.update(registration_3d_figures(w))

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(w, mats[0], args.cpu_budget, checks)

    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "LC2 evals/s (3D affine, 7^3 patch)",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 warp/Sobel, f64 moments/solve/sums",
        "data": "synthetic (seeded phantom + simulated ultrasound)",
        "config": {"workload": f"3D affine LC2: MR 256x256x192 {args.mr_storage} (1 mm, moving) vs "
                               "US 192x192x160 u8 (0.5 mm, fixed grid), patch 7^3",
                   "poses_per_step": P, "virtual_grid": list(pair.virt_dims),
                   "zchunk": zchunk_batch, "graph_replay": args.graph},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "hbm_gbs_algorithmic": achieved_gbs,
        "step_latency_ms": {"p50": float(np.percentile(lat_ms, 50)),
                            "p99": float(np.percentile(lat_ms, 99))},
        "lc2_at_true_pose": float(res.lc2[0]),
        **extra,
    }
    print(json.dumps(out), flush=True)
    pair.close()
    fv.close()
    mv.close()
    eng.close()


if __name__ == "__main__":
    main()
