"""Throughput benchmark of the ray/beam-tracing hot path on one MI355X.

Workloads (one "step" = one pass of the hot path over one batch of input):

* sweep (default): the heterogeneous batched sweep -- 4096 independent 2D TL
  problems drawn from one seed (10-200 SSP knots, c-linear or PCHIP; flat or
  1-500-segment bathymetry, depth 100-5000 m; 25-1000 Hz; 500-20000 hat
  beams; 51-501 x 201-5001 receivers), FP64. All problems are resident in
  HBM; a step traces a slice of 512 of them in one persistent launch
  (bhc_sweep_launch_range), and steps cycle through the slices (8 steps = the
  whole sweep), so that any --steps / --warmup fits a few minutes. The plain
  per-environment GPU launch (one resident problem per launch) is timed on a
  sample of the same problems and reported beside it. Progress lines go to
  stderr.
* seamount: one range-dependent TL field -- 40-knot spline SSP, 500-segment
  bathymetry, 20001 ray-centred Gaussian beams, 501 x 5001 receivers, FP64.
* munk: one Munk-profile TL field, 2001 hat beams, 201 x 1001 receivers.
* c3: one full-3D Gaussian-beam TL field -- hexahedral SSP, 201 x 201
  bathymetry grid, 181 x 1001 beams, 36 bearings x 101 depths x 1001 ranges.
* surrogate: the surrogate-screened sweep -- an MLP ensemble trained on the
  modal coefficients of tracer fields of a 14-parameter family answers
  held-out environments where its members agree, the batched tracer where
  they do not (--sur-family global: every parameter over its full range;
  local: an uncertainty sweep around one nominal environment).

Inputs are resident in HBM before the timed region. Prints one JSON line:
fields/s (value), rays/s, the roofline of the dominant field kernel with the
other kernel and the whole path beside it (algorithmic FP64 operations
F = f_I R + f_S S + f_C C, f_* from the op-counting instantiation of the same
core, R, S, C counted by the kernels; the trace kernel does f_I R + f_S S, the
scatter f_C C; kernel times from HIP events around every launch) and the CPU
baseline (the oracle's thread pool on a bounded sample of rays of the same
workload).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 vector
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector


def single_workload(name, precision):
    import bellhopcuda_amd as bh
    if name == "seamount":
        return bh.c2_env(precision=precision), "seamount 2D range-dependent TL, 20001 'b' beams, 501x5001 receivers"
    if name == "munk":
        return bh.munk_env(precision=precision), "Munk 2D TL, 2001 'G' beams, 201x1001 receivers"
    raise SystemExit("unknown workload " + name)


def cpu_rays_per_s(envs, target_s, cores):
    """Oracle (plain-C restatement, thread pool over `cores` host threads) on
    chunks of rays spread over each environment's fan, cycling over `envs`,
    doubling the chunk until about `target_s` seconds were spent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    nchunks, per, done, t_total, rounds = 8, 8, 0, 0.0, 0
    while t_total < target_s and rounds < 12:
        for env in envs:
            n = env.nrays
            p = min(per, max(1, n // nchunks))
            t0 = time.time()
            for k in range(nchunks):
                b = max(0, min(n - p, int((k + 0.5) * n / nchunks) - p // 2))
                oracle.field_run(env, nthreads=cores, job_begin=b, job_end=b + p, stats=False)
                done += p
            t_total += time.time() - t0
            if t_total >= target_s:
                break
        per *= 2
        rounds += 1
    return done / t_total, done, t_total


def opcounts(L, _abi, probs, nsample):
    """Mean per-unit operation counts over `probs` (op-counting core)."""
    acc = [0.0, 0.0, 0.0]
    for p in probs:
        oc = (C.c_double * 6)()
        if L.bhc_opcount2d(C.byref(p.view), nsample, oc) != 0:
            raise SystemExit("op count failed")
        for k in range(3):
            acc[k] += oc[k]
    return [v / len(probs) for v in acc]


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` as measured by the committed PMC
    summary (FETCH_SIZE and WRITE_SIZE passes scaled by the copy-kernel
    calibration), or None if it holds no figure for that kernel."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_sweep4096_f64.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch "):
            name, value = line[len("traffic_bytes_per_launch "):].rsplit(" ", 1)
            if name == kernel:
                return float(value)
    return None


def kernel_roofline(precision, kernels, F_path, path_ms):
    """Roofline object for the dominant kernel (most time) of a launch.

    kernels: dicts with name, ms (summed over its launches), launches and flop
    (algorithmic FP ops of those launches: f_I R + f_S S for the trace kernel,
    which also forms each step's influence prelude, f_C C for the scatter).
    The whole path (trace + scatter) is reported beside it."""
    peak = FP64_PEAK_TFLOPS if precision == 64 else FP32_PEAK_TFLOPS
    out = []
    for k in kernels:
        ms = max(k["ms"], 1e-9)
        t = k["flop"] / (ms * 1e-3) / 1e12
        out.append(dict(k, ms_per_launch=ms / max(k["launches"], 1), flop_per_launch=k["flop"] / max(k["launches"], 1),
                        achieved=t, frac=t / peak))
    for k in out:
        t = pmc_traffic(k["name"])
        if t is not None:
            k["traffic_bytes_per_launch"] = t
            k["hbm_tb_per_s"] = t / (k["ms_per_launch"] * 1e-3) / 1e12
    dom = max(out, key=lambda k: k["ms"])
    path = F_path / (path_ms * 1e-3) / 1e12
    traffic = dom.get("traffic_bytes_per_launch")
    return {"bound": "valu", "pipe": "FP%d vector ALU (the tracer has no matrix work)" % precision,
            "kernel": dom["name"], "achieved": dom["achieved"], "peak": peak, "unit": "TFLOP/s",
            "frac": dom["frac"], "traffic": traffic,
            "traffic_source": "profiles/pmc_traffic_sweep4096_f64.md (rocprofv3 FETCH_SIZE / WRITE_SIZE x copy-kernel "
                              "calibration)" if traffic is not None else None,
            "launches": dom["launches"],
            "ms_per_launch": dom["ms_per_launch"], "flop_per_launch": dom["flop_per_launch"],
            "kernels": out,
            "path": {"kernels": "trace + scatter", "ms": path_ms, "flop": F_path, "achieved": path,
                     "frac": path / peak}}


def run_single(a):
    import bellhopcuda_amd as bh
    from bellhopcuda_amd import _abi
    L = bh.lib()
    env, desc = single_workload(a.workload, a.precision)
    prob = bh.Problem(env)
    f_I, f_S, f_C = opcounts(L, _abi, [prob], 40)
    ctx = bh.Context(0)
    res = bh.Resident(prob, ctx)
    for _ in range(a.warmup):
        res.launch()
        L.bhc_ctx_synchronize(ctx._h)
    kms, tms, sms = [], [], []
    L.bhc_ctx_synchronize(ctx._h)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res.launch()
        L.bhc_ctx_synchronize(ctx._h)
        kms.append(ctx.last_kernel_ms())
        tr, sc, npass = C.c_double(), C.c_double(), C.c_int()
        L.bhc_last_phase_ms(ctx._h, C.byref(tr), C.byref(sc), C.byref(npass))
        tms.append(tr.value)
        sms.append(sc.value)
    L.bhc_ctx_synchronize(ctx._h)
    wall = time.perf_counter() - t0
    out = res.download()
    fields_s = a.steps / wall
    R, S, Cn = prob.nrays, out.total_steps, out.total_contrib
    F = f_I * R + f_S * S + f_C * Cn
    k_ms = sum(kms) / len(kms)
    nl = len(tms) * npass.value
    rl = kernel_roofline(a.precision, [
        {"name": "k_trace2d", "ms": sum(tms), "launches": nl,
         "flop": (f_I * R + f_S * S) * len(tms)},
        {"name": "scatter (k_scatter_tile with k_bin_count / k_bin_scan / k_bin_fill, or k_scatter2d / "
                 "k_scatter2d_lanes, as autotuned)", "ms": sum(sms), "launches": nl, "flop": f_C * Cn * len(tms)},
    ], F, k_ms)
    rl.update({"f_I": f_I, "f_S": f_S, "f_C": f_C, "rays": R, "steps": S, "contributions": Cn})
    line = {
        "metric": "TL fields/s + rays/s on 1 MI355X",
        "value": fields_s, "unit": "TL fields/s", "rays_per_s": fields_s * R,
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == 64 else "f32", "data": "synthetic (seeded environment)",
        "config": {"workload": desc, "rays": R, "receivers": int(env.field_shape[1] * env.field_shape[2]),
                   "parallelism": "single GPU"},
        "roofline": rl,
        "errors": {"err": out.err, "warn": out.warn},
    }
    if not a.no_cpu:
        cores = os.cpu_count() or 1
        rs, done, t = cpu_rays_per_s([env], a.cpu_seconds, cores)
        line["cpu_baseline"] = {"value": rs / R, "unit": "TL fields/s", "rays_per_s": rs, "cores": cores,
                                "kind": "port",
                                "sample": "%d rays in chunks spread over the %d-ray fan, full receiver grid, %.1f s"
                                          % (done, R, t)}
    return line


def scatter3_name():
    """The Cartesian 3D scatter kernel the library launches for a TL run whose
    receiver tables fit LDS: BHC_SCATTER3_MAP 0 one wavefront per segment, 1
    one lane per segment, 2 (default) segment groups per wavefront."""
    m = os.environ.get("BHC_SCATTER3_MAP")
    try:
        k = 2 if m is None else min(max(int(m.strip()), 0), 2)
    except ValueError:
        k = 0   # the library reads the value with atoi
    return ("k_scatter3d", "k_scatter3d_lanes", "k_scatter3d_flat")[k]


def run_c3(a):
    """3D Gaussian-beam TL on the shelf-canyon environment: hexahedral SSP,
    201 x 201 bathymetry grid, 181 bearings x 1001 declinations, 36 bearings x
    101 depths x 1001 receiver ranges (bhc_field_run3d)."""
    import numpy as np
    import bellhopcuda_amd as bh
    from bellhopcuda_amd.env3d import c3_env
    L = bh.lib()
    env = c3_env(precision=a.precision)
    prob = bh.Problem3D(env)
    R = prob.nrays
    oc = (C.c_double * 6)()
    if L.bhc_opcount3d(C.byref(prob.view), 64, oc) != 0:
        raise SystemExit("3D op count failed")
    f_I, f_S, f_C = oc[0], oc[1], oc[2]
    ctx = bh.Context(0)
    for _ in range(a.warmup):
        bh.field_run3d(prob, ctx=ctx)
    kms, tms, sms, walls = [], [], [], []
    for _ in range(a.steps):
        t1 = time.perf_counter()
        out = bh.field_run3d(prob, ctx=ctx)
        walls.append(time.perf_counter() - t1)
        kms.append(ctx.last_kernel_ms())
        tr, sc, npass = C.c_double(), C.c_double(), C.c_int()
        L.bhc_last_phase_ms(ctx._h, C.byref(tr), C.byref(sc), C.byref(npass))
        tms.append(tr.value)
        sms.append(sc.value)
    k_ms = sum(kms) / len(kms)
    S, Cn = out.total_steps, out.total_contrib
    F = f_I * R + f_S * S + f_C * Cn
    nl = len(tms) * npass.value
    rl = kernel_roofline(a.precision, [
        {"name": "k_trace3d<%s>" % ("double" if a.precision == 64 else "float"), "ms": sum(tms), "launches": nl,
         "flop": (f_I * R + f_S * S) * len(tms)},
        {"name": "%s<%s, true>" % (scatter3_name(), "double" if a.precision == 64 else "float"), "ms": sum(sms),
         "launches": nl, "flop": f_C * Cn * len(tms)},
    ], F, k_ms)
    rl.update({"f_I": f_I, "f_S": f_S, "f_C": f_C, "rays": R, "steps": S, "contributions": Cn,
               "passes_per_field": npass.value})
    fields_s = 1e3 / k_ms
    line = {
        "metric": "TL fields/s + rays/s on 1 MI355X",
        "value": fields_s, "unit": "TL fields/s", "rays_per_s": fields_s * R,
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": k_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == 64 else "f32", "data": "synthetic (seeded 3D environment, seed 20260003)",
        "config": {"workload": "3D Gaussian-beam TL: 201x201 shelf-canyon bathymetry over 40x40 km, hexahedral "
                               "21x21x50 SSP, 200 Hz, 181 bearings x 1001 declinations, receivers 36 bearings x "
                               "101 depths x 1001 ranges to 20 km",
                   "rays": R, "receivers": int(np.prod(prob.field_shape)), "parallelism": "single GPU"},
        "timing": "value and ms_per_step from HIP events around the field kernels (inputs resident in HBM); "
                  "wall_ms_per_step adds the per-call upload, allocation and download",
        "wall_ms_per_step": 1e3 * sum(walls) / len(walls),
        "roofline": rl,
        "errors": {"err": out.err, "warn": out.warn},
    }
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        cores = os.cpu_count() or 1
        nch, per, done, t_total = 8, 16, 0, 0.0
        while t_total < a.cpu_seconds and per <= R // nch:
            t0 = time.time()
            for k in range(nch):
                b = max(0, min(R - per, int((k + 0.5) * R / nch) - per // 2))
                oracle.field_run3d(env, nthreads=cores, job_begin=b, job_end=b + per)
                done += per
            t_total += time.time() - t0
            per *= 2
        rs = done / t_total
        line["cpu_baseline"] = {"value": rs / R, "unit": "TL fields/s", "rays_per_s": rs, "cores": cores,
                                "kind": "port",
                                "sample": "%d rays in chunks spread over the %d-ray fan, full receiver grid, %.1f s"
                                          % (done, R, t_total)}
    return line


def log(msg):
    """Progress line on stderr (the JSON result line is the only stdout)."""
    print("[bench] %s" % msg, file=sys.stderr, flush=True)


def cpu_fields_per_s(envs, target_s, cores):
    """CPU baseline of a sweep: the oracle (plain-C restatement, thread pool of
    `cores` host threads pulling rays from one atomic counter) on each sampled
    problem's rays in chunks spread over its fan. Each problem's full time is
    extrapolated from its own sample (time x rays / sampled rays), and
    fields/s = problems / sum of their estimated times, so that problems of
    different sizes are weighted by their own cost."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    per_env = target_s / len(envs)
    est, done, spent = 0.0, 0, 0.0
    for env in envs:
        n = env.nrays
        k, t, per = 0, 0.0, 8
        while t < per_env and k < n:
            p = min(per, n - k)
            nch = max(1, min(8, p))
            q = max(1, p // nch)
            t0 = time.time()
            for c in range(nch):
                b = max(0, min(n - q, int((c + 0.5) * n / nch) - q // 2))
                oracle.field_run(env, nthreads=cores, job_begin=b, job_end=b + q, stats=False)
            t += time.time() - t0
            k += nch * q
            per *= 2
        est += t * n / k
        done += k
        spent += t
    return len(envs) / est, done, spent


def run_sweep(a):
    import bellhopcuda_amd as bh
    from bellhopcuda_amd import _abi
    L = bh.lib()
    envs = bh.sweep_envs(n=a.sweep_n, precision=a.precision, scale=a.sweep_scale)
    t0 = time.perf_counter()
    probs = [bh.Problem(e) for e in envs]
    t_pre = time.perf_counter() - t0
    log("preprocessed %d problems in %.1f s" % (len(probs), t_pre))
    ctx = bh.Context(0)
    if a.copy_calibration_mb > 0:
        cms = C.c_double()
        if L.bhc_copy_calibrate(ctx._h, a.copy_calibration_mb << 20, C.byref(cms)) != 0:
            raise SystemExit("copy calibration failed: " + L.bhc_last_error().decode())
        log("copy calibration: %d MiB in %.3f ms" % (a.copy_calibration_mb, cms.value))
    nwarm = min(len(probs), 64)
    warm = bh.Sweep(probs[:nwarm], ctx)
    warm.launch()
    warm.release()
    sw = bh.Sweep(probs, ctx)
    if a.sweep_config:
        b, l, m = (int(v) for v in a.sweep_config.split(","))
        sw.set_launch(b, l, m)
        tuned = {"scatter_blocks": b, "lds_bytes": l, "scatter_lanes": m, "source": "--sweep-config"}
        log("launch configuration given: %s" % json.dumps(tuned))
    else:
        log("uploaded; autotuning on the first 64 problems")
        tuned = sw.autotune()
        log("autotuned: %s" % json.dumps(tuned))
    # kernel-by-kernel passes, so that HIP events time the trace and the scatter
    # kernels of every pass separately (the graph-replayed passes are the same
    # kernels; each pass ends in a host synchronisation either way)
    sw.set_graph(False)
    nchunk = max(1, min(a.sweep_chunk, len(probs)))
    chunks = [(i, min(i + nchunk, len(probs))) for i in range(0, len(probs), nchunk)]
    rays = [p.nrays for p in probs]
    for w in range(a.warmup):
        c0, c1 = chunks[w % len(chunks)]
        sw.launch(c0, c1)
    L.bhc_ctx_synchronize(ctx._h)
    log("warm-up done (%d launches of %d problems)" % (a.warmup, nchunk))
    tr_ms, sc_ms, npasses, nprob_done, R, S, Cn = [], [], 0, 0, 0, 0, 0
    errs = []
    t0 = time.perf_counter()
    for k in range(a.steps):
        c0, c1 = chunks[k % len(chunks)]
        sw.launch(c0, c1)
        st = sw.stats()
        tr_ms.append(st["trace_ms"])
        sc_ms.append(st["scatter_ms"])
        npasses += st["passes"]
        S += st["steps"]
        Cn += st["contrib"]
        R += sum(rays[c0:c1])
        nprob_done += c1 - c0
        if k < len(chunks):
            errs += [sw.download(i) for i in range(c0, c1, max(1, (c1 - c0) // 8))]
        log("step %d: problems [%d, %d) in %.0f ms (%d passes)" % (k, c0, c1, st["trace_ms"] + st["scatter_ms"],
                                                                   st["passes"]))
    L.bhc_ctx_synchronize(ctx._h)
    wall = time.perf_counter() - t0
    nerr = sum(1 for r in errs if r.err)
    sw.release()
    fields_s = nprob_done / wall

    # plain per-environment GPU launches on a sample of the same problems
    step = max(1, len(probs) // max(a.plain_sample, 1))
    sample = probs[::step][:a.plain_sample]
    t_plain = 0.0
    modes = []
    for p in sample:
        r = bh.Resident(p, ctx)
        modes.append(r.autotune())
        L.bhc_ctx_synchronize(ctx._h)
        t1 = time.perf_counter()
        r.launch()
        L.bhc_ctx_synchronize(ctx._h)
        t_plain += time.perf_counter() - t1
        r.release()
    plain_fields_s = len(sample) / t_plain if sample else None
    if sample:
        log("plain per-environment launches: %.2f fields/s on %d problems" % (plain_fields_s, len(sample)))

    f_I, f_S, f_C = opcounts(L, _abi, probs[::max(1, len(probs) // 16)][:16], 16)
    F = f_I * R + f_S * S + f_C * Cn
    k_ms = (sum(tr_ms) + sum(sc_ms)) / len(tr_ms)
    real_t = "double" if a.precision == 64 else "float"
    if tuned["scatter_lanes"] == 3:
        # fused: one kernel traces and evaluates the receivers (f_I R + f_S S + f_C C)
        kern = [{"name": "k_trace_batch<%s, fused>" % real_t, "ms": sum(tr_ms) + sum(sc_ms), "launches": npasses,
                 "flop": F}]
    else:
        kern = [
            {"name": "k_trace_batch<%s>" % real_t, "ms": sum(tr_ms), "launches": npasses, "flop": f_I * R + f_S * S},
            {"name": {0: "k_scatter_batch<%s, true>", 1: "k_scatter_lanes<%s, true>", 2: "k_scatter_flat<%s, true>"}[
                tuned["scatter_lanes"]] % real_t, "ms": sum(sc_ms), "launches": npasses, "flop": f_C * Cn}]
    rl = kernel_roofline(a.precision, kern, F, k_ms * a.steps)
    rl.update({"passes": npasses, "f_I": f_I, "f_S": f_S, "f_C": f_C, "rays": R, "steps": S,
               "contributions": Cn})
    line = {
        "metric": "TL fields/s + rays/s on 1 MI355X, batched/surrogate vs plain GPU",
        "value": fields_s, "unit": "TL fields/s", "rays_per_s": R / wall,
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == 64 else "f32", "data": "synthetic (seeded sweep, seed 20260004)",
        "config": {"workload": "batched sweep of %d independent 2D TL problems (10-200 knots c-linear/PCHIP, "
                               "flat or 1-500-segment bathymetry, 25-1000 Hz, 500-20000 hat beams, "
                               "51-501 x 201-5001 receivers)%s, %s; all problems resident in HBM, one step = one "
                               "persistent launch over %d of them (steps cycle through the sweep, %d steps = the "
                               "whole sweep)" % (
                                   len(probs), "" if a.sweep_scale == 1.0 else ", beams/grids scaled %.3g" % a.sweep_scale,
                                   "f64" if a.precision == 64 else "f32", nchunk, len(chunks)),
                   "problems": len(probs), "problems_per_step": nchunk, "problems_timed": nprob_done,
                   "rays_timed": R, "parallelism": "single GPU"},
        "plain_per_env": {"value": plain_fields_s, "unit": "TL fields/s",
                          "sample": "%d of the %d problems, one resident launch each with its autotuned scatter "
                                    "mapping (tiles / wavefront / lane: %d / %d / %d)" % (
                                        len(sample), len(probs), modes.count(1), modes.count(2), modes.count(3))},
        "batched_over_plain": fields_s / plain_fields_s if plain_fields_s else None,
        "autotuned": tuned,
        "roofline": rl,
        "errors": {"problems_with_errors_in_sample": nerr, "sampled": len(errs)},
        "host_preprocess_s": t_pre,
    }
    if not a.no_cpu:
        cores = os.cpu_count() or 1
        cs = envs[::max(1, len(envs) // 16)][:16]
        cf, done, t = cpu_fields_per_s(cs, a.cpu_seconds, cores)
        line["cpu_baseline"] = {"value": cf, "unit": "TL fields/s", "cores": cores, "kind": "port",
                                "sample": "%d rays of 16 of the sweep's problems (chunks spread over each fan, full "
                                          "receiver grids), %.1f s; each problem's time extrapolated from its own "
                                          "rays, fields/s = problems / summed estimated times" % (done, t)}
        log("cpu baseline: %.4f fields/s" % cf)
    return line


def run_surrogate(a):
    """Surrogate-screened sweep: fit the modal basis to tracer fields of the
    environment family and train the ensemble on their coefficients, then
    answer held-out environments with the ensemble's field where its members
    agree and with the batched tracer where they do not."""
    import numpy as np
    import bellhopcuda_amd as bh
    from bellhopcuda_amd import surrogate as sg
    ctx = bh.Context(0)
    fam = a.sur_family
    th_tr = sg.draw_params(a.sur_train, 20260005, fam)
    th_ev = sg.draw_params(a.sur_eval, 20260006, fam)
    t0 = time.perf_counter()
    f_tr, e_tr, t_tr = sg.trace_fields(th_tr, ctx)
    t_train_fields = time.perf_counter() - t0
    log("traced %d training fields in %.1f s" % (len(f_tr), t_train_fields))
    t0 = time.perf_counter()
    sur = sg.Surrogate(th_tr, f_tr, fam, k=a.sur_modes, members=4, seed=11, max_batch=a.sur_batch)
    t_basis = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = sur.train(a.sur_steps, a.sur_batch, a.sur_lr)
    t_train = time.perf_counter() - t0
    train_ms = sur.train_ms
    step_flops = sur.mlp.step_flops(a.sur_batch)
    log("basis %.1f s, training %.1f s (%d steps)" % (t_basis, t_train, a.sur_steps))
    # ground truth of the held-out environments (batched tracer)
    f_ev, e_ev, t_ev = sg.trace_fields(th_ev, ctx)
    truth = np.stack(f_ev)
    rz, rr = e_ev[0].Rz, e_ev[0].Rr_km
    # plain per-environment tracer on a sample
    t_plain, ns = 0.0, 0
    for e in e_ev[::max(1, len(e_ev) // 32)][:32]:
        r = bh.Resident(bh.Problem(e), ctx)
        r.launch()
        bh.lib().bhc_ctx_synchronize(ctx._h)
        t1 = time.perf_counter()
        r.launch()
        bh.lib().bhc_ctx_synchronize(ctx._h)
        t_plain += time.perf_counter() - t1
        ns += 1
        r.release()
    plain_fields_s = ns / t_plain
    # the screened answer (timed): prediction and screen, then the tracer for
    # the environments the screen sends back
    final, fallback, rms_spread, t_pred, t_fb = sg.screened_answer(sur, th_ev, e_ev, ctx, a.sur_spread)
    rms, p95, per = sg.field_errors(final, truth, th_ev, rz, rr)
    acc = ~fallback
    rms_acc, p95_acc, _ = (sg.field_errors(final[acc], truth[acc], th_ev[acc], rz, rr) if acc.any()
                           else (0.0, 0.0, None))
    # diagnostics outside the timed region: the network's field everywhere,
    # in sample, and the mean training field as the method without a network
    net = sur.predict(th_ev).mean(0)
    net_rms, net_p95, net_per = sg.field_errors(net, truth, th_ev, rz, rr)
    base = np.broadcast_to(np.clip(sur.basis.ref.reshape(truth.shape[1:]), sg.TL_MIN, sg.TL_CAP), truth.shape)
    base_rms, base_p95, base_per = sg.field_errors(base, truth, th_ev, rz, rr)
    k_in = list(range(0, len(th_tr), max(1, len(th_tr) // 100)))[:100]
    pin = sur.predict(th_tr[k_in]).mean(0)
    in_rms, _, _ = sg.field_errors(pin, np.stack([f_tr[i] for i in k_in]), th_tr[k_in], rz, rr)
    spread_err = float(np.corrcoef(rms_spread, net_per)[0, 1]) if len(net_per) > 2 else None
    screened_fields_s = len(th_ev) / (t_pred + t_fb)
    achieved = step_flops * a.sur_steps / (train_ms * 1e-3) / 1e12
    peak = 2500.0
    sur.close()
    return {
        "metric": "TL fields/s on 1 MI355X, surrogate-screened sweep vs plain GPU",
        "value": screened_fields_s, "unit": "TL fields/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * (t_pred + t_fb), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 (network), f64 (tracer)",
        "data": "synthetic (seeded environment family '%s', seeds 20260005 / 20260006)" % fam,
        "config": {"workload": "surrogate-screened sweep, family '%s': %d training fields, %d held-out "
                               "environments, 51 x 201 receivers, 1001 'B' beams, incoherent TL; %d modes, "
                               "4 x MLP(16 inputs + 24 Fourier pairs, 6 x 256 GELU, %d outputs)"
                               % (fam, len(th_tr), len(th_ev), a.sur_modes, a.sur_modes),
                   "parallelism": "single GPU"},
        "plain_per_env": {"value": plain_fields_s, "unit": "TL fields/s"},
        "screened_over_plain": screened_fields_s / plain_fields_s,
        "fallback_fraction": float(np.mean(fallback)),
        "spread_threshold_db": a.sur_spread,
        "error_db": {"note": "water-column receivers with tracer TL below %g dB" % sg.TL_CAP,
                     "rms": rms, "p95": p95, "accepted_rms": rms_acc, "accepted_p95": p95_acc,
                     "network_only_rms": net_rms, "network_only_p95": net_p95,
                     "network_env_frac_rms_below_1p5": float(np.mean(net_per <= 1.5)),
                     "network_in_sample_rms": in_rms,
                     "mean_field_baseline_rms": base_rms, "mean_field_baseline_p95": base_p95,
                     "median_rms_spread": float(np.median(rms_spread)),
                     "corr_spread_vs_network_error": spread_err},
        "training": {"fields": len(th_tr), "modes": a.sur_modes, "steps": a.sur_steps, "batch": a.sur_batch,
                     "final_loss": [float(v) for v in loss], "seconds": t_train, "basis_seconds": t_basis,
                     "tracer_seconds_for_training_fields": t_train_fields},
        "roofline": {"bound": "mfma", "kernel": "BF16 MLP training step (forward + backward products)",
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": None, "step_ms": train_ms / a.sur_steps, "flop_per_step": step_flops},
        "timing_s": {"predict_and_screen": t_pred, "fallback_trace": t_fb, "eval_truth_trace": t_ev},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sweep", choices=["sweep", "seamount", "munk", "surrogate", "c3"])
    ap.add_argument("--sur-train", type=int, default=20000)
    ap.add_argument("--sur-eval", type=int, default=1000)
    ap.add_argument("--sur-steps", type=int, default=5000)
    ap.add_argument("--sur-batch", type=int, default=2048)
    ap.add_argument("--sur-lr", type=float, default=1e-3)
    ap.add_argument("--sur-modes", type=int, default=128)
    ap.add_argument("--sur-spread", type=float, default=1.0)
    ap.add_argument("--sur-family", default="global", choices=["global", "local"])
    ap.add_argument("--precision", type=int, default=64)
    ap.add_argument("--sweep-n", type=int, default=4096)
    ap.add_argument("--sweep-scale", type=float, default=1.0)
    ap.add_argument("--sweep-chunk", type=int, default=512,
                    help="problems per step (one persistent launch each); steps cycle through the sweep")
    ap.add_argument("--plain-sample", type=int, default=64, help="problems timed as plain launches (0: none)")
    ap.add_argument("--sweep-config", default="",
                    help="scatter_blocks,lds_bytes,mapping: use this launch configuration instead of autotuning")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="first run one launch of the library's plain copy kernel over this many MiB (the "
                         "known-size kernel for scaling profiler byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the tracer has no multi-GPU path (problems and ray ranges would be "
              "sharded per GPU with no exchange). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if a.workload == "sweep":
        line = run_sweep(a)
    elif a.workload == "surrogate":
        line = run_surrogate(a)
    elif a.workload == "c3":
        line = run_c3(a)
    else:
        line = run_single(a)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
