"""Throughput benchmark of the 2D ray/beam-tracing hot path on one MI355X.

Default workload (one "step" = one complete TL field): the range-dependent
seamount problem -- 40-knot spline SSP, 500-segment bathymetry, acousto-elastic
halfspace, source 50 m, 100 Hz, 20001 ray-centred Gaussian beams over +-60 deg,
501 x 5001 receivers to 200 km, FP64. Inputs are resident in HBM before the
timed region; each step zeroes the field and traces + scatters every ray.

Prints one JSON line: fields/s (value), rays/s, the FP64 roofline of the field
kernels (algorithmic operations F = f_I R + f_S S + f_C C, with f_* measured by
the op-counting instantiation of the same core and R, S, C counted by the
kernels, divided by the kernels' time from HIP events), and the CPU baseline
(the oracle's std-thread pool on a bounded ray sample of the same workload).
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


def workload(name, precision):
    import bellhopcuda_amd as bh
    if name == "seamount":
        return bh.c2_env(precision=precision), "seamount 2D range-dependent TL, 20001 'b' beams, 501x5001 receivers"
    if name == "munk":
        return bh.munk_env(precision=precision), "Munk 2D TL, 2001 'G' beams, 201x1001 receivers"
    raise SystemExit("unknown workload " + name)


def cpu_baseline(env, nrays_total, target_s):
    """Oracle (plain-C restatement, std::thread-style pool over all host cores)
    on chunks of rays spread over the fan; returns rays/s and TL fields/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cores = os.cpu_count() or 1
    nchunks = 8
    per = 8
    done = 0
    t_total = 0.0
    while t_total < target_s and per <= nrays_total // nchunks:
        t0 = time.time()
        for k in range(nchunks):
            b = max(0, int((k + 0.5) * nrays_total / nchunks) - per // 2)
            oracle.field_run(env, nthreads=cores, job_begin=b, job_end=b + per, stats=False)
            done += per
        t_total += time.time() - t0
        per *= 2
    rays_s = done / t_total
    return {"value": rays_s / nrays_total, "unit": "TL fields/s", "rays_per_s": rays_s, "cores": cores,
            "kind": "port",
            "sample": "%d rays in %d chunks spread over the %d-ray fan, full receiver grid, %.1f s" % (
                done, nchunks, nrays_total, t_total)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="seamount")
    ap.add_argument("--precision", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (the reference kernel for "
                         "scaling profiler byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the tracer has no multi-GPU path (problems and ray ranges would be "
              "sharded per GPU with no exchange; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import bellhopcuda_amd as bh
    from bellhopcuda_amd import _abi
    L = bh.lib()
    L.bhc_last_phase_ms.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.bhc_opcount2d.argtypes = [C.POINTER(_abi.Problem2D), C.c_int, C.POINTER(C.c_double)]

    if a.copy_calibration_mb > 0:
        import torch
        src = torch.ones(a.copy_calibration_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    env, desc = workload(a.workload, a.precision)
    prob = bh.Problem(env)
    # per-unit algorithmic operation counts (same core, op-counting scalar type)
    oc = (C.c_double * 6)()
    if L.bhc_opcount2d(C.byref(prob.view), 40, oc) != 0:
        raise SystemExit("op count failed")
    f_I, f_S, f_C = oc[0], oc[1], oc[2]

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
    ms_per_step = 1e3 * wall / a.steps
    fields_s = a.steps / wall
    R, S, Cn = prob.nrays, out.total_steps, out.total_contrib
    F = f_I * R + f_S * S + f_C * Cn
    k_ms = sum(kms) / len(kms)
    peak = FP64_PEAK_TFLOPS if a.precision == 64 else FP32_PEAK_TFLOPS
    achieved = F / (k_ms * 1e-3) / 1e12
    line = {
        "metric": "TL fields/s + rays/s on 1 MI355X",
        "value": fields_s,
        "unit": "TL fields/s",
        "rays_per_s": fields_s * R,
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == 64 else "f32",
        "data": "synthetic (seeded environment)",
        "config": {"workload": desc, "rays": R, "receivers": int(env.field_shape[1] * env.field_shape[2]),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "pipe": "FP64 vector ALU (the tracer has no matrix work)",
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel_ms": k_ms, "trace_ms": sum(tms) / len(tms), "scatter_ms": sum(sms) / len(sms),
                     "flop_per_field": F, "f_I": f_I, "f_S": f_S, "f_C": f_C,
                     "rays": R, "steps": S, "contributions": Cn},
        "errors": {"err": out.err, "warn": out.warn},
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(env, R, a.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
