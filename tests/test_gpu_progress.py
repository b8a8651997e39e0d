"""GPU: progress counters and non-blocking runs at the boundary.

The reference counts finished rays (completedRayCount / totalJobs, read by
get_percent_progress) and, in non-blocking mode, calls completedCallback once
the last worker ends (src/mode/fieldimpl.cpp.in:43-71). Here: bhc_progress
read while a non-blocking run is in flight rises monotonically to the run's
ray count, the callback fires once, other calls on the busy context fail
cleanly, and the result equals the blocking run's.
"""
import sys
import time

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd.env import munk_env, c2_env, sweep_envs
from parity import tl_db

pytestmark = pytest.mark.gpu


def _announce(name):
    print("\n[gpu-case]", name, flush=True)
    sys.stdout.flush()


def test_async_field_progress_and_callback(gpu_ctx):
    _announce("async_field_2d")
    env = c2_env()                       # 20001 ray-centred beams, 501 x 5001 receivers
    prob = bh.Problem(env)
    ref = bh.field_run(prob, ctx=gpu_ctx, stats=True)
    assert gpu_ctx.progress() == (prob.nrays, prob.nrays)
    ctx = bh.Context(0)
    try:
        run = bh.AsyncRun(prob, ctx, stats=True)
        if ctx.busy():
            with pytest.raises(bh.BhcError):
                bh.field_run(munk_env(nbeams=11, nrz=3, nrr=3), ctx=ctx)
        samples = []
        t0 = time.time()
        while ctx.busy():
            samples.append(ctx.progress())
            assert time.time() - t0 < 120.0, "non-blocking run did not finish"
            time.sleep(0.0002)
        res = run.wait()
        samples.append(ctx.progress())
    finally:
        ctx.close()
    done = [d for d, t in samples]
    print("progress samples %d, distinct %d, first %s, last %s" % (len(samples), len(set(done)), samples[0],
                                                                     samples[-1]), flush=True)
    assert all(t == prob.nrays for d, t in samples)
    assert all(b >= a for a, b in zip(done, done[1:]))
    assert samples[-1] == (prob.nrays, prob.nrays)
    assert any(0 < d < prob.nrays for d in done), "no intermediate progress was seen"
    assert run.calls == 1
    assert (res.err, res.warn, res.total_steps, res.total_contrib) == (ref.err, ref.warn, ref.total_steps,
                                                                     ref.total_contrib)
    assert np.array_equal(res.nsteps, ref.nsteps)
    ta, tb = tl_db(prob.scale_pressure(res.field)), tl_db(prob.scale_pressure(ref.field))
    m = tb < 120
    assert np.max(np.abs(ta - tb)[m]) < 0.01


def test_async_runs_back_to_back_and_3d(gpu_ctx):
    """Several non-blocking runs in a row on one context (each waits for the
    previous), a 3D one among them, each equal to its blocking run."""
    from bellhopcuda_amd.env3d import munk_env3d
    _announce("async_back_to_back")
    p2 = bh.Problem(munk_env(nbeams=301, nrz=51, nrr=201))
    p3 = bh.Problem3D(munk_env3d(nalpha=41, nbeta=5, ntheta=2, nrz=11, nrr=41))
    r2 = bh.field_run(p2, ctx=gpu_ctx)
    r3 = bh.field_run3d(p3, ctx=gpu_ctx)
    ctx = bh.Context(0)
    try:
        for k in range(3):
            prob, ref = (p3, r3) if k == 1 else (p2, r2)
            flag = []
            run = bh.AsyncRun(prob, ctx, on_done=lambda: flag.append(1))
            res = run.wait()
            assert flag == [1] and run.calls == 1
            assert ctx.progress() == (prob.nrays, prob.nrays)
            assert res.total_steps == ref.total_steps and res.total_contrib == ref.total_contrib, k
            ta, tb = tl_db(prob.scale_pressure(res.field)), tl_db(prob.scale_pressure(ref.field))
            m = tb < 120
            assert m.any() and np.max(np.abs(ta - tb)[m]) < 0.01, k
    finally:
        ctx.close()


def test_sweep_progress_counts_every_ray(gpu_ctx):
    _announce("sweep_progress")
    probs = [bh.Problem(e) for e in sweep_envs(n=6, scale=0.05, seed=5)]
    sw = bh.Sweep(probs, gpu_ctx)
    sw.launch()
    total = sum(p.nrays for p in probs)
    assert gpu_ctx.progress() == (total, total)
    sw.launch(2, 5)
    part = sum(p.nrays for p in probs[2:5])
    assert gpu_ctx.progress() == (part, part)
    sw.release()
