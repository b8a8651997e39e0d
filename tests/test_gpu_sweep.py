"""GPU: batched sweeps (bhc_sweep_*) against single-problem runs and the oracle.

A sweep mixes SSP types, boundaries, frequencies, fans and receiver grids; each
problem's rays must take exactly the steps of its single-problem run, and its
TL field must agree (the FP32 atomic accumulation order differs, nothing else).
Suspension (segment buffer full -> save rays -> scatter -> resume) is forced
with a small segment buffer and must not change any result.
"""
import os
import sys

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd.env import munk_env, c2_env, sweep_envs
import oracle
from parity import tl_db

pytestmark = pytest.mark.gpu


def _announce(name):
    print("\n[gpu-case]", name, flush=True)
    sys.stdout.flush()


def _mixed(precision):
    envs = sweep_envs(n=10, scale=0.05, precision=precision)
    envs += [munk_env(nbeams=201, nrz=41, nrr=81, precision=precision),
             munk_env(nbeams=151, nrz=31, nrr=61, ssp_type="S", precision=precision),
             munk_env(nbeams=151, nrz=31, nrr=61, ssp_type="N", run_type="IG", precision=precision),
             c2_env(nbeams=201, nrz=31, nrr=101, precision=precision)]
    return envs


def _tl_close(a, b, tol):
    ta, tb = tl_db(a), tl_db(b)
    m = ta < 120
    return (np.max(np.abs(ta - tb)[m]) if m.any() else 0.0) <= tol


@pytest.mark.parametrize("precision", [64, 32])
def test_sweep_matches_single_runs(precision, gpu_ctx):
    _announce("sweep_mixed_f%d" % precision)
    envs = _mixed(precision)
    probs = [bh.Problem(e) for e in envs]
    sw = bh.Sweep(probs, gpu_ctx, stats=True)
    sw.launch()
    st = sw.stats()
    total = 0
    for i, p in enumerate(probs):
        g = sw.download(i)
        s = bh.field_run(p, ctx=gpu_ctx, stats=True)
        assert (g.err, g.warn) == (s.err, s.warn), i
        assert np.array_equal(g.nsteps, s.nsteps), i
        assert np.array_equal(g.ntop, s.ntop) and np.array_equal(g.nbot, s.nbot), i
        if precision == 64:
            assert np.array_equal(g.xend, s.xend), i
        assert g.total_steps == s.total_steps, i
        total += s.total_contrib
        assert _tl_close(p.scale_pressure(g.field), p.scale_pressure(s.field), 0.01), i
    assert st["contrib"] == total
    sw.release()


def test_sweep_against_oracle(gpu_ctx):
    _announce("sweep_vs_oracle")
    envs = sweep_envs(n=4, scale=0.04) + [munk_env(nbeams=101, nrz=21, nrr=41, ssp_type="P")]
    probs = [bh.Problem(e) for e in envs]
    sw = bh.Sweep(probs, gpu_ctx, stats=True)
    sw.launch()
    for i, (e, p) in enumerate(zip(envs, probs)):
        g = sw.download(i)
        o = oracle.field_run(e)
        assert np.array_equal(g.nsteps, o.nsteps), i
        assert np.array_equal(g.xend, o.xend), i
        assert _tl_close(p.scale_pressure(g.field), oracle.scale_pressure(e, o.field), 0.1), i
    sw.release()


def test_sweep_suspend_and_resume():
    """A 3 GiB segment buffer holds ~10.9 M FP64 segments; the suspension
    margin (one poll interval of every wavefront) leaves room for ~6.8 M, so
    24 Munk fields with 1001 receiver ranges (~23 M steps, most of them
    crossing a receiver range and so handed to the scatter) need at least two
    passes."""
    _announce("sweep_suspend")
    old = os.environ.get("BHC_SEG_BUFFER_MB")
    os.environ["BHC_SEG_BUFFER_MB"] = "3000"
    try:
        small = bh.Context(0)
    finally:
        if old is None:
            del os.environ["BHC_SEG_BUFFER_MB"]
        else:
            os.environ["BHC_SEG_BUFFER_MB"] = old
    big = bh.Context(0)
    envs = [munk_env(nbeams=2001 - 10 * k, nrz=51, nrr=1001) for k in range(24)]
    probs = [bh.Problem(e) for e in envs]
    a = bh.Sweep(probs, small, stats=True)
    a.launch()
    sa = a.stats()
    b = bh.Sweep(probs, big, stats=True)
    b.launch()
    sb = b.stats()
    print("passes", sa["passes"], sb["passes"], flush=True)
    assert sa["passes"] >= 2 and sb["passes"] == 1
    assert sa["steps"] == sb["steps"] and sa["contrib"] == sb["contrib"]
    for i, p in enumerate(probs):
        x, y = a.download(i), b.download(i)
        assert np.array_equal(x.nsteps, y.nsteps) and np.array_equal(x.xend, y.xend), i
        assert _tl_close(p.scale_pressure(x.field), p.scale_pressure(y.field), 0.01), i
    a.release()
    b.release()
    small.close()
    big.close()


def test_sweep_graph_replays_match_eager(gpu_ctx):
    """Each pass is replayed from a captured HIP graph: three replays, with
    eager work and synchronisations in between, must each give the eager
    (kernel-by-kernel) result."""
    _announce("sweep_graph_replay")
    envs = sweep_envs(n=6, scale=0.05, seed=77) + [munk_env(nbeams=101, nrz=21, nrr=41)]
    probs = [bh.Problem(e) for e in envs]
    sw = bh.Sweep(probs, gpu_ctx, stats=True)
    sw.set_graph(False)
    sw.launch()
    ref = [sw.download(i) for i in range(len(probs))]
    ref_st = sw.stats()
    tuned = sw.autotune()
    print("autotuned", tuned, flush=True)
    sw.set_graph(True)
    for rep in range(3):
        sw.launch()
        st = sw.stats()
        assert st["steps"] == ref_st["steps"] and st["contrib"] == ref_st["contrib"], rep
        for i, p in enumerate(probs):
            g = sw.download(i)
            assert np.array_equal(g.nsteps, ref[i].nsteps), (rep, i)
            assert _tl_close(p.scale_pressure(g.field), p.scale_pressure(ref[i].field), 0.01), (rep, i)
        bh.field_run(probs[rep], ctx=gpu_ctx)   # eager work between replays
    sw.release()


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("mapping", [0, 1, 2, 3])
def test_sweep_scatter_mappings(mapping, precision, gpu_ctx):
    """Every scatter mapping of the batched passes (a wavefront per segment, a
    lane per segment, 64 segments' receivers flattened over a wavefront, and
    the fused trace that evaluates each step's receivers itself) gives the
    single-problem runs' contribution counts and TL; the mix holds Cartesian
    hat / Gaussian and ray-centred Gaussian problems."""
    _announce("sweep_scatter_mapping_%d_f%d" % (mapping, precision))
    envs = _mixed(precision)
    probs = [bh.Problem(e) for e in envs]
    sw = bh.Sweep(probs, gpu_ctx, stats=True)
    sw.set_scatter(mapping)
    sw.launch()
    st = sw.stats()
    total = 0
    for i, p in enumerate(probs):
        g = sw.download(i)
        s = bh.field_run(p, ctx=gpu_ctx)
        total += s.total_contrib
        assert _tl_close(p.scale_pressure(g.field), p.scale_pressure(s.field), 0.01), i
    assert st["contrib"] == total
    sw.release()
