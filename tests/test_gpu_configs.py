"""GPU parity of whole workload configurations against the oracle.

* the plain Munk TL field at its full size (2001 hat beams, 201 x 1001
  receivers, FP64);
* the batched sweep in FP32 against the FP32 oracle (the FP64 half is in
  test_gpu_sweep.py), including a PCHIP problem;
* a sweep launched slice by slice (bhc_sweep_launch_range) against the same
  sweep launched whole;
* the reference-arithmetic phasor (BHC_EXACT_PHASOR=1: complex FP64
  exponential per contribution, as src/influence.hpp:432-473);
* the 2D eigenray re-trace (bhc_ray_run2d_list, src/mode/eigen.cpp:24-48)
  against the oracle's re-trace of the same hits.

Each case prints its name before touching the GPU.
"""
import os
import sys

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd.env import munk_env, c2_env, sweep_envs
import oracle
from parity import compare_field, tl_db

pytestmark = pytest.mark.gpu


def _announce(name):
    print("\n[gpu-case]", name, flush=True)
    sys.stdout.flush()


def test_munk_full_size_against_oracle(gpu_ctx):
    _announce("munk_full_size_f64")
    env = munk_env()
    assert env.nrays == 2001 and env.field_shape == (1, 201, 1001)
    o = oracle.field_run(env)
    g = bh.field_run(env, ctx=gpu_ctx, stats=True)
    assert (g.err, g.warn) == (o.err, o.warn)
    ok, rep = compare_field(o, g, oracle.scale_pressure(env, o.field), g.problem.scale_pressure(g.field))
    print(rep, flush=True)
    assert ok, rep
    # FP64: every ray's step and bounce counts exact, not only 99.9 %
    assert rep["steps_exact"] == 1.0 and rep["bounces_exact"] == 1.0
    assert g.total_steps == o.total_steps and g.total_contrib == o.total_contrib


def test_sweep_f32_against_oracle(gpu_ctx):
    """FP32 batched sweep (the reference's BHC_USE_FLOATS arithmetic) against
    the FP32 oracle problem by problem: per-ray step and bounce counts exact,
    TL within 0.1 dB where TL < 120 dB."""
    _announce("sweep_f32_vs_oracle")
    envs = sweep_envs(n=8, scale=0.05, precision=32) + [
        munk_env(nbeams=151, nrz=21, nrr=41, ssp_type="P", precision=32)]
    kinds = sorted(set(e.ssp_type for e in envs))
    assert "P" in kinds and "C" in kinds, kinds
    probs = [bh.Problem(e) for e in envs]
    sw = bh.Sweep(probs, gpu_ctx, stats=True)
    sw.launch()
    worst = 0.0
    for i, (e, p) in enumerate(zip(envs, probs)):
        g = sw.download(i)
        o = oracle.field_run(e)
        assert (g.err, g.warn) == (o.err, o.warn), i
        assert np.array_equal(g.nsteps, o.nsteps), i
        assert np.array_equal(g.ntop, o.ntop) and np.array_equal(g.nbot, o.nbot), i
        ta, tb = tl_db(p.scale_pressure(g.field)), tl_db(oracle.scale_pressure(e, o.field))
        m = tb < 120
        if m.any():
            worst = max(worst, float(np.max(np.abs(ta - tb)[m])))
    print("sweep f32 worst TL difference %.4f dB" % worst, flush=True)
    assert worst <= 0.1
    sw.release()


def test_sweep_range_launch_matches_whole(gpu_ctx):
    """Launching a sweep slice by slice gives, problem by problem, the steps,
    errors and field of the whole launch; a slice leaves the other problems'
    fields untouched."""
    _announce("sweep_range_launch")
    envs = sweep_envs(n=9, scale=0.05, seed=91)
    probs = [bh.Problem(e) for e in envs]
    sw = bh.Sweep(probs, gpu_ctx, stats=True)
    sw.launch()
    whole = [sw.download(i) for i in range(len(probs))]
    total = sw.stats()
    steps = contrib = 0
    for c0, c1 in ((0, 4), (4, 7), (7, 9)):
        sw.launch(c0, c1)
        st = sw.stats()
        steps += st["steps"]
        contrib += st["contrib"]
    assert steps == total["steps"] and contrib == total["contrib"]
    for i, p in enumerate(probs):
        g = sw.download(i)
        assert (g.err, g.warn, g.total_steps) == (whole[i].err, whole[i].warn, whole[i].total_steps), i
        assert np.array_equal(g.nsteps, whole[i].nsteps), i
        ta, tb = tl_db(p.scale_pressure(g.field)), tl_db(p.scale_pressure(whole[i].field))
        m = tb < 120
        assert not m.any() or np.max(np.abs(ta - tb)[m]) < 0.01, i
    # a slice does not touch fields outside it
    before = sw.download(0).field.copy()
    sw.launch(5, 9)
    assert np.array_equal(sw.download(0).field, before)
    with pytest.raises(bh.BhcError):
        sw.launch(3, 3)
    with pytest.raises(bh.BhcError):
        sw.launch(0, len(probs) + 1)
    sw.release()


@pytest.mark.parametrize("name", ["munk_hat_cart", "seamount_gauss_raycen"])
def test_exact_phasor_against_oracle(name):
    """With BHC_EXACT_PHASOR=1 the scatter evaluates each contribution with
    the reference's complex FP64 exponential; the FP32 field then differs from
    the oracle's only by the order of the FP32 atomic sums."""
    _announce("exact_phasor_" + name)
    env = munk_env(nbeams=401, nrz=101, nrr=201) if name == "munk_hat_cart" else c2_env(nbeams=401, nrz=61, nrr=401)
    old = os.environ.get("BHC_EXACT_PHASOR")
    os.environ["BHC_EXACT_PHASOR"] = "1"
    try:
        ctx = bh.Context(0)
    finally:
        if old is None:
            del os.environ["BHC_EXACT_PHASOR"]
        else:
            os.environ["BHC_EXACT_PHASOR"] = old
    try:
        g = bh.field_run(env, ctx=ctx, stats=True)
    finally:
        ctx.close()
    o = oracle.field_run(env)
    assert g.total_contrib == o.total_contrib
    so, sg = oracle.scale_pressure(env, o.field), g.problem.scale_pressure(g.field)
    ok, rep = compare_field(o, g, so, sg)
    print(name, rep, flush=True)
    assert ok, rep
    # per receiver, |p| to the FP32 rounding of a sum of contributions
    m = tl_db(so) < 120
    rel = np.abs(sg - so)[m] / np.maximum(np.abs(so)[m], 1e-30)
    print(name, "median relative field difference %.2e" % float(np.median(rel)), flush=True)
    assert float(np.median(rel)) < 1e-5


@pytest.mark.parametrize("precision", [64, 32])
def test_eigen_retrace_2d_against_oracle(precision, gpu_ctx):
    """Eigenray run: the GPU's eigen hits re-traced on the GPU
    (bhc_ray_run2d_list, each ray stopped after the hit's step count as
    EigenModePostWorker does) equal the oracle's re-trace of the same hits:
    step counts, positions and bounce counts, every point."""
    _announce("eigen_retrace_2d_f%d" % precision)
    env = munk_env(nbeams=201, nrz=5, nrr=11, run_type="EG", precision=precision)
    g = bh.field_run(env, ctx=gpu_ctx, eigen_memsize=200000)
    n = int(g.neigen[0])
    assert n > 0
    hits = np.ctypeslib.as_array(g.hits)[:n]
    ia = hits["ialpha"].astype(np.int32)
    iz = hits["isz"].astype(np.int32)
    ns = hits["is_"].astype(np.int32)
    pts_g, nst_g, eg = bh.ray_run(env, ctx=gpu_ctx, max_points=4000, ialpha=ia, isz=iz, nsteps_in=ns)
    pts_o, nst_o, eo = oracle.ray_run(env, max_points=4000, ialpha=ia, isz=iz, nsteps_in=ns)
    assert eg == eo
    assert np.array_equal(nst_g, nst_o)
    assert np.all(nst_g >= 1)
    for k in range(n):
        m = nst_o[k]
        a, b = pts_g[k, :m], pts_o[k, :m]
        if precision == 64:
            assert np.array_equal(a["x"], b["x"]), k
        else:
            assert np.allclose(a["x"], b["x"], rtol=1e-5, atol=1e-3), k
        assert np.array_equal(a["NumTopBnc"], b["NumTopBnc"]) and np.array_equal(a["NumBotBnc"], b["NumBotBnc"]), k
