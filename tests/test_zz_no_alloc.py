"""GPU: the allocation-free launch (bhc_field_launch_noalloc) against the
ordinary resident launch (bhc_field_launch), both called through the C ABI.

Two calls in a row share the device's static scratch array, the second on a
smaller problem: each must give the ordinary launch's per-ray step and bounce
counts, error / warning words, step and contribution totals exactly, and its
TL to 0.01 dB (the FP32 field sums run in another order when the rays are
split over more passes). The first problem needs several scratch passes.
"""
import sys

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd.env import munk_env, c2_env
from parity import tl_db

pytestmark = pytest.mark.gpu


def _announce(name):
    print("\n[gpu-case]", name, flush=True)
    sys.stdout.flush()


def _launch(prob, ctx, noalloc):
    L = bh.lib()
    res = bh.Resident(prob, ctx, stats=True)
    try:
        f = L.bhc_field_launch_noalloc if noalloc else L.bhc_field_launch
        rc = f(ctx._h, res._h)
        assert rc == 0, L.bhc_last_error().decode()
        out = res.download()
        passes = ctx.last_phase()[2]
    finally:
        res.release()
    return out, passes


def test_noalloc_launch_matches_launch_twice():
    _announce("noalloc_two_calls")
    ctx = bh.Context(0)
    try:
        cases = [("munk_2001_hat", munk_env(nbeams=2001, nrz=101, nrr=501)),
                 ("seamount_401_gauss_raycen_smaller", c2_env(nbeams=401, nrz=31, nrr=201))]
        for k, (name, env) in enumerate(cases):
            prob = bh.Problem(env)
            a, pa = _launch(prob, ctx, False)
            b, pb = _launch(prob, ctx, True)
            print(name, "passes", pa, pb, "steps", a.total_steps, b.total_steps, "contributions", a.total_contrib,
                  b.total_contrib, flush=True)
            if k == 0:
                assert pb >= 3          # 1 M steps do not fit one 64 MiB pass
            assert (a.err, a.warn, a.err_count, a.warn_count) == (b.err, b.warn, b.err_count, b.warn_count)
            assert a.total_steps == b.total_steps and a.total_contrib == b.total_contrib
            assert np.array_equal(a.nsteps, b.nsteps)
            assert np.array_equal(a.ntop, b.ntop) and np.array_equal(a.nbot, b.nbot)
            ta, tb = tl_db(prob.scale_pressure(a.field)), tl_db(prob.scale_pressure(b.field))
            m = ta < 120
            assert m.any() and np.max(np.abs(ta - tb)[m]) < 0.01
    finally:
        ctx.close()


def test_noalloc_launch_rejects_other_run_types():
    _announce("noalloc_rejects_arrivals")
    ctx = bh.Context(0)
    try:
        prob = bh.Problem(munk_env(nbeams=21, nrz=3, nrr=5, run_type="AG"))
        res = bh.Resident(prob, ctx, max_narr=10)
        try:
            L = bh.lib()
            assert L.bhc_field_launch_noalloc(ctx._h, res._h) != 0
            assert b"TL run types" in L.bhc_last_error()
        finally:
            res.release()
    finally:
        ctx.close()
