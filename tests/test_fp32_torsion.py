"""FP32 3D fields: one-ulp differences in the torsion's trigonometry, shown on
the CPU.

The product's 3D core, built on the host by tests/harness/core_cpu.cpp,
reproduces the FP32 oracle bit for bit (tests/test_oracle3d.py). The same
core built with BHC_TORSION_ULP_TEST (tests/harness/_core_cpu_ulp.so) moves
every single-precision cos / sin / acos of the torsion (RayNormal, Reflect)
by one ulp. That alone leaves every trajectory exact (the torsion enters
only the beam quantities p, q), changes the contribution count by about
0.1 % (receivers at a beam's edge flip in or out) and moves TL at a few
percent of the receivers, by up to a few dB. This is the pattern of the
FP32 GPU fields against the oracle, whose only departures from the host's
arithmetic in these cases are the device's single-precision cos / sin / acos
(tests/test_gpu_parity3d.py holds FP32 3D TL to trajectories exact,
contributions within 0.5 % and >= 99 % of receivers within 0.1 dB); the
ulp-moved build is a stress bound (every call moved, where the device and
glibc differ only in some).
"""
import ctypes as C
import os

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd import _abi
import oracle
from parity import tl_db
from test_oracle3d import CORE_CASES

_H = None


def _ulp_harness():
    global _H
    if _H is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "harness", "_core_cpu_ulp.so")
        _H = C.CDLL(path)
        _H.core_field_run3d.restype = C.c_int
        _H.core_field_run3d.argtypes = [C.POINTER(_abi.Problem3D), C.POINTER(_abi.FieldOut)]
    return _H


def _run(env):
    p = bh.Problem3D(env)
    n = p.nrays
    field = np.zeros(p.field_shape, np.complex64)
    ns, nt, nb = (np.zeros(n, np.int32) for _ in range(3))
    xe = np.zeros((n, 3))
    out = _abi.FieldOut()
    out.field = field.ctypes.data_as(_abi.c_float_p)
    out.ray_nsteps = ns.ctypes.data_as(_abi.c_int32_p)
    out.ray_ntop = nt.ctypes.data_as(_abi.c_int32_p)
    out.ray_nbot = nb.ctypes.data_as(_abi.c_int32_p)
    out.ray_xend = xe.ctypes.data_as(_abi.c_double_p)
    assert _ulp_harness().core_field_run3d(C.byref(p.view), C.byref(out)) == 0
    return p, field, ns, nt, nb, xe, int(out.total_contrib)


@pytest.mark.parametrize("name", ["shelf_canyon_hexahedral_f32", "shelf_canyon_raycen_hat_f32"])
def test_one_ulp_torsion_trig_moves_edges_not_trajectories(name):
    env = CORE_CASES[name]()
    o = oracle.field_run3d(env)
    p, field, ns, nt, nb, xe, contrib = _run(env)
    # trajectories do not see the torsion
    assert np.array_equal(ns, o.nsteps) and np.array_equal(nt, o.ntop) and np.array_equal(nb, o.nbot)
    assert np.array_equal(xe, o.xend)
    # receivers at beam edges flip: a fraction of a percent of the contributions
    rel = abs(contrib - o.total_contrib) / o.total_contrib
    assert 0 < rel <= 0.005
    to = tl_db(oracle.scale_pressure3d(env, o.field))
    tg = tl_db(p.scale_pressure(field))
    m = to < 120.0
    d = np.abs(to - tg)[m]
    frac = float(np.mean(d <= 0.1))
    print(name, "contributions %+.2e, within 0.1 dB %.4f, median %.1e dB, max %.2f dB" % (
        (contrib - o.total_contrib) / o.total_contrib, frac, np.median(d), d.max()))
    # most receivers are untouched, a few percent move by up to a few dB
    assert np.median(d) < 0.01 and 0.9 <= frac < 1.0 and 0.1 < d.max() < 10.0


def test_one_ulp_torsion_trig_without_reflection():
    """A free-space run never reflects (no acos) and its torsion stays zero:
    the ulp-moved build leaves its contributions and TL as the oracle's (TL
    to the rounding of the field sums, whose order the oracle's threads vary)."""
    env = CORE_CASES["free_incoherent_f32"]()
    o = oracle.field_run3d(env)
    p, field, ns, nt, nb, xe, contrib = _run(env)
    assert np.array_equal(ns, o.nsteps) and contrib == o.total_contrib
    to = tl_db(oracle.scale_pressure3d(env, o.field))
    tg = tl_db(p.scale_pressure(field))
    m = to < 120.0
    assert np.max(np.abs(to - tg)[m]) < 1e-3
