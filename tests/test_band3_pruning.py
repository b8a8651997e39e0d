"""Brute-force check of the 3D Cartesian scatter's pruning on the CPU.

The host build of the product core tests every step of a sample of rays
against every receiver the reference's row test admits (no bearing
prefilter, no depth band) and evaluates each with the beam-window test:
every accepted receiver must have passed the bearing prefilter and lie in
its row's band (the circle bound intersected with the per-axis bounds of
the scaled q), in FP64 and FP32, for Gaussian and hat beams. The bands must
also prune: they hold a fraction of the receivers of the full windows.
"""
import ctypes as C

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd.env3d import free_env3d, munk_env3d, c3_env
from bellhopcuda_amd import _abi
from test_oracle3d import harness

CASES = {
    "shelf_canyon_gauss_f64": (lambda: c3_env(nalpha=201, nbeta=37, ntheta=12, nrz=51, nrr=201), 3),
    "shelf_canyon_gauss_f32": (lambda: c3_env(nalpha=201, nbeta=37, ntheta=12, nrz=51, nrr=201, precision=32), 3),
    "shelf_canyon_hat_f64": (lambda: c3_env(nalpha=201, nbeta=37, ntheta=12, nrz=51, nrr=201, run_type="CG   3"), 3),
    "shelf_canyon_hat_f32": (lambda: c3_env(nalpha=201, nbeta=37, ntheta=12, nrz=51, nrr=201, run_type="CG   3",
                                            precision=32), 3),
    "munk_gauss_bearings": (lambda: munk_env3d(nalpha=151, nbeta=9, beta=(-40.0, 40.0), ntheta=8, ssp_type="S"), 1),
    "free_gauss": (lambda: free_env3d(), 1),
    "free_hat_f32": (lambda: free_env3d(run_type="CG   3", precision=32), 1),
}


@pytest.mark.parametrize("name", list(CASES))
def test_band_pruning_keeps_every_accepted_receiver(name):
    mk, stride = CASES[name]
    h = harness()
    h.core_band_check3d.restype = C.c_int
    h.core_band_check3d.argtypes = [C.POINTER(_abi.Problem3D), C.c_int, C.POINTER(C.c_longlong)]
    p = bh.Problem3D(mk())
    out = (C.c_longlong * 5)()
    assert h.core_band_check3d(C.byref(p.view), stride, out) == 0
    accepted, pre_lost, band_lost, full, band = list(out)
    assert accepted > 1000
    assert pre_lost == 0 and band_lost == 0, list(out)
    assert band < 0.5 * full, list(out)
