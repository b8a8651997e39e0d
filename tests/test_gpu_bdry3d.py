"""GPU parity on curvilinear ('C') and province ('L') 3D boundary grids, full
3D and Nx2D, against the oracle: the bars of the 3D and Nx2D parity tests
(steps, bounces and end points exact in FP64, TL within 0.1 dB, equal
contribution counts). Each case prints its name before it touches the GPU."""
import sys

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd.env3d import munk_env3d
import oracle
from parity import compare_field
from test_oracle3d_bdry import _trough, _prov_grid

pytestmark = pytest.mark.gpu


def _env(rt, grid, precision=64, nalpha=201, nbeta=21):
    e = munk_env3d(nalpha=nalpha, nbeta=nbeta if rt.endswith("3") else 1, beta=(-10.0, 10.0), ntheta=1, nrz=41,
                   nrr=101, run_type=rt, precision=precision)
    if rt.endswith("2"):
        e.theta_deg = np.array([0.0, 10.0, 20.0, 300.0])
    else:
        e.theta_deg = np.array([0.0, 5.0])
    e.bty = grid
    e.box_x_km = e.box_y_km = 30.0
    return e


CASES = {
    "trough_curvilinear_3d": lambda: _env("CB   3", _trough(a=3e-5, z0=3500.0, nx=81, ny=9, kind="CS", x0=5.0,
                                                             x1=25.0)[0]),
    "provinces_curvilinear_3d": lambda: _env("CG   3", _prov_grid("CL")),
    "provinces_3d_f32": lambda: _env("CB   3", _prov_grid("RL"), precision=32),
    "trough_curvilinear_nx2d": lambda: _env("CB   2", _trough(a=3e-5, z0=3500.0, nx=81, ny=9, kind="CS", x0=5.0,
                                                              x1=25.0)[0]),
    "provinces_curvilinear_nx2d": lambda: _env("CG   2", _prov_grid("CL")),
    "provinces_nx2d_f32": lambda: _env("IB   2", _prov_grid("RL"), precision=32),
}


@pytest.mark.parametrize("name", list(CASES))
def test_field_parity_bdry3d(name, gpu_ctx):
    print("\n[gpu-case-bdry3d]", name, flush=True)
    sys.stdout.flush()
    env = CASES[name]()
    o = oracle.field_run3d(env)
    prob = bh.Problem3D(env)
    g = bh.field_run3d(prob, ctx=gpu_ctx, stats=True)
    so = oracle.scale_pressure3d(env, o.field)
    sg = prob.scale_pressure(g.field)
    fp64 = env.precision == 64
    ok, rep = compare_field(o, g, so, sg, fp64=fp64)
    rep["contrib"] = (int(g.total_contrib), int(o.total_contrib))
    print(name, rep, "bounces", int(o.nbot.sum()), flush=True)
    assert o.nbot.sum() > 0
    assert g.err == o.err and g.total_steps == o.total_steps
    if fp64 or env.nx2d:
        assert ok, rep
        assert g.total_contrib == o.total_contrib
    else:
        # FP32 3D: trajectories exact; beam widths through the torsion's
        # single-precision trigonometry (see test_gpu_parity3d)
        assert rep["steps_exact"] == 1.0 and rep["bounces_exact"] == 1.0, rep
        assert abs(g.total_contrib - o.total_contrib) <= 0.005 * o.total_contrib, rep


def test_arrivals_on_provinces_nx2d(gpu_ctx):
    print("\n[gpu-case-bdry3d] arrivals_provinces_nx2d", flush=True)
    env = _env("AB   2", _prov_grid("RL"), nalpha=101)
    env.Rz = np.linspace(0.0, 5000.0, 11).astype(np.float32)
    env.Rr_km = np.linspace(0.0, 20.0, 21)
    o = oracle.field_run3d(env, nthreads=1, max_narr=20, allow_merging=1)
    g = bh.field_run3d(env, ctx=gpu_ctx, max_narr=20, allow_merging=1)
    on, gn = np.minimum(o.narr, 20), np.minimum(g.narr, 20)
    assert np.array_equal(on, gn) and on.sum() > 0
    A = np.ctypeslib.as_array(o.arr).reshape(-1, 20)
    B = np.ctypeslib.as_array(g.arr).reshape(-1, 20)
    for r in np.nonzero(on)[0]:
        for f in ("a", "Phase", "delay", "NTopBnc", "NBotBnc"):
            assert np.array_equal(A[r, :on[r]][f], B[r, :on[r]][f]), (r, f)
