"""GPU parity of the Nx2D path: the HIP tracer (bhc_field_runnx2d /
bhc_ray_runnx2d, 2D rays on the bearing planes of a 3D environment) against
the Nx2D oracle on the same inputs.

Bars as for the 2D and 3D paths: per-ray step and bounce counts equal on
>= 99.9 % of rays, ray end points within 1e-9 relative (FP64), TL within
0.1 dB where TL < 120 dB, equal receiver-contribution counts; arrivals equal
value for value (FP64) after the (ray, step)-ordered merge; eigen hits equal as
sets. Each case prints its name before it touches the GPU.
"""
import sys

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd.env3d import munk_env3d, c3_env
import oracle
from parity import compare_field, tl_db, small_segment_context
from test_oracle_nx2d import _rotated_munk, _wedge_pair

pytestmark = pytest.mark.gpu


def _announce(tag, name):
    print("\n[%s]" % tag, name, flush=True)
    sys.stdout.flush()


def _refl_table(env, n):
    """Bottom reflection coefficients from a table ('F') of n angles (staged
    in LDS by the trace kernel when they fit)."""
    from bellhopcuda_amd.env import ReflTable
    th = np.linspace(0.0, 90.0, n)
    env.brc = ReflTable(th, 1.0 - 0.6 * th / 90.0, 180.0 - 0.5 * th)
    return env


CASES = {
    "munk_gauss_refl_table": lambda: _refl_table(munk_env3d(nalpha=301, nbeta=1, ntheta=1, nrz=51, nrr=101,
                                                            run_type="CB   2", bot="F"), 12),
    "munk_hat_bearing0": lambda: munk_env3d(nalpha=401, nbeta=1, ntheta=1, nrz=51, nrr=201, run_type="CG   2"),
    "munk_gauss_rotated": lambda: _rotated_munk(),
    "munk_gauss_rotated_f32": lambda: _rotated_munk(precision=32),
    "munk_raycen_hat_rotated": lambda: _rotated_munk(rt="Cg   2"),
    "munk_semicoherent_n2linear": lambda: munk_env3d(nalpha=301, nbeta=1, ntheta=1, nrz=51, nrr=101,
                                                     run_type="SB   2", ssp_type="N"),
    "wedge_x_halfspace": lambda: _wedge_pair(0)[1],
    "wedge_y_halfspace": lambda: _wedge_pair(1)[1],
    "shelf_canyon_hexahedral": lambda: c3_env(nalpha=201, nbeta=36, ntheta=36, nrz=51, nrr=101, nxy=81,
                                              run_type="CB   2"),
    "shelf_canyon_hexahedral_f32": lambda: c3_env(nalpha=201, nbeta=36, ntheta=36, nrz=51, nrr=101, nxy=81,
                                                  run_type="CB   2", precision=32),
    "shelf_canyon_raycen_incoherent": lambda: c3_env(nalpha=101, nbeta=18, ntheta=18, nrz=31, nrr=81, nxy=81,
                                                     run_type="Ib   2"),
}


@pytest.mark.parametrize("name", list(CASES))
def test_field_parity_nx2d(name, gpu_ctx):
    _announce("gpu-case-nx2d", name)
    env = CASES[name]()
    o = oracle.field_run3d(env)
    prob = bh.Problem3D(env)
    assert prob.nx2d
    g = bh.field_run3d(prob, ctx=gpu_ctx, stats=True)
    so = oracle.scale_pressure3d(env, o.field)
    sg = prob.scale_pressure(g.field)
    ok, rep = compare_field(o, g, so, sg, fp64=env.precision == 64)
    rep["contrib"] = (int(g.total_contrib), int(o.total_contrib))
    print(name, rep, "err", g.err, o.err, "warn", g.warn, o.warn, flush=True)
    assert g.err == o.err and g.warn == o.warn
    assert g.total_steps == o.total_steps
    assert ok, rep
    assert g.total_contrib == o.total_contrib


def test_split_passes_nx2d(gpu_ctx):
    """A 16 MiB segment buffer splits the run into trace / scatter passes,
    some retraced after an overflow: counts and field unchanged."""
    _announce("gpu-case-nx2d", "split_passes")
    env = c3_env(nalpha=201, nbeta=36, ntheta=36, nrz=51, nrr=101, nxy=81, run_type="CG   2")
    small = small_segment_context(16)
    try:
        g = bh.field_run3d(env, ctx=small, stats=True)
        passes = small.last_phase()[2]
    finally:
        small.close()
    o = oracle.field_run3d(env)
    assert passes > 1
    assert g.total_steps == o.total_steps and g.total_contrib == o.total_contrib
    assert np.array_equal(g.nsteps, o.nsteps)
    so, sg = oracle.scale_pressure3d(env, o.field), bh.Problem3D(env).scale_pressure(g.field)
    tlo, tlg = tl_db(so), tl_db(sg)
    m = tlo < 120.0
    assert np.max(np.abs(tlo - tlg)[m]) <= 0.1


RAY_CASES = {
    "munk_rays_rotated": lambda: _ray_env(),
    "shelf_canyon_rays": lambda: c3_env(nalpha=21, nbeta=12, ntheta=4, nrz=11, nrr=21, nxy=51, run_type="R    2"),
    "shelf_canyon_rays_f32": lambda: c3_env(nalpha=21, nbeta=12, ntheta=4, nrz=11, nrr=21, nxy=51,
                                            run_type="R    2", precision=32),
}


def _ray_env():
    e = munk_env3d(nalpha=41, nbeta=1, nrz=5, nrr=5, run_type="R    2")
    e.beta_deg = np.array([0.0, 37.0, 143.0, 200.0])
    e.Sx_km, e.Sy_km = np.array([3.0]), np.array([-2.0])
    e.box_x_km = e.box_y_km = 30.0
    return e


@pytest.mark.parametrize("name", list(RAY_CASES))
def test_ray_mode_nx2d_parity(name, gpu_ctx):
    """Nx2D ray run type: step counts, origins and every point's position,
    tangent, p, q and bounces equal the oracle's; the amplitude and phase,
    which pass through the halfspace coefficient's |R| (hypot) and arg R
    (atan2) at each bottom reflection, within 1e-12 relative in FP64 and 1e-6
    in FP32 (the device's and glibc's hypot / atan2 may differ by an ulp)."""
    _announce("gpu-case-nx2d-ray", name)
    env = RAY_CASES[name]()
    M = 3000
    po, no, oo = oracle.ray_runnx2d(env, max_points=M)
    prob, pg, ng, og = bh.ray_runnx2d(env, ctx=gpu_ctx, max_points=M)
    assert np.array_equal(no, ng)
    assert np.array_equal(oo, og)
    ao = np.ctypeslib.as_array(po).reshape(len(no), M)
    ag = np.ctypeslib.as_array(pg).reshape(len(ng), M)
    for k in range(len(no)):
        n = min(no[k], M)
        for f in ("x", "t", "p", "q", "NumTopBnc", "NumBotBnc"):
            assert np.array_equal(ao[k, :n][f], ag[k, :n][f]), (k, f)
        tol = 1e-12 if env.precision == 64 else 1e-6
        for f in ("Amp", "Phase"):
            assert np.allclose(ao[k, :n][f], ag[k, :n][f], rtol=tol, atol=tol), (k, f)


ARR_CASES = {
    "munk_gauss_merging": (lambda: _arr_env("AB   2"), 20, 1),
    "munk_hat_first_maxnarr": (lambda: _arr_env("aG   2"), 6, 0),
    "shelf_canyon_gauss_merging": (lambda: c3_env(nalpha=61, nbeta=8, ntheta=8, nrz=11, nrr=21, nxy=41,
                                                  run_type="AB   2"), 30, 1),
    "shelf_canyon_raycen_hat_f32": (lambda: c3_env(nalpha=61, nbeta=8, ntheta=8, nrz=11, nrr=21, nxy=41,
                                                   run_type="Ag   2", precision=32), 30, 1),
}


def _arr_env(rt):
    e = munk_env3d(nalpha=101, nbeta=1, ntheta=1, nrz=11, nrr=21, run_type=rt, r_max_km=10.0)
    e.theta_deg = np.array([0.0, 25.0, 250.0])
    return e


@pytest.mark.parametrize("name", list(ARR_CASES))
def test_arrivals_nx2d_parity(name, gpu_ctx):
    """Nx2D arrivals (AddArr<false> on the bearing plane's receivers, azimuths
    = the bearing): counts, amplitudes, phases, delays and bounces equal the
    oracle's in FP64 (angles to 1e-6 relative: device atan2), within 1e-4 in
    FP32."""
    _announce("gpu-case-nx2d-arr", name)
    mk, max_narr, merging = ARR_CASES[name]
    env = mk()
    o = oracle.field_run3d(env, nthreads=1, max_narr=max_narr, allow_merging=merging)
    g = bh.field_run3d(env, ctx=gpu_ctx, max_narr=max_narr, allow_merging=merging)
    assert g.err == o.err and g.total_contrib == o.total_contrib
    on, gn = np.minimum(o.narr, max_narr), np.minimum(g.narr, max_narr)
    assert np.array_equal(on, gn) and on.sum() > 0
    A = np.ctypeslib.as_array(o.arr).reshape(-1, max_narr)
    B = np.ctypeslib.as_array(g.arr).reshape(-1, max_narr)
    fp64 = env.precision == 64
    for r in np.nonzero(on)[0]:
        a, b = A[r, :on[r]], B[r, :on[r]]
        assert np.array_equal(a["NTopBnc"], b["NTopBnc"]) and np.array_equal(a["NBotBnc"], b["NBotBnc"]), r
        for f in ("a", "Phase", "delay"):
            if fp64:
                assert np.array_equal(a[f], b[f]), (r, f)
            else:
                assert np.allclose(a[f], b[f], rtol=1e-4, atol=1e-9), (r, f)
        for f in ("SrcDeclAngle", "SrcAzimAngle", "RcvrDeclAngle", "RcvrAzimAngle"):
            assert np.allclose(a[f], b[f], rtol=1e-6 if fp64 else 1e-4, atol=1e-5), (r, f)


def test_eigen_hits_nx2d_parity(gpu_ctx):
    """Nx2D eigen hits (itheta = ibeta, source indices): the oracle's set."""
    _announce("gpu-case-nx2d-eig", "munk_and_shelf_eigen")
    for env in (_arr_env("EG   2"), c3_env(nalpha=61, nbeta=8, ntheta=8, nrz=11, nrr=21, nxy=41, run_type="EB   2")):
        o = oracle.field_run3d(env, eigen_memsize=200000)
        g = bh.field_run3d(env, ctx=gpu_ctx, eigen_memsize=200000)
        no, ng = int(o.neigen[0]), int(g.neigen[0])
        assert no == ng > 0
        key = ["is_", "itheta", "ir", "iz", "isx", "isy", "isz", "ialpha", "ibeta"]
        assert np.array_equal(np.sort(np.ctypeslib.as_array(o.hits)[:no], order=key),
                              np.sort(np.ctypeslib.as_array(g.hits)[:ng], order=key))


def test_async_nx2d_equals_blocking(gpu_ctx):
    _announce("gpu-case-nx2d", "async")
    prob = bh.Problem3D(_rotated_munk())
    ref = bh.field_run3d(prob, ctx=gpu_ctx, stats=True)
    ctx = bh.Context(0)
    try:
        flag = []
        run = bh.AsyncRun(prob, ctx, on_done=lambda: flag.append(1), stats=True)
        res = run.wait()
        assert flag == [1] and ctx.progress() == (prob.nrays, prob.nrays)
    finally:
        ctx.close()
    assert res.total_steps == ref.total_steps and res.total_contrib == ref.total_contrib
    assert np.array_equal(res.nsteps, ref.nsteps)
    ta, tb = tl_db(prob.scale_pressure(res.field)), tl_db(prob.scale_pressure(ref.field))
    m = tb < 120
    assert np.max(np.abs(ta - tb)[m]) < 0.01


def test_nx2d_entry_points_reject_3d_problems(gpu_ctx):
    e3 = munk_env3d(nalpha=11, nbeta=3, nrz=3, nrr=3, run_type="CG   3")
    nx = munk_env3d(nalpha=11, nbeta=1, nrz=3, nrr=3, run_type="CG   2")
    from bellhopcuda_amd import _abi
    import ctypes as C
    p3 = bh.Problem3D(e3)
    res = bh.FieldResult()
    out = bh._make_out3d(p3, res, False, 0, 1, 0)
    assert bh.lib().bhc_field_runnx2d(gpu_ctx._h, C.byref(p3.view), C.byref(out)) != 0
    pn = bh.Problem3D(nx)
    out = bh._make_out3d(pn, res, False, 0, 1, 0)
    assert bh.lib().bhc_field_run3d(gpu_ctx._h, C.byref(pn.view), C.byref(out)) != 0
