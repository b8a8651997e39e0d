"""GPU parity of the full-3D path: the HIP tracer (bhc_field_run3d) against the
3D oracle on the same inputs.

Bars: per-ray step and bounce counts equal on >= 99.9 % of rays, ray end
points within 1e-9 relative (FP64), TL within 0.1 dB where TL < 120 dB, and
equal receiver-contribution counts (the scatter's depth windows skip only
receivers the reference's tests reject). Each case prints its name before it
touches the GPU.
"""
import sys

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd.env3d import free_env3d, munk_env3d, c3_env, analysis3d_env
import oracle
from parity import compare_field, tl_db

pytestmark = pytest.mark.gpu

def _refl_table(env, n):
    """Bottom reflection coefficients read from a table ('F') of n angles:
    12 angles are staged in LDS by the trace kernel, 300 (900 reals) are read
    from HBM."""
    from bellhopcuda_amd.env import ReflTable
    th = np.linspace(0.0, 90.0, n)
    env.brc = ReflTable(th, 1.0 - 0.6 * th / 90.0, 180.0 - 0.5 * th)
    return env


CASES = {
    "munk_gauss_refl_table": lambda: _refl_table(munk_env3d(nalpha=151, nbeta=7, bot="F"), 12),
    "munk_gauss_refl_table_unstaged": lambda: _refl_table(munk_env3d(nalpha=151, nbeta=7, bot="F"), 300),
    "free_gauss_f64": lambda: free_env3d(),
    "free_hat_f64": lambda: free_env3d(run_type="CG   3"),
    "free_gauss_incoherent_f32": lambda: free_env3d(run_type="IB   3", precision=32),
    "munk_hat_clinear": lambda: munk_env3d(nalpha=151, nbeta=7, run_type="CG   3"),
    "munk_gauss_semicoherent_pchip": lambda: munk_env3d(nalpha=151, nbeta=7, run_type="SB   3", ssp_type="P"),
    "munk_gauss_spline_bearings": lambda: munk_env3d(nalpha=101, nbeta=9, beta=(-40.0, 40.0), ntheta=8,
                                                     ssp_type="S"),
    "shelf_canyon_hexahedral_f64": lambda: c3_env(nalpha=61, nbeta=36, ntheta=8, nrz=31, nrr=81, nxy=51),
    "shelf_canyon_hexahedral_hat": lambda: c3_env(nalpha=61, nbeta=36, ntheta=8, nrz=31, nrr=81, nxy=51,
                                                  run_type="CG   3"),
    "shelf_canyon_hexahedral_f32": lambda: c3_env(nalpha=61, nbeta=36, ntheta=8, nrz=31, nrr=81, nxy=51,
                                                  precision=32),
    # 8000 receiver ranges: the tables exceed the LDS budget, the scatter reads them from HBM
    "free_gauss_unstaged_tables": lambda: free_env3d(nalpha=45, nbeta=19, ntheta=2, nrz=5, nrr=8000),
    "free_raycen_gauss_incoherent": lambda: free_env3d(run_type="Ib   3"),
    "munk_raycen_hat": lambda: munk_env3d(nalpha=151, nbeta=7, run_type="Cg   3"),
    "shelf_canyon_raycen_gauss": lambda: c3_env(nalpha=41, nbeta=24, ntheta=6, nrz=21, nrr=51, nxy=51,
                                                run_type="Cb   3"),
    "shelf_canyon_raycen_hat_f32": lambda: c3_env(nalpha=41, nbeta=24, ntheta=6, nrz=21, nrr=51, nxy=51,
                                                  run_type="Cg   3", precision=32),
}


@pytest.mark.parametrize("name", list(CASES))
def test_field_parity3d(name, gpu_ctx):
    print("\n[gpu-case-3d]", name, flush=True)
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
    tlo, tlg = tl_db(so), tl_db(sg)
    m = tlo < 120.0
    rep["tl_frac_within_0.1dB"] = float(np.mean(np.abs(tlo - tlg)[m] <= 0.1))
    print(name, rep, "err", g.err, o.err, "warn", g.warn, o.warn, flush=True)
    assert g.err == o.err
    assert g.total_steps == o.total_steps
    if fp64:
        assert ok, rep
        assert g.total_contrib == o.total_contrib
    else:
        # FP32: trajectories are exact; the beam widths pass through the
        # torsion's single-precision cos / sin / acos, whose device and host
        # libm results differ by an ulp, so receivers at a beam's edge can
        # flip in or out: held to a statistical bar
        assert rep["steps_exact"] == 1.0 and rep["bounces_exact"] == 1.0, rep
        assert abs(g.total_contrib - o.total_contrib) <= 0.005 * o.total_contrib, rep
        assert rep["tl_frac_within_0.1dB"] >= 0.99, rep


def test_run3d_rejects_unsupported(gpu_ctx):
    env = free_env3d(run_type="CR   3")
    with pytest.raises(bh.BhcError):
        bh.Problem3D(env)


RAY_CASES = {
    "munk_rays": lambda: munk_env3d(nalpha=41, nbeta=5, beta=(-30.0, 30.0), run_type="R    3"),
    "shelf_canyon_rays": lambda: c3_env(nalpha=21, nbeta=12, ntheta=4, nrz=11, nrr=21, nxy=51, run_type="R    3"),
    "shelf_canyon_rays_f32": lambda: c3_env(nalpha=21, nbeta=12, ntheta=4, nrz=11, nrr=21, nxy=51,
                                            run_type="R    3", precision=32),
}


@pytest.mark.parametrize("name", list(RAY_CASES))
def test_ray_mode3d_parity(name, gpu_ctx):
    """Ray run type in 3D: step counts and every point's position and tangent
    equal the oracle's (FP64 and FP32: they involve no transcendental); p, q
    and the torsion to 1e-6 relative in FP64 (device cos / sin / acos of the
    torsion) and 1e-2 in FP32, where one ulp of the torsion's single-precision
    trigonometry compounds over the reflections."""
    print("\n[gpu-case-3d-ray]", name, flush=True)
    env = RAY_CASES[name]()
    M = 3000
    po, no = oracle.ray_run3d(env, max_points=M)
    prob, pg, ng = bh.ray_run3d(env, ctx=gpu_ctx, max_points=M)
    assert np.array_equal(no, ng)
    ao = np.ctypeslib.as_array(po).reshape(len(no), M)
    ag = np.ctypeslib.as_array(pg).reshape(len(ng), M)
    for k in range(len(no)):
        n = min(no[k], M)
        assert np.array_equal(ao[k, :n]["x"], ag[k, :n]["x"])
        assert np.array_equal(ao[k, :n]["t"], ag[k, :n]["t"])
        assert np.array_equal(ao[k, :n]["NumBotBnc"], ag[k, :n]["NumBotBnc"])
        tol = 1e-6 if env.precision == 64 else 1e-2
        for f in ("p", "q"):
            a, b = ao[k, :n][f], ag[k, :n][f]
            assert np.allclose(a, b, rtol=tol, atol=tol * max(1.0, float(np.max(np.abs(a))))), (k, f)


ARR_CASES = {
    # (env builder, max_narr, allow_merging, oracle threads)
    "analysis3d_hat_f64": (lambda: analysis3d_env(), 10, 1, 1),
    "analysis3d_hat_f32": (lambda: analysis3d_env(precision=32), 10, 1, 1),
    "munk_gauss_merging": (lambda: munk_env3d(nalpha=61, nbeta=5, beta=(-20.0, 20.0), ntheta=2, nrz=11, nrr=21,
                                              run_type="AB   3"), 20, 1, 1),
    "munk_hat_first_maxnarr": (lambda: munk_env3d(nalpha=61, nbeta=5, beta=(-20.0, 20.0), ntheta=2, nrz=11,
                                                  nrr=21, run_type="aG   3"), 6, 0, 1),
    "shelf_canyon_gauss_merging": (lambda: c3_env(nalpha=31, nbeta=12, ntheta=4, nrz=11, nrr=21, nxy=41,
                                                  run_type="AB   3"), 30, 1, 1),
    "shelf_canyon_raycen_hat": (lambda: c3_env(nalpha=31, nbeta=12, ntheta=4, nrz=11, nrr=21, nxy=41,
                                               run_type="Ag   3"), 30, 1, 1),
}


@pytest.mark.parametrize("name", list(ARR_CASES))
def test_arrivals3d_parity(name, gpu_ctx):
    """3D arrivals: candidates merged per receiver in (ray, step) order replay
    the single-threaded AddArr, so counts, amplitudes, phases, delays and
    bounces equal the oracle's (FP64; angles within 1e-6 relative, device
    atan2); FP32 runs match counts and bounces and the values to 1e-4."""
    print("\n[gpu-case-3d-arr]", name, flush=True)
    mk, max_narr, merging, nth = ARR_CASES[name]
    env = mk()
    o = oracle.field_run3d(env, nthreads=nth, max_narr=max_narr, allow_merging=merging)
    g = bh.field_run3d(env, ctx=gpu_ctx, max_narr=max_narr, allow_merging=merging)
    assert g.err == o.err and g.total_contrib == o.total_contrib
    gn = np.minimum(g.narr, max_narr)
    on = np.minimum(o.narr, max_narr)
    assert np.array_equal(gn, on)
    assert on.sum() > 0
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


def test_eigen_hits3d_parity(gpu_ctx):
    """3D eigen hits: the same set of (step, bearing, range, depth, ray) hits as
    the oracle, in any order (both record them with an atomic counter)."""
    print("\n[gpu-case-3d-eig] munk_eigen", flush=True)
    env = munk_env3d(nalpha=61, nbeta=5, beta=(-20.0, 20.0), ntheta=2, nrz=11, nrr=21, run_type="EG   3")
    o = oracle.field_run3d(env, eigen_memsize=200000)
    g = bh.field_run3d(env, ctx=gpu_ctx, eigen_memsize=200000)
    no, ng = int(o.neigen[0]), int(g.neigen[0])
    assert no == ng > 0
    key = ("is_", "itheta", "ir", "iz", "isx", "isy", "isz", "ialpha", "ibeta")
    ho = np.sort(np.ctypeslib.as_array(o.hits)[:no], order=list(key))
    hg = np.sort(np.ctypeslib.as_array(g.hits)[:ng], order=list(key))
    assert np.array_equal(ho, hg)


def test_split_passes3d_keep_counts_and_field(gpu_ctx):
    """A 16 MiB segment buffer splits a 1057-ray 3D Munk field (~118 k steps)
    into passes, with overflows after completed passes: step and contribution
    counts and the field equal the oracle's."""
    from parity import small_segment_context
    print("\n[gpu-case-3d] split_passes3d", flush=True)
    env = munk_env3d(nalpha=151, nbeta=7, run_type="CG   3")
    small = small_segment_context(16)
    try:
        g = bh.field_run3d(env, ctx=small, stats=True)
    finally:
        small.close()
    o = oracle.field_run3d(env)
    assert g.total_steps == o.total_steps == int(g.nsteps.sum())
    assert g.total_contrib == o.total_contrib
    assert np.array_equal(g.nsteps, o.nsteps)
    so, sg = oracle.scale_pressure3d(env, o.field), bh.Problem3D(env).scale_pressure(g.field)
    tlo, tlg = tl_db(so), tl_db(sg)
    m = tlo < 120.0
    assert np.max(np.abs(tlo - tlg)[m]) <= 0.1


def test_eigen_retrace3d_parity(gpu_ctx):
    """3D eigenray re-trace on the GPU (bhc_ray_run3d_list) equals the oracle's
    re-trace of the same hits: step counts, positions, tangents and bounces
    exact, p and q to 1e-6."""
    print("\n[gpu-case-3d-eig] munk_eigen_retrace", flush=True)
    env = munk_env3d(nalpha=61, nbeta=5, beta=(-20.0, 20.0), ntheta=2, nrz=11, nrr=21, run_type="EG   3")
    o = oracle.field_run3d(env, eigen_memsize=200000)
    n = min(int(o.neigen[0]), 400)
    M = 3000
    po, no = oracle.ray_run3d_list(env, o.hits, n, max_points=M)
    _, pg, ng = bh.ray_run3d_list(env, o.hits, n, ctx=gpu_ctx, max_points=M)
    assert np.array_equal(no, ng)
    ao = np.ctypeslib.as_array(po).reshape(n, M)
    ag = np.ctypeslib.as_array(pg).reshape(n, M)
    for k in range(n):
        m = min(no[k], M)
        for f in ("x", "t", "NumTopBnc", "NumBotBnc"):
            assert np.array_equal(ao[k, :m][f], ag[k, :m][f]), (k, f)
        for f in ("p", "q"):
            a, b = ao[k, :m][f], ag[k, :m][f]
            assert np.allclose(a, b, rtol=1e-6, atol=1e-6 * max(1.0, float(np.max(np.abs(a))))), (k, f)


def test_arrivals_and_eigen3d_parity(gpu_ctx):
    """Run type 'V' in 3D: arrivals and eigen hits from the same run
    (IsAlsoEigenraysRun), each equal to the oracle's."""
    print("\n[gpu-case-3d-arr] munk_arrivals_and_eigen", flush=True)
    env = munk_env3d(nalpha=61, nbeta=5, beta=(-20.0, 20.0), ntheta=2, nrz=11, nrr=21, run_type="VG   3")
    o = oracle.field_run3d(env, max_narr=20, allow_merging=1, eigen_memsize=200000)
    g = bh.field_run3d(env, ctx=gpu_ctx, max_narr=20, allow_merging=1, eigen_memsize=200000)
    assert np.array_equal(np.minimum(o.narr, 20), np.minimum(g.narr, 20)) and o.narr.sum() > 0
    A = np.ctypeslib.as_array(o.arr).reshape(-1, 20)
    B = np.ctypeslib.as_array(g.arr).reshape(-1, 20)
    for r in np.nonzero(o.narr)[0]:
        n = min(o.narr[r], 20)
        for f in ("a", "Phase", "delay", "NTopBnc", "NBotBnc"):
            assert np.array_equal(A[r, :n][f], B[r, :n][f]), (r, f)
    no, ng = int(o.neigen[0]), int(g.neigen[0])
    assert no == ng > 0
    key = ["is_", "itheta", "ir", "iz", "isx", "isy", "isz", "ialpha", "ibeta"]
    assert np.array_equal(np.sort(np.ctypeslib.as_array(o.hits)[:no], order=key),
                          np.sort(np.ctypeslib.as_array(g.hits)[:ng], order=key))


@pytest.mark.parametrize("run_type,precision", [("CB   3", 64), ("AB   3", 64), ("CB   3", 32)])
def test_scatter3_mappings_agree(run_type, precision, gpu_ctx):
    """The three mappings of the Cartesian 3D scatter (one wavefront per
    segment, one lane per segment, segment groups flattened per wavefront)
    make the same contributions: equal counts (FP64: the oracle's), TL within
    1e-3 dB of each other (FP32 atomic order) and, for arrivals, equal merged
    lists. In FP32 the three run the same float arithmetic on the same
    segments, so they agree with each other exactly in count."""
    from parity import env_context
    print("\n[gpu-case-3d] scatter3 mappings", run_type, precision, flush=True)
    env = c3_env(nalpha=41, nbeta=24, ntheta=6, nrz=21, nrr=51, nxy=51, run_type=run_type, precision=precision)
    arr = run_type[0] == "A"
    o = oracle.field_run3d(env, max_narr=20) if arr else oracle.field_run3d(env)
    res = {}
    for m in (0, 1, 2):
        ctx = env_context(BHC_SCATTER3_MAP=m, BHC_SEG_BUFFER_MB=512)
        try:
            res[m] = bh.field_run3d(env, ctx=ctx, max_narr=20) if arr else bh.field_run3d(env, ctx=ctx)
        finally:
            ctx.close()
        assert res[m].err == o.err, m
        if precision == 64:
            assert res[m].total_contrib == o.total_contrib, (m, res[m].total_contrib)
        else:
            assert res[m].total_contrib == res[0].total_contrib, (m, res[m].total_contrib)
            assert abs(res[m].total_contrib - o.total_contrib) <= 0.005 * o.total_contrib
    if arr:
        for m in (0, 1):
            assert np.array_equal(res[m].narr, res[2].narr)
            A = np.ctypeslib.as_array(res[m].arr).reshape(-1, 20)
            B = np.ctypeslib.as_array(res[2].arr).reshape(-1, 20)
            n = np.minimum(res[2].narr, 20)
            assert n.sum() > 0
            for r in np.nonzero(n)[0]:
                for f in ("a", "Phase", "delay", "NTopBnc", "NBotBnc"):
                    assert np.array_equal(A[r, :n[r]][f], B[r, :n[r]][f]), (m, r, f)
    else:
        prob = bh.Problem3D(env)
        tl = {m: tl_db(prob.scale_pressure(r.field)) for m, r in res.items()}
        mask = tl[2] < 120.0
        assert mask.sum() > 1000
        for m in (0, 1):
            assert np.max(np.abs(tl[m] - tl[2])[mask]) <= 1e-3, m
