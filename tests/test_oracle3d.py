"""The 3D oracle and the product's 3D core on the CPU.

What pins the 3D oracle (oracle/bellhop3d.c):
  * free space (isovelocity, absorbing halfspaces above and below): coherent
    hat beams and incoherent Gaussian beams reproduce spherical spreading,
    TL = 20 log10 R, to a fraction of a dB -- this fixes the 3D amplitude
    (Ratio1 with sqrt(Dalpha Dbeta) / c, rcp_q0, rcp_qhat0), the q-matrix
    evolution, the beam coordinates n1', n2' and the Gaussian scale 2 pi;
  * on a depth-only Munk environment the 3D field on a bearing equals the
    independently pinned 2D oracle's field (the 3D ray equations reduce to
    2D ones, reflections included) -- this fixes steps, boundary crossings,
    the fluid-halfspace reflection and the coherent phase;
  * a 3D bathymetry grid that is flat reproduces the flat default boundary;
  * a hexahedral grid with no lateral variation reproduces the same field as
    the equivalent c-linear depth profile.
  * arrivals: the reference's own 3D arrivals fixture
    (test/individual/analysis_3d/example_analysis_3d.arr.txt, copied to
    tests/golden/) is reproduced end to end -- oracle AddArr with azimuths,
    PostProcessArrivals 3D, the 3D .arr writer -- to within one float ulp of
    every amplitude and delay (the fixture's receiver ranges are single
    precision, e.g. 900.000061 m) and exactly in counts, phases, angles and
    bounces; eigen hits bracket the receivers they name.
The product's 3D core (core3d.h, run serially by tests/harness/core_cpu.cpp
with the scatter kernel's row / depth-window enumeration) is compared with
the oracle bit for bit on trajectories and contribution counts. Its
preprocessing (host3d.cpp) is compared value for value.
"""
import ctypes as C
import os

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd import _abi
from bellhopcuda_amd.env import munk_env, munk_c
from bellhopcuda_amd import files
from bellhopcuda_amd.env3d import free_env3d, munk_env3d, c3_env, analysis3d_env, Grid3D
import oracle
from parity import tl_db

_H = None


def harness():
    global _H
    if _H is None:
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "harness", "_core_cpu.so")
        _H = C.CDLL(path)
        _H.core_field_run3d.restype = C.c_int
        _H.core_field_run3d.argtypes = [C.POINTER(_abi.Problem3D), C.POINTER(_abi.FieldOut)]
    return _H


class _Res:
    pass


def core_run3d(env):
    p = bh.Problem3D(env)
    n = p.nrays
    r = _Res()
    r.field = np.zeros(p.field_shape, np.complex64)
    r.nsteps, r.ntop, r.nbot = (np.zeros(n, np.int32) for _ in range(3))
    r.xend = np.zeros((n, 3))
    out = _abi.FieldOut()
    out.field = r.field.ctypes.data_as(_abi.c_float_p)
    out.ray_nsteps = r.nsteps.ctypes.data_as(_abi.c_int32_p)
    out.ray_ntop = r.ntop.ctypes.data_as(_abi.c_int32_p)
    out.ray_nbot = r.nbot.ctypes.data_as(_abi.c_int32_p)
    out.ray_xend = r.xend.ctypes.data_as(_abi.c_double_p)
    assert harness().core_field_run3d(C.byref(p.view), C.byref(out)) == 0
    r.total_contrib, r.total_steps, r.err, r.warn = out.total_contrib, out.total_steps, out.err, out.warn
    return r


def _free_tl_error(env, f):
    Rz = env.Rz.astype(np.float64)
    Rr = env.Rr_km * 1000.0
    R = np.sqrt((Rz[:, None] - float(env.Sz[0])) ** 2 + Rr[None, :] ** 2)
    m = R > 500.0
    return tl_db(f)[:, m] - 20.0 * np.log10(R[m])[None, :]


@pytest.mark.parametrize("run_type,tol_med,tol_p95", [("CG   3", 0.05, 0.2), ("IB   3", 0.1, 0.25)])
def test_free_space_spherical_spreading(run_type, tol_med, tol_p95):
    env = free_env3d(nalpha=181, nbeta=72, run_type=run_type)
    o = oracle.field_run3d(env)
    f = oracle.scale_pressure3d(env, o.field)[0, 0, 0]   # [Ntheta, NRz, NRr]
    d = _free_tl_error(env, f)
    assert o.err == 0
    assert abs(np.median(d)) < tol_med, np.median(d)
    assert np.percentile(np.abs(d), 95) < tol_p95, np.percentile(np.abs(d), 95)
    # the medium is axially symmetric: every bearing sees the same field
    assert np.allclose(tl_db(f[0]), tl_db(f[-1]), atol=0.05)


@pytest.mark.parametrize("rt3,rt2,tol_med", [("CG   3", "CG", 0.1), ("IB   3", "IB", 0.15)])
def test_depth_only_munk_matches_2d(rt3, rt2, tol_med):
    e3 = munk_env3d(nalpha=801, nbeta=41, beta=(-4.0, 4.0), nrz=51, nrr=201, run_type=rt3)
    e2 = munk_env(nbeams=801, nrz=51, nrr=201, run_type=rt2, r_max_km=20.0)
    o3 = oracle.field_run3d(e3)
    o2 = oracle.field_run(e2)
    t3 = tl_db(oracle.scale_pressure3d(e3, o3.field)[0, 0, 0, 0])
    t2 = tl_db(oracle.scale_pressure(e2, o2.field)[0])
    m = (t2 < 90) & (t3 < 90)
    m[:, :10] = False
    d = np.abs(t3 - t2)[m]
    assert m.sum() > 3000
    assert np.median(d) < tol_med, np.median(d)
    # the in-plane ray (beta = 0) takes the 2D ray's steps and bounces
    r3 = o3.nsteps.reshape(41, 801)[20]
    r2 = o2.nsteps
    assert np.mean(r3 == r2) > 0.99


def test_flat_grid_equals_default_boundary():
    base = munk_env3d(nalpha=101, nbeta=5, nrz=21, nrr=41)
    grid = munk_env3d(nalpha=101, nbeta=5, nrz=21, nrr=41)
    x = np.linspace(-25.0, 25.0, 11)
    grid.bty = Grid3D(x_km=x, y_km=x, z=np.full((11, 11), 5000.0))
    a, b = oracle.field_run3d(base), oracle.field_run3d(grid)
    # the grid's cell edges and diagonals add step ends (TopBotSegCrossing,
    # TriDiagCrossing), so step counts grow; bounces and the field do not change
    assert np.all(b.nsteps >= a.nsteps)
    assert np.array_equal(a.nbot, b.nbot) and np.array_equal(a.ntop, b.ntop)
    ta, tb = tl_db(a.field), tl_db(b.field)
    m = ta < 100
    assert np.median(np.abs(ta - tb)[m]) < 0.05


def test_laterally_uniform_hexahedral_equals_clinear():
    e1 = munk_env3d(nalpha=101, nbeta=5, nrz=21, nrr=41, bot="V")
    e2 = munk_env3d(nalpha=101, nbeta=5, nrz=21, nrr=41, bot="V")
    z = e1.ssp_z
    e2.ssp_type = "H"
    e2.hex_x_km = np.array([-30.0, 30.0])
    e2.hex_y_km = np.array([-30.0, 30.0])
    e2.hex_z = z
    e2.hex_c = np.broadcast_to(munk_c(z)[None, None, :], (2, 2, len(z))).copy()
    a, b = oracle.field_run3d(e1), oracle.field_run3d(e2)
    assert np.mean(a.nsteps == b.nsteps) > 0.99
    ta, tb = tl_db(a.field), tl_db(b.field)
    m = ta < 100
    assert np.median(np.abs(ta - tb)[m]) < 0.01


CORE_CASES = {
    "free_gauss": lambda: free_env3d(),
    "free_hat": lambda: free_env3d(run_type="CG   3"),
    "free_incoherent_f32": lambda: free_env3d(run_type="IB   3", precision=32),
    "munk_hat": lambda: munk_env3d(nalpha=101, nbeta=5, run_type="CG   3"),
    "munk_semicoherent_pchip": lambda: munk_env3d(nalpha=101, nbeta=5, run_type="SB   3", ssp_type="P"),
    "munk_n2linear_bearings": lambda: munk_env3d(nalpha=61, nbeta=9, beta=(-40.0, 40.0), ntheta=8,
                                                 ssp_type="N"),
    "shelf_canyon_hexahedral": lambda: c3_env(nalpha=41, nbeta=24, ntheta=6, nrz=21, nrr=51, nxy=41),
    "shelf_canyon_hexahedral_f32": lambda: c3_env(nalpha=41, nbeta=24, ntheta=6, nrz=21, nrr=51, nxy=41,
                                                  precision=32),
    "free_raycen_gauss_incoherent": lambda: free_env3d(run_type="Ib   3", nalpha=45, nbeta=19, nrz=11, nrr=26),
    "free_raycen_hat": lambda: free_env3d(run_type="Cg   3", nalpha=45, nbeta=19, nrz=11, nrr=26),
    "munk_raycen_hat": lambda: munk_env3d(nalpha=101, nbeta=5, run_type="Cg   3"),
    "shelf_canyon_raycen_gauss": lambda: c3_env(nalpha=31, nbeta=16, ntheta=4, nrz=11, nrr=41, nxy=41,
                                                run_type="Cb   3"),
    "shelf_canyon_raycen_hat_f32": lambda: c3_env(nalpha=31, nbeta=16, ntheta=4, nrz=11, nrr=41, nxy=41,
                                                  run_type="Cg   3", precision=32),
}


@pytest.mark.parametrize("name", list(CORE_CASES))
def test_core3d_matches_oracle(name):
    env = CORE_CASES[name]()
    o = oracle.field_run3d(env)
    g = core_run3d(env)
    assert np.array_equal(g.nsteps, o.nsteps)
    assert np.array_equal(g.ntop, o.ntop) and np.array_equal(g.nbot, o.nbot)
    assert np.array_equal(g.xend, o.xend)
    assert g.total_contrib == o.total_contrib
    assert g.err == o.err and g.warn == o.warn
    to, tg = tl_db(o.field), tl_db(g.field)
    m = to < 120
    assert np.max(np.abs(to - tg)[m]) < 1e-3


@pytest.mark.parametrize("precision", [64, 32])
def test_preprocess3d_matches_oracle(precision):
    env = c3_env(nalpha=21, nbeta=12, ntheta=6, nrz=11, nrr=21, nxy=31, precision=precision)
    ref = oracle.preprocess_dump3d(env)
    p = bh.Problem3D(env)
    v = p.view
    got = []
    for nodes, nx, ny in ((v.bty, v.bty_nx, v.bty_ny), (v.ati, v.ati_nx, v.ati_ny)):
        for i in range(nx * ny):
            b = nodes[i]
            got += list(b.x) + list(b.n1) + list(b.n2)
    nz = v.hex_nz
    got += [v.hex_cz[i] for i in range(v.hex_nx * v.hex_ny * (nz - 1))]
    got += [v.t_rcvr[i] for i in range(2 * v.Ntheta)]
    got += [v.beta[i] for i in range(v.Nbeta)]
    got += [v.Sx[i] for i in range(v.NSx)] + [v.Sy[i] for i in range(v.NSy)]
    got += [v.deltas, v.box_x, v.box_y, v.box_z, v.Dbeta]
    got = np.array(got)
    assert got.shape == ref.shape
    # the bottom grid's normals are only defined for cells (the last row /
    # column of nodes carries none): compare where the oracle holds a value
    ok = np.isfinite(ref)
    assert np.array_equal(got[ok], ref[ok])


def test_preprocess3d_rejects_unsupported():
    for rt in ("CR   3", "CC   3", "XB   3", "CS   3"):
        with pytest.raises(bh.BhcError):
            bh.Problem3D(free_env3d(run_type=rt))
    e = free_env3d()
    e.ssp_type = "Q"
    with pytest.raises(bh.BhcError):
        bh.Problem3D(e)


def test_abi_exports_3d_symbols():
    L = bh.lib()
    for name in ("bhc_preprocess3d", "bhc_problem3d_view", "bhc_problem3d_free", "bhc_host3d_last_error",
                 "bhc_field_run3d", "bhc_scale_pressure3d", "bhc_opcount3d", "bhc_ray_run3d", "bhc_write_ray3d",
                 "bhc_postprocess_arrivals3d", "bhc_write_arr3d", "bhc_ray_run3d_list"):
        assert hasattr(L, name)


def test_ray_mode3d_straight_rays_and_2d_plane():
    # isovelocity: every ray is straight along its launch direction
    env = free_env3d(nalpha=11, nbeta=8, run_type="R    3")
    pts, ns = oracle.ray_run3d(env, max_points=400)
    M = 400
    for k in range(env.nrays):
        n = ns[k]
        assert n >= 2
        x = np.array([list(pts[k * M + i].x) for i in range(n)])
        d = x[-1] - x[0]
        ia, ib = k % 11, k // 11
        a, b = np.radians(env.alpha_deg[ia]), np.radians(env.beta_deg[ib])
        u = np.array([np.cos(a) * np.cos(b), np.cos(a) * np.sin(b), np.sin(a)])
        assert np.allclose(d / np.linalg.norm(d), u, atol=1e-12)
    # depth-only Munk: the beta = 0 ray stays in its vertical plane and follows the 2D ray
    e3 = munk_env3d(nalpha=21, nbeta=3, beta=(-1.0, 1.0), run_type="R    3")
    pts3, ns3 = oracle.ray_run3d(e3, max_points=3000)
    e2 = munk_env(nbeams=21, run_type="R ", r_max_km=20.0)
    arr2, ns2, _ = oracle.ray_run(e2, max_points=3000)
    for ia in range(21):
        k = 1 * 21 + ia
        n3 = ns3[k]
        assert n3 == ns2[ia]
        x3 = np.array([list(pts3[k * 3000 + i].x) for i in range(n3)])
        x2 = np.array(arr2[ia, :n3]["x"])
        assert np.max(np.abs(x3[:, 1])) == 0.0
        assert np.max(np.abs(x3[:, 0] - x2[:, 0])) < 1e-6 and np.max(np.abs(x3[:, 2] - x2[:, 1])) < 1e-6


def test_write_ray3d(tmp_path):
    env = free_env3d(nalpha=5, nbeta=4, run_type="R    3")
    pts, ns = oracle.ray_run3d(env, max_points=400)
    prob = bh.Problem3D(env)
    path = str(tmp_path / "free3d.ray")
    bh.write_ray3d(path, prob, pts, ns, 400, title="free space 3D")
    lines = open(path).read().splitlines()
    assert lines[0].strip().strip("'") == "free space 3D"
    assert lines[2].split() == ["1", "1", "1"] and lines[3].split() == ["5", "4"]
    assert lines[6].strip().strip("'") == "xyz"
    # first ray: declination in radians, then count and bounces, then n points of 3 values
    assert abs(float(lines[7]) - np.radians(env.alpha_deg[0])) < 1e-12
    n, ntop, nbot = map(int, lines[8].split())
    assert n == ns[0] and len(lines[9].split()) == 3
    assert len(lines) == 7 + sum(2 + int(v) for v in ns)


def test_core_ray_mode3d_matches_oracle():
    """The product's 3D ray core on the host equals the oracle point for point
    (positions, tangents, p, q, torsion), FP64 and FP32."""
    H = harness()
    H.core_ray_run3d.restype = C.c_int
    H.core_ray_run3d.argtypes = [C.POINTER(_abi.Problem3D), C.POINTER(_abi.RayPt3D), _abi.c_int32_p, C.c_int]
    M = 3000
    for env in (munk_env3d(nalpha=41, nbeta=5, beta=(-30.0, 30.0), run_type="R    3"),
                c3_env(nalpha=11, nbeta=8, ntheta=4, nrz=11, nrr=21, nxy=41, run_type="R    3", precision=32)):
        po, no = oracle.ray_run3d(env, max_points=M)
        p = bh.Problem3D(env)
        n = p.nrays
        pg = (_abi.RayPt3D * (n * M))()
        ng = np.zeros(n, np.int32)
        assert H.core_ray_run3d(C.byref(p.view), C.cast(pg, C.POINTER(_abi.RayPt3D)),
                                ng.ctypes.data_as(_abi.c_int32_p), M) == 0
        assert np.array_equal(no, ng)
        ao = np.ctypeslib.as_array(po).reshape(n, M)
        ag = np.ctypeslib.as_array(pg).reshape(n, M)
        for k in range(n):
            m = min(no[k], M)
            for f in ("x", "t", "p", "q", "phi", "tau", "Amp"):
                assert np.array_equal(ao[k, :m][f], ag[k, :m][f]), (k, f)


GOLDEN_ARR3 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "example_analysis_3d.arr.txt")


def _ulps32(a, b):
    a, b = np.float32(a), np.float32(b)
    return abs(float(a) - float(b)) / float(np.spacing(np.float32(max(abs(a), abs(b)))))


@pytest.mark.parametrize("binary", [False, True])
def test_oracle_arrivals3d_reproduce_reference_fixture(tmp_path, binary):
    env = analysis3d_env()
    o = oracle.field_run3d(env, max_narr=10, allow_merging=1)
    prob = bh.Problem3D(env)
    narr, maxn = files.postprocess_arrivals3d(prob, o.arr, o.narr, 10)
    path = str(tmp_path / ("a3d.barr" if binary else "a3d.arr"))
    files.write_arr3d(path, prob, o.arr, narr, 10, maxn, binary=binary)
    got = files.read_arr3d(path, binary=binary)
    ref = files.read_arr3d(GOLDEN_ARR3)
    assert got["maxn"] == ref["maxn"] == [1]
    assert float(np.float32(got["freq0"])) == ref["freq0"]
    for k in ("Sx", "Sy", "Sz", "Rz", "theta"):
        assert np.array_equal(got[k], ref[k]), k
    assert np.allclose(got["Rr"], ref["Rr"], rtol=1e-7)
    assert len(got["arrivals"][0]) == len(ref["arrivals"][0]) == 15
    for g, r in zip(got["arrivals"][0], ref["arrivals"][0]):
        assert len(g) == len(r) == 1
        for f in ("phase_deg", "delay_im", "src_decl", "src_azim", "rcvr_decl", "rcvr_azim", "ntop", "nbot"):
            assert g[f][0] == r[f][0], f
        assert _ulps32(g["a"][0], r["a"][0]) <= 2.0
        assert _ulps32(g["delay_re"][0], r["delay_re"][0]) <= 2.0


def test_oracle_arrivals3d_fp32_close_to_fixture():
    """Single precision: the same arrivals within the FP32 build's accumulated
    step error (up to about 2e-4 relative over the 1129 steps of each ray)."""
    env = analysis3d_env(precision=32)
    o = oracle.field_run3d(env, max_narr=10, allow_merging=1)
    ref = files.read_arr3d(GOLDEN_ARR3)["arrivals"][0]
    A = np.ctypeslib.as_array(o.arr).reshape(-1, 10)
    assert list(o.narr) == [1] * 15
    for ir in range(15):
        a, r = A[ir, 0], ref[ir][0]
        assert abs(a["a"] - r["a"]) <= 5e-4 * r["a"]
        assert abs(a["delay"][0] - r["delay_re"]) <= 5e-4 * r["delay_re"]
        assert abs(a["SrcAzimAngle"] - r["src_azim"]) < 1e-4 and a["NBotBnc"] == 0 and a["NTopBnc"] == 0
        assert abs(a["RcvrAzimAngle"] - r["rcvr_azim"]) < 1e-3


def test_oracle_eigen_hits3d_bracket_receivers():
    """Eigen hits on the isovelocity case: the central ray (declination 0,
    bearing 34 deg = the receivers' bearing) hits every receiver, from the
    step whose segment brackets the receiver's range."""
    env = analysis3d_env(run_type="EG   3")
    o = oracle.field_run3d(env, eigen_memsize=1000)
    n = int(o.neigen[0])
    hits = np.ctypeslib.as_array(o.hits)[:n]
    central = hits[(hits["ialpha"] == 1) & (hits["ibeta"] == 0)]
    assert sorted(central["ir"].tolist()) == list(range(15))
    assert np.all(hits["itheta"] == 0) and np.all(hits["iz"] == 0)
    x = env.Rr_km * 1000.0
    ds = 1.7646
    for h in central:
        # the hit step's segment [is * ds, (is + 1) * ds] brackets the receiver range
        assert h["is_"] * ds <= x[h["ir"]] + 1e-6 and x[h["ir"]] <= (h["is_"] + 1) * ds + 1e-6


def test_arr3d_writer_roundtrip(tmp_path):
    """3D .arr ASCII and binary layouts read back to what was written."""
    env = munk_env3d(nalpha=31, nbeta=4, beta=(-20.0, 20.0), ntheta=2, nrz=5, nrr=11, run_type="AB   3")
    o = oracle.field_run3d(env, max_narr=8, allow_merging=1)
    prob = bh.Problem3D(env)
    narr, maxn = files.postprocess_arrivals3d(prob, o.arr, o.narr, 8)
    A = np.ctypeslib.as_array(o.arr).reshape(-1, 8)
    assert narr.sum() > 0
    for binary in (False, True):
        path = str(tmp_path / ("m.barr" if binary else "m.arr"))
        files.write_arr3d(path, prob, o.arr, narr, 8, maxn, binary=binary)
        got = files.read_arr3d(path, binary=binary)
        assert got["maxn"] == list(maxn)
        assert np.allclose(got["theta"], env.theta_deg)
        for b, arrs in enumerate(got["arrivals"][0]):
            assert len(arrs) == narr[b]
            for k in range(narr[b]):
                a = A[b, k]
                assert np.float32(arrs["a"][k]) == a["a"]
                assert np.float32(arrs["delay_re"][k]) == a["delay"][0]
                assert np.float32(arrs["src_azim"][k]) == a["SrcAzimAngle"]
                assert np.float32(arrs["rcvr_azim"][k]) == a["RcvrAzimAngle"]
                assert arrs["nbot"][k] == a["NBotBnc"]


def test_depth_window_guess_equals_search():
    """The scatter's depth window (index guessed from the mean spacing, then
    verified) equals the plain search on uniform, perturbed and irregular
    receiver grids, for bounds inside, outside, on and between the depths."""
    H = harness()
    H.core_depth_window.restype = None
    H.core_depth_window.argtypes = [_abi.c_float_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]
    rng = np.random.default_rng(5)
    grids = [np.linspace(0.0, 5000.0, 501).astype(np.float32),
             np.linspace(10.0, 20.0, 2).astype(np.float32),
             np.sort(rng.uniform(0, 4000, 300)).astype(np.float32),
             (np.linspace(0.0, 1000.0, 101) + rng.normal(0, 3.0, 101)).astype(np.float32),
             np.geomspace(1.0, 5000.0, 64).astype(np.float32)]
    for rz in grids:
        rz = np.unique(rz)
        n = len(rz)
        rdz = (n - 1) / (float(rz[-1]) - float(rz[0]))
        pts = np.concatenate([rz.astype(np.float64), rz + 1e-3, rz - 1e-3, rng.uniform(-100, 5100, 200),
                              [-1e30, 1e30, np.nan]])
        for zlo in pts[::3]:
            for zhi in (zlo, zlo + 7.5, zlo + 300.0):
                a, b = C.c_int(), C.c_int()
                H.core_depth_window(rz.ctypes.data_as(_abi.c_float_p), n, rdz, zlo, zhi, C.byref(a), C.byref(b))
                lo = int(np.searchsorted(rz.astype(np.float64), zlo, "left")) if zlo == zlo else 0
                hi = int(np.searchsorted(rz.astype(np.float64), zhi, "right")) if zhi == zhi else 0
                assert (a.value, b.value) == (max(lo - 1, 0), min(hi, n - 1)), (n, zlo, zhi)


def test_oracle_eigen_retrace3d():
    """Eigenray re-trace in 3D: each hit's ray is the ray run type's ray cut
    after the hit's step (a prefix of it, point for point), and its last
    segment brackets the receiver's range."""
    env = analysis3d_env(run_type="EG   3")
    o = oracle.field_run3d(env, eigen_memsize=1000)
    n = int(o.neigen[0])
    M = 2000
    pts, ns = oracle.ray_run3d_list(env, o.hits, n, max_points=M)
    full_env = analysis3d_env(run_type="R    3")
    fp, fn = oracle.ray_run3d(full_env, max_points=M)
    hits = np.ctypeslib.as_array(o.hits)[:n]
    A = np.ctypeslib.as_array(pts).reshape(n, M)
    F = np.ctypeslib.as_array(fp).reshape(len(fn), M)
    rr = env.Rr_km * 1000.0
    nb = len(env.beta_deg)
    for k, h in enumerate(hits):
        assert ns[k] == h["is_"] + 2
        job = h["ibeta"] * len(env.alpha_deg) + h["ialpha"]
        m = ns[k]
        for f in ("x", "t", "p", "q", "tau"):
            assert np.array_equal(A[k, :m][f], F[job, :m][f]), (k, f)
        ra = np.hypot(A[k, h["is_"]]["x"][0], A[k, h["is_"]]["x"][1])
        rb = np.hypot(A[k, h["is_"] + 1]["x"][0], A[k, h["is_"] + 1]["x"][1])
        assert min(ra, rb) <= rr[h["ir"]] <= max(ra, rb) + 1e-9
    assert nb == 2


@pytest.mark.parametrize("run_type,tol_med,tol_p95", [("Cg   3", 0.05, 0.2), ("Ib   3", 0.1, 0.25)])
def test_free_space_spherical_spreading_raycen(run_type, tol_med, tol_p95):
    """Ray-centred 3D beams (Step_InfluenceGeoRayCen, R3D) in free space:
    TL = 20 log10 R on every bearing, to the Cartesian beams' bars. This fixes
    the ray-normal receiver-line intercepts (m, n, r), the range brackets and
    the interpolation of n1, n2 between the segment's ends."""
    env = free_env3d(nalpha=181, nbeta=72, run_type=run_type)
    o = oracle.field_run3d(env)
    f = oracle.scale_pressure3d(env, o.field)[0, 0, 0]
    d = _free_tl_error(env, f)
    assert o.err == 0
    assert abs(np.median(d)) < tol_med, np.median(d)
    assert np.percentile(np.abs(d), 95) < tol_p95, np.percentile(np.abs(d), 95)


def test_raycen_hat_equals_cartesian_hat_munk():
    """On a depth-only Munk environment ray-centred and Cartesian hat beams
    give the same TL (they differ only in how receivers are bracketed)."""
    a = oracle.field_run3d(munk_env3d(nalpha=151, nbeta=7, run_type="Cg   3"))
    b = oracle.field_run3d(munk_env3d(nalpha=151, nbeta=7, run_type="CG   3"))
    ta, tb = tl_db(a.field)[0, 0, 0, 0], tl_db(b.field)[0, 0, 0, 0]
    m = tb < 100.0
    assert np.median(np.abs(ta - tb)[m]) < 0.02
    assert np.percentile(np.abs(ta - tb)[m], 90) < 0.2
