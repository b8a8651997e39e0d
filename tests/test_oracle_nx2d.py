"""The Nx2D oracle (oracle/bellhop3d.c, O3D = true, R3D = false) on the CPU.

Nx2D traces 2D rays (range along a bearing, depth) through a 3D environment:
the step limits, boundary cells and SSP come from the 3D ocean point
RayToOceanX(x), the SSP gradient is projected on the bearing, and the step
end is mapped back by OceanToRayX with its one-ulp snapping
(src/common_run.hpp:221-317, src/step.hpp:716-765, src/ssp.hpp:449-462,
src/reflect.hpp:94-140). What pins the restatement:

  * on a depth-only environment with flat boundaries, the bearing-0 plane
    through a source at the origin is the 2D problem: steps and bounces equal
    the independently pinned 2D oracle's on every ray, TL to 1e-2 dB (the
    halfspace coefficient takes the 3D formula, which agrees with the 2D one
    to rounding);
  * rotation: bearings whose plane does not cross the diagonal of the flat
    default boundary cell give the bearing-0 field (TL to 1e-3 dB) -- this
    fixes the projections of x, t and the SSP gradient; a bearing that does
    cross the diagonal gets one extra step end there (TriDiagCrossing) and a
    field within a few hundredths of a dB;
  * a bottom sloping along x (resp. y), traced at bearing 0 (resp. 90 deg),
    equals the 2D oracle's wedge -- this fixes the projected boundary normal,
    the recomputed boundary tangent of the reflection and the halfspace;
  * a hexahedral SSP with no lateral variation equals the c-linear profile;
  * ray mode: the points of the bearing-0 rays equal the 2D oracle's ray
    points; every ray's ocean points lie on its bearing's vertical plane;
    rays that end on an oblique beam-box face end at or just past the face in
    ocean coordinates (OceanToRayX snapping);
  * arrivals and eigen hits carry the bearing: azimuths equal the bearing,
    itheta = ibeta, and the bearing-0 arrivals equal the 2D oracle's.
"""
import ctypes as C
import warnings

import numpy as np
import pytest

from bellhopcuda_amd import _abi
from bellhopcuda_amd.env import munk_env, Boundary
from bellhopcuda_amd.env3d import munk_env3d, Grid3D
import oracle
from parity import tl_db


def _nx(env3):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r = oracle.field_run3d(env3)
        return r, tl_db(oracle.scale_pressure3d(env3, r.field))


def _tl2(env2):
    r = oracle.field_run(env2)
    return r, tl_db(oracle.scale_pressure(env2, r.field))


@pytest.mark.parametrize("rt", ["CG", "IB", "Cg", "SB"])
def test_nx2d_bearing0_equals_2d(rt):
    e3 = munk_env3d(nalpha=401, nbeta=1, ntheta=1, theta0=0.0, nrz=51, nrr=201, run_type=rt + "   2")
    e2 = munk_env(nbeams=401, nrz=51, nrr=201, run_type=rt, r_max_km=20.0)
    o3, t3 = _nx(e3)
    o2, t2 = _tl2(e2)
    assert o3.err == 0 and o2.err == 0
    assert np.array_equal(o3.nsteps, o2.nsteps)
    assert np.array_equal(o3.ntop, o2.ntop) and np.array_equal(o3.nbot, o2.nbot)
    # the ray's range is the ocean x (source at the origin, bearing 0)
    assert np.allclose(o3.xend[:, 0], o2.xend[:, 0], rtol=1e-12, atol=1e-9)
    assert np.allclose(o3.xend[:, 2], o2.xend[:, 1], rtol=1e-12, atol=1e-9)
    assert np.all(o3.xend[:, 1] == 0.0)
    t3 = t3[0, 0, 0, 0]
    m = (t2[0] < 100) & (t3 < 100)
    assert m.sum() > 5000
    assert np.max(np.abs(t3 - t2[0])[m]) < 1e-2


def test_nx2d_rotation_invariance():
    e3 = munk_env3d(nalpha=201, nbeta=1, ntheta=1, nrz=41, nrr=101, run_type="CB   2")
    # source at (3, -2) km: the planes at 37, 270 and 315 deg stay below the
    # default cell's diagonal y = x; 90 and 143 deg cross it
    e3.theta_deg = np.array([0.0, 37.0, 270.0, 315.0, 90.0, 143.0])
    e3.Sx_km, e3.Sy_km = np.array([3.0]), np.array([-2.0])
    e3.box_x_km = e3.box_y_km = 30.0
    o, t = _nx(e3)
    assert o.err == 0
    t = t[0, 0, 0]
    for k in range(1, 6):
        m = (t[0] < 90) & (t[k] < 90)
        d = np.abs(t[k] - t[0])[m]
        assert m.sum() > 2000
        if k <= 3:
            assert np.percentile(d, 99.5) < 1e-3, (e3.theta_deg[k], np.percentile(d, 99.5))
        else:
            assert np.median(d) < 0.01 and np.percentile(d, 99) < 0.5, e3.theta_deg[k]
    # every ray's end point lies in its bearing's vertical plane
    xe = o.xend.reshape(6, 201, 3)
    for k, th in enumerate(e3.theta_deg):
        b = np.deg2rad(th)
        rx, ry = xe[k, :, 0] - 3000.0, xe[k, :, 1] + 2000.0
        assert np.max(np.abs(rx * np.sin(b) - ry * np.cos(b))) < 1e-6 * 30000.0


def _wedge_pair(axis):
    """A bottom plane from 3000 m at -1 km to 1200 m at 30 km along `axis`,
    as a 2D bathymetry and as a 3D grid (2 nodes across, spanning -1 ..
    1000 km so that the cell diagonal meets the ray's plane behind the
    source)."""
    zs = np.array([3000.0, 1200.0])
    e2 = munk_env(nbeams=301, nrz=41, nrr=101, run_type="CB", r_max_km=20.0)
    e2.bot.bc = "A"
    e2.bty = Boundary(r_km=np.array([-1.0, 30.0]), z=zs)
    e2.box_r_km = 25.0
    e3 = munk_env3d(nalpha=301, nbeta=1, ntheta=1, nrz=41, nrr=101, run_type="CB   2")
    e3.box_x_km = e3.box_y_km = 25.0
    along, across = np.array([-1.0, 30.0]), np.array([-1.0, 1000.0])
    if axis == 0:
        e3.bty = Grid3D(x_km=along, y_km=across, z=np.array([[zs[0], zs[0]], [zs[1], zs[1]]]))
        e3.theta_deg = np.array([0.0])
    else:
        e3.bty = Grid3D(x_km=across, y_km=along, z=np.array([[zs[0], zs[1]], [zs[0], zs[1]]]))
        e3.theta_deg = np.array([90.0])
    return e2, e3


@pytest.mark.parametrize("axis", [0, 1])
def test_nx2d_wedge_equals_2d_slope(axis):
    e2, e3 = _wedge_pair(axis)
    o2, t2 = _tl2(e2)
    o3, t3 = _nx(e3)
    assert o2.err == 0 and o3.err == 0
    assert o2.nbot.sum() > 300
    assert np.mean(o3.nsteps == o2.nsteps) > 0.99
    assert np.array_equal(o3.nbot, o2.nbot) and np.array_equal(o3.ntop, o2.ntop)
    t3 = t3[0, 0, 0, 0]
    m = (t2[0] < 90) & (t3 < 90)
    assert m.sum() > 1500
    d = np.abs(t3 - t2[0])[m]
    assert np.median(d) < 1e-3 and np.percentile(d, 99) < 0.05, (np.median(d), np.percentile(d, 99))


def test_nx2d_hexahedral_laterally_uniform_equals_clinear():
    e1 = munk_env3d(nalpha=151, nbeta=1, ntheta=1, nrz=21, nrr=61, run_type="CB   2", bot="V")
    e2 = munk_env3d(nalpha=151, nbeta=1, ntheta=1, nrz=21, nrr=61, run_type="CB   2", bot="V")
    e1.theta_deg = e2.theta_deg = np.array([0.0, 300.0])
    e2.ssp_type = "H"
    e2.hex_x_km = np.array([-30.0, 30.0])
    e2.hex_y_km = np.array([-30.0, 30.0])
    e2.hex_z = e1.ssp_z.copy()
    e2.hex_c = np.broadcast_to(e1.ssp_c, (2, 2, len(e1.ssp_z))).copy()
    a, ta = _nx(e1)
    b, tb = _nx(e2)
    assert a.err == 0 and b.err == 0
    assert np.array_equal(a.nsteps, b.nsteps)
    m = ta < 100
    assert np.max(np.abs(ta - tb)[m]) < 1e-3


def test_nx2d_ray_mode_matches_2d_and_stays_in_plane():
    e3 = munk_env3d(nalpha=21, nbeta=1, nrz=5, nrr=5, run_type="R    2")
    e3.beta_deg = np.array([0.0, 37.0, 200.0])
    e2 = munk_env(nbeams=21, nrz=5, nrr=5, run_type="R", r_max_km=20.0)
    M = 3000
    p3, n3, org = oracle.ray_runnx2d(e3, max_points=M)
    # 2D oracle ray run in job order
    out = _abi.RayOut()
    p2 = (_abi.RayPt2D * (21 * M))()
    ns2 = np.zeros(21, np.int32)
    out.points = C.cast(p2, C.POINTER(_abi.RayPt2D))
    out.nsteps = ns2.ctypes.data_as(_abi.c_int32_p)
    out.max_points_per_ray = M
    ec, keep = e2.to_c()
    assert oracle.lib().oracle_ray_run2d_d(C.byref(ec), C.byref(out), None, None, None, 21) == 0
    assert np.array_equal(n3[:21], ns2)
    A3 = np.ctypeslib.as_array(p3).reshape(3, 21, M)
    A2 = np.ctypeslib.as_array(p2).reshape(21, M)
    for k in range(21):
        n = ns2[k]
        x3 = np.array([(A3[0, k, i]["x"][0], A3[0, k, i]["x"][1]) for i in range(n)])
        x2 = np.array([(A2[k, i]["x"][0], A2[k, i]["x"][1]) for i in range(n)])
        assert np.allclose(x3, x2, rtol=1e-12, atol=1e-8)
    # origins and planes
    assert np.allclose(org[:, 0], 0.0) and np.allclose(org[:, 1], 0.0)
    for j, b in enumerate(np.deg2rad(e3.beta_deg)):
        assert np.allclose(org[21 * j:21 * (j + 1), 2], np.cos(b)) and np.allclose(org[21 * j:21 * (j + 1), 3],
                                                                                 np.sin(b))
    # the same declination traces the same (range, depth) path on every
    # bearing of a laterally uniform ocean until the box
    for k in (3, 10, 17):
        n = min(n3[k], n3[21 + k]) - 2
        xa = np.array([A3[0, k, i]["x"][0] for i in range(n)])
        xb = np.array([A3[1, k, i]["x"][0] for i in range(n)])
        assert np.allclose(xa, xb, rtol=1e-9, atol=1e-6)


def test_nx2d_rays_end_on_oblique_box_face():
    """Rays leaving through the x face of the box at bearing 37 deg: the
    snapped last step puts the ocean point at or past the face (by at most
    the minimum step), never short of it."""
    e3 = munk_env3d(nalpha=41, nbeta=1, nrz=5, nrr=5, run_type="R    2", bot="R", r_max_km=10.0)
    e3.beta_deg = np.array([37.0])
    e3.box_x_km = 10.0
    e3.box_y_km = 50.0
    M = 4000
    p3, n3, org = oracle.ray_runnx2d(e3, max_points=M)
    A = np.ctypeslib.as_array(p3).reshape(41, M)
    cb, sb = org[0, 2], org[0, 3]
    hits = 0
    for k in range(41):
        last = A[k, n3[k] - 1]["x"]
        xo = org[k, 0] + last[0] * cb
        if last[0] * cb > 9000.0:
            hits += 1
            assert xo >= 10000.0 - 1e-9 and xo < 10000.0 + 1.0, (k, xo)
    assert hits > 20


def test_nx2d_arrivals_and_eigen_carry_the_bearing():
    e3 = munk_env3d(nalpha=101, nbeta=1, ntheta=1, nrz=5, nrr=11, run_type="A    2", r_max_km=10.0)
    e3.theta_deg = np.array([0.0, 25.0])
    e2 = munk_env(nbeams=101, nrz=5, nrr=11, run_type="A", r_max_km=10.0)
    a3 = oracle.field_run3d(e3, max_narr=20, allow_merging=1)
    a2 = oracle.field_run(e2, max_narr=20, allow_merging=1)
    nr = 5 * 11
    n3 = a3.narr.reshape(2, nr)
    assert np.array_equal(n3[0], a2.narr)
    arr3 = np.ctypeslib.as_array(a3.arr).reshape(2, nr, 20)
    arr2 = np.ctypeslib.as_array(a2.arr).reshape(nr, 20)
    for r in range(nr):
        for j in range(a2.narr[r]):
            x, y = arr3[0, r, j], arr2[r, j]
            assert abs(x["a"] - y["a"]) <= 1e-5 * abs(y["a"]) and abs(x["delay"][0] - y["delay"][0]) < 1e-6
            assert x["SrcAzimAngle"] == 0.0 and x["RcvrAzimAngle"] == 0.0
        for j in range(n3[1, r]):
            z = arr3[1, r, j]
            assert abs(z["SrcAzimAngle"] - 25.0) < 1e-4 and z["SrcAzimAngle"] == z["RcvrAzimAngle"]
    e3.run_type = "E    2"
    h = oracle.field_run3d(e3, eigen_memsize=20000)
    nh = int(h.neigen[0])
    assert nh > 100
    hits = np.ctypeslib.as_array(h.hits)[:min(nh, 20000)]
    assert set(np.unique(hits["itheta"])) == {0, 1}
    assert np.array_equal(hits["itheta"], hits["ibeta"])
    assert np.all(hits["isx"] == 0) and np.all(hits["isy"] == 0) and np.all(hits["isz"] == 0)


# ---------------------------------------------------------------------------
# The product's Nx2D core (bellhopcuda_amd/csrc/nx2d.h with the 2D influence
# descriptors and receiver loop), run serially on the CPU by
# tests/harness/core_cpu.cpp, against the oracle: bit for bit on trajectories,
# bounces, end points and contribution counts.
# ---------------------------------------------------------------------------
import bellhopcuda_amd as bh
from bellhopcuda_amd.env3d import c3_env
from test_oracle3d import harness


class _Res:
    pass


def core_runnx2d(env):
    p = bh.Problem3D(env)
    assert p.nx2d
    H = harness()
    if not getattr(H, "_nx", False):
        H.core_field_runnx2d.restype = C.c_int
        H.core_field_runnx2d.argtypes = [C.POINTER(_abi.Problem3D), C.POINTER(_abi.FieldOut)]
        H._nx = True
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
    assert H.core_field_runnx2d(C.byref(p.view), C.byref(out)) == 0
    r.total_contrib, r.total_steps, r.err, r.warn = out.total_contrib, out.total_steps, out.err, out.warn
    r.problem = p
    return r


def _rotated_munk(precision=64, rt="CB   2"):
    e = munk_env3d(nalpha=151, nbeta=1, ntheta=1, nrz=31, nrr=81, run_type=rt, precision=precision)
    e.theta_deg = np.array([0.0, 37.0, 90.0, 143.0, 200.0, 315.0])
    e.Sx_km, e.Sy_km = np.array([3.0]), np.array([-2.0])
    e.box_x_km = e.box_y_km = 30.0
    return e


NX_CASES = {
    "munk_hat_bearing0": lambda: munk_env3d(nalpha=201, nbeta=1, ntheta=1, nrz=31, nrr=81, run_type="CG   2"),
    "munk_gauss_rotated": lambda: _rotated_munk(),
    "munk_gauss_rotated_f32": lambda: _rotated_munk(precision=32),
    "munk_raycen_hat_rotated": lambda: _rotated_munk(rt="Cg   2"),
    "munk_semicoherent_n2linear": lambda: munk_env3d(nalpha=151, nbeta=1, ntheta=1, nrz=31, nrr=81,
                                                     run_type="SB   2", ssp_type="N"),
    "wedge_x": lambda: _wedge_pair(0)[1],
    "wedge_y": lambda: _wedge_pair(1)[1],
    "shelf_canyon_hexahedral": lambda: c3_env(nalpha=61, nbeta=6, ntheta=6, nrz=21, nrr=51, nxy=41,
                                              run_type="CB   2"),
    "shelf_canyon_hexahedral_f32": lambda: c3_env(nalpha=61, nbeta=6, ntheta=6, nrz=21, nrr=51, nxy=41,
                                                  run_type="CB   2", precision=32),
    "shelf_canyon_raycen_incoherent": lambda: c3_env(nalpha=41, nbeta=6, ntheta=6, nrz=21, nrr=51, nxy=41,
                                                     run_type="Ib   2"),
}


@pytest.mark.parametrize("name", list(NX_CASES))
def test_core_nx2d_matches_oracle(name):
    env = NX_CASES[name]()
    o = oracle.field_run3d(env)
    g = core_runnx2d(env)
    assert np.array_equal(g.nsteps, o.nsteps)
    assert np.array_equal(g.ntop, o.ntop) and np.array_equal(g.nbot, o.nbot)
    assert np.array_equal(g.xend, o.xend)
    assert g.total_contrib == o.total_contrib
    assert g.err == o.err and g.warn == o.warn
    to, tg = tl_db(o.field), tl_db(g.field)
    m = to < 120
    assert m.sum() > 100
    assert np.max(np.abs(to - tg)[m]) < 1e-3


def test_shelf_canyon_nx2d_exercises_snapping_and_diagonals():
    """The sloping, canyoned grid makes the Nx2D rays cross cell edges and
    diagonals off the bearing axes, so the OceanToRayX snap runs for x, y and
    diagonal step ends; the rays still end in their bearing planes."""
    env = NX_CASES["shelf_canyon_hexahedral"]()
    o = oracle.field_run3d(env)
    assert o.err == 0
    assert o.nbot.sum() > 0
    xe = o.xend.reshape(6, 61, 3)
    for k, th in enumerate(env.theta_deg):
        b = np.deg2rad(th)
        assert np.max(np.abs(xe[k, :, 0] * np.sin(b) - xe[k, :, 1] * np.cos(b))) < 1e-6 * 40000.0


@pytest.mark.parametrize("precision", [64, 32])
def test_preprocess_nx2d_beams_on_receiver_bearings(precision):
    env = c3_env(nalpha=21, nbeta=12, ntheta=6, nrz=11, nrr=21, nxy=31, precision=precision, run_type="CB   2")
    p = bh.Problem3D(env)
    v = p.view
    assert p.nx2d and v.Nbeta == 6
    ref = oracle.preprocess_dump3d(env)
    got = np.array([v.beta[i] for i in range(v.Nbeta)])
    # the dump holds nodes, hexahedral cz, bearings, beta, Sx, Sy, ...: locate beta after the bearings
    nb = (v.bty_nx * v.bty_ny + v.ati_nx * v.ati_ny) * 9 + v.hex_nx * v.hex_ny * (v.hex_nz - 1) + 2 * v.Ntheta
    assert np.array_equal(ref[nb:nb + 6], got)
    dr = np.float64(np.pi) / 180.0 if precision == 64 else np.float32(np.pi) / np.float32(180.0)
    assert np.allclose(got, np.asarray(env.theta_deg) * dr, rtol=1e-6)
    # a ray run keeps the beta fan
    env.run_type = "R    2"
    assert bh.Problem3D(env).view.Nbeta == 12


def test_scale_pressure_nx2d_matches_oracle():
    for rt in ("CB   2", "IB   2"):
        env = munk_env3d(nalpha=51, nbeta=1, ntheta=1, nrz=11, nrr=21, run_type=rt)
        env.theta_deg = np.array([0.0, 45.0])
        p = bh.Problem3D(env)
        rng = np.random.default_rng(3)
        f = (rng.random(p.field_shape) + 1j * rng.random(p.field_shape)).astype(np.complex64)
        a = p.scale_pressure(f)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            b = oracle.scale_pressure3d(env, f)
        assert np.array_equal(a, b)


def test_abi_exports_nx2d_symbols():
    L = bh.lib()
    for name in ("bhc_field_runnx2d", "bhc_field_runnx2d_async", "bhc_ray_runnx2d"):
        assert hasattr(L, name)
    with pytest.raises(bh.BhcError):
        bh.ray_runnx2d(munk_env3d(nalpha=3, nbeta=1, run_type="R    3"))
