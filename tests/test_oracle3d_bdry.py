"""3D boundary grids beyond flat triangles: curvilinear ('C') grids and
bottom provinces ('L'), in the 3D and Nx2D oracles, the product's
preprocessing and the product's core on the CPU.

What pins the oracle's restatement (oracle/bellhop3d.c bdry3_curvature,
refl_bdry3; src/module/boundary.hpp:152-197, 527-537, 650-830;
src/trace.hpp:318-345; src/boundary.hpp:134-140):
  * curvature formulas against closed forms on a parabolic trough
    z = z0 + a x^2: the node normals are (-2 a x, 0, 1) exactly (centred
    differences are exact on a quadratic), z_xx = 2a / |(-2ax, 0, 1)|, and
    kappa_xx, the change of the normal's angle over the chord, is the
    trough's curvature 2a / (1 + 4a^2x^2)^(3/2) to second order in the
    spacing;
  * a planar tilted grid read as curvilinear traces as the same grid read as
    flat triangles (node normals equal the triangles' normals, no curvature);
  * Nx2D on a trough along the bearing reproduces the 2D oracle's curvilinear
    bottom of the same shape (each restates its own curvature formula:
    agreement to a fraction of a dB);
  * provinces: one province holding the halfspace's own values is the
    single-set grid (same steps, end points and contributions); a bottom split into two provinces along x,
    traced in Nx2D on the bearing-0 plane, is the 2D oracle's long-format
    bottom with the same per-segment geoacoustics.
The product's preprocessing equals the oracle's value for value, and the
product's 3D and Nx2D cores equal the oracle bit for bit on these grids.
"""
import ctypes as C
import warnings

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd import _abi
from bellhopcuda_amd.env import munk_env, Boundary
from bellhopcuda_amd.env3d import munk_env3d, Grid3D
import oracle
from parity import tl_db
from test_oracle3d import core_run3d
from test_oracle_nx2d import core_runnx2d


def _tl3(env):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r = oracle.field_run3d(env)
        return r, tl_db(oracle.scale_pressure3d(env, r.field))


def _trough(a=2.0e-6, z0=3000.0, nx=41, ny=5, along=0, kind="CS", x0=-5.0, x1=35.0):
    """Bottom z = z0 + a (x - 15 km)^2 along `along` (0: x, 1: y), constant
    across; x, y in km, depth in m."""
    u = np.linspace(x0, x1, nx)
    v = np.linspace(-1.0, 1000.0, ny)
    zu = z0 + a * ((u - 15.0) * 1000.0) ** 2
    if along == 0:
        return Grid3D(x_km=u, y_km=v, z=np.repeat(zu[:, None], ny, axis=1), type=kind), zu
    return Grid3D(x_km=v, y_km=u, z=np.repeat(zu[None, :], ny, axis=0), type=kind), zu


def test_curvature_closed_forms_on_a_parabolic_trough():
    a = 2.0e-6
    g, zu = _trough(a=a)
    env = munk_env3d(nalpha=11, nbeta=1, ntheta=1, nrz=5, nrr=5, run_type="CB   2")
    env.bty = g
    p = bh.Problem3D(env)
    v = p.view
    nx, ny = v.bty_nx, v.bty_ny
    x = np.array([v.bty[i * ny].x[0] for i in range(nx)]) - 15000.0
    nodes = [v.bty[i * ny + 2] for i in range(nx)]
    Nu = np.stack([np.array([-2 * a * xi, 0.0, 1.0]) for xi in x])
    Nn = Nu / np.linalg.norm(Nu, axis=1)[:, None]
    got = np.array([list(n.Noden) for n in nodes])
    assert np.allclose(got[1:-1], Nn[1:-1], rtol=0, atol=1e-12)
    zxx = np.array([n.z_xx for n in nodes])[1:-2]
    assert np.allclose(zxx, 2 * a / np.linalg.norm(Nu[1:-2], axis=1), rtol=1e-9)
    # kappa_xx: change of the normal angle over the chord, against the curvature at the cell's midpoint
    xm = 0.5 * (x[1:-2] + x[2:-1])
    kap = 2 * a / (1 + (2 * a * xm) ** 2) ** 1.5
    kxx = np.array([n.kappa_xx for n in nodes])[1:-2]
    assert np.allclose(kxx, kap, rtol=2e-3)
    # z_yy, kappa_yy vanish: no curvature across
    assert np.max(np.abs([n.z_yy for n in nodes[1:-2]])) < 1e-15
    assert np.max(np.abs([n.kappa_yy for n in nodes[1:-2]])) < 1e-12


@pytest.mark.parametrize("precision", [64, 32])
def test_preprocess_curvilinear_and_provinces_matches_oracle(precision):
    env = munk_env3d(nalpha=11, nbeta=3, ntheta=2, nrz=5, nrr=5, precision=precision)
    g, _ = _trough(nx=21, ny=7, kind="CL")
    g.province = np.where(np.arange(21)[:, None] < 10, 1, 2) * np.ones((1, 7), np.int32)
    g.prov = np.array([[1600.0, 0.0, 1.8, 0.8, 0.0], [1750.0, 300.0, 2.0, 0.5, 0.2]])
    env.bty = g
    ref = oracle.preprocess_dump3d(env)
    p = bh.Problem3D(env)
    v = p.view
    got = []
    for nodes, nx, ny in ((v.bty, v.bty_nx, v.bty_ny), (v.ati, v.ati_nx, v.ati_ny)):
        for i in range(nx * ny):
            b = nodes[i]
            got += list(b.x) + list(b.n1) + list(b.n2)
    for i in range(v.bty_nx * v.bty_ny):
        b = v.bty[i]
        got += list(b.Noden) + [b.z_xx, b.z_xy, b.z_yy, b.kappa_xx, b.kappa_xy, b.kappa_yy, b.province]
    got += [v.bty_prov[i] for i in range(5 * v.bty_nprov)]
    got = np.array(got)
    ref = ref[:len(got)]
    ok = np.isfinite(ref)
    assert np.array_equal(got[ok], ref[ok])
    assert bytes(v.bty_type) == b"CL" and v.bty_nprov == 2


def test_planar_grid_curvilinear_equals_flat_triangles():
    x = np.linspace(-30.0, 30.0, 13)
    y = np.linspace(-30.0, 30.0, 9)
    X, Y = np.meshgrid(x, y, indexing="ij")
    z = 4000.0 + 20.0 * X + 8.0 * Y
    e1 = munk_env3d(nalpha=101, nbeta=5, beta=(-10.0, 10.0), nrz=21, nrr=41, run_type="CG   3")
    e2 = munk_env3d(nalpha=101, nbeta=5, beta=(-10.0, 10.0), nrz=21, nrr=41, run_type="CG   3")
    e1.bty = Grid3D(x_km=x, y_km=y, z=z, type="RS")
    e2.bty = Grid3D(x_km=x, y_km=y, z=z, type="CS")
    a, ta = _tl3(e1)
    b, tb = _tl3(e2)
    assert a.err == 0 and b.err == 0 and a.nbot.sum() > 100
    assert np.array_equal(a.nsteps, b.nsteps) and np.array_equal(a.nbot, b.nbot)
    m = ta < 100
    assert np.max(np.abs(ta - tb)[m]) < 1e-3


@pytest.mark.parametrize("rt", ["CB", "CG"])
def test_nx2d_trough_matches_2d_curvilinear_bottom(rt):
    g, zu = _trough(a=1.0e-6, nx=161, x0=-5.0, x1=35.0)
    e3 = munk_env3d(nalpha=301, nbeta=1, ntheta=1, nrz=41, nrr=101, run_type=rt + "   2")
    e3.bty = g
    e3.box_x_km = e3.box_y_km = 30.0
    e2 = munk_env(nbeams=301, nrz=41, nrr=101, run_type=rt, r_max_km=20.0)
    e2.bty = Boundary(type="CS", r_km=np.asarray(g.x_km, float), z=zu)
    e2.box_r_km = 30.0
    o3, t3 = _tl3(e3)
    o2 = oracle.field_run(e2)
    t2 = tl_db(oracle.scale_pressure(e2, o2.field))
    assert o3.err == 0 and o2.err == 0 and o2.nbot.sum() > 150
    assert np.mean(o3.nbot == o2.nbot) > 0.99
    t3 = t3[0, 0, 0, 0]
    m = (t2[0] < 90) & (t3 < 90)
    d = np.abs(t3 - t2[0])[m]
    assert m.sum() > 1500
    assert np.median(d) < 0.02 and np.percentile(d, 95) < 0.3, (np.median(d), np.percentile(d, 95))


def test_single_province_equals_single_set_grid():
    x = np.linspace(-30.0, 30.0, 7)
    z = np.full((7, 7), 5000.0)
    base = munk_env3d(nalpha=101, nbeta=5, nrz=21, nrr=41, run_type="CB   3")
    base.bot.alphaR, base.bot.rho, base.bot.alphaI = 1600.0, 1.8, 0.8
    base.bty = Grid3D(x_km=x, y_km=x, z=z, type="RS")
    prov = munk_env3d(nalpha=101, nbeta=5, nrz=21, nrr=41, run_type="CB   3")
    prov.bty = Grid3D(x_km=x, y_km=x, z=z, type="RL", province=np.ones((7, 7), np.int32),
                      prov=np.array([[1600.0, 0.0, 1.8, 0.8, 0.0]]))
    a, ta = _tl3(base)
    b, tb = _tl3(prov)
    assert np.array_equal(a.nsteps, b.nsteps) and np.array_equal(a.xend, b.xend)
    assert a.total_contrib == b.total_contrib
    # the oracle's threads add to the field in a varying order: equal to rounding
    m = ta < 120
    assert np.max(np.abs(ta - tb)[m]) < 1e-4


def test_nx2d_provinces_match_2d_long_format_bottom():
    """Bottom at 4000 m split at x = 12 km into a fast and a slow province;
    the 2D oracle's long-format flat bottom carries the same values per
    segment (its cells take the lower node's geoacoustics)."""
    xs = np.array([-1.0, 12.0, 40.0])
    p1, p2 = [1550.0, 0.0, 1.5, 0.5, 0.0], [1900.0, 0.0, 2.1, 0.2, 0.0]
    e3 = munk_env3d(nalpha=301, nbeta=1, ntheta=1, nrz=41, nrr=101, run_type="CB   2")
    e3.bty = Grid3D(x_km=xs, y_km=np.array([-1.0, 1000.0]), z=np.full((3, 2), 4000.0), type="RL",
                    province=np.array([[1, 1], [2, 2], [2, 2]], np.int32), prov=np.array([p1, p2]))
    e3.box_x_km = e3.box_y_km = 30.0
    e2 = munk_env(nbeams=301, nrz=41, nrr=101, run_type="CB", r_max_km=20.0)
    e2.bty = Boundary(type="LL", r_km=xs, z=np.full(3, 4000.0),
                      alphaR=np.array([p1[0], p2[0], p2[0]]), betaR=np.zeros(3),
                      rho=np.array([p1[2], p2[2], p2[2]]), alphaI=np.array([p1[3], p2[3], p2[3]]),
                      betaI=np.zeros(3))
    e2.box_r_km = 30.0
    o3, t3 = _tl3(e3)
    o2 = oracle.field_run(e2)
    t2 = tl_db(oracle.scale_pressure(e2, o2.field))
    assert o3.err == 0 and o2.err == 0 and o2.nbot.sum() > 100
    assert np.array_equal(o3.nbot, o2.nbot)
    t3 = t3[0, 0, 0, 0]
    m = (t2[0] < 90) & (t3 < 90)
    assert m.sum() > 1500
    d = np.abs(t3 - t2[0])[m]
    assert np.percentile(d, 99) < 0.01, np.percentile(d, 99)
    # and the provinces matter: a single slow province differs
    e3.bty.province = np.ones((3, 2), np.int32)
    _, t1 = _tl3(e3)
    dd = np.abs(t1[0, 0, 0, 0] - t3)[m]
    assert np.mean(dd) > 0.1 and np.max(dd) > 1.0


CORE_CASES = {
    "trough_3d_curvilinear": lambda: _env3("CB   3", _trough(nx=41, ny=9, kind="CS")[0]),
    "trough_nx2d_curvilinear": lambda: _env3("CB   2", _trough(nx=41, ny=5, kind="CS")[0]),
    "provinces_3d": lambda: _env3("CB   3", _prov_grid()),
    "provinces_nx2d_curvilinear_f32": lambda: _env3("CB   2", _prov_grid("CL"), precision=32),
}


def _env3(rt, grid, precision=64):
    e = munk_env3d(nalpha=81, nbeta=5 if rt.endswith("3") else 1, beta=(-10.0, 10.0), ntheta=1, nrz=21, nrr=41,
                   run_type=rt, precision=precision)
    if rt.endswith("2"):
        e.theta_deg = np.array([0.0, 20.0])
    e.bty = grid
    e.box_x_km = e.box_y_km = 30.0
    return e


def _prov_grid(kind="RL"):
    x = np.linspace(-5.0, 35.0, 9)
    y = np.linspace(-20.0, 20.0, 5)
    X, Y = np.meshgrid(x, y, indexing="ij")
    z = 4000.0 - 15.0 * X + 3.0 * Y ** 2
    pv = np.where(X < 12.0, 1, 2).astype(np.int32)
    return Grid3D(x_km=x, y_km=y, z=z, type=kind, province=pv,
                  prov=np.array([[1550.0, 0.0, 1.5, 0.5, 0.0], [1900.0, 250.0, 2.1, 0.2, 0.1]]))


@pytest.mark.parametrize("name", list(CORE_CASES))
def test_core_matches_oracle_on_curvilinear_and_province_grids(name):
    env = CORE_CASES[name]()
    o = oracle.field_run3d(env)
    g = core_runnx2d(env) if env.nx2d else core_run3d(env)
    assert o.nbot.sum() > 0
    assert np.array_equal(g.nsteps, o.nsteps)
    assert np.array_equal(g.ntop, o.ntop) and np.array_equal(g.nbot, o.nbot)
    assert np.array_equal(g.xend, o.xend)
    assert g.total_contrib == o.total_contrib
    assert g.err == o.err and g.warn == o.warn
    to, tg = tl_db(o.field), tl_db(g.field)
    m = to < 120
    assert np.max(np.abs(to - tg)[m]) < 1e-3


def test_3d_curvature_matches_nx2d_on_a_trough():
    """A trough of 25 km radius along x, source on its axis: the 3D field on
    the bearing-0 plane (a narrow bearing fan) and the Nx2D field agree to a
    few hundredths of a dB with curvature ('C') as without ('R'), while the
    curvature itself moves the field by over a dB at a tenth of the receivers --
    so the 3D reflection's curvature matrix (z_xx in DMat) and the Nx2D
    projected curvature (kappa_xx) act alike."""
    t = {}
    for kind in ("CS", "RS"):
        g, _ = _trough(a=3e-5, z0=3500.0, nx=161, ny=5, kind=kind, x0=5.0, x1=25.0)
        e3 = munk_env3d(nalpha=401, nbeta=41, beta=(-4.0, 4.0), nrz=41, nrr=101, run_type="CG   3")
        en = munk_env3d(nalpha=401, nbeta=1, ntheta=1, nrz=41, nrr=101, run_type="CG   2")
        for e in (e3, en):
            e.bty = g
            e.box_x_km = e.box_y_km = 30.0
            e.Sx_km = np.array([6.0])
        t[kind] = (_tl3(e3)[1][0, 0, 0, 0], _tl3(en)[1][0, 0, 0, 0])
    m = np.ones_like(t["CS"][0], bool)
    for a in t.values():
        for x in a:
            m &= x < 90
    m[:, :10] = False
    assert m.sum() > 1500
    for kind in ("CS", "RS"):
        d = np.abs(t[kind][0] - t[kind][1])[m]
        assert np.median(d) < 0.05 and np.percentile(d, 90) < 0.15, (kind, np.median(d), np.percentile(d, 90))
    assert np.percentile(np.abs(t["CS"][0] - t["RS"][0])[m], 90) > 1.0
