"""Pins the oracle (oracle/bellhop2d.c) against what the mathematics fixes.

The reference ships no 2D golden vectors (its only fixture,
test/individual/analysis_3d/example_analysis_3d.arr.txt, is a 3D isovelocity
arrivals file) and cannot be built here, so each part of the oracle is checked
against closed forms, invariants and textbook results, chosen so that a
plausible mistake (a dropped term, a wrong sign or index) fails one of them.

What pins what (oracle function -> test):
  RayInit launch direction, Step in isovelocity (x' = c t, q' = c p, tau' = 1/c),
  boundary/box step reduction, flat Reflect (vacuum +pi, rigid 0)
      -> test_isovelocity_rays_are_folded_straight_lines,
         test_isovelocity_step_counts_closed_form
  Step gradient terms (t' = -grad c / c^2) and the Heun blend
      -> test_linear_gradient_ray_is_circular_arc
  Range-independent invariance (t.x bitwise constant, flat reflections)
      -> test_tx_invariant_range_independent
  EvaluateSSP c-linear / n2-linear / spline (not-a-knot) / Munk analytic
      -> test_ssp_values_along_rays
  PCHIP: knot interpolation and monotonicity; exact reproduction of a
  monotone quadratic (3-point end slopes, clamped-spline seeds, Hermite
  coefficients); equality with SciPy's clamped cubic spline where the
  Fritsch-Carlson bounds hold; the projection onto those bounds, zero slopes
  at extrema and the end-slope limiter where they do not
      -> test_pchip_properties, test_pchip_reproduces_monotone_quadratic,
         test_pchip_equals_clamped_spline_where_monotone,
         test_pchip_projection_where_the_spline_overshoots
  Quad SSP: bilinear values (SciPy) and, on a sampled linear field, rays that
  are circular arcs about the line c = 0 (fixes both gradient components)
      -> test_quad_ssp_is_bilinear, test_quad_ssp_linear_field_rays_are_circular_arcs
  Reflect, acousto-elastic halfspace (fluid, cs = 0) -> Rayleigh coefficient
      -> test_halfspace_reflection_rayleigh
  Reflect, elastic halfspace (cs > 0) -> Brekhovskikh's fluid-solid coefficient
      -> test_elastic_halfspace_reflection_brekhovskikh
  Reflect with beam shift ('S') -> Goos-Haenchen lateral shift of the Rayleigh
  phase, its delay and width terms
      -> test_beam_shift_is_goos_haenchen_shift
  Curvilinear boundaries ('C'): curvature of a parabola sampled at equal
  steps (exact for the chord scheme), node tangents / normals, the
  interpolated reflecting normal
      -> test_curvilinear_bottom_parabola
  Reflect, tabulated coefficient ('F') -> linear interpolation in grazing angle
      -> test_reflection_table_interpolation
  Influence geometric hat / Gaussian, Cartesian and ray-centred, coherent sum
  and ScalePressure -> point source in free space, Lloyd mirror
      -> test_free_space_point_source, test_lloyd_mirror
  ScalePressure factors -> test_scale_pressure_closed_form
  Incoherent / semicoherent sums -> test_incoherent_and_semicoherent_closed_form
  Arrivals (AddArr) delay and amplitude in free space
      -> test_free_space_arrivals
  AddArr merging (pair test on the last slot, weakest-slot replacement,
  drops, float averaging) -> a replay of the rules on the oracle's raw calls
      -> test_arrivals_merge_replays_addarr_rules
  Eigen hits (RecordEigenHit) -> test_eigen_hits_bracket_receivers
  crci, every attenuation unit (N, M, m, F, W, Q, L), Thorp, biological layers
      -> test_volume_attenuation_closed_form, test_biological_attenuation_closed_form
Parity unpinned (held only by faithfulness to the reference source text):
exact-amplitude ties in the AddArr weakest-slot search (they do not arise in
the replayed runs), and Francois-Garrison volume attenuation beyond its use
in the product-versus-oracle comparisons.
"""
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import oracle
from bellhopcuda_amd.env import Env2D, Halfspace, ReflTable, Boundary, subtab, munk_c, munk_env


def iso_env(D=1000.0, c=1500.0, zs=300.0, alpha=(-30.0, 30.0), nb=61, top="V", bot="R", rmax_km=5.0,
            run_type="R", deltas=137.0, precision=64, rz=None, rr_km=None, freq=100.0, bot_hs=None):
    return Env2D(
        title="iso", precision=precision, freq0=freq, ssp_type="C",
        ssp_z=np.array([0.0, D]), ssp_c=np.array([c, c]),
        top=Halfspace(top, depth=0.0), bot=bot_hs or Halfspace(bot, depth=D),
        Sz=np.array([zs], np.float32),
        Rz=np.asarray(rz if rz is not None else [0.0], np.float32),
        Rr_km=np.asarray(rr_km if rr_km is not None else [0.0, rmax_km], np.float64),
        alpha_deg=subtab(alpha[0], alpha[1], nb), run_type=run_type, deltas=deltas,
        box_r_km=rmax_km, box_z=D + 100.0)


def _rays(env, max_points=6000):
    pts, nst, ew = oracle.ray_run(env, max_points=max_points)
    assert ew == (0, 0)
    return [pts[i][:nst[i]] for i in range(len(nst))]


def _fold(u, D):
    u = np.mod(u, 2 * D)
    return np.where(u > D, 2 * D - u, u)


# --------------------------------------------------------------------------
# straight rays
# --------------------------------------------------------------------------

def test_isovelocity_rays_are_folded_straight_lines():
    D, zs, c = 1000.0, 300.0, 1500.0
    env = iso_env(D=D, zs=zs, c=c, top="V", bot="R")
    for a, r in zip(env.alpha_deg, _rays(env)):
        al = math.radians(a)
        x, z = r["x"][:, 0], r["x"][:, 1]
        # unfolded depth along the launch direction, folded by the mirrors at 0 and D
        assert np.allclose(z, _fold(zs + x * math.tan(al), D), rtol=0, atol=1e-9 * D)
        s = x / math.cos(al)                     # path length
        assert np.allclose(r["tau"][:, 0], s / c, rtol=1e-12, atol=1e-15)
        assert np.all(r["tau"][:, 1] == 0.0)
        assert np.allclose(r["q"][:, 0], c * s, rtol=1e-12, atol=1e-9)   # q' = c p, p = 1
        assert np.all(r["q"][:, 1] == 1.0) and np.all(r["p"][:, 0] == 1.0)
        assert np.allclose(np.hypot(r["t"][:, 0], r["t"][:, 1]) * c, 1.0, rtol=1e-14)
        assert np.all(r["Amp"] == 1.0)
        # vacuum top adds pi per bounce, rigid bottom nothing
        assert np.allclose(r["Phase"], math.pi * r["NumTopBnc"], rtol=0, atol=1e-12)
        # bounce counts: boundary crossings of the unfolded line before the end
        u_end = zs + x[-1] * math.tan(al)
        k = np.arange(-50, 51)
        crossings_top = np.sum(((2 * k * D) > min(zs, u_end)) & ((2 * k * D) < max(zs, u_end)))
        crossings_bot = np.sum((((2 * k + 1) * D) > min(zs, u_end)) & (((2 * k + 1) * D) < max(zs, u_end)))
        assert r["NumTopBnc"][-1] == crossings_top
        assert r["NumBotBnc"][-1] == crossings_bot


def _expected_points(alpha_deg, D, zs, ds, box_r):
    """Points a straight ray stores: one per step of min(ds, distance to the
    next boundary / the box edge), plus a duplicated point at each reflection."""
    a = math.radians(alpha_deg)
    ca, sa = math.cos(a), math.sin(a)
    s_box = box_r / ca
    events = []
    if sa != 0.0:
        k = math.floor(zs / D) + 1 if sa > 0 else math.ceil(zs / D) - 1
        while True:
            s = (k * D - zs) / sa
            if s >= s_box:
                break
            events.append(s)
            k += 1 if sa > 0 else -1
    legs = np.diff([0.0] + events + [s_box])
    assert np.all(np.abs(legs / ds - np.round(legs / ds)) > 1e-6), "ambiguous case"
    steps = int(sum(math.ceil(leg / ds) for leg in legs))
    return steps + 1 + len(events)


@pytest.mark.parametrize("deltas,zs", [(137.0, 300.0), (61.0, 777.0)])
def test_isovelocity_step_counts_closed_form(deltas, zs):
    D = 1000.0
    env = iso_env(D=D, zs=zs, deltas=deltas, nb=61, alpha=(-31.0, 29.0))
    pts, nst, _ = oracle.ray_run(env, max_points=6000)
    exp = [_expected_points(a, D, zs, deltas, 5000.0) for a in env.alpha_deg]
    assert list(nst) == exp


# --------------------------------------------------------------------------
# curved rays
# --------------------------------------------------------------------------

def test_linear_gradient_ray_is_circular_arc():
    """c(z) = c0 + g z: rays are circles centred at the depth where c would
    vanish, radius 1 / (g xi) with xi = cos(theta)/c the ray parameter."""
    c0, g, D = 1500.0, 0.05, 4000.0
    env = Env2D(title="grad", precision=64, freq0=50.0, ssp_type="C",
                ssp_z=np.array([0.0, D]), ssp_c=np.array([c0, c0 + g * D]),
                top=Halfspace("V", depth=0.0), bot=Halfspace("R", depth=D),
                Sz=np.array([1000.0], np.float32), Rz=np.array([0.0], np.float32),
                Rr_km=np.array([0.0, 20.0]), alpha_deg=subtab(-20.0, 20.0, 9), run_type="R",
                deltas=20.0, box_r_km=20.0, box_z=D + 100.0)
    zc = -c0 / g
    for r in _rays(env):
        nb = r["NumTopBnc"] + r["NumBotBnc"]
        seg = r[nb == 0]                          # the arc before the first bounce
        xi = seg["t"][0, 0]
        R = 1.0 / (g * xi)
        # centre range from the launch point: (x0 - rc)^2 + (z0 - zc)^2 = R^2,
        # on the side the launch direction fixes
        t0, x0 = seg["t"][0], seg["x"][0]
        dz = x0[1] - zc
        rc = x0[0] + np.sign(t0[1]) * math.sqrt(max(R * R - dz * dz, 0.0))
        dist = np.hypot(seg["x"][:, 0] - rc, seg["x"][:, 1] - zc)
        assert np.max(np.abs(dist - R)) / R < 1e-7
        # xi = t.x is conserved exactly (grad c has no range component)
        assert np.all(seg["t"][:, 0] == xi)


def test_tx_invariant_range_independent():
    env = munk_env(nbeams=41, nrz=3, nrr=3, run_type="R")
    for r in _rays(env, max_points=20000):
        assert np.all(r["t"][:, 0] == r["t"][0, 0])


# --------------------------------------------------------------------------
# sound speed
# --------------------------------------------------------------------------

def _ssp_env(ssp_type, z, c, D):
    return Env2D(title="ssp", precision=64, freq0=50.0, ssp_type=ssp_type,
                 ssp_z=z, ssp_c=c, top=Halfspace("V", depth=0.0), bot=Halfspace("R", depth=D),
                 Sz=np.array([900.0], np.float32), Rz=np.array([0.0], np.float32),
                 Rr_km=np.array([0.0, 30.0]), alpha_deg=subtab(-14.0, 14.0, 15), run_type="R",
                 deltas=97.0, box_r_km=30.0, box_z=D + 100.0)


@pytest.mark.parametrize("ssp_type", ["C", "N", "S", "A"])
def test_ssp_values_along_rays(ssp_type):
    D = 5000.0
    # the analytic profile still walks the tabulated layers for its step
    # reduction (a single layer keeps every interface on a boundary)
    z = np.linspace(0.0, D, 26 if ssp_type != "A" else 2)
    c = munk_c(z)
    env = _ssp_env(ssp_type, z, c, D)
    for r in _rays(env, max_points=20000):
        zz, cc = r["x"][:, 1], r["c"]
        if ssp_type == "C":
            ref = np.interp(zz, z, c)
            tol = 1e-12
        elif ssp_type == "N":      # 1/c^2 linear between knots
            ref = 1.0 / np.sqrt(np.interp(zz, z, 1.0 / c ** 2))
            tol = 1e-12
        elif ssp_type == "S":      # cubic spline, not-a-knot ends
            ref = CubicSpline(z, c, bc_type="not-a-knot")(zz)
            tol = 1e-10
        else:                      # Munk's canonical profile, closed form; the
            # reference holds its constants 1300 and 0.00737 as floats
            eps = float(np.float32(0.00737))
            eta = 2.0 * (zz - 1300.0) / 1300.0
            ref = 1500.0 * (1.0 + eps * (eta - 1.0 + np.exp(-eta)))
            tol = 1e-12
        assert np.max(np.abs(cc - ref) / ref) < tol


def test_pchip_properties():
    D = 3000.0
    z = np.array([0.0, 200.0, 500.0, 800.0, 1500.0, 2200.0, 3000.0])
    c = np.array([1540.0, 1520.0, 1495.0, 1490.0, 1492.0, 1510.0, 1530.0])
    env = _ssp_env("P", z, c, D)
    env.Sz = np.array([700.0], np.float32)
    seen = 0
    for r in _rays(env, max_points=20000):
        zz, cc = r["x"][:, 1], r["c"]
        for i in range(len(z) - 1):
            m = (zz >= z[i]) & (zz <= z[i + 1])
            if not m.any():
                continue
            seen += m.sum()
            lo, hi = min(c[i], c[i + 1]), max(c[i], c[i + 1])
            # shape preservation: between monotone neighbours the interpolant
            # stays within the bracket of the two knot values
            monotone = (i == 0 or (c[i] - c[i - 1]) * (c[i + 1] - c[i]) > 0) and \
                       (i + 2 >= len(c) or (c[i + 2] - c[i + 1]) * (c[i + 1] - c[i]) > 0)
            if monotone:
                assert np.all(cc[m] >= lo - 1e-9) and np.all(cc[m] <= hi + 1e-9)
            at = m & np.isin(zz, z)
            for zk, ck in zip(zz[at], cc[at]):
                assert abs(ck - c[np.where(z == zk)[0][0]]) < 1e-9
    assert seen > 100


def _pchip_coefs(z, c, D=None):
    """PCHIP polynomial coefficients c0..c3 per knot (c0 + c1 s + c2 s^2 +
    c3 s^3, s = z - z_k) from the oracle's preprocessing."""
    env = _ssp_env("P", np.asarray(z, float), np.asarray(c, float), D or float(z[-1]))
    d = oracle.preprocess_dump(env)
    n = len(z)
    coef = d[4 * n:12 * n].reshape(4, n, 2)
    assert np.all(coef[..., 1] == 0.0)          # no attenuation: real coefficients
    return coef[..., 0]


def test_pchip_reproduces_monotone_quadratic():
    """src/curves.hpp:299-372: the end derivatives are 3-point differences,
    exact for a quadratic; the interior seeds come from the cubic spline
    clamped to them, which reproduces a quadratic exactly; the monotone
    projection leaves them alone (they lie between the adjacent secants). So
    on a monotone quadratic every knot holds the quadratic's own Taylor
    coefficients: value, slope, half the curvature, and no cubic term."""
    z = np.array([0.0, 150.0, 400.0, 700.0, 1200.0, 1500.0, 2300.0, 3000.0])
    b, q = 0.016, 3.0e-6
    c = 1490.0 + b * z + q * z ** 2
    k = _pchip_coefs(z, c)
    scale = np.array([1500.0, b, q, q / 3000.0])
    for j in range(len(z) - 1):
        exp = np.array([c[j], b + 2 * q * z[j], q, 0.0])
        assert np.all(np.abs(k[:, j] - exp) <= 1e-9 * scale), (j, k[:, j], exp)


def test_pchip_equals_clamped_spline_where_monotone():
    """Smooth strictly monotone data: the spline's knot slopes already satisfy
    the Fritsch-Carlson bounds 0 <= d <= 3 min(secants), so the PCHIP is the
    cubic spline clamped to the 3-point end slopes (SciPy's CubicSpline with
    first-derivative end conditions): slopes and the interval polynomials
    agree."""
    from scipy.interpolate import CubicSpline
    z = np.array([0.0, 120.0, 300.0, 650.0, 900.0, 1400.0, 1800.0, 2500.0, 3000.0])
    c = 1480.0 + 35.0 * np.tanh((z - 1300.0) / 900.0)
    h1, h2 = z[1] - z[0], z[2] - z[1]
    d1, d2 = (c[1] - c[0]) / h1, (c[2] - c[1]) / h2
    d0 = ((2 * h1 + h2) * d1 - h1 * d2) / (h1 + h2)
    h1, h2 = z[-2] - z[-3], z[-1] - z[-2]
    d1, d2 = (c[-2] - c[-3]) / h1, (c[-1] - c[-2]) / h2
    dn = (-h2 * d1 + (h1 + 2 * h2) * d2) / (h1 + h2)
    sp = CubicSpline(z, c, bc_type=((1, d0), (1, dn)))
    sec = np.diff(c) / np.diff(z)
    slopes = sp(z, 1)
    assert np.all(slopes[1:-1] >= 0) and np.all(slopes[1:-1] <= 3 * np.minimum(sec[:-1], sec[1:]))
    k = _pchip_coefs(z, c)
    for j in range(len(z) - 1):
        exp = np.array([c[j], slopes[j], sp(z[j], 2) / 2.0, sp(z[j], 3) / 6.0])
        # sp(z, 3) at a knot is the right-hand interval's third derivative
        assert np.allclose(k[:, j], exp, rtol=1e-9, atol=1e-12), (j, k[:, j], exp)


def test_pchip_projection_where_the_spline_overshoots():
    """Where adjacent secants change sign the knot slope is 0 (an extremum of
    the interpolant); where the spline's slope leaves [0, 3 min(secants)]
    (Fritsch-Carlson) it is projected onto that interval; an end slope whose
    sign differs from the end secant's is 0. Checked against SciPy's clamped
    spline and the projection, knot by knot, on data with a sharp step."""
    from scipy.interpolate import CubicSpline
    z = np.array([0.0, 100.0, 200.0, 300.0, 400.0, 500.0, 600.0, 700.0, 800.0])
    c = np.array([1500.0, 1499.0, 1499.5, 1520.0, 1521.0, 1521.2, 1515.0, 1516.0, 1530.0])
    sec = np.diff(c) / np.diff(z)

    def end(d1, d2, f):
        if d1 * f <= 0:
            return 0.0
        if d1 * d2 <= 0 and abs(f) > abs(3 * d1):
            return 3 * d1
        return f
    h1, h2 = z[1] - z[0], z[2] - z[1]
    d0 = end(sec[0], sec[1], ((2 * h1 + h2) * sec[0] - h1 * sec[1]) / (h1 + h2))
    h1, h2 = z[-2] - z[-3], z[-1] - z[-2]
    dn = end(sec[-1], sec[-2], (-h2 * sec[-2] + (h1 + 2 * h2) * sec[-1]) / (h1 + h2))
    sp = CubicSpline(z, c, bc_type=((1, d0), (1, dn)))
    exp = np.zeros(len(z))
    exp[0], exp[-1] = d0, dn
    projected = extrema = 0
    for j in range(1, len(z) - 1):
        a, b = sec[j - 1], sec[j]
        f = sp(z[j], 1)
        if a * b > 0:
            lo, hi = (0.0, 3 * min(a, b)) if a > 0 else (3 * max(a, b), 0.0)
            exp[j] = min(max(f, lo), hi)
            projected += exp[j] != f
        else:
            exp[j] = 0.0
            extrema += 1
    assert projected >= 1 and extrema >= 2
    k = _pchip_coefs(z, c)
    assert np.allclose(k[1], exp, rtol=1e-9, atol=1e-12), (k[1], exp)
    # the Hermite cubic of each interval from the knot values and slopes
    h = np.diff(z)
    f1, f2, p1, p2 = c[:-1], c[1:], exp[:-1], exp[1:]
    c2 = (3 * (f2 - f1) - h * (2 * p1 + p2)) / h ** 2
    c3 = (h * (p1 + p2) - 2 * (f2 - f1)) / h ** 3
    assert np.allclose(k[2, :-1], c2, rtol=1e-9, atol=1e-12)
    assert np.allclose(k[3, :-1], c3, rtol=1e-9, atol=1e-15)


def _quad_env(z, r_km, cmat, D, alpha=(-14.0, 14.0), nb=15, zs=900.0, deltas=97.0, rmax_km=30.0):
    e = _ssp_env("Q", np.asarray(z, float), np.asarray(cmat, float)[:, 0].copy(), D)
    e.ssp_r_km = np.asarray(r_km, float)
    e.ssp_cmat = np.asarray(cmat, float)
    e.Sz = np.array([zs], np.float32)
    e.alpha_deg = subtab(alpha[0], alpha[1], nb)
    e.deltas = deltas
    e.Rr_km = np.array([0.0, rmax_km])
    e.box_r_km = rmax_km
    return e


def test_quad_ssp_is_bilinear():
    """Quad SSP (src/ssp.hpp:157-214): inside each (range, depth) cell the
    sound speed is the bilinear interpolation of the four corner values."""
    from scipy.interpolate import RegularGridInterpolator
    D = 3000.0
    z = np.array([0.0, 300.0, 800.0, 1500.0, 2200.0, 3000.0])
    r_km = np.array([-1.0, 4.0, 9.0, 17.0, 31.0])
    rng = np.random.default_rng(7)
    cmat = 1490.0 + 0.012 * z[:, None] + rng.uniform(-8.0, 8.0, (len(z), len(r_km)))
    env = _quad_env(z, r_km, cmat, D)
    f = RegularGridInterpolator((z, r_km * 1000.0), cmat, bounds_error=False, fill_value=None)
    n = 0
    for r in _rays(env, max_points=20000):
        ref = f(np.stack([r["x"][:, 1], r["x"][:, 0]], axis=1))
        assert np.max(np.abs(r["c"] - ref) / ref) < 1e-12
        n += len(r)
    assert n > 500


def test_quad_ssp_linear_field_rays_are_circular_arcs():
    """A quad SSP sampling c = c0 + a r + b z (a linear field, exact on the
    grid): rays are circles of radius 1 / (|g| xi), xi = t . m the ray
    parameter along the unit vector m normal to the gradient g = (a, b), with
    centres on the line c = 0. This fixes the gradient the quad evaluation
    returns (both components and signs), not only its values."""
    D = 3000.0
    c0, a, b = 1500.0, 0.002, 0.03
    z = np.linspace(0.0, D, 7)
    r_km = np.linspace(-1.0, 21.0, 12)
    cmat = c0 + a * r_km[None, :] * 1000.0 + b * z[:, None]
    env = _quad_env(z, r_km, cmat, D, alpha=(-12.0, 12.0), nb=9, zs=1200.0, deltas=20.0, rmax_km=20.0)
    g = np.array([a, b])
    gn = np.linalg.norm(g)
    m = np.array([g[1], -g[0]]) / gn
    checked = 0
    for r in _rays(env, max_points=20000):
        nb = r["NumTopBnc"] + r["NumBotBnc"]
        seg = r[nb == 0]
        t, x = seg["t"], seg["x"]
        xi = t[:, 0] * m[0] + t[:, 1] * m[1]
        assert np.max(np.abs(xi - xi[0])) <= 1e-12 * abs(xi[0])
        R = 1.0 / (gn * abs(xi[0]))
        c_at = c0 + g @ x[0]
        that = t[0] * c_at
        kap = -(g - (g @ that) * that) / c_at          # curvature vector
        centre = x[0] + kap / np.dot(kap, kap)
        assert abs(c0 + g @ centre) < 1e-6 * c0         # on the line c = 0
        dist = np.hypot(x[:, 0] - centre[0], x[:, 1] - centre[1])
        assert np.max(np.abs(dist - R)) / R < 1e-7
        checked += len(seg)
    assert checked > 200


# --------------------------------------------------------------------------
# reflection
# --------------------------------------------------------------------------

def _first_bottom_bounce(r):
    k = int(np.argmax(r["NumBotBnc"] == 1))
    assert k > 0
    return r[k - 1], r[k]


def test_halfspace_reflection_rayleigh():
    """Fluid halfspace: R = (rho2 kz1 - rho1 kz2) / (rho2 kz1 + rho1 kz2), with
    the e^{+i omega t} convention of the field (contributions carry
    e^{-i omega tau}), so the evanescent kz2 has a negative imaginary part."""
    D, c1, c2, rho2, f = 1000.0, 1500.0, 1700.0, 1.5, 100.0
    env = iso_env(D=D, c=c1, zs=500.0, alpha=(5.0, 85.0), nb=33, rmax_km=10.0, deltas=50.0, freq=f,
                  bot_hs=Halfspace("A", depth=D, alphaR=c2, rho=rho2, zTemp=D))
    w = 2 * math.pi * f
    n_total = 0
    for a, r in zip(env.alpha_deg, _rays(env)):
        before, after = _first_bottom_bounce(r)
        g = math.radians(a)
        kx, kz1 = w * math.cos(g) / c1, w * math.sin(g) / c1
        kz2 = np.sqrt(complex((w / c2) ** 2 - kx ** 2))
        if kz2.imag > 0:
            kz2 = -kz2
        R = (rho2 * kz1 - kz2) / (rho2 * kz1 + kz2)
        n_total += abs(R) > 0.999999
        assert after["Amp"] / before["Amp"] == pytest.approx(abs(R), rel=1e-10)
        assert after["Phase"] - before["Phase"] == pytest.approx(np.angle(R), abs=1e-10)
    assert 0 < n_total < len(env.alpha_deg)   # both total and partial reflection covered


def test_elastic_halfspace_reflection_brekhovskikh():
    """Acousto-elastic halfspace with shear (src/reflect.hpp:316-350): for
    angles where both the compressional and the shear wave propagate in the
    bottom, the plane-wave coefficient of a fluid over an elastic solid
    (Brekhovskikh) is
        R = (Zp cos^2 2ts + Zs sin^2 2ts - Z1) / (Zp cos^2 2ts + Zs sin^2 2ts + Z1),
    Z1 = rho1 c1 / cos t1, Zp = rho2 cp / cos tp, Zs = rho2 cs / cos ts, angles
    from the normal and linked by Snell's law."""
    D, c1, cp, cs, rho2, f = 1000.0, 1500.0, 1800.0, 600.0, 1.9, 100.0
    env = iso_env(D=D, c=c1, zs=500.0, alpha=(60.0, 85.0), nb=11, rmax_km=10.0, deltas=50.0, freq=f,
                  bot_hs=Halfspace("A", depth=D, alphaR=cp, betaR=cs, rho=rho2, zTemp=D))
    for a, r in zip(env.alpha_deg, _rays(env)):
        before, after = _first_bottom_bounce(r)
        t1 = math.radians(90.0 - a)                      # from the normal
        stp, sts = cp / c1 * math.sin(t1), cs / c1 * math.sin(t1)
        assert stp < 1.0 and sts < 1.0
        tp, ts = math.asin(stp), math.asin(sts)
        Z1, Zp, Zs = c1 / math.cos(t1), rho2 * cp / math.cos(tp), rho2 * cs / math.cos(ts)
        num = Zp * math.cos(2 * ts) ** 2 + Zs * math.sin(2 * ts) ** 2
        R = (num - Z1) / (num + Z1)
        assert after["Amp"] / before["Amp"] == pytest.approx(abs(R), rel=1e-9)
        dphi = (after["Phase"] - before["Phase"]) % (2 * math.pi)
        assert min(dphi, 2 * math.pi - dphi) == pytest.approx(0.0 if R > 0 else math.pi, abs=1e-9)


def test_curvilinear_bottom_parabola():
    """Curvilinear bathymetry ('C', src/module/boundary.hpp:600-890 and
    src/trace.hpp:350-362) on a parabola z = z0 + q (r - rc)^2 sampled at equal
    range steps. A parabola's chord slopes are exact, so the reference's
    segment curvature (difference of chord slopes over the chord width, times
    the chord's tangent x-component cubed) equals the graph curvature form
    z'' cos^3 = 2 q t_x^3 exactly; node tangents are the mean of the adjacent
    unit chord tangents and node normals their rotation; at a bounce the
    reflecting normal is the node normals interpolated linearly in the
    along-chord position (not renormalised), so the reflected tangent is
    t - 2 (t.n) n with that n. All expectations are computed from the raw
    points."""
    D, c = 2500.0, 1500.0
    rr = np.linspace(0.0, 30.0, 31)
    q, rc = 400.0 / 15000.0 ** 2, 15000.0
    zb = 1500.0 + q * (rr * 1000.0 - rc) ** 2
    env = _ssp_env("C", np.array([0.0, D]), np.array([c, c]), D)
    env.bty = Boundary("CS", r_km=rr, z=zb)
    env.Sz = np.array([500.0], np.float32)
    env.alpha_deg = subtab(5.0, 40.0, 15)
    env.deltas = 50.0
    env.Rr_km = np.array([0.0, 29.0])
    env.box_r_km = 29.0
    d = oracle.preprocess_dump(env)
    N = 2
    off = 17 * N + len(env.Rr_km) + len(env.alpha_deg)
    nodes = d[off:off + 13 * (len(rr) + 2)].reshape(-1, 13)[1:-1]     # drop the two extension nodes
    x = np.stack([rr * 1000.0, zb], axis=1)
    assert np.array_equal(nodes[:, 0:2], x)
    ch = np.diff(x, axis=0)
    L = np.hypot(ch[:, 0], ch[:, 1])
    t = ch / L[:, None]
    # chords 0 .. n-3 have a parabola chord on both sides
    kap = nodes[:-2, 11]
    assert np.allclose(kap, 2 * q * t[:-1, 0] ** 3, rtol=1e-9, atol=0)
    nodet = 0.5 * (t[:-1] + t[1:])
    assert np.allclose(nodes[1:-1, 8:10], nodet, rtol=0, atol=1e-15)
    noden = np.stack([-nodet[:, 1], nodet[:, 0]], axis=1)
    assert np.allclose(nodes[1:-1, 6:8], noden, rtol=0, atol=1e-15)
    bounces = 0
    for r in _rays(env, max_points=20000):
        k = np.nonzero(np.diff(r["NumBotBnc"]) == 1)[0]
        for j in k:
            a, b = r[j], r[j + 1]
            xr = a["x"]
            i = int(np.searchsorted(x[:, 0], xr[0]) - 1)
            if i < 1 or i > len(rr) - 3:
                continue                      # next to an end node: tangent (1, 0) by definition
            sss = np.dot(xr - x[i], t[i]) / L[i]
            n = (1 - sss) * noden[i - 1] + sss * noden[i]
            tin = a["t"]
            exp = tin - 2.0 * np.dot(tin, n) * n
            assert np.allclose(b["t"], exp, rtol=1e-11, atol=1e-18), (b["t"], exp)
            bounces += 1
    assert bounces >= 10


def test_reflection_table_interpolation():
    th = np.array([0.0, 20.0, 45.0, 90.0])
    rr = np.array([1.0, 0.8, 0.5, 0.2])
    ph = np.array([0.0, 10.0, 25.0, 40.0])         # degrees in the table
    D = 1000.0
    env = iso_env(D=D, zs=500.0, alpha=(3.0, 80.0), nb=17, rmax_km=12.0, deltas=50.0,
                  bot_hs=Halfspace("F", depth=D))
    env.brc = ReflTable(th, rr, ph)
    for a, r in zip(env.alpha_deg, _rays(env)):
        before, after = _first_bottom_bounce(r)
        assert after["Amp"] / before["Amp"] == pytest.approx(np.interp(a, th, rr), rel=1e-10)
        assert after["Phase"] - before["Phase"] == pytest.approx(math.radians(np.interp(a, th, ph)), abs=1e-10)


# --------------------------------------------------------------------------
# influence and field
# --------------------------------------------------------------------------

def _field_env(D, zs, rz, rr_km, nb, alpha, run_type, top="V", bot="V", freq=200.0):
    return Env2D(title="fs", precision=64, freq0=freq, ssp_type="C", ssp_z=np.array([0.0, D]),
                 ssp_c=np.array([1500.0, 1500.0]), top=Halfspace(top, depth=0.0), bot=Halfspace(bot, depth=D),
                 Sz=np.array([zs], np.float32), Rz=np.asarray(rz, np.float32),
                 Rr_km=np.asarray(rr_km, np.float64), alpha_deg=subtab(alpha[0], alpha[1], nb),
                 run_type=run_type, deltas=0.0, box_r_km=rr_km[-1] * 1.01, box_z=D + 100.0)


@pytest.mark.parametrize("run_type,tol_db", [("CG", 0.01), ("Cg", 0.01), ("CB", 0.5), ("Cb", 0.5)])
def test_free_space_point_source(run_type, tol_db):
    """Deep isovelocity water, receivers far from both boundaries: the scaled
    field is -exp(-ikR)/R (BELLHOP's sign and time convention)."""
    D, zs, f = 20000.0, 10000.0, 200.0
    rz = subtab(9000.0, 11000.0, 41, np.float32)
    rr = subtab(0.2, 3.0, 57)
    env = _field_env(D, zs, rz, rr, 2001, (-80.0, 80.0), run_type, freq=f)
    o = oracle.field_run(env, nthreads=0, stats=False)
    p = oracle.scale_pressure(env, o.field)[0]
    R = np.hypot(rr[None, :] * 1000.0, rz[:, None].astype(np.float64) - zs)
    m = np.ones_like(R, bool)
    if run_type[1] in "gb":
        m[:, 0] = False      # ray-centred influence never reaches the first receiver range
    if run_type[1] in "Bb":
        m &= rr[None, :] >= 1.0   # Gaussian beams: far field only
    tl = -20 * np.log10(np.abs(p[m]))
    assert np.max(np.abs(tl - 20 * np.log10(R[m]))) < tol_db
    if run_type[1] in "Gg":
        k = 2 * math.pi * f / 1500.0
        ph = np.angle(-p * np.exp(1j * k * R))[m]
        assert np.max(np.abs(ph)) < 0.01


def test_lloyd_mirror():
    """Vacuum surface: p = -(e^{-ikR1}/R1 - e^{-ikR2}/R2) with the image source at -zs."""
    zs, f = 50.0, 100.0
    rz = subtab(10.0, 200.0, 39, np.float32)
    rr = subtab(0.5, 5.0, 91)
    env = _field_env(20000.0, zs, rz, rr, 4001, (-80.0, 80.0), "CG", freq=f)
    o = oracle.field_run(env, nthreads=0, stats=False)
    p = oracle.scale_pressure(env, o.field)[0]
    k = 2 * math.pi * f / 1500.0
    z = rz[:, None].astype(np.float64)
    r = rr[None, :] * 1000.0
    R1, R2 = np.hypot(r, z - zs), np.hypot(r, z + zs)
    ref = -(np.exp(-1j * k * R1) / R1 - np.exp(-1j * k * R2) / R2)
    m = np.abs(ref) * R1 > 0.3           # away from the interference nulls
    assert m.sum() > 1000
    d = np.abs(20 * np.log10(np.abs(p[m]) / np.abs(ref[m])))
    assert d.max() < 0.01
    assert np.max(np.abs(np.angle(p[m] * np.conj(ref[m])))) < 0.01


def test_scale_pressure_closed_form():
    """Geometric beams: cnst = -1, point source cylindrical spreading 1/sqrt(r),
    zero at r = 0; incoherent runs take sqrt of the intensity first."""
    rr = np.array([0.0, 0.5, 2.0, 7.5])
    for rt in ("CG", "IG"):
        env = iso_env(rz=[10.0, 20.0, 30.0], rr_km=rr, run_type=rt)
        u = (np.arange(12).reshape(1, 3, 4) + 1.0 + 0.5j).astype(np.complex64)
        if rt[0] == "I":
            u = u.real.astype(np.complex64)
        got = oracle.scale_pressure(env, u)
        base = u if rt[0] == "C" else np.sqrt(u.real).astype(np.complex64)
        fac = -1.0 / np.sqrt(np.where(rr == 0.0, np.inf, rr * 1000.0))
        assert np.allclose(got, base * fac[None, None, :].astype(np.float32), rtol=1e-6, atol=0)


def test_incoherent_and_semicoherent_closed_form():
    """Incoherent runs add beam intensities: with hat beams (a partition of
    unity) a vacuum-surface Lloyd mirror gives |p|^2 = 1/R1^2 + 1/R2^2.
    Semicoherent runs weight each ray by the Lloyd factor 2 sin^2(k zs sin a)
    and add intensities; under an impedance-matched surface (no reflection)
    |p|^2 = 2 sin^2(k zs sin a) / R^2 with a the direct-path launch angle."""
    f, k = 200.0, 2 * math.pi * 200.0 / 1500.0
    rz = subtab(100.0, 400.0, 31, np.float32)
    rr = subtab(0.5, 3.0, 51)
    z = rz[:, None].astype(np.float64)
    r = rr[None, :] * 1000.0
    zs = 60.0
    R1, R2 = np.hypot(r, z - zs), np.hypot(r, z + zs)
    env = _field_env(20000.0, zs, rz, rr, 4001, (-80.0, 80.0), "IG", freq=f)
    p = oracle.scale_pressure(env, oracle.field_run(env, stats=False).field)[0]
    assert np.all(p.imag == 0)
    ref = np.sqrt(1.0 / R1 ** 2 + 1.0 / R2 ** 2)
    assert np.max(np.abs(20 * np.log10(np.abs(p) / ref))) < 0.01

    zs = 20.0
    R = np.hypot(r, z - zs)
    env = _field_env(20000.0, zs, rz, rr, 4001, (-80.0, 80.0), "SG", freq=f)
    env.top = Halfspace("A", depth=0.0, alphaR=1500.0, rho=1.0, zTemp=0.0)
    p = oracle.scale_pressure(env, oracle.field_run(env, stats=False).field)[0]
    sin_a = (z - zs) / R
    ref = np.sqrt(2.0 * np.sin(k * zs * sin_a) ** 2) / R
    m = np.abs(np.sin(k * zs * sin_a)) > 0.3
    assert m.sum() > 500
    assert np.max(np.abs(20 * np.log10(np.abs(p[m]) / ref[m]))) < 0.05


# --------------------------------------------------------------------------
# arrivals and eigen hits
# --------------------------------------------------------------------------

def test_free_space_arrivals():
    """One direct arrival per receiver with delay R/c (straight-ray travel
    time to the receiver's projection on the beam, within the beam width)."""
    D, zs = 20000.0, 10000.0
    rz = subtab(9500.0, 10500.0, 11, np.float32)
    rr = subtab(1.0, 3.0, 11)
    env = _field_env(D, zs, rz, rr, 801, (-40.0, 40.0), "AG", freq=200.0)
    o = oracle.field_run(env, stats=False, max_narr=20, allow_merging=1)
    arr = np.ctypeslib.as_array(o.arr).reshape(1, len(rz), len(rr), 20)
    narr = o.narr.reshape(1, len(rz), len(rr))
    R = np.hypot(rr[None, :] * 1000.0, rz[:, None].astype(np.float64) - zs)
    assert np.all(narr >= 1)
    for iz in range(len(rz)):
        for ir in range(len(rr)):
            a = arr[0, iz, ir, :narr[0, iz, ir]]
            assert np.all(a["NTopBnc"] == 0) and np.all(a["NBotBnc"] == 0)
            assert np.allclose(a["delay"][:, 0], R[iz, ir] / 1500.0, rtol=1e-5)
            decl = math.degrees(math.atan2(rz[iz] - zs, rr[ir] * 1000.0))
            assert np.all(np.abs(a["SrcDeclAngle"] - decl) < 0.5)


def test_eigen_hits_bracket_receivers():
    """Each eigen hit names a ray step whose segment brackets the receiver
    range and passes within the beam width of the receiver."""
    D, zs = 20000.0, 10000.0
    rz = subtab(9800.0, 10200.0, 5, np.float32)
    rr = subtab(1.0, 2.0, 5)
    env = _field_env(D, zs, rz, rr, 201, (-20.0, 20.0), "EG", freq=200.0)
    o = oracle.field_run(env, stats=False, eigen_memsize=100000)
    n = int(o.neigen[0])
    assert n > 0
    hits = np.ctypeslib.as_array(o.hits)[:n]
    env_r = _field_env(D, zs, rz, rr, 201, (-20.0, 20.0), "R", freq=200.0)
    rays = _rays(env_r)
    dalpha = math.radians(40.0 / 200)
    for h in hits:
        r = rays[h["ialpha"]]
        xa, xb = r["x"][h["is_"]], r["x"][h["is_"] + 1]
        rec = np.array([rr[h["ir"]] * 1000.0, float(rz[h["iz"]])])
        assert min(xa[0], xb[0]) <= rec[0] <= max(xa[0], xb[0])
        d = xb - xa
        u = d / np.linalg.norm(d)
        normal_dist = abs(u[0] * (rec - xa)[1] - u[1] * (rec - xa)[0])
        # hat beam half-width in isovelocity: q / q0 = (distance travelled) * dalpha
        width = np.linalg.norm(rec - np.array([0.0, zs])) * dalpha
        assert normal_dist <= width * 1.01


@pytest.mark.parametrize("unit,level", [("N ", 2e-5), ("M ", 1e-4), ("m ", 1e-4), ("F ", 0.5), ("W ", 0.2),
                                        ("Q ", 300.0), ("L ", 1e-5), ("WT", 0.0)])
def test_volume_attenuation_closed_form(unit, level):
    """crci (src/module/atten.cpp:99-194): in free space the attenuation adds
    8.686 alpha_T s dB along the path s, alpha_T in nepers/m from the unit's
    definition: N nepers/m, M dB/m, F dB/(m kHz), W dB/wavelength, Q quality
    factor (omega / (2 c Q)), L loss parameter (alpha omega / c), m dB/m with
    a power law in f / freq0 (at the run frequency freq0 the factor is 1), and with a
    'T' second letter Thorp's formula in dB/km,
    3.3e-3 + 0.11 f^2/(1+f^2) + 44 f^2/(4100+f^2) + 3e-4 f^2 (f in kHz)."""
    D, zs, f, c = 20000.0, 10000.0, 200.0, 1500.0
    rz = subtab(9500.0, 10500.0, 11, np.float32)
    rr = subtab(0.5, 3.0, 11)       # before the steepest surface-reflected beams arrive (3.3 km)

    def tl(u, lev):
        env = _field_env(D, zs, rz, rr, 2001, (-80.0, 80.0), "CG", freq=f)
        env.atten_unit = u
        env.ssp_alphaI = np.full(2, lev)
        o = oracle.field_run(env, nthreads=0, stats=False)
        return -20 * np.log10(np.abs(oracle.scale_pressure(env, o.field)[0]))

    w = 2 * math.pi * f
    alpha_np = {"N": level, "M": level / 8.6858896, "m": level / 8.6858896, "F": level * f / 8685.8896,
                "W": level * f / (8.6858896 * c),
                "Q": w / (2.0 * c * level) if level else 0.0, "L": level * w / c}[unit[0]]
    if unit[1] == "T":
        fk2 = (f / 1000.0) ** 2
        alpha_np += (3.3e-3 + 0.11 * fk2 / (1 + fk2) + 44.0 * fk2 / (4100.0 + fk2) + 3e-4 * fk2) / 8685.8896
    R = np.hypot(rr[None, :] * 1000.0, rz[:, None].astype(np.float64) - zs)
    excess = tl(unit, level) - tl("N ", 0.0)
    expect = 20.0 * np.log10(math.e) * alpha_np * R
    assert np.max(np.abs(excess - expect)) < 1e-3 + 1e-3 * np.max(expect)
    assert np.max(expect) > 0.02


@pytest.mark.parametrize("layer,gated", [((-100.0, 25000.0), False), ((-100.0, 100.0), True)])
def test_biological_attenuation_closed_form(layer, gated):
    """Biological layers ('B', src/module/atten.cpp:169-180): a knot inside a
    layer [z1, z2] gains a0 / ((1 - (f0/f)^2)^2 + 1/Q^2) dB/km. With both
    knots of a 20 km deep isovelocity profile inside the layer the excess
    attenuation is uniform; with only the top knot inside, the imaginary sound
    speed falls linearly to zero at the bottom knot, so the excess is the
    path integral of a(1 - z/D) (10 m steps there: inside a step the delay is
    interpolated linearly, which takes the step's mean attenuation)."""
    D, zs, f, c = 20000.0, 10000.0, 200.0, 1500.0
    f0, Q, a0 = 180.0, 4.0, 0.3
    rz = subtab(9500.0, 10500.0, 11, np.float32)
    rr = subtab(0.5, 3.0, 11)

    def tl(u, bio):
        env = _field_env(D, zs, rz, rr, 2001, (-80.0, 80.0), "CG", freq=f)
        env.atten_unit = u
        env.ssp_alphaI = np.zeros(2)
        env.bio = bio
        if gated:
            env.deltas = 10.0
        o = oracle.field_run(env, nthreads=0, stats=False)
        return -20 * np.log10(np.abs(oracle.scale_pressure(env, o.field)[0]))

    a = a0 / ((1.0 - (f0 / f) ** 2) ** 2 + 1.0 / Q ** 2) / 8685.8896     # nepers/m
    bio = np.array([[layer[0], layer[1], f0, Q, a0], [30000.0, 40000.0, f0, Q, 100.0]])  # second layer below
    R = np.hypot(rr[None, :] * 1000.0, rz[:, None].astype(np.float64) - zs)
    if gated:
        zr = rz[:, None].astype(np.float64)
        expect_np = a * R * (1.0 - 0.5 * (zs + zr) / D)      # mean of a (1 - z/D) along the straight path
    else:
        expect_np = a * R
    excess = tl("NB", bio) - tl("N ", None)
    expect = 20.0 * np.log10(math.e) * expect_np
    assert np.max(np.abs(excess - expect)) < 1e-3 + 1e-3 * np.max(expect)
    assert np.max(expect) > 0.02


def test_beam_shift_is_goos_haenchen_shift():
    """Beam shift on a fluid-halfspace reflection (RunType[6] 'S',
    src/reflect.hpp:383-442): below the critical angle the reflected point
    moves forward along the boundary by the lateral (Goos-Haenchen) shift
    Delta = (d phi / d theta) / (k sin theta) of the Rayleigh phase phi, the
    delay grows by Delta cos(theta) / c, and q by |d Delta / d theta| sin(theta)
    c p; above it the shift is zero (the width term is 0/0 there, so the
    reference's q becomes NaN for a lossless halfspace). Isovelocity water keeps
    |c t| = 1, so the closed form applies at the ray's own angle."""
    D, c1, c2, rho2, f = 1000.0, 1500.0, 1700.0, 1.5, 100.0
    env = iso_env(D=D, c=c1, zs=500.0, alpha=(5.0, 80.0), nb=31, rmax_km=10.0, deltas=50.0, freq=f,
                  bot_hs=Halfspace("A", depth=D, alphaR=c2, rho=rho2, zTemp=D))
    env.run_type = "R     S"
    w = 2 * math.pi * f

    def phase(g):
        kx, kz1 = w * math.cos(g) / c1, w * math.sin(g) / c1
        kz2 = np.sqrt(complex((w / c2) ** 2 - kx ** 2))
        if kz2.imag > 0:
            kz2 = -kz2
        return np.angle((rho2 * kz1 - kz2) / (rho2 * kz1 + kz2))

    def shift(g, h=1e-6):
        return (phase(g + h) - phase(g - h)) / (2 * h) / (w / c1 * math.sin(g))

    crit = math.acos(c1 / c2)
    below = above = 0
    for a, r in zip(env.alpha_deg, _rays(env)):
        before, after = _first_bottom_bounce(r)
        g = math.radians(a)
        dx = after["x"][0] - before["x"][0]
        dtau = after["tau"][0] - before["tau"][0]
        dq = after["q"][0] - before["q"][0]
        assert after["x"][1] == before["x"][1]          # flat bottom: the shift is along it
        if g < crit - math.radians(1.0):
            below += 1
            d = abs(shift(g))
            assert dx == pytest.approx(d, rel=1e-7)
            assert dtau == pytest.approx(d * math.cos(g) / c1, rel=1e-7)
            dd = (abs(shift(g + 1e-5)) - abs(shift(g - 1e-5))) / 2e-5
            assert dq == pytest.approx(abs(dd) * math.sin(g) * c1 * before["p"][0], rel=2e-4)
        elif g > crit + math.radians(1.0):
            above += 1
            assert dx == 0.0 and dtau == 0.0
            assert np.isnan(dq)
    assert below >= 5 and above >= 5


def _replay_addarr(log, nrcv, max_narr):
    """AddArr with merging (src/arrivals.hpp:31-130) replayed on the raw calls
    in call order, with the reference's roundings: slots hold floats, the
    pair test and the weakest-arrival search compare the unrounded call
    values with the stored floats, a merge averages in float with the
    amplitude weights. Only the last slot is tested for a pair; the weakest
    slot is the first one strictly below the new amplitude (ties keep the
    earlier slot); a new arrival weaker than every stored one is dropped."""
    f32 = np.float32
    A = np.zeros((nrcv, max_narr), f32)
    Ph, Dr, Di, Sd, Rd = (np.zeros((nrcv, max_narr), f32) for _ in range(5))
    Nt_, Nb_ = np.zeros((nrcv, max_narr), np.int32), np.zeros((nrcv, max_narr), np.int32)
    narr = np.zeros(nrcv, np.int32)
    tol = float(f32(0.05))
    st = {"new": 0, "merged": 0, "replaced": 0, "dropped": 0, "ties": 0}
    for row in log:
        k = int(row[0])
        amp, phase, dre, dim, om, sdec, rdec = (float(v) for v in row[1:8])
        ntop, nbot = int(row[8]), int(row[9])
        n = int(narr[k])
        second = (n >= 1 and om * abs(complex(dre, dim) - complex(float(Dr[k, n - 1]), float(Di[k, n - 1]))) < tol
                  and abs(float(Ph[k, n - 1]) - phase) < tol)
        if not second:
            if n >= max_narr:
                i, weakest = -1, amp
                for j in range(max_narr):
                    if float(A[k, j]) < weakest:
                        weakest, i = float(A[k, j]), j
                    elif float(A[k, j]) == weakest:
                        st["ties"] += 1          # the earlier slot (or the new arrival's drop) wins
                if i < 0:
                    st["dropped"] += 1
                    continue
                st["replaced"] += 1
            else:
                i = n
                narr[k] = n + 1
                st["new"] += 1
            A[k, i], Ph[k, i], Dr[k, i], Di[k, i] = f32(amp), f32(phase), f32(dre), f32(dim)
            Sd[k, i], Rd[k, i], Nt_[k, i], Nb_[k, i] = f32(sdec), f32(rdec), ntop, nbot
        else:
            st["merged"] += 1
            i = n - 1
            tot = f32(A[k, i] + f32(amp))
            w1, w2 = f32(A[k, i] / tot), f32(f32(amp) / tot)
            Dr[k, i] = f32(f32(w1 * Dr[k, i]) + f32(w2 * f32(dre)))
            Di[k, i] = f32(f32(w1 * Di[k, i]) + f32(w2 * f32(dim)))
            A[k, i] = tot
            Sd[k, i] = f32(f32(w1 * Sd[k, i]) + f32(w2 * f32(sdec)))
            Rd[k, i] = f32(f32(w1 * Rd[k, i]) + f32(w2 * f32(rdec)))
    return narr, A, Ph, Dr, Di, Sd, Rd, Nt_, Nb_, st


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("max_narr", [2, 3, 6])
def test_arrivals_merge_replays_addarr_rules(max_narr, precision):
    """The oracle's arrivals merge equals a replay of the reference's AddArr
    rules on its own raw calls, bit for bit, with few enough slots that pairs
    merge, weaker arrivals are dropped and stored ones are replaced. (Exact
    amplitude ties between distinct arrivals do not arise here: mirror rays of
    a waveguide symmetric about the source share delay and phase, so the pair
    rule merges them first.)"""
    D = 1000.0
    env = iso_env(D=D, zs=500.0, alpha=(-40.0, 40.0), nb=81, top="R", bot="R", rmax_km=6.0, deltas=97.0,
                  rz=np.array([250.0, 500.0, 750.0], np.float32), rr_km=np.array([1.0, 2.5, 4.0, 5.5]),
                  run_type="AG", precision=precision)
    res, log = oracle.arrivals_with_log(env, max_narr)
    nrcv = 12
    narr, A, Ph, Dr, Di, Sd, Rd, Ntop, Nbot, st = _replay_addarr(log, nrcv, max_narr)
    print(st)
    assert st["merged"] > 0 and st["replaced"] > 0 and st["dropped"] > 0
    arr = np.ctypeslib.as_array(res.arr).reshape(nrcv, max_narr)
    assert np.array_equal(res.narr.reshape(-1), narr)
    for k in range(nrcv):
        n = narr[k]
        got = arr[k, :n]
        assert np.array_equal(got["a"], A[k, :n]) and np.array_equal(got["Phase"], Ph[k, :n])
        assert np.array_equal(got["delay"][:, 0], Dr[k, :n]) and np.array_equal(got["delay"][:, 1], Di[k, :n])
        assert np.array_equal(got["SrcDeclAngle"], Sd[k, :n]) and np.array_equal(got["RcvrDeclAngle"], Rd[k, :n])
        assert np.array_equal(got["NTopBnc"], Ntop[k, :n]) and np.array_equal(got["NBotBnc"], Nbot[k, :n])
