"""CPU tests of the product's host side and of its tracer core (no GPU).

* The C-ABI library loads and exports every function include/bhc_hip.h
  declares; the ctypes mirrors have the C layouts (checked against gcc).
* Host preprocessing (bhc_preprocess2d, C++) is bit-identical to the
  oracle's own preprocessing.
* The tracer core that the HIP kernels run (csrc/core2d.h), instantiated on
  the CPU by tests/harness/core_cpu.cpp, reproduces the oracle bit for bit:
  per-ray steps, bounces, end points, contribution counts and the field.
* ScalePressure and the .shd writer match the reference's layouts.
* Without a GPU the run calls raise (there is no CPU execution path).
"""
import ctypes as C
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd import _abi
from bellhopcuda_amd.env import munk_env, c2_env, subtab, Boundary
import oracle
import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bhc_hip.h")
HARNESS_SRC = os.path.join(ROOT, "tests", "harness", "core_cpu.cpp")
HARNESS_SO = os.path.join(ROOT, "tests", "harness", "_core_cpu.so")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bhc_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = bh.lib()
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_record_layouts(tmp_path):
    """sizeof / offsetof of the C structs (gcc) equal the ctypes mirrors, and
    the output records have the reference's sizes: Arrival 40 B, EigenHit 36 B,
    2D ray point 112 B (include/bhc/structs.hpp:376-380, 397-424)."""
    structs = {
        "bhc_arrival": _abi.Arrival, "bhc_eigen_hit": _abi.EigenHit, "bhc_ray_pt2d": _abi.RayPt2D,
        "bhc_halfspace_in": _abi.HalfspaceIn, "bhc_bdry_in": _abi.BdryIn, "bhc_refl_in": _abi.ReflIn,
        "bhc_env2d": _abi.Env2D_C, "bhc_bdry_node2d": _abi.BdryNode2D, "bhc_hs2d": _abi.HS2D,
        "bhc_problem2d": _abi.Problem2D, "bhc_field_out": _abi.FieldOut, "bhc_ray_out": _abi.RayOut,
    }
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void){"]
    for cname, cls in structs.items():
        src.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in cls._fields_:
            if fname.startswith("pad"):
                continue
            cf = "is" if fname == "is_" else fname
            src.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, cf))
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(c), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            if fname.startswith("pad"):
                continue
            assert int(got["%s.%s" % (cname, fname)]) == getattr(cls, fname).offset, (cname, fname)
    assert C.sizeof(_abi.Arrival) == 40
    assert C.sizeof(_abi.EigenHit) == 36
    assert C.sizeof(_abi.RayPt2D) == 112


# --------------------------------------------------------------------------
# preprocessing
# --------------------------------------------------------------------------

def _view_arrays(p):
    v = p.view
    N = v.ssp_npts
    a = lambda ptr, n: np.ctypeslib.as_array(ptr, (n,)).copy() if n else np.zeros(0)  # noqa: E731
    out = [a(v.ssp_c, 2 * N), a(v.ssp_cz, 2 * N), a(v.ssp_coef, 8 * N), a(v.ssp_n2, 2 * N),
           a(v.ssp_n2z, 2 * N), a(v.ssp_rho, N), a(v.Rr, v.NRr), a(v.alpha, v.Nalpha)]
    nodes = []
    for i in range(v.bty_npts):
        b = v.bty[i]
        nodes += [b.x[0], b.x[1], b.t[0], b.t[1], b.n[0], b.n[1], b.Noden[0], b.Noden[1],
                  b.Nodet[0], b.Nodet[1], b.Len, b.kappa, 0.0]
    out.append(np.array(nodes))
    out.append(np.array([v.deltas, v.box_r, v.Dalpha, v.Delta_r, v.bot_hs.cP[0], v.bot_hs.cP[1],
                         v.bot_hs.cS[0], v.bot_hs.cS[1]]))
    return np.concatenate(out)


PRE_CASES = {
    "munk_C": lambda: munk_env(nbeams=101, nrz=11, nrr=21),
    "munk_S": lambda: munk_env(nbeams=101, nrz=11, nrr=21, ssp_type="S"),
    "munk_P": lambda: munk_env(nbeams=101, nrz=11, nrr=21, ssp_type="P"),
    "munk_N": lambda: munk_env(nbeams=101, nrz=11, nrr=21, ssp_type="N"),
    "munk_C_f32": lambda: munk_env(nbeams=101, nrz=11, nrr=21, precision=32),
    "seamount": lambda: c2_env(nbeams=101, nrz=11, nrr=21),
    "seamount_f32": lambda: c2_env(nbeams=101, nrz=11, nrr=21, precision=32),
    "biological_layers": lambda: cases.biological_layers(),
    "biological_layers_f32": lambda: cases.biological_layers(32),
}


@pytest.mark.parametrize("name", list(PRE_CASES))
def test_preprocessing_matches_oracle_bitwise(name):
    env = PRE_CASES[name]()
    ref = oracle.preprocess_dump(env)
    got = _view_arrays(bh.Problem(env))
    if env.precision == 32:
        # FP32 build: the oracle's values are floats widened to double
        got = got.astype(np.float32).astype(np.float64)
    assert ref.shape == got.shape
    assert np.array_equal(ref, got, equal_nan=True), np.flatnonzero(~((ref == got) | (np.isnan(ref) & np.isnan(got))))[:10]


@pytest.mark.parametrize("precision", [64, 32])
def test_power_law_attenuation_equals_db_per_m_at_freq0(precision):
    """Attenuation unit 'm' (dB/m with a power law in f / freq0,
    src/module/atten.cpp:120-128) preprocesses bit for bit as 'M' at the run
    frequency freq0, in the SSP and in the halfspaces."""
    a, b = cases.attenuation_units(precision, "m "), cases.attenuation_units(precision, "M ")
    va, vb = _view_arrays(bh.Problem(a)), _view_arrays(bh.Problem(b))
    assert np.array_equal(va, vb, equal_nan=True)
    c = cases.attenuation_units(precision, "N ")
    assert not np.array_equal(va, _view_arrays(bh.Problem(c)), equal_nan=True)


def test_preprocess_rejects_unsupported_inputs():
    env = munk_env(nbeams=11, nrz=3, nrr=3, run_type="CR")       # Cerveny beams: not on this path
    with pytest.raises(bh.BhcError):
        bh.Problem(env)
    env = munk_env(nbeams=11, nrz=3, nrr=3)
    env.ssp_z = env.ssp_z[::-1].copy()                               # depths must increase
    with pytest.raises(bh.BhcError):
        bh.Problem(env)


# --------------------------------------------------------------------------
# tracer core on the CPU
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def harness():
    if not os.path.exists(HARNESS_SO) or os.path.getmtime(HARNESS_SO) < max(
            os.path.getmtime(HARNESS_SRC), os.path.getmtime(os.path.join(ROOT, "bellhopcuda_amd", "csrc",
                                                                         "core2d.h"))):
        subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                        HARNESS_SRC, "-o", HARNESS_SO], check=True)
    H = C.CDLL(HARNESS_SO)
    H.core_field_run2d.argtypes = [C.POINTER(_abi.Problem2D), C.POINTER(_abi.FieldOut)]
    H.core_field_run2d.restype = C.c_int
    return H


CORE_CASES = {
    "munk_hat_cart": lambda: munk_env(nbeams=301, nrz=51, nrr=201),
    "munk_hat_cart_f32": lambda: munk_env(nbeams=301, nrz=51, nrr=201, precision=32),
    "munk_gauss_cart": lambda: munk_env(nbeams=301, nrz=51, nrr=201, run_type="CB"),
    "munk_hat_raycen": lambda: munk_env(nbeams=301, nrz=51, nrr=201, run_type="Cg"),
    "munk_gauss_raycen": lambda: munk_env(nbeams=301, nrz=51, nrr=201, run_type="Cb"),
    "munk_incoherent": lambda: munk_env(nbeams=301, nrz=51, nrr=201, run_type="IG"),
    "munk_semicoherent": lambda: munk_env(nbeams=301, nrz=51, nrr=201, run_type="SG"),
    "munk_spline": lambda: munk_env(nbeams=301, nrz=51, nrr=201, ssp_type="S"),
    "munk_pchip": lambda: munk_env(nbeams=301, nrz=51, nrr=201, ssp_type="P"),
    "munk_n2linear": lambda: munk_env(nbeams=301, nrz=51, nrr=201, ssp_type="N"),
    "seamount": lambda: c2_env(nbeams=201, nrz=41, nrr=201),
    "seamount_f32": lambda: c2_env(nbeams=201, nrz=41, nrr=201, precision=32),
}



for _n, _f in cases.FEATURES.items():
    CORE_CASES[_n] = _f
    CORE_CASES[_n + "_f32"] = (lambda f: lambda: f(32))(_f)


@pytest.mark.parametrize("name", list(CORE_CASES))
def test_core_matches_oracle_bitwise(name, harness):
    env = CORE_CASES[name]()
    o = oracle.field_run(env, nthreads=0)
    p = bh.Problem(env)
    res = bh.FieldResult()
    out = bh._make_out(p, res, True, 0, 1, 0)
    assert harness.core_field_run2d(C.byref(p.view), C.byref(out)) == 0
    assert (out.err, out.warn) == (o.err, o.warn)
    assert np.array_equal(res.nsteps, o.nsteps)
    assert np.array_equal(res.ntop, o.ntop) and np.array_equal(res.nbot, o.nbot)
    assert np.array_equal(res.xend, o.xend)
    assert out.total_steps == o.total_steps and out.total_contrib == o.total_contrib
    # the field is a sum of FP32 atomics in thread order: equal up to summation order
    a = np.abs(o.field)
    assert np.allclose(res.field, o.field, rtol=2e-5, atol=2e-6 * a.max())


# --------------------------------------------------------------------------
# post-processing and files
# --------------------------------------------------------------------------

@pytest.mark.parametrize("run_type", ["CG", "IG", "Cb"])
def test_scale_pressure_matches_oracle(run_type):
    env = munk_env(nbeams=11, nrz=7, nrr=9, run_type=run_type)
    rng = np.random.default_rng(1)
    u = (rng.random((1, 7, 9)) + 1j * rng.random((1, 7, 9))).astype(np.complex64)
    if run_type[0] == "I":
        u = u.real.astype(np.complex64)
    assert np.array_equal(bh.Problem(env).scale_pressure(u), oracle.scale_pressure(env, u))


def _read_shd(path):
    """Reader of the .shd layout (ReadOutTL, src/mode/tl.cpp:263-363)."""
    raw = open(path, "rb").read()
    recl = struct.unpack_from("<i", raw, 0)[0] * 4
    rec = lambda r: raw[r * recl:(r + 1) * recl]  # noqa: E731
    title = rec(0)[4:84].decode()
    plot = rec(1)[:10].decode()
    nfreq, ntheta, nsx, nsy, nsz, nrz, nrr = struct.unpack_from("<7i", rec(2), 0)
    return dict(recl=recl, size=len(raw), title=title, plot=plot, dims=(nfreq, ntheta, nsx, nsy, nsz, nrz, nrr),
                rec=rec)


@pytest.mark.parametrize("precision", [64, 32])
def test_shd_file_layout(tmp_path, precision):
    env = munk_env(nbeams=11, nrz=5, nrr=13, precision=precision)
    p = bh.Problem(env)
    rng = np.random.default_rng(2)
    field = (rng.random((1, 5, 13)) + 1j * rng.random((1, 5, 13))).astype(np.complex64)
    path = str(tmp_path / "t.shd")
    p.write_shd(path, field, title="layout test")
    f = _read_shd(path)
    rs = 8 if precision == 64 else 4
    rt = "<d" if precision == 64 else "<f"
    assert f["recl"] == max(84, 13 * 8)
    assert f["title"].rstrip() == "layout test"
    assert f["plot"] == "rectilin  "
    assert f["dims"] == (1, 1, 1, 1, 1, 5, 13)
    assert struct.unpack_from(rt, f["rec"](2), 28)[0] == pytest.approx(50.0)
    assert struct.unpack_from(rt, f["rec"](3), 0)[0] == pytest.approx(50.0)
    assert struct.unpack_from("<f", f["rec"](7), 0)[0] == np.float32(1000.0)
    assert np.array_equal(np.frombuffer(f["rec"](8)[:20], "<f4"), env.Rz)
    rr = np.frombuffer(f["rec"](9)[:13 * rs], rt)
    assert np.allclose(rr, env.Rr_km * 1000.0, rtol=1e-7 if precision == 32 else 1e-15)
    for irz in range(5):
        row = np.frombuffer(f["rec"](10 + irz)[:13 * 8], "<c8")
        assert np.array_equal(row, field[0, irz])
    assert f["size"] == (10 + 5) * f["recl"]


def test_no_cpu_fallback_without_gpu():
    """On a machine without a HIP device the run calls raise."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(bh.BhcError):
        bh.Context(0)
    with pytest.raises(bh.BhcError):
        bh.field_run(munk_env(nbeams=11, nrz=3, nrr=3))


def test_subtab_matches_fma_rule():
    """SubTab (src/common_setup.hpp:282-316): x[i] = fma(i, (x1 - x0)/(N-1), x0)."""
    from bellhopcuda_amd.env import _libm
    x = subtab(-20.0, 20.0, 2001)
    d = 40.0 / 2000
    for i in (0, 1, 777, 1000, 2000):
        assert x[i] == _libm.fma(float(i), d, -20.0)
    assert np.all(np.diff(x) > 0) and abs(x[-1] - 20.0) < 1e-12
    xf = subtab(0.0, 5000.0, 201, np.float32)
    assert xf.dtype == np.float32 and xf[-1] == np.float32(5000.0)


def test_sweep_environments_are_deterministic():
    a = bh.sweep_envs(n=8, scale=0.05)
    b = bh.sweep_envs(n=8, scale=0.05)
    for x, y in zip(a, b):
        assert x.ssp_type == y.ssp_type and np.array_equal(x.ssp_c, y.ssp_c)
        assert np.array_equal(x.bty.z, y.bty.z) and x.freq0 == y.freq0
    assert {e.ssp_type for e in bh.sweep_envs(n=64, scale=0.01)} == {"C", "P"}
    assert isinstance(a[0].bty, Boundary)


@pytest.mark.parametrize("name", ["munk_hat_cart", "munk_gauss_raycen", "munk_incoherent", "seamount",
                                  "seamount_f32"])
def test_core_fast_phasor_tl(name, harness):
    """The FP32-accuracy phasor (FP64 range reduction + FP32 sincos / exp) of
    the default GPU scatter keeps the TL within 0.01 dB of the oracle where
    TL < 120 dB; trajectories and contribution counts are unaffected."""
    harness.core_field_run2d_fast.argtypes = [C.POINTER(_abi.Problem2D), C.POINTER(_abi.FieldOut)]
    env = CORE_CASES[name]()
    o = oracle.field_run(env, nthreads=0)
    p = bh.Problem(env)
    res = bh.FieldResult()
    out = bh._make_out(p, res, True, 0, 1, 0)
    assert harness.core_field_run2d_fast(C.byref(p.view), C.byref(out)) == 0
    assert np.array_equal(res.nsteps, o.nsteps) and out.total_contrib == o.total_contrib
    so = oracle.scale_pressure(env, o.field)
    sg = p.scale_pressure(res.field)
    tlo = -20 * np.log10(np.maximum(np.abs(so), 1e-30))
    tlg = -20 * np.log10(np.maximum(np.abs(sg), 1e-30))
    m = tlo < 120
    assert m.sum() > 100
    assert np.max(np.abs(tlo - tlg)[m]) < 0.01


def test_entry_points_reject_bad_arguments_before_any_device_call():
    """Argument checks of the sweep-mapping setter and the 3D re-trace fail
    with a message and touch no device (run here without one)."""
    L = bh.lib()
    assert L.bhc_sweep_set_scatter(None, 1) != 0
    assert L.bhc_sweep_set_scatter(None, 7) != 0
    assert b"bhc_sweep_set_scatter" in L.bhc_last_error()
    assert L.bhc_ray_run3d_list(None, None, None, 0, None) != 0
    assert b"bhc_ray_run3d_list" in L.bhc_last_error()


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)
