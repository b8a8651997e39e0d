"""Output files (.arr, .ray, .shd) against the reference's layouts, on the CPU.

The list-directed writer is pinned byte for byte by the reference's own
arrivals fixture (test/individual/analysis_3d/example_analysis_3d.arr.txt,
copied to tests/golden/): every token of that file is re-emitted with the
type its width shows (12-column integers, 15-column floats, 24-column
doubles, quoted strings) and the bytes must be identical -- this covers the
fixed / scientific switch, the sign column, the trailing blanks and the
format flags that carry over between items. The 2D writers are then checked
by round trips through the readers.
"""
import math
import os
import re

import numpy as np
import pytest

import bellhopcuda_amd as bh
from bellhopcuda_amd import files
from bellhopcuda_amd.env import munk_env, subtab
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "example_analysis_3d.arr.txt")


def _kind(tok):
    if tok.startswith("'"):
        return "str"
    if "." not in tok and "E" not in tok:
        return "i32"
    mant = tok.split("E")[0].lstrip("-")
    digits = len(mant.replace(".", "").lstrip("0")) if mant.replace(".", "").strip("0") else len(mant) - 1
    if "E" in tok:
        return "f64" if len(tok.split("E")[1].lstrip("+-")) == 3 else "f32"
    return "f64" if digits > 10 else "f32"


def test_list_directed_writer_matches_reference_fixture(tmp_path):
    ref = open(GOLDEN, "rb").read()
    L = bh.lib()
    out = str(tmp_path / "re.arr")
    f = L.bhc_ld_open(out.encode())
    assert f
    for line in ref.decode().split("\n")[:-1]:
        for tok in re.findall(r"\S+", line):
            k = _kind(tok)
            if k == "str":
                L.bhc_ld_str(f, tok.strip("'").encode())
            elif k == "i32":
                L.bhc_ld_i32(f, int(tok))
            elif k == "f32":
                L.bhc_ld_f32(f, float(np.float32(float(tok))))
            else:
                L.bhc_ld_f64(f, float(tok))
        L.bhc_ld_char(f, b"\n")
    assert L.bhc_ld_close(f) == 0
    got = open(out, "rb").read()
    assert got == ref


def _oracle_arrivals(env, max_narr):
    o = oracle.field_run(env, nthreads=1, max_narr=max_narr, allow_merging=1)
    return o.arr, o.narr


def test_postprocess_arrivals_cylindrical_spreading():
    env = munk_env(nbeams=101, nrz=5, nrr=11, run_type="AG")
    p = bh.Problem(env)
    arr, narr = _oracle_arrivals(env, 20)
    before = np.ctypeslib.as_array(arr).reshape(1, 5, 11, 20)["a"].copy()
    narr2, maxn = files.postprocess_arrivals(p, arr, narr, 20)
    after = np.ctypeslib.as_array(arr).reshape(1, 5, 11, 20)["a"]
    rr = env.Rr_km * 1000.0
    fac = np.where(rr == 0.0, 1e5, 1.0 / np.sqrt(np.where(rr == 0.0, 1.0, rr))).astype(np.float32)
    n = narr2.reshape(1, 5, 11)
    for iz in range(5):
        for ir in range(11):
            k = n[0, iz, ir]
            assert np.array_equal(after[0, iz, ir, :k], (before[0, iz, ir, :k] * fac[ir]).astype(np.float32))
    assert maxn[0] == n.max()


@pytest.mark.parametrize("binary", [False, True])
@pytest.mark.parametrize("precision", [64, 32])
def test_arr_file_roundtrip(tmp_path, binary, precision):
    env = munk_env(nbeams=201, nrz=7, nrr=13, run_type="Ab" if not binary else "ab", precision=precision)
    p = bh.Problem(env)
    arr, narr = _oracle_arrivals(env, 30)
    narr, maxn = files.postprocess_arrivals(p, arr, narr, 30)
    path = str(tmp_path / "t.arr")
    files.write_arr(path, p, arr, narr, 30, maxn, binary=binary)
    r = files.read_arr(path, binary=binary)
    assert r["freq0"] == pytest.approx(env.freq0)
    assert np.allclose(r["Rr"], env.Rr_km * 1000.0) and np.allclose(r["Rz"], env.Rz)
    assert r["maxn"] == [int(maxn[0])]
    A = np.ctypeslib.as_array(arr).reshape(1, 7, 13, 30)
    n = narr.reshape(1, 7, 13)
    rad_deg = np.float32(180.0 / math.pi) if precision == 64 else np.float32(180.0) / np.float32(math.pi)
    for iz in range(7):
        for ir in range(13):
            got = r["arrivals"][0][iz * 13 + ir]
            k = n[0, iz, ir]
            assert len(got) == k
            a = A[0, iz, ir, :k]
            # nine significant digits (15-column floats) return each float32 exactly
            assert np.array_equal(got["a"].astype(np.float32), a["a"])
            assert np.array_equal(got["delay_re"].astype(np.float32), a["delay"][:, 0])
            assert np.array_equal(got["src_decl"].astype(np.float32), a["SrcDeclAngle"])
            assert np.array_equal(got["ntop"], a["NTopBnc"]) and np.array_equal(got["nbot"], a["NBotBnc"])
            if not binary:
                assert np.array_equal(got["phase_deg"].astype(np.float32), (rad_deg * a["Phase"]).astype(np.float32))


@pytest.mark.parametrize("precision", [64, 32])
def test_ray_file_roundtrip(tmp_path, precision):
    env = munk_env(nbeams=21, nrz=3, nrr=3, run_type="R", precision=precision)
    p = bh.Problem(env)
    pts, nst, _ = oracle.ray_run(env, max_points=3000)
    path = str(tmp_path / "t.ray")
    files.write_ray(path, p, pts, nst, env.alpha_deg)
    r = files.read_ray(path)
    assert r["title"] == env.title and r["Nalpha"] == 21 and r["NSz"] == 1
    assert r["top"] == 0.0 and r["bot"] == 5000.0
    assert len(r["rays"]) == 21
    for k, (decl, ntop, nbot, xy) in enumerate(r["rays"]):
        n = nst[k]
        assert decl == pytest.approx(env.alpha_deg[k], abs=1e-12 if precision == 64 else 1e-5)
        assert xy.shape == (n, 2)
        assert ntop == pts[k, n - 1]["NumTopBnc"] and nbot == pts[k, n - 1]["NumBotBnc"]
        if precision == 64:
            # 17 significant digits return every double exactly
            assert np.array_equal(xy, pts[k, :n]["x"])
        else:
            assert np.array_equal(xy.astype(np.float32), pts[k, :n]["x"].astype(np.float32))


def test_shd_file_roundtrip(tmp_path):
    env = munk_env(nbeams=11, nrz=9, nrr=17)
    p = bh.Problem(env)
    rng = np.random.default_rng(3)
    field = (rng.random((1, 9, 17)) + 1j * rng.random((1, 9, 17))).astype(np.complex64)
    path = str(tmp_path / "t.shd")
    p.write_shd(path, field, title="roundtrip")
    hdr, f = files.read_shd(path)
    assert hdr["title"] == "roundtrip" and hdr["dims"] == (1, 1, 1, 1, 1, 9, 17)
    assert hdr["freq0"] == pytest.approx(50.0)
    assert np.array_equal(hdr["Rz"], env.Rz) and np.allclose(hdr["Rr"], env.Rr_km * 1000.0)
    assert np.array_equal(f, field)
