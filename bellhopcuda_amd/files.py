"""Output files in the reference's layouts: writers (through libbhc_hip.so)
and readers, 2D and (for arrivals) 3D.

* .shd  -- Fortran direct-access TL file (src/mode/tl.cpp:38-101, 211-248;
           read back as ReadOutTL does, :263-363);
* .arr  -- arrivals, ASCII or Fortran-unformatted binary (src/mode/arr.cpp:88-238);
* .ray  -- ray paths, ASCII (src/mode/ray.hpp:168-244).
"""
import ctypes as C
import struct

import numpy as np

from . import _abi, lib, BhcError


# --------------------------------------------------------------------------
# writers
# --------------------------------------------------------------------------

def postprocess_arrivals(problem, arr, narr, max_narr):
    """PostProcessArrivals in place; returns the maximum count per source."""
    maxn = np.zeros(problem.view.NSz, np.int32)
    narr_c = np.ascontiguousarray(narr, np.int32)
    rc = lib().bhc_postprocess_arrivals2d(C.byref(problem.view), C.cast(arr, C.POINTER(_abi.Arrival)),
                                          narr_c.ctypes.data_as(_abi.c_int32_p), max_narr,
                                          maxn.ctypes.data_as(_abi.c_int32_p))
    if rc != 0:
        raise BhcError("postprocess_arrivals failed")
    return narr_c, maxn


def postprocess_arrivals3d(problem, arr, narr, max_narr):
    """PostProcessArrivals for a 3D problem in place (counts clamped, amplitudes
    unscaled); returns (narr, maximum count per source [NSz * NSx * NSy])."""
    v = problem.view
    maxn = np.zeros(v.NSz * v.NSx * v.NSy, np.int32)
    narr_c = np.ascontiguousarray(narr, np.int32)
    rc = lib().bhc_postprocess_arrivals3d(C.byref(v), C.cast(arr, C.POINTER(_abi.Arrival)),
                                          narr_c.ctypes.data_as(_abi.c_int32_p), max_narr,
                                          maxn.ctypes.data_as(_abi.c_int32_p))
    if rc != 0:
        raise BhcError("postprocess_arrivals3d failed")
    return narr_c, maxn


def write_arr3d(path, problem, arr, narr, max_narr, maxn, binary=False):
    narr_c = np.ascontiguousarray(narr, np.int32)
    maxn_c = np.ascontiguousarray(maxn, np.int32)
    rc = lib().bhc_write_arr3d(path.encode(), C.byref(problem.view), C.cast(arr, C.POINTER(_abi.Arrival)),
                               narr_c.ctypes.data_as(_abi.c_int32_p), max_narr,
                               maxn_c.ctypes.data_as(_abi.c_int32_p), 1 if binary else 0)
    if rc != 0:
        raise BhcError("writing %s failed" % path)


def write_arr(path, problem, arr, narr, max_narr, maxn, binary=False):
    narr_c = np.ascontiguousarray(narr, np.int32)
    maxn_c = np.ascontiguousarray(maxn, np.int32)
    rc = lib().bhc_write_arr2d(path.encode(), C.byref(problem.view), C.cast(arr, C.POINTER(_abi.Arrival)),
                               narr_c.ctypes.data_as(_abi.c_int32_p), max_narr,
                               maxn_c.ctypes.data_as(_abi.c_int32_p), 1 if binary else 0)
    if rc != 0:
        raise BhcError("writing %s failed" % path)


def write_ray(path, problem, points, nsteps, decl_deg, title=None):
    """points: RayPt2D array of shape [nrays, max_points]."""
    pts = np.ascontiguousarray(points)
    nrays, maxp = pts.shape
    ns = np.ascontiguousarray(nsteps, np.int32)
    dd = np.ascontiguousarray(decl_deg, np.float64)
    t = (title if title is not None else problem.env.title).encode()
    rc = lib().bhc_write_ray2d(path.encode(), C.byref(problem.view), t,
                               pts.ctypes.data_as(C.POINTER(_abi.RayPt2D)), ns.ctypes.data_as(_abi.c_int32_p),
                               maxp, dd.ctypes.data_as(_abi.c_double_p), nrays)
    if rc != 0:
        raise BhcError("writing %s failed" % path)


# --------------------------------------------------------------------------
# readers
# --------------------------------------------------------------------------

def read_shd(path):
    """ReadOutTL: header dict and the complex field [NSz][NRz_per_range][NRr]
    (NRz_per_range = NRz for rectilinear grids, 1 for irregular ones)."""
    raw = open(path, "rb").read()
    recl = struct.unpack_from("<i", raw, 0)[0] * 4
    rec = lambda r: raw[r * recl:(r + 1) * recl]  # noqa: E731
    title = rec(0)[4:84].decode().rstrip()
    plot = rec(1)[:10].decode()
    nfreq, ntheta, nsx, nsy, nsz, nrz, nrr = struct.unpack_from("<7i", rec(2), 0)
    # reals are 8 bytes in 64-bit files: the record-2 tail tells (freq0, atten)
    wide = struct.unpack_from("<d", rec(3), 0)[0]
    narrow = struct.unpack_from("<f", rec(3), 0)[0]
    f8 = abs(wide) > 1e-30 and abs(wide) < 1e12 and np.isfinite(wide)
    rt = "<f8" if f8 else "<f4"
    freq0 = wide if f8 else narrow
    sz = np.frombuffer(rec(7)[:4 * nsz], "<f4").copy()
    rz = np.frombuffer(rec(8)[:4 * nrz], "<f4").copy()
    rr = np.frombuffer(rec(9)[:np.dtype(rt).itemsize * nrr], rt).astype(np.float64)
    nzr = 1 if plot.startswith("irregular") else nrz
    field = np.zeros((nsz, nzr, nrr), np.complex64)
    for isz in range(nsz):
        for iz in range(nzr):
            field[isz, iz] = np.frombuffer(rec(10 + isz * nzr + iz)[:8 * nrr], "<c8")
    return {"title": title, "plot_type": plot, "freq0": freq0, "Sz": sz, "Rz": rz, "Rr": rr,
            "dims": (nfreq, ntheta, nsx, nsy, nsz, nrz, nrr)}, field


def read_arr(path, binary=False):
    """Arrivals file -> dict(freq0, Sz, Rz, Rr, maxn, arrivals), arrivals as a
    list over [isz][iz][ir] of record arrays (a, phase_deg, delay_re, delay_im,
    src_decl, rcvr_decl, ntop, nbot)."""
    dt = np.dtype([("a", "f8"), ("phase_deg", "f8"), ("delay_re", "f8"), ("delay_im", "f8"),
                   ("src_decl", "f8"), ("rcvr_decl", "f8"), ("ntop", "i4"), ("nbot", "i4")])
    if not binary:
        lines = open(path).read().split("\n")
        pos = 0

        def nxt():
            nonlocal pos
            v = lines[pos].split()
            pos += 1
            return v
        assert nxt()[0] == "'2D'"
        freq0 = float(nxt()[0])
        v = nxt()
        sz = np.array(v[1:int(v[0]) + 1], np.float64)
        v = nxt()
        rz = np.array(v[1:int(v[0]) + 1], np.float64)
        v = nxt()
        rr = np.array(v[1:int(v[0]) + 1], np.float64)
        out, maxn = [], []
        for _ in range(len(sz)):
            maxn.append(int(nxt()[0]))
            per = []
            for _ in range(len(rz) * len(rr)):
                n = int(nxt()[0])
                a = np.zeros(n, dt)
                for k in range(n):
                    t = nxt()
                    a[k] = (float(t[0]), float(t[1]), float(t[2]), float(t[3]), float(t[4]), float(t[5]),
                            int(t[6]), int(t[7]))
                per.append(a)
            out.append(per)
        return {"freq0": freq0, "Sz": sz, "Rz": rz, "Rr": rr, "maxn": maxn, "arrivals": out}
    raw = open(path, "rb").read()
    recs, pos = [], 0
    while pos < len(raw):
        n = struct.unpack_from("<i", raw, pos)[0]
        recs.append(raw[pos + 4:pos + 4 + n])
        assert struct.unpack_from("<i", raw, pos + 4 + n)[0] == n
        pos += n + 8
    assert recs[0] == b"'2D'"
    freq0 = struct.unpack("<f", recs[1])[0]
    nsz = struct.unpack_from("<i", recs[2], 0)[0]
    sz = np.frombuffer(recs[2][4:4 + 4 * nsz], "<f4").astype(np.float64)
    nrz = struct.unpack_from("<i", recs[3], 0)[0]
    rz = np.frombuffer(recs[3][4:4 + 4 * nrz], "<f4").astype(np.float64)
    nrr = struct.unpack_from("<i", recs[4], 0)[0]
    w = (len(recs[4]) - 4) // max(nrr, 1)
    rr = np.frombuffer(recs[4][4:4 + w * nrr], "<f8" if w == 8 else "<f4").astype(np.float64)
    k = 5
    out, maxn = [], []
    for _ in range(nsz):
        maxn.append(struct.unpack("<i", recs[k])[0])
        k += 1
        per = []
        for _ in range(nrz * nrr):
            n = struct.unpack("<i", recs[k])[0]
            k += 1
            a = np.zeros(n, dt)
            for j in range(n):
                f = struct.unpack("<8f", recs[k])
                k += 1
                a[j] = (f[0], f[1], f[2], f[3], f[4], f[5], int(f[6]), int(f[7]))
            per.append(a)
        out.append(per)
    return {"freq0": freq0, "Sz": sz, "Rz": rz, "Rr": rr, "maxn": maxn, "arrivals": out}


ARR3_DTYPE = np.dtype([("a", "f8"), ("phase_deg", "f8"), ("delay_re", "f8"), ("delay_im", "f8"),
                       ("src_decl", "f8"), ("src_azim", "f8"), ("rcvr_decl", "f8"), ("rcvr_azim", "f8"),
                       ("ntop", "i4"), ("nbot", "i4")])


def read_arr3d(path, binary=False):
    """3D arrivals file -> dict(freq0, Sx, Sy, Sz, Rz, Rr, theta, maxn,
    arrivals), arrivals a list over sources [isz][isx][isy] (flattened) of lists
    over receivers [itheta][iz][ir] (flattened) of ARR3_DTYPE record arrays."""
    if not binary:
        toks = open(path).read().split("\n")
        pos = 0

        def nxt():
            nonlocal pos
            v = toks[pos].split()
            pos += 1
            return v

        assert nxt()[0] == "'3D'"
        freq0 = float(nxt()[0])
        vecs = []
        for _ in range(6):
            v = nxt()
            vecs.append(np.array(v[1:int(v[0]) + 1], np.float64))
        sx, sy, sz, rz, rr, th = vecs
        nrcv = len(th) * len(rz) * len(rr)
        out, maxn = [], []
        for _ in range(len(sx) * len(sy) * len(sz)):
            maxn.append(int(nxt()[0]))
            per = []
            for _ in range(nrcv):
                n = int(nxt()[0])
                a = np.zeros(n, ARR3_DTYPE)
                for k in range(n):
                    t = nxt()
                    a[k] = tuple(float(x) for x in t[:8]) + (int(t[8]), int(t[9]))
                per.append(a)
            out.append(per)
        return {"freq0": freq0, "Sx": sx, "Sy": sy, "Sz": sz, "Rz": rz, "Rr": rr, "theta": th, "maxn": maxn,
                "arrivals": out}
    raw = open(path, "rb").read()
    recs, pos = [], 0
    while pos < len(raw):
        n = struct.unpack_from("<i", raw, pos)[0]
        recs.append(raw[pos + 4:pos + 4 + n])
        assert struct.unpack_from("<i", raw, pos + 4 + n)[0] == n
        pos += n + 8
    assert recs[0] == b"'3D'"
    freq0 = struct.unpack("<f", recs[1])[0]

    def vec(rec, float_only=False):
        n = struct.unpack_from("<i", rec, 0)[0]
        w = 4 if float_only else (len(rec) - 4) // max(n, 1)
        return np.frombuffer(rec[4:4 + w * n], "<f8" if w == 8 else "<f4").astype(np.float64)

    sx, sy = vec(recs[2]), vec(recs[3])
    sz, rz = vec(recs[4], True), vec(recs[5], True)
    rr, th = vec(recs[6]), vec(recs[7])
    k = 8
    nrcv = len(th) * len(rz) * len(rr)
    out, maxn = [], []
    for _ in range(len(sx) * len(sy) * len(sz)):
        maxn.append(struct.unpack("<i", recs[k])[0])
        k += 1
        per = []
        for _ in range(nrcv):
            n = struct.unpack("<i", recs[k])[0]
            k += 1
            a = np.zeros(n, ARR3_DTYPE)
            for j in range(n):
                f = struct.unpack("<10f", recs[k])
                k += 1
                a[j] = tuple(f[:8]) + (int(f[8]), int(f[9]))
            per.append(a)
        out.append(per)
    return {"freq0": freq0, "Sx": sx, "Sy": sy, "Sz": sz, "Rz": rz, "Rr": rr, "theta": th, "maxn": maxn,
            "arrivals": out}


def read_ray(path):
    """Ray file -> dict(title, freq0, NSz, Nalpha, top, bot, rays), rays a list
    of (decl_deg, ntop, nbot, points[n, 2])."""
    lines = open(path).read().split("\n")
    title = lines[0].strip().strip("'")
    freq0 = float(lines[1].split()[0])
    nsx, nsy, nsz = (int(v) for v in lines[2].split())
    nalpha, nbeta = (int(v) for v in lines[3].split())
    top = float(lines[4].split()[0])
    bot = float(lines[5].split()[0])
    assert lines[6].strip() == "'rz'"
    rays, pos = [], 7
    while pos < len(lines) and lines[pos].strip():
        decl = float(lines[pos].split()[0])
        n, ntop, nbot = (int(v) for v in lines[pos + 1].split())
        pts = np.array([[float(v) for v in lines[pos + 2 + k].split()] for k in range(n)], np.float64)
        rays.append((decl, ntop, nbot, pts))
        pos += 2 + n
    return {"title": title, "freq0": freq0, "NSz": nsz, "Nalpha": nalpha, "top": top, "bot": bot,
            "rays": rays}
