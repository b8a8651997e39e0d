"""ORACLE loader -- test infrastructure only (tests/, smoke(), bench.py cpu_baseline).

ctypes binding of liboracle.so, the plain-C restatement of the reference's 2D
hot path in oracle/bellhop2d.c. Never imported by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from bellhopcuda_amd import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        for sfx in ("d", "f"):
            f = getattr(L, "oracle_field_run2d_" + sfx)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(_abi.Env2D_C), C.POINTER(_abi.FieldOut), C.c_int, C.c_int, C.c_int]
            f = getattr(L, "oracle_ray_run2d_" + sfx)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(_abi.Env2D_C), C.POINTER(_abi.RayOut), _abi.c_int32_p,
                          _abi.c_int32_p, _abi.c_int32_p, C.c_int]
            f = getattr(L, "oracle_scale_pressure2d_" + sfx)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(_abi.Env2D_C), _abi.c_float_p]
            f = getattr(L, "oracle_preprocess_dump2d_" + sfx)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(_abi.Env2D_C), _abi.c_double_p, C.c_int64, C.POINTER(C.c_int64)]
            f = getattr(L, "oracle_arrlog_" + sfx)
            f.restype = C.c_int
            f.argtypes = [_abi.c_double_p, C.c_int64]
            f = getattr(L, "oracle_arrlog_count_" + sfx)
            f.restype = C.c_int64
            f.argtypes = []
        _LIB = L
    return _LIB


def _sfx(env):
    return "d" if env.precision == 64 else "f"


class FieldResult:
    pass


def field_run(env, nthreads=0, job_begin=0, job_end=-1, stats=True, max_narr=0, allow_merging=1,
              eigen_memsize=0):
    """Run the oracle on an Env2D. Returns a FieldResult with .field (complex64,
    shape [NSz, NRz_per_range, NRr], pre-ScalePressure) and per-ray stats."""
    ec, keep = env.to_c()
    nsz, nrz, nrr = env.field_shape
    nj = env.nrays if job_end < 0 else job_end - job_begin
    res = FieldResult()
    res.field = np.zeros((nsz, nrz, nrr), np.complex64)
    out = _abi.FieldOut()
    out.field = res.field.ctypes.data_as(_abi.c_float_p)
    rt = env.run_type[0]
    if rt in "AaVv":
        res.max_narr = max_narr
        res.arr = (_abi.Arrival * (nsz * nrz * nrr * max_narr))()
        res.narr = np.zeros(nsz * nrz * nrr, np.int32)
        out.arr = C.cast(res.arr, C.POINTER(_abi.Arrival))
        out.narr = res.narr.ctypes.data_as(_abi.c_int32_p)
        out.max_narr = max_narr
        out.allow_merging = allow_merging
    if rt in "EVv":
        res.hits = (_abi.EigenHit * max(eigen_memsize, 1))()
        res.neigen = np.zeros(1, np.int32)
        out.hits = C.cast(res.hits, C.POINTER(_abi.EigenHit))
        out.neigen = res.neigen.ctypes.data_as(_abi.c_int32_p)
        out.eigen_memsize = eigen_memsize
    if stats:
        res.nsteps = np.zeros(nj, np.int32)
        res.ntop = np.zeros(nj, np.int32)
        res.nbot = np.zeros(nj, np.int32)
        res.xend = np.zeros((nj, 2), np.float64)
        out.ray_nsteps = res.nsteps.ctypes.data_as(_abi.c_int32_p)
        out.ray_ntop = res.ntop.ctypes.data_as(_abi.c_int32_p)
        out.ray_nbot = res.nbot.ctypes.data_as(_abi.c_int32_p)
        out.ray_xend = res.xend.ctypes.data_as(_abi.c_double_p)
    f = getattr(lib(), "oracle_field_run2d_" + _sfx(env))
    rc = f(C.byref(ec), C.byref(out), nthreads, job_begin, job_end)
    if rc != 0:
        raise RuntimeError("oracle_field_run2d failed")
    res.err, res.warn = out.err, out.warn
    res.total_steps, res.total_contrib = out.total_steps, out.total_contrib
    del keep
    return res


def scale_pressure(env, field):
    ec, keep = env.to_c()
    f = np.ascontiguousarray(field, np.complex64).copy()
    getattr(lib(), "oracle_scale_pressure2d_" + _sfx(env))(C.byref(ec), f.ctypes.data_as(_abi.c_float_p))
    return f


def ray_run(env, max_points=20000, ialpha=None, isz=None, nsteps_in=None):
    ec, keep = env.to_c()
    n = env.nrays if ialpha is None else len(ialpha)
    pts = (_abi.RayPt2D * (n * max_points))()
    nst = np.zeros(n, np.int32)
    out = _abi.RayOut()
    out.points = C.cast(pts, C.POINTER(_abi.RayPt2D))
    out.nsteps = nst.ctypes.data_as(_abi.c_int32_p)
    out.max_points_per_ray = max_points
    ia = isz_a = ns = None
    if ialpha is not None:
        ia = np.ascontiguousarray(ialpha, np.int32)
        isz_a = np.ascontiguousarray(isz if isz is not None else np.zeros(n), np.int32)
    if nsteps_in is not None:
        ns = np.ascontiguousarray(nsteps_in, np.int32)
    f = getattr(lib(), "oracle_ray_run2d_" + _sfx(env))
    rc = f(C.byref(ec), C.byref(out),
           ia.ctypes.data_as(_abi.c_int32_p) if ia is not None else None,
           isz_a.ctypes.data_as(_abi.c_int32_p) if isz_a is not None else None,
           ns.ctypes.data_as(_abi.c_int32_p) if ns is not None else None, n)
    if rc != 0:
        raise RuntimeError("oracle_ray_run2d failed")
    arr = np.ctypeslib.as_array(pts).reshape(n, max_points)
    return arr, nst, (out.err, out.warn)


def preprocess_dump(env, cap=1 << 22):
    ec, keep = env.to_c()
    buf = np.zeros(cap, np.float64)
    used = C.c_int64(0)
    getattr(lib(), "oracle_preprocess_dump2d_" + _sfx(env))(C.byref(ec), buf.ctypes.data_as(_abi.c_double_p),
                                                           cap, C.byref(used))
    return buf[:used.value].copy()


def _lib3():
    L = lib()
    if not getattr(L, "_bind3", False):
        from bellhopcuda_amd import _abi as A
        for sfx in ("d", "f"):
            f = getattr(L, "oracle_field_run3d_" + sfx)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(A.Env3D_C), C.POINTER(A.FieldOut), C.c_int, C.c_int, C.c_int]
            f = getattr(L, "oracle_field_runnx2d_" + sfx)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(A.Env3D_C), C.POINTER(A.FieldOut), C.c_int, C.c_int, C.c_int]
            f = getattr(L, "oracle_ray_runnx2d_" + sfx)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(A.Env3D_C), C.POINTER(A.RayPt2D), A.c_int32_p, A.c_double_p, C.c_int]
            f = getattr(L, "oracle_preprocess_dump3d_" + sfx)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(A.Env3D_C), A.c_double_p, C.c_int64, C.POINTER(C.c_int64)]
        L._bind3 = True
    return L


def field_run3d(env, nthreads=0, job_begin=0, job_end=-1, max_narr=0, allow_merging=1, eigen_memsize=0):
    """Run the 3D oracle on an Env3D. Returns a FieldResult with .field
    (complex64, shape env.field_shape, pre-ScalePressure) and per-ray stats
    (.nsteps, .ntop, .nbot, .xend [rays, 3]); arrivals runs ('A', 'a') fill
    .arr / .narr ([receivers][max_narr], receivers in field order), eigen
    runs ('E') .hits / .neigen."""
    L = _lib3()
    ec, keep = env.to_c()
    nj = env.nrays
    if job_end < 0 or job_end > nj:
        job_end = nj
    n = job_end - job_begin
    nrcv = int(np.prod(env.field_shape))
    fld = np.zeros(nrcv * 2, np.float32)
    r = FieldResult()
    out = _abi.FieldOut()
    rt = env.run_type[0]
    if rt in "AaVv":
        r.max_narr = max_narr
        r.arr = (_abi.Arrival * (nrcv * max_narr))()
        r.narr = np.zeros(nrcv, np.int32)
        out.arr = C.cast(r.arr, C.POINTER(_abi.Arrival))
        out.narr = r.narr.ctypes.data_as(_abi.c_int32_p)
        out.max_narr = max_narr
        out.allow_merging = allow_merging
    if rt in "EVv":
        r.hits = (_abi.EigenHit * max(eigen_memsize, 1))()
        r.neigen = np.zeros(1, np.int32)
        out.hits = C.cast(r.hits, C.POINTER(_abi.EigenHit))
        out.neigen = r.neigen.ctypes.data_as(_abi.c_int32_p)
        out.eigen_memsize = eigen_memsize
    ns = np.zeros(max(n, 1), np.int32)
    nt = np.zeros(max(n, 1), np.int32)
    nb = np.zeros(max(n, 1), np.int32)
    xe = np.zeros((max(n, 1), 3), np.float64)
    out.field = fld.ctypes.data_as(_abi.c_float_p)
    out.ray_nsteps = ns.ctypes.data_as(_abi.c_int32_p)
    out.ray_ntop = nt.ctypes.data_as(_abi.c_int32_p)
    out.ray_nbot = nb.ctypes.data_as(_abi.c_int32_p)
    out.ray_xend = xe.ctypes.data_as(_abi.c_double_p)
    name = "oracle_field_runnx2d_" if env.nx2d else "oracle_field_run3d_"
    f = getattr(L, name + _sfx(env))
    rc = f(C.byref(ec), C.byref(out), nthreads, job_begin, job_end)
    if rc != 0:
        raise RuntimeError(name + " failed (unsupported environment)")
    r.field = fld.view(np.complex64).reshape(env.field_shape)
    r.nsteps, r.ntop, r.nbot, r.xend = ns[:n], nt[:n], nb[:n], xe[:n]
    r.total_steps, r.total_contrib = out.total_steps, out.total_contrib
    r.err, r.warn = out.err, out.warn
    return r


def scale_pressure3d(env, field):
    """ScalePressure for 3D geometric beams (src/influence.hpp:1556-1627):
    sqrt of the intensity for incoherent / semicoherent runs. Nx2D runs
    (R3D = false) then take the 2D cylindrical spreading -1 / sqrt(|r|) per
    receiver range (0 at r = 0, -4 sqrt(pi) for a line source)."""
    f = np.array(field, np.complex64, copy=True)
    if env.run_type[0] != "C":
        f = np.sqrt(f.real.astype(np.float32)).astype(np.complex64)
    if env.nx2d:
        line = (env.run_type + " " * 6)[3] == "X"
        if env.precision == 64:
            r = np.asarray(env.Rr_km, np.float64) * 1000.0
            fac = np.where(np.abs(r) < np.finfo(np.float64).eps, 0.0, -1.0 / np.sqrt(np.abs(r)))
            if line:
                fac[:] = np.float64(-4.0 * np.float32(np.sqrt(np.float32(3.14159265))) * -1.0)
        else:
            r = (np.asarray(env.Rr_km, np.float32) * np.float32(1000.0)).astype(np.float32)
            with np.errstate(divide="ignore"):
                fac = np.where(np.abs(r) < np.finfo(np.float32).eps, np.float32(0.0),
                               np.float32(-1.0) / np.sqrt(np.abs(r))).astype(np.float32)
            if line:
                fac[:] = np.float32(-4.0) * np.sqrt(np.float32(3.14159265)) * np.float32(-1.0)
        f = (f * fac.astype(np.float32)).astype(np.complex64)
    return f


def preprocess_dump3d(env, cap=1 << 22):
    L = _lib3()
    ec, keep = env.to_c()
    dst = np.zeros(cap, np.float64)
    used = C.c_int64()
    f = getattr(L, "oracle_preprocess_dump3d_" + _sfx(env))
    if f(C.byref(ec), dst.ctypes.data_as(_abi.c_double_p), cap, C.byref(used)) != 0:
        raise RuntimeError("oracle_preprocess_dump3d failed")
    return dst[:min(used.value, cap)]


def ray_run3d(env, max_points=4000):
    """3D ray run of the oracle: (points [rays, max_points] of bhc_ray_pt3d, nsteps [rays])."""
    L = _lib3()
    from bellhopcuda_amd import _abi as A
    f = getattr(L, "oracle_ray_run3d_" + _sfx(env))
    f.restype = C.c_int
    f.argtypes = [C.POINTER(A.Env3D_C), C.POINTER(A.RayPt3D), A.c_int32_p, C.c_int]
    ec, keep = env.to_c()
    n = env.nrays
    pts = (A.RayPt3D * (n * max_points))()
    ns = np.zeros(n, np.int32)
    if f(C.byref(ec), pts, ns.ctypes.data_as(A.c_int32_p), max_points) != 0:
        raise RuntimeError("oracle_ray_run3d failed")
    return pts, ns


def ray_run3d_list(env, hits, nhits, max_points=4000):
    """3D eigenray re-trace of the oracle: ray k from hit k, stopped after its
    step. Returns (points [nhits, max_points], nsteps [nhits])."""
    L = _lib3()
    from bellhopcuda_amd import _abi as A
    f = getattr(L, "oracle_ray_run3d_list_" + _sfx(env))
    f.restype = C.c_int
    f.argtypes = [C.POINTER(A.Env3D_C), C.POINTER(A.EigenHit), C.c_int32, C.POINTER(A.RayPt3D), A.c_int32_p, C.c_int]
    ec, keep = env.to_c()
    pts = (A.RayPt3D * (nhits * max_points))()
    ns = np.zeros(nhits, np.int32)
    if f(C.byref(ec), C.cast(hits, C.POINTER(A.EigenHit)), nhits, pts, ns.ctypes.data_as(A.c_int32_p),
         max_points) != 0:
        raise RuntimeError("oracle_ray_run3d_list failed")
    return pts, ns


def arrivals_with_log(env, max_narr, allow_merging=1, cap=1 << 20):
    """Single-threaded arrivals run that also returns the raw AddArr calls in
    call order ([n, 10]: receiver, Amp, Phase, delay re, delay im, omega,
    source and receiver declination, top and bottom bounces)."""
    buf = np.zeros((cap, 10), np.float64)
    L = lib()
    sfx = _sfx(env)
    getattr(L, "oracle_arrlog_" + sfx)(buf.ctypes.data_as(_abi.c_double_p), cap)
    try:
        res = field_run(env, nthreads=1, max_narr=max_narr, allow_merging=allow_merging)
        n = getattr(L, "oracle_arrlog_count_" + sfx)()
    finally:
        getattr(L, "oracle_arrlog_" + sfx)(None, 0)
    if n > cap:
        raise RuntimeError("arrivals log overflow: %d calls" % n)
    return res, buf[:n].copy()


def ray_runnx2d(env, max_points=4000):
    """Nx2D ray run of the oracle: (points [rays, max_points] of bhc_ray_pt2d in
    ray coordinates, nsteps [rays], origin [rays, 4] = xs.x, xs.y, cos beta,
    sin beta)."""
    L = _lib3()
    from bellhopcuda_amd import _abi as A
    f = getattr(L, "oracle_ray_runnx2d_" + _sfx(env))
    ec, keep = env.to_c()
    n = env.nrays
    pts = (A.RayPt2D * (n * max_points))()
    ns = np.zeros(n, np.int32)
    org = np.zeros((n, 4), np.float64)
    if f(C.byref(ec), pts, ns.ctypes.data_as(A.c_int32_p), org.ctypes.data_as(A.c_double_p), max_points) != 0:
        raise RuntimeError("oracle_ray_runnx2d failed")
    return pts, ns, org
