"""bellhopcuda_amd -- MI355X-native BELLHOP ray/beam tracing hot path.

Python host binding of libbhc_hip.so (include/bhc_hip.h). The flow mirrors the
reference's library API for one field run (bhc::setup -> bhc::run ->
bhc::writeout, include/bhc/bhc.hpp:58-374): an `Env2D` (what the .env file
holds) is preprocessed on the host (C++), handed to the HIP launch layer that
replaces RunFieldModesImpl (src/mode/fieldimpl.hpp:27-28), and the results
come back as numpy arrays in the reference's layouts.

There is no CPU execution path: if the HIP library is missing or no GPU is
visible, the run calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from .env import Env2D, munk_env, c2_env, sweep_envs, subtab  # noqa: F401
from .env3d import Env3D, c3_env, free_env3d, munk_env3d, analysis3d_env  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbhc_hip.so")
_LIB = None


class BhcError(RuntimeError):
    pass


# completion callback of the non-blocking runs: void (*)(void *user)
COMPLETED_FN = C.CFUNCTYPE(None, C.c_void_p)


def lib():
    """Load libbhc_hip.so (built by bellhopcuda_amd.build / __graft_entry__.build)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise BhcError("libbhc_hip.so is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    sig = {
        "bhc_preprocess2d": (C.c_int, [P(_abi.Env2D_C), P(C.c_void_p)]),
        "bhc_problem_view": (P(_abi.Problem2D), [C.c_void_p]),
        "bhc_problem_free": (None, [C.c_void_p]),
        "bhc_host_last_error": (C.c_char_p, []),
        "bhc_last_error": (C.c_char_p, []),
        "bhc_device_info": (C.c_int, [C.c_int, C.c_char_p, C.c_int, P(C.c_int), P(C.c_int)]),
        "bhc_ctx_create": (C.c_int, [C.c_int, P(C.c_void_p)]),
        "bhc_ctx_destroy": (None, [C.c_void_p]),
        "bhc_ctx_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "bhc_ctx_synchronize": (C.c_int, [C.c_void_p]),
        "bhc_copy_calibrate": (C.c_int, [C.c_void_p, C.c_longlong, P(C.c_double)]),
        "bhc_last_kernel_ms": (C.c_int, [C.c_void_p, P(C.c_double)]),
        "bhc_field_run2d": (C.c_int, [C.c_void_p, P(_abi.Problem2D), P(_abi.FieldOut)]),
        "bhc_ray_run2d": (C.c_int, [C.c_void_p, P(_abi.Problem2D), P(_abi.RayOut)]),
        "bhc_ray_run2d_list": (C.c_int, [C.c_void_p, P(_abi.Problem2D), P(_abi.RayOut), _abi.c_int32_p,
                                         _abi.c_int32_p, _abi.c_int32_p, C.c_int32]),
        "bhc_problem_upload2d": (C.c_int, [C.c_void_p, P(_abi.Problem2D), P(_abi.FieldOut), P(C.c_void_p)]),
        "bhc_field_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
        "bhc_field_launch_noalloc": (C.c_int, [C.c_void_p, C.c_void_p]),
        "bhc_field_download": (C.c_int, [C.c_void_p, C.c_void_p, P(_abi.FieldOut)]),
        "bhc_problem_release": (None, [C.c_void_p]),
        "bhc_problem_autotune": (C.c_int, [C.c_void_p, C.c_void_p, P(C.c_int)]),
        "bhc_problem_set_scatter": (C.c_int, [C.c_void_p, C.c_int]),
        "bhc_scale_pressure2d": (C.c_int, [P(_abi.Problem2D), _abi.c_float_p]),
        "bhc_write_shd2d": (C.c_int, [C.c_char_p, P(_abi.Problem2D), C.c_char_p, _abi.c_float_p]),
        "bhc_sweep_upload2d": (C.c_int, [C.c_void_p, P(P(_abi.Problem2D)), C.c_int, C.c_int, P(C.c_void_p)]),
        "bhc_sweep_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
        "bhc_sweep_launch_range": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
        "bhc_progress": (C.c_int, [C.c_void_p, P(C.c_int64), P(C.c_int64)]),
        "bhc_field_run2d_async": (C.c_int, [C.c_void_p, P(_abi.Problem2D), P(_abi.FieldOut), COMPLETED_FN,
                                            C.c_void_p]),
        "bhc_field_run3d_async": (C.c_int, [C.c_void_p, P(_abi.Problem3D), P(_abi.FieldOut), COMPLETED_FN,
                                            C.c_void_p]),
        "bhc_ctx_busy": (C.c_int, [C.c_void_p]),
        "bhc_ctx_wait": (C.c_int, [C.c_void_p]),
        "bhc_sweep_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, P(_abi.FieldOut)]),
        "bhc_sweep_stats": (C.c_int, [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_int), P(C.c_longlong),
                                      P(C.c_longlong)]),
        "bhc_sweep_release": (None, [C.c_void_p]),
        "bhc_sweep_autotune": (C.c_int, [C.c_void_p, C.c_void_p, P(C.c_int), P(C.c_int), P(C.c_int)]),
        "bhc_sweep_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
        "bhc_sweep_set_scatter": (C.c_int, [C.c_void_p, C.c_int]),
        "bhc_sweep_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "bhc_last_phase_ms": (C.c_int, [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_int)]),
        "bhc_opcount2d": (C.c_int, [P(_abi.Problem2D), C.c_int, P(C.c_double)]),
        "bhc_subtab": (C.c_int, [C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_void_p]),
        "bhc_postprocess_arrivals2d": (C.c_int, [P(_abi.Problem2D), P(_abi.Arrival), _abi.c_int32_p, C.c_int32,
                                                 _abi.c_int32_p]),
        "bhc_write_arr2d": (C.c_int, [C.c_char_p, P(_abi.Problem2D), P(_abi.Arrival), _abi.c_int32_p, C.c_int32,
                                      _abi.c_int32_p, C.c_int]),
        "bhc_write_ray2d": (C.c_int, [C.c_char_p, P(_abi.Problem2D), C.c_char_p, P(_abi.RayPt2D), _abi.c_int32_p,
                                      C.c_int32, _abi.c_double_p, C.c_int32]),
        "bhc_ld_open": (C.c_void_p, [C.c_char_p]),
        "bhc_ld_i32": (None, [C.c_void_p, C.c_int32]),
        "bhc_ld_f32": (None, [C.c_void_p, C.c_float]),
        "bhc_ld_f64": (None, [C.c_void_p, C.c_double]),
        "bhc_ld_str": (None, [C.c_void_p, C.c_char_p]),
        "bhc_ld_char": (None, [C.c_void_p, C.c_char]),
        "bhc_ld_close": (C.c_int, [C.c_void_p]),
        "bhc_preprocess3d": (C.c_int, [P(_abi.Env3D_C), P(C.c_void_p)]),
        "bhc_problem3d_view": (P(_abi.Problem3D), [C.c_void_p]),
        "bhc_problem3d_free": (None, [C.c_void_p]),
        "bhc_host3d_last_error": (C.c_char_p, []),
        "bhc_field_run3d": (C.c_int, [C.c_void_p, P(_abi.Problem3D), P(_abi.FieldOut)]),
        "bhc_field_runnx2d": (C.c_int, [C.c_void_p, P(_abi.Problem3D), P(_abi.FieldOut)]),
        "bhc_field_runnx2d_async": (C.c_int, [C.c_void_p, P(_abi.Problem3D), P(_abi.FieldOut), COMPLETED_FN,
                                              C.c_void_p]),
        "bhc_ray_runnx2d": (C.c_int, [C.c_void_p, P(_abi.Problem3D), P(_abi.RayOut), _abi.c_double_p]),
        "bhc_scale_pressure3d": (C.c_int, [P(_abi.Problem3D), _abi.c_float_p]),
        "bhc_opcount3d": (C.c_int, [P(_abi.Problem3D), C.c_int, P(C.c_double)]),
        "bhc_ray_run3d": (C.c_int, [C.c_void_p, P(_abi.Problem3D), P(_abi.RayOut3D)]),
        "bhc_ray_run3d_list": (C.c_int, [C.c_void_p, P(_abi.Problem3D), P(_abi.EigenHit), C.c_int32,
                                         P(_abi.RayOut3D)]),
        "bhc_write_ray3d": (C.c_int, [C.c_char_p, P(_abi.Problem3D), C.c_char_p, P(_abi.RayPt3D), _abi.c_int32_p,
                                      C.c_int32, C.c_int32]),
        "bhc_postprocess_arrivals3d": (C.c_int, [P(_abi.Problem3D), P(_abi.Arrival), _abi.c_int32_p, C.c_int32,
                                                 _abi.c_int32_p]),
        "bhc_write_arr3d": (C.c_int, [C.c_char_p, P(_abi.Problem3D), P(_abi.Arrival), _abi.c_int32_p, C.c_int32,
                                      _abi.c_int32_p, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def _check(rc, host=False):
    if rc != 0:
        msg = (lib().bhc_host_last_error() if host else lib().bhc_last_error()) or b"?"
        raise BhcError(msg.decode())


class Problem:
    """A preprocessed 2D problem (bhc_problem2d view owned by the C++ side)."""

    def __init__(self, env):
        self.env = env
        ec, self._keep = env.to_c()
        h = C.c_void_p()
        _check(lib().bhc_preprocess2d(C.byref(ec), C.byref(h)), host=True)
        self._h = h
        self.view = lib().bhc_problem_view(h).contents

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.bhc_problem_free(self._h)
            self._h = None

    @property
    def nrays(self):
        v = self.view
        return (1 if v.iSingle_alpha >= 1 else v.Nalpha) * v.NSz

    @property
    def field_shape(self):
        v = self.view
        return (v.NSz, v.NRz_per_range, v.NRr)

    def scale_pressure(self, field):
        f = np.ascontiguousarray(field, np.complex64).copy()
        _check(lib().bhc_scale_pressure2d(C.byref(self.view), f.ctypes.data_as(_abi.c_float_p)), host=True)
        return f

    def write_shd(self, path, field, title=None):
        f = np.ascontiguousarray(field, np.complex64)
        t = (title or self.env.title).encode()
        rc = lib().bhc_write_shd2d(path.encode(), C.byref(self.view), t, f.ctypes.data_as(_abi.c_float_p))
        if rc != 0:
            raise BhcError("writing %s failed" % path)


class Context:
    """One HIP device context (stream + events)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().bhc_ctx_create(device, C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().bhc_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.bhc_ctx_destroy(self._h)

    def set_launch(self, blocks_per_cu=0):
        _check(lib().bhc_ctx_set_launch(self._h, 0, blocks_per_cu))

    def progress(self):
        """(rays finished, rays of the run) of the current or last run."""
        d, t = C.c_int64(), C.c_int64()
        _check(lib().bhc_progress(self._h, C.byref(d), C.byref(t)))
        return d.value, t.value

    def busy(self):
        return bool(lib().bhc_ctx_busy(self._h))

    def wait(self):
        """Join a non-blocking run; raises if it failed."""
        _check(lib().bhc_ctx_wait(self._h))

    def last_kernel_ms(self):
        v = C.c_double()
        _check(lib().bhc_last_kernel_ms(self._h, C.byref(v)))
        return v.value

    def last_phase(self):
        """(trace ms, scatter ms, passes) of the last single-problem launch."""
        t, sc, p = C.c_double(), C.c_double(), C.c_int()
        _check(lib().bhc_last_phase_ms(self._h, C.byref(t), C.byref(sc), C.byref(p)))
        return t.value, sc.value, p.value


class FieldResult:
    pass


def _make_out(prob, res, stats, max_narr, allow_merging, eigen_memsize):
    nsz, nrz, nrr = prob.field_shape
    out = _abi.FieldOut()
    rt = prob.env.run_type[0]
    if rt in "CSI":
        res.field = np.zeros((nsz, nrz, nrr), np.complex64)
        out.field = res.field.ctypes.data_as(_abi.c_float_p)
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
        n = prob.nrays
        res.nsteps = np.zeros(n, np.int32)
        res.ntop = np.zeros(n, np.int32)
        res.nbot = np.zeros(n, np.int32)
        res.xend = np.zeros((n, 2), np.float64)
        out.ray_nsteps = res.nsteps.ctypes.data_as(_abi.c_int32_p)
        out.ray_ntop = res.ntop.ctypes.data_as(_abi.c_int32_p)
        out.ray_nbot = res.nbot.ctypes.data_as(_abi.c_int32_p)
        out.ray_xend = res.xend.ctypes.data_as(_abi.c_double_p)
    return out


def field_run(env_or_prob, ctx=None, stats=False, max_narr=0, allow_merging=1, eigen_memsize=0):
    """One field run (TL / arrivals / eigen hits) on the GPU."""
    prob = env_or_prob if isinstance(env_or_prob, Problem) else Problem(env_or_prob)
    own = ctx is None
    ctx = ctx or Context(0)
    res = FieldResult()
    out = _make_out(prob, res, stats, max_narr, allow_merging, eigen_memsize)
    try:
        _check(lib().bhc_field_run2d(ctx._h, C.byref(prob.view), C.byref(out)))
        res.kernel_ms = ctx.last_kernel_ms()
    finally:
        if own:
            ctx.close()
    res.err, res.warn = out.err, out.warn
    res.err_count, res.warn_count = out.err_count, out.warn_count
    res.total_steps, res.total_contrib = out.total_steps, out.total_contrib
    res.problem = prob
    return res


class AsyncRun:
    """A non-blocking field run (bhc_field_run2d_async / bhc_field_run3d_async):
    returns at once; `on_done` (no arguments) is called on the context's
    worker thread when the outputs are written. wait() joins it and returns
    the FieldResult."""

    def __init__(self, prob, ctx, on_done=None, stats=False, max_narr=0, allow_merging=1, eigen_memsize=0):
        self.prob, self.ctx = prob, ctx
        self.res = FieldResult()
        mk = _make_out3d if isinstance(prob, Problem3D) else _make_out
        self.out = mk(prob, self.res, stats, max_narr, allow_merging, eigen_memsize)
        self.calls = 0

        def done(_user):
            self.calls += 1
            if on_done is not None:
                on_done()
        self._cb = COMPLETED_FN(done)      # kept alive until the run has ended
        if isinstance(prob, Problem3D):
            f = lib().bhc_field_runnx2d_async if prob.nx2d else lib().bhc_field_run3d_async
        else:
            f = lib().bhc_field_run2d_async
        _check(f(ctx._h, C.byref(prob.view), C.byref(self.out), self._cb, None))

    def wait(self):
        self.ctx.wait()
        r, out = self.res, self.out
        r.err, r.warn, r.err_count, r.warn_count = out.err, out.warn, out.err_count, out.warn_count
        r.total_steps, r.total_contrib = out.total_steps, out.total_contrib
        r.problem = self.prob
        return r


class Resident:
    """A problem whose inputs and outputs stay in HBM across launches."""

    def __init__(self, prob, ctx, stats=False, max_narr=0, allow_merging=1, eigen_memsize=0):
        self.prob, self.ctx = prob, ctx
        self.res = FieldResult()
        self.out = _make_out(prob, self.res, stats, max_narr, allow_merging, eigen_memsize)
        h = C.c_void_p()
        _check(lib().bhc_problem_upload2d(ctx._h, C.byref(prob.view), C.byref(self.out), C.byref(h)))
        self._h = h

    def launch(self):
        _check(lib().bhc_field_launch(self.ctx._h, self._h))

    def autotune(self):
        """Pick the fastest scatter mapping; returns 1 (tiles), 2 (wavefront) or 3 (lane)."""
        m = C.c_int()
        _check(lib().bhc_problem_autotune(self.ctx._h, self._h, C.byref(m)))
        return m.value

    def set_scatter(self, mode):
        _check(lib().bhc_problem_set_scatter(self._h, mode))

    def download(self):
        _check(lib().bhc_field_download(self.ctx._h, self._h, C.byref(self.out)))
        r = self.res
        r.err, r.warn = self.out.err, self.out.warn
        r.err_count, r.warn_count = self.out.err_count, self.out.warn_count
        r.total_steps, r.total_contrib = self.out.total_steps, self.out.total_contrib
        return r

    def release(self):
        if self._h:
            lib().bhc_problem_release(self._h)
            self._h = None


def ray_run(env_or_prob, ctx=None, max_points=20000, ialpha=None, isz=None, nsteps_in=None):
    """Ray run type (or eigenray re-trace with per-ray step limits)."""
    prob = env_or_prob if isinstance(env_or_prob, Problem) else Problem(env_or_prob)
    own = ctx is None
    ctx = ctx or Context(0)
    n = prob.nrays if ialpha is None else len(ialpha)
    pts = (_abi.RayPt2D * (n * max_points))()
    nst = np.zeros(n, np.int32)
    out = _abi.RayOut()
    out.points = C.cast(pts, C.POINTER(_abi.RayPt2D))
    out.nsteps = nst.ctypes.data_as(_abi.c_int32_p)
    out.max_points_per_ray = max_points
    try:
        if ialpha is None:
            _check(lib().bhc_ray_run2d(ctx._h, C.byref(prob.view), C.byref(out)))
        else:
            ia = np.ascontiguousarray(ialpha, np.int32)
            iz = np.ascontiguousarray(isz if isz is not None else np.zeros(n), np.int32)
            ns = None if nsteps_in is None else np.ascontiguousarray(nsteps_in, np.int32)
            _check(lib().bhc_ray_run2d_list(ctx._h, C.byref(prob.view), C.byref(out),
                                            ia.ctypes.data_as(_abi.c_int32_p), iz.ctypes.data_as(_abi.c_int32_p),
                                            ns.ctypes.data_as(_abi.c_int32_p) if ns is not None else None, n))
    finally:
        if own:
            ctx.close()
    arr = np.ctypeslib.as_array(pts).reshape(n, max_points)
    return arr, nst, (out.err, out.warn)


class Sweep:
    """Many TL problems traced in one persistent launch (bhc_sweep_*)."""

    def __init__(self, problems, ctx, stats=False):
        self.problems = list(problems)
        self.ctx = ctx
        n = len(self.problems)
        arr = (C.POINTER(_abi.Problem2D) * n)(*[C.pointer(p.view) for p in self.problems])
        h = C.c_void_p()
        _check(lib().bhc_sweep_upload2d(ctx._h, arr, n, 1 if stats else 0, C.byref(h)))
        self._h = h
        self.with_stats = stats

    def launch(self, p0=None, p1=None):
        """Trace every problem, or problems [p0, p1) only."""
        if p0 is None and p1 is None:
            _check(lib().bhc_sweep_launch(self.ctx._h, self._h))
        else:
            p0 = 0 if p0 is None else p0
            p1 = len(self.problems) if p1 is None else p1
            _check(lib().bhc_sweep_launch_range(self.ctx._h, self._h, p0, p1))

    def autotune(self):
        b, l, m = C.c_int(), C.c_int(), C.c_int()
        _check(lib().bhc_sweep_autotune(self.ctx._h, self._h, C.byref(b), C.byref(l), C.byref(m)))
        return {"scatter_blocks": b.value, "lds_bytes": l.value, "scatter_lanes": m.value}

    def set_graph(self, on):
        _check(lib().bhc_sweep_set_graph(self._h, 1 if on else 0))

    def set_scatter(self, mapping):
        """0: one wavefront per segment, 1: one lane per segment, 2: 64 segments'
        receivers flattened over a wavefront."""
        _check(lib().bhc_sweep_set_scatter(self._h, mapping))

    def set_launch(self, scatter_blocks, lds_bytes, mapping=None):
        """Launch configuration without a search (as autotune() reports it)."""
        _check(lib().bhc_sweep_set_launch(self._h, scatter_blocks, lds_bytes))
        if mapping is not None:
            self.set_scatter(mapping)

    def stats(self):
        t, s, p = C.c_double(), C.c_double(), C.c_int()
        st, co = C.c_longlong(), C.c_longlong()
        _check(lib().bhc_sweep_stats(self._h, C.byref(t), C.byref(s), C.byref(p), C.byref(st), C.byref(co)))
        return {"trace_ms": t.value, "scatter_ms": s.value, "passes": p.value, "steps": st.value,
                "contrib": co.value}

    def download(self, i):
        prob = self.problems[i]
        res = FieldResult()
        out = _make_out(prob, res, self.with_stats, 0, 1, 0)
        _check(lib().bhc_sweep_download(self.ctx._h, self._h, i, C.byref(out)))
        res.err, res.warn = out.err, out.warn
        res.total_steps = out.total_steps
        res.problem = prob
        return res

    def release(self):
        if self._h:
            lib().bhc_sweep_release(self._h)
            self._h = None

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.bhc_sweep_release(self._h)
            self._h = None


class Problem3D:
    """A preprocessed 3D problem (bhc_problem3d view owned by the C++ side)."""

    def __init__(self, env):
        self.env = env
        ec, self._keep = env.to_c()
        h = C.c_void_p()
        if lib().bhc_preprocess3d(C.byref(ec), C.byref(h)) != 0:
            raise BhcError((lib().bhc_host3d_last_error() or b"?").decode())
        self._h = h
        self.view = lib().bhc_problem3d_view(h).contents

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.bhc_problem3d_free(self._h)
            self._h = None

    @property
    def nrays(self):
        v = self.view
        na = 1 if v.iSingle_alpha >= 1 else v.Nalpha
        nb = 1 if v.iSingle_beta >= 1 else v.Nbeta
        return na * nb * v.NSx * v.NSy * v.NSz

    @property
    def field_shape(self):
        v = self.view
        return (v.NSz, v.NSx, v.NSy, v.Ntheta, v.NRz, v.NRr)

    @property
    def run_type(self):
        return bytes(self.view.RunType).decode("latin-1")

    @property
    def nx2d(self):
        """Nx2D problem (RunType[5] '2'): 2D rays on the bearing planes."""
        return self.run_type[5] == "2"

    def scale_pressure(self, field):
        f = np.ascontiguousarray(field, np.complex64).copy()
        if lib().bhc_scale_pressure3d(C.byref(self.view), f.ctypes.data_as(_abi.c_float_p)) != 0:
            raise BhcError((lib().bhc_host3d_last_error() or b"?").decode())
        return f


def _make_out3d(prob, res, stats, max_narr, allow_merging, eigen_memsize):
    nrays = prob.nrays
    res.field = np.zeros(prob.field_shape, np.complex64)
    out = _abi.FieldOut()
    out.field = res.field.ctypes.data_as(_abi.c_float_p)
    rt = prob.run_type[0]
    nrcv = int(np.prod(prob.field_shape))
    if rt in "AaVv":
        res.max_narr = max_narr
        res.arr = (_abi.Arrival * (nrcv * max(max_narr, 1)))()
        res.narr = np.zeros(nrcv, np.int32)
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
        res.nsteps = np.zeros(nrays, np.int32)
        res.ntop = np.zeros(nrays, np.int32)
        res.nbot = np.zeros(nrays, np.int32)
        res.xend = np.zeros((nrays, 3), np.float64)
        out.ray_nsteps = res.nsteps.ctypes.data_as(_abi.c_int32_p)
        out.ray_ntop = res.ntop.ctypes.data_as(_abi.c_int32_p)
        out.ray_nbot = res.nbot.ctypes.data_as(_abi.c_int32_p)
        out.ray_xend = res.xend.ctypes.data_as(_abi.c_double_p)
    return out


def field_run3d(env_or_prob, ctx=None, stats=False, max_narr=0, allow_merging=1, eigen_memsize=0):
    """3D field run on the GPU (RunFieldModesImpl<CFG, true, true>). Returns
    a FieldResult: .field (complex64, [NSz, NSx, NSy, Ntheta, NRz, NRr],
    pre-ScalePressure) for TL runs, .arr / .narr ([receivers][max_narr] in
    field order) for arrivals runs ('A', 'a'), .hits / .neigen for eigenray
    runs ('E'); .total_steps, .total_contrib, .err, .warn and, with
    stats=True, per-ray .nsteps / .ntop / .nbot / .xend [rays, 3]."""
    prob = env_or_prob if isinstance(env_or_prob, Problem3D) else Problem3D(env_or_prob)
    own = ctx is None
    if own:
        ctx = Context(0)
    try:
        res = FieldResult()
        out = _make_out3d(prob, res, stats, max_narr, allow_merging, eigen_memsize)
        f = lib().bhc_field_runnx2d if prob.nx2d else lib().bhc_field_run3d
        _check(f(ctx._h, C.byref(prob.view), C.byref(out)))
        res.total_steps, res.total_contrib = out.total_steps, out.total_contrib
        res.err, res.warn = out.err, out.warn
        return res
    finally:
        if own:
            ctx.close()


def ray_runnx2d(env_or_prob, ctx=None, max_points=4000):
    """Nx2D ray run type on the GPU (bhc_ray_runnx2d). Returns (problem,
    points [rays * max_points] of bhc_ray_pt2d in ray coordinates (range
    along the bearing, depth), nsteps [rays], origin [rays, 4] = source x, y
    and cos, sin of the bearing)."""
    prob = env_or_prob if isinstance(env_or_prob, Problem3D) else Problem3D(env_or_prob)
    if not prob.nx2d:
        raise BhcError("ray_runnx2d: not an Nx2D problem (run type character 6 must be '2')")
    own = ctx is None
    if own:
        ctx = Context(0)
    try:
        n = prob.nrays
        pts = (_abi.RayPt2D * (n * max_points))()
        ns = np.zeros(n, np.int32)
        org = np.zeros((n, 4), np.float64)
        out = _abi.RayOut()
        out.points = C.cast(pts, C.POINTER(_abi.RayPt2D))
        out.nsteps = ns.ctypes.data_as(_abi.c_int32_p)
        out.max_points_per_ray = max_points
        _check(lib().bhc_ray_runnx2d(ctx._h, C.byref(prob.view), C.byref(out), org.ctypes.data_as(_abi.c_double_p)))
        return prob, pts, ns, org
    finally:
        if own:
            ctx.close()


def ray_run3d(env_or_prob, ctx=None, max_points=4000):
    """3D ray run type on the GPU (RunRayMode). Returns (problem, points
    [rays * max_points] of bhc_ray_pt3d, nsteps [rays])."""
    prob = env_or_prob if isinstance(env_or_prob, Problem3D) else Problem3D(env_or_prob)
    own = ctx is None
    if own:
        ctx = Context(0)
    try:
        n = prob.nrays
        pts = (_abi.RayPt3D * (n * max_points))()
        ns = np.zeros(n, np.int32)
        out = _abi.RayOut3D()
        out.points = C.cast(pts, C.POINTER(_abi.RayPt3D))
        out.nsteps = ns.ctypes.data_as(_abi.c_int32_p)
        out.max_points_per_ray = max_points
        _check(lib().bhc_ray_run3d(ctx._h, C.byref(prob.view), C.byref(out)))
        return prob, pts, ns
    finally:
        if own:
            ctx.close()


def ray_run3d_list(env_or_prob, hits, nhits, ctx=None, max_points=4000):
    """3D eigenray re-trace on the GPU (bhc_ray_run3d_list): one ray per hit,
    stopped after the hit's step. Returns (problem, points, nsteps)."""
    prob = env_or_prob if isinstance(env_or_prob, Problem3D) else Problem3D(env_or_prob)
    own = ctx is None
    if own:
        ctx = Context(0)
    try:
        pts = (_abi.RayPt3D * (nhits * max_points))()
        ns = np.zeros(nhits, np.int32)
        out = _abi.RayOut3D()
        out.points = C.cast(pts, C.POINTER(_abi.RayPt3D))
        out.nsteps = ns.ctypes.data_as(_abi.c_int32_p)
        out.max_points_per_ray = max_points
        _check(lib().bhc_ray_run3d_list(ctx._h, C.byref(prob.view), C.cast(hits, C.POINTER(_abi.EigenHit)), nhits,
                                        C.byref(out)))
        return prob, pts, ns
    finally:
        if own:
            ctx.close()


def write_ray3d(path, prob, pts, nsteps, max_points, title="bellhop-mi355x"):
    """.ray file of a 3D ray run (bhc_write_ray3d)."""
    rc = lib().bhc_write_ray3d(path.encode(), C.byref(prob.view), title.encode(),
                               C.cast(pts, C.POINTER(_abi.RayPt3D)), np.ascontiguousarray(nsteps, np.int32)
                               .ctypes.data_as(_abi.c_int32_p), max_points, len(nsteps))
    if rc != 0:
        raise BhcError("writing %s failed" % path)
