"""Learned TL surrogate with tracer fallback.

A field of an environment family is written as the mean training field plus
K orthonormal modes (the leading principal components of the training
fields, receiver by receiver) with coefficients that an ensemble of MLPs
predicts from the environment's parameters (libbhc_hip.so, bhc_mlp_*: 16
normalised parameters -> the inputs and 24 Fourier pairs -> 6 x 256 GELU -> K
coefficients, BF16 matrix-core forward/backward, Adam). TL is learned capped
at 120 dB, the level above which receivers count as shadowed. On new
environments the ensemble's spread decides, environment by environment,
whether its field is the answer or the batched tracer runs instead; errors
against the tracer are counted at water-column receivers whose tracer TL is
below 120 dB.

A coordinate network (TL of one receiver from its position and the
parameters, one output) uses the same kernels with n_out = 1.

Environment family (14 parameters):
  source depth 20-500 m; frequency 50-500 Hz (log-uniform); 8 sound-speed
  EOFs -- sin(k pi z / D), k = 1..8, amplitude N(0, 1) x 5 m/s -- on the Munk
  profile over D = 4000 m; and 4 seamount parameters: height 0-2000 m, centre
  10-40 km, near-flank width 2-10 km and far-flank width 0.5-2 times it (an
  asymmetric Gaussian); bottom an acousto-elastic halfspace of 1600 m/s.
  Receivers: 51 depths x 201 ranges to 50 km; 1001 Gaussian beams over
  +-30 degrees, incoherent TL. Family "global" draws every parameter over its
  full range; family "local" is an uncertainty sweep around one nominal
  environment (each [0, 1] parameter within +-0.05 of its nominal value, EOF
  amplitudes N(0, 0.2^2)).
"""
import ctypes as C
import math
import time

import numpy as np

from . import _abi, lib, BhcError, Problem, Sweep, Resident
from .env import Env2D, Halfspace, Boundary, subtab, munk_c

D_IN = 16
TL_MIN, TL_MAX, TL_CAP = 40.0, 140.0, 120.0
DEPTH = 4000.0
R_MAX_KM = 50.0
N_PARAMS = 14
# family "local": nominal parameters and half-widths (uniform parameters) or
# standard deviations (EOF amplitudes, N(0, 1) units)
LOCAL_CENTRE = np.array([0.3, 0.5] + [0.0] * 8 + [0.5] * 4)
LOCAL_SPREAD = np.array([0.05, 0.05] + [0.2] * 8 + [0.05] * 4)


def _mlp_lib():
    L = lib()
    if not getattr(L, "_mlp_sigs", False):
        P = C.POINTER
        fp = P(C.c_float)
        for name, res, args in [
            ("bhc_mlp_create", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_float, P(C.c_void_p)]),
            ("bhc_mlp_create_ex", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_float,
                                            P(C.c_void_p)]),
            ("bhc_mlp_destroy", None, [C.c_void_p]),
            ("bhc_mlp_last_error", C.c_char_p, []),
            ("bhc_mlp_num_params", C.c_int, []),
            ("bhc_mlp_param_count", C.c_longlong, [C.c_void_p]),
            ("bhc_mlp_get_params", C.c_int, [C.c_void_p, fp, fp]),
            ("bhc_mlp_set_params", C.c_int, [C.c_void_p, fp]),
            ("bhc_mlp_set_fourier", C.c_int, [C.c_void_p, fp]),
            ("bhc_mlp_predict", C.c_int, [C.c_void_p, fp, C.c_longlong, fp]),
            ("bhc_mlp_gradients", C.c_int, [C.c_void_p, fp, fp, C.c_int, fp, fp]),
            ("bhc_mlp_set_dataset", C.c_int, [C.c_void_p, fp, fp, C.c_longlong]),
            ("bhc_mlp_train", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_ulonglong, fp]),
            ("bhc_mlp_last_ms", C.c_double, [C.c_void_p]),
            ("bhc_mlp_step_flops", C.c_double, [C.c_void_p, C.c_int]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._mlp_sigs = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """Ensemble of `members` networks (bhc_mlp_*) with n_out outputs and
    feature front `front` (0: Fourier features only; 1: the inputs and 24
    Fourier pairs)."""

    def __init__(self, members=4, max_batch=16384, seed=0, sigma=1.0, device=0, n_out=1, front=0):
        L = _mlp_lib()
        h = C.c_void_p()
        if L.bhc_mlp_create_ex(device, members, max_batch, n_out, front, seed, sigma, C.byref(h)) != 0:
            raise BhcError(L.bhc_mlp_last_error().decode())
        self._h, self.members, self.max_batch = h, members, max_batch
        self.n_out, self.front = n_out, front
        self.nparam = int(L.bhc_mlp_param_count(h))

    def _check(self, rc):
        if rc != 0:
            raise BhcError(_mlp_lib().bhc_mlp_last_error().decode())

    def params(self):
        P = np.zeros((self.members, self.nparam), np.float32)
        Bf = np.zeros((32, 16), np.float32)
        self._check(_mlp_lib().bhc_mlp_get_params(self._h, _fp(P), _fp(Bf)))
        return P, Bf

    def set_fourier(self, B):
        B = np.ascontiguousarray(B, np.float32)
        assert B.shape == (32, 16)
        self._check(_mlp_lib().bhc_mlp_set_fourier(self._h, _fp(B)))

    def set_params(self, P):
        P = np.ascontiguousarray(P, np.float32)
        assert P.shape == (self.members, self.nparam)
        self._check(_mlp_lib().bhc_mlp_set_params(self._h, _fp(P)))

    def predict(self, x):
        """[members][n] for one output, [members][n][n_out] otherwise."""
        x = np.ascontiguousarray(x, np.float32)
        assert x.ndim == 2 and x.shape[1] == D_IN
        out = np.zeros((self.members, x.shape[0], self.n_out), np.float32)
        self._check(_mlp_lib().bhc_mlp_predict(self._h, _fp(x), x.shape[0], _fp(out)))
        return out[..., 0] if self.n_out == 1 else out

    def _targets(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        assert x.ndim == 2 and x.shape[1] == D_IN and y.size == x.shape[0] * self.n_out
        return x, y

    def gradients(self, x, y):
        x, y = self._targets(x, y)
        g = np.zeros((self.members, self.nparam), np.float32)
        loss = np.zeros(self.members, np.float32)
        self._check(_mlp_lib().bhc_mlp_gradients(self._h, _fp(x), _fp(y), x.shape[0], _fp(g), _fp(loss)))
        return g, loss

    def set_dataset(self, X, Y):
        self._X, self._Y = self._targets(X, Y)
        self._check(_mlp_lib().bhc_mlp_set_dataset(self._h, _fp(self._X), _fp(self._Y), self._X.shape[0]))

    def train(self, steps, batch, lr=1e-3, seed=0):
        loss = np.zeros(self.members, np.float32)
        self._check(_mlp_lib().bhc_mlp_train(self._h, steps, batch, lr, seed, _fp(loss)))
        return loss

    def last_ms(self):
        return _mlp_lib().bhc_mlp_last_ms(self._h)

    def step_flops(self, batch):
        return _mlp_lib().bhc_mlp_step_flops(self._h, batch)

    def close(self):
        if self._h:
            _mlp_lib().bhc_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------
# environment family
# --------------------------------------------------------------------------

def draw_params(n, seed, family="global"):
    """n parameter vectors [n, 14] ordered (zs, logf, eof1..8, h, rc, w, asym):
    zs, logf and the seamount's four in [0, 1], the EOF amplitudes in N(0, 1)
    units (family "global"), or within the local sweep around the nominal
    environment (family "local")."""
    rng = np.random.default_rng(seed)
    th = np.zeros((n, N_PARAMS))
    if family == "global":
        th[:, 0:2] = rng.random((n, 2))
        th[:, 2:10] = rng.normal(size=(n, 8))
        th[:, 10:14] = rng.random((n, 4))
    elif family == "local":
        th[:] = LOCAL_CENTRE
        th[:, 0:2] += LOCAL_SPREAD[0:2] * (2.0 * rng.random((n, 2)) - 1.0)
        th[:, 2:10] += LOCAL_SPREAD[2:10] * rng.normal(size=(n, 8))
        th[:, 10:14] += LOCAL_SPREAD[10:14] * (2.0 * rng.random((n, 4)) - 1.0)
    else:
        raise ValueError("family must be 'global' or 'local'")
    return th


def environment(theta, precision=64, nbeams=1001, nrz=51, nrr=201):
    zs = 20.0 + 480.0 * theta[0]
    freq = 50.0 * 10.0 ** theta[1]
    z = np.linspace(0.0, DEPTH, 41)
    c = munk_c(z * (5000.0 / DEPTH))
    for k in range(8):
        c = c + 5.0 * theta[2 + k] * np.sin((k + 1) * np.pi * z / DEPTH)
    rb = np.linspace(0.0, R_MAX_KM + 1.0, 201)
    db = seabed_depth(theta, rb)
    cp = 1600.0
    return Env2D(
        title="surrogate family", precision=precision, freq0=float(freq), ssp_type="C", atten_unit="W ",
        ssp_z=z, ssp_c=c, top=Halfspace("V", depth=0.0),
        bot=Halfspace("A", depth=DEPTH, alphaR=cp, betaR=0.0, rho=1.8, alphaI=0.5, zTemp=DEPTH),
        bty=Boundary("LS", r_km=rb, z=db), Sz=np.array([zs], np.float32),
        Rz=subtab(0.0, DEPTH, nrz, np.float32), Rr_km=subtab(R_MAX_KM / (nrr - 1), R_MAX_KM, nrr),
        alpha_deg=subtab(-30.0, 30.0, nbeams), run_type="IB", deltas=0.0,
        box_r_km=R_MAX_KM + 1.0, box_z=DEPTH + 100.0)


def seabed_depth(theta, r_km):
    """Bathymetry (m) of environment `theta` at ranges r_km: an asymmetric
    Gaussian seamount on the flat 4000 m bottom."""
    h = 2000.0 * theta[10]
    rc = 10.0 + 30.0 * theta[11]
    w1 = 2.0 + 8.0 * theta[12]
    w2 = w1 * (0.5 + 1.5 * theta[13])
    r = np.asarray(r_km, np.float64)
    w = np.where(r < rc, w1, w2)
    return DEPTH - h * np.exp(-0.5 * ((r - rc) / w) ** 2)


def param_inputs(thetas, family="global"):
    """Network inputs [n, 16]: the 14 parameters scaled to O(1) for the family
    (uniform parameters to [-1, 1], EOF amplitudes in standard deviations),
    then two zeros."""
    th = np.atleast_2d(np.asarray(thetas, np.float64))
    if family == "local":
        t = (th - LOCAL_CENTRE) / LOCAL_SPREAD
    elif family == "global":
        t = th.copy()
        t[:, 0:2] = 2.0 * t[:, 0:2] - 1.0
        t[:, 10:14] = 2.0 * t[:, 10:14] - 1.0
    else:
        raise ValueError("family must be 'global' or 'local'")
    x = np.zeros((th.shape[0], D_IN), np.float32)
    x[:, :N_PARAMS] = t
    return x


def water_mask(theta, rz, rr_km):
    """Receivers [nrz, nrr] in the water column (at or above the seabed)."""
    return np.asarray(rz, np.float64)[:, None] <= seabed_depth(theta, rr_km)[None, :]


def fourier_matrix(seed, sigma=0.05):
    """Projection B [32][16] of the Fourier features: N(0, sigma^2) entries on
    the 14 parameter inputs (small: the parameters act smoothly), zero on the
    two padding inputs."""
    rng = np.random.default_rng(seed)
    B = np.zeros((32, 16), np.float32)
    B[:, :N_PARAMS] = rng.normal(scale=sigma, size=(32, N_PARAMS))
    return B


def tl_db(problem, field):
    p = problem.scale_pressure(field)[0]
    with np.errstate(divide="ignore"):
        tl = -20.0 * np.log10(np.maximum(np.abs(p).astype(np.float64), 1e-30))
    return np.clip(tl, TL_MIN, TL_MAX)


def trace_fields(thetas, ctx, chunk=256):
    """TL (dB, clipped) of every environment, through batched sweeps of
    `chunk` problems. Returns (list of [nrz, nrr] arrays, envs, seconds of
    GPU sweep time)."""
    out, envs, t_sweep = [], [], 0.0
    for c0 in range(0, len(thetas), chunk):
        es = [environment(th) for th in thetas[c0:c0 + chunk]]
        ps = [Problem(e) for e in es]
        sw = Sweep(ps, ctx)
        t0 = time.perf_counter()
        sw.launch()
        t_sweep += time.perf_counter() - t0
        for i, p in enumerate(ps):
            out.append(tl_db(p, sw.download(i).field))
        sw.release()
        envs += es
    return out, envs, t_sweep


# --------------------------------------------------------------------------
# modal basis and the surrogate
# --------------------------------------------------------------------------

class ModalBasis:
    """Mean field and K orthonormal modes of a set of training fields (TL in
    dB capped at TL_CAP), from a randomised range finder (two power
    iterations) and the SVD of the projected fields."""

    def __init__(self, fields, k, seed=0, oversample=32, power=2):
        T = np.minimum(np.stack(fields), TL_CAP)
        self.shape = T.shape[1:]
        X = T.reshape(T.shape[0], -1).astype(np.float32)
        self.ref = X.mean(0)
        X -= self.ref
        rng = np.random.default_rng(seed)
        l = min(k + oversample, X.shape[0], X.shape[1])
        Q = X @ rng.standard_normal((X.shape[1], l)).astype(np.float32)
        for _ in range(power):
            Q, _ = np.linalg.qr(Q)
            Q = X @ (X.T @ Q)
        Q, _ = np.linalg.qr(Q)
        _, sv, Vt = np.linalg.svd((Q.T @ X).astype(np.float64), full_matrices=False)
        self.k = k
        self.modes = np.ascontiguousarray(Vt[:k], np.float32)        # [k][n_rcv]
        self.singular = sv[:k]
        c = X @ self.modes.T
        self.scale = float(np.std(c[:, 0])) or 1.0                   # coefficient unit of the network
        self.train_coeffs = (c / self.scale).astype(np.float32)

    def coeffs(self, fields):
        X = np.minimum(np.stack(fields), TL_CAP).reshape(len(fields), -1).astype(np.float32) - self.ref
        return (X @ self.modes.T) / self.scale

    def fields(self, coeffs):
        """Fields [..., nrz, nrr] (dB, clipped to [TL_MIN, TL_CAP]) of network coefficients [..., k]."""
        c = np.asarray(coeffs, np.float32)
        f = (c.reshape(-1, self.k) * self.scale) @ self.modes + self.ref
        return np.clip(f, TL_MIN, TL_CAP).reshape(c.shape[:-1] + self.shape)


class Surrogate:
    """The modal surrogate: a ModalBasis of the training fields and an
    ensemble of `members` networks mapping parameters to its coefficients."""

    def __init__(self, thetas, fields, family="global", k=128, members=4, seed=11, sigma=0.05, max_batch=2048):
        self.family = family
        self.basis = ModalBasis(fields, k, seed=seed)
        self.mlp = Mlp(members=members, max_batch=max_batch, seed=seed, sigma=sigma, n_out=k, front=1)
        self.mlp.set_fourier(fourier_matrix(seed, sigma))
        self.X = param_inputs(thetas, family)
        self.mlp.set_dataset(self.X, self.basis.train_coeffs)
        self.steps, self.train_ms = 0, 0.0

    def train(self, steps, batch=2048, lr=1e-3, chunk=500, floor=0.02):
        """Adam with a cosine learning-rate decay from lr to floor * lr, one
        value per chunk of steps. Returns the last per-member loss."""
        done, loss = 0, None
        while done < steps:
            n = min(chunk, steps - done)
            f = floor + (1.0 - floor) * 0.5 * (1.0 + math.cos(math.pi * (done + 0.5 * n) / steps))
            loss = self.mlp.train(n, batch, lr * f, seed=self.steps + done)
            self.train_ms += self.mlp.last_ms()
            done += n
        self.steps += steps
        return loss

    def predict(self, thetas):
        """Every member's fields [members, n, nrz, nrr] (dB, clipped)."""
        c = self.mlp.predict(param_inputs(thetas, self.family))
        return self.basis.fields(c)

    def close(self):
        self.mlp.close()


def screen(pred, masks, spread_db=1.0):
    """Per environment: ensemble mean and whether it falls back to the tracer
    -- the RMS over water-column receivers (masks [n, nrz, nrr]) of the
    ensemble spread (the members' sample standard deviation about their mean)
    above `spread_db`."""
    mean = pred.mean(0)
    spread = pred.std(0, ddof=1)
    m = np.asarray(masks, bool)
    rms_spread = np.sqrt(np.sum(np.where(m, spread, 0.0) ** 2, axis=(1, 2)) / np.maximum(m.sum(axis=(1, 2)), 1))
    return mean, rms_spread > spread_db, rms_spread


def screened_answer(sur, thetas, envs, ctx, spread_db=1.0):
    """The screened sweep's answer for environments `thetas` / `envs`: per
    environment, the ensemble mean where the members agree (RMS of the
    ensemble spread at or below `spread_db`) and the batched tracer's field
    elsewhere. Returns
    (fields [n, nrz, nrr] in dB, fallback flags, spread RMS per environment,
    seconds predicting and screening, seconds tracing the fallbacks)."""
    t0 = time.perf_counter()
    pred = sur.predict(thetas)
    masks = np.stack([water_mask(th, e.Rz, e.Rr_km) for th, e in zip(thetas, envs)])
    final, fallback, rms_spread = screen(pred, masks, spread_db)
    final = final.astype(np.float64)
    t_pred = time.perf_counter() - t0
    fb = [i for i in range(len(thetas)) if fallback[i]]
    t_fb = 0.0
    if fb:
        t0 = time.perf_counter()
        f_fb, _, _ = trace_fields(np.asarray(thetas)[fb], ctx)
        t_fb = time.perf_counter() - t0
        for k, i in enumerate(fb):
            final[i] = f_fb[k]
    return final, fallback, rms_spread, t_pred, t_fb


def field_errors(answer, truth, thetas, rz, rr_km):
    """Errors (dB) of fields [n, nrz, nrr] against the tracer's, at
    water-column receivers whose tracer TL is below TL_CAP: (RMS, p95 of |e|,
    per-environment RMS)."""
    errs, per = [], []
    for a, t, th in zip(answer, truth, thetas):
        m = water_mask(th, rz, rr_km) & (t < TL_CAP)
        e = (np.minimum(a, TL_CAP) - t)[m]
        errs.append(e)
        per.append(float(np.sqrt(np.mean(e ** 2))) if e.size else 0.0)
    e = np.concatenate(errs) if errs else np.zeros(0)
    if e.size == 0:
        return 0.0, 0.0, np.asarray(per)
    return float(np.sqrt(np.mean(e ** 2))), float(np.percentile(np.abs(e), 95)), np.asarray(per)
