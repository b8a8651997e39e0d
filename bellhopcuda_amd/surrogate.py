"""Learned TL surrogate with tracer fallback.

An ensemble of coordinate MLPs (libbhc_hip.so, bhc_mlp_*: 16 inputs -> 64
Fourier features -> 6 x 256 GELU -> TL, BF16 matrix-core forward/backward,
Adam) predicts the TL (dB) of a receiver from its position and the
environment's parameters. It is trained on fields the batched tracer computes
for a family of environments; on new environments the ensemble's spread
decides where its prediction is trusted and where the tracer runs instead.

Environment family (14 parameters, each mapped from [0, 1] or N(0, 1)):
  source depth 20-500 m; frequency 50-500 Hz (log-uniform); 8 sound-speed
  EOFs -- sin(k pi z / D), k = 1..8, amplitude N(0, 1) x 5 m/s -- on the Munk
  profile over D = 4000 m; and 4 seamount parameters: height 0-2000 m, centre
  10-40 km, near-flank width 2-10 km and far-flank width 0.5-2 times it (an
  asymmetric Gaussian); bottom an acousto-elastic halfspace of 1600 m/s.
  Receivers: 51 depths x 201 ranges to 50 km; 1001 Gaussian beams over
  +-30 degrees, incoherent TL (the smooth field a coordinate network can
  represent; coherent interference is left to the tracer).
"""
import ctypes as C
import math
import time

import numpy as np

from . import _abi, lib, BhcError, Problem, Sweep, Resident
from .env import Env2D, Halfspace, Boundary, subtab, munk_c

D_IN = 16
TL_MIN, TL_MAX, TL_SCALE = 40.0, 140.0, 25.0
DEPTH = 4000.0
R_MAX_KM = 50.0


def _mlp_lib():
    L = lib()
    if not getattr(L, "_mlp_sigs", False):
        P = C.POINTER
        fp = P(C.c_float)
        for name, res, args in [
            ("bhc_mlp_create", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_float, P(C.c_void_p)]),
            ("bhc_mlp_destroy", None, [C.c_void_p]),
            ("bhc_mlp_last_error", C.c_char_p, []),
            ("bhc_mlp_num_params", C.c_int, []),
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
    """Ensemble of `members` networks (bhc_mlp_*)."""

    def __init__(self, members=4, max_batch=16384, seed=0, sigma=1.0, device=0):
        L = _mlp_lib()
        h = C.c_void_p()
        if L.bhc_mlp_create(device, members, max_batch, seed, sigma, C.byref(h)) != 0:
            raise BhcError(L.bhc_mlp_last_error().decode())
        self._h, self.members, self.max_batch = h, members, max_batch
        self.nparam = L.bhc_mlp_num_params()

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
        self._check(_mlp_lib().bhc_mlp_set_params(self._h, _fp(P)))

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros((self.members, x.shape[0]), np.float32)
        self._check(_mlp_lib().bhc_mlp_predict(self._h, _fp(x), x.shape[0], _fp(out)))
        return out

    def gradients(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros((self.members, self.nparam), np.float32)
        loss = np.zeros(self.members, np.float32)
        self._check(_mlp_lib().bhc_mlp_gradients(self._h, _fp(x), _fp(y), x.shape[0], _fp(g), _fp(loss)))
        return g, loss

    def set_dataset(self, X, Y):
        self._X = np.ascontiguousarray(X, np.float32)
        self._Y = np.ascontiguousarray(Y, np.float32)
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

def draw_params(n, seed):
    """n parameter vectors: [zs, logf, h, rc, w, asym] in [0, 1] and 8 EOF
    amplitudes N(0, 1) -> array [n, 14] ordered (zs, logf, eof1..8, h, rc, w, asym)."""
    rng = np.random.default_rng(seed)
    th = np.zeros((n, 14))
    th[:, 0:2] = rng.random((n, 2))
    th[:, 2:10] = rng.normal(size=(n, 8))
    th[:, 10:14] = rng.random((n, 4))
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


# Frequency scale of each input in the Fourier projection: range and depth
# carry the field's structure, the environment parameters modulate it
# smoothly (low frequencies there keep the network from keying on individual
# training environments).
INPUT_SCALES = np.array([2.0, 2.0, 0.5, 0.5] + [0.25] * 8 + [0.5] * 4)


def fourier_matrix(seed, kind="random"):
    """Projection B [32][16] of the features sin / cos(2 pi B x). "random":
    Gaussian rows scaled per input (INPUT_SCALES); "structured": 10 range
    and 8 depth rows at geometrically spaced frequencies (0.5-32 and 0.5-16
    cycles over the normalised span), the other 14 rows random mixtures."""
    rng = np.random.default_rng(seed)
    if kind == "random":
        return (rng.normal(size=(32, 16)) * INPUT_SCALES[None, :]).astype(np.float32)
    if kind != "structured":
        raise ValueError(kind)
    B = np.zeros((32, 16), np.float32)
    B[np.arange(10), 0] = np.geomspace(0.5, 32.0, 10)
    B[10 + np.arange(8), 1] = np.geomspace(0.5, 16.0, 8)
    B[18:] = rng.normal(size=(14, 16)) * np.array([1.0, 1.0] + [0.5] * 14)[None, :]
    return B


def inputs(theta, rz, rr_km):
    """Network inputs [n_rcv, 16] of one environment's receiver grid."""
    Z, R = np.meshgrid(np.asarray(rz, np.float64), np.asarray(rr_km, np.float64), indexing="ij")
    n = Z.size
    x = np.zeros((n, D_IN), np.float32)
    x[:, 0] = (R.reshape(-1) / R_MAX_KM)
    x[:, 1] = (Z.reshape(-1) / DEPTH)
    x[:, 2] = theta[1]
    x[:, 3] = theta[0]
    x[:, 4:12] = np.asarray(theta[2:10]) / 3.0
    x[:, 12:16] = np.asarray(theta[10:14])
    return x


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


def reference_field(fields):
    """Mean TL over the training fields, receiver by receiver: the network
    learns each environment's deviation from it."""
    return np.mean(np.stack(fields), axis=0)


def dataset(thetas, fields, envs, points_per_field, seed, ref):
    rng = np.random.default_rng(seed)
    X, Y = [], []
    for th, tl, e in zip(thetas, fields, envs):
        x = inputs(th, e.Rz, e.Rr_km)
        y = ((tl - ref).reshape(-1) / TL_SCALE).astype(np.float32)
        k = rng.choice(x.shape[0], size=min(points_per_field, x.shape[0]), replace=False)
        X.append(x[k])
        Y.append(y[k])
    return np.concatenate(X), np.concatenate(Y)


def predict_fields(mlp, thetas, envs, ref):
    """Ensemble predictions (dB) [members, n_env, nrz, nrr]."""
    xs = [inputs(th, e.Rz, e.Rr_km) for th, e in zip(thetas, envs)]
    shape = (len(envs[0].Rz), len(envs[0].Rr_km))
    pred = mlp.predict(np.concatenate(xs)).reshape(mlp.members, len(thetas), *shape)
    pred = pred * TL_SCALE + ref[None, None]
    return np.clip(pred, TL_MIN, TL_MAX)


def screen_sparse(mlp, thetas, envs, ref, spread_db=1.0, stride=(5, 8)):
    """The screen's decision from a receiver subsample (every stride[0]-th
    depth and stride[1]-th range): per environment, the RMS of the members'
    spread there, and whether it exceeds `spread_db` (fall back to the
    tracer). Costs 1 / (stride[0] * stride[1]) of a full prediction."""
    rz, rr = np.asarray(envs[0].Rz)[::stride[0]], np.asarray(envs[0].Rr_km)[::stride[1]]
    x = np.concatenate([inputs(th, rz, rr) for th in thetas])
    pred = mlp.predict(x).reshape(mlp.members, len(thetas), len(rz), len(rr))
    pred = np.clip(pred * TL_SCALE + ref[::stride[0], ::stride[1]][None, None], TL_MIN, TL_MAX)
    spread = pred.max(0) - pred.min(0)
    rms_spread = np.sqrt(np.mean(spread ** 2, axis=(1, 2)))
    return rms_spread > spread_db, rms_spread


def screened_answer(mlp, thetas, envs, ref, ctx, spread_db=1.0, stride=(5, 8)):
    """The screened sweep's answer for environments `thetas` / `envs`: the
    ensemble's spread on a receiver subsample decides, per environment,
    between the network's full field (ensemble mean) and the batched tracer.
    Returns (fields [n, nrz, nrr] in dB, fallback flags, spread RMS per
    environment, seconds predicting, seconds tracing the fallbacks)."""
    t0 = time.perf_counter()
    fallback, rms_spread = screen_sparse(mlp, thetas, envs, ref, spread_db, stride)
    shape = (len(envs[0].Rz), len(envs[0].Rr_km))
    final = np.zeros((len(thetas),) + shape)
    keep = [i for i in range(len(thetas)) if not fallback[i]]
    if keep:
        final[keep] = predict_fields(mlp, thetas[keep], [envs[i] for i in keep], ref).mean(0)
    t_pred = time.perf_counter() - t0
    fb = [i for i in range(len(thetas)) if fallback[i]]
    t_fb = 0.0
    if fb:
        t0 = time.perf_counter()
        f_fb, _, _ = trace_fields(thetas[fb], ctx)
        t_fb = time.perf_counter() - t0
        for k, i in enumerate(fb):
            final[i] = f_fb[k]
    return final, fallback, rms_spread, t_pred, t_fb


def screen(pred, spread_db=1.0):
    """Per environment: ensemble mean and whether it falls back to the tracer
    (RMS over receivers of the members' spread above `spread_db`)."""
    mean = pred.mean(0)
    spread = pred.max(0) - pred.min(0)
    rms_spread = np.sqrt(np.mean(spread ** 2, axis=(1, 2)))
    return mean, rms_spread > spread_db, rms_spread
