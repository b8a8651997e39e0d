"""GPU: the surrogate MLP's matrix-core forward/backward against the CPU FP64
restatement (tests/mlp_ref.py), and training that learns."""
import ctypes as C
import sys

import numpy as np
import pytest

import bellhopcuda_amd as bh
import mlp_ref

pytestmark = pytest.mark.gpu


def _announce(name):
    print("\n[gpu-case]", name, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [96, 4096])
def test_mlp_gradients_match_fp64_restatement(n):
    """n = 4096 takes the split-K path of the weight-gradient products."""
    _announce("mlp_gradients_%d" % n)
    from bellhopcuda_amd import surrogate
    m = surrogate.Mlp(members=2, max_batch=max(256, n), seed=7)
    rng = np.random.default_rng(1)
    x = rng.random((n, 16), dtype=np.float32)
    y = rng.normal(size=n).astype(np.float32)
    P, Bf = m.params()
    # the output layer starts at zero: give it values so that every layer's
    # gradient is exercised
    off, _ = mlp_ref.layout()
    o = off["Wout"][0]
    P[:, o:o + 257] = rng.normal(scale=0.1, size=(2, 257)).astype(np.float32)
    m.set_params(P)
    g, loss = m.gradients(x, y)
    for e in range(2):
        out_r, loss_r, g_r = mlp_ref.forward_backward(P[e], Bf, x, y.astype(np.float64))
        assert loss[e] == pytest.approx(loss_r, rel=2e-4)
        scale = np.max(np.abs(g_r))
        err = np.max(np.abs(g[e] - g_r)) / scale
        print("member", e, "loss", loss[e], loss_r, "max rel grad err", err, flush=True)
        assert err < 2e-3
        # a hidden activation on a BF16 rounding boundary can round the other
        # way from FP32 pre-activations: most outputs agree to 1e-3, all to 1e-2
        pred = m.predict(x)
        tol = np.max(np.abs(out_r))
        close = np.isclose(pred[e], out_r, rtol=1e-3, atol=1e-3 * tol)
        assert np.mean(close) >= 0.9 and np.max(np.abs(pred[e] - out_r)) <= 1e-2 * tol


def test_mlp_training_reduces_loss():
    _announce("mlp_training")
    from bellhopcuda_amd import surrogate
    rng = np.random.default_rng(2)
    X = rng.random((20000, 16), dtype=np.float32)
    Y = (np.sin(3 * X[:, 0]) + X[:, 1] ** 2 - 0.5 * X[:, 2]).astype(np.float32)
    m = surrogate.Mlp(members=4, max_batch=1024, seed=3)
    m.set_dataset(X, Y)
    l0 = m.train(steps=5, batch=1024, lr=1e-3, seed=1)
    l1 = m.train(steps=400, batch=1024, lr=1e-3, seed=2)
    print("loss", l0, l1, flush=True)
    assert np.all(l1 < 0.1 * l0)
    pred = m.predict(X[:2000])
    rms = np.sqrt(np.mean((pred.mean(0) - Y[:2000]) ** 2))
    assert rms < 0.25 * np.std(Y)


def test_sparse_screen_matches_full_screen_at_unit_stride():
    """The subsampled screen (ensemble spread on every stride-th receiver) is
    the full screen when the stride is 1, and with a stride its spread
    estimate stays close to the full one."""
    _announce("surrogate_sparse_screen")
    from bellhopcuda_amd import surrogate as sg
    m = sg.Mlp(members=4, max_batch=16384, seed=5)
    m.set_fourier(sg.fourier_matrix(3))
    P, _ = m.params()
    rng = np.random.default_rng(4)
    m.set_params((P + rng.normal(scale=0.02, size=P.shape)).astype(np.float32))  # members differ
    th = sg.draw_params(6, 99)
    envs = [sg.environment(t, nbeams=11, nrz=11, nrr=41) for t in th]
    ref = np.full((11, 41), 80.0)
    pred = sg.predict_fields(m, th, envs, ref)
    _, fb_full, rms_full = sg.screen(pred, 1.0)
    fb_1, rms_1 = sg.screen_sparse(m, th, envs, ref, 1.0, stride=(1, 1))
    assert np.array_equal(fb_full, fb_1) and np.allclose(rms_full, rms_1, rtol=1e-5)
    _, rms_s = sg.screen_sparse(m, th, envs, ref, 1.0, stride=(2, 4))
    assert np.allclose(rms_s, rms_full, rtol=0.5)
