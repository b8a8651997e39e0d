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


def test_screened_sweep_bookkeeping_and_errors(gpu_ctx):
    """The surrogate-screened sweep end to end on a small family sample: 256
    tracer fields train the ensemble, 24 held-out environments are answered.
    Fallback environments get the batched tracer's field exactly and the
    others the ensemble mean; the screen's flags follow the spread; the
    reported RMS / p95 against tracer truth are those of the mixed answer;
    training learns (in-sample error below the mean-field baseline). The
    held-out network error and its mean-field baseline are printed: at this
    data size the network does not beat the baseline on held-out environments,
    which is why the screen exists."""
    _announce("surrogate_screened_sweep")
    from bellhopcuda_amd import surrogate as sg
    th_tr = sg.draw_params(256, 20260005)
    th_ev = sg.draw_params(24, 20260006)
    f_tr, e_tr, _ = sg.trace_fields(th_tr, gpu_ctx)
    ref = sg.reference_field(f_tr)
    X, Y = sg.dataset(th_tr, f_tr, e_tr, 2048, 1, ref)
    mlp = sg.Mlp(members=4, max_batch=8192, seed=11)
    try:
        mlp.set_fourier(sg.fourier_matrix(12, "structured"))
        mlp.set_dataset(X, Y)
        loss0 = mlp.train(1, 8192, 1e-3, seed=0)
        loss = mlp.train(1500, 8192, 1e-3, seed=1)
        assert np.all(np.asarray(loss) < np.asarray(loss0))
        f_ev, e_ev, _ = sg.trace_fields(th_ev, gpu_ctx)
        truth = np.stack(f_ev)
        for thr in (1.0, 1e9, 0.0):
            final, fb, rms_spread, t_pred, t_fb = sg.screened_answer(mlp, th_ev, e_ev, ref, gpu_ctx, thr)
            assert np.array_equal(fb, rms_spread > thr)
            net = sg.predict_fields(mlp, th_ev, e_ev, ref).mean(0)
            for i in range(len(th_ev)):
                if fb[i]:
                    assert np.array_equal(final[i], truth[i]), i   # the tracer answers exactly
                else:
                    assert np.allclose(final[i], net[i], atol=1e-4), i
            err = np.abs(final - truth)
            rms, p95 = float(np.sqrt(np.mean(err ** 2))), float(np.percentile(err, 95))
            if fb.all():
                assert rms == 0.0 and p95 == 0.0
            print("threshold %g: fallback %d / %d, RMS %.3f dB, p95 %.3f dB" % (thr, fb.sum(), len(fb), rms, p95))
        assert fb.all()     # threshold 0: everything falls back
        net = sg.predict_fields(mlp, th_ev, e_ev, ref).mean(0)
        net_rms = float(np.sqrt(np.mean((net - truth) ** 2)))
        base_rms = float(np.sqrt(np.mean((ref[None] - truth) ** 2)))
        k = list(range(0, 256, 8))
        pin = sg.predict_fields(mlp, th_tr[k], [e_tr[i] for i in k], ref).mean(0)
        tin = np.stack([f_tr[i] for i in k])
        in_rms = float(np.sqrt(np.mean((pin - tin) ** 2)))
        in_base = float(np.sqrt(np.mean((ref[None] - tin) ** 2)))
        print("network held-out RMS %.2f dB vs mean-field %.2f dB; in-sample %.2f vs %.2f" % (net_rms, base_rms, in_rms,
                                                                                           in_base), flush=True)
        assert in_rms < in_base
        assert np.isfinite(net_rms) and np.isfinite(base_rms)
    finally:
        mlp.close()
