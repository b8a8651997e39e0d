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


@pytest.mark.parametrize("n,n_out,front", [(96, 1, 0), (4096, 1, 0), (96, 64, 1), (4096, 64, 1)])
def test_mlp_gradients_match_fp64_restatement(n, n_out, front):
    """n = 4096 takes the split-K path of the weight-gradient products; n_out =
    64 the matrix-core output layer (forward, dWout, dH5) with the
    input-carrying feature front."""
    _announce("mlp_gradients_%d_%d" % (n, n_out))
    from bellhopcuda_amd import surrogate
    m = surrogate.Mlp(members=2, max_batch=max(256, n), seed=7, n_out=n_out, front=front)
    try:
        rng = np.random.default_rng(1)
        x = rng.random((n, 16), dtype=np.float32)
        y = rng.normal(size=(n, n_out) if n_out > 1 else n).astype(np.float32)
        P, Bf = m.params()
        # the output layer starts at zero: give it values so that every layer's
        # gradient is exercised
        off, _ = mlp_ref.layout(n_out)
        o = off["Wout"][0]
        nw = 256 * n_out + n_out
        P[:, o:o + nw] = rng.normal(scale=0.1, size=(2, nw)).astype(np.float32)
        m.set_params(P)
        g, loss = m.gradients(x, y)
        pred = m.predict(x)
        for e in range(2):
            out_r, loss_r, g_r = mlp_ref.forward_backward(P[e], Bf, x, y.astype(np.float64), n_out, front)
            assert loss[e] == pytest.approx(loss_r, rel=2e-4)
            scale = np.max(np.abs(g_r))
            err = np.max(np.abs(g[e] - g_r)) / scale
            print("member", e, "loss", loss[e], loss_r, "max rel grad err", err, flush=True)
            assert err < 2e-3
            # a hidden activation on a BF16 rounding boundary can round the other
            # way from FP32 pre-activations: most outputs agree to 1e-3, all to 1e-2
            tol = np.max(np.abs(out_r))
            close = np.isclose(pred[e], out_r, rtol=1e-3, atol=1e-3 * tol)
            assert np.mean(close) >= 0.9 and np.max(np.abs(pred[e] - out_r)) <= 1e-2 * tol
    finally:
        m.close()


@pytest.mark.parametrize("n_out", [1, 32])
def test_mlp_training_reduces_loss(n_out):
    _announce("mlp_training_%d" % n_out)
    from bellhopcuda_amd import surrogate
    rng = np.random.default_rng(2)
    X = rng.random((20000, 16), dtype=np.float32)
    f = (np.sin(3 * X[:, 0]) + X[:, 1] ** 2 - 0.5 * X[:, 2]).astype(np.float32)
    Y = f if n_out == 1 else (f[:, None] * np.linspace(-1.0, 1.0, n_out)[None, :] + X[:, 3:4]).astype(np.float32)
    m = surrogate.Mlp(members=4, max_batch=1024, seed=3, n_out=n_out, front=0 if n_out == 1 else 1)
    try:
        m.set_dataset(X, Y)
        l0 = m.train(steps=5, batch=1024, lr=1e-3, seed=1)
        l1 = m.train(steps=400, batch=1024, lr=1e-3, seed=2)
        print("loss", l0, l1, flush=True)
        assert np.all(l1 < 0.1 * l0)
        pred = m.predict(X[:2000])
        rms = np.sqrt(np.mean((pred.mean(0) - Y[:2000]) ** 2))
        assert rms < 0.25 * np.std(Y)
    finally:
        m.close()


def test_screened_sweep_bookkeeping_and_errors(gpu_ctx):
    """The surrogate-screened sweep end to end on a small sample of the local
    family: 512 tracer fields give the modal basis and train the ensemble, 32
    held-out environments are answered. Fallback environments get the batched
    tracer's field (to the float rounding of its field atomics) and the others
    the ensemble mean; the screen's flags follow the spread; RMS / p95 against
    tracer truth are those of the mixed answer; the network's held-out error is
    below the mean training field's (the method without the network)."""
    _announce("surrogate_screened_sweep")
    from bellhopcuda_amd import surrogate as sg
    th_tr = sg.draw_params(512, 20260005, "local")
    th_ev = sg.draw_params(32, 20260006, "local")
    f_tr, e_tr, _ = sg.trace_fields(th_tr, gpu_ctx)
    sur = sg.Surrogate(th_tr, f_tr, "local", k=64, members=4, seed=11, max_batch=256)
    try:
        loss0 = sur.mlp.train(1, 256, 1e-3, seed=0)
        loss = sur.train(3000, 256, 1e-3)
        assert np.all(np.asarray(loss) < np.asarray(loss0))
        f_ev, e_ev, _ = sg.trace_fields(th_ev, gpu_ctx)
        truth = np.stack(f_ev)
        rz, rr = e_ev[0].Rz, e_ev[0].Rr_km
        net = sur.predict(th_ev).mean(0)
        for thr in (1.0, 1e9, 0.0):
            final, fb, rms_spread, t_pred, t_fb = sg.screened_answer(sur, th_ev, e_ev, gpu_ctx, thr)
            assert np.array_equal(fb, rms_spread > thr)
            for i in range(len(th_ev)):
                if fb[i]:   # the tracer answers: equal up to the order of its float atomics
                    assert np.max(np.abs(final[i] - truth[i])) < 1e-3, i
                else:
                    assert np.allclose(final[i], net[i], atol=1e-4), i
            rms, p95, per = sg.field_errors(final, truth, th_ev, rz, rr)
            if fb.all():
                assert rms < 1e-3 and p95 < 1e-3
            print("threshold %g: fallback %d / %d, RMS %.3f dB, p95 %.3f dB" % (thr, fb.sum(), len(fb), rms, p95))
        assert fb.all()     # threshold 0: everything falls back
        net_rms, net_p95, net_per = sg.field_errors(net, truth, th_ev, rz, rr)
        base = np.broadcast_to(np.clip(sur.basis.ref.reshape(truth.shape[1:]), sg.TL_MIN, sg.TL_CAP), truth.shape)
        base_rms, base_p95, _ = sg.field_errors(base, truth, th_ev, rz, rr)
        k = list(range(0, 512, 16))
        pin = sur.predict(th_tr[k]).mean(0)
        in_rms, _, _ = sg.field_errors(pin, np.stack([f_tr[i] for i in k]), th_tr[k], rz, rr)
        print("network held-out RMS %.2f dB (p95 %.2f) vs mean field %.2f dB (p95 %.2f); in sample %.2f dB; "
              "spread RMS median %.2f dB" % (net_rms, net_p95, base_rms, base_p95, in_rms, np.median(rms_spread)),
              flush=True)
        assert in_rms < net_rms < base_rms
    finally:
        sur.close()
