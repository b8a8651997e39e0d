"""CPU checks of the surrogate's test restatement and host-side pieces."""
import numpy as np
import pytest

import mlp_ref
from bellhopcuda_amd import surrogate as sg
import bellhopcuda_amd as bh


def test_mlp_reference_gradients_by_finite_differences():
    """The FP64 restatement's backward pass (without BF16 rounding) against
    central differences of its own loss."""
    mlp_ref.ROUND = False
    try:
        rng = np.random.default_rng(0)
        off, n = mlp_ref.layout()
        p = rng.normal(scale=0.05, size=n)
        Bf = rng.normal(scale=2.0, size=(32, 16))
        x = rng.random((8, 16))
        y = rng.normal(size=8)
        _, loss, g = mlp_ref.forward_backward(p, Bf, x, y)
        for name in ("W0", "b2", "W5", "Wout", "bout"):
            o, shape = off[name]
            size = int(np.prod(shape))
            for k in rng.choice(size, size=min(3, size), replace=False):
                i = o + int(k)
                h = 1e-5
                pp, pm = p.copy(), p.copy()
                pp[i] += h
                pm[i] -= h
                fd = (mlp_ref.forward_backward(pp, Bf, x, y)[1] - mlp_ref.forward_backward(pm, Bf, x, y)[1]) / (2 * h)
                assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), (name, k)
    finally:
        mlp_ref.ROUND = True


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5e-3], np.float32)
    r = mlp_ref.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 4 * 2 ** -8
    assert r[3] == 1.0
    assert abs(r[4] - (-2.5e-3)) / 2.5e-3 < 2 ** -8


def test_environment_family_and_inputs():
    th = sg.draw_params(5, 3)
    assert th.shape == (5, 14)
    for t in th:
        e = sg.environment(t)
        p = bh.Problem(e)
        assert p.field_shape == (1, 51, 201)
        x = sg.inputs(t, e.Rz, e.Rr_km)
        assert x.shape == (51 * 201, 16)
        assert np.all(x[:, 0] > 0) and np.all(x[:, 0] <= 1.0) and np.all(x[:, 1] >= 0) and np.all(x[:, 1] <= 1)
        assert 50.0 <= e.freq0 <= 500.0 and 20.0 <= e.Sz[0] <= 500.0
        assert np.all(np.diff(e.ssp_z) > 0) and np.all(e.bty.z <= sg.DEPTH) and np.all(e.bty.z > 1000.0)


def test_screen_rule():
    pred = np.zeros((4, 3, 2, 2))
    pred[:, 1] = np.array([0.0, 0.5, 1.0, 3.0])[:, None, None]   # spread 3 dB everywhere
    pred[:, 2] = np.array([0.0, 0.2, 0.4, 0.9])[:, None, None]   # spread 0.9 dB
    mean, fb, rms = sg.screen(pred, 1.0)
    assert list(fb) == [False, True, False]
    assert rms[1] == pytest.approx(3.0) and mean[1, 0, 0] == pytest.approx(1.125)
