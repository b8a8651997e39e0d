"""CPU checks of the surrogate's test restatement and host-side pieces."""
import numpy as np
import pytest

import mlp_ref
from bellhopcuda_amd import surrogate as sg
import bellhopcuda_amd as bh


@pytest.mark.parametrize("n_out,front", [(1, 0), (32, 1)])
def test_mlp_reference_gradients_by_finite_differences(n_out, front):
    """The FP64 restatement's backward pass (without BF16 rounding) against
    central differences of its own loss, for the coordinate network and the
    multi-output network with the input-carrying front."""
    mlp_ref.ROUND = False
    try:
        rng = np.random.default_rng(0)
        off, n = mlp_ref.layout(n_out)
        p = rng.normal(scale=0.05, size=n)
        Bf = rng.normal(scale=2.0, size=(32, 16))
        x = rng.random((8, 16))
        y = rng.normal(size=8) if n_out == 1 else rng.normal(size=(8, n_out))
        _, loss, g = mlp_ref.forward_backward(p, Bf, x, y, n_out, front)
        for name in ("W0", "b2", "W5", "Wout", "bout"):
            o, shape = off[name]
            size = int(np.prod(shape))
            for k in rng.choice(size, size=min(3, size), replace=False):
                i = o + int(k)
                h = 1e-5
                pp, pm = p.copy(), p.copy()
                pp[i] += h
                pm[i] -= h
                fd = (mlp_ref.forward_backward(pp, Bf, x, y, n_out, front)[1]
                      - mlp_ref.forward_backward(pm, Bf, x, y, n_out, front)[1]) / (2 * h)
                assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), (name, k)
    finally:
        mlp_ref.ROUND = True


def test_feature_fronts():
    """Front 1 carries the 16 inputs unchanged, then the sin / cos pairs of the
    first 24 projections; front 0 interleaves the pairs of all 32."""
    mlp_ref.ROUND = False
    try:
        rng = np.random.default_rng(3)
        x = rng.random((5, 16))
        Bf = rng.normal(size=(32, 16))
        s = x.astype(np.float32) @ Bf.astype(np.float32).T
        F0 = mlp_ref.features(x, Bf, 0)
        F1 = mlp_ref.features(x, Bf, 1)
        assert np.allclose(F0[:, 0::2], np.sin(2 * np.pi * s), atol=1e-5)
        assert np.allclose(F0[:, 1::2], np.cos(2 * np.pi * s), atol=1e-5)
        assert np.allclose(F1[:, :16], x.astype(np.float32))
        assert np.allclose(F1[:, 16::2], np.sin(2 * np.pi * s[:, :24]), atol=1e-5)
        assert np.allclose(F1[:, 17::2], np.cos(2 * np.pi * s[:, :24]), atol=1e-5)
    finally:
        mlp_ref.ROUND = True


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5e-3], np.float32)
    r = mlp_ref.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 4 * 2 ** -8
    assert r[3] == 1.0
    assert abs(r[4] - (-2.5e-3)) / 2.5e-3 < 2 ** -8


@pytest.mark.parametrize("family", ["global", "local"])
def test_environment_family_and_inputs(family):
    th = sg.draw_params(200, 3, family)
    assert th.shape == (200, 14)
    x = sg.param_inputs(th, family)
    assert x.shape == (200, 16) and np.all(x[:, 14:] == 0.0)
    # uniform parameters land in [-1, 1], EOF amplitudes in standard deviations
    u = np.r_[0:2, 10:14]
    assert np.all(np.abs(x[:, u]) <= 1.0) and np.all(np.abs(x[:, 2:10]).max(0) > 1.0)
    assert 0.7 < np.std(x[:, 2:10]) < 1.3
    if family == "local":
        assert np.all(np.abs(th[:, u] - sg.LOCAL_CENTRE[u]) <= 0.05 + 1e-12)
    for t in th[:5]:
        e = sg.environment(t)
        p = bh.Problem(e)
        assert p.field_shape == (1, 51, 201)
        assert 50.0 <= e.freq0 <= 500.0 and 20.0 <= e.Sz[0] <= 500.0
        assert np.all(np.diff(e.ssp_z) > 0) and np.all(e.bty.z <= sg.DEPTH) and np.all(e.bty.z > 1000.0)
        m = sg.water_mask(t, e.Rz, e.Rr_km)
        d = sg.seabed_depth(t, e.Rr_km)
        assert m.shape == (51, 201) and np.array_equal(m, np.asarray(e.Rz)[:, None] <= d[None, :])


def test_screen_rule():
    """The spread is the members' sample standard deviation, its RMS taken
    over the water-column receivers only; the fallback flag is that RMS above
    the threshold."""
    pred = np.zeros((4, 3, 2, 2))
    pred[:, 1] = np.array([0.0, 1.0, 2.0, 3.0])[:, None, None]   # std sqrt(5/3) everywhere
    pred[:, 2] = np.array([0.0, 0.2, 0.4, 0.6])[:, None, None]   # std sqrt(0.2/3)
    pred[3, 2, 1, 1] = 50.0                                       # outside the water column
    masks = np.ones((3, 2, 2), bool)
    masks[2, 1, 1] = False
    mean, fb, rms = sg.screen(pred, masks, 1.0)
    assert list(fb) == [False, True, False]
    assert rms[1] == pytest.approx(np.sqrt(5.0 / 3.0)) and mean[1, 0, 0] == pytest.approx(1.5)
    assert rms[2] == pytest.approx(np.sqrt(0.2 / 3.0)) and rms[0] == 0.0


def test_modal_basis_reconstructs_low_rank_fields():
    """Fields in a rank-r affine family are reproduced by r modes (to float
    rounding); coefficients are those of the training fields; capping at the
    shadow level is applied before the fit."""
    rng = np.random.default_rng(7)
    nrz, nrr, r = 11, 23, 5
    base = 80.0 + 5.0 * rng.random((nrz, nrr))
    comps = rng.normal(size=(r, nrz, nrr))
    a = rng.normal(size=(60, r))
    fields = [base + np.tensordot(ai, comps, 1) for ai in a]
    B = sg.ModalBasis(fields, r, seed=1)
    assert np.allclose(B.modes @ B.modes.T, np.eye(r), atol=1e-5)
    rec = B.fields(B.train_coeffs)
    assert np.max(np.abs(rec - np.stack(fields))) < 1e-3
    assert np.allclose(B.coeffs(fields[:3]), B.train_coeffs[:3], atol=1e-4)
    capped = [np.full((nrz, nrr), 130.0)] * 3
    B2 = sg.ModalBasis(capped, 1, seed=1)
    assert np.allclose(B2.ref, sg.TL_CAP)


def test_field_errors_mask():
    """Errors count water-column receivers whose tracer TL is below the cap;
    answers above the cap are compared as the cap."""
    th = sg.draw_params(1, 5)[0]
    th[10] = 0.9  # a tall seamount: some receivers lie below the seabed
    rz = np.linspace(0.0, sg.DEPTH, 51)
    rr = np.linspace(0.25, 50.0, 201)
    truth = np.full((1, 51, 201), 90.0)
    truth[0, :, :10] = 130.0                       # shadowed: not counted
    ans = truth + 2.0
    ans[0, ~sg.water_mask(th, rz, rr)] = 0.0       # below the seabed: not counted
    rms, p95, per = sg.field_errors(ans, truth, th[None], rz, rr)
    assert rms == pytest.approx(2.0) and p95 == pytest.approx(2.0) and per[0] == pytest.approx(2.0)
