"""Misfit, adjoint source, preconditioning and model-update steps (host side)."""

import numpy as np
import pytest
import scipy.ndimage

from stride_amd.optim import L2DistanceLoss, apply_prec, GradientDescent, LBFGS


def test_l2_loss_and_adjoint_source():
    rng = np.random.default_rng(0)
    mod = rng.standard_normal((5, 40)).astype(np.float32)
    obs = rng.standard_normal((5, 40)).astype(np.float32)
    loss = L2DistanceLoss()
    f, res = loss.forward(mod, obs)
    assert f == pytest.approx(0.5 * np.sum((mod.astype(np.float64) - obs) ** 2), rel=1e-6)
    g_mod, g_obs = loss.adjoint(2.0, True, True)
    np.testing.assert_allclose(g_mod, 2.0 * (mod - obs), rtol=1e-6)
    np.testing.assert_allclose(g_obs, -2.0 * (mod - obs), rtol=1e-6)
    # the adjoint source is the derivative of f: directional check
    m64, o64 = mod.astype(np.float64), obs.astype(np.float64)
    d = rng.standard_normal(mod.shape)
    f1, _ = L2DistanceLoss().forward(m64 - 1e-4 * d, o64)
    f2, _ = L2DistanceLoss().forward(m64 + 1e-4 * d, o64)
    assert (f2 - f1) / 2e-4 == pytest.approx(np.sum(g_mod / 2.0 * d), rel=1e-4)


def test_l2_loss_torch():
    torch = pytest.importorskip('torch')
    mod = torch.randn(4, 30)
    obs = torch.randn(4, 30)
    f, res = L2DistanceLoss().forward(mod, obs)
    assert f == pytest.approx(0.5 * float(((mod - obs) ** 2).sum()), rel=1e-5)


def test_apply_prec_matches_reference_formula():
    rng = np.random.default_rng(1)
    g = rng.standard_normal((20, 24)).astype(np.float32)
    p = rng.random((20, 24)) * 5
    out = apply_prec(g, p, prec_scale=4.0, prec_smooth=1.0)
    ps = scipy.ndimage.gaussian_filter(p, 1.0)
    ps = ps * (4.0 * ps.size / np.sum(np.abs(ps))) + 1
    np.testing.assert_allclose(out, g / ps, rtol=1e-5)
    np.testing.assert_array_equal(apply_prec(g, None), g)


def test_gradient_descent_in_slowness():
    vp = np.full((4, 4), 1500., np.float32)
    direction = np.full((4, 4), 1e-6)
    gd = GradientDescent(step_size=2.0, transform=lambda v: 1 / v, inverse=lambda s: 1 / s, bounds=(1400, 1700))
    out = gd.step(vp, direction)
    np.testing.assert_allclose(out, 1 / (1 / 1500. - 2e-6), rtol=1e-6)


def test_lbfgs_converges_on_quadratic():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((30, 30))
    A = A @ A.T + 30 * np.eye(30)
    b = rng.standard_normal(30)
    xs = np.linalg.solve(A, b)

    def fg(x):
        return 0.5 * x @ A @ x - b @ x, A @ x - b

    opt = LBFGS(memory=6, first_step=0.1)
    x, f = opt.minimize(np.zeros(30), fg, iterations=40)
    assert np.linalg.norm(x - xs) < 1e-5 * np.linalg.norm(xs)
    assert opt.calls <= 60
    fs = [h[1] for h in opt.history]
    assert all(b <= a + 1e-12 for a, b in zip(fs, fs[1:]))
