"""Host logic of the learned-preconditioner inversion (stride_amd.learned) without a GPU: the
solver is replaced by a quadratic misfit in slowness with its exact gradient, the network by
fixed proposals. A proposal is kept only when the misfit drops; otherwise the plain gradient
step (in slowness, backtracked) is taken; passes are counted (a gradient two, a check one)."""

import numpy as np
import pytest

from stride_amd import learned


class _Space:
    shape = (4, 4)


class _FakeFWI:
    """J(vp) = 0.5 sum (1/vp - 1/true)^2 * k; gradient in slowness."""

    def __init__(self, true, k=1e12):
        self.true = true
        self.k = k
        self.space = _Space()
        self.passes = 0
        self.dev = 'cpu'

    def observe(self, true, freq):
        pass

    def _f(self, vp):
        r = 1 / vp.astype(np.float64) - 1 / self.true
        return 0.5 * self.k * float(np.sum(r ** 2)), self.k * r

    def misfit(self, vp, freq):
        self.passes += 1
        return self._f(vp)[0]

    def gradient(self, vp, freq, below=None):
        import torch
        f, g = self._f(vp)
        if below is not None and not f < below:
            self.passes += 1
            return f, None
        self.passes += 2
        return f, torch.tensor(g, dtype=torch.float32)


class _FakeNet:
    def __init__(self, proposals):
        self.proposals = list(proposals)

    def forward(self, g, scale=1.0, model=None, model_scale=1.0):
        import torch
        return torch.tensor(self.proposals.pop(0)[None] / learned.TARGET_SCALE, dtype=torch.float32)


def test_gated_updates():
    true = np.full((4, 4), 1550., np.float32)
    good = np.full((4, 4), 40., np.float32)      # toward the truth: kept
    bad = np.full((4, 4), -60., np.float32)      # away from it: rejected, plain step taken
    fwi = _FakeFWI(true)
    net = _FakeNet([good, bad, np.full((4, 4), 3., np.float32)])
    vp, hist, accepted = learned.invert_learned(fwi, net, {500e3: 1.0}, true, [500e3], iters_per_freq=3,
                                                first_step=5.0, proposal_scales=(1.0,))
    assert accepted == 2
    fs = [h[1] for h in hist]
    assert all(b < a for a, b in zip(fs, fs[1:]))
    # iteration 1: gradient + check; iteration 2: gradient + rejected check + one plain step
    # (it lowers the misfit at once); iteration 3: gradient + check
    assert [h[0] for h in hist] == [3, 7, 10]
    # the rejected proposal left the model where the plain step put it, closer to the truth
    assert np.all(np.abs(vp - true) < np.abs(1500 + 40 - true) + 1e-3)


def test_plain_step_is_in_slowness():
    vp = np.full((2, 2), 1500., np.float32)
    g = np.full((2, 2), 1e-6)
    out = learned._gradient_step(vp, g, 1.0)
    np.testing.assert_allclose(out, 1 / (1 / 1500. - 1e-6), rtol=1e-6)
    # bounds
    assert learned._gradient_step(vp, np.full((2, 2), -1.0), 1.0).max() == pytest.approx(learned.VP_BOUNDS[0])


def test_phantom_and_partial_model():
    p = learned.phantom(3)
    assert p.shape == (512, 512) and p.dtype == np.float32
    assert p.min() >= 1400 - 1e-3 and p.max() <= 1650 + 1e-3
    assert p[0, 0] == 1500. and p[256, 256] != 1500.
    rng = np.random.default_rng(0)
    m = learned.partial_model(p, rng)
    assert m.shape == p.shape and abs(float(m[0, 0]) - 1500.) < 1.0
    # between water and the phantom on average
    d = (p - 1500.).astype(np.float64)
    frac = float(np.sum((m - 1500.) * d) / np.sum(d * d))
    assert -0.01 < frac < 0.81


def test_input_scales():
    import torch
    pairs = [(torch.full((2, 2), v), None, f, None) for v, f in ((2.0, 1.0), (4.0, 1.0), (8.0, 2.0))]
    s = learned.input_scales(pairs)
    assert s[1.0] == pytest.approx(1 / 3.0) and s[2.0] == pytest.approx(1 / 8.0)


def test_lbfgs_gated_and_first_direction():
    """L-BFGS with gated trials (a rejected trial returns no gradient) and a proposed first step:
    a good proposal is kept and lands at once, a bad one costs one trial and the scaled gradient
    step follows; both converge on a quadratic like the plain run."""
    from stride_amd.optim import LBFGS
    A = np.diag(np.linspace(1.0, 5.0, 6))
    xs = np.linspace(-1, 1, 6)
    calls = {'full': 0, 'gated': 0}

    def fg(x, threshold=None):
        f = 0.5 * float((x - xs) @ A @ (x - xs))
        if threshold is not None and not f <= threshold:
            calls['gated'] += 1
            return f, None
        calls['full'] += 1
        return f, A @ (x - xs)

    x0 = np.zeros(6)
    plain = LBFGS(memory=5, first_step=0.1)
    xp, fp = plain.minimize(x0, fg, 12, gated=True)
    good = LBFGS(memory=5, first_step=0.1)
    xg, fgood = good.minimize(x0, fg, 1, first_direction=lambda x, g: x - xs, gated=True)
    assert good.proposals_kept == 1 and fgood < 1e-20
    calls['gated'] = 0
    bad = LBFGS(memory=5, first_step=0.1)
    xb, fb = bad.minimize(x0, fg, 12, first_direction=lambda x, g: -10 * (x - xs) + 0 * g, gated=True)
    assert bad.proposals_kept == 0 and fb < 1e-6 and fp < 1e-6
    # the bad proposal is not a descent direction: it is never evaluated
    bad2 = LBFGS(memory=5, first_step=0.1)
    calls['gated'] = 0
    bad2.minimize(x0, fg, 1, first_direction=lambda x, g: 3 * (x - xs) + 20 * g, gated=True)
    assert bad2.proposals_kept == 0 and calls['gated'] >= 1


def test_learned_lbfgs_first_step():
    """invert_learned_lbfgs: the network's update replaces the first step of a band when it lowers
    the misfit enough; the band otherwise runs as invert_plain, pass for pass."""
    true = np.full((4, 4), 1550., np.float32)
    good = np.full((4, 4), 49., np.float32)
    bad = np.full((4, 4), -60., np.float32)
    fp = _FakeFWI(true)
    _, hp = learned.invert_plain(fp, true, [500e3], iters_per_freq=3, first_step=5.0)
    fgood = _FakeFWI(true)
    _, hg, kept = learned.invert_learned_lbfgs(fgood, _FakeNet([good]), {500e3: 1.0}, true, [500e3],
                                               iters_per_freq=3, first_step=5.0)
    assert kept == 1 and hg[0][1] < hp[0][1] and hg[0][0] == hp[0][0] == 4
    # an overshoot (toward the truth, too far): tried, rejected at one forward pass, then the
    # plain run's steps
    over = np.full((4, 4), 200., np.float32)
    fo = _FakeFWI(true)
    _, ho, kept = learned.invert_learned_lbfgs(fo, _FakeNet([over]), {500e3: 1.0}, true, [500e3],
                                               iters_per_freq=3, first_step=5.0)
    assert kept == 0
    assert [h[0] for h in ho] == [h[0] + 1 for h in hp]
    assert np.allclose([h[1] for h in ho], [h[1] for h in hp])
    # away from the truth: not a descent direction, never evaluated
    fbad = _FakeFWI(true)
    _, hb, kept = learned.invert_learned_lbfgs(fbad, _FakeNet([bad]), {500e3: 1.0}, true, [500e3],
                                               iters_per_freq=3, first_step=5.0)
    assert kept == 0 and [h[0] for h in hb] == [h[0] for h in hp]
