"""Pins of the pipeline-step oracle (oracle/pipelines_oracle.py) by closed forms, invariants
and brute force. The reference ships no tests or fixtures for these steps, so each check
below is chosen to fail on a plausible mistake (a wrong centre or offset, a missing flip, a
dropped normalisation, a wrong edge mode)."""

import numpy as np
import pytest
import scipy.signal

import pipelines_oracle as po


def test_cos_table_closed_form():
    np.testing.assert_allclose(po.cos_table(3), [0.25, 0.5, 0.25], atol=1e-15)
    for n in (5, 9, 21):
        t = po.cos_table(n)
        assert t.sum() == pytest.approx(1.0, abs=1e-14)
        np.testing.assert_allclose(t, t[::-1], atol=1e-15)
        assert np.argmax(t) == n // 2


def test_cos_filter_impulse_responses():
    nt, t0 = 40, 17
    x = np.zeros((1, nt))
    x[0, t0] = 1.0
    f_max = 0.1  # period round(1 / 0.2) = 5, length 11
    table = po.cos_table(11)
    causal = po.lowpass_filter_cos(x, f_max, zero_phase=False)
    np.testing.assert_allclose(causal[0, t0:t0 + 11], table, atol=1e-15)
    assert np.abs(causal[0, :t0]).max() == 0 and np.abs(causal[0, t0 + 11:]).max() == 0
    centred = po.lowpass_filter_cos(x, f_max, zero_phase=True)
    np.testing.assert_allclose(centred[0, t0 - 5:t0 + 6], table, atol=1e-15)


@pytest.mark.parametrize('kind', ['cos', 'sos'])
def test_filter_adjoint_dot_product(kind):
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 64))
    y = rng.standard_normal((3, 64))
    if kind == 'cos':
        fx = po.lowpass_filter_cos(x, 0.07, zero_phase=False)
        fty = po.lowpass_filter_cos(y, 0.07, zero_phase=False, adjoint=True)
    else:
        sos = po.butterworth_sos(0.05, 0.2)
        fx = po.sosfilt(sos, x)
        fty = po.sosfilt(sos, y, adjoint=True)
    assert np.sum(fx * y) == pytest.approx(np.sum(x * fty), rel=1e-12)


def test_butterworth_response():
    # Butterworth: |H| = 1 in the pass band centre, 1/sqrt(2) at each cut-off
    lo = po.butterworth_sos(f_max=0.1)
    assert lo.shape == (4, 6)
    w, h = scipy.signal.sosfreqz(lo, worN=[0.0, 2 * np.pi * 0.1])
    np.testing.assert_allclose(np.abs(h), [1.0, 2 ** -0.5], rtol=1e-9)
    band = po.butterworth_sos(0.05, 0.2)
    assert band.shape == (8, 6)
    _, h = scipy.signal.sosfreqz(band, worN=2 * np.pi * np.array([0.05, 0.2]))
    np.testing.assert_allclose(np.abs(h), 2 ** -0.5, rtol=1e-6)
    hp = po.butterworth_sos(f_min=0.15)
    _, h = scipy.signal.sosfreqz(hp, worN=2 * np.pi * np.array([0.15, 0.5]))
    np.testing.assert_allclose(np.abs(h), [2 ** -0.5, 1.0], rtol=1e-9)


def test_sosfilt_brute_force_difference_equation():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 50))
    sos = po.butterworth_sos(0.05, 0.2)
    y = x.copy()
    for b0, b1, b2, a0, a1, a2 in sos:  # direct form I, zero initial state, one section at a time
        out = np.zeros_like(y)
        for n in range(y.shape[1]):
            xm1 = y[:, n - 1] if n >= 1 else 0.0
            xm2 = y[:, n - 2] if n >= 2 else 0.0
            ym1 = out[:, n - 1] if n >= 1 else 0.0
            ym2 = out[:, n - 2] if n >= 2 else 0.0
            out[:, n] = (b0 * y[:, n] + b1 * xm1 + b2 * xm2 - a1 * ym1 - a2 * ym2) / a0
        y = out
    np.testing.assert_allclose(po.sosfilt(sos, x), y, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(po.sosfilt(sos, x, adjoint=True), po.sosfilt(sos, x[:, ::-1])[:, ::-1], atol=1e-12)


def test_norms():
    rng = np.random.default_rng(2)
    d = rng.standard_normal((6, 30)) * 7
    out, norm = po.norm_per_shot(d)
    assert np.sum(out ** 2) / 6 == pytest.approx(1.0, rel=1e-12)
    assert norm == pytest.approx(np.sqrt(np.sum(d ** 2) / 6))
    np.testing.assert_allclose(np.sum(po.norm_per_trace(d) ** 2, axis=-1), 1.0, rtol=1e-12)
    z, _ = po.norm_per_shot(np.zeros((2, 5)))
    assert np.all(z == 0)


def test_l2_closed_form():
    f, r = po.l2(np.array([[3.0, 4.0]]), np.zeros((1, 2)))
    assert f == 12.5
    np.testing.assert_array_equal(r, [[3.0, 4.0]])


def test_check_traces():
    d = np.ones((4, 5), np.float32)
    d[1, 2] = np.nan
    d[3, 0] = np.inf
    out, bad = po.check_traces(d, filter_incorrect=True)
    np.testing.assert_array_equal(bad, [False, True, False, True])
    assert np.all(out[[1, 3]] == 0) and np.all(out[[0, 2]] == 1)


def test_mute_closed_form():
    obs = np.ones((2, 10))
    obs[1, 5:] = 0
    mod = np.full((2, 10), 2.0)
    out = po.mute_traces(mod, obs)  # default f_max 0.5: table [1/4, 1/2, 1/4], zero outside
    np.testing.assert_allclose(out[0], [1.5] + [2.0] * 8 + [1.5])
    np.testing.assert_allclose(out[1], [1.5, 2, 2, 2, 1.5, 0.5, 0, 0, 0, 0])


def test_rampoff_mask_values():
    m = po.rampoff_mask((25,), 4)
    # pos = (ramp - index - 1) / (ramp - 1); value 1 - cos(pi/2 (1 - pos))
    expect = [0.0, 1 - np.cos(np.pi / 6), 0.5, 1.0]
    np.testing.assert_allclose(m[:4], expect, atol=1e-7)
    np.testing.assert_allclose(m[::-1][:4], expect, atol=1e-7)
    assert np.all(m[4:21] == 1)
    assert np.all(po.rampoff_mask((7,), 4) == 1)  # 2 ramp > n: no ramp on that axis
    m2 = po.rampoff_mask((12, 12), 4)
    assert m2[0, 5] == 0 and m2[5, 0] == 0 and m2[5, 5] == 1
    # in a corner the profile of the nearer edge holds (not the product of the two profiles)
    assert m2[1, 2] == pytest.approx(1 - np.cos(np.pi / 6)) and m2[2, 1] == pytest.approx(1 - np.cos(np.pi / 6))
    assert m2[2, 2] == pytest.approx(0.5) and m2[2, 6] == pytest.approx(0.5)


def test_mask_field_zero_outside_inner():
    f = np.ones((30, 30), np.float32)
    out = po.mask_field(f, (slice(5, 25), slice(4, 26)), ramp=3)
    assert np.all(out[:5] == 0) and np.all(out[25:] == 0) and np.all(out[:, :4] == 0)
    assert np.all(out[5:25, 4:26] == 1)


def test_smooth_closed_forms():
    c = np.full((9, 11), 3.25)
    np.testing.assert_allclose(po.smooth_field(c, 1.5), c, rtol=1e-13)
    np.testing.assert_allclose(po.smooth_field(c, 1.5, mode='reflect'), c, rtol=1e-13)
    d = np.zeros((7, 7))
    d[3, 3] = 1
    e = np.exp(-8.0)  # sigma 0.25: radius 1, weights [e^-8, 1, e^-8] / (1 + 2 e^-8)
    w = np.array([e, 1.0, e]) / (1 + 2 * e)
    np.testing.assert_allclose(po.smooth_field(d, 0.25)[2:5, 2:5], np.outer(w, w), rtol=1e-12)
    # nearest vs reflect differ at an edge: a ramp, sigma 1
    r = np.arange(8.0)[None, :].repeat(3, 0)
    assert po.smooth_field(r, (0, 1.0))[0, 0] != po.smooth_field(r, (0, 1.0), mode='reflect')[0, 0]


def test_norm_field_closed_form():
    rng = np.random.default_rng(3)
    f = rng.standard_normal((5, 6))
    assert np.max(np.abs(po.norm_field(f))) == pytest.approx(1.0)
    var = np.full((5, 6), 1500.0)
    var[0, 0] = 1700.0
    out = po.norm_field(f, var, transform=lambda v: 1 / v)
    assert np.max(np.abs(out)) == pytest.approx(1 / 1600.0 * 0.5 / 100)


def test_apply_prec_closed_form():
    rng = np.random.default_rng(4)
    g = rng.standard_normal((8, 9))
    out = po.apply_prec(g, np.full((8, 9), 7.0), prec_scale=4.0)
    np.testing.assert_allclose(out, g / 5.0, rtol=1e-12)
    np.testing.assert_allclose(po.apply_prec(g, np.zeros((8, 9))), g)


def test_descent_closed_form():
    x = np.array([1500.0, 1600.0, 1400.0])
    d = np.array([1e-5, -2e-5, 0.0])
    v = po.descent(x, d, 0.5)
    np.testing.assert_allclose(v, 1 / (1 / x - 0.5 * d), rtol=1e-14)
    assert v[0] > 1500 and v[1] < 1600 and v[2] == 1400
    np.testing.assert_allclose(po.descent(x, d, 0.5, slowness=False), x - 0.5 * d)
    np.testing.assert_allclose(po.descent(x, d * 10, 1.0, lo=1450, hi=1550), [1550, 1450, 1450])


def test_fir_design_and_filter():
    lo = po.fir_taps(f_max=0.1)
    assert len(lo) % 2 == 1 and lo.sum() == pytest.approx(1.0, abs=1e-12)  # unit DC gain
    np.testing.assert_allclose(lo, lo[::-1], atol=1e-15)  # linear phase
    hp = po.fir_taps(f_min=0.2)
    assert abs(np.sum(hp * (-1.0) ** np.arange(len(hp)))) == pytest.approx(1.0, abs=1e-12)  # unit Nyquist gain
    x = np.zeros((1, 100))
    x[0, 10] = 1
    np.testing.assert_allclose(po.fir_filter(lo, x)[0, 10:10 + len(lo)], lo, atol=1e-15)  # causal response
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal((2, 3, 90))
    band = po.fir_taps(0.05, 0.2)
    assert np.sum(po.fir_filter(band, a) * b) == pytest.approx(np.sum(a * po.fir_filter(band, b, adjoint=True)),
                                                               rel=1e-12)
