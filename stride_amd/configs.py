"""Synthetic inputs for the five benchmark / parity configurations.

Each generator returns a dict with ``space``, ``time``, medium arrays on the
extended grid, ``wavelet`` and source / receiver coordinates (metres, inner
frame). Seeds: N1 = 1001, N2 = 1002, hemisphere = 3001, shots = 3002,
phantoms = 4000 + k. "Gaussian-filtered noise" is
``gaussian_filter(rng.standard_normal(shape), sigma / dx)`` scaled to [-1, 1].

Every generator takes a ``scale`` argument that shrinks the grid (and the time
axis with it) for parity tests; ``scale=1`` is the full configuration.
"""

import numpy as np
import scipy.ndimage

from .problem import Space, Time
from .utils import tone_burst, elliptical, ellipsoidal

__all__ = ['gaussian_noise', 'config1', 'config2', 'config3', 'config4', 'config5_phantom', 'shot_schedule']


def gaussian_noise(shape, sigma_cells, seed):
    rng = np.random.default_rng(seed)
    n = scipy.ndimage.gaussian_filter(rng.standard_normal(shape), sigma_cells)
    lo, hi = n.min(), n.max()
    return 2 * (n - lo) / (hi - lo) - 1


def _extend(space, inner):
    pad = [(e, e) for e in space.extra]
    return np.pad(inner, pad, mode='edge').astype(np.float32)


def _centre(space):
    return tuple(l / 2 for l in space.limit)


def config1(scale=1.0, nt=None):
    """2D forward: 200^2, dx 0.5 mm, sponge 20, Gaussian blob (peak +100 m/s, sigma 8 mm),
    128-element 40 mm ring, 500 kHz 3-cycle burst, dt 0.08 us x 1000."""
    n = max(int(round(200 * scale)), 24)
    dx = 0.5e-3 / scale if scale < 1 else 0.5e-3
    space = Space(shape=(n, n), spacing=dx, extra=(20, 20), absorbing=(20, 20))
    steps = nt or max(int(round(1000 * scale)), 60)
    time = Time(start=0., step=0.08e-6, num=steps)
    x = np.arange(n) * dx
    cx, cy = _centre(space)
    r2 = (x[:, None] - cx) ** 2 + (x[None, :] - cy) ** 2
    vp = 1500. + 100. * np.exp(-r2 / (2 * (8e-3) ** 2))
    radius = 40e-3 * min(scale, 1.0) if scale < 1 else 40e-3
    ring = elliptical(128, (radius, radius), (cx, cy))
    return dict(space=space, time=time, vp=_extend(space, vp), rho=np.full(space.extended_shape, 1000., np.float32),
                wavelet=tone_burst(500e3, 3, time.num, time.step), sources=ring[:1], receivers=ring)


def config2(scale=1.0, nt=None):
    """2D FWI gradient: 512^2, dx 0.2 mm, CFS-PML 30; vp = 1500 + 150 N1 in a 45 mm disc;
    rho = 1000 + 50 N2; alpha 0.5 dB/cm/MHz^2 (power 2); 128-element 48 mm ring; dt 0.04 us."""
    n = max(int(round(512 * scale)), 32)
    dx = 0.2e-3
    space = Space(shape=(n, n), spacing=dx, extra=(30, 30), absorbing=(30, 30))
    steps = nt or max(int(round(2500 * scale)), 80)
    time = Time(start=0., step=0.04e-6, num=steps)
    cx, cy = _centre(space)
    x = np.arange(n) * dx
    r = np.sqrt((x[:, None] - cx) ** 2 + (x[None, :] - cy) ** 2)
    disc_r = 45e-3 * n / 512
    mask = r <= disc_r
    n1 = gaussian_noise((n, n), 3e-3 / dx * scale, 1001) * mask
    n2 = gaussian_noise((n, n), 3e-3 / dx * scale, 1002) * mask
    vp = 1500. + 150. * n1
    rho = 1000. + 50. * n2
    alpha = np.full((n, n), 0.5)
    ring_r = 48e-3 * n / 512
    ring = elliptical(128, (ring_r, ring_r), (cx, cy))
    return dict(space=space, time=time, vp=_extend(space, vp), rho=_extend(space, rho),
                alpha=_extend(space, alpha), wavelet=tone_burst(500e3, 3, time.num, time.step),
                sources=ring, receivers=ring)


def _hemisphere_medium(space, seed=3001, sigma=4e-3, radius=55e-3):
    n = space.shape
    dx = space.spacing[0]
    c = _centre(space)
    axes = [np.arange(s) * dx - cc for s, cc in zip(n, c)]
    X, Y, Z = np.meshgrid(*axes, indexing='ij')
    rr = radius * n[0] / 320
    inside = (X ** 2 + Y ** 2 + Z ** 2 <= rr ** 2) & (Z >= 0)
    noise = gaussian_noise(n, sigma / dx * n[0] / 320, seed)
    vp = np.where(inside, 1400. + (noise + 1) / 2 * 250., 1500.)
    return vp, inside


def shot_schedule(num_shots=32, seed=3002, nt_range=(1500, 3000), nrec_range=(256, 1024)):
    rng = np.random.default_rng(seed)
    nts = rng.integers(nt_range[0], nt_range[1] + 1, size=num_shots)
    nrecs = rng.integers(nrec_range[0], nrec_range[1] + 1, size=num_shots)
    return nts, nrecs


def config3(scale=1.0, nt=None, nrec=1024):
    """3D forward: 320^3, dx 0.4 mm, extra/absorbing 20, smooth random vp (1400-1650 m/s) in a
    55 mm hemisphere, 1500 m/s water elsewhere, constant rho; 400 kHz burst, dt 0.08 us."""
    n = max(int(round(320 * scale)), 24)
    dx = 0.4e-3
    space = Space(shape=(n, n, n), spacing=dx, extra=(20, 20, 20), absorbing=(20, 20, 20))
    steps = nt or max(int(round(2500 * scale)), 60)
    time = Time(start=0., step=0.08e-6, num=steps)
    vp, _ = _hemisphere_medium(space)
    c = _centre(space)
    rr = 55e-3 * n / 320
    rec = ellipsoidal(nrec, (rr, rr, rr), c, threshold=0.5)
    return dict(space=space, time=time, vp=_extend(space, vp), rho=np.full(space.extended_shape, 1000., np.float32),
                wavelet=tone_burst(400e3, 3, time.num, time.step), sources=rec[:1], receivers=rec)


def config4(scale=1.0, nt=None, nrec=1024):
    """3D FWI gradient: config 3's grid, 2500 steps, variable vp / rho / alpha."""
    cfg = config3(scale=scale, nt=nt, nrec=nrec)
    space = cfg['space']
    n = space.shape
    dx = space.spacing[0]
    _, inside = _hemisphere_medium(space)
    rho = 1000. + 50. * gaussian_noise(n, 4e-3 / dx * n[0] / 320, 1002) * inside
    alpha = np.where(inside, 0.5, 0.)
    cfg['rho'] = _extend(space, rho)
    cfg['alpha'] = _extend(space, alpha)
    return cfg


def config5_phantom(k, n=512, dx=0.2e-3):
    """Random smooth phantom k: 1400-1650 m/s, sigma 2-5 mm, seed 4000 + k."""
    rng = np.random.default_rng(4000 + k)
    sigma = rng.uniform(2e-3, 5e-3)
    noise = gaussian_noise((n, n), sigma / dx, 4000 + k)
    return 1400. + (noise + 1) / 2 * 250.
