"""Source signals, transducer geometries and spectral bandwidth.

Host-side numpy restatements of the helpers that feed the solver:

- ``tone_burst``    stride/utils/wavelets.py:6-53  (Gaussian-windowed sine, Tukey(0.05) taper)
- ``ricker``        stride/utils/wavelets.py:56-84
- ``elliptical``    stride/utils/geometries.py:115-142 (2D ring)
- ``ellipsoidal``   stride/utils/geometries.py:145-214 (3D Fibonacci sphere / hemisphere)
- ``magnitude_spectrum`` / ``bandwidth``  stride/utils/fft.py:7-60, 99-141
"""

import numpy as np
import scipy.linalg
import scipy.signal

__all__ = ['tone_burst', 'ricker', 'elliptical', 'ellipsoidal', 'magnitude_spectrum',
           'bandwidth', 'tukey']


def tukey(n, alpha):
    """Tukey window as the reference builds it, ``scipy.signal.get_window(('tukey', a), n, False)``
    (third argument ``fftbins=False``: the symmetric window)."""
    return scipy.signal.get_window(('tukey', alpha), n, False)


def tone_burst(centre_freq, n_cycles, n_samples, dt, envelope='gaussian', offset=0):
    duration = n_cycles / centre_freq
    n_tone = int(duration // dt + 1)
    t = np.linspace(0, duration, n_tone, endpoint=False)
    signal = np.sin(2 * np.pi * centre_freq * t)
    if envelope == 'gaussian':
        x = np.linspace(-3, 3, n_tone)
        signal = signal * np.exp(-x ** 2 / 2)
    elif envelope != 'rectangular':
        raise Exception('Envelope type not implemented')
    signal = signal * tukey(n_tone, 0.05)
    return np.pad(signal, ((offset, n_samples - offset - n_tone),), mode='constant')


def ricker(centre_freq, n_samples, dt, offset=0):
    duration = 3 * np.sqrt(6) / (centre_freq * np.pi)
    t = np.linspace(-duration / 2, duration / 2, int(duration // dt + 1), endpoint=False)
    a = (np.pi * centre_freq * t) ** 2
    signal = (1 - 2 * a) * np.exp(-a)
    signal = signal * tukey(signal.shape[0], 0.05)
    return np.pad(signal, ((offset, n_samples - offset - signal.shape[0]),), mode='constant')


def elliptical(num, radius, centre):
    angles = np.linspace(0, 2 * np.pi, num, endpoint=False)
    out = np.empty((num, 2))
    out[:, 0] = radius[0] * np.cos(angles) + centre[0]
    out[:, 1] = radius[1] * np.sin(angles) + centre[1]
    return out


def _rotation(axis, theta):
    axis = np.asarray(axis, dtype=np.float64)
    return scipy.linalg.expm(np.cross(np.eye(3), axis / np.linalg.norm(axis) * theta))


def ellipsoidal(num, radius, centre, theta=0., axis=None, threshold=0., flip=False,
                angle_range=np.pi):
    """Fibonacci-sphere points; ``threshold`` removes the lower (or upper if
    ``flip``) fraction of the sphere, so ``threshold=0.5`` gives a hemisphere."""
    golden = np.pi * (3. - np.sqrt(5.))
    start_offset = np.sin(np.pi / 2 - angle_range / 2)
    threshold = threshold or start_offset
    total = int(np.round(num / (1 - threshold)))
    offset = 2. / total
    rot = _rotation(axis or [1, 0, 0], theta)
    points = []
    for sample in range(total):
        z = (sample * offset - 1) + offset / 2
        r = np.sqrt(1 - z ** 2)
        phi = ((sample + 1) % total) * golden
        y = np.cos(phi) * r
        x = np.sin(phi) * r
        if not flip and z + 1 < threshold * 2:
            continue
        if flip and z + 1 > (1 - threshold) * 2:
            continue
        p = rot @ np.array([x * radius[0], y * radius[1], z * radius[2]])
        points.append(p + np.asarray(centre, dtype=np.float64))
    return np.array(points).reshape(-1, 3)


def magnitude_spectrum(signal, dt, db=True):
    num = signal.shape[-1]
    freqs = np.fft.rfftfreq(num, dt)
    mag = np.abs(np.fft.rfft(signal, axis=-1, norm='forward'))
    if db:
        mag = 20 * np.log10((mag + 1e-31) / (np.max(mag) + 1e-31))
    return freqs, mag


def bandwidth(signal, dt, cutoff=-10):
    """(f_min, f_centre, f_max): first / last rFFT bin above ``cutoff`` dB and the peak bin."""
    freqs, mag = magnitude_spectrum(signal, dt, db=True)
    if mag.ndim > 1:
        mag = np.mean(mag, axis=0)
    above = np.nonzero(mag > cutoff)[0]
    f_min = freqs[above[0]] if above.size else 0
    f_max = freqs[above[-1]] if above.size else mag.shape[-1]
    f_centre = freqs[np.argmax(mag)]
    return f_min, f_centre, f_max
