"""Absorbing-boundary profiles, computed on the host.

``damping`` restates ``Boundary.damping`` (stride/physics/boundaries/boundary.py:59-171)
for the whole extended grid; the sponge (``SpongeBoundary2``,
stride/physics/boundaries/devito.py:37-69) and CFS-PML
(``ComplexFrequencyShiftPML2``, devito.py:72-193) parameter sets are built from it.

The device never receives the full damping volume: the sponge kernels evaluate
the layer rule on the fly from one 1-D table per axis (``sponge_tables``), so the
damping costs no HBM traffic. ``damping`` (whole volume) is the independent
statement the on-the-fly rule is checked against.
"""

import numpy as np

__all__ = ['reflection_coefficient', 'damping', 'sponge_tables', 'sponge_damp', 'pml_profiles',
           'SPONGE', 'PML', 'boundary_code']

SPONGE = 0
PML = 1

_BOUNDARIES = {
    'sponge_boundary_2': SPONGE,
    'complex_frequency_shift_PML_2': PML,
}


def boundary_code(boundary_type):
    try:
        return _BOUNDARIES[boundary_type]
    except KeyError:
        raise ValueError('Unknown boundary type %s' % boundary_type)


def reflection_coefficient(absorbing):
    """R = 10^(-(log10 N - 1)/log10 2 - 3) with N the largest layer thickness
    (stride/physics/boundaries/devito.py:51)."""
    n = max(*absorbing) if len(absorbing) > 1 else absorbing[0]
    return 10 ** (-(np.log10(n) - 1) / np.log10(2) - 3)


def _axis_coefficient(n_abs, h, power_degree, refl):
    if n_abs > 15:
        return (power_degree + 1) / 2 * np.log(1.0 / refl) / (n_abs * h)
    return 0.67 / h


def _layer_value(kind, n_abs, index, coefficient, mask, power_degree):
    pos = abs((n_abs - index - 1) / float(n_abs - 1))
    if kind == 'sine':
        pos = pos - np.sin(2 * np.pi * pos) / (2 * np.pi)
    elif kind == 'power':
        pos = pos ** power_degree
    elif kind == 'cosine':
        pos = np.cos(np.pi / 2 * (1 - pos))
        return -pos if mask else pos
    else:
        raise ValueError('Allowed dumping type are (`sine`, `power`)')
    if mask:
        pos = -pos
    return coefficient * pos


def damping(space, dimensions=None, damping_coefficient=None, mask=False, damping_type='sine',
            velocity=1.0, power_degree=2, reflection_coefficient=1e-3, assign=False):
    """Whole-grid damping field, following boundary.py:59-171 step by step:
    per axis and layer index the value is written into a slab that spans
    ``[index, n - index)`` of the other axes, so corner cells take the value of
    the outermost layer they belong to (and diagonal cells the sum of both)."""
    dimensions = tuple(range(space.dim)) if dimensions is None else dimensions
    shape = np.array(space.extended_shape).take(dimensions)
    if mask:
        damp = np.ones(shape, dtype=np.float32)
        if damping_coefficient is not None:
            damp *= damping_coefficient
    else:
        damp = np.zeros(shape, dtype=np.float32)

    for axis, dimension in enumerate(dimensions):
        n_abs = space.absorbing[dimension]
        coefficient = damping_coefficient
        if coefficient is None:
            coefficient = _axis_coefficient(n_abs, space.spacing[dimension], power_degree,
                                            reflection_coefficient)
        for index in range(n_abs):
            val = _layer_value(damping_type, n_abs, index, coefficient, mask, power_degree)
            for side in (0, 1):
                sl = [slice(index, s - index) for s in shape]
                pos = index if side == 0 else damp.shape[axis] - index - 1
                sl[axis] = slice(pos, pos + 1)
                if assign:
                    damp[tuple(sl)] = val
                else:
                    damp[tuple(sl)] += val

    if damping_coefficient is None:
        velocity = np.asarray(velocity)
        if damp.shape == velocity.shape:
            damp *= velocity
        else:
            damp *= np.max(velocity)
    return damp


def sponge_tables(space):
    """Per-axis 1-D layer values ``val_d[l]`` (l = distance from the outer edge)
    of the sine sponge profile before the velocity factor. Returned as one
    float32 array of shape (dim, max_layers)."""
    refl = reflection_coefficient(space.absorbing)
    width = max(space.absorbing) if space.dim else 0
    tables = np.zeros((space.dim, max(width, 1)), dtype=np.float32)
    for d in range(space.dim):
        n_abs = space.absorbing[d]
        coefficient = _axis_coefficient(n_abs, space.spacing[d], 2, refl)
        for index in range(n_abs):
            tables[d, index] = _layer_value('sine', n_abs, index, coefficient, False, 2)
    return tables


def sponge_damp(space, vp_extended, dt):
    """``damp`` Function of SpongeBoundary2: 7 * damping(velocity=vp) * dt
    (stride/physics/boundaries/devito.py:56-58)."""
    refl = reflection_coefficient(space.absorbing)
    return 7 * damping(space, velocity=vp_extended, reflection_coefficient=refl) * dt


def pml_profiles(space, vp_extended, f_centre):
    """CFS-PML per-axis profiles (stride/physics/boundaries/devito.py:100-131):
    sigma (power 3, times max vp), alpha (power 2 mask profile with coefficient
    0.01*pi*f_centre) and their centred 2nd-order first derivatives over the
    edge-padded arrays. Returns float32 array (dim, 4, max_n): rows
    sigma, alpha, d(sigma)/dx, d(alpha)/dx."""
    refl = reflection_coefficient(space.absorbing)
    n_max = max(space.extended_shape)
    out = np.zeros((space.dim, 4, n_max), dtype=np.float32)
    for d in range(space.dim):
        sigma = damping(space, dimensions=(d,), velocity=vp_extended, damping_type='power',
                        power_degree=3, reflection_coefficient=refl, assign=True)
        alpha = damping(space, dimensions=(d,), velocity=vp_extended, damping_type='power',
                        power_degree=2, damping_coefficient=0.01 * np.pi * f_centre,
                        mask=True, assign=True)
        h = space.spacing[d]
        n = sigma.shape[0]
        for row, prof in ((0, sigma), (1, alpha)):
            padded = np.pad(prof, ((1, 1),), mode='edge').astype(np.float32)
            out[d, row, :n] = prof
            out[d, row + 2, :n] = (padded[2:] - padded[:-2]) / np.float32(2 * h)
    return out
