"""Off-grid point operators: interpolation taps for sources and receivers.

Each sparse point becomes a fixed-size list of (grid cell, weight) taps:

- ``linear``: multilinear 2^d taps from ``(x - pml_origin) / h`` (Devito's
  ``SparseTimeFunction`` default interpolation, used by
  stride/physics/common/devito.py:631-632).
- ``hicks``: Kaiser-windowed sinc (Hicks 2002, Geophysics 67(1)), half-width 3,
  b = 4.14, window extended to 3/0.99, tensor product of 7 taps per axis; sources
  get the 5-tap smoothing blend (stride/physics/common/devito.py:843-885).
  Support offsets are -3..+3 around ``floor`` of the grid coordinate.
- ``sinc``: the reference hands it to Devito (``SparseTimeFunction(interpolation='sinc',
  r=7)``, stride/physics/common/devito.py:634-638). Devito's ``SincInterpolator`` (Devito >= 4.8;
  the reference pins ``devito>=4.6``) is absent here, so its published method is restated:
  Kaiser-windowed sinc of Hicks (2002), 2r taps per axis at ``floor(x) - r + 1 .. floor(x) + r``,
  weight ``sinc(d) I0(b sqrt(1 - (d / r)^2)) / I0(b)`` for the offset d of the tap from the
  point, b from Hicks's Table 1 as Devito tabulates it (r = 7: 7.51). No smoothing.

Taps that fall outside the extended grid are dropped (weight 0, index -1).
Coordinates are shifted by 1e-3*h first (stride/physics/iso_acoustic/devito.py:307-309).
"""

import numpy as np
import scipy.special

__all__ = ['linear_taps', 'hicks_coefficients', 'smooth_source_coefficients', 'hicks_taps', 'sinc_coefficients',
           'sinc_taps', 'point_taps', 'INTERPOLATION']

INTERPOLATION = ('linear', 'hicks', 'sinc')

# Kaiser window parameter per half-width r (Hicks 2002, Table 1, as Devito's SincInterpolator
# tabulates it)
SINC_B = {2: 2.94, 3: 4.53, 4: 4.14, 5: 5.26, 6: 6.40, 7: 7.51, 8: 8.56, 9: 9.56, 10: 10.64}


def _grid_coordinates(space, coordinates):
    coords = np.asarray(coordinates, dtype=np.float32).astype(np.float64)
    return (coords - np.array(space.pml_origin)) / np.array(space.spacing)


def _flatten(space, cells, weights):
    """cells (n, taps, dim) int64, weights (n, taps) -> linear indices, dropping
    out-of-grid taps."""
    shape = np.array(space.extended_shape)
    inside = np.all((cells >= 0) & (cells < shape), axis=-1)
    strides = np.array([int(np.prod(shape[d + 1:])) for d in range(space.dim)], dtype=np.int64)
    index = np.where(inside, (np.clip(cells, 0, shape - 1) * strides).sum(-1), -1)
    weights = np.where(inside, weights, 0.)
    return index.astype(np.int64), weights.astype(np.float32)


def linear_taps(space, coordinates):
    g = _grid_coordinates(space, coordinates)
    base = np.floor(g).astype(np.int64)
    frac = (g - base).astype(np.float32)
    dim = space.dim
    corners = np.array(np.meshgrid(*[[0, 1]] * dim, indexing='ij')).reshape(dim, -1).T
    cells = base[:, None, :] + corners[None, :, :]
    w = np.ones((g.shape[0], corners.shape[0]), dtype=np.float32)
    for d in range(dim):
        wd = np.where(corners[None, :, d] == 1, frac[:, d:d + 1], np.float32(1) - frac[:, d:d + 1])
        w = w * wd
    return _flatten(space, cells, w)


def smooth_source_coefficients(coef):
    """The 5-tap blend applied to source coefficients (n, dim, 8) in place
    (stride/physics/common/devito.py:874-883), with n = 2 the column of offset -1.

    The reference binds a, b, c as numpy views of columns n, n+1, n+2
    (common/devito.py:876-878), so each assignment is seen by the lines after it: the
    n+1 line reads the new column n, the n+2 line the new column n+1 and the n+3 line
    the new column n+2. The columns are blended in that order, in place."""
    n = 2
    col = [coef[..., k] for k in range(coef.shape[-1])]  # views, as in the reference
    col[n - 1] += col[n] * 0.01
    col[n][...] = col[n] * 0.98 + col[n + 1] * 0.03
    col[n + 1][...] = col[n + 1] * 0.94 + (col[n] + col[n + 2]) * 0.01
    col[n + 2][...] = col[n + 2] * 0.98 + col[n + 1] * 0.03
    col[n + 3] += col[n + 2] * 0.01
    return coef


def hicks_coefficients(space, coordinates, smooth=False):
    """(reference cells (n, dim) int32, coefficients (n, dim, 8) float64), the
    8th coefficient always zero."""
    g = _grid_coordinates(space, coordinates)
    ref = np.floor(g).astype(np.int32)
    offsets = g - ref
    b = 4.14
    half = 3
    den = scipy.special.iv(0, b)
    width = half / 0.99
    coef = np.zeros((g.shape[0], space.dim, 2 * half + 2))
    for k in range(-half, half + 1):
        x = offsets - k
        w = np.minimum((x / width) ** 2, 1.)
        coef[:, :, half + k] = np.sinc(x) * scipy.special.iv(0, b * np.sqrt(1 - w)) / den
    if smooth:
        smooth_source_coefficients(coef)
    return ref, coef


def hicks_taps(space, coordinates, smooth=False):
    ref, coef = hicks_coefficients(space, coordinates, smooth=smooth)
    dim = space.dim
    k = np.arange(-3, 4)
    grids = np.array(np.meshgrid(*[k] * dim, indexing='ij')).reshape(dim, -1).T  # (7^d, dim)
    cells = ref[:, None, :].astype(np.int64) + grids[None, :, :]
    w = np.ones((ref.shape[0], grids.shape[0]))
    for d in range(dim):
        w = w * coef[:, d, grids[:, d] + 3]
    return _flatten(space, cells, w)


def sinc_coefficients(space, coordinates, r=7):
    """(first tap cells (n, dim) int64 = floor(x) - r + 1, coefficients (n, dim, 2r) float64)."""
    if r not in SINC_B:
        raise ValueError('sinc interpolation radius %d not in %s' % (r, sorted(SINC_B)))
    g = _grid_coordinates(space, coordinates)
    ref = np.floor(g)
    frac = g - ref
    b = SINC_B[r]
    coef = np.zeros((g.shape[0], space.dim, 2 * r))
    for i in range(2 * r):
        d = i - r + 1 - frac  # tap minus point, in cells: within [-r, r]
        coef[:, :, i] = np.sinc(d) * scipy.special.i0(b * np.sqrt(np.maximum(1 - (d / r) ** 2, 0.))) \
            / scipy.special.i0(b)
    return ref.astype(np.int64) - r + 1, coef


def sinc_taps(space, coordinates, r=7):
    first, coef = sinc_coefficients(space, coordinates, r)
    dim = space.dim
    k = np.arange(2 * r)
    grids = np.array(np.meshgrid(*[k] * dim, indexing='ij')).reshape(dim, -1).T  # ((2r)^d, dim)
    cells = first[:, None, :] + grids[None, :, :]
    w = np.ones((first.shape[0], grids.shape[0]))
    for d in range(dim):
        w = w * coef[:, d, grids[:, d]]
    return _flatten(space, cells, w)


def point_taps(space, coordinates, interpolation_type='linear', smooth=False):
    """Shift by 1e-3*h and return (index (n, taps) int64, weight (n, taps) float32)."""
    coords = np.asarray(coordinates, dtype=np.float32) + 1e-3 * np.array(space.spacing).reshape(1, -1)
    if interpolation_type == 'linear':
        return linear_taps(space, coords)
    if interpolation_type == 'hicks':
        return hicks_taps(space, coords, smooth=smooth)
    if interpolation_type == 'sinc':
        return sinc_taps(space, coords)
    raise ValueError('Only "linear" and "hicks" interpolations are allowed.')
