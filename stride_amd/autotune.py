"""Tile autotuner for the specialised 3D update.

For a grid and scheme, times short forward runs of the update kernel (HIP events
around every launch) for each candidate — kernel version (7: 4 waves, 8: 8 waves; an
LDS plane ring filled by direct-to-LDS loads, u read once, axis-0 terms pushed into
partial sums), rows per lane (tile 64 x waves*rows cells) and march chunk lengths that split the
slowest axis into 3..16 pieces — and returns the fastest. Results are cached per
(shape, absorbing, boundary) for the life of the process. ``autotune_batch`` does the same for
the two-shot update and ``autotune_density`` for the fused update of the density schemes
(tile rows x waves and chunk length); ``tune_solver`` applies the tuned tile of whichever
update a scheme runs.
"""

import numpy as np

from ._native import Solver

__all__ = ['candidates', 'autotune', 'candidates_batch', 'autotune_batch', 'candidates_density',
           'autotune_density', 'uses_fast_update', 'uses_density_update', 'tune_solver']

_CACHE = {}


def candidates(shape, versions=(7, 8)):
    nx = shape[0]
    chunks = sorted({max(1, -(-nx // k)) for k in (3, 4, 5, 6, 7, 8, 10, 12, 16)})
    # the march runs whole 12-plane periods: chunk lengths that are multiples of 12
    chunks12 = sorted({12 * max(1, -(-c // 12)) for c in chunks})
    rows_of = {7: (2,), 8: (1, 2)}
    return [(rows, chunk, version) for version in versions for rows in rows_of[version] for chunk in chunks12]


def autotune(plan, steps=24, device=0, verbose=False, cands=None):
    """(best (rows, chunk, version), [(milliseconds per launch, rows, chunk, version), ...]) for ``plan``'s grid."""
    key = (plan.shape, plan.absorbing, plan.boundary)
    if key in _CACHE:
        return _CACHE[key]
    nt = steps + 2
    s = Solver(plan.dim, plan.shape, plan.spacing, plan.absorbing, plan.dt, nt, device=device)
    s.set_scheme(2, 0, -2, plan.boundary, 0, 4, 0, 0, nt)
    s.set_medium(plan.vp)
    if plan.sponge_tables is not None:
        s.set_sponge(plan.sponge_tables)
    s.set_sources(plan.src_idx, plan.src_w)
    s.set_receivers(plan.rec_idx[:1], plan.rec_w[:1])
    src = np.zeros((nt, plan.nsrc), np.float32)
    src[1:, :] = 1.0
    results = []
    for rows, chunk, version in (cands or candidates(plan.shape)):
        s.set_tile(rows, chunk, version)
        s.forward(src)  # warm
        s.timing(True)
        s.forward(src)
        ms, n = s.timing()
        s.timing(False)
        results.append((ms / max(n, 1), rows, chunk, version))
        if verbose:
            print('tile rows=%d chunk=%d version=%d: %.1f us/launch' % (rows, chunk, version, 1e3 * ms / max(n, 1)),
                  flush=True)
    s.close()
    results.sort()
    best = results[0][1:]
    _CACHE[key] = (best, results)
    return best, results


def candidates_batch(shape):
    nx = shape[0]
    chunks = sorted({12 * max(1, -(-max(1, -(-nx // k)) // 12)) for k in (3, 4, 5, 6, 8, 10, 12, 16)})
    return [(rows, chunk, waves) for (rows, waves) in ((1, 8), (2, 8), (2, 4)) for chunk in chunks]


def autotune_batch(plan, plan1, steps=24, device=0, verbose=False, cands=None):
    """As ``autotune`` for the two-shot update: (best (rows, chunk, waves), results)."""
    key = ('b2', plan.shape, plan.absorbing, plan.boundary)
    if key in _CACHE:
        return _CACHE[key]
    nt = steps + 2
    s = Solver(plan.dim, plan.shape, plan.spacing, plan.absorbing, plan.dt, nt, device=device)
    s.set_scheme(2, 0, -2, plan.boundary, 0, 4, 0, 0, nt)
    s.set_medium(plan.vp)
    if plan.sponge_tables is not None:
        s.set_sponge(plan.sponge_tables)
    s.batch_create()
    for k, p in enumerate((plan, plan1)):
        s.batch_set_taps(k, p.src_idx, p.src_w, p.rec_idx[:1], p.rec_w[:1])
    src = [np.ones((nt, p.nsrc), np.float32) for p in (plan, plan1)]
    results = []
    for rows, chunk, waves in (cands or candidates_batch(plan.shape)):
        s.batch_set_tile(rows, chunk, waves)
        s.batch_forward(src)  # warm
        s.timing(True)
        s.batch_forward(src)
        ms, n = s.timing()
        s.timing(False)
        results.append((ms / max(n, 1), rows, chunk, waves))
        if verbose:
            print('two-shot tile rows=%d chunk=%d waves=%d: %.1f us/launch' % (rows, chunk, waves,
                                                                              1e3 * ms / max(n, 1)), flush=True)
    s.close()
    results.sort()
    best = results[0][1:]
    _CACHE[key] = (best, results)
    return best, results


def candidates_density(shape):
    nx = shape[0]
    chunks = sorted({max(1, -(-nx // k)) for k in (3, 4, 5, 6, 8)})
    return [(rows, chunk, waves) for (rows, waves) in ((1, 8), (1, 4), (2, 4)) for chunk in chunks]


def autotune_density(plan, steps=24, device=0, verbose=False, cands=None):
    """As ``autotune`` for the fused density update (``plan`` with rho): (best (rows, chunk, waves), results)."""
    key = ('rho', plan.shape, plan.absorbing, plan.boundary, plan.rho_mode, plan.att_power)
    if key in _CACHE:
        return _CACHE[key]
    nt = steps + 2
    s = Solver(plan.dim, plan.shape, plan.spacing, plan.absorbing, plan.dt, nt, device=device)
    s.set_scheme(2, plan.rho_mode, plan.att_power, plan.boundary, 0, 4, 0, 0, nt, 0, 0, 0)
    s.set_medium(plan.vp, plan.rho, plan.alpha)
    if plan.sponge_tables is not None:
        s.set_sponge(plan.sponge_tables)
    s.set_sources(plan.src_idx, plan.src_w)
    s.set_receivers(plan.rec_idx[:1], plan.rec_w[:1])
    src = np.zeros((nt, plan.nsrc), np.float32)
    src[1:, :] = 1.0
    results = []
    for rows, chunk, waves in (cands or candidates_density(plan.shape)):
        s.set_tile_density(rows, chunk, waves)
        s.forward(src)  # warm
        s.timing(True)
        s.forward(src)
        ms, n = s.timing()
        s.timing(False)
        results.append((ms / max(n, 1), rows, chunk, waves))
        if verbose:
            print('density tile rows=%d chunk=%d waves=%d: %.1f us/launch' % (rows, chunk, waves,
                                                                              1e3 * ms / max(n, 1)), flush=True)
    s.close()
    results.sort()
    best = results[0][1:]
    _CACHE[key] = (best, results)
    return best, results


def uses_fast_update(plan, save_mode):
    """The scheme runs on the specialised 3D update (3D, OT2, no density, no PML; any attenuation
    form); the tile is tuned on the plain update."""
    return (plan.dim == 3 and plan.kernel_order == 2 and plan.rho_mode == 0 and plan.boundary != 1
            and save_mode in (0, 1))


def uses_density_update(plan):
    """The scheme runs on the fused 3D density update (3D, OT2, constant or variable density, no
    PML; any attenuation form)."""
    return plan.dim == 3 and plan.kernel_order == 2 and plan.rho_mode != 0 and plan.boundary != 1


def tune_solver(solver, plan, save_mode, device=0):
    """Time the candidate tiles of the update ``plan``'s scheme runs on (cached per grid and
    scheme) and set the fastest on ``solver``. Returns the tile, or None for schemes on the
    general kernels."""
    if uses_fast_update(plan, save_mode):
        tile, _ = autotune(plan, device=device)
        solver.set_tile(*tile)
        return tile
    if uses_density_update(plan):
        tile, _ = autotune_density(plan, device=device)
        solver.set_tile_density(*tile)
        return tile
    return None
