"""The specialised 3D update against the general kernels and the oracle: awkward
sizes (not multiples of the tile or of the 12-plane period), unequal layers per
axis, both kernel versions (7: 4 waves, 8: 8 waves) and their tile shapes, fused save and
gradient."""

import os

import numpy as np
import pytest

import oracle
from helpers import small_case, solver_for, rel_l2
from stride_amd.physics.plan import adjoint_source_samples

pytestmark = pytest.mark.gpu


def _run(plan, res_s, general, tile=None):
    os.environ['STRIDE_AMD_GENERAL_ONLY'] = '1' if general else '0'
    try:
        s = solver_for(plan, save_mode=1)
        if tile:
            s.set_tile(*tile)
        rec = s.forward(plan.src_data)
        g = s.adjoint(res_s, want_vp=True)
        return rec, g
    finally:
        os.environ.pop('STRIDE_AMD_GENERAL_ONLY', None)


@pytest.mark.parametrize('n,nabs,tile', [
    (30, 6, (2, 7, 7)), (37, 5, (2, 13, 7)), (41, 7, (2, 41, 7)), (37, 5, (2, 1, 7)), (70, 8, (2, 24, 7)),
    (66, 6, (2, 12, 7)), (70, 8, (2, 30, 7)), (30, 6, (1, 7, 8)), (37, 5, (2, 13, 8)), (41, 7, (1, 41, 8)),
    (37, 5, (1, 1, 8)), (70, 8, (1, 36, 8)), (66, 6, (1, 12, 8)), (66, 6, (2, 5, 8)), (41, 7, (2, 41, 8)),
    (70, 8, (2, 30, 8)),
])
def test_fast_matches_general_and_oracle(gpu, n, nabs, tile):
    plan, _ = small_case(dim=3, n=n, nabs=nabs, nt=90, kernel='OT2', grad_vp=True, nrec=20)
    print('fast3d n=%d nabs=%d tile=%s' % (n, nabs, tile), flush=True)
    rng = np.random.default_rng(4)
    res_s = adjoint_source_samples(plan, rng.standard_normal((plan.nrec, plan.nt)))
    rec_f, g_f = _run(plan, res_s, general=False, tile=tile)
    rec_g, g_g = _run(plan, res_s, general=True)
    ref, saved, _ = oracle.forward(plan, save=True)
    gref = oracle.adjoint_compat(plan, res_s, saved)
    assert rel_l2(rec_f, rec_g) < 1e-6
    assert rel_l2(rec_f.T, ref) < 1e-5
    for key in ('grad_vp', 'prec_vp'):
        assert rel_l2(g_f[key], g_g[key]) < 1e-5, key
        assert rel_l2(g_f[key], gref[key]) < 1e-4, key


@pytest.mark.parametrize('tile', [(2, 11, 7), (1, 13, 8), (2, 13, 8)])
def test_fast_unequal_layers(gpu, tile):
    from stride_amd.problem import Space, Time
    from stride_amd.utils import tone_burst, ellipsoidal
    from stride_amd.physics.plan import build_plan
    space = Space(shape=(26, 30, 34), spacing=0.5e-3, extra=(4, 7, 9), absorbing=(4, 7, 9))
    time = Time(start=0., step=0.08e-6, num=90)
    rng = np.random.default_rng(0)
    vp = (1500 + 30 * rng.standard_normal(space.extended_shape)).astype(np.float32)
    c = tuple(l / 2 for l in space.limit)
    geom = ellipsoidal(16, (4e-3,) * 3, c, threshold=0.5)
    plan = build_plan(space, time, tone_burst(0.5e6, 3, 90, 0.08e-6)[None], vp, source_coordinates=geom[:1],
                      receiver_coordinates=geom, kernel='OT2', grad_vp=True)
    res_s = adjoint_source_samples(plan, rng.standard_normal((plan.nrec, plan.nt)))
    print('fast3d unequal layers tile=%s' % (tile,), flush=True)
    rec_f, g_f = _run(plan, res_s, general=False, tile=tile)
    ref, saved, _ = oracle.forward(plan, save=True)
    assert rel_l2(rec_f.T, ref) < 1e-5
    gref = oracle.adjoint_compat(plan, res_s, saved)
    assert rel_l2(g_f['grad_vp'], gref['grad_vp']) < 1e-4


@pytest.mark.parametrize('tile', [(2, 24, 7), (1, 13, 8)])
def test_fast_anisotropic_spacing(gpu, tile):
    from stride_amd.problem import Space, Time
    from stride_amd.utils import tone_burst, ellipsoidal
    from stride_amd.physics.plan import build_plan
    space = Space(shape=(30, 34, 38), spacing=(0.5e-3, 0.45e-3, 0.4e-3), extra=(6, 6, 6), absorbing=(6, 6, 6))
    time = Time(start=0., step=0.06e-6, num=140)
    rng = np.random.default_rng(6)
    vp = (1500 + 30 * rng.standard_normal(space.extended_shape)).astype(np.float32)
    c = tuple(l / 2 for l in space.limit)
    geom = ellipsoidal(16, (4e-3,) * 3, c, threshold=0.5)
    plan = build_plan(space, time, tone_burst(0.5e6, 3, 140, 0.06e-6)[None], vp, source_coordinates=geom[:1],
                      receiver_coordinates=geom, kernel='OT2', grad_vp=True)
    res_s = adjoint_source_samples(plan, rng.standard_normal((plan.nrec, plan.nt)))
    print('fast3d anisotropic spacing tile=%s' % (tile,), flush=True)
    rec_f, g_f = _run(plan, res_s, general=False, tile=tile)
    ref, saved, _ = oracle.forward(plan, save=True)
    assert rel_l2(rec_f.T, ref) < 1e-5
    gref = oracle.adjoint_compat(plan, res_s, saved)
    assert rel_l2(g_f['grad_vp'], gref['grad_vp']) < 1e-4


@pytest.mark.parametrize('n,nabs,tile,power', [
    (37, 5, (2, 13, 7), 2), (41, 7, (1, 24, 8), 2), (30, 6, (2, 12, 8), 2),
    (37, 5, (2, 13, 7), 0), (41, 7, (1, 24, 8), 0), (30, 6, (2, 24, 7), None),
])
def test_fast_attenuation(gpu, n, nabs, tile, power):
    """Attenuation on the specialised update against the general passes and the oracle (traces and
    the as-written vp gradient): power 0 (dt alpha vp in the damping), power None (the step scaled
    by alpha) and power 2 (the Cbeta (lap u - lap other) term, the Laplacian of u handed to the
    next step)."""
    if power is None:
        plan, _ = small_case(dim=3, n=n, nabs=nabs, nt=90, kernel='OT2', grad_vp=True, nrec=20,
                             alpha='none-power')
        assert plan.att_power == -1
    else:
        plan, _ = small_case(dim=3, n=n, nabs=nabs, nt=90, kernel='OT2', grad_vp=True, nrec=20, alpha='p',
                             attenuation_power=power)
        assert plan.att_power == power
    print('fast3d attenuation %s n=%d nabs=%d tile=%s' % (power, n, nabs, tile), flush=True)
    rng = np.random.default_rng(6)
    res_s = adjoint_source_samples(plan, rng.standard_normal((plan.nrec, plan.nt)))
    rec_f, g_f = _run(plan, res_s, general=False, tile=tile)
    rec_g, g_g = _run(plan, res_s, general=True)
    ref, saved, _ = oracle.forward(plan, save=True)
    gref = oracle.adjoint_compat(plan, res_s, saved)
    assert rel_l2(rec_f, rec_g) < 1e-6
    assert rel_l2(rec_f.T, ref) < 1e-5
    for key in ('grad_vp', 'prec_vp'):
        assert rel_l2(g_f[key], g_g[key]) < 1e-5, key
        assert rel_l2(g_f[key], gref[key]) < 1e-4, key


def test_tile_versions(gpu):
    """Only versions 7 (rows 2) and 8 (rows 1, 2) exist; other tiles are refused."""
    plan, _ = small_case(dim=3, n=30, nabs=6, nt=90, kernel="OT2", grad_vp=True, nrec=4)
    s = solver_for(plan, save_mode=1)
    for bad in ((2, 24, 5), (2, 24, 3), (1, 24, 7), (4, 24, 8), (2, 0, 7)):
        with pytest.raises(ValueError):
            s.set_tile(*bad)
    s.set_tile(2, 24, 8)
    s.close()
