"""Shots on two solvers with their own streams and host threads (stride_amd.workers) give the
same traces and gradients, bit for bit, as the same shots run one after the other on one
solver: the 2D gradient scheme mix (CFS-PML, variable density, power-2 attenuation)."""

import numpy as np
import pytest

from helpers import small_case, solver_for
from stride_amd.physics.plan import build_plan, adjoint_source_samples

pytestmark = pytest.mark.gpu


def test_two_workers_match_sequential_shots(gpu):
    from stride_amd.workers import ShotWorkers
    plan, ctx = small_case(dim=2, n=100, nabs=12, nt=260, rho='var', alpha='p2', attenuation_power=2,
                           boundary_type='complex_frequency_shift_PML_2', kernel='OT2', grad_vp=True,
                           grad_rho=True, nrec=24)
    geom = ctx['kw']['receiver_coordinates']
    plans = [build_plan(ctx['space'], ctx['time'], ctx['wavelet'], ctx['vp'], rho=ctx['rho'], alpha=ctx['alpha'],
                        source_coordinates=geom[k:k + 1], receiver_coordinates=geom, kernel='OT2',
                        attenuation_power=2, boundary_type='complex_frequency_shift_PML_2', grad_vp=True,
                        grad_rho=True) for k in (0, 5, 11, 17)]
    rng = np.random.default_rng(23)
    res = [adjoint_source_samples(p, rng.standard_normal((p.nrec, p.nt)).astype(np.float32)) for p in plans]
    print('workers: %d shots on %s' % (len(plans), plan.shape), flush=True)

    def shot(s, k):
        s.set_sources(plans[k].src_idx, plans[k].src_w)
        rec = s.forward(plans[k].src_data)
        g = s.adjoint(res[k], want_vp=True, want_rho=True)
        return rec, g

    seq = solver_for(plan, save_mode=2)
    want = [shot(seq, k) for k in range(len(plans))]
    seq.close()
    with ShotWorkers(lambda: solver_for(plan, save_mode=2), n=2) as pool:
        got = pool.map(shot, range(len(plans)))
    for k in range(len(plans)):
        assert np.array_equal(got[k][0], want[k][0]), k
        for key in ('grad_vp', 'prec_vp', 'grad_rho', 'prec_rho'):
            assert np.array_equal(got[k][1][key], want[k][1][key]), (k, key)
