"""Shots on two solvers with their own streams and host threads (stride_amd.workers) give the
same traces and gradients, bit for bit, as the same shots run one after the other on one
solver, and match the oracle: the 2D gradient scheme mix (CFS-PML, variable density, power-2
attenuation)."""

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
    # and both against the oracle (traces 1e-5, raw gradient accumulators 1e-4), two of the shots
    import oracle
    from helpers import rel_l2
    for k in (0, 3):
        ref, saved, _ = oracle.forward(plans[k], save=True)
        assert rel_l2(got[k][0].T, ref) < 1e-5, k
        gref = oracle.adjoint_compat(plans[k], res[k], saved)
        for key in ('grad_vp', 'grad_rho'):
            assert rel_l2(got[k][1][key], gref[key]) < 1e-4, (k, key)


def test_dev_shots_with_medium_changes_match_sequential(gpu):
    """Device-pointer shots (sa_forward_dev / sa_adjoint_dev return at once) on workers whose
    streams are non-blocking, with a different medium set before every shot: each shot's
    traces and vp gradient equal the synchronous one-solver run bit for bit (the setters wait
    for the work already queued on the solver's stream)."""
    import torch
    from stride_amd.workers import ShotWorkers
    plan, ctx = small_case(dim=2, n=100, nabs=12, nt=260, kernel='OT2', grad_vp=True, nrec=24)
    geom = ctx['kw']['receiver_coordinates']
    factors = (1.0, 1.04, 0.97, 1.06)
    plans = [build_plan(ctx['space'], ctx['time'], ctx['wavelet'], (ctx['vp'] * f).astype(np.float32),
                        source_coordinates=geom[k:k + 1], receiver_coordinates=geom, kernel='OT2', grad_vp=True)
             for k, f in enumerate(factors)]
    rng = np.random.default_rng(29)
    res = [adjoint_source_samples(p, rng.standard_normal((p.nrec, p.nt)).astype(np.float32)) for p in plans]
    print('workers: %d device-pointer shots with medium changes' % len(plans), flush=True)

    seq = solver_for(plan, save_mode=1)
    want = []
    for k, p in enumerate(plans):
        seq.set_medium(p.vp)
        seq.set_sources(p.src_idx, p.src_w)
        rec = seq.forward(p.src_data)
        want.append((rec, seq.adjoint(res[k], want_vp=True)['grad_vp']))
    seq.close()

    dev = torch.device('cuda', 0)
    src_d = [torch.from_numpy(np.ascontiguousarray(p.src_data)).to(dev) for p in plans]
    res_d = [torch.from_numpy(np.ascontiguousarray(r)).to(dev) for r in res]
    rec_d = [torch.zeros((p.nt, p.nrec), dtype=torch.float32, device=dev) for p in plans]
    grad_d = [torch.zeros(plan.shape, dtype=torch.float32, device=dev) for _ in plans]
    torch.cuda.synchronize()

    def shot(s, k):
        s.set_medium(plans[k].vp)
        s.set_sources(plans[k].src_idx, plans[k].src_w)
        s.forward_dev(src_d[k].data_ptr(), rec_d[k].data_ptr())
        s.adjoint_dev(res_d[k].data_ptr(), want=1)
        s.read_gradient_dev(0, grad_d[k].data_ptr())

    with ShotWorkers(lambda: solver_for(plan, save_mode=1), n=2) as pool:
        pool.map(shot, range(len(plans)))
    for k in range(len(plans)):
        assert np.array_equal(rec_d[k].cpu().numpy(), want[k][0]), k
        assert np.array_equal(grad_d[k].cpu().numpy(), want[k][1]), k


def test_sixteen_fresh_workers_match_sequential(gpu):
    """Sixteen fresh solvers, one shot each, all allocating their buffers at once on their own
    non-blocking streams: every shot equals the one-solver run bit for bit. (The buffers a solver
    allocates are cleared on the null stream, which non-blocking streams do not wait for; the
    clear is waited for before the buffer is used, or a late clear zeroes a wavefield already
    being computed.) Each worker also checks its own misfit against the oracle's traces."""
    import oracle
    from stride_amd.workers import ShotWorkers
    plan, ctx = small_case(dim=2, n=100, nabs=12, nt=260, kernel='OT2', grad_vp=True, nrec=32,
                           boundary_type='complex_frequency_shift_PML_2')
    geom = ctx['kw']['receiver_coordinates']
    plans = [build_plan(ctx['space'], ctx['time'], ctx['wavelet'], ctx['vp'], source_coordinates=geom[k:k + 1],
                        receiver_coordinates=geom, kernel='OT2', boundary_type='complex_frequency_shift_PML_2',
                        grad_vp=True) for k in range(0, 32, 2)]
    rng = np.random.default_rng(31)
    res = [adjoint_source_samples(p, rng.standard_normal((p.nrec, p.nt)).astype(np.float32)) for p in plans]
    print('workers: 16 fresh solvers, one shot each', flush=True)

    def shot(s, k):
        s.set_sources(plans[k].src_idx, plans[k].src_w)
        rec = s.forward(plans[k].src_data)
        return rec, s.adjoint(res[k], want_vp=True)['grad_vp']

    seq = solver_for(plan, save_mode=1)
    want = [shot(seq, k) for k in range(len(plans))]
    seq.close()
    with ShotWorkers(lambda: solver_for(plan, save_mode=1), n=16) as pool:
        got = pool.map(shot, range(len(plans)))
    for k in range(len(plans)):
        assert np.array_equal(got[k][0], want[k][0]), k
        assert np.array_equal(got[k][1], want[k][1]), k
    ref, _, _ = oracle.forward(plans[3])
    assert np.linalg.norm(got[3][0].T - ref) / np.linalg.norm(ref) < 1e-5
