"""Benchmark: 3D 320^3 (360^3 extended) SO10 OT2 forward + adjoint vp gradient.

One step = one shot: the state run over nt time steps saving p.dt2 every 4th step
(interior), then the as-written adjoint accumulating the vp gradient and
preconditioner on the guarded steps -- the hot path of IsoAcousticDevito's
forward/adjoint with vp.needs_grad. Inputs (medium, taps, source samples,
residual) are resident in HBM before the timed region; nothing crosses PCIe
inside it.

Prints one JSON line (rank 0) with the headline Gpts/s (extended-grid point
updates per second, forward and adjoint updates counted separately), the
roofline of the dominant kernel (the wavefield update, timed with HIP events on
the solver stream) and the CPU oracle timed on this host on a bounded sample.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def workload(scale=1.0, nt=2500, nrec=1024, undersampling=4):
    from stride_amd.configs import config3
    from stride_amd.physics.plan import build_plan
    cfg = config3(scale=scale, nt=nt, nrec=nrec)
    plan = build_plan(cfg['space'], cfg['time'], cfg['wavelet'][None], cfg['vp'],
                      source_coordinates=cfg['sources'], receiver_coordinates=cfg['receivers'],
                      kernel='OT2', save_undersampling=undersampling, grad_vp=True)
    return cfg, plan


def make_solver(plan):
    from stride_amd._native import Solver
    s = Solver(plan.dim, plan.shape, plan.spacing, plan.absorbing, plan.dt, plan.nt)
    s.set_scheme(plan.kernel_order, plan.rho_mode, plan.att_power, plan.boundary, plan.diff_source,
                 plan.undersampling, 1, plan.tb0, plan.tb1, 1, 0, 0)
    s.set_medium(plan.vp)
    s.set_sponge(plan.sponge_tables)
    s.set_sources(plan.src_idx, plan.src_w)
    s.set_receivers(plan.rec_idx, plan.rec_w)
    return s


def algorithmic_bytes(plan):
    """Bytes the update kernel must move per shot: 16 B per point update (read u, other
    level, vp; write the new level); +4 B per interior point on save steps (the p.dt2
    snapshot); +20 B per interior point on gradient steps (read the snapshot, read and
    write grad and prec)."""
    n = int(np.prod(plan.shape))
    interior = int(np.prod([s - 2 * a for s, a in zip(plan.shape, plan.absorbing)]))
    f = plan.undersampling
    fwd_steps = plan.tb1 - 2
    adj_steps = plan.tb1 - 1 - f - plan.tb0
    saves = sum(1 for t in range(1, plan.tb1 - 1) if t % f == 0)
    grads = sum(1 for t in range(plan.tb1 - 1 - f, plan.tb0, -1) if t % f == 0)
    total = 16 * n * (fwd_steps + adj_steps) + 4 * interior * saves + 20 * interior * grads
    return total, fwd_steps, adj_steps, n, interior


def cpu_baseline(plan_full, seconds_budget=20.0):
    """The oracle (C/OpenMP, test infrastructure) on the same grid and medium, a bounded
    number of forward + adjoint steps, timed on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    from stride_amd.configs import config3
    from stride_amd.physics.plan import build_plan, adjoint_source_samples
    from stride_amd.problem import Time
    from stride_amd.utils import tone_burst
    nt = 40
    cfg = config3(scale=1.0, nt=200, nrec=1024)
    cfg['time'] = Time(start=0., step=cfg['time'].step, num=nt)
    cfg['wavelet'] = tone_burst(400e3, 3, 200, cfg['time'].step)[:nt]
    plan = build_plan(cfg['space'], cfg['time'], cfg['wavelet'][None], cfg['vp'],
                      source_coordinates=cfg['sources'], receiver_coordinates=cfg['receivers'],
                      kernel='OT2', save_undersampling=4, grad_vp=True)
    res = adjoint_source_samples(plan, np.random.default_rng(0).standard_normal((plan.nrec, nt)))
    t0 = time.perf_counter()
    _, saved, _ = oracle.forward(plan, save=True)
    t1 = time.perf_counter()
    oracle.adjoint_compat(plan, res, saved)
    t2 = time.perf_counter()
    n = int(np.prod(plan.shape))
    fwd = nt - 2
    adj = nt - 1 - plan.undersampling
    gpts = n * (fwd + adj) / (t2 - t0) / 1e9
    threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    return dict(value=round(gpts, 4), unit='Gpts/s', cores=threads, kind='port',
                sample='oracle (C/OpenMP) on the same 360^3 grid: %d forward steps with p.dt2 save + %d '
                       'as-written adjoint steps with vp gradient, %.1f s' % (fwd, adj, t2 - t0))


def traffic_from_profile():
    path = os.path.join(ROOT, 'profiles', 'fast3d_traffic.json')
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get('bytes_per_launch')
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--scale', type=float, default=1.0, help='grid scale (1.0 = 320^3 inner)')
    ap.add_argument('--nt', type=int, default=2500)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--tile', type=str, default=None, help='rows,chunk,version of the 3D update tile')
    ap.add_argument('--no-autotune', action='store_true')
    ap.add_argument('--copy-check', type=int, default=0, metavar='MB',
                    help='only run 10 device-to-device copies of MB megabytes (calibrates the byte counters '
                         'of a profiler against a transfer of known size) and exit')
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py measures one GPU: the solver shards by shot, and multi-GPU runs are out of scope '
              'for this build (see DESIGN.md, sharding).', file=sys.stderr)
        sys.exit(2)

    if args.copy_check:
        from stride_amd._native import Solver
        s = Solver(3, (8, 8, 8), (1e-3,) * 3, (0, 0, 0), 1e-7, 3)
        nbytes = args.copy_check * 1000000
        a, b = s.device_alloc(nbytes), s.device_alloc(nbytes)
        for _ in range(10):
            s.memcpy(b, a, nbytes, 3)
        s.synchronize()
        s.device_free(a)
        s.device_free(b)
        s.close()
        print(json.dumps({'copy_check_bytes': nbytes, 'copies': 10}))
        return

    cfg, plan = workload(scale=args.scale, nt=args.nt)
    print('[bench] workload %s ext, nt %d, f %d, nrec %d' % (plan.shape, plan.nt, plan.undersampling, plan.nrec),
          file=sys.stderr, flush=True)
    tile = None
    if args.tile:
        tile = tuple(int(v) for v in args.tile.split(','))
    elif not args.no_autotune:
        from stride_amd.autotune import autotune
        t0 = time.perf_counter()
        tile, table = autotune(plan)
        print('[bench] autotune %.1f s: best rows=%d chunk=%d version=%d (%.1f us/launch); worst %.1f us'
              % ((time.perf_counter() - t0,) + tuple(tile) + (table[0][0] * 1e3, table[-1][0] * 1e3)),
              file=sys.stderr, flush=True)
    s = make_solver(plan)
    if tile:
        s.set_tile(*tile)
    # device-resident inputs
    nbytes_src = plan.src_data.nbytes
    rng = np.random.default_rng(0)
    res = rng.standard_normal((plan.nt, plan.nrec)).astype(np.float32) * 1e-3
    d_src = s.device_alloc(nbytes_src)
    d_rec = s.device_alloc(plan.nt * plan.nrec * 4)
    d_res = s.device_alloc(res.nbytes)
    s.memcpy(d_src, plan.src_data.ctypes.data, nbytes_src, 1)
    s.memcpy(d_res, res.ctypes.data, res.nbytes, 1)

    def one_shot():
        s.forward_dev(d_src, d_rec)
        s.adjoint_dev(d_res, want=1)

    for _ in range(args.warmup):
        one_shot()
    s.synchronize()
    s.timing(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        one_shot()
        print('[bench] step %d issued' % k, file=sys.stderr, flush=True)
    s.synchronize()
    t1 = time.perf_counter()
    kernel_ms, launches = s.timing()
    s.timing(False)

    total_bytes, fwd_steps, adj_steps, n, interior = algorithmic_bytes(plan)
    wall = (t1 - t0) / args.steps
    updates = n * (fwd_steps + adj_steps)
    gpts = updates / wall / 1e9
    avg_launch_s = kernel_ms / 1e3 / max(launches, 1)
    bytes_per_launch = total_bytes / (fwd_steps + adj_steps)
    achieved = bytes_per_launch / avg_launch_s / 1e9

    grad = s.read_gradient(0)
    if not np.isfinite(grad).all():
        print('[bench] non-finite gradient', file=sys.stderr)
        sys.exit(1)

    out = {
        'metric': 'Gpts/s fwd+adjoint 3D 320^3 SO10 (% HBM peak)',
        'value': round(gpts, 3),
        'unit': 'Gpts/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(wall * 1e3, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (smooth random vp in a 55 mm hemisphere, seeds 3001/3002)',
        'config': {
            'workload': '3D %dx%dx%d inner (%dx%dx%d extended), dx 0.4 mm, SO10 OT2, sponge %d, nt %d, '
                        'fwd with p.dt2 save every %d + as-written adjoint with vp gradient, 1 source, %d receivers, '
                        'one shot per step' % (tuple(cfg['space'].shape) + tuple(plan.shape) +
                                               (plan.absorbing[0], plan.nt, plan.undersampling, plan.nrec)),
            'point_updates_per_step': updates,
            'inner_gpts': round(interior * (fwd_steps + adj_steps) / wall / 1e9, 3),
            'hbm_roofline_frac_of_value': round(gpts * 1e9 * (total_bytes / updates) / 1e9 / HBM_PEAK_GBS, 4),
        },
        'roofline': {
            'bound': 'hbm',
            'achieved': round(achieved, 1),
            'peak': HBM_PEAK_GBS,
            'unit': 'GB/s',
            'frac': round(achieved / HBM_PEAK_GBS, 4),
            'traffic': traffic_from_profile(),
            'kernel': 'k_fast3d (wavefield update)',
            'tile': list(tile) if tile else None,
            'avg_launch_us': round(avg_launch_s * 1e6, 2),
            'launches': launches,
            'algorithmic_bytes_per_launch': int(bytes_per_launch),
        },
    }
    if not args.no_cpu:
        try:
            out['cpu_baseline'] = cpu_baseline(plan)
        except Exception as exc:  # the baseline is reported, never required
            out['cpu_baseline'] = {'error': repr(exc)}
    for p in (d_src, d_rec, d_res):
        s.device_free(p)
    s.close()
    print(json.dumps(out))


if __name__ == '__main__':
    main()
