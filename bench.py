"""Benchmark: 3D 320^3 (360^3 extended) SO10 OT2 forward + adjoint vp gradient.

One step = one batch of shots (default 2, `--shots 1` for one): the state run over nt
time steps saving p.dt2 every 4th step (interior), then the as-written adjoint
accumulating the vp gradient and preconditioner on the guarded steps -- the hot path of
IsoAcousticDevito's forward/adjoint with vp.needs_grad. Two shots advance together
through one update kernel (interleaved wavefields, one vp read, one gradient
read-modify-write for both). Inputs (medium, taps, source samples, residuals) are
resident in HBM before the timed region; nothing crosses PCIe inside it.

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
# the timed region brackets every KERNEL_SAMPLE-th update launch with HIP events (7 is prime to
# the 4-step save / gradient period, so every launch kind is sampled in proportion); events on
# every launch add about 1.7 % to the step
KERNEL_SAMPLE = 7


def workload(scale=1.0, nt=2500, nrec=1024, undersampling=4):
    from stride_amd.configs import config3
    from stride_amd.physics.plan import build_plan
    cfg = config3(scale=scale, nt=nt, nrec=nrec)
    plan = build_plan(cfg['space'], cfg['time'], cfg['wavelet'][None], cfg['vp'],
                      source_coordinates=cfg['sources'], receiver_coordinates=cfg['receivers'],
                      kernel='OT2', save_undersampling=undersampling, grad_vp=True)
    return cfg, plan


def second_shot(cfg, plan):
    """The second shot of a two-shot step: the source at the receiver diametrically across the
    hemisphere from the first source, the same 1024 receivers."""
    from stride_amd.physics.plan import build_plan
    rec = cfg['receivers']
    d = np.linalg.norm(rec - cfg['sources'][0], axis=1)
    return build_plan(cfg['space'], cfg['time'], cfg['wavelet'][None], cfg['vp'],
                      source_coordinates=rec[int(np.argmax(d))][None], receiver_coordinates=rec,
                      kernel='OT2', save_undersampling=plan.undersampling, grad_vp=True)


def make_batch_solver(plan, plan1):
    from stride_amd._native import Solver
    s = Solver(plan.dim, plan.shape, plan.spacing, plan.absorbing, plan.dt, plan.nt)
    s.set_scheme(plan.kernel_order, plan.rho_mode, plan.att_power, plan.boundary, plan.diff_source,
                 plan.undersampling, 1, plan.tb0, plan.tb1, 1, 0, 0)
    s.set_medium(plan.vp)
    s.set_sponge(plan.sponge_tables)
    s.batch_create()
    for k, p in enumerate((plan, plan1)):
        s.batch_set_taps(k, p.src_idx, p.src_w, p.rec_idx, p.rec_w)
    return s


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


def algorithmic_bytes(plan, shots=1):
    """Bytes the update kernel must move per step of `shots` shots advanced together.
    One shot: 16 B per point update (read u, other level, vp; write the new level); +4 B per
    interior point on save steps (the p.dt2 snapshot); +20 B per interior point on gradient
    steps (read the snapshot, read and write grad and prec). Two interleaved shots: 28 B per
    point and pass (u, other level, new level of both shots, one vp); +8 B on save steps;
    +24 B on gradient steps (both snapshots, one grad and one prec read-modify-write)."""
    n = int(np.prod(plan.shape))
    interior = int(np.prod([s - 2 * a for s, a in zip(plan.shape, plan.absorbing)]))
    f = plan.undersampling
    fwd_steps = plan.tb1 - 2
    adj_steps = plan.tb1 - 1 - f - plan.tb0
    saves = sum(1 for t in range(1, plan.tb1 - 1) if t % f == 0)
    grads = sum(1 for t in range(plan.tb1 - 1 - f, plan.tb0, -1) if t % f == 0)
    if shots == 1:
        total = 16 * n * (fwd_steps + adj_steps) + 4 * interior * saves + 20 * interior * grads
    else:
        total = 28 * n * (fwd_steps + adj_steps) + 8 * interior * saves + 24 * interior * grads
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
    model, nproc = host_cpu()
    return dict(value=round(gpts, 4), unit='Gpts/s', cores=threads, kind='port', cpu_model=model, nproc=nproc,
                sample='oracle (C/OpenMP) on the same 360^3 grid: %d forward steps with p.dt2 save + %d '
                       'as-written adjoint steps with vp gradient, %.1f s' % (fwd, adj, t2 - t0))


def traffic_from_profile(shots=1, tile=None):
    """Calibrated HBM bytes per launch of the update kernel from the committed counter run, or
    None when that run measured another tile than the one timed (the figure would be stale)."""
    path = os.path.join(ROOT, 'profiles', 'fast3d_traffic.json' if shots == 1 else 'fast3d_b2_traffic.json')
    if not os.path.exists(path):
        return None
    try:
        rec = json.load(open(path))
    except Exception:
        return None
    if tile is None:
        return None
    by_tile = rec.get('by_tile', {})
    key = ','.join(str(int(v)) for v in tile)
    if key in by_tile:
        return by_tile[key].get('bytes_per_launch')
    if rec.get('tile') is None or list(rec['tile']) != list(tile):
        return None
    return rec.get('bytes_per_launch')


def host_cpu():
    """The host's CPU model and logical core count (what the CPU baseline ran on)."""
    model = None
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def bench_pipelines(plan, reps=5):
    """Wall time of each device step either side of the solver at the full workload: traces of
    nrec x nt (Traces rows), fields on the extended grid (gradient steps on the inner box). Each
    step returns with its work finished; `gbs` is the step's algorithmic bytes (inputs read once,
    outputs written once) over its time."""
    import torch
    from stride_amd import pipelines as pl
    from stride_amd._native import Solver
    s = Solver(plan.dim, plan.shape, plan.spacing, plan.absorbing, plan.dt, plan.nt)
    dt = plan.dt
    rng = np.random.default_rng(1)
    ntr, nt = plan.nrec, plan.nt
    mod = torch.from_numpy(rng.standard_normal((ntr, nt)).astype(np.float32)).cuda()
    obs = torch.from_numpy(rng.standard_normal((ntr, nt)).astype(np.float32)).cuda()
    obs[:, :200] = 0
    tm = torch.from_numpy(rng.standard_normal((nt, ntr)).astype(np.float32)).cuda()
    window = np.ones(nt, np.float32)
    ext = tuple(plan.shape)
    inner = tuple((a, n - a) for a, n in zip(plan.absorbing, ext))
    ishape = tuple(b - a for a, b in inner)
    g_ext = torch.from_numpy(rng.standard_normal(ext).astype(np.float32)).cuda()
    g_in = torch.from_numpy(rng.standard_normal(ishape).astype(np.float32)).cuda()
    p_in = torch.from_numpy(rng.random(ishape).astype(np.float32)).cuda()
    vp_in = (1500 + 20 * torch.rand(ishape, device='cuda')).contiguous()
    d_in = (1e-6 * torch.randn(ishape, device='cuda')).contiguous()
    band = pl.FilterTraces(f_min=0.2e6, f_max=0.8e6)
    low = pl.FilterTraces(f_max=0.8e6)
    fir = pl.FilterTraces(f_min=0.2e6, f_max=0.8e6, filter_type='fir')
    T, F = ntr * nt * 4, int(np.prod(ext)) * 4
    I = int(np.prod(ishape)) * 4
    steps = [
        ('check_traces', lambda: pl.CheckTraces().forward(s, mod), 2 * T),
        ('mute_traces', lambda: pl.MuteTraces(f_max=0.8e6).forward(s, mod, obs, dt), 4 * T),
        ('filter butterworth band (8 sections)', lambda: band.forward(s, mod, dt=dt), 2 * T),
        ('filter butterworth band adjoint', lambda: band.adjoint(s, mod, dt=dt), 2 * T),
        ('filter cos low-pass', lambda: low.forward(s, mod, dt=dt), 2 * T),
        ('filter fir band', lambda: fir.forward(s, mod, dt=dt), 2 * T),
        ('norm_per_shot', lambda: pl.NormPerShot().forward(s, mod), 2 * T),
        ('l2 misfit + residual', lambda: pl.L2DistanceLoss().forward(s, mod, obs), 3 * T),
        ('traces transpose', lambda: pl.traces_from_solver(s, tm), 2 * T),
        ('mask_field (extended)', lambda: pl.MaskField().forward(s, g_ext, inner), 2 * F),
        ('smooth_field sigma 0.25 (extended)', lambda: pl.SmoothField().forward(s, g_ext), 2 * F),
        ('norm_field (extended)', lambda: pl.NormField().forward(s, g_ext), 3 * F),
        ('apply_prec sigma 1 (inner)', lambda: pl.apply_prec(s, g_in, p_in), 3 * I),
        ('descent + clip (inner)', lambda: pl.GradientDescent(min=1400., max=1600.).step(s, vp_in, d_in), 3 * I),
    ]
    out = {}
    for name, fn, nbytes in steps:
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        out[name] = {'ms': round(ms, 3), 'gbs': round(nbytes / ms / 1e6, 1)}
        print('[pipelines] %-40s %9.3f ms %8.1f GB/s' % (name, ms, nbytes / ms / 1e6), file=sys.stderr, flush=True)
    s.close()
    return out


def bench_schedule(tile=None, num_shots=32, timed=False):
    """Batched 3D forward over uneven shots (the multi-shot schedule): the 360^3 grid and
    hemisphere medium, 32 shots whose durations (1500-3000 steps) and receiver counts (256-1024
    hemisphere points) come from the seeded schedule, each with its own time axis. Shots are
    paired longest first; every step is one update launch over all pairs still running plus one
    injection and one interpolation launch over all shots still running, and the whole schedule
    is one HIP graph launch. Counted work: each shot's own forward updates, (nt_i - 2) x grid
    points. The wall time includes capturing and instantiating the graph (reported apart).
    timed: also run the same launches one by one with HIP events around every update launch,
    for the update kernel's time per pair and step."""
    from stride_amd.configs import config3, shot_schedule
    from stride_amd.physics.plan import build_plan
    from stride_amd.problem import Time
    from stride_amd.utils import ellipsoidal
    from stride_amd._native import Solver
    nts, nrecs = shot_schedule(num_shots)
    nt_max = int(nts.max())
    cfg = config3(scale=1.0, nt=nt_max, nrec=1024)
    space, dt = cfg['space'], cfg['time'].step
    c = tuple(l / 2 for l in space.limit)
    plans = []
    for i in range(num_shots):
        rec = ellipsoidal(int(nrecs[i]), (55e-3,) * 3, c, threshold=0.5)
        plans.append(build_plan(space, Time(start=0., step=dt, num=int(nts[i])), cfg['wavelet'][None, :int(nts[i])],
                                cfg['vp'], source_coordinates=rec[(7 * i) % len(rec)][None], receiver_coordinates=rec,
                                kernel='OT2'))
    p0 = plans[0]
    s = Solver(p0.dim, p0.shape, p0.spacing, p0.absorbing, p0.dt, nt_max)
    s.set_medium(p0.vp)
    s.set_sponge(p0.sponge_tables)
    s.set_scheme(p0.kernel_order, p0.rho_mode, p0.att_power, p0.boundary, p0.diff_source, p0.undersampling, 0,
                 0, nt_max, 0, 0, 0)
    order = np.argsort(-nts, kind='stable')
    if tile is None:
        from stride_amd.autotune import autotune_batch
        tile, _ = autotune_batch(plans[int(order[0])], plans[int(order[1])])
    s.multi_create(num_shots)
    for i, p in enumerate(plans):
        s.multi_set_shot(i, p.nt, p.src_idx, p.src_w, p.rec_idx, p.rec_w)
    s.batch_set_tile(*tile)
    d_src = [s._to_device(p.src_data) for p in plans]
    d_rec = [s.device_alloc(p.nt * p.nrec * 4) for p in plans]
    s.synchronize()
    t0 = time.perf_counter()
    s.multi_forward_dev(d_src, d_rec, graph=True)
    s.synchronize()
    wall = time.perf_counter() - t0
    graph_ms = s.multi_graph_ms()
    n = int(np.prod(p0.shape))
    own_steps = int(sum(int(t) - 2 for t in nts))
    sorted_nt = nts[order]
    pair_steps = int(sum(int(sorted_nt[2 * j]) - 2 for j in range((num_shots + 1) // 2)))
    updates = own_steps * n
    out = dict(metric='Gpts/s batched 3D forward, %d uneven shots' % num_shots, value=round(updates / wall / 1e9, 3),
               unit='Gpts/s', seconds=round(wall, 3), graph_capture_ms=round(graph_ms, 1), shots=num_shots,
               pairs=(num_shots + 1) // 2, nt_range=[int(nts.min()), int(nts.max())],
               nrec_range=[int(nrecs.min()), int(nrecs.max())], point_updates=updates,
               padded_steps_fraction=round((2 * pair_steps - own_steps) / float(own_steps), 4), tile=list(tile),
               config={'workload': '3D 320x320x320 inner (360x360x360 extended), dx 0.4 mm, SO10 OT2, sponge 20, '
                                   'forward only; shots paired longest first, one update launch per step over all '
                                   'running pairs, one HIP graph for the whole schedule'})
    if timed:
        s.timing(True)
        s.multi_forward_dev(d_src, d_rec, graph=False)
        s.synchronize()
        kernel_ms, launches = s.timing()
        s.timing(False)
        out['update_kernel_ms'] = round(kernel_ms, 1)
        out['update_us_per_pair_step'] = round(kernel_ms * 1e3 / pair_steps, 2)
        out['update_launches'] = launches
    for d in d_src + d_rec:
        s.device_free(d)
    s.close()
    return out


def unet_flops(N, H, W):
    """Multiply-add flops (x2) of one forward pass of the preconditioner's U-Net; a training step
    (forward, backward data, weight gradient) is three times that."""
    from stride_amd.unet import UNet, LEVEL_CHANNELS
    c1 = LEVEL_CHANNELS[0]
    level = {'e1': 0, 'e2': 1, 'e3': 2, 'b': 3, 'd3': 2, 'd2': 1, 'd1': 0}
    shapes = [('e1a', c1, c1), ('e1', c1, c1), ('e2a', 32, 64), ('e2', 64, 64), ('e3a', 64, 128), ('e3', 128, 128),
              ('ba', 128, 256), ('b', 256, 256), ('d3a', 384, 128), ('d3', 128, 128), ('d2a', 192, 64),
              ('d2', 64, 64), ('d1a', 96, 32), ('d1', 32, 32)]
    total = 0
    for name, cin, cout in shapes:
        lv = level[name.rstrip('a')] if name.rstrip('a') in level else level[name]
        px = N * (H >> lv) * (W >> lv)
        total += 2 * 9 * cin * cout * px
    return total


def bench_learned(ntrain=30, epochs=40, neval=4, iters_per_freq=4, nsrc=16, log=None, save_net=None, load_net=None,
                  eval_start=0):
    """The learned-preconditioner workload on the 2D gradient grid: solver-generated training
    pairs (every gradient of plain inversions of ntrain training phantoms), U-Net training, then
    inversions of unseen phantoms with frequency continuation (300 - 700 kHz): L-BFGS whose first
    step in each band is the network's update when the misfit drops enough (the scaled gradient
    step otherwise) against plain L-BFGS; solver calls counted as passes over all shots (a
    gradient is two: forward and adjoint; a rejected line-search trial one, both methods)."""
    import torch
    from stride_amd import learned
    log = log or (lambda m: print('[learned] ' + m, file=sys.stderr, flush=True))
    freqs = [300e3, 400e3, 500e3, 600e3, 700e3]
    fwi = learned.RingFWI(nsrc=nsrc, workers=nsrc)
    info = {}
    if load_net:
        net, scales = learned.load_network(load_net)
        info.update(network=os.path.basename(load_net))
    else:
        t0 = time.perf_counter()
        pairs = learned.make_trajectory_pairs(fwi, range(ntrain), freqs, iters_per_freq, log=log)
        t_pairs = time.perf_counter() - t0
        scales = learned.input_scales(pairs)
        t0 = time.perf_counter()
        net, losses = learned.train(pairs, scales, epochs=epochs, log=log)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        steps = epochs * (len(pairs) // 4)
        train_flops = 3 * unet_flops(4, 512, 512) * steps
        # inference throughput of the network alone
        g = torch.stack([pairs[j][0] * scales[pairs[j][2]] for j in range(4)])
        m = torch.stack([pairs[j][3] for j in range(4)])
        for _ in range(2):
            net.forward(g, model=m)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            net.forward(g, model=m)
        torch.cuda.synchronize()
        t_inf = (time.perf_counter() - t0) / 10
        info.update(training_phantoms=ntrain, training_pairs=len(pairs), epochs=epochs,
                    pair_seconds=round(t_pairs, 1), train_seconds=round(t_train, 1),
                    train_tflops=round(train_flops / t_train / 1e12, 2), final_train_loss=losses[-1],
                    first_train_loss=losses[0], inference_ms_batch4=round(t_inf * 1e3, 2),
                    inference_tflops=round(unet_flops(4, 512, 512) / t_inf / 1e12, 2))
        del pairs
        if save_net:
            learned.save_network(net, scales, save_net)
    results = []
    for e in range(eval_start, eval_start + neval):
        k = 100000 + e
        true = learned.phantom(k)
        fwi.passes = 0
        vp_p, hp = learned.invert_plain(fwi, true, freqs, iters_per_freq)
        log('phantom %d plain: %d passes, misfit %.4g, model error %.4f' % (k, hp[-1][0], hp[-1][1], hp[-1][2]))
        last = freqs[-1]
        plain_final = hp[-1][1]
        row = dict(phantom=k, plain_passes=hp[-1][0], plain_misfit=plain_final, plain_error=hp[-1][2],
                   plain_history=[[h[0], h[1], h[2]] for h in hp])
        # the learned inversion on the same schedule, and on a schedule of half the iterations per band
        for tag, ipf in (('learned', iters_per_freq), ('learned_half', max(1, iters_per_freq // 2))):
            fwi.passes = 0
            vp_l, hl, acc = learned.invert_learned_lbfgs(fwi, net, scales, true, freqs, ipf)
            log('phantom %d %s (%d per band): %d passes, misfit %.4g, model error %.4f, %d of %d band-first steps '
                'from the network' % (k, tag, ipf, hl[-1][0], hl[-1][1], hl[-1][2], acc, len(freqs)))
            reach = [h[0] for h in hl if h[3] == last and h[1] <= plain_final]
            reach_err = [h[0] for h in hl if h[2] <= hp[-1][2]]
            row.update({tag + '_passes': hl[-1][0], tag + '_misfit': hl[-1][1], tag + '_error': hl[-1][2],
                        tag + '_passes_to_plain_misfit': reach[0] if reach else None,
                        tag + '_passes_to_plain_error': reach_err[0] if reach_err else None,
                        tag + '_network_updates': acc, tag + '_history': [[h[0], h[1], h[2]] for h in hl]})
        results.append(row)
    fwi.close()

    def ratios(key):
        return [r[key] / r['plain_passes'] for r in results if r[key] is not None]

    summary = {}
    for tag in ('learned', 'learned_half'):
        summary[tag] = dict(passes_ratio_to_plain_final_misfit=ratios(tag + '_passes_to_plain_misfit'),
                            passes_ratio_to_plain_final_error=ratios(tag + '_passes_to_plain_error'),
                            final_misfit_below_plain=sum(r[tag + '_misfit'] < r['plain_misfit'] for r in results),
                            final_error_below_plain=sum(r[tag + '_error'] < r['plain_error'] for r in results))
    return dict(metric='FWI misfit drop per solver call, learned preconditioner vs plain L-BFGS',
                unit='passes over all shots', sources=nsrc, **info,
                eval=results, summary=summary, phantoms=neval, first_phantom=100000 + eval_start,
                config={'workload': '2D 512x512 inner (572x572 extended), dx 0.2 mm, CFS-PML 30, vp only, '
                                    '%d ring sources / 128 receivers, 2500 steps, 5 bands 300-700 kHz x %d '
                                    'iterations' % (nsrc, iters_per_freq)})


def cpu_baseline_density(cfg, nt=24):
    """The oracle (C/OpenMP, test infrastructure) on the same medium: a bounded forward with raw
    saves + as-written adjoint with vp and rho gradients, timed on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    from stride_amd.physics.plan import build_plan, adjoint_source_samples
    from stride_amd.problem import Time
    space = cfg['space']
    tm = Time(start=0., step=cfg['time'].step, num=nt)
    plan = build_plan(space, tm, cfg['wavelet'][None, :nt], cfg['vp'], rho=cfg['rho'], alpha=cfg['alpha'],
                      source_coordinates=cfg['sources'], receiver_coordinates=cfg['receivers'], kernel='OT2',
                      save_undersampling=4, grad_vp=True, grad_rho=True)
    res = adjoint_source_samples(plan, np.random.default_rng(0).standard_normal((plan.nrec, nt)))
    t0 = time.perf_counter()
    _, saved, _ = oracle.forward(plan, save=True)
    oracle.adjoint_compat(plan, res, saved)
    t1 = time.perf_counter()
    n = int(np.prod(plan.shape))
    fwd, adj = nt - 2, nt - 1 - plan.undersampling
    threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    return dict(value=round(n * (fwd + adj) / (t1 - t0) / 1e9, 4), unit='Gpts/s', cores=threads, kind='port',
                sample='oracle (C/OpenMP) on the same 360^3 medium: %d forward steps with raw saves + %d as-written '
                       'adjoint steps with vp and rho gradients, %.1f s' % (fwd, adj, t1 - t0))


def bench_density(nt=2500, tile=None, with_cpu=True, checkpoint=0, att_power=0, kernel='OT2'):
    """One 3D shot on the variable vp / rho / alpha medium (360^3 extended, attenuation power 0,
    sponge 20, OT2, 1024 receivers): forward with the raw wavefield saved every 4th step over the
    whole domain, then the as-written adjoint with the vp and rho gradients. The update runs as the
    fused density kernel; `roofline` covers it (algorithmic bytes: read u, the other level, b, A, B,
    G, write the new level = 28 B per point update, + 4 B per point on save steps; att_power=2 adds
    Cbeta and the Laplacians of both levels: 40 B, and the step's time then includes the launch that
    forms the Laplacian of u). checkpoint=K:
    the raw snapshots recomputed from one state per K snapshots instead (save mode 5); the rate
    then counts the shot's own updates only, and the recomputation's launches are extra work."""
    from stride_amd.configs import config4
    from stride_amd.physics.plan import build_plan
    from stride_amd._native import Solver
    cfg = config4(scale=1.0, nt=nt, nrec=1024)
    space, time_ = cfg['space'], cfg['time']
    plan = build_plan(space, time_, cfg['wavelet'][None], cfg['vp'], rho=cfg['rho'], alpha=cfg['alpha'],
                      attenuation_power=att_power, source_coordinates=cfg['sources'],
                      receiver_coordinates=cfg['receivers'], kernel=kernel, save_undersampling=4, grad_vp=True,
                      grad_rho=True)
    s = Solver(plan.dim, plan.shape, plan.spacing, plan.absorbing, plan.dt, plan.nt)
    s.set_scheme(plan.kernel_order, plan.rho_mode, plan.att_power, plan.boundary, plan.diff_source,
                 plan.undersampling, 5 if checkpoint else 2, plan.tb0, plan.tb1, 1, 1, 0, checkpoint=checkpoint)
    s.set_medium(plan.vp, plan.rho, plan.alpha)
    s.set_sponge(plan.sponge_tables)
    s.set_sources(plan.src_idx, plan.src_w)
    s.set_receivers(plan.rec_idx, plan.rec_w)
    if tile:
        s.set_tile_density(*tile)
    rng = np.random.default_rng(0)
    d_src = s._to_device(plan.src_data)
    d_rec = s.device_alloc(plan.nt * plan.nrec * 4)
    d_res = s._to_device(rng.standard_normal((plan.nt, plan.nrec)).astype(np.float32) * 1e-3)
    want = 3  # vp and rho gradients

    def shot():
        s.forward_dev(d_src, d_rec)
        s.adjoint_dev(d_res, want)
        s.synchronize()

    shot()  # allocations (saved wavefield, gradient accumulators)
    s.timing(True)
    t0 = time.perf_counter()
    shot()
    wall = time.perf_counter() - t0
    ms, launches = s.timing()
    s.timing(False)
    n = int(np.prod(plan.shape))
    f = plan.undersampling
    nsave = len(range(0, plan.nt - 1, f))
    # power 2: + Cbeta and the two Laplacians read (the Laplacian of u is its own launch)
    # OT4: two passes per step (u, b, m in, K u and m K u out = 20 B; then m K u, b, the other
    # level, A, B, G, u, K u in, the new level out = 36 B)
    per_update = 40.0 if att_power == 2 else (56.0 if kernel == 'OT4' else 28.0)
    alg = launches * per_update * n + (0 if checkpoint else nsave * 4.0 * n)
    achieved = alg / (ms * 1e-3) / 1e9
    own = (plan.tb1 - 2) + (plan.tb1 - 1 - f - plan.tb0)  # the shot's forward and adjoint steps
    per_gb = plan.shape[0] * plan.shape[1] * (-(-plan.shape[2] // 32) * 32) * 4 / 1e9  # one snapshot
    nsnap = plan.tb1 // f + 1
    if checkpoint:
        last = (plan.tb1 - 2) // f
        k = min(checkpoint, last + 1)
        stored = 2 * (last // k) + k + 2  # segment states + the recomputation ring
    else:
        stored = nsnap
    for d in (d_src, d_rec, d_res):
        s.device_free(d)
    s.close()
    cpu = cpu_baseline_density(cfg) if with_cpu else None
    return dict(metric='Gpts/s fwd+adjoint 3D 320^3 SO10, variable vp/rho/alpha, vp+rho gradients',
                value=round(own * n / wall / 1e9, 3), unit='Gpts/s', seconds=round(wall, 3),
                snapshots='recomputed, segments of %d' % checkpoint if checkpoint else 'stored',
                snapshot_store_gb=round(stored * per_gb, 1),
                update_launches=int(launches), update_kernel_us=round(ms * 1e3 / launches, 1),
                update_kernel_share=round(ms * 1e-3 / wall, 3),
                roofline={'bound': 'hbm', 'kernel': 'k_rho3d', 'achieved': round(achieved, 1),
                          'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4)},
                cpu_baseline=cpu, dtype='f32', data='synthetic',
                config={'workload': '3D 320x320x320 inner (360x360x360 extended), dx 0.4 mm, SO10 OT2, sponge 20, '
                                    'variable vp / rho (Gaussian-filtered noise in the hemisphere) and alpha 0.5 '
                                    '(power %d), one shot, nt %d, raw saves every 4th step, 1024 receivers, %s'
                                    % (att_power, plan.nt, kernel)})


def bench_density_pair(nt=2500, tile=None):
    """Two 3D shots on the density workload's medium advanced together (interleaved wavefields,
    one read of b and the coefficients per pass for both): forward with the raw wavefields saved
    every 4th step, then the as-written adjoint with the summed vp and rho gradients. The second
    shot is sourced at the receiver farthest from the first source. The rate counts each shot's own
    updates; `roofline` covers the two-shot update (algorithmic bytes per cell and pass: read u and
    the other level of both shots, b, A, B, G, write both new levels = 40 B, i.e. 20 B per point
    update; + 8 B per cell on save steps). The gradient pass runs once per shot on guarded steps."""
    from stride_amd.configs import config4
    from stride_amd.physics.plan import build_plan
    from stride_amd._native import Solver
    cfg = config4(scale=1.0, nt=nt, nrec=1024)
    space, time_ = cfg['space'], cfg['time']
    src0 = cfg['sources']
    far = cfg['receivers'][np.argmax(np.linalg.norm(cfg['receivers'] - src0[0], axis=1))][None]
    plans = [build_plan(space, time_, cfg['wavelet'][None], cfg['vp'], rho=cfg['rho'], alpha=cfg['alpha'],
                        attenuation_power=0, source_coordinates=sc, receiver_coordinates=cfg['receivers'],
                        kernel='OT2', save_undersampling=4, grad_vp=True, grad_rho=True) for sc in (src0, far)]
    plan = plans[0]
    s = Solver(plan.dim, plan.shape, plan.spacing, plan.absorbing, plan.dt, plan.nt)
    s.set_scheme(plan.kernel_order, plan.rho_mode, plan.att_power, plan.boundary, plan.diff_source,
                 plan.undersampling, 2, plan.tb0, plan.tb1, 1, 1, 0)
    s.set_medium(plan.vp, plan.rho, plan.alpha)
    s.set_sponge(plan.sponge_tables)
    s.batch_create()
    for k, p in enumerate(plans):
        s.batch_set_taps(k, p.src_idx, p.src_w, p.rec_idx, p.rec_w)
    if tile:
        s.batch_set_tile_density(*tile)
    rng = np.random.default_rng(0)
    d_src = [s._to_device(p.src_data) for p in plans]
    d_rec = [s.device_alloc(p.nt * p.nrec * 4) for p in plans]
    d_res = [s._to_device(rng.standard_normal((p.nt, p.nrec)).astype(np.float32) * 1e-3) for p in plans]

    def pair():
        s.batch_forward_dev(d_src, d_rec)
        s.batch_adjoint_dev(d_res)
        s.synchronize()

    pair()  # allocations (saved wavefields, gradient accumulators)
    s.timing(True)
    t0 = time.perf_counter()
    pair()
    wall = time.perf_counter() - t0
    ms, launches = s.timing()
    s.timing(False)
    n = int(np.prod(plan.shape))
    f = plan.undersampling
    nsave = len(range(0, plan.nt - 1, f))
    alg = launches * 40.0 * n + nsave * 8.0 * n
    achieved = alg / (ms * 1e-3) / 1e9
    own = 2 * ((plan.tb1 - 2) + (plan.tb1 - 1 - f - plan.tb0))
    per_gb = plan.shape[0] * plan.shape[1] * (-(-plan.shape[2] // 32) * 32) * 4 / 1e9
    for d in d_src + d_rec + d_res:
        s.device_free(d)
    s.close()
    return dict(metric='Gpts/s fwd+adjoint 3D 320^3 SO10, variable vp/rho/alpha, vp+rho gradients, two shots per pass',
                value=round(own * n / wall / 1e9, 3), unit='Gpts/s', seconds=round(wall, 3), shots=2,
                snapshot_store_gb=round(2 * (plan.tb1 // f + 1) * per_gb, 1),
                update_launches=int(launches), update_kernel_us=round(ms * 1e3 / launches, 1),
                update_kernel_share=round(ms * 1e-3 / wall, 3),
                roofline={'bound': 'hbm', 'kernel': 'k_rho3d (two shots)', 'achieved': round(achieved, 1),
                          'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4)},
                config={'workload': '3D 320x320x320 inner (360x360x360 extended), dx 0.4 mm, SO10 OT2, sponge 20, '
                                    'variable vp / rho and alpha 0.5 (power 0), two shots per pass (interleaved '
                                    'wavefields), nt %d, raw saves every 4th step, 1024 receivers per shot'
                                    % plan.nt})


def bench_pml3d(nt=500, general=False):
    """One 3D shot with the CFS-PML boundary (20 cells, the reference's complex_frequency_shift_PML_2)
    on the 360^3 grid and medium of the headline workload: forward with p.dt2 saved every 4th step,
    then the as-written adjoint with the vp gradient. The update runs as the specialised one-shot
    kernel without a boundary term, the auxiliary fields of every axis are advanced by one launch
    over the absorbing layers before it and their term is applied by one after it; general=True
    runs the separate general passes (the Laplacian, k_pml per axis, k_update) instead."""
    from stride_amd.configs import config3
    from stride_amd.physics.plan import build_plan
    from stride_amd._native import Solver
    if general:
        os.environ['STRIDE_AMD_GENERAL_ONLY'] = '1'
    try:
        cfg = config3(scale=1.0, nt=nt, nrec=1024)
        plan = build_plan(cfg['space'], cfg['time'], cfg['wavelet'][None], cfg['vp'],
                          boundary_type='complex_frequency_shift_PML_2', source_coordinates=cfg['sources'],
                          receiver_coordinates=cfg['receivers'], kernel='OT2', save_undersampling=4, grad_vp=True)
        s = Solver(plan.dim, plan.shape, plan.spacing, plan.absorbing, plan.dt, plan.nt)
        s.set_scheme(plan.kernel_order, plan.rho_mode, plan.att_power, plan.boundary, plan.diff_source,
                     plan.undersampling, 1, plan.tb0, plan.tb1, 1, 0, 0)
        s.set_medium(plan.vp)
        s.set_pml(plan.pml)
        s.set_sources(plan.src_idx, plan.src_w)
        s.set_receivers(plan.rec_idx, plan.rec_w)
        d_src = s._to_device(plan.src_data)
        d_rec = s.device_alloc(plan.nt * plan.nrec * 4)
        d_res = s._to_device(np.random.default_rng(0).standard_normal((plan.nt, plan.nrec)).astype(np.float32) * 1e-3)

        def shot():
            s.forward_dev(d_src, d_rec)
            s.adjoint_dev(d_res, 1)
            s.synchronize()

        shot()
        t0 = time.perf_counter()
        shot()
        wall = time.perf_counter() - t0
        for d in (d_src, d_rec, d_res):
            s.device_free(d)
        s.close()
    finally:
        if general:
            os.environ.pop('STRIDE_AMD_GENERAL_ONLY', None)
    n = int(np.prod(plan.shape))
    fwd, adj = plan.tb1 - 2, plan.tb1 - 1 - plan.undersampling - plan.tb0
    return dict(metric='Gpts/s fwd+adjoint 3D 320^3 SO10, CFS-PML 20, vp gradient',
                value=round((fwd + adj) * n / wall / 1e9, 3), unit='Gpts/s', seconds=round(wall, 3), nt=plan.nt,
                path='general passes' if general else 'specialised update + PML layer passes')


def bench_ot4_3d(nt=500, general=False):
    """One 3D shot with fourth-order time stepping (kernel='OT4') on the 360^3 grid, medium and
    sponge of the headline workload: forward with p.dt2 saved every 4th step, then the as-written
    adjoint with the vp gradient. The fused path runs two marches per step (K u and w = vp^2 K u;
    then the update staging w, so its Laplacian is K(m K u)); general=True runs the separate general
    passes (the Laplacian, K of the scaled K u, the product, k_update) instead."""
    from stride_amd.configs import config3
    from stride_amd.physics.plan import build_plan
    if general:
        os.environ['STRIDE_AMD_GENERAL_ONLY'] = '1'
    try:
        cfg = config3(scale=1.0, nt=nt, nrec=1024)
        plan = build_plan(cfg['space'], cfg['time'], cfg['wavelet'][None], cfg['vp'],
                          source_coordinates=cfg['sources'], receiver_coordinates=cfg['receivers'], kernel='OT4',
                          save_undersampling=4, grad_vp=True)
        s = make_solver(plan)
        d_src = s._to_device(plan.src_data)
        d_rec = s.device_alloc(plan.nt * plan.nrec * 4)
        d_res = s._to_device(np.random.default_rng(0).standard_normal((plan.nt, plan.nrec)).astype(np.float32) * 1e-3)

        def shot():
            s.forward_dev(d_src, d_rec)
            s.adjoint_dev(d_res, 1)
            s.synchronize()

        shot()
        t0 = time.perf_counter()
        shot()
        wall = time.perf_counter() - t0
        for d in (d_src, d_rec, d_res):
            s.device_free(d)
        s.close()
    finally:
        if general:
            os.environ.pop('STRIDE_AMD_GENERAL_ONLY', None)
    n = int(np.prod(plan.shape))
    fwd, adj = plan.tb1 - 2, plan.tb1 - 1 - plan.undersampling - plan.tb0
    return dict(metric='Gpts/s fwd+adjoint 3D 320^3 SO10 OT4, vp gradient',
                value=round((fwd + adj) * n / wall / 1e9, 3), unit='Gpts/s', seconds=round(wall, 3), nt=plan.nt,
                kernel_order=plan.kernel_order, path='general passes' if general else 'two marches per step')


def bench_power2(nt=500, general=False):
    """One 3D shot with power-2 attenuation (alpha 0.5 dB/cm/MHz, the 2D gradient workload's
    value) on the 360^3 grid of the headline workload: forward with p.dt2 saved every 4th step,
    then the as-written adjoint with the vp gradient. The update runs as the specialised one-shot
    kernel with the attenuation term fused (algorithmic bytes: read u, the other level, vp, alpha
    and the other level's Laplacian, write the new level and the Laplacian of u = 28 B per point
    update; + 4 B per interior point on save steps, + 20 B on gradient steps); general=True runs
    the separate general passes instead, for comparison."""
    from stride_amd.configs import config3
    from stride_amd.physics.plan import build_plan
    from stride_amd.autotune import tune_solver
    if general:
        os.environ['STRIDE_AMD_GENERAL_ONLY'] = '1'
    try:
        cfg = config3(scale=1.0, nt=nt, nrec=1024)
        alpha = np.full(cfg['space'].extended_shape, 0.5, np.float32)
        plan = build_plan(cfg['space'], cfg['time'], cfg['wavelet'][None], cfg['vp'], alpha=alpha,
                          attenuation_power=2, source_coordinates=cfg['sources'],
                          receiver_coordinates=cfg['receivers'], kernel='OT2', save_undersampling=4, grad_vp=True)
        s = make_solver(plan)
        s.set_medium(plan.vp, None, plan.alpha)
        if not general:
            tune_solver(s, plan, 1)
        d_src = s._to_device(plan.src_data)
        d_rec = s.device_alloc(plan.nt * plan.nrec * 4)
        d_res = s._to_device(np.random.default_rng(0).standard_normal((plan.nt, plan.nrec)).astype(np.float32) * 1e-3)

        def shot():
            s.forward_dev(d_src, d_rec)
            s.adjoint_dev(d_res, 1)
            s.synchronize()

        shot()
        s.timing(True)
        t0 = time.perf_counter()
        shot()
        wall = time.perf_counter() - t0
        ms, launches = s.timing()
        s.timing(False)
        for d in (d_src, d_rec, d_res):
            s.device_free(d)
        s.close()
    finally:
        if general:
            os.environ.pop('STRIDE_AMD_GENERAL_ONLY', None)
    n = int(np.prod(plan.shape))
    interior = int(np.prod([a - 2 * b for a, b in zip(plan.shape, plan.absorbing)]))
    f = plan.undersampling
    fwd, adj = plan.tb1 - 2, plan.tb1 - 1 - f - plan.tb0
    saves = sum(1 for t in range(1, plan.tb1 - 1) if t % f == 0)
    grads = sum(1 for t in range(plan.tb1 - 1 - f, plan.tb0, -1) if t % f == 0)
    out = dict(metric='Gpts/s fwd+adjoint 3D 320^3 SO10, power-2 attenuation, vp gradient',
               value=round((fwd + adj) * n / wall / 1e9, 3), unit='Gpts/s', seconds=round(wall, 3), nt=plan.nt,
               path='general passes' if general else 'specialised update (attenuation fused)',
               kernel_launches=int(launches), kernel_ms=round(ms, 2))
    if not general:
        alg = 28.0 * n * (fwd + adj) + 4.0 * interior * saves + 20.0 * interior * grads
        achieved = alg / (ms * 1e-3) / 1e9
        out['roofline'] = {'bound': 'hbm', 'kernel': 'k_fast3d_v7 (power 2)', 'achieved': round(achieved, 1),
                           'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4),
                           'avg_launch_us': round(ms * 1e3 / launches, 1)}
    return out


def bench_autograd(nt=2500, checkpoint=0):
    """One shot of the 3D gradient workload through the torch.autograd.Function (exact adjoint):
    variable vp / rho / alpha (attenuation power 0) on the 360^3 grid, device tensors, the forward
    keeping one pair of levels every K steps (K = ceil(sqrt(2 nt)) unless given) and the backward
    recomputing each K-step segment; dJ/dvp, dJ/drho, dJ/dalpha of J = 1/2 |traces - obs|^2. The
    rate counts the shot's forward and adjoint updates (the recomputation is extra work)."""
    import torch
    from stride_amd.autograd import Geometry, iso_acoustic_traces
    from stride_amd.configs import config4
    cfg = config4(scale=1.0, nt=nt, nrec=1024)
    space, time_ = cfg['space'], cfg['time']
    K = checkpoint or int(np.ceil(np.sqrt(2 * nt)))
    geom = Geometry(space, time_, cfg['sources'], cfg['receivers'], checkpoint_every=K, kernel='OT2')
    dev = torch.device('cuda')
    vp = torch.tensor(np.asarray(cfg['vp'], np.float32), device=dev, requires_grad=True)
    rho = torch.tensor(np.asarray(cfg['rho'], np.float32), device=dev, requires_grad=True)
    alpha = torch.tensor(np.asarray(cfg['alpha'], np.float32), device=dev, requires_grad=True)
    wav = torch.tensor(cfg['wavelet'][None].astype(np.float32), device=dev)
    obs = torch.zeros((cfg['receivers'].shape[0], nt), device=dev)

    def shot():
        for t in (vp, rho, alpha):
            t.grad = None
        tr = iso_acoustic_traces(vp, rho, alpha, wav, geom)
        loss = 0.5 * torch.sum((tr - obs) ** 2)
        loss.backward()
        torch.cuda.synchronize()
        return float(loss.detach())

    t0 = time.perf_counter()
    shot()  # plan, allocations
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = shot()
    wall = time.perf_counter() - t0
    n = int(np.prod(space.extended_shape))
    ok = all(torch.isfinite(t.grad).all().item() and t.grad.abs().max().item() > 0 for t in (vp, rho, alpha))
    shp = space.extended_shape
    box_gb = shp[0] * shp[1] * (-(-shp[2] // 32) * 32) * 4 / 1e9   # a stored level (box layout)
    pad_gb = (shp[0] + 20) * (shp[1] + 20) * (-(-(shp[2] + 42) // 32) * 32) * 4 / 1e9  # padded layout
    return dict(metric='Gpts/s fwd+adjoint 3D 320^3 SO10, exact adjoint through torch.autograd, '
                       'vp/rho/alpha gradients',
                value=round(2 * (nt - 1) * n / wall / 1e9, 3), unit='Gpts/s', seconds=round(wall, 3),
                first_call_seconds=round(first, 3), checkpoint_every=K,
                # checkpoint pairs + the segment's levels and their K u (padded)
                wavefield_store_gb=round(2 * (nt // K + 1) * box_gb + 2 * (K + 1) * pad_gb, 1), loss=loss,
                gradients_finite_nonzero=bool(ok), dtype='f32', data='synthetic',
                config={'workload': '3D 320x320x320 inner (360x360x360 extended), dx 0.4 mm, SO10 OT2, sponge 20, '
                                    'variable vp / rho / alpha (power 0), one shot, nt %d, 1024 receivers, '
                                    'checkpoints every %d steps' % (nt, K)})


def cpu_baseline_2d(cfg, nt=500):
    """The oracle (C/OpenMP, test infrastructure) on the 2D gradient workload's medium and scheme: a
    bounded forward with raw saves + as-written adjoint with vp and rho gradients."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    from stride_amd.physics.plan import build_plan, adjoint_source_samples
    from stride_amd.problem import Time
    tm = Time(start=0., step=cfg['time'].step, num=nt)
    plan = build_plan(cfg['space'], tm, cfg['wavelet'][None, :nt], cfg['vp'], rho=cfg['rho'], alpha=cfg['alpha'],
                      source_coordinates=cfg['sources'][:1], receiver_coordinates=cfg['receivers'],
                      boundary_type='complex_frequency_shift_PML_2', attenuation_power=2, grad_vp=True,
                      grad_rho=True)
    res = adjoint_source_samples(plan, np.random.default_rng(0).standard_normal((plan.nrec, nt)))
    t0 = time.perf_counter()
    _, saved, _ = oracle.forward(plan, save=True)
    oracle.adjoint_compat(plan, res, saved)
    t1 = time.perf_counter()
    n = int(np.prod(plan.shape))
    fwd, adj = nt - 2, nt - 1 - plan.undersampling
    threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    return dict(value=round(n * (fwd + adj) / (t1 - t0) / 1e9, 4), unit='Gpts/s', cores=threads, kind='port',
                sample='oracle (C/OpenMP), one shot of %d steps: forward with raw saves + as-written adjoint '
                       'with vp and rho gradients, %.1f s' % (nt, t1 - t0))


def bench_gradient_2d(shots=8, workers=2, with_cpu=True):
    """The 2D gradient workload: 512^2 inner (572^2 extended), dx 0.2 mm, CFS-PML 30, variable vp /
    rho, alpha 0.5 with power 2, 128 ring receivers, nt 2500. Each shot: forward with the raw
    wavefield saved every f-th step, then the as-written adjoint with the vp and rho gradients
    (sources at ring elements 0, 16, 32, ...). The update runs as the fused 2D kernel (injection,
    interpolation, the power-2 Laplacian and the gradients folded in). `value`: the ``shots`` shots
    advanced in lockstep, one solver per shot and one launch per time step for all of them; the
    line also gives the rate of the same shots on ``workers`` solvers, each on its own stream and
    host thread (as that many Mosaic workers sharing the GPU would), and on one worker."""
    from stride_amd.configs import config2
    from stride_amd.workers import ShotWorkers
    from stride_amd.physics.plan import build_plan
    from stride_amd._native import Solver
    cfg = config2()
    plans = [build_plan(cfg['space'], cfg['time'], cfg['wavelet'][None], cfg['vp'], rho=cfg['rho'],
                        alpha=cfg['alpha'], source_coordinates=cfg['sources'][16 * k:16 * k + 1],
                        receiver_coordinates=cfg['receivers'], boundary_type='complex_frequency_shift_PML_2',
                        attenuation_power=2, grad_vp=True, grad_rho=True) for k in range(shots)]
    p = plans[0]

    def make_solver():
        sv = Solver(p.dim, p.shape, p.spacing, p.absorbing, p.dt, p.nt)
        sv.set_scheme(p.kernel_order, p.rho_mode, p.att_power, p.boundary, p.diff_source, p.undersampling, 2,
                      p.tb0, p.tb1, 1, 1, 0)
        sv.set_medium(p.vp, p.rho, p.alpha)
        sv.set_pml(p.pml)
        sv.set_receivers(p.rec_idx, p.rec_w)
        return sv

    res = np.random.default_rng(0).standard_normal((p.nt, p.nrec)).astype(np.float32) * 1e-3
    buffers = {}

    def stage(sv):
        buffers[id(sv)] = (sv._to_device(res), sv.device_alloc(p.nt * p.nrec * 4),
                           [sv._to_device(q.src_data) for q in plans])

    def release(sv):
        d_res, d_rec, d_src = buffers.pop(id(sv))
        for d in d_src + [d_rec, d_res]:
            sv.device_free(d)

    def shot(sv, k):
        d_res, d_rec, d_src = buffers[id(sv)]
        sv.set_sources(plans[k].src_idx, plans[k].src_w)
        sv.forward_dev(d_src[k], d_rec)
        sv.adjoint_dev(d_res, 3)

    # one worker on the null stream
    s = make_solver()
    stage(s)
    shot(s, 0)
    s.synchronize()
    t0 = time.perf_counter()
    for k in range(shots):
        shot(s, k)
    s.synchronize()
    wall1 = time.perf_counter() - t0
    # kernel times from a separate pass: events around every launch slow these short steps
    s.timing(True)
    t1 = time.perf_counter()
    for k in range(shots):
        shot(s, k)
    s.synchronize()
    wall_timed = time.perf_counter() - t1
    ms, launches = s.timing()
    s.timing(False)
    release(s)
    s.close()
    # the same shots on `workers` solvers with their own streams and host threads
    pool = ShotWorkers(make_solver, n=workers)
    for sv in pool.solvers:
        stage(sv)
    pool.map(shot, range(len(pool)))
    t2 = time.perf_counter()
    pool.map(shot, range(shots))
    wall = time.perf_counter() - t2
    for sv in pool.solvers:
        release(sv)
    pool.close()
    # the same shots in lockstep: one solver per shot, one launch per time step for all of them
    from stride_amd._native import shots2d_forward_dev, shots2d_adjoint_dev
    group = []
    for k in range(shots):
        sv = make_solver()
        sv.set_sources(plans[k].src_idx, plans[k].src_w)
        group.append(sv)
    gbuf = [(sv._to_device(plans[k].src_data), sv.device_alloc(p.nt * p.nrec * 4), sv._to_device(res))
            for k, sv in enumerate(group)]

    def run_group():
        shots2d_forward_dev(group, [b[0] for b in gbuf], [b[1] for b in gbuf])
        shots2d_adjoint_dev(group, [b[2] for b in gbuf], 3)
        group[0].synchronize()

    run_group()  # allocations
    t3 = time.perf_counter()
    run_group()
    wall_ls = time.perf_counter() - t3
    group[0].timing(True)
    run_group()
    ms_ls, launches_ls = group[0].timing()
    group[0].timing(False)
    for sv, b in zip(group, gbuf):
        for d in b:
            sv.device_free(d)
        sv.close()
    n = int(np.prod(p.shape))
    cpu = cpu_baseline_2d(cfg) if with_cpu else None
    return dict(metric='Gpts/s fwd+adjoint 2D 512^2 SO10, CFS-PML, variable vp/rho, power-2 attenuation, '
                       'vp+rho gradients', value=round(launches * n / wall_ls / 1e9, 3), unit='Gpts/s',
                seconds_per_shot=round(wall_ls / shots, 4),
                lockstep={'shots_per_launch': shots, 'launches': int(launches_ls),
                          'launch_us': round(ms_ls * 1e3 / max(launches_ls, 1), 2)},
                value_workers=round(launches * n / wall / 1e9, 3), seconds_per_shot_workers=round(wall / shots, 4),
                workers=workers,
                value_one_worker=round(launches * n / wall1 / 1e9, 3),
                seconds_per_shot_one_worker=round(wall1 / shots, 4), shots=shots, update_launches=int(launches),
                update_kernel_us=round(ms * 1e3 / launches, 2),
                update_kernel_share_one_worker=round(ms * 1e-3 / wall_timed, 3),
                cpu_baseline=cpu, dtype='f32', data='synthetic',
                config={'workload': '2D 512x512 inner (572x572 extended), dx 0.2 mm, SO10 OT2, CFS-PML 30, '
                                    'vp 1500 + 150 N1 in a 45 mm disc, rho 1000 + 50 N2, alpha 0.5 (power 2), '
                                    '128-element 48 mm ring, nt %d, %d shots advanced in lockstep (one launch '
                                    'per time step for all shots); also on %d workers (solvers with their own '
                                    'streams and host threads) and on one' % (p.nt, shots, workers)})


def bench_forward_2d(reps=20):
    """The reference's own CPU-runnable case: 2D 200^2 inner (240^2 extended), dx 0.5 mm, sponge 20,
    Gaussian inclusion, one source on the 128-element 40 mm ring, 1000 steps (OT2), traces
    (128, 1000). The GPU run (fused 2D update, one launch per step) is timed over ``reps`` shots;
    the oracle (C/OpenMP, test infrastructure) runs the same whole shot once on the host cores, and
    its traces are compared."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    from stride_amd.configs import config1
    from stride_amd.physics.plan import build_plan
    from stride_amd._native import Solver
    cfg = config1()
    p = build_plan(cfg['space'], cfg['time'], cfg['wavelet'][None], cfg['vp'], source_coordinates=cfg['sources'],
                   receiver_coordinates=cfg['receivers'])
    s = Solver(p.dim, p.shape, p.spacing, p.absorbing, p.dt, p.nt)
    s.set_scheme(p.kernel_order, p.rho_mode, p.att_power, p.boundary, p.diff_source, p.undersampling, 0,
                 p.tb0, p.tb1, 0, 0, 0)
    s.set_medium(p.vp)
    s.set_sponge(p.sponge_tables)
    s.set_sources(p.src_idx, p.src_w)
    s.set_receivers(p.rec_idx, p.rec_w)
    d_src = s._to_device(p.src_data)
    d_rec = s.device_alloc(p.nt * p.nrec * 4)
    s.forward_dev(d_src, d_rec)
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        s.forward_dev(d_src, d_rec)
    s.synchronize()
    wall = (time.perf_counter() - t0) / reps
    rec = np.zeros((p.nt, p.nrec), np.float32)
    s.memcpy(rec.ctypes.data, d_rec, rec.nbytes, 2)
    s.device_free(d_src)
    s.device_free(d_rec)
    s.close()
    t1 = time.perf_counter()
    ref, _, _ = oracle.forward(p)
    cpu = time.perf_counter() - t1
    n = int(np.prod(p.shape))
    steps = p.nt - 2
    err = float(np.linalg.norm(rec.T - ref) / np.linalg.norm(ref))
    threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    return dict(metric='Gpts/s 2D 200^2 SO10 forward (the reference\'s CPU-runnable case)',
                value=round(n * steps / wall / 1e9, 3), unit='Gpts/s', seconds_per_shot=round(wall, 5),
                us_per_step=round(wall / steps * 1e6, 2), traces_rel_l2_vs_oracle=err,
                cpu_baseline={'value': round(n * steps / cpu / 1e9, 4), 'unit': 'Gpts/s', 'cores': threads,
                              'kind': 'port', 'sample': 'oracle (C/OpenMP), the whole shot, %.2f s' % cpu},
                dtype='f32', data='synthetic',
                config={'workload': '2D 200x200 inner (240x240 extended), dx 0.5 mm, SO10 OT2, sponge 20, '
                                    'Gaussian inclusion, 1 source, 128 receivers, nt %d' % p.nt})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--scale', type=float, default=1.0, help='grid scale (1.0 = 320^3 inner)')
    ap.add_argument('--nt', type=int, default=2500)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-schedule', action='store_true',
                    help='leave out the secondary 32-shot batched-forward line')
    ap.add_argument('--no-density', action='store_true',
                    help='leave out the secondary 3D variable-density gradient shot')
    ap.add_argument('--no-gradient-2d', action='store_true',
                    help='leave out the secondary 2D gradient workload (8 shots in lockstep)')
    ap.add_argument('--tile', type=str, default=None,
                    help='rows,chunk,version of the 3D update tile (rows,chunk,waves with --shots 2)')
    ap.add_argument('--shots', type=int, default=2, choices=(1, 2),
                    help='shots advanced together per step (2: interleaved wavefields, shared vp reads)')
    ap.add_argument('--no-autotune', action='store_true')
    ap.add_argument('--copy-check', type=int, default=0, metavar='MB',
                    help='only run 10 device-to-device copies of MB megabytes (calibrates the byte counters '
                         'of a profiler against a transfer of known size) and exit')
    ap.add_argument('--schedule', action='store_true',
                    help='only run the batched 3D forward over 32 shots of uneven duration and receiver count '
                         '(seeded schedule) and exit')
    ap.add_argument('--checkpoint', type=int, default=0,
                    help='with --density: recompute the raw snapshots from one state per this many snapshots')
    ap.add_argument('--autograd', action='store_true',
                    help='one 3D gradient-workload shot through the autograd exact adjoint with checkpoints '
                         '(--checkpoint K, default ceil(sqrt(2 nt))) and exit')
    ap.add_argument('--att-power', type=int, default=0, choices=(0, 2),
                    help='with --density: attenuation power of the medium (default 0)')
    ap.add_argument('--power2', action='store_true',
                    help='one 3D shot with power-2 attenuation on the fused update and on the general passes '
                         '(use with --nt) and exit')
    ap.add_argument('--ot4', action='store_true',
                    help='one 3D OT4 shot on the fused marches and on the general passes (use --nt 500); with '
                         '--density, the density shot with OT4')
    ap.add_argument('--pml', action='store_true',
                    help='one 3D shot with the CFS-PML boundary on the fused update and on the general passes '
                         '(use --nt 500)')
    ap.add_argument('--density-pair', action='store_true',
                    help='two 3D density shots per pass (interleaved), fwd + adjoint with summed vp / rho '
                         'gradients (--tile rows,chunk,waves)')
    ap.add_argument('--density', action='store_true',
                    help='only run one 3D shot on the variable vp / rho / alpha medium (fused density update, '
                         'vp and rho gradients) and exit')
    ap.add_argument('--gradient-2d', action='store_true',
                    help='only run shots of the 2D gradient workload (572^2 extended, CFS-PML, variable density, '
                         'power-2 attenuation; vp and rho gradients) and exit')
    ap.add_argument('--forward-2d', action='store_true',
                    help="only run the reference's CPU-runnable case (2D 240^2 forward, 1000 steps) on the GPU and "
                         "the CPU oracle and exit")
    ap.add_argument('--workers', type=int, default=2,
                    help='with --gradient-2d: solvers running shots concurrently, each on its own stream')
    ap.add_argument('--learned', action='store_true',
                    help='only run the learned-preconditioner workload (training pairs from the solver, U-Net '
                         'training, misfit-gated inversions against plain L-BFGS) and exit')
    ap.add_argument('--train-phantoms', type=int, default=30)
    ap.add_argument('--epochs', type=int, default=40)
    ap.add_argument('--eval', type=int, default=4)
    ap.add_argument('--eval-start', type=int, default=0, help='with --learned: first unseen phantom (100000 + this)')
    ap.add_argument('--save-net', type=str, default=None, help='with --learned: write the trained network here')
    ap.add_argument('--load-net', type=str, default=None,
                    help='with --learned: evaluate this saved network instead of generating pairs and training')
    ap.add_argument('--pipelines', action='store_true',
                    help='only time the device steps either side of the solver (trace processing, misfit, '
                         'gradient post-processing, model update) at the full workload and exit')
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py measures one GPU: shots shard across GPUs as independent solvers, one process per '
              'GPU, and that multi-GPU run is not implemented here; run with --gpus 1.', file=sys.stderr)
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

    if args.learned:
        print(json.dumps(bench_learned(ntrain=args.train_phantoms, epochs=args.epochs, neval=args.eval,
                                   save_net=args.save_net, load_net=args.load_net, eval_start=args.eval_start)))
        return
    if args.forward_2d:
        print(json.dumps(bench_forward_2d()))
        return
    if args.gradient_2d:
        print(json.dumps(bench_gradient_2d(workers=args.workers, with_cpu=not args.no_cpu)))
        return
    if args.autograd:
        print(json.dumps(bench_autograd(args.nt, args.checkpoint)))
        return
    if args.ot4 and not args.density:
        print(json.dumps({'fused': bench_ot4_3d(args.nt), 'general': bench_ot4_3d(args.nt, general=True)}))
        return
    if args.pml:
        print(json.dumps({'fused': bench_pml3d(args.nt), 'general': bench_pml3d(args.nt, general=True)}))
        return
    if args.power2:
        print(json.dumps({'fused': bench_power2(args.nt), 'general': bench_power2(args.nt, general=True)}))
        return
    if args.density_pair:
        tile = tuple(int(v) for v in args.tile.split(',')) if args.tile else None
        print(json.dumps(bench_density_pair(args.nt, tile)))
        return
    if args.density:
        tile = tuple(int(v) for v in args.tile.split(',')) if args.tile else None
        print(json.dumps(bench_density(args.nt, tile, with_cpu=not args.no_cpu, checkpoint=args.checkpoint,
                                       att_power=args.att_power, kernel='OT4' if args.ot4 else 'OT2')))
        return
    if args.schedule:
        tile = tuple(int(v) for v in args.tile.split(',')) if args.tile else None
        print(json.dumps(bench_schedule(tile, timed=True)))
        return
    cfg, plan = workload(scale=args.scale, nt=args.nt)
    if args.pipelines:
        res = bench_pipelines(plan)
        print(json.dumps({'pipelines': res, 'traces': [plan.nrec, plan.nt], 'grid': list(plan.shape),
                          'peak_gbs': HBM_PEAK_GBS}))
        return
    print('[bench] workload %s ext, nt %d, f %d, nrec %d' % (plan.shape, plan.nt, plan.undersampling, plan.nrec),
          file=sys.stderr, flush=True)
    shots = args.shots
    plan1 = second_shot(cfg, plan) if shots == 2 else None
    tile = None
    if args.tile:
        tile = tuple(int(v) for v in args.tile.split(','))
    elif not args.no_autotune:
        from stride_amd.autotune import autotune, autotune_batch
        t0 = time.perf_counter()
        tile, table = autotune_batch(plan, plan1) if shots == 2 else autotune(plan)
        print('[bench] autotune %.1f s: best %s (%.1f us/launch); worst %.1f us'
              % (time.perf_counter() - t0, tile, table[0][0] * 1e3, table[-1][0] * 1e3), file=sys.stderr, flush=True)
    rng = np.random.default_rng(0)
    if shots == 2:
        s = make_batch_solver(plan, plan1)
        if tile:
            s.batch_set_tile(*tile)
    else:
        s = make_solver(plan)
        if tile:
            s.set_tile(*tile)
    # device-resident inputs: source samples and residuals of each shot
    plans = (plan, plan1) if shots == 2 else (plan,)
    d_src, d_rec, d_res = [], [], []
    for p in plans:
        res = rng.standard_normal((p.nt, p.nrec)).astype(np.float32) * 1e-3
        d_src.append(s.device_alloc(p.src_data.nbytes))
        d_rec.append(s.device_alloc(p.nt * p.nrec * 4))
        d_res.append(s.device_alloc(res.nbytes))
        s.memcpy(d_src[-1], p.src_data.ctypes.data, p.src_data.nbytes, 1)
        s.memcpy(d_res[-1], res.ctypes.data, res.nbytes, 1)

    def one_step():
        if shots == 2:
            s.batch_forward_dev(d_src, d_rec)
            s.batch_adjoint_dev(d_res)
        else:
            s.forward_dev(d_src[0], d_rec[0])
            s.adjoint_dev(d_res[0], want=1)

    for _ in range(args.warmup):
        one_step()
    s.synchronize()
    s.timing(KERNEL_SAMPLE)
    t0 = time.perf_counter()
    for k in range(args.steps):
        one_step()
        print('[bench] step %d issued' % k, file=sys.stderr, flush=True)
    s.synchronize()
    t1 = time.perf_counter()
    kernel_ms, launches = s.timing()
    s.timing(False)

    total_bytes, fwd_steps, adj_steps, n, interior = algorithmic_bytes(plan, shots)
    wall = (t1 - t0) / args.steps
    updates = shots * n * (fwd_steps + adj_steps)
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
        'data': 'synthetic (smooth random vp in a 55 mm hemisphere, seeds 3001/3002; second shot sourced at the '
                'receiver farthest from the first source)',
        'config': {
            'workload': '3D %dx%dx%d inner (%dx%dx%d extended), dx 0.4 mm, SO10 OT2, sponge %d, nt %d, '
                        'fwd with p.dt2 save every %d + as-written adjoint with vp gradient, 1 source and %d '
                        'receivers per shot, %s' % (
                            tuple(cfg['space'].shape) + tuple(plan.shape) +
                            (plan.absorbing[0], plan.nt, plan.undersampling, plan.nrec,
                             'one shot per step' if shots == 1 else
                             '2 shots per step advanced together (interleaved wavefields, shared vp reads, '
                             'summed vp gradient)')),
            'shots_per_step': shots,
            'point_updates_per_step': updates,
            'inner_gpts': round(shots * interior * (fwd_steps + adj_steps) / wall / 1e9, 3),
            'hbm_roofline_frac_of_value': round(gpts * 1e9 * (total_bytes / updates) / 1e9 / HBM_PEAK_GBS, 4),
        },
        'roofline': {
            'bound': 'hbm',
            'achieved': round(achieved, 1),
            'peak': HBM_PEAK_GBS,
            'unit': 'GB/s',
            'frac': round(achieved / HBM_PEAK_GBS, 4),
            'traffic': traffic_from_profile(shots, tile),

            'kernel': 'k_fast3d_b2 (two-shot wavefield update)' if shots == 2 else 'k_fast3d (wavefield update)',
            'tile': list(tile) if tile else None,
            'avg_launch_us': round(avg_launch_s * 1e6, 2),
            'launches_timed': launches,
            'timed_every': KERNEL_SAMPLE,
            'algorithmic_bytes_per_launch': int(bytes_per_launch),
        },
    }
    for p in d_src + d_rec + d_res:
        s.device_free(p)
    s.close()
    if not args.no_schedule:
        # secondary line: the 32 uneven shots of the batched-forward workload in one graph launch
        print('[bench] 32-shot schedule', file=sys.stderr, flush=True)
        out['schedule'] = bench_schedule(tuple(tile) if (tile and shots == 2) else None)
    if not args.no_density:
        # secondary line: one shot of the 3D gradient workload's variable vp / rho / alpha medium
        print('[bench] 3D density shot', file=sys.stderr, flush=True)
        out['density'] = bench_density(args.nt, None, with_cpu=False)
        # the same shot with the raw snapshots recomputed from checkpoints (K = ceil(sqrt(2N)))
        print('[bench] 3D density shot, checkpointed snapshots', file=sys.stderr, flush=True)
        k = int(np.ceil(np.sqrt(2 * (args.nt // 4 + 1))))
        cp = bench_density(args.nt, None, with_cpu=False, checkpoint=k)
        out['density']['checkpointed'] = {key: cp[key] for key in ('value', 'seconds', 'snapshots', 'snapshot_store_gb',
                                                                   'update_launches', 'update_kernel_us')}
    if not args.no_gradient_2d:
        # secondary line: the 2D gradient workload (572^2, CFS-PML, variable density, power 2), eight
        # shots in lockstep; reported, never required
        print('[bench] 2D gradient workload, 8 shots in lockstep', file=sys.stderr, flush=True)
        try:
            g2 = bench_gradient_2d(with_cpu=False)
            out['gradient_2d'] = {key: g2[key] for key in ('metric', 'value', 'unit', 'seconds_per_shot', 'lockstep',
                                                           'value_workers', 'value_one_worker', 'update_kernel_us')}
        except Exception as exc:
            out['gradient_2d'] = {'error': repr(exc)}
    if not args.no_cpu:
        try:
            out['cpu_baseline'] = cpu_baseline(plan)
        except Exception as exc:  # the baseline is reported, never required
            out['cpu_baseline'] = {'error': repr(exc)}
    print(json.dumps(out))


if __name__ == '__main__':
    main()
