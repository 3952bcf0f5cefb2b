"""Learned gradient preconditioning for 2D ring FWI.

The workload is the 2D gradient geometry (512^2 inner grid, dx 0.2 mm, CFS-PML 30, 128
transducers on a 48 mm ring, dt 0.04 us, 2500 steps) with vp as the only unknown. A U-Net
(``stride_amd.unet``) maps the raw vp gradient of an iteration to a model update. It is trained
on pairs (gradient at a model m_b, true perturbation m_true - m_b) generated by this solver on
random smooth phantoms (1400 - 1650 m/s, sigma 2 - 5 mm, inside a 45 mm disc), where m_b is the
water background plus a random fraction of the smoothed perturbation (a partly recovered
model). During an inversion the network's update is kept only if the forward solver's misfit
drops; otherwise the plain step is taken. Two forms: ``invert_learned`` proposes at every
iteration (plain gradient step with backtracking as the fallback); ``invert_learned_lbfgs`` runs
the plain inversion's L-BFGS and replaces only the first step of each frequency band by the
network's update when that passes the line search's sufficient-decrease test. The plain
inversion it is compared with is L-BFGS on the same misfit and gradient. With one worker per
shot, a trial point's adjoint runs only when its forward misfit passes the test, so a rejected
trial costs one pass over the shots, in the plain and the learned inversions alike.

No counterpart in the reference: Stride's inversion loop (stride/__init__.py:172-494,
stride/optimisation/optimisation_loop.py) steps with the gradient of ``get_grad_vp``
(stride/physics/iso_acoustic/devito.py:1045-1111), which is what ``RingFWI.gradient`` returns
(the slowness rescale included), summed over shots as stride/core.py:565-633 does.

Everything per shot stays in HBM: source samples, observed traces, modelled traces, residuals
(``sa_traces_l2``) and gradients (``sa_read_gradient_dev``, ``sa_field_axpy``). Shots run on
``workers`` solvers, each on its own stream and host thread (``stride_amd.workers``).
"""

import time

import numpy as np
import scipy.ndimage

from ._native import Solver
from .configs import gaussian_noise
from .physics.plan import build_plan
from .problem import Space, Time
from .unet import UNet
from .utils import tone_burst, elliptical
from .workers import ShotWorkers

__all__ = ['RingFWI', 'phantom', 'make_pairs', 'make_trajectory_pairs', 'train', 'invert_plain',
           'invert_learned', 'invert_learned_lbfgs', 'save_network', 'load_network']

PML = 'complex_frequency_shift_PML_2'
VP_BOUNDS = (1300., 1800.)
TARGET_SCALE = 150.   # m/s per network output unit


def _torch():
    import torch
    return torch


def phantom(k, n=512, dx=0.2e-3, disc=45e-3):
    """Random smooth phantom k (seed 4000 + k): 1400 - 1650 m/s, sigma 2 - 5 mm, inside a disc of
    45 mm around the centre, 1500 m/s water outside (inner grid)."""
    rng = np.random.default_rng(4000 + k)
    sigma = rng.uniform(2e-3, 5e-3)
    noise = gaussian_noise((n, n), sigma / dx, 4000 + k)
    x = (np.arange(n) - (n - 1) / 2) * dx
    inside = x[:, None] ** 2 + x[None, :] ** 2 <= (disc * n / 512) ** 2
    return np.where(inside, 1400. + (noise + 1) / 2 * 250., 1500.).astype(np.float32)


class RingFWI:
    """Misfit and vp gradient over the ring's shots for any vp model (inner grid), on the device.

    ``nsrc`` transducers, evenly spread over the ring, act as sources; all 128 record."""

    def __init__(self, n=512, dx=0.2e-3, nt=2500, dt=0.04e-6, nsrc=16, workers=2, device=0, absorbing=30):
        torch = _torch()
        self.device = device
        self.dev = torch.device('cuda', device)
        self.space = Space(shape=(n, n), spacing=dx, extra=(absorbing, absorbing), absorbing=(absorbing, absorbing))
        self.time = Time(start=0., step=dt, num=nt)
        c = tuple(l / 2 for l in self.space.limit)
        r = 48e-3 * n / 512
        ring = elliptical(128, (r, r), c)
        self.receivers = ring
        self.sources = ring[::128 // nsrc][:nsrc]
        self.nsrc = nsrc
        self.nt = nt
        self.nrec = len(ring)
        self.inner = self.space.inner
        shape = self.space.extended_shape
        def make():
            return Solver(2, shape, self.space.spacing, self.space.absorbing, dt, nt, device=device)

        self.pool = ShotWorkers(make, n=workers, device=device)
        nw = len(self.pool.solvers)
        ext = int(np.prod(shape))
        self.buf = [dict(rec=torch.zeros((nt, self.nrec), device=self.dev),
                         res=torch.zeros((nt, self.nrec), device=self.dev),
                         g=torch.zeros(shape, device=self.dev),
                         total=torch.zeros(shape, device=self.dev)) for _ in range(nw)]
        self.ext_points = ext
        self.passes = 0        # forward or adjoint passes over all shots (solver calls)
        self._obs = {}

    # ---- plans and data -----------------------------------------------------------------------
    def _wavelet(self, freq):
        return tone_burst(freq, 3, self.nt, self.time.step)[None]

    def _plan(self, vp_inner, freq, grad=False):
        vp = np.pad(np.asarray(vp_inner, np.float32), [(e, e) for e in self.space.extra], mode='edge')
        wav = np.repeat(self._wavelet(freq), self.nsrc, 0)
        return build_plan(self.space, self.time, wav, vp, source_coordinates=self.sources,
                          receiver_coordinates=self.receivers, kernel='OT2', boundary_type=PML, grad_vp=grad)

    def _configure(self, plan, save_mode):
        for s in self.pool.solvers:
            s.set_scheme(plan.kernel_order, plan.rho_mode, plan.att_power, plan.boundary, plan.diff_source,
                         plan.undersampling, save_mode, plan.tb0, plan.tb1, 1, 0, 0)
            s.set_medium(plan.vp)
            s.set_pml(plan.pml)
            s.set_receivers(plan.rec_idx, plan.rec_w)
        torch = _torch()
        # per-shot source samples (nt, 1) on the device
        return [torch.from_numpy(np.ascontiguousarray(plan.src_data[:, k:k + 1])).to(self.dev)
                for k in range(self.nsrc)]

    def observe(self, vp_true, freq):
        """Observed traces of every shot at `freq` (device tensors, kept for later calls)."""
        torch = _torch()
        plan = self._plan(vp_true, freq)
        src = self._configure(plan, 0)
        obs = [torch.zeros((self.nt, self.nrec), device=self.dev) for _ in range(self.nsrc)]
        torch.cuda.synchronize(self.dev)

        def shot(s, k):
            s.set_sources(plan.src_idx[k:k + 1], plan.src_w[k:k + 1])
            s.forward_dev(src[k].data_ptr(), obs[k].data_ptr())

        self.pool.map(shot, range(self.nsrc))
        self._obs[freq] = obs
        return obs

    # ---- misfit and gradient ------------------------------------------------------------------
    def misfit(self, vp_inner, freq):
        plan = self._plan(vp_inner, freq)
        src = self._configure(plan, 0)
        obs = self._obs[freq]
        _torch().cuda.synchronize(self.dev)
        wid = {id(s): i for i, s in enumerate(self.pool.solvers)}

        def shot(s, k):
            b = self.buf[wid[id(s)]]
            s.set_sources(plan.src_idx[k:k + 1], plan.src_w[k:k + 1])
            s.forward_dev(src[k].data_ptr(), b['rec'].data_ptr())
            return s.traces_l2(b['rec'].data_ptr(), obs[k].data_ptr(), b['res'].data_ptr(), self.nt, self.nrec)

        f = float(sum(self.pool.map(shot, range(self.nsrc))))
        self.passes += 1
        return f

    def gradient(self, vp_inner, freq, below=None):
        """(misfit, vp gradient on the inner grid as a device tensor): the shots' vp gradients
        summed, with get_grad_vp's slowness rescale (x 1 / vp). With `below`, when every shot has a
        worker of its own (workers >= nsrc, so each keeps its forward wavefield): the forward
        runs over all shots first and, when the misfit is not below `below`, (misfit, None) is
        returned after that one pass, without the adjoint."""
        torch = _torch()
        plan = self._plan(vp_inner, freq, grad=True)
        src = self._configure(plan, 1)
        obs = self._obs[freq]
        for b in self.buf:
            b['total'].zero_()
        torch.cuda.synchronize(self.dev)
        wid = {id(s): i for i, s in enumerate(self.pool.solvers)}

        def forward(s, k):
            b = self.buf[wid[id(s)]]
            s.set_sources(plan.src_idx[k:k + 1], plan.src_w[k:k + 1])
            s.forward_dev(src[k].data_ptr(), b['rec'].data_ptr())
            return s.traces_l2(b['rec'].data_ptr(), obs[k].data_ptr(), b['res'].data_ptr(), self.nt, self.nrec)

        def adjoint(s, k):
            b = self.buf[wid[id(s)]]
            # the residual is zero at t = 0 and nt - 1 (no receiver samples there), where the
            # adjoint source window of devito.py:863-869 is zero: the window changes nothing
            s.adjoint_dev(b['res'].data_ptr(), want=1)
            s.read_gradient_dev(0, b['g'].data_ptr())
            s.field_axpy(b['total'].data_ptr(), b['g'].data_ptr(), self.ext_points, 1.0)

        if below is not None and len(self.pool.solvers) >= self.nsrc:
            f = float(sum(self.pool.map(forward, range(self.nsrc))))
            self.passes += 1
            if not f < below:
                return f, None
            self.pool.map(adjoint, range(self.nsrc))
            self.passes += 1
        else:
            def shot(s, k):
                f = forward(s, k)
                adjoint(s, k)
                return f

            f = float(sum(self.pool.map(shot, range(self.nsrc))))
            self.passes += 2
        s0 = self.pool.solvers[0]
        total = self.buf[0]['total']
        for b in self.buf[1:]:
            s0.field_axpy(total.data_ptr(), b['total'].data_ptr(), self.ext_points, 1.0)
        n = self.space.shape
        g = torch.empty(n, device=self.dev)
        s0.field_crop(total.data_ptr(), g.data_ptr(), self.space.extended_shape,
                      [e for e in self.space.extra], [e + m for e, m in zip(self.space.extra, n)])
        vp = torch.from_numpy(np.ascontiguousarray(vp_inner, dtype=np.float32)).to(self.dev)
        s0.field_grad_vp_finish(g.data_ptr(), None, vp.data_ptr(), g.numel(), slowness=True)
        s0.synchronize()
        return f, g

    def close(self):
        self.pool.close()


# ---- training pairs and training ----------------------------------------------------------------

def partial_model(true, rng, smooth_cells=10.0):
    """Water plus a random fraction (0 - 0.8) of the smoothed perturbation: a partly recovered
    model at which a training gradient is taken."""
    beta = rng.uniform(0.0, 0.8)
    pert = scipy.ndimage.gaussian_filter(true.astype(np.float64) - 1500., smooth_cells)
    return (1500. + beta * pert).astype(np.float32)


def make_pairs(fwi, ks, freqs, seed=0, log=None):
    """Solver-generated pairs for phantoms ks: (gradient (device), true perturbation / 150 m/s
    (device), frequency, the model the gradient was taken at, (m - 1500) / 150 m/s (device))."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    pairs = []
    t0 = time.perf_counter()
    for i, k in enumerate(ks):
        true = phantom(k, n=fwi.space.shape[0], dx=fwi.space.spacing[0])
        freq = float(freqs[rng.integers(len(freqs))])
        base = partial_model(true, rng)
        fwi.observe(true, freq)
        _, g = fwi.gradient(base, freq)
        target = torch.from_numpy((true - base) / TARGET_SCALE).to(fwi.dev)
        model = torch.from_numpy((base - 1500.) / TARGET_SCALE).to(fwi.dev)
        pairs.append((g, target, freq, model))
        if log and (i + 1) % 25 == 0:
            log('pairs %d / %d (%.1f s)' % (i + 1, len(ks), time.perf_counter() - t0))
    return pairs


def input_scales(pairs):
    """Per frequency: 1 / median of max |gradient| over its training pairs (a fixed linear
    normalisation, so the network sees the gradient's amplitude)."""
    out = {}
    for f in sorted(set(p[2] for p in pairs)):
        m = [float(p[0].abs().max()) for p in pairs if p[2] == f]
        out[f] = 1.0 / max(float(np.median(m)), 1e-30)
    return out


def train(pairs, scales, epochs=60, batch=4, lr=1e-3, seed=0, device=0, log=None):
    """Adam with a cosine-decayed rate on the mean-square error of the predicted perturbation; the
    network sees the scaled gradient and the model it was taken at."""
    torch = _torch()
    net = UNet(device=device, seed=seed)
    rng = np.random.default_rng(seed)
    n = len(pairs)
    losses = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(n)
        ep_loss = []
        for i in range(0, n - batch + 1, batch):
            idx = order[i:i + batch]
            g = torch.stack([pairs[j][0] * scales[pairs[j][2]] for j in idx])
            y = torch.stack([pairs[j][1] for j in idx])
            m = torch.stack([pairs[j][3] for j in idx])
            rate = lr * 0.5 * (1 + np.cos(np.pi * ep / epochs))   # cosine decay
            ep_loss.append(net.train_step(g, y, scale=1.0, lr=rate, model=m))
        losses.append(float(np.mean(ep_loss)))
        if log:
            log('epoch %d loss %.4g (%.1f s)' % (ep, losses[-1], time.perf_counter() - t0))
    return net, losses


def save_network(net, scales, path):
    """The trained network's fp32 parameters and the input scales, as an .npz of plain arrays."""
    arrays = {'w:' + k: v for k, v in net.state().items()}
    freqs = sorted(scales)
    np.savez(path, scale_freqs=np.array(freqs, np.float64), scale_values=np.array([scales[f] for f in freqs]),
             **arrays)


def load_network(path, device=0):
    """(network, input scales) from save_network's file."""
    with np.load(path, allow_pickle=False) as z:
        net = UNet(device=device)
        net.load_state({k[2:]: z[k] for k in z.files if k.startswith('w:')})
        scales = dict(zip(z['scale_freqs'].tolist(), z['scale_values'].tolist()))
    return net, scales


# ---- inversions ------------------------------------------------------------------------------------

def _model_error(vp, true):
    return float(np.linalg.norm(vp - true) / np.linalg.norm(true - 1500.))


def make_trajectory_pairs(fwi, ks, freqs, iters_per_freq=4, log=None):
    """Solver-generated pairs along plain inversions of the training phantoms ks: every gradient
    the plain L-BFGS evaluates (line searches included), with the true perturbation from the model
    it was taken at: (gradient (device), (m_true - m) / 150 m/s, frequency, (m - 1500) / 150 m/s).
    These are the models an inversion meets, partly recovered by the inversion itself."""
    torch = _torch()
    pairs = []
    t0 = time.perf_counter()
    for i, k in enumerate(ks):
        true = phantom(k, n=fwi.space.shape[0], dx=fwi.space.spacing[0])

        def record(g, vp, freq):
            pairs.append((g, torch.from_numpy((true - vp) / TARGET_SCALE).to(fwi.dev), freq,
                          torch.from_numpy((vp - 1500.) / TARGET_SCALE).to(fwi.dev)))

        invert_plain(fwi, true, freqs, iters_per_freq, record=record)
        if log:
            log('trajectory %d / %d: %d pairs (%.1f s)' % (i + 1, len(ks), len(pairs), time.perf_counter() - t0))
    return pairs


def _fun_and_grad(fwi, shape, freq, record=None):
    """L-BFGS objective in slowness: (misfit, gradient) at x, or (misfit, None) when the misfit
    exceeds the trial's threshold (the adjoint skipped)."""
    def fg(xs, threshold=None):
        vp = (1 / xs).reshape(shape).astype(np.float32)
        f, g = fwi.gradient(vp, freq, below=threshold)
        if g is None:
            return f, None
        if record is not None:
            record(g, vp, freq)
        return f, g.double().cpu().numpy().reshape(-1)
    return fg


def invert_plain(fwi, true, freqs, iters_per_freq=4, first_step=20.0, record=None):
    """L-BFGS per frequency band (restarted at each band) from water, in slowness, the space the
    gradient of get_grad_vp belongs to with vp's default transform (the reference steps in the
    transformed variable, stride/optimisation/optimisers/optimiser.py:242-330). The first step
    changes vp by at most about first_step m/s. Returns (vp, history of (solver passes, misfit at
    the band, model error, band)); record(gradient, vp, band), when given, sees every gradient."""
    from .optim import LBFGS
    shape = fwi.space.shape
    x = np.full(int(np.prod(shape)), 1 / 1500., np.float64)
    hist = []
    base = fwi.passes
    bounds = (1 / VP_BOUNDS[1], 1 / VP_BOUNDS[0])
    for freq in freqs:
        fwi.observe(true, freq)
        opt = LBFGS(memory=5, first_step=first_step / 1500. ** 2, bounds=bounds)

        def cb(xs, f):
            hist.append((fwi.passes - base, f, _model_error((1 / xs).reshape(shape), true), freq))

        x, _ = opt.minimize(x, _fun_and_grad(fwi, shape, freq, record), iters_per_freq, callback=cb, gated=True)
    return (1 / x).reshape(shape).astype(np.float32), hist


def invert_learned_lbfgs(fwi, net, scales, true, freqs, iters_per_freq=4, first_step=20.0):
    """L-BFGS per band exactly as invert_plain, except that the first step of every band is the
    network's update from that band's first gradient and the current model: it is kept when the
    misfit drops enough (the line search's sufficient-decrease test at full length), otherwise
    the scaled gradient step of invert_plain is taken. Returns (vp, history, proposals kept)."""
    from .optim import LBFGS
    torch = _torch()
    shape = fwi.space.shape
    x = np.full(int(np.prod(shape)), 1 / 1500., np.float64)
    hist = []
    kept = 0
    base = fwi.passes
    bounds = (1 / VP_BOUNDS[1], 1 / VP_BOUNDS[0])
    for freq in freqs:
        fwi.observe(true, freq)
        opt = LBFGS(memory=5, first_step=first_step / 1500. ** 2, bounds=bounds)
        scale = scales[min(scales, key=lambda k: abs(k - freq))]

        def proposal(xs, g):
            vp = (1 / xs).reshape(shape)
            gt = torch.from_numpy(g.reshape(shape).astype(np.float32)).to(fwi.dev)
            model = torch.from_numpy(((vp - 1500.) / TARGET_SCALE).astype(np.float32)).to(fwi.dev)
            d = (net.forward((gt * scale)[None], scale=1.0, model=model[None])[0] * TARGET_SCALE).cpu().numpy()
            new = np.clip(vp + d, *VP_BOUNDS)
            return (xs - 1 / new.reshape(-1)).astype(np.float64)  # the step is x -> 1 / new

        def cb(xs, f):
            hist.append((fwi.passes - base, f, _model_error((1 / xs).reshape(shape), true), freq))

        x, _ = opt.minimize(x, _fun_and_grad(fwi, shape, freq), iters_per_freq, callback=cb,
                            first_direction=proposal, gated=True)
        kept += opt.proposals_kept
    return (1 / x).reshape(shape).astype(np.float32), hist, kept


def _gradient_step(vp, g, step):
    """The plain step in slowness: 1 / vp - step g, mapped back to vp and clipped."""
    return np.clip(1 / (1 / vp.astype(np.float64) - step * g), *VP_BOUNDS).astype(np.float32)


def invert_learned(fwi, net, scales, true, freqs, iters_per_freq=4, first_step=20.0, max_backtracks=3,
                   proposal_scales=(1.0, 0.5)):
    """Each iteration: the gradient at the current model and the network's update d from it and
    the model; m + s d is kept for the first s of proposal_scales whose misfit is lower than the
    current one; if none is, the plain gradient step in slowness (vp changed by at most about
    first_step m/s, halved until the misfit drops). Returns (vp, history of (solver passes,
    misfit, model error, band), number of network updates kept)."""
    torch = _torch()
    vp = np.full(fwi.space.shape, 1500., np.float32)
    hist = []
    accepted = 0
    base = fwi.passes
    for freq in freqs:
        fwi.observe(true, freq)
        for _ in range(iters_per_freq):
            f, g = fwi.gradient(vp, freq)
            scale = scales[min(scales, key=lambda k: abs(k - freq))]
            model = torch.from_numpy((vp - 1500.) / TARGET_SCALE).to(g.device)
            d = (net.forward((g * scale)[None], scale=1.0, model=model[None])[0] * TARGET_SCALE).cpu().numpy()
            kept = False
            for ps in proposal_scales:
                prop = np.clip(vp + ps * d, *VP_BOUNDS).astype(np.float32)
                fp = fwi.misfit(prop, freq)
                if fp < f:
                    vp, f = prop, fp
                    accepted += 1
                    kept = True
                    break
            if not kept:
                gn = g.double().cpu().numpy()
                step = first_step / 1500. ** 2 / max(np.abs(gn).max(), 1e-30)
                for _ in range(max_backtracks):
                    cand = _gradient_step(vp, gn, step)
                    fc = fwi.misfit(cand, freq)
                    if fc < f:
                        vp, f = cand, fc
                        break
                    step *= 0.5
            hist.append((fwi.passes - base, f, _model_error(vp, true), freq))
    return vp, hist, accepted
