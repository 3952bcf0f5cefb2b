"""ctypes binding of the C ABI in include/stride_amd.h.

The HIP library is mandatory: when libstride_amd.so is missing or does not load,
every solver call raises ``NativeLibraryError``. There is no CPU fallback.
"""

import ctypes
import os

import numpy as np

__all__ = ['lib', 'NativeLibraryError', 'SolverError', 'Solver', 'library_path']

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, '_lib', 'libstride_amd.so')


class NativeLibraryError(RuntimeError):
    pass


class SolverError(RuntimeError):
    def __init__(self, code, message):
        super().__init__('stride_amd error %d: %s' % (code, message))
        self.code = code


class _Grid(ctypes.Structure):
    _fields_ = [('dim', ctypes.c_int), ('shape', ctypes.c_int * 3), ('spacing', ctypes.c_double * 3),
                ('absorbing', ctypes.c_int * 3), ('dt', ctypes.c_double), ('nt', ctypes.c_int),
                ('device', ctypes.c_int)]


class _Scheme(ctypes.Structure):
    _fields_ = [('kernel', ctypes.c_int), ('rho_mode', ctypes.c_int), ('att_power', ctypes.c_int),
                ('boundary', ctypes.c_int), ('diff_source', ctypes.c_int), ('undersampling', ctypes.c_int),
                ('save_mode', ctypes.c_int), ('tb0', ctypes.c_int), ('tb1', ctypes.c_int),
                ('grad_vp', ctypes.c_int), ('grad_rho', ctypes.c_int), ('grad_src', ctypes.c_int),
                ('checkpoint', ctypes.c_int)]


EXPORTS = ['sa_last_error', 'sa_version', 'sa_create', 'sa_destroy', 'sa_set_stream', 'sa_set_scheme',
           'sa_set_medium', 'sa_set_sponge', 'sa_set_pml', 'sa_set_sources', 'sa_set_receivers',
           'sa_forward', 'sa_forward_dev', 'sa_adjoint', 'sa_adjoint_dev', 'sa_read_gradient',
           'sa_adjoint_exact', 'sa_get_wavefield', 'sa_free_saved', 'sa_set_tile', 'sa_set_tile_density',
           'sa_timing_enable', 'sa_timing_read', 'sa_stream_create', 'sa_stream_destroy',
           'sa_device_alloc', 'sa_device_free', 'sa_memcpy', 'sa_synchronize', 'sa_batch_create',
           'sa_batch_set_taps', 'sa_batch_forward_dev', 'sa_batch_adjoint_dev', 'sa_batch_get_wavefield',
           'sa_batch_set_tile', 'sa_batch_set_tile_density', 'sa_shots2d_forward_dev', 'sa_shots2d_adjoint_dev',
           'sa_traces_check', 'sa_traces_mute', 'sa_traces_fir', 'sa_traces_sosfilt',
           'sa_traces_norm', 'sa_traces_l2', 'sa_traces_l2_noalloc', 'sa_multi_create', 'sa_multi_set_shot',
           'sa_multi_forward_dev', 'sa_multi_get_wavefield', 'sa_multi_graph_ms', 'sa_nn_conv3x3',
           'sa_nn_flip_weights', 'sa_nn_wgrad3x3', 'sa_nn_bias_grad', 'sa_nn_avgpool', 'sa_nn_avgpool_back',
           'sa_nn_upsample', 'sa_nn_upsample_back', 'sa_nn_copy_channels', 'sa_nn_pack_input', 'sa_nn_head',
           'sa_nn_head_back', 'sa_nn_mse', 'sa_nn_adam', 'sa_nn_to_bf16', 'sa_nn_from_bf16', 'sa_field_mask', 'sa_field_gaussian', 'sa_field_stats', 'sa_field_axpb',
           'sa_field_div_nonzero', 'sa_field_axpy', 'sa_set_medium_dev', 'sa_adjoint_exact_dev',
           'sa_exact_gradients_dev', 'sa_field_descent', 'sa_read_gradient_dev', 'sa_saved_count',
           'sa_saved_to_host', 'sa_saved_from_host', 'sa_set_dump', 'sa_dump_count', 'sa_dump_to_host',
           'sa_traces_transpose', 'sa_field_crop', 'sa_field_embed', 'sa_field_grad_vp_finish']

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise NativeLibraryError('%s is not built: run python -m stride_amd.build' % library_path)
        try:
            handle = ctypes.CDLL(library_path)
        except OSError as exc:
            raise NativeLibraryError('cannot load %s: %s' % (library_path, exc))
        vp = ctypes.c_void_p
        handle.sa_last_error.restype = ctypes.c_char_p
        handle.sa_version.restype = ctypes.c_char_p
        for name in EXPORTS[2:]:
            getattr(handle, name).restype = ctypes.c_int
        handle.sa_create.argtypes = [ctypes.POINTER(_Grid), ctypes.POINTER(vp)]
        _lib = handle
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


class Solver:
    """Owner of one ``sa_solver`` (one GPU, one grid)."""

    def __init__(self, dim, shape, spacing, absorbing, dt, nt, device=0):
        self._lib = lib()
        g = _Grid()
        g.dim = dim
        for i in range(3):
            g.shape[i] = shape[i] if i < dim else 1
            g.spacing[i] = spacing[i] if i < dim else 1.0
            g.absorbing[i] = absorbing[i] if i < dim else 0
        g.dt = dt
        g.nt = nt
        g.device = device
        self.key = (dim, tuple(shape), tuple(spacing), tuple(absorbing), float(dt), int(nt), int(device))
        self.shape = tuple(shape)
        self.nt = nt
        self.device = int(device)
        self._h = ctypes.c_void_p()
        self._check(self._lib.sa_create(ctypes.byref(g), ctypes.byref(self._h)))
        self.nsrc = self.nrec = 0
        self.src_ntap = 0

    def _check(self, rc):
        if rc != 0:
            msg = self._lib.sa_last_error().decode(errors='replace')
            if rc == -1:
                raise ValueError(msg)
            raise SolverError(rc, msg)

    def close(self):
        if self._h:
            self._lib.sa_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_scheme(self, kernel, rho_mode, att_power, boundary, diff_source, undersampling, save_mode,
                   tb0, tb1, grad_vp=0, grad_rho=0, grad_src=0, checkpoint=0):
        s = _Scheme(kernel, rho_mode, att_power, boundary, int(diff_source), undersampling, save_mode,
                    tb0, tb1, int(grad_vp), int(grad_rho), int(grad_src), int(checkpoint))
        self._check(self._lib.sa_set_scheme(self._h, ctypes.byref(s)))

    def set_medium(self, vp, rho=None, alpha=None):
        vp, rho, alpha = _f32(vp), _f32(rho), _f32(alpha)
        for a in (vp, rho, alpha):
            if a is not None and a.shape != self.shape:
                raise ValueError('medium array of shape %s, expected %s' % (a.shape, self.shape))
        self._check(self._lib.sa_set_medium(self._h, _p(vp), _p(rho), _p(alpha)))

    def set_sponge(self, tables):
        tables = _f32(tables)
        self._check(self._lib.sa_set_sponge(self._h, _p(tables), ctypes.c_int(tables.shape[-1])))

    def set_pml(self, profiles):
        profiles = _f32(profiles)
        self._check(self._lib.sa_set_pml(self._h, _p(profiles), ctypes.c_int(profiles.shape[-1])))

    def set_sources(self, idx, w):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        w = _f32(w)
        self._check(self._lib.sa_set_sources(self._h, ctypes.c_int(idx.shape[0]), ctypes.c_int(idx.shape[1]),
                                             _p(idx), _p(w)))
        self.nsrc, self.src_ntap = idx.shape

    def set_receivers(self, idx, w):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        w = _f32(w)
        self._check(self._lib.sa_set_receivers(self._h, ctypes.c_int(idx.shape[0]), ctypes.c_int(idx.shape[1]),
                                               _p(idx), _p(w)))
        self.nrec = idx.shape[0]

    def forward(self, src):
        src = _f32(src)
        if src.shape != (self.nt, self.nsrc):
            raise ValueError('source samples of shape %s, expected %s' % (src.shape, (self.nt, self.nsrc)))
        rec = np.zeros((self.nt, self.nrec), dtype=np.float32)
        self._check(self._lib.sa_forward(self._h, _p(src), _p(rec)))
        return rec

    def forward_dev(self, src_ptr, rec_ptr):
        self._check(self._lib.sa_forward_dev(self._h, ctypes.c_void_p(src_ptr), ctypes.c_void_p(rec_ptr)))

    def adjoint(self, res, want_vp=True, want_rho=False, want_src=False):
        res = _f32(res)
        if res.shape != (self.nt, self.nrec):
            raise ValueError('adjoint source of shape %s, expected %s' % (res.shape, (self.nt, self.nrec)))
        n = self.shape
        gv = np.zeros(n, np.float32) if want_vp else None
        pv = np.zeros(n, np.float32) if want_vp else None
        gr = np.zeros(n, np.float32) if want_rho else None
        pr = np.zeros(n, np.float32) if want_rho else None
        gs = np.zeros((self.nt, self.nsrc), np.float32) if want_src else None
        self._check(self._lib.sa_adjoint(self._h, _p(res), _p(gv), _p(pv), _p(gr), _p(pr), _p(gs)))
        return dict(grad_vp=gv, prec_vp=pv, grad_rho=gr, prec_rho=pr, grad_src=gs)

    def adjoint_dev(self, res_ptr, want=1, grad_src_ptr=None):
        self._check(self._lib.sa_adjoint_dev(self._h, ctypes.c_void_p(res_ptr), ctypes.c_int(want),
                                             ctypes.c_void_p(grad_src_ptr) if grad_src_ptr else None))

    def read_gradient(self, which):
        out = np.zeros(self.shape, np.float32)
        self._check(self._lib.sa_read_gradient(self._h, ctypes.c_int(which), _p(out)))
        return out

    def adjoint_exact(self, res, want_m=False, want_b=False):
        res = _f32(res)
        n = self.shape
        out = {k: np.zeros(n, np.float32) for k in ('gA', 'gB', 'gG')}
        out['gm'] = np.zeros(n, np.float32) if want_m else None
        out['gb'] = np.zeros(n, np.float32) if want_b else None
        lam = np.zeros((self.nt, self.nsrc * self.src_ntap), np.float32)
        self._check(self._lib.sa_adjoint_exact(self._h, _p(res), _p(out['gA']), _p(out['gB']), _p(out['gG']),
                                               _p(out['gm']), _p(out['gb']), _p(lam)))
        out['lam_src'] = lam.reshape(self.nt, self.nsrc, self.src_ntap)
        return out

    def set_medium_dev(self, vp_ptr, rho_ptr=None, alpha_ptr=None):
        """Medium from device pointers (unpadded extended arrays)."""
        vp_ = lambda x: ctypes.c_void_p(x) if x else None
        self._check(self._lib.sa_set_medium_dev(self._h, vp_(vp_ptr), vp_(rho_ptr), vp_(alpha_ptr)))

    def adjoint_exact_dev(self, res_ptr, gA, gB, gG, gm=None, gb=None, lam_ptr=None):
        vp_ = lambda x: ctypes.c_void_p(x) if x else None
        self._check(self._lib.sa_adjoint_exact_dev(self._h, vp_(res_ptr), vp_(gA), vp_(gB), vp_(gG), vp_(gm),
                                                   vp_(gb), vp_(lam_ptr)))

    def exact_gradients_dev(self, gA, gB, gG, gm=None, gb=None, lam_ptr=None, src_ptr=None, g_vp=None, g_rho=None,
                            g_alpha=None, g_src=None, alpha_scale=1.0):
        vp_ = lambda x: ctypes.c_void_p(x) if x else None
        self._check(self._lib.sa_exact_gradients_dev(self._h, vp_(gA), vp_(gB), vp_(gG), vp_(gm), vp_(gb),
                                                     vp_(lam_ptr), vp_(src_ptr), vp_(g_vp), vp_(g_rho),
                                                     vp_(g_alpha), vp_(g_src), ctypes.c_double(alpha_scale)))

    def wavefield(self, which=0):
        out = np.zeros(self.shape, np.float32)
        self._check(self._lib.sa_get_wavefield(self._h, ctypes.c_int(which), _p(out)))
        return out

    def free_saved(self):
        self._check(self._lib.sa_free_saved(self._h))

    # ---- stored snapshots in host memory (cache_forward) and wavefield dumps ----
    def _margin(self, margin):
        if margin is None:
            return None
        m = (ctypes.c_int * 3)(*[int(x) for x in margin] + [0] * (3 - len(margin)))
        return m

    def saved_count(self, margin=None):
        n = ctypes.c_longlong(0)
        self._check(self._lib.sa_saved_count(self._h, self._margin(margin), ctypes.byref(n)))
        return n.value

    def saved_to_host(self, margin=None, release=True):
        """The stored snapshots as a flat float32 array (nsave x region), the device copy released
        (release=True). margin: cells left out at each end of every axis (None: the whole box)."""
        out = np.empty(self.saved_count(margin), np.float32)
        self._check(self._lib.sa_saved_to_host(self._h, _p(out), ctypes.c_longlong(out.size), self._margin(margin),
                                               ctypes.c_int(int(release))))
        return out

    def saved_from_host(self, data, margin=None):
        data = np.ascontiguousarray(data, np.float32).ravel()
        self._check(self._lib.sa_saved_from_host(self._h, _p(data), ctypes.c_longlong(data.size),
                                                 self._margin(margin)))

    def set_dump(self, forward_factor=0, adjoint_factor=0):
        self._check(self._lib.sa_set_dump(self._h, ctypes.c_int(int(forward_factor)), ctypes.c_int(int(adjoint_factor))))

    def dump_to_host(self, which=0):
        """(slots, *extended shape) of the forward (0) or adjoint (1) wavefield dump."""
        n = ctypes.c_longlong(0)
        self._check(self._lib.sa_dump_count(self._h, ctypes.c_int(which), ctypes.byref(n)))
        out = np.empty(n.value, np.float32)
        self._check(self._lib.sa_dump_to_host(self._h, ctypes.c_int(which), _p(out), ctypes.c_longlong(out.size)))
        return out.reshape((-1,) + tuple(self.shape))

    # ---- two shots per pass (interleaved wavefields, shared vp, summed vp gradient) ----
    def batch_create(self):
        self._check(self._lib.sa_batch_create(self._h, ctypes.c_int(2)))
        self._batch = [dict(), dict()]

    def batch_set_taps(self, shot, src_idx, src_w, rec_idx, rec_w):
        si = np.ascontiguousarray(src_idx, dtype=np.int64)
        ri = np.ascontiguousarray(rec_idx, dtype=np.int64)
        sw, rw = _f32(src_w), _f32(rec_w)
        self._check(self._lib.sa_batch_set_taps(self._h, ctypes.c_int(shot), ctypes.c_int(si.shape[0]),
                                                ctypes.c_int(si.shape[1]), _p(si), _p(sw), ctypes.c_int(ri.shape[0]),
                                                ctypes.c_int(ri.shape[1]), _p(ri), _p(rw)))
        self._batch[shot] = dict(nsrc=si.shape[0], nrec=ri.shape[0])

    def batch_set_tile(self, rows, chunk, waves):
        self._check(self._lib.sa_batch_set_tile(self._h, ctypes.c_int(rows), ctypes.c_int(chunk), ctypes.c_int(waves)))

    def batch_set_tile_density(self, rows, chunk, waves):
        self._check(self._lib.sa_batch_set_tile_density(self._h, ctypes.c_int(rows), ctypes.c_int(chunk),
                                                        ctypes.c_int(waves)))

    def batch_forward_dev(self, src_ptrs, rec_ptrs):
        sp = (ctypes.c_void_p * 2)(*src_ptrs)
        rp = (ctypes.c_void_p * 2)(*rec_ptrs)
        self._check(self._lib.sa_batch_forward_dev(self._h, sp, rp))

    def batch_adjoint_dev(self, res_ptrs):
        rp = (ctypes.c_void_p * 2)(*res_ptrs)
        self._check(self._lib.sa_batch_adjoint_dev(self._h, rp))

    def _to_device(self, a):
        a = _f32(a)
        d = self.device_alloc(max(a.nbytes, 4))
        self.memcpy(d, a.ctypes.data, a.nbytes, 1)
        return d

    def batch_forward(self, srcs):
        """Host convenience: source samples (nt, nsrc_k) of the two shots -> their traces (nt, nrec_k)."""
        ds = [self._to_device(srcs[k]) for k in range(2)]
        recs = [np.zeros((self.nt, self._batch[k]['nrec']), np.float32) for k in range(2)]
        dr = [self.device_alloc(max(r.nbytes, 4)) for r in recs]
        try:
            self.batch_forward_dev(ds, dr)
            self.synchronize()
            for k in range(2):
                self.memcpy(recs[k].ctypes.data, dr[k], recs[k].nbytes, 2)
        finally:
            for d in ds + dr:
                self.device_free(d)
        return recs

    def batch_adjoint(self, residuals, density=False):
        """Host convenience: adjoint sources (nt, nrec_k) of the two shots -> summed vp gradient
        and preconditioner (extended grid, as read_gradient(0) / read_gradient(1)); with
        ``density`` (a density scheme) also the summed rho gradient and preconditioner."""
        dres = [self._to_device(residuals[k]) for k in range(2)]
        try:
            self.batch_adjoint_dev(dres)
            self.synchronize()
        finally:
            for d in dres:
                self.device_free(d)
        out = dict(grad_vp=self.read_gradient(0), prec_vp=self.read_gradient(1))
        if density:
            out.update(grad_rho=self.read_gradient(2), prec_rho=self.read_gradient(3))
        return out

    def batch_wavefield(self, shot, which=0):
        out = np.zeros(self.shape, np.float32)
        self._check(self._lib.sa_batch_get_wavefield(self._h, ctypes.c_int(shot), ctypes.c_int(which), _p(out)))
        return out

    # ---- many shots of uneven duration (forward) ----
    def multi_create(self, nshot):
        self._check(self._lib.sa_multi_create(self._h, ctypes.c_int(nshot)))
        self._multi = [None] * nshot

    def multi_set_shot(self, shot, nt, src_idx, src_w, rec_idx, rec_w):
        si = np.ascontiguousarray(src_idx, dtype=np.int64)
        ri = np.ascontiguousarray(rec_idx, dtype=np.int64)
        sw, rw = _f32(src_w), _f32(rec_w)
        self._check(self._lib.sa_multi_set_shot(self._h, ctypes.c_int(shot), ctypes.c_int(nt),
                                                ctypes.c_int(si.shape[0]), ctypes.c_int(si.shape[1]), _p(si), _p(sw),
                                                ctypes.c_int(ri.shape[0]), ctypes.c_int(ri.shape[1]), _p(ri), _p(rw)))
        self._multi[shot] = dict(nt=int(nt), nsrc=si.shape[0], nrec=ri.shape[0])

    def multi_forward_dev(self, src_ptrs, rec_ptrs, graph=True):
        n = len(self._multi)
        sp = (ctypes.c_void_p * n)(*src_ptrs)
        rp = (ctypes.c_void_p * n)(*rec_ptrs)
        self._check(self._lib.sa_multi_forward_dev(self._h, sp, rp, ctypes.c_int(int(graph))))

    def multi_forward(self, srcs, graph=True):
        """Host convenience: source samples (nt_k, nsrc_k) of every shot -> traces (nt_k, nrec_k)."""
        ds = [self._to_device(x) for x in srcs]
        recs = [np.zeros((m['nt'], m['nrec']), np.float32) for m in self._multi]
        dr = [self.device_alloc(max(r.nbytes, 4)) for r in recs]
        try:
            self.multi_forward_dev(ds, dr, graph)
            self.synchronize()
            for k, r in enumerate(recs):
                self.memcpy(r.ctypes.data, dr[k], r.nbytes, 2)
        finally:
            for d in ds + dr:
                self.device_free(d)
        return recs

    def multi_wavefield(self, shot, which=0):
        out = np.zeros(self.shape, np.float32)
        self._check(self._lib.sa_multi_get_wavefield(self._h, ctypes.c_int(shot), ctypes.c_int(which), _p(out)))
        return out

    def multi_graph_ms(self):
        ms = ctypes.c_double()
        self._check(self._lib.sa_multi_graph_ms(self._h, ctypes.byref(ms)))
        return ms.value

    # ---- steps around the solver on device arrays (pointers) ----
    def traces_check(self, ptr, ntr, nt, zero_bad=False):
        bad = np.zeros(max(ntr, 1), np.int32)
        nbad = ctypes.c_int()
        self._check(self._lib.sa_traces_check(self._h, ctypes.c_void_p(ptr), ctypes.c_int(ntr), ctypes.c_int(nt),
                                              ctypes.c_int(int(zero_bad)), _p(bad), ctypes.byref(nbad)))
        return bad[:ntr].astype(bool), nbad.value

    def traces_mute(self, mod_ptr, obs_ptr, ntr, nt, table):
        t = _f32(table)
        self._check(self._lib.sa_traces_mute(self._h, ctypes.c_void_p(mod_ptr), ctypes.c_void_p(obs_ptr),
                                             ctypes.c_int(ntr), ctypes.c_int(nt), ctypes.c_int(t.size), _p(t)))

    def traces_fir(self, ptr, ntr, nt, table, off, adjoint=False):
        t = _f32(table)
        self._check(self._lib.sa_traces_fir(self._h, ctypes.c_void_p(ptr), ctypes.c_int(ntr), ctypes.c_int(nt),
                                            ctypes.c_int(t.size), _p(t), ctypes.c_int(off),
                                            ctypes.c_int(int(adjoint))))

    def traces_sosfilt(self, ptr, ntr, nt, sos, adjoint=False):
        sos = np.ascontiguousarray(sos, dtype=np.float64)
        self._check(self._lib.sa_traces_sosfilt(self._h, ctypes.c_void_p(ptr), ctypes.c_int(ntr), ctypes.c_int(nt),
                                                ctypes.c_int(sos.shape[0]), _p(sos), ctypes.c_int(int(adjoint))))

    def traces_norm(self, ptr, ntr, nt, per_trace=False):
        norm = ctypes.c_double()
        self._check(self._lib.sa_traces_norm(self._h, ctypes.c_void_p(ptr), ctypes.c_int(ntr), ctypes.c_int(nt),
                                             ctypes.c_int(int(per_trace)), ctypes.byref(norm)))
        return norm.value

    def traces_l2(self, mod_ptr, obs_ptr, res_ptr, ntr, nt):
        fun = ctypes.c_double()
        self._check(self._lib.sa_traces_l2(self._h, ctypes.c_void_p(mod_ptr), ctypes.c_void_p(obs_ptr),
                                           ctypes.c_void_p(res_ptr), ctypes.c_int(ntr), ctypes.c_int(nt),
                                           ctypes.byref(fun)))
        return fun.value

    def traces_transpose(self, in_ptr, out_ptr, rows, cols, col_scale=None):
        w = _f32(col_scale)
        if w is not None and w.size != cols:
            raise ValueError('column scale of length %d, expected %d' % (w.size, cols))
        self._check(self._lib.sa_traces_transpose(self._h, ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr),
                                                  ctypes.c_int(rows), ctypes.c_int(cols), _p(w)))

    def field_mask(self, ptr, shape, lo, hi, ramp):
        d = len(shape)
        arr = lambda v: (ctypes.c_int * d)(*[int(x) for x in v])
        self._check(self._lib.sa_field_mask(self._h, ctypes.c_void_p(ptr), ctypes.c_int(d), arr(shape), arr(lo),
                                            arr(hi), ctypes.c_int(ramp)))

    def field_gaussian(self, ptr, shape, sigma, mode='nearest'):
        d = len(shape)
        sg = sigma if np.iterable(sigma) else (sigma,) * d
        self._check(self._lib.sa_field_gaussian(self._h, ctypes.c_void_p(ptr), ctypes.c_int(d),
                                                (ctypes.c_int * d)(*[int(x) for x in shape]),
                                                (ctypes.c_double * d)(*[float(x) for x in sg]),
                                                ctypes.c_int({'nearest': 0, 'reflect': 1}[mode])))

    def field_stats(self, ptr, n):
        lo, hi, ab = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self._lib.sa_field_stats(self._h, ctypes.c_void_p(ptr), ctypes.c_int64(n), ctypes.byref(lo),
                                             ctypes.byref(hi), ctypes.byref(ab)))
        return lo.value, hi.value, ab.value

    def field_axpb(self, ptr, n, a, b):
        self._check(self._lib.sa_field_axpb(self._h, ctypes.c_void_p(ptr), ctypes.c_int64(n), ctypes.c_double(a),
                                            ctypes.c_double(b)))

    def field_axpy(self, y_ptr, x_ptr, n, a=1.0):
        self._check(self._lib.sa_field_axpy(self._h, ctypes.c_void_p(y_ptr), ctypes.c_void_p(x_ptr),
                                            ctypes.c_int64(n), ctypes.c_double(a)))

    def field_div_nonzero(self, g_ptr, prec_ptr, n):
        self._check(self._lib.sa_field_div_nonzero(self._h, ctypes.c_void_p(g_ptr), ctypes.c_void_p(prec_ptr),
                                                   ctypes.c_int64(n)))

    def field_descent(self, x_ptr, dir_ptr, n, step, slowness=True, lo=None, hi=None):
        clip = lo is not None or hi is not None
        self._check(self._lib.sa_field_descent(self._h, ctypes.c_void_p(x_ptr), ctypes.c_void_p(dir_ptr),
                                               ctypes.c_int64(n), ctypes.c_double(step), ctypes.c_int(int(slowness)),
                                               ctypes.c_int(int(clip)), ctypes.c_double(-np.inf if lo is None else lo),
                                               ctypes.c_double(np.inf if hi is None else hi)))

    def _box(self, fn, a_ptr, b_ptr, shape, lo, hi):
        d = len(shape)
        arr = lambda v: (ctypes.c_int * d)(*[int(x) for x in v])
        self._check(fn(self._h, ctypes.c_void_p(a_ptr), ctypes.c_void_p(b_ptr), ctypes.c_int(d), arr(shape), arr(lo),
                       arr(hi)))

    def field_crop(self, grid_ptr, box_ptr, shape, lo, hi):
        self._box(self._lib.sa_field_crop, grid_ptr, box_ptr, shape, lo, hi)

    def field_embed(self, box_ptr, grid_ptr, shape, lo, hi):
        self._box(self._lib.sa_field_embed, box_ptr, grid_ptr, shape, lo, hi)

    def field_grad_vp_finish(self, grad_ptr, prec_ptr, vp_ptr, n, slowness=True):
        self._check(self._lib.sa_field_grad_vp_finish(self._h, ctypes.c_void_p(grad_ptr) if grad_ptr else None,
                                                      ctypes.c_void_p(prec_ptr) if prec_ptr else None,
                                                      ctypes.c_void_p(vp_ptr), ctypes.c_int64(n),
                                                      ctypes.c_int(int(slowness))))

    def read_gradient_dev(self, which, dev_ptr):
        self._check(self._lib.sa_read_gradient_dev(self._h, ctypes.c_int(which), ctypes.c_void_p(dev_ptr)))

    def set_tile(self, rows, chunk, version=7):
        self._check(self._lib.sa_set_tile(self._h, ctypes.c_int(rows), ctypes.c_int(chunk), ctypes.c_int(version)))

    @staticmethod
    def stream_create(device=0):
        """A new non-blocking stream of default priority on ``device`` (a hipStream_t handle)."""
        h = ctypes.c_void_p()
        code = lib().sa_stream_create(ctypes.c_int(device), ctypes.byref(h))
        if code != 0:
            raise SolverError(code, lib().sa_last_error().decode())
        return h.value

    @staticmethod
    def stream_destroy(stream):
        code = lib().sa_stream_destroy(ctypes.c_void_p(stream))
        if code != 0:
            raise SolverError(code, lib().sa_last_error().decode())

    def set_stream(self, stream):
        """HIP stream (a hipStream_t handle, e.g. torch.cuda.Stream().cuda_stream) the solver issues
        its work on; None: the null stream."""
        self._check(self._lib.sa_set_stream(self._h, ctypes.c_void_p(stream) if stream else None))

    def set_tile_density(self, rows, chunk=60, waves=8):
        """Tile of the fused 3D density update; rows=0 switches the fused 3D density and 2D updates
        off (the separate general passes run)."""
        self._check(self._lib.sa_set_tile_density(self._h, ctypes.c_int(rows), ctypes.c_int(chunk),
                                                  ctypes.c_int(waves)))

    def timing(self, enable=None):
        if enable is not None:
            self._check(self._lib.sa_timing_enable(self._h, ctypes.c_int(int(enable))))
            return None
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self._lib.sa_timing_read(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def device_alloc(self, nbytes):
        ptr = ctypes.c_void_p()
        self._check(self._lib.sa_device_alloc(self._h, ctypes.c_int64(nbytes), ctypes.byref(ptr)))
        return ptr.value

    def device_free(self, ptr):
        self._check(self._lib.sa_device_free(self._h, ctypes.c_void_p(ptr)))

    def memcpy(self, dst, src, nbytes, kind):
        self._check(self._lib.sa_memcpy(self._h, ctypes.c_void_p(dst), ctypes.c_void_p(src),
                                        ctypes.c_int64(nbytes), ctypes.c_int(kind)))

    def synchronize(self):
        self._check(self._lib.sa_synchronize(self._h))


# ---- several 2D shots per launch (lockstep; one solver per shot) ----
def _handles(solvers):
    return (ctypes.c_void_p * len(solvers))(*[s._h.value for s in solvers])


def shots2d_forward_dev(solvers, src_ptrs, rec_ptrs):
    """Forward runs of the 2D shots held by ``solvers`` (same grid and scheme), one launch per
    time step for all of them: device source samples / traces per shot."""
    s0 = solvers[0]
    n = len(solvers)
    s0._check(s0._lib.sa_shots2d_forward_dev(_handles(solvers), ctypes.c_int(n), (ctypes.c_void_p * n)(*src_ptrs),
                                             (ctypes.c_void_p * n)(*rec_ptrs)))


def shots2d_adjoint_dev(solvers, res_ptrs, want=1):
    """As-written adjoint runs of the 2D shots (after shots2d_forward_dev or one-shot forward runs
    with saves), one launch per time step; each solver keeps its own shot's gradients."""
    s0 = solvers[0]
    n = len(solvers)
    s0._check(s0._lib.sa_shots2d_adjoint_dev(_handles(solvers), ctypes.c_int(n), (ctypes.c_void_p * n)(*res_ptrs),
                                             ctypes.c_int(want)))


def shots2d_forward(solvers, srcs):
    """Host convenience: source samples (nt, nsrc_k) per shot -> traces (nt, nrec_k) per shot."""
    s0 = solvers[0]
    ds = [s._to_device(x) for s, x in zip(solvers, srcs)]
    recs = [np.zeros((s.nt, s.nrec), np.float32) for s in solvers]
    dr = [s0.device_alloc(max(r.nbytes, 4)) for r in recs]
    try:
        shots2d_forward_dev(solvers, ds, dr)
        s0.synchronize()
        for r, d in zip(recs, dr):
            s0.memcpy(r.ctypes.data, d, r.nbytes, 2)
    finally:
        for s, d in zip(solvers, ds):
            s.device_free(d)
        for d in dr:
            s0.device_free(d)
    return recs


def shots2d_adjoint(solvers, residuals, want_vp=True, want_rho=False):
    """Host convenience: adjoint sources (nt, nrec_k) per shot -> per-shot gradient dicts."""
    s0 = solvers[0]
    dres = [s._to_device(r) for s, r in zip(solvers, residuals)]
    try:
        shots2d_adjoint_dev(solvers, dres, (1 if want_vp else 0) | (2 if want_rho else 0))
        s0.synchronize()
    finally:
        for s, d in zip(solvers, dres):
            s.device_free(d)
    out = []
    for s in solvers:
        g = {}
        if want_vp:
            g.update(grad_vp=s.read_gradient(0), prec_vp=s.read_gradient(1))
        if want_rho:
            g.update(grad_rho=s.read_gradient(2), prec_rho=s.read_gradient(3))
        out.append(g)
    return out
