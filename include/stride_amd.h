/*
 * stride_amd: MI355X (gfx950) iso-acoustic forward / adjoint solver, C ABI.
 *
 * One solver object holds one shot's worth of device state on one GPU: the
 * medium, boundary description, source / receiver taps, the wavefield buffers
 * and the saved forward wavefield that the adjoint consumes. It replaces the
 * three Devito operators of Stride's IsoAcousticDevito (acoustic_iso_state,
 * acoustic_iso_state_save, acoustic_iso_adjoint;
 * stride/physics/iso_acoustic/devito.py:151-162) and the buffers GridDevito
 * keeps for them (stride/physics/common/devito.py:197-885).
 *
 * Conventions
 *  - Arrays are float32, C order, last axis fastest, on the extended grid
 *    (inner shape + 2*extra), exactly as ScalarField.extended_data
 *    (stride/problem/data.py:284-293).
 *  - Time series are time-major: (nt, npoint).
 *  - Functions return 0 on success, a negative SA_E* code on failure;
 *    sa_last_error() gives the message (thread-local).
 *  - "_dev" entry points take device pointers and do not synchronise; the
 *    others take host pointers, copy, run and synchronise.
 *  - All work is issued on the solver's stream (sa_set_stream; default the
 *    null stream).
 */
#ifndef STRIDE_AMD_H
#define STRIDE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SA_OK 0
#define SA_E_ARG -1      /* invalid argument (ValueError in the reference) */
#define SA_E_STATE -2    /* call out of order (e.g. adjoint before a saving forward) */
#define SA_E_HIP -3      /* HIP runtime error */
#define SA_E_NOMEM -4    /* device allocation failed */
#define SA_E_UNSUPPORTED -5

typedef struct sa_solver sa_solver;

/* Grid and time axis: Space/Time (stride/problem/domain.py:48-74, 285-330).
 * absorbing: boundary layer thickness; the interior is middle(absorbing)
 * (stride/physics/common/devito.py:35-47, 244). */
typedef struct {
    int dim;             /* 2 or 3 */
    int shape[3];        /* extended shape */
    double spacing[3];   /* metres */
    int absorbing[3];
    double dt;           /* seconds */
    int nt;              /* time.num */
    int device;          /* HIP device ordinal */
} sa_grid;

/* Scheme selection (IsoAcousticDevito kwargs, devito.py:40-115 and _check_conditions :1354-1464). */
typedef struct {
    int kernel;          /* 2: OT2, 4: OT4 */
    int rho_mode;        /* 0: no density, 1: constant density (div grad), 2: variable density */
    int att_power;       /* -2: no attenuation, 0, 2, -1: attenuation_power=None */
    int boundary;        /* -1: none, 0: sponge_boundary_2, 1: complex_frequency_shift_PML_2 */
    int diff_source;     /* source injected as is (0) or already differentiated (1): changes the dt power */
    int undersampling;   /* forward wavefield save every f steps */
    int save_mode;       /* 0: no save, 1: p.dt2 (vp gradient), 2: raw p (rho / wavelet gradients),
                            3: raw p every step (exact adjoint), 4: exact adjoint by checkpointed
                            recomputation (undersampling = levels per segment), 5: the raw p snapshots
                            of mode 2 by checkpointed recomputation (`checkpoint` snapshots per
                            segment): the forward keeps one state per segment, the adjoint recomputes
                            each segment's snapshots from it; the gradients equal mode 2's */
    int tb0, tb1;        /* time_bounds */
    int grad_vp, grad_rho, grad_src;
    int checkpoint;      /* save_mode 5: snapshots per recomputation segment (>= 1) */
} sa_scheme;

const char *sa_last_error(void);
const char *sa_version(void);

/* IsoAcousticDevito.__init__ / GridDevito (devito.py:122-183): allocate the solver on grid->device. */
int sa_create(const sa_grid *grid, sa_solver **out);
/* deallocate everything (deallocate=True path, devito.py:673-683, 969-979) */
int sa_destroy(sa_solver *s);
/* stream the solver issues work on (a hipStream_t); NULL is the null stream */
int sa_set_stream(sa_solver *s, void *stream);

/* _check_problem / _stencil choices (devito.py:1266-1352, 1466-1608) */
int sa_set_scheme(sa_solver *s, const sa_scheme *scheme);

/* Medium (devito.py:476-497): vp required; rho used for rho_mode 2; alpha (already in
 * nepers, i.e. times db_to_neper) used when att_power != -2. Host arrays, extended shape. */
int sa_set_medium(sa_solver *s, const float *vp, const float *rho, const float *alpha);

/* SpongeBoundary2 layer values per axis before the velocity factor: tables[d * len + l] for
 * layer l counted from the outer edge (stride/physics/boundaries/boundary.py:59-171). */
int sa_set_sponge(sa_solver *s, const float *tables, int len);

/* ComplexFrequencyShiftPML2 profiles: profiles[(d * 4 + r) * stride + x], r = sigma, alpha,
 * dsigma/dx, dalpha/dx (stride/physics/boundaries/devito.py:100-131). */
int sa_set_pml(sa_solver *s, const float *profiles, int stride);

/* Sparse points: npoint x ntap linear indices into the extended grid (-1 = no tap) and
 * interpolation weights (src.inject / rec.interpolate, devito.py:338-364). */
int sa_set_sources(sa_solver *s, int npoint, int ntap, const int64_t *index, const float *weight);
int sa_set_receivers(sa_solver *s, int npoint, int ntap, const int64_t *index, const float *weight);

/* State run (run_forward, devito.py:515-581): time steps 1 .. tb1-2, src (nt, nsrc) already
 * windowed, rec (nt, nrec) out. Saves the wavefield per scheme->save_mode. */
int sa_forward(sa_solver *s, const float *src, float *rec);
int sa_forward_dev(sa_solver *s, const float *src_dev, float *rec_dev);

/* As-written adjoint (run_adjoint, devito.py:875-920): steps tb1-1-f down to tb0+1,
 * residual res (nt, nrec) already windowed. Gradient and preconditioner accumulators
 * are returned on the whole extended grid (non-zero on the interior only), before the
 * get_grad_vp rescale; any output pointer may be NULL. grad_src (nt, nsrc). */
int sa_adjoint(sa_solver *s, const float *res, float *grad_vp, float *prec_vp, float *grad_rho,
               float *prec_rho, float *grad_src);

#define SA_WANT_VP 1
#define SA_WANT_RHO 2
/* Device-resident form: res_dev and grad_src_dev (nt, nsrc; may be NULL) are device pointers;
 * the gradients requested in `want` stay in the solver until read with sa_read_gradient. */
int sa_adjoint_dev(sa_solver *s, const float *res_dev, int want, float *grad_src_dev);
/* which: 0 grad_vp, 1 prec_vp, 2 grad_rho, 3 prec_rho -> host array (extended grid) */
int sa_read_gradient(sa_solver *s, int which, float *out);

/* Exact adjoint of the state run for J = sum_t <res[t], rec[t]> (no reference counterpart:
 * the reference adjoint above is not the discrete transpose). Requires save_mode 3.
 * Outputs (extended grid): sensitivities of J to the per-point step coefficients
 * A, B, G, m and to the buoyancy b (see stride_amd.physics.gradients); lam_src
 * (nt, nsrc * ntap) holds the adjoint field at each source tap. */
int sa_adjoint_exact(sa_solver *s, const float *res, float *gA, float *gB, float *gG, float *gm,
                     float *gb, float *lam_src);

/* Device-resident forms for callers that keep the medium, residuals and gradients in HBM (the
 * torch.autograd path; no counterpart in the reference). sa_set_medium_dev: as sa_set_medium
 * with device pointers (unpadded extended arrays; the padding and 1/rho run on the device).
 * sa_adjoint_exact_dev: as sa_adjoint_exact with res, the accumulators (unpadded extended grid)
 * and lam_src device pointers (any output may be NULL); queued on the solver's stream.
 * sa_exact_gradients_dev: the chain rule from those accumulators to dJ/dvp, dJ/drho, dJ/dalpha
 * (alpha_scale: the medium's alpha units per caller unit, e.g. the dB-to-neper factor), by
 * complex-step derivatives of the step coefficients in double precision on the device; with
 * lam_src and the source samples src (nt, nsrc) the source taps' term is added to dJ/dvp and
 * dJ/dsrc (nt, nsrc) is written to g_src. Outputs device pointers, unpadded, any may be NULL. */
int sa_set_medium_dev(sa_solver *s, const float *vp, const float *rho, const float *alpha);
int sa_adjoint_exact_dev(sa_solver *s, const float *res, float *gA, float *gB, float *gG, float *gm, float *gb,
                         float *lam_src);
int sa_exact_gradients_dev(sa_solver *s, const float *gA, const float *gB, const float *gG, const float *gm,
                           const float *gb, const float *lam_src, const float *src, float *g_vp, float *g_rho,
                           float *g_alpha, float *g_src, double alpha_scale);

/* Copy a wavefield level to the host (extended grid). which = 0: last computed level,
 * 1: the level before it. Used for wavefield dumps (dump_forward_wavefield). */
int sa_get_wavefield(sa_solver *s, int which, float *out);

/* deallocate_wavefield (devito.py:191-196): release the saved forward wavefield. */
int sa_free_saved(sa_solver *s);

/* cache_forward (devito.py:632-668 after the forward run, :820-842 before the adjoint): the
 * stored snapshots (save_mode 1 or 2) to host memory and back. Host layout: (nsave, region),
 * region = the stored box without margin[a] cells at each end of axis a (grid->dim values;
 * NULL keeps the whole box; the reference keeps `inner_wavefield`, margin = space.extra).
 * sa_saved_count gives the float count of that layout; sa_saved_to_host with release = 1 frees
 * the device copy (the adjoint then refuses to run until sa_saved_from_host);
 * sa_saved_from_host re-allocates it, zero outside the region. */
int sa_saved_count(sa_solver *s, const int *margin, long long *n);
int sa_saved_to_host(sa_solver *s, float *host, long long n, const int *margin, int release);
int sa_saved_from_host(sa_solver *s, const float *host, long long n, const int *margin);

/* dump_forward_wavefield / dump_adjoint_wavefield (devito.py:421-432, 765-778): keep the raw level
 * of every factor-th step (0: off) over the whole extended grid, (tb1 - tb0 + f) / f + 1 slots
 * at slot t / f, as undersampled_time_function (common/devito.py:552-578) sizes them. which = 0
 * forward, 1 adjoint; sa_dump_to_host copies (slots, extended shape) after the run. */
int sa_set_dump(sa_solver *s, int forward_factor, int adjoint_factor);
int sa_dump_count(sa_solver *s, int which, long long *n);
int sa_dump_to_host(sa_solver *s, int which, float *host, long long n);

/* Tile of the specialised 3D update (3D, no density, save mode 0 or 1; OT2 with sponge, no
 * boundary or CFS-PML (the layers' auxiliary fields in a pass of their own) and any attenuation
 * form; OT4 in two marches with sponge or no boundary and attenuation off / power 0 / None):
 * 64 x waves*rows cells per workgroup, marching `chunk` planes of the slowest
 * axis. version 7 (LDS ring of planes filled by direct-to-LDS loads two planes ahead, u read
 * once: axis-0 terms pushed into partial sums; 4 waves, rows 2) or 8 (as 7 with 8 waves; rows
 * 1, 2). Default: version 7, rows 2, chunk 24. Used by the autotuner (stride_amd.autotune).
 * Grids whose fields reach 2 GiB run the general kernels. */
int sa_set_tile(sa_solver *s, int rows, int chunk, int version);

/* Tile of the fused 3D update for the div(b grad) forms (constant or variable density; OT2,
 * sponge, no boundary or CFS-PML, any attenuation form), which replaces the staggered
 * derivative, divergence and update passes of the operator built by
 * stride/physics/iso_acoustic/devito.py:1466-1608, 1646-1666 for those schemes. The tile is
 * 64 (axis 2) x waves*rows (axis 1) cells marched over `chunk` planes of axis 0; rows 1 or 2
 * with 4 or 8 waves, rows 1 with 16 waves. rows = 0 switches the fused updates off, this one
 * and the one-launch 2D update (the separate general passes run instead). Default: rows 1,
 * chunk 60, waves 8. */
int sa_set_tile_density(sa_solver *s, int rows, int chunk, int waves);

/* A non-blocking stream of default priority on `device`, for sa_set_stream (several solvers of
 * one GPU on their own streams overlap their work; no reference counterpart: the reference runs
 * one Devito operator per worker process). */
int sa_stream_create(int device, void **stream);
int sa_stream_destroy(void *stream);

/* Kernel timing: when enabled, launches of the wavefield-update kernel are bracketed by HIP
 * events on the solver stream: every launch for enable = 1, every enable-th launch for
 * enable > 1 (a sample that leaves the other launches without events). sa_timing_read returns
 * the summed kernel milliseconds and the count of timed launches since the last reset. */
int sa_timing_enable(sa_solver *s, int enable);
int sa_timing_read(sa_solver *s, double *kernel_ms, int64_t *launches);

/* Device buffer helpers for callers without another allocator. */
int sa_device_alloc(sa_solver *s, int64_t bytes, void **out);
int sa_device_free(sa_solver *s, void *ptr);
int sa_memcpy(sa_solver *s, void *dst, const void *src, int64_t bytes, int kind /* 1 h2d, 2 d2h, 3 d2d */);
int sa_synchronize(sa_solver *s);

/* Two shots per pass. No counterpart in the reference, which runs one shot per worker call
 * (`IsoAcousticDevito.forward/adjoint`, stride/physics/iso_acoustic/devito.py:515-581, 875-920;
 * shots dispatched by stride/__init__.py:110-141, 325-334) and sums per-shot gradients in the
 * optimisation loop (stride/core.py:565-633). Two shots on the same grid, scheme, medium and
 * boundary advance together: their wavefields are interleaved per cell so one pass of the
 * specialised 3D update (3D, OT2, no density, sponge or no boundary, no attenuation; save mode
 * 0 or 1) serves both with one read of vp, and the adjoint accumulates the SUM of the two shots'
 * vp gradients and preconditioners (read with sa_read_gradient 0 / 1). With density (3D, OT2,
 * sponge or no boundary, attenuation power 0 / None / off; save mode 0 or 2) one pass of the
 * fused density update serves both with one read of b and the coefficients, and the adjoint
 * (variable density) sums the two shots' vp and rho gradients and preconditioners
 * (sa_read_gradient 0 .. 3; stride/physics/iso_acoustic/devito.py:1115-1150). Per-shot taps,
 * samples, traces and residuals. Other schemes return SA_E_UNSUPPORTED. */
int sa_batch_create(sa_solver *s, int nshot /* 2 */);
int sa_batch_set_taps(sa_solver *s, int shot, int nsrc, int src_ntap, const int64_t *src_index,
                      const float *src_weight, int nrec, int rec_ntap, const int64_t *rec_index,
                      const float *rec_weight);
/* src[k]: device (nt, nsrc_k) source samples of shot k; rec[k]: device (nt, nrec_k) traces */
int sa_batch_forward_dev(sa_solver *s, const float *const *src, float *const *rec);
/* res[k]: device (nt, nrec_k) adjoint source of shot k, as for sa_adjoint_dev */
int sa_batch_adjoint_dev(sa_solver *s, const float *const *res);
int sa_batch_get_wavefield(sa_solver *s, int shot, int which, float *out);
/* tile of the two-shot update: waves 8 with rows 1 or 2, or waves 4 with rows 2 (default rows 2,
 * chunk 24, waves 4) */
int sa_batch_set_tile(sa_solver *s, int rows, int chunk, int waves);
/* tile of the two-shot density update: rows 1 with waves 4 or 8 (default rows 1, chunk 60,
 * waves 4) */
int sa_batch_set_tile_density(sa_solver *s, int rows, int chunk, int waves);

/* Several 2D shots per launch (lockstep). No counterpart in the reference, which runs one shot
 * per worker call (IsoAcousticDevito.forward / adjoint, devito.py:515-581, 875-920; the shots of
 * an iteration are spread over workers by stride/__init__.py:117-146). Shots on the same 2D grid
 * and scheme (fused 2D update: OT2, any density form, sponge or CFS-PML, any attenuation; save
 * mode 0, 1 or 2; no wavefield dumps), each set up on its own solver v[k] (medium, taps,
 * nt), advance together: one launch per time step covers every shot's tiles. Each shot's traces,
 * saved wavefield and gradients are those of its one-shot run (sa_forward_dev / sa_adjoint_dev
 * on v[k]), and stay with v[k]. The work runs on v[0]'s stream after what is queued on each
 * solver's stream; each solver's stream waits for it. */
/* src[k], rec[k]: device (nt, nsrc_k) samples and (nt, nrec_k) traces of shot k */
int sa_shots2d_forward_dev(sa_solver *const *v, int n, const float *const *src, float *const *rec);
/* res[k]: device (nt, nrec_k) adjoint source of shot k; want: SA_WANT_VP | SA_WANT_RHO */
int sa_shots2d_adjoint_dev(sa_solver *const *v, int n, const float *const *res, int want);

/* Many shots of uneven duration and receiver count, forward only. No counterpart in the
 * reference, which runs one shot per worker call (IsoAcousticDevito.forward,
 * stride/physics/iso_acoustic/devito.py:515-581) over the shots of an iteration
 * (stride/__init__.py:117-146, runtime.async_for). All shots share the grid, scheme (the
 * specialised 3D update: OT2, no density, sponge or no boundary, no attenuation, save mode 0),
 * medium and boundary; each has its own nt (<= the grid's nt), source and receiver taps.
 * Shots are paired longest first; one launch per time step advances every pair still running
 * (its two shots interleaved per cell, vp read once per pair; the launch's tail is shared by all
 * its pairs), one launch injects and one interpolates for every shot still running. Shot k
 * runs its own steps 1 .. nt_k - 2 (run_forward with time bounds (0, nt_k)). */
int sa_multi_create(sa_solver *s, int nshot);
int sa_multi_set_shot(sa_solver *s, int shot, int nt, int nsrc, int src_ntap, const int64_t *src_index,
                      const float *src_weight, int nrec, int rec_ntap, const int64_t *rec_index,
                      const float *rec_weight);
/* src[k]: device (nt_k, nsrc_k) windowed samples of shot k; rec[k]: device (nt_k, nrec_k) traces.
 * graph != 0: the whole schedule is captured into one HIP graph and launched once (level clears
 * stay outside the graph); graph == 0: the same launches are issued one by one. Returns once
 * the work is queued on the solver's stream. */
int sa_multi_forward_dev(sa_solver *s, const float *const *src, float *const *rec, int graph);
/* wavefield `which` (0: last level, 1: the level before) of shot k's pair, shot k's half; the
 * shorter shot of a pair keeps being advanced until its partner's last step. */
int sa_multi_get_wavefield(sa_solver *s, int shot, int which, float *out);
/* capture + instantiation time of the last graph (milliseconds) */
int sa_multi_graph_ms(sa_solver *s, double *ms);

/* ---- Steps either side of the solver, on device arrays (the solver supplies device and
 * stream). Traces are (ntr, nt) float32 rows; fields are extended-grid float32 arrays in C
 * order. Reductions are deterministic. sa_field_axpb, sa_field_div_nonzero and
 * sa_field_descent are queued on the solver's stream and return at once (sa_synchronize waits
 * for them); the other calls return with their work finished. */

/* CheckTraces (stride/optimisation/pipelines/steps/check_traces.py:52-83): flags rows with a
 * NaN / Inf (bad_rows: host array of ntr ints, may be NULL), counts them, zeroes them if
 * zero_bad. */
int sa_traces_check(sa_solver *s, float *d, int ntr, int nt, int zero_bad, int *bad_rows, int *nbad);
/* MuteTraces (mute_traces.py:26-60): mod *= conv(obs != 0, table) (centred, zero outside;
 * table: host, odd length, e.g. the cosine low-pass of stride/utils/filters.py:401-465). */
int sa_traces_mute(sa_solver *s, float *mod, const float *obs, int ntr, int nt, int len, const float *table);
/* FIR along time: out[i] = sum_j w[j] d[i + j - off]; adjoint: sum_j w[j] d[i - j + off]
 * (cosine low-pass of FilterTraces, filters.py:414-465; table: host). In place. */
int sa_traces_fir(sa_solver *s, float *d, int ntr, int nt, int len, const float *table, int off, int adjoint);
/* Butterworth second-order sections as scipy.signal.sosfilt (filters.py:14-175 with
 * zero_phase False, filter_traces.py:88); adjoint runs backwards in time. sos: host, nsec x 6
 * (b0 b1 b2 1 a1 a2). In place. */
int sa_traces_sosfilt(sa_solver *s, float *d, int ntr, int nt, int nsec, const double *sos, int adjoint);
/* NormPerShot / NormPerTrace (norm_traces.py:41-86): d /= sqrt(mean over rows of sum d^2)
 * + 1e-31, or per row sqrt(sum d^2) + 1e-31; norm: the per-shot value. In place. */
int sa_traces_norm(sa_solver *s, float *d, int ntr, int nt, int per_trace, double *norm);
/* L2DistanceLoss (stride/optimisation/loss/l2_distance.py:36-66): res = mod - obs,
 * fun = 0.5 sum res^2. */
int sa_traces_l2(sa_solver *s, const float *mod, const float *obs, float *res, int ntr, int nt, double *fun);
/* sa_traces_l2 with no device allocation. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The per-row partial sums and their total live in a static
 * device array of the library (65537 doubles, 512 KiB, one copy per GPU, held until the
 * process ends), so ntr is at most 65536 (SA_E_ARG beyond). Calls that overlap on one GPU
 * share that array and race; the array keeps the last call's contents. Same results as
 * sa_traces_l2, bit for bit. */
int sa_traces_l2_noalloc(sa_solver *s, const float *mod, const float *obs, float *res, int ntr, int nt,
                         double *fun);
/* out[c, r] = in[r, c] x col_scale[c] (col_scale: host, length cols, or null). Converts the
 * solver's time-major receiver traces (nt, nrec) to Traces rows (nrec, nt) (data.py:1423-1430),
 * and back for an adjoint source with the Tukey window of devito.py:863-869 as col_scale. */
int sa_traces_transpose(sa_solver *s, const float *in, float *out, int rows, int cols, const float *col_scale);
/* MaskField (mask_field.py:8-67): f *= [lo <= c < hi] x the ramp-off profile of width ramp. */
int sa_field_mask(sa_solver *s, float *f, int dim, const int *shape, const int *lo, const int *hi, int ramp);
/* SmoothField / apply_prec smoothing (smooth_field.py:28-47, data.py:409-446):
 * scipy.ndimage.gaussian_filter (truncate 4), mode 0 nearest, 1 reflect. In place. */
int sa_field_gaussian(sa_solver *s, float *f, int dim, const int *shape, const double *sigma, int mode);
/* min, max and sum |f| (NormField, norm_field.py:24-52; apply_prec's normalisation). */
int sa_field_stats(sa_solver *s, const float *f, int64_t n, double *vmin, double *vmax, double *abssum);
int sa_field_axpb(sa_solver *s, float *f, int64_t n, double a, double b); /* f = a f + b */
/* y += a x (the sum of per-shot gradients that Stride's optimisation loop forms,
 * stride/core.py:565-633); queued on the stream */
int sa_field_axpy(sa_solver *s, float *y, const float *x, int64_t n, double a);
int sa_field_div_nonzero(sa_solver *s, float *g, const float *prec, int64_t n); /* g /= prec where != 0 */
/* GradientDescent step in the variable's transformed space (optimiser.py:242-330) with Clip
 * (clip.py): transform 0 identity, 1 slowness (x -> 1/x). */
int sa_field_descent(sa_solver *s, float *x, const float *dir, int64_t n, double step, int transform, int clip,
                     double lo, double hi);
/* Box [lo, hi) of a C-order grid of `shape` copied out to a dense array (crop), or a dense
 * array copied into that box of the grid (embed; the rest of the grid is left as it is): the
 * inner / extended views of a ScalarField (data.py:284-293). Queued on the stream. */
int sa_field_crop(sa_solver *s, const float *grid, float *box, int dim, const int *shape, const int *lo,
                  const int *hi);
int sa_field_embed(sa_solver *s, const float *box, float *grid, int dim, const int *shape, const int *lo,
                   const int *hi);
/* get_grad_vp's post-scale (devito.py:1076-1081): slowness 1 (vp's default transform):
 * grad /= vp, prec /= vp^2; else grad *= -2 / vp^3, prec *= (2 / vp)^9 (as written). grad or
 * prec may be NULL. Queued on the stream. */
int sa_field_grad_vp_finish(sa_solver *s, float *grad, float *prec, const float *vp, int64_t n, int slowness);
/* Gradient accumulator `which` (as sa_read_gradient) copied to an unpadded device array. */
int sa_read_gradient_dev(sa_solver *s, int which, float *dev_out);

/* ---- Learned gradient preconditioner: U-Net kernels (BF16 on the matrix cores, FP32
 * accumulation). No counterpart in the reference: the network maps the raw vp gradient of an
 * FWI iteration (get_grad_vp, stride/physics/iso_acoustic/devito.py:1045-1111) to an update
 * direction that is kept only if the forward solver's data misfit drops (the plain gradient
 * step of stride/optimisation/optimisers/optimiser.py:242-330 otherwise).
 * All buffers are device pointers; work is queued on `stream` (a hipStream_t, NULL the null
 * stream) and the calls return at once, except sa_nn_mse with a loss pointer, which waits.
 * Activations are NHWC bf16 (N images, H x W pixels, C channels innermost); 3x3 weights are
 * [ky][kx][co][ci] bf16; biases and gradients fp32. H % 4 == 0, W % 64 == 0, and the channel
 * counts of the 3x3 convolutions are multiples of 32. */

/* y = act(conv3x3(x, wt) + bias), zero padding; relu 0 / 1; bias may be NULL. amask (NULL or
 * shaped as x): x elements whose amask is <= 0 read as zero (backward data pass, with amask the
 * output of the layer whose gradient x is). */
int sa_nn_conv3x3(void *stream, const void *x, const void *amask, const void *wt, const float *bias, void *y,
                  int N, int H, int W, int Cin, int Cout, int relu);
/* wb[ky][kx][ci][co] = wt[2-ky][2-kx][co][ci]: the weights of the backward data pass */
int sa_nn_flip_weights(void *stream, const void *wt, void *wb, int Cin, int Cout);
/* dw[ky][kx][co][ci] (+)= sum over pixels of dy[p][co] [ymask[p][co] > 0] x[p + (ky-1, kx-1)][ci];
 * part: nsplit * 9 * Cin * Cout floats of scratch, nsplit = ceil(N H W / px_per_split),
 * px_per_split a multiple of 32. */
int sa_nn_wgrad3x3(void *stream, const void *dy, const void *ymask, const void *x, float *part, float *dw,
                   int N, int H, int W, int Cin, int Cout, int px_per_split, int accumulate);
/* db[c] (+)= sum_p dy[p][c] [ymask[p][c] > 0]; part: ceil(npx / px_per_block) * C floats */
int sa_nn_bias_grad(void *stream, const void *dy, const void *ymask, float *part, float *db, long long npx, int C,
                    int px_per_block, int accumulate);
/* 2 x 2 average pool and its backward (dx (+)= dy / 4 on each child) */
int sa_nn_avgpool(void *stream, const void *x, void *y, int N, int H, int W, int C);
int sa_nn_avgpool_back(void *stream, const void *dy, void *dx, int N, int H, int W, int C, int accumulate);
/* nearest x2 upsample of x (N,H,W,C) into channels [coff, coff + C) of y (N,2H,2W,Cy), and its
 * backward (dx = sum of the four children) */
int sa_nn_upsample(void *stream, const void *x, void *y, int N, int H, int W, int C, int Cy, int coff);
int sa_nn_upsample_back(void *stream, const void *dy, void *dx, int N, int H, int W, int C, int Cy, int coff);
/* dst[p][doff + c] (+)= src[p][soff + c], c < C: concatenation along channels and its split */
int sa_nn_copy_channels(void *stream, const void *src, int Cs, int soff, void *dst, int Cd, int doff, int C,
                        long long npx, int accumulate);
/* x[p][0] = bf16(s0 g0[p]), x[p][1] = bf16(s1 g1[p]) (zero if g1 is NULL), other channels zero:
 * the network input from one or two fp32 images */
int sa_nn_pack_input(void *stream, const float *g0, float s0, const float *g1, float s1, void *x, long long npx,
                     int C);
/* out[p] = sum_c w[c] x[p][c] + b[0] (fp32): the 1 x 1 output projection, and its backward:
 * dx = dout w masked by x > 0; dw[0..C-1] (+)= sum_p dout x, dw[C] (+)= sum_p dout;
 * part: ceil(npx / px_per_block) * (C + 1) floats */
int sa_nn_head(void *stream, const void *x, const float *w, const float *b, float *out, long long npx, int C);
int sa_nn_head_back(void *stream, const float *dout, const float *w, const void *x, void *dx, float *part, float *dw,
                    long long npx, int C, int px_per_block, int accumulate);
/* mean-square loss 0.5 sum (out - target)^2 inv_n (to *loss, a host pointer; NULL: no wait) and
 * d = (out - target) inv_n (may be NULL); part: ceil(n / 256) + 1 doubles of scratch */
int sa_nn_mse(void *stream, const float *out, const float *target, float *d, long long n, float inv_n, double *part,
              double *loss);
/* Adam step `step` (>= 1) on fp32 masters w with moments m, v; wb (may be NULL) gets bf16(w) */
int sa_nn_adam(void *stream, float *w, const float *g, float *m, float *v, void *wb, long long n, float lr, float b1,
               float b2, float eps, int step);
int sa_nn_to_bf16(void *stream, const float *a, void *b, long long n);
int sa_nn_from_bf16(void *stream, const void *a, float *b, long long n);

#ifdef __cplusplus
}
#endif
#endif
