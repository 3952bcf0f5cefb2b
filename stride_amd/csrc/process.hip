// Trace processing and gradient post-processing on the device: the steps either side of the
// wave solver in an FWI iteration (stride/optimisation/pipelines/default_pipelines.py:83-175,
// stride/optimisation/loss/l2_distance.py:36-66, stride/problem/data.py:409-446,
// stride/optimisation/optimisers/optimiser.py:242-330). Traces are (ntr, nt) float32 rows;
// fields are extended-grid float32 arrays in C order. Every reduction is deterministic: one
// block per row in a fixed tree order, rows combined in index order in double precision.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <vector>

#include "internal.h"

namespace {

constexpr int kT = 256;

#define PTRY(x)                                                                        \
    do {                                                                               \
        if ((x) != hipSuccess) return sa_fail(SA_E_HIP, "HIP call failed: " #x);        \
    } while (0)

struct Scratch {  // device scratch for one call
    void *p = nullptr;
    ~Scratch() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes) {
        if (hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            return sa_fail(SA_E_NOMEM, "device scratch allocation failed");
        }
        return SA_OK;
    }
};

// ---- row reductions -----------------------------------------------------------------------

__device__ double block_sum(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    return sh[0];
}

// out[r] = sum_t (a[r,t] - b[r,t])^2 (b may be null); optionally res = a - b
__global__ void k_row_sumsq(const float *__restrict__ a, const float *__restrict__ b, float *__restrict__ res,
                            int nt, double *__restrict__ out) {
    __shared__ double sh[kT];
    const long long base = (long long)blockIdx.x * nt;
    double acc = 0.0;
    for (int t = threadIdx.x; t < nt; t += kT) {
        const float d = b ? a[base + t] - b[base + t] : a[base + t];
        if (res) res[base + t] = d;
        acc += (double)d * d;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// sum of n doubles in a fixed order (thread i sums i, i + kT, ... then a fixed tree)
__global__ void k_ordered_sum(const double *__restrict__ v, int n, double *__restrict__ out) {
    __shared__ double sh[kT];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kT) acc += v[i];
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) *out = t;
}

__global__ void k_row_bad(const float *__restrict__ d, int nt, int *__restrict__ bad) {
    __shared__ int sh;
    if (threadIdx.x == 0) sh = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * nt;
    int any = 0;
    for (int t = threadIdx.x; t < nt; t += kT) any |= !isfinite(d[base + t]);
    if (any) atomicOr(&sh, 1);  // flag only: the result does not depend on the order
    __syncthreads();
    if (threadIdx.x == 0) bad[blockIdx.x] = sh;
}

__global__ void k_zero_rows(float *__restrict__ d, int nt, const int *__restrict__ bad) {
    if (!bad[blockIdx.x]) return;
    const long long base = (long long)blockIdx.x * nt;
    for (int t = threadIdx.x; t < nt; t += kT) d[base + t] = 0.f;
}

// d[r, :] *= scale[r] (per row) or scale[0] (all rows): 1 / (sqrt(sum) + 1e-31)
__global__ void k_scale_rows(float *__restrict__ d, int nt, const double *__restrict__ sums, int per_row, int nrow) {
    const double s = per_row ? sums[blockIdx.x] : sums[0] / nrow;
    const float f = (float)(1.0 / (std::sqrt(s) + 1e-31));
    const long long base = (long long)blockIdx.x * nt;
    for (int t = threadIdx.x; t < nt; t += kT) d[base + t] *= f;
}

// ---- filters ------------------------------------------------------------------------------

// out[r, i] = sum_j w[j] x[r, i + j - off] (zero outside); adjoint: sum_j w[j] x[r, i - j + off]
__global__ void k_fir(const float *__restrict__ x, float *__restrict__ out, int nt, const float *__restrict__ w,
                      int L, int off, int adjoint) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= nt) return;
    const long long base = (long long)blockIdx.y * nt;
    double acc = 0.0;
    for (int j = 0; j < L; ++j) {
        const int t = adjoint ? i - j + off : i + j - off;
        if (t >= 0 && t < nt) acc += (double)w[j] * x[base + t];
    }
    out[base + i] = (float)acc;
}

// mute: mod[r, i] *= sum_j w[j] [obs[r, i + j - c] != 0] (centred, zero outside)
__global__ void k_mute(float *__restrict__ mod, const float *__restrict__ obs, int nt, const float *__restrict__ w,
                       int L) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= nt) return;
    const long long base = (long long)blockIdx.y * nt;
    const int c = L / 2;
    double m = 0.0;
    for (int j = 0; j < L; ++j) {
        const int t = i + j - c;
        if (t >= 0 && t < nt && obs[base + t] != 0.f) m += (double)w[j];
    }
    mod[base + i] = (float)(m * mod[base + i]);
}

// Cascaded second-order sections (b0 b1 b2 a0 a1 a2, a0 = 1), transposed direct form II, in
// double precision, forward in time or (adjoint) backward. Systolic: a group of G lanes
// filters one row, lane k running section k (lanes k >= nsec pass their input through); at
// iteration n lane k works on sample n - k, taking lane k-1's output of iteration n-1, so the
// per-sample chain is one section instead of nsec (the hand-over is a DPP lane shift; the
// shift across group boundaries reaches lane 0 of the next group, which ignores it). Lane 0
// reads the row 16 samples ahead in
// registers; lane G-1 writes sample n - G + 1 (already read: the filter runs in place). Each
// section sees the same operations in the same order as a sequential cascade.
// value of the lane one below (DPP wave_shr:1; lane 0 gets 0), both halves of a double
__device__ __forceinline__ double wave_shr1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x138, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Cascaded second-order sections (b0 b1 b2 a0 a1 a2, a0 = 1), transposed direct form II, in
// double precision, forward in time or (adjoint) backward. Systolic: a group of G lanes
// filters one row, lane k running section k (lanes k >= nsec pass their input through); at
// iteration n lane k works on sample n - k, taking lane k-1's output of iteration n-1 (a DPP
// lane shift; across a group boundary it reaches lane 0 of the next group, which ignores it),
// so the per-sample chain is one section instead of nsec. The block stages its rows through
// LDS in chunks of C samples, loading chunk c+1 with coalesced reads while it filters chunk c;
// lane G-1 writes sample n - G + 1, already staged (the filter runs in place). Each section
// sees the same operations in the same order as a sequential cascade.
template <int G>
__global__ void __launch_bounds__(kT) k_sosfilt(float *__restrict__ d, int ntr, int nt,
                                                const double *__restrict__ sos, int nsec, int adjoint) {
    constexpr int RB = kT / G;        // rows per block
    constexpr int C = 128;            // samples per chunk
    constexpr int PER = RB * C / kT;  // staged samples per thread and chunk
    __shared__ float buf[2][RB][C + 1];
    const int k = threadIdx.x % G;
    const int lr = threadIdx.x / G;
    const int r = blockIdx.x * RB + lr;
    const bool live = r < ntr;
    const bool pass = k >= nsec;
    double b0 = 0, b1 = 0, b2 = 0, a1 = 0, a2 = 0;
    if (!pass) {
        const double *c = sos + 6 * k;
        b0 = c[0], b1 = c[1], b2 = c[2], a1 = c[4], a2 = c[5];
    }
    auto at = [&](int row, int t) { return (long long)row * nt + (adjoint ? nt - 1 - t : t); };
    float stage[PER];
    auto fetch = [&](int ch) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = threadIdx.x + kT * j, row = blockIdx.x * RB + e / C, t = ch * C + e % C;
            stage[j] = ((row < ntr) & (t < nt)) ? d[at(row, t)] : 0.f;
        }
    };
    auto put = [&](int b) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = threadIdx.x + kT * j;
            buf[b][e / C][e % C] = stage[j];
        }
    };
    fetch(0);
    put(0);
    __syncthreads();
    double z1 = 0.0, z2 = 0.0, y = 0.0;
    const int nch = (nt + G - 1 + C - 1) / C;
    for (int ch = 0; ch < nch; ++ch) {
        const int cb = ch & 1;
        fetch(ch + 1);
        for (int i0 = 0; i0 < C; i0 += 8) {
            float xs[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) xs[u] = buf[cb][lr][i0 + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int smp = ch * C + i0 + u - k;
                const double prev = wave_shr1(y);
                const double x = k == 0 ? (double)xs[u] : prev;
                const double o = b0 * x + z1;
                const double nz1 = b1 * x - a1 * o + z2;
                const double nz2 = b2 * x - a2 * o;
                const bool in = (smp >= 0) & (smp < nt);
                const bool upd = in & !pass;
                z1 = upd ? nz1 : z1;
                z2 = upd ? nz2 : z2;
                y = in ? (pass ? x : o) : y;
                if (live & (k == G - 1) & in) d[at(r, smp)] = (float)y;
            }
        }
        put(cb ^ 1);
        __syncthreads();
    }
}

// out[c, r] = in[r, c] * scale[c] (scale may be null): 64 x 64 tiles through LDS, reads and
// writes both along rows
__global__ void k_transpose(const float *__restrict__ in, float *__restrict__ out, int rows, int cols,
                            const float *__restrict__ scale) {
    __shared__ float tile[64][65];
    const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4 threads
    for (int k = ty; k < 64; k += 4) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = in[(long long)r * cols + c] * (scale ? scale[c] : 1.f);
    }
    __syncthreads();
    for (int k = ty; k < 64; k += 4) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) out[(long long)c * rows + r] = tile[tx][k];
    }
}

// ---- fields ------------------------------------------------------------------------------

struct Shape {
    int n[3];
};


// Field kernels take one thread per cell in C order (32-bit indices: fields below 2^31 cells).

// Mask of stride's MaskField (mask_field.py:8-67): 1 inside [lo, hi) x the ramp-off profile.
// The reference writes, per axis d (ascending) and index i < ramp, tab[i] on the two faces
// c[d] = i and c[d] = n_d - 1 - i restricted to i <= c[e] <= n_e - i on the other axes; later
// writes win. Axes with 2 ramp > n_d are skipped. A cell matches at most one index per axis,
// i = min(c[d], n_d - 1 - c[d]), so its value is tab[i] of the last axis whose write covers
// it, or 1.
constexpr int kMaskRamp = 1024;  // longest ramp-off (cells)

constexpr int kEl = 4;  // cells per thread in the field kernels below (strided by kT)

__device__ __forceinline__ void cell_coords(const Shape &s, unsigned p, int (&c)[3]) {
    const unsigned q = p / (unsigned)s.n[2];
    c[0] = (int)(q / (unsigned)s.n[1]);
    c[1] = (int)(q % (unsigned)s.n[1]);
    c[2] = (int)(p - q * (unsigned)s.n[2]);
}

__global__ void k_mask(float *__restrict__ f, Shape s, int dim, Shape lo, Shape hi, int ramp,
                       const float *__restrict__ tab, unsigned n) {
    __shared__ float lt[kMaskRamp];
    for (int i = threadIdx.x; i < ramp; i += kT) lt[i] = tab[i];
    const unsigned p0 = blockIdx.x * (kT * kEl) + threadIdx.x;
    float v[kEl];
#pragma unroll
    for (int u = 0; u < kEl; ++u) v[u] = p0 + u * kT < n ? f[p0 + u * kT] : 0.f;
    __syncthreads();
    const int off = 3 - dim;  // leading unit axes of a lower-dimensional field
#pragma unroll
    for (int u = 0; u < kEl; ++u) {
        const unsigned p = p0 + u * kT;
        if (p >= n) break;
        int c[3];
        cell_coords(s, p, c);
        bool inner = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) inner = inner & ((a < off) | ((c[a] >= lo.n[a]) & (c[a] < hi.n[a])));
        float m = 1.f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int i = min(c[d], s.n[d] - 1 - c[d]);
            bool hit = (d >= off) & (2 * ramp <= s.n[d]) & (i < ramp);
#pragma unroll
            for (int e = 0; e < 3; ++e)
                if (e != d) hit = hit & ((e < off) | ((c[e] >= i) & (c[e] <= s.n[e] - i)));
            if (hit) m = lt[i];
        }
        f[p] = v[u] * (inner ? m : 0.f);
    }
}

// One axis of a separable correlation with weights w[0 .. 2R] (centre R), boundary mode
// 0 nearest (edge value), 1 reflect (d c b a | a b c d | d c b a); the boundary index map is
// evaluated only within R of an edge. Blocks take contiguous runs of kT cells (whole cache
// lines, whatever the row length). Along axis 0 the 2R + 1 planes a block reads are shared with
// the blocks at the same place in the neighbouring planes: block b runs on XCD b % 8, and each
// XCD takes a contiguous run of (plane-fastest) chunks, so those blocks share its L2.
template <int RM, int EL>  // RM > 0: R <= RM, taps unrolled; RM = 0: any R. EL cells per thread
__global__ void k_axis_filter(const float *__restrict__ in, float *__restrict__ out, Shape s, int axis,
                              const double *__restrict__ w, int R, int mode, unsigned cpp, unsigned per_xcd) {
    constexpr int CH = kT * EL;  // cells per block
    const unsigned P = (unsigned)s.n[1] * (unsigned)s.n[2];
    const unsigned n = P * (unsigned)s.n[0];
    unsigned p0, lim;  // first cell of this thread, end of its run
    if (axis == 0) {
        const unsigned b = blockIdx.x;
        const unsigned L = (b & 7) * per_xcd + (b >> 3);
        const unsigned z = L % (unsigned)s.n[0], j = L / (unsigned)s.n[0];
        if (j >= cpp) return;
        p0 = z * P + j * CH + threadIdx.x;
        lim = z * P + P;
    } else {
        p0 = blockIdx.x * CH + threadIdx.x;
        lim = n;
    }
    const unsigned stride = axis == 2 ? 1u : (axis == 1 ? (unsigned)s.n[2] : P);
    const int len = s.n[axis];
    auto src = [&](int t) {  // boundary map of the line index t
        if (mode == 0) return t < 0 ? 0 : (t >= len ? len - 1 : t);
        const int period = 2 * len;
        t = ((t % period) + period) % period;
        return t >= len ? period - 1 - t : t;
    };
    if (RM > 0) {
        constexpr int NT = RM > 0 ? 2 * RM + 1 : 1;
        float v[EL][NT];
#pragma unroll
        for (int u = 0; u < EL; ++u) {  // every load of the thread first, then the sums
            const unsigned p = p0 + u * kT;
            const bool live = p < lim;
            int c[3];
            cell_coords(s, live ? p : 0u, c);
            const int ca = c[axis];
            const float *q0 = in + ((live ? p : 0u) - (unsigned)ca * stride);
            const bool inside = (ca >= R) & (ca + R < len);
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const int t = ca - R + j;
                v[u][j] = (live & (j <= 2 * R)) ? q0[(unsigned)(inside ? t : src(t)) * stride] : 0.f;
            }
        }
#pragma unroll
        for (int u = 0; u < EL; ++u) {
            const unsigned p = p0 + u * kT;
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < NT; ++j)
                if (j <= 2 * R) acc += w[j] * v[u][j];
            if (p < lim) out[p] = (float)acc;
        }
    } else {
        for (int u = 0; u < EL; ++u) {
            const unsigned p = p0 + u * kT;
            if (p >= lim) break;
            int c[3];
            cell_coords(s, p, c);
            const int ca = c[axis];
            const float *q0 = in + (p - (unsigned)ca * stride);
            double acc = 0.0;
            for (int j = -R; j <= R; ++j) acc += w[j + R] * q0[(unsigned)src(ca + j) * stride];
            out[p] = (float)acc;
        }
    }
}

// box of a grid, one thread per box cell in order: to_box 1: box[b] = grid[lo + b]; 0:
// grid[lo + b] = box[b]
__global__ void k_box_copy(float *__restrict__ grid, float *__restrict__ box, Shape s, Shape lo, Shape bn,
                           unsigned n, int to_box) {
    const unsigned b = blockIdx.x * kT + threadIdx.x;
    if (b >= n) return;
    const unsigned q = b / (unsigned)bn.n[2];
    const int c[3] = {(int)(q / (unsigned)bn.n[1]), (int)(q % (unsigned)bn.n[1]), (int)(b - q * (unsigned)bn.n[2])};
    const long long g = ((long long)(c[0] + lo.n[0]) * s.n[1] + (c[1] + lo.n[1])) * s.n[2] + (c[2] + lo.n[2]);
    if (to_box)
        box[b] = grid[g];
    else
        grid[g] = box[b];
}


// compat-mode vp gradient post-scale: slowness transform: grad /= vp, prec /= vp^2; else
// grad *= -2 / vp^3, prec *= (2 / vp)^9
__global__ void k_grad_vp_finish(float *__restrict__ grad, float *__restrict__ prec, const float *__restrict__ vp,
                                 long long n, int slowness) {
    const long long p = (long long)blockIdx.x * kT + threadIdx.x;
    if (p >= n) return;
    const float v = vp[p];
    if (slowness) {
        const float r = 1.f / v;
        if (grad) grad[p] *= r;
        if (prec) prec[p] *= r * r;
    } else {
        if (grad) grad[p] *= -2.f / (v * v * v);
        if (prec) {
            const float q = 2.f / v, q3 = q * q * q;
            prec[p] *= q3 * q3 * q3;
        }
    }
}

__global__ void k_partial_minmax_abssum(const float *__restrict__ f, long long n, float *__restrict__ pmin,
                                        float *__restrict__ pmax, double *__restrict__ pabs) {
    __shared__ float smin[kT], smax[kT];
    __shared__ double sabs[kT];
    float lo = INFINITY, hi = -INFINITY;
    double a = 0.0;
    const long long chunk = (n + gridDim.x - 1) / gridDim.x;  // contiguous, fixed split
    const long long p0 = (long long)blockIdx.x * chunk, p1 = p0 + chunk < n ? p0 + chunk : n;
    for (long long p = p0 + threadIdx.x; p < p1; p += kT) {
        const float v = f[p];
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
        a += fabs((double)v);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    sabs[threadIdx.x] = a;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + w]);
            sabs[threadIdx.x] += sabs[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pmin[blockIdx.x] = smin[0];
        pmax[blockIdx.x] = smax[0];
        pabs[blockIdx.x] = sabs[0];
    }
}

__global__ void k_axpb(float *__restrict__ f, long long n, float a, float b) {
    const long long p = (long long)blockIdx.x * kT + threadIdx.x;
    if (p < n) f[p] = a * f[p] + b;
}

__global__ void k_axpy(float *__restrict__ y, const float *__restrict__ x, long long n, float a) {
    const long long p = (long long)blockIdx.x * kT + threadIdx.x;
    if (p < n) y[p] += a * x[p];
}

// grad[p] *= 1 / prec[p] where prec[p] != 0
__global__ void k_div_nonzero(float *__restrict__ g, const float *__restrict__ prec, long long n) {
    const long long p = (long long)blockIdx.x * kT + threadIdx.x;
    if (p < n && prec[p] != 0.f) g[p] *= 1.f / prec[p];
}

// x <- inverse(transform(x) - step * dir), clipped; transform 0 identity, 1 slowness (1/x)
__global__ void k_descent(float *__restrict__ x, const float *__restrict__ dir, long long n, float step,
                          int transform, float lo, float hi, int clip) {
    const long long p = (long long)blockIdx.x * kT + threadIdx.x;
    if (p >= n) return;
    float v = x[p];
    if (transform == 1) {
        v = 1.f / (1.f / v - step * dir[p]);
    } else {
        v = v - step * dir[p];
    }
    if (clip) v = fminf(fmaxf(v, lo), hi);
    x[p] = v;
}

int blocks(long long n) { return (int)((n + kT - 1) / kT); }

int prep(sa_solver *s) {
    if (!s) return sa_fail(SA_E_ARG, "null solver");
    PTRY(hipSetDevice(sa_solver_device(s)));
    return SA_OK;
}

int rows_sumsq(sa_solver *s, const float *a, const float *b, float *res, int ntr, int nt, double *rows,
               double *total) {
    hipStream_t st = sa_solver_stream(s);
    k_row_sumsq<<<ntr, kT, 0, st>>>(a, b, res, nt, rows);
    k_ordered_sum<<<1, kT, 0, st>>>(rows, ntr, total);
    PTRY(hipGetLastError());
    return SA_OK;
}

}  // namespace

extern "C" int sa_traces_check(sa_solver *s, float *d, int ntr, int nt, int zero_bad, int *bad_rows, int *nbad) {
    int rc = prep(s);
    if (rc) return rc;
    if (!d || ntr < 0 || nt < 1) return sa_fail(SA_E_ARG, "bad trace arguments");
    if (nbad) *nbad = 0;
    if (ntr == 0) return SA_OK;
    Scratch sc;
    if ((rc = sc.alloc(ntr * sizeof(int)))) return rc;
    int *bad = (int *)sc.p;
    hipStream_t st = sa_solver_stream(s);
    k_row_bad<<<ntr, kT, 0, st>>>(d, nt, bad);
    if (zero_bad) k_zero_rows<<<ntr, kT, 0, st>>>(d, nt, bad);
    PTRY(hipGetLastError());
    std::vector<int> h(ntr);
    PTRY(hipMemcpyAsync(h.data(), bad, ntr * sizeof(int), hipMemcpyDeviceToHost, st));
    PTRY(hipStreamSynchronize(st));
    int count = 0;
    for (int r = 0; r < ntr; ++r) count += h[r] != 0;
    if (bad_rows)
        for (int r = 0; r < ntr; ++r) bad_rows[r] = h[r];
    if (nbad) *nbad = count;
    return SA_OK;
}

extern "C" int sa_traces_mute(sa_solver *s, float *mod, const float *obs, int ntr, int nt, int len,
                              const float *table) {
    int rc = prep(s);
    if (rc) return rc;
    if (!mod || !obs || !table || ntr < 0 || nt < 1 || len < 1 || len % 2 == 0)
        return sa_fail(SA_E_ARG, "bad mute arguments (odd table length)");
    if (ntr == 0) return SA_OK;
    Scratch sc;
    if ((rc = sc.alloc(len * sizeof(float)))) return rc;
    hipStream_t st = sa_solver_stream(s);
    PTRY(hipMemcpyAsync(sc.p, table, len * sizeof(float), hipMemcpyHostToDevice, st));
    k_mute<<<dim3(blocks(nt), ntr), kT, 0, st>>>(mod, obs, nt, (const float *)sc.p, len);
    PTRY(hipGetLastError());
    PTRY(hipStreamSynchronize(st));
    return SA_OK;
}

extern "C" int sa_traces_fir(sa_solver *s, float *d, int ntr, int nt, int len, const float *table, int off,
                             int adjoint) {
    int rc = prep(s);
    if (rc) return rc;
    if (!d || !table || ntr < 0 || nt < 1 || len < 1) return sa_fail(SA_E_ARG, "bad filter arguments");
    if (ntr == 0) return SA_OK;
    Scratch sc;
    const size_t nb = (size_t)ntr * nt * sizeof(float);
    if ((rc = sc.alloc(nb + len * sizeof(float)))) return rc;
    float *tmp = (float *)sc.p, *w = tmp + (size_t)ntr * nt;
    hipStream_t st = sa_solver_stream(s);
    PTRY(hipMemcpyAsync(w, table, len * sizeof(float), hipMemcpyHostToDevice, st));
    PTRY(hipMemcpyAsync(tmp, d, nb, hipMemcpyDeviceToDevice, st));
    k_fir<<<dim3(blocks(nt), ntr), kT, 0, st>>>(tmp, d, nt, w, len, off, adjoint);
    PTRY(hipGetLastError());
    PTRY(hipStreamSynchronize(st));
    return SA_OK;
}

extern "C" int sa_traces_sosfilt(sa_solver *s, float *d, int ntr, int nt, int nsec, const double *sos,
                                 int adjoint) {
    int rc = prep(s);
    if (rc) return rc;
    if (!d || !sos || ntr < 0 || nt < 1 || nsec < 1 || nsec > 32) return sa_fail(SA_E_ARG, "bad sos arguments");
    if (ntr == 0) return SA_OK;
    Scratch sc;
    if ((rc = sc.alloc(6 * nsec * sizeof(double)))) return rc;
    hipStream_t st = sa_solver_stream(s);
    PTRY(hipMemcpyAsync(sc.p, sos, 6 * nsec * sizeof(double), hipMemcpyHostToDevice, st));
    const double *c = (const double *)sc.p;
    const int G = nsec <= 8 ? 8 : (nsec <= 16 ? 16 : 32);  // lanes per row
    const dim3 g((ntr + kT / G - 1) / (kT / G));
    if (G == 8)
        k_sosfilt<8><<<g, kT, 0, st>>>(d, ntr, nt, c, nsec, adjoint);
    else if (G == 16)
        k_sosfilt<16><<<g, kT, 0, st>>>(d, ntr, nt, c, nsec, adjoint);
    else
        k_sosfilt<32><<<g, kT, 0, st>>>(d, ntr, nt, c, nsec, adjoint);
    PTRY(hipGetLastError());
    PTRY(hipStreamSynchronize(st));
    return SA_OK;
}

extern "C" int sa_traces_norm(sa_solver *s, float *d, int ntr, int nt, int per_trace, double *norm) {
    int rc = prep(s);
    if (rc) return rc;
    if (!d || ntr < 1 || nt < 1) return sa_fail(SA_E_ARG, "bad norm arguments");
    Scratch sc;
    if ((rc = sc.alloc((ntr + 1) * sizeof(double)))) return rc;
    double *rows = (double *)sc.p, *total = rows + ntr;
    if ((rc = rows_sumsq(s, d, nullptr, nullptr, ntr, nt, rows, total))) return rc;
    hipStream_t st = sa_solver_stream(s);
    k_scale_rows<<<ntr, kT, 0, st>>>(d, nt, per_trace ? rows : total, per_trace, ntr);
    PTRY(hipGetLastError());
    double t = 0.0;
    PTRY(hipMemcpyAsync(&t, total, sizeof(double), hipMemcpyDeviceToHost, st));
    PTRY(hipStreamSynchronize(st));
    if (norm) *norm = std::sqrt(t / ntr);
    return SA_OK;
}

extern "C" int sa_traces_l2(sa_solver *s, const float *mod, const float *obs, float *res, int ntr, int nt,
                            double *fun) {
    int rc = prep(s);
    if (rc) return rc;
    if (!mod || !obs || !res || ntr < 1 || nt < 1) return sa_fail(SA_E_ARG, "bad misfit arguments");
    Scratch sc;
    if ((rc = sc.alloc((ntr + 1) * sizeof(double)))) return rc;
    double *rows = (double *)sc.p, *total = rows + ntr;
    if ((rc = rows_sumsq(s, mod, obs, res, ntr, nt, rows, total))) return rc;
    hipStream_t st = sa_solver_stream(s);
    double t = 0.0;
    PTRY(hipMemcpyAsync(&t, total, sizeof(double), hipMemcpyDeviceToHost, st));
    PTRY(hipStreamSynchronize(st));
    if (fun) *fun = 0.5 * t;
    return SA_OK;
}

// ---- misfit with no device allocation ------------------------------------------------------
// sa_traces_l2_noalloc keeps the per-row partial sums and their total in `scratch`, a
// file-scope device array (one copy on every GPU the library runs on, held until the process
// ends), instead of a buffer allocated per call. Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol): the device address of each GPU's
// copy is looked up once per device, with that device current, and kept in scratch_dev[].
constexpr int kNoAllocRows = 1 << 16;  // at most 65536 trace rows: 512 KiB of doubles
static __device__ __align__(256) double scratch[kNoAllocRows + 1];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double *scratch_dev[kMaxDevices] = {};

// device address of this GPU's copy of `scratch` (device must be current)
int scratch_for(int device, double **out) {
    if (device < 0 || device >= kMaxDevices) return sa_fail(SA_E_ARG, "device ordinal out of range for scratch");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return sa_fail(SA_E_HIP, "hipGetSymbolAddress(scratch) failed");
        scratch_dev[device] = static_cast<double *>(p);
    }
    *out = scratch_dev[device];
    return SA_OK;
}

}  // namespace

extern "C" int sa_traces_l2_noalloc(sa_solver *s, const float *mod, const float *obs, float *res, int ntr, int nt,
                                    double *fun) {
    int rc = prep(s);
    if (rc) return rc;
    if (!mod || !obs || !res || ntr < 1 || nt < 1) return sa_fail(SA_E_ARG, "bad misfit arguments");
    if (ntr > kNoAllocRows) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "sa_traces_l2_noalloc: %d trace rows exceed the maximum of %d", ntr,
                      kNoAllocRows);
        return sa_fail(SA_E_ARG, msg);
    }
    double *rows = nullptr;
    if ((rc = scratch_for(sa_solver_device(s), &rows))) return rc;
    // every row sum and the total are written before they are read: nothing to clear
    double *total = rows + ntr;
    if ((rc = rows_sumsq(s, mod, obs, res, ntr, nt, rows, total))) return rc;
    hipStream_t st = sa_solver_stream(s);
    double t = 0.0;
    PTRY(hipMemcpyAsync(&t, total, sizeof(double), hipMemcpyDeviceToHost, st));
    PTRY(hipStreamSynchronize(st));
    if (fun) *fun = 0.5 * t;
    return SA_OK;
}

extern "C" int sa_traces_transpose(sa_solver *s, const float *in, float *out, int rows, int cols,
                                   const float *col_scale) {
    int rc = prep(s);
    if (rc) return rc;
    if (!in || !out || rows < 0 || cols < 0) return sa_fail(SA_E_ARG, "bad transpose arguments");
    if (in == out && rows > 1 && cols > 1) return sa_fail(SA_E_ARG, "transpose needs distinct arrays");
    if (!rows || !cols) return SA_OK;
    Scratch sc;
    const float *w = nullptr;
    hipStream_t st = sa_solver_stream(s);
    if (col_scale) {
        if ((rc = sc.alloc(cols * sizeof(float)))) return rc;
        PTRY(hipMemcpyAsync(sc.p, col_scale, cols * sizeof(float), hipMemcpyHostToDevice, st));
        w = (const float *)sc.p;
    }
    k_transpose<<<dim3((cols + 63) / 64, (rows + 63) / 64), kT, 0, st>>>(in, out, rows, cols, w);
    PTRY(hipGetLastError());
    PTRY(hipStreamSynchronize(st));
    return SA_OK;
}

extern "C" int sa_field_mask(sa_solver *s, float *f, int dim, const int *shape, const int *lo, const int *hi,
                             int ramp) {
    int rc = prep(s);
    if (rc) return rc;
    if (!f || !shape || !lo || !hi || dim < 1 || dim > 3 || ramp < 0) return sa_fail(SA_E_ARG, "bad mask arguments");
    Shape sh{{1, 1, 1}}, l{{0, 0, 0}}, h{{1, 1, 1}};
    for (int a = 0; a < dim; ++a) {
        sh.n[3 - dim + a] = shape[a];
        l.n[3 - dim + a] = lo[a];
        h.n[3 - dim + a] = hi[a];
    }
    if (ramp > kMaskRamp) return sa_fail(SA_E_ARG, "mask ramp longer than 1024 cells");
    if (ramp < 2) ramp = 0;  // one-cell ramps divide by zero in the reference: no ramp
    std::vector<float> tab(std::max(ramp, 1), 1.f);
    for (int i = 0; i < ramp; ++i) {  // mask_field.py:25-27 in double, stored as float32
        const double pos = std::fabs((double)(ramp - i - 1) / (double)(ramp - 1));
        tab[i] = (float)(1.0 - std::cos(M_PI / 2.0 * (1.0 - pos)));
    }
    Scratch sc;
    if ((rc = sc.alloc(tab.size() * sizeof(float)))) return rc;
    hipStream_t st = sa_solver_stream(s);
    PTRY(hipMemcpyAsync(sc.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, st));
    const long long n = (long long)sh.n[0] * sh.n[1] * sh.n[2];
    if (n >= (1LL << 31)) return sa_fail(SA_E_ARG, "field larger than 2^31 cells");
    k_mask<<<(n + kT * kEl - 1) / (kT * kEl), kT, 0, st>>>(f, sh, dim, l, h, ramp, (const float *)sc.p, (unsigned)n);
    PTRY(hipGetLastError());
    PTRY(hipStreamSynchronize(sa_solver_stream(s)));
    return SA_OK;
}

static int box_args(int dim, const int *shape, const int *lo, const int *hi, Shape &s, Shape &l, Shape &bn) {
    if (!shape || !lo || !hi || dim < 1 || dim > 3) return sa_fail(SA_E_ARG, "bad box arguments");
    s = Shape{{1, 1, 1}};
    l = Shape{{0, 0, 0}};
    bn = Shape{{1, 1, 1}};
    for (int a = 0; a < dim; ++a) {
        const int k = 3 - dim + a;
        if (shape[a] < 1 || lo[a] < 0 || hi[a] > shape[a] || lo[a] > hi[a])
            return sa_fail(SA_E_ARG, "box outside the grid");
        s.n[k] = shape[a];
        l.n[k] = lo[a];
        bn.n[k] = hi[a] - lo[a];
    }
    if ((long long)bn.n[0] * bn.n[1] * bn.n[2] >= (1LL << 31)) return sa_fail(SA_E_ARG, "box larger than 2^31 cells");
    return SA_OK;
}

extern "C" int sa_field_crop(sa_solver *s, const float *grid, float *box, int dim, const int *shape, const int *lo,
                             const int *hi) {
    int rc = prep(s);
    if (rc) return rc;
    Shape sh, l, bn;
    if (!grid || !box) return sa_fail(SA_E_ARG, "null array");
    if ((rc = box_args(dim, shape, lo, hi, sh, l, bn))) return rc;
    const long long n = (long long)bn.n[0] * bn.n[1] * bn.n[2];
    if (n) k_box_copy<<<blocks(n), kT, 0, sa_solver_stream(s)>>>((float *)grid, box, sh, l, bn, (unsigned)n, 1);
    PTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_field_embed(sa_solver *s, const float *box, float *grid, int dim, const int *shape, const int *lo,
                              const int *hi) {
    int rc = prep(s);
    if (rc) return rc;
    Shape sh, l, bn;
    if (!grid || !box) return sa_fail(SA_E_ARG, "null array");
    if ((rc = box_args(dim, shape, lo, hi, sh, l, bn))) return rc;
    const long long n = (long long)bn.n[0] * bn.n[1] * bn.n[2];
    if (n) k_box_copy<<<blocks(n), kT, 0, sa_solver_stream(s)>>>(grid, (float *)box, sh, l, bn, (unsigned)n, 0);
    PTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_field_grad_vp_finish(sa_solver *s, float *grad, float *prec, const float *vp, int64_t n,
                                       int slowness) {
    int rc = prep(s);
    if (rc) return rc;
    if (!vp || n < 0) return sa_fail(SA_E_ARG, "bad gradient arguments");
    if (n) k_grad_vp_finish<<<blocks(n), kT, 0, sa_solver_stream(s)>>>(grad, prec, vp, n, slowness);
    PTRY(hipGetLastError());
    return SA_OK;
}

// separable Gaussian smoothing as scipy.ndimage.gaussian_filter (truncate 4): per axis the
// weights exp(-x^2 / (2 sigma^2)), x = -R .. R, R = int(4 sigma + 0.5), normalised. Axis
// passes alternate between the field and a scratch copy.
extern "C" int sa_field_gaussian(sa_solver *s, float *f, int dim, const int *shape, const double *sigma, int mode) {
    int rc = prep(s);
    if (rc) return rc;
    if (!f || !shape || !sigma || dim < 1 || dim > 3 || mode < 0 || mode > 1)
        return sa_fail(SA_E_ARG, "bad smoothing arguments");
    Shape sh{{1, 1, 1}};
    for (int a = 0; a < dim; ++a) {
        if (shape[a] < 1) return sa_fail(SA_E_ARG, "bad smoothing arguments");
        sh.n[3 - dim + a] = shape[a];
    }
    const long long n = (long long)sh.n[0] * sh.n[1] * sh.n[2];
    if (n >= (1LL << 31)) return sa_fail(SA_E_ARG, "field larger than 2^31 cells");
    std::vector<double> h;  // weights of every smoothed axis, back to back
    int axis[3], R[3], at[3], npass = 0;
    for (int a = 0; a < dim; ++a) {
        const double sg = sigma[a];
        if (!(sg > 0)) continue;
        const int r = (int)(4.0 * sg + 0.5);
        if (2 * r + 1 > 4096) return sa_fail(SA_E_ARG, "smoothing radius too large");
        at[npass] = (int)h.size();
        double sum = 0.0;
        for (int j = -r; j <= r; ++j) {
            h.push_back(std::exp(-0.5 * (double)j * j / (sg * sg)));
            sum += h.back();
        }
        for (int j = at[npass]; j < (int)h.size(); ++j) h[j] /= sum;
        axis[npass] = 3 - dim + a;
        R[npass++] = r;
    }
    if (!npass) return SA_OK;
    const long long nf = (n + 1) / 2 * 2;  // keeps the weights 8-byte aligned
    Scratch sc;
    if ((rc = sc.alloc(nf * sizeof(float) + h.size() * sizeof(double)))) return rc;
    float *tmp = (float *)sc.p;
    double *w = (double *)(tmp + nf);
    hipStream_t st = sa_solver_stream(s);
    PTRY(hipMemcpyAsync(w, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st));
    float *src = f, *dst = tmp;
    for (int k = 0; k < npass; ++k) {
        const int el = R[k] <= 2 ? 4 : (R[k] <= 4 ? 2 : 1);  // cells per thread
        const unsigned P = (unsigned)sh.n[1] * sh.n[2], cpp = (P + kT * el - 1) / (kT * el);
        const unsigned nblk = axis[k] == 0 ? cpp * sh.n[0] : (unsigned)((n + kT * el - 1) / (kT * el));
        const unsigned per_xcd = (nblk + 7) / 8;
        const unsigned nb = axis[k] == 0 ? per_xcd * 8 : nblk;
        const double *wk = w + at[k];
        if (R[k] <= 2)
            k_axis_filter<2, 4><<<nb, kT, 0, st>>>(src, dst, sh, axis[k], wk, R[k], mode, cpp, per_xcd);
        else if (R[k] <= 4)
            k_axis_filter<4, 2><<<nb, kT, 0, st>>>(src, dst, sh, axis[k], wk, R[k], mode, cpp, per_xcd);
        else
            k_axis_filter<0, 1><<<nb, kT, 0, st>>>(src, dst, sh, axis[k], wk, R[k], mode, cpp, per_xcd);
        std::swap(src, dst);
    }
    PTRY(hipGetLastError());
    if (src != f) PTRY(hipMemcpyAsync(f, src, n * sizeof(float), hipMemcpyDeviceToDevice, st));
    PTRY(hipStreamSynchronize(st));  // the host weights and the scratch live until here
    return SA_OK;
}

// min, max and sum of |f| (deterministic: fixed contiguous split, fixed tree, ordered combine)
extern "C" int sa_field_stats(sa_solver *s, const float *f, int64_t n, double *vmin, double *vmax, double *abssum) {
    int rc = prep(s);
    if (rc) return rc;
    if (!f || n < 1) return sa_fail(SA_E_ARG, "bad field arguments");
    const int nb = (int)std::min<int64_t>(4096, (n + kT - 1) / kT);
    Scratch sc;
    if ((rc = sc.alloc(nb * (2 * sizeof(float) + sizeof(double))))) return rc;
    double *pabs = (double *)sc.p;
    float *pmin = (float *)(pabs + nb), *pmax = pmin + nb;
    hipStream_t st = sa_solver_stream(s);
    k_partial_minmax_abssum<<<nb, kT, 0, st>>>(f, n, pmin, pmax, pabs);
    PTRY(hipGetLastError());
    std::vector<float> lo(nb), hi(nb);
    std::vector<double> ab(nb);
    PTRY(hipMemcpyAsync(lo.data(), pmin, nb * sizeof(float), hipMemcpyDeviceToHost, st));
    PTRY(hipMemcpyAsync(hi.data(), pmax, nb * sizeof(float), hipMemcpyDeviceToHost, st));
    PTRY(hipMemcpyAsync(ab.data(), pabs, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    PTRY(hipStreamSynchronize(st));
    double a = 0.0, l = INFINITY, h = -INFINITY;
    for (int b = 0; b < nb; ++b) {
        a += ab[b];
        l = std::min<double>(l, lo[b]);
        h = std::max<double>(h, hi[b]);
    }
    if (vmin) *vmin = l;
    if (vmax) *vmax = h;
    if (abssum) *abssum = a;
    return SA_OK;
}

extern "C" int sa_field_axpb(sa_solver *s, float *f, int64_t n, double a, double b) {
    int rc = prep(s);
    if (rc) return rc;
    if (!f || n < 0) return sa_fail(SA_E_ARG, "bad field arguments");
    if (n) k_axpb<<<blocks(n), kT, 0, sa_solver_stream(s)>>>(f, n, (float)a, (float)b);
    PTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_field_axpy(sa_solver *s, float *y, const float *x, int64_t n, double a) {
    int rc = prep(s);
    if (rc) return rc;
    if (!y || !x || n < 0) return sa_fail(SA_E_ARG, "bad field arguments");
    if (n) k_axpy<<<blocks(n), kT, 0, sa_solver_stream(s)>>>(y, x, n, (float)a);
    PTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_field_div_nonzero(sa_solver *s, float *g, const float *prec, int64_t n) {
    int rc = prep(s);
    if (rc) return rc;
    if (!g || !prec || n < 0) return sa_fail(SA_E_ARG, "bad field arguments");
    if (n) k_div_nonzero<<<blocks(n), kT, 0, sa_solver_stream(s)>>>(g, prec, n);
    PTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_field_descent(sa_solver *s, float *x, const float *dir, int64_t n, double step, int transform,
                                int clip, double lo, double hi) {
    int rc = prep(s);
    if (rc) return rc;
    if (!x || !dir || n < 0 || transform < 0 || transform > 1) return sa_fail(SA_E_ARG, "bad update arguments");
    if (n) k_descent<<<blocks(n), kT, 0, sa_solver_stream(s)>>>(x, dir, n, (float)step, transform, (float)lo,
                                                               (float)hi, clip);
    PTRY(hipGetLastError());
    return SA_OK;
}
