// Trace processing and gradient post-processing on the device: the steps either side of the
// wave solver in an FWI iteration (stride/optimisation/pipelines/default_pipelines.py:83-175,
// stride/optimisation/loss/l2_distance.py:36-66, stride/problem/data.py:409-446,
// stride/optimisation/optimisers/optimiser.py:242-330). Traces are (ntr, nt) float32 rows;
// fields are extended-grid float32 arrays in C order. Every reduction is deterministic: one
// block per row in a fixed tree order, rows combined in index order in double precision.

#include <hip/hip_runtime.h>

#include <cmath>
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

__global__ void k_ordered_sum(const double *__restrict__ v, int n, double *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += v[i];
        *out = s;
    }
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

// cascaded second-order sections (b0 b1 b2 a0 a1 a2, a0 = 1), transposed direct form II, in
// double precision; one lane per row, forward in time or (adjoint) backward
__global__ void k_sosfilt(float *__restrict__ d, int ntr, int nt, const double *__restrict__ sos, int nsec,
                          int adjoint) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ntr) return;
    double z1[32], z2[32];
    for (int k = 0; k < nsec; ++k) z1[k] = z2[k] = 0.0;
    const long long base = (long long)r * nt;
    for (int n = 0; n < nt; ++n) {
        const int t = adjoint ? nt - 1 - n : n;
        double x = d[base + t];
        for (int k = 0; k < nsec; ++k) {
            const double *c = sos + 6 * k;
            const double y = c[0] * x + z1[k];
            z1[k] = c[1] * x - c[4] * y + z2[k];
            z2[k] = c[2] * x - c[5] * y;
            x = y;
        }
        d[base + t] = (float)x;
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

__device__ __forceinline__ void unflat(const Shape &s, long long p, int (&c)[3]) {
    c[2] = (int)(p % s.n[2]);
    const long long q = p / s.n[2];
    c[1] = (int)(q % s.n[1]);
    c[0] = (int)(q / s.n[1]);
}

// mask of stride's MaskField (mask_field.py:8-67): 1 inside [lo, hi) x ramp-off profile
// assigned per axis, index and side in the reference's loop order (later writes win)
__global__ void k_mask(float *__restrict__ f, Shape s, int dim, Shape lo, Shape hi, int ramp) {
    const long long n = (long long)s.n[0] * s.n[1] * s.n[2];
    const long long p = (long long)blockIdx.x * kT + threadIdx.x;
    if (p >= n) return;
    int c[3];
    unflat(s, p, c);
    const int off = 3 - dim;  // leading unit axes of a lower-dimensional field
    bool inner = true;
    for (int a = off; a < 3; ++a) inner = inner && c[a] >= lo.n[a] && c[a] < hi.n[a];
    float m = 1.f;
    for (int d = off; d < 3; ++d) {
        if (2 * ramp > s.n[d]) continue;
        for (int index = 0; index < ramp; ++index) {
            const float pos = fabsf((float)(ramp - index - 1) / (float)(ramp - 1));
            const float val = 1.f - cosf(3.14159265358979f / 2.f * (1.f - pos));
            bool others = true;
            for (int e = off; e < 3; ++e)
                if (e != d) others = others && c[e] >= index && c[e] <= s.n[e] - index;
            if (!others) continue;
            if (c[d] == index) m = val;
            if (c[d] == s.n[d] - index - 1) m = val;
        }
    }
    f[p] *= inner ? m : 0.f;
}

// one axis of a separable correlation with weights w[0 .. 2R] (centre R), boundary mode
// 0 nearest (edge value), 1 reflect (d c b a | a b c d | d c b a)
__global__ void k_axis_filter(const float *__restrict__ in, float *__restrict__ out, Shape s, int axis,
                              const double *__restrict__ w, int R, int mode) {
    const long long n = (long long)s.n[0] * s.n[1] * s.n[2];
    const long long p = (long long)blockIdx.x * kT + threadIdx.x;
    if (p >= n) return;
    int c[3];
    unflat(s, p, c);
    const long long stride = axis == 2 ? 1 : (axis == 1 ? s.n[2] : (long long)s.n[1] * s.n[2]);
    const int len = s.n[axis];
    double acc = 0.0;
    for (int j = -R; j <= R; ++j) {
        int t = c[axis] + j;
        if (mode == 0) {
            t = t < 0 ? 0 : (t >= len ? len - 1 : t);
        } else {
            const int period = 2 * len;
            t = ((t % period) + period) % period;
            if (t >= len) t = period - 1 - t;
        }
        acc += w[j + R] * in[p + (long long)(t - c[axis]) * stride];
    }
    out[p] = (float)acc;
}

// box of a grid: to_box 1: box[b] = grid[lo + b]; 0: grid[lo + b] = box[b]
__global__ void k_box_copy(float *__restrict__ grid, float *__restrict__ box, Shape s, Shape lo, Shape bn,
                           int to_box) {
    const long long n = (long long)bn.n[0] * bn.n[1] * bn.n[2];
    const long long b = (long long)blockIdx.x * kT + threadIdx.x;
    if (b >= n) return;
    int c[3];
    unflat(bn, b, c);
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
    k_ordered_sum<<<1, 1, 0, st>>>(rows, ntr, total);
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
    k_sosfilt<<<(ntr + 63) / 64, 64, 0, st>>>(d, ntr, nt, (const double *)sc.p, nsec, adjoint);
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
    const long long n = (long long)sh.n[0] * sh.n[1] * sh.n[2];
    k_mask<<<blocks(n), kT, 0, sa_solver_stream(s)>>>(f, sh, dim, l, h, ramp < 2 ? 0 : ramp);
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
    if (n) k_box_copy<<<blocks(n), kT, 0, sa_solver_stream(s)>>>((float *)grid, box, sh, l, bn, 1);
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
    if (n) k_box_copy<<<blocks(n), kT, 0, sa_solver_stream(s)>>>(grid, (float *)box, sh, l, bn, 0);
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
// weights exp(-x^2 / (2 sigma^2)), x = -R .. R, R = int(4 sigma + 0.5), normalised
extern "C" int sa_field_gaussian(sa_solver *s, float *f, int dim, const int *shape, const double *sigma, int mode) {
    int rc = prep(s);
    if (rc) return rc;
    if (!f || !shape || !sigma || dim < 1 || dim > 3 || mode < 0 || mode > 1)
        return sa_fail(SA_E_ARG, "bad smoothing arguments");
    Shape sh{{1, 1, 1}};
    for (int a = 0; a < dim; ++a) sh.n[3 - dim + a] = shape[a];
    const long long n = (long long)sh.n[0] * sh.n[1] * sh.n[2];
    Scratch sc;
    const long long nf = (n + 1) / 2 * 2;  // keeps the weights 8-byte aligned
    if ((rc = sc.alloc(nf * sizeof(float) + 4096 * sizeof(double)))) return rc;
    float *tmp = (float *)sc.p;
    double *w = (double *)(tmp + nf);
    hipStream_t st = sa_solver_stream(s);
    for (int a = 0; a < dim; ++a) {
        const double sg = sigma[a];
        if (!(sg > 0)) continue;
        const int R = (int)(4.0 * sg + 0.5);
        if (2 * R + 1 > 4096) return sa_fail(SA_E_ARG, "smoothing radius too large");
        std::vector<double> h(2 * R + 1);
        double sum = 0.0;
        for (int j = -R; j <= R; ++j) sum += (h[j + R] = std::exp(-0.5 * (double)j * j / (sg * sg)));
        for (double &v : h) v /= sum;
        PTRY(hipMemcpyAsync(w, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st));
        k_axis_filter<<<blocks(n), kT, 0, st>>>(f, tmp, sh, 3 - dim + a, w, R, mode);
        PTRY(hipMemcpyAsync(f, tmp, n * sizeof(float), hipMemcpyDeviceToDevice, st));
        PTRY(hipStreamSynchronize(st));  // the host weights are reused for the next axis
    }
    PTRY(hipGetLastError());
    return SA_OK;
}

// min, max and sum of |f| (deterministic: fixed contiguous split, fixed tree, ordered combine)
extern "C" int sa_field_stats(sa_solver *s, const float *f, int64_t n, double *vmin, double *vmax, double *abssum) {
    int rc = prep(s);
    if (rc) return rc;
    if (!f || n < 1) return sa_fail(SA_E_ARG, "bad field arguments");
    const int nb = (int)std::min<int64_t>(1024, (n + kT - 1) / kT);
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
