// stride_amd solver: C ABI (include/stride_amd.h), host orchestration of the time
// loops, and the general kernels covering every scheme option.
//
// The equations are those Stride hands to Devito for IsoAcousticDevito
// (stride/physics/iso_acoustic/devito.py:1466-1608); the time loops follow
// run_forward / run_adjoint (devito.py:515-581, 875-920). Every time step is
//
//   out = A u + B other + G (K u [+ dt^2/12 K(m K u)]) + Cbeta (lap u - lap other)
//         - Cpml sum_d u1_d ,                                               (1)
//
// with per-point coefficients built once per shot by k_coefficients:
//   a = dt (alpha vp [power 0] + vp^2 D [sponge]),  e = dt^2 vp^2 D^2 [sponge],
//   D = 7 dt vp s(x) (sponge layer rule, boundaries/boundary.py:59-171),
//   W = alpha [power None] else 1,  ia = 1 / (1 + a),
//   A = W (2 - e) ia,  B = -W (1 - a) ia,  Cq = W dt^2 ia,  m = vp^2 [rho],  G = Cq m,
//   Cbeta = Cq 2 alpha vp^3 / dt [power 2],  Cpml = Cq vp^2 [PML].
// This is the solve() of p.dt2 = L + vp^2 (attenuation) - vp^2 (boundary term)
// for p[t+1] (forward, other = p[t-1]) or p[t-1] (adjoint, other = p[t+1]); the
// time-reversed form is the same because the adjoint uses the transposed centred
// first time derivative (boundaries/devito.py:65-67, devito.py:1545).
// K is the SO10 Laplacian (no density), or sum_d D-_d (bbar_d D+_d) with SO10
// staggered derivatives (constant density: bbar = 1; variable density: bbar the
// half-point mean of the buoyancy).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/stride_amd.h"
#include "kernels.h"
#include "fast3d.h"
#include "rho3d.h"
#include "flat2d.h"

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(x)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail(SA_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

#define CHECK_LAUNCH() HIP_TRY(hipGetLastError())

// SO10 weights (Fornberg closed forms)
static const double kW2[6] = {-5269.0 / 1800.0, 5.0 / 3.0, -5.0 / 21.0, 5.0 / 126.0, -5.0 / 1008.0, 1.0 / 3150.0};
static const double kW1[6] = {0.0, 5.0 / 6.0, -5.0 / 21.0, 5.0 / 84.0, -5.0 / 504.0, 1.0 / 1260.0};
static const double kWS[5] = {19845.0 / 16384.0, -735.0 / 8192.0, 567.0 / 40960.0, -405.0 / 229376.0,
                              35.0 / 294912.0};

// ------------------------------------------------------------------------------------------
// kernels

#define DOMAIN_INDEX                                         \
    int k = blockIdx.x * blockDim.x + threadIdx.x;           \
    int j = blockIdx.y * blockDim.y + threadIdx.y;           \
    int i = blockIdx.z;                                      \
    if (k >= g.n[2] || j >= g.n[1] || i >= g.n[0]) return; \
    long long p = sa_pidx(g, i, j, k);

// per-point step coefficients (1)
__global__ void k_coefficients(Geo g, const float *__restrict__ vp, const float *__restrict__ rho,
                               const float *__restrict__ alpha, const float *__restrict__ tables, int tlen,
                               int boundary, int att_power, int rho_mode, float dt, float *A, float *B, float *G,
                               float *M, float *Cbeta, float *Cpml) {
    DOMAIN_INDEX
    float v = vp[p];
    float v2 = v * v;
    float m = rho_mode == 2 ? v2 * rho[p] : v2;
    float a = 0.f, e = 0.f, W = 1.f;
    if (att_power == 0) a += dt * alpha[p] * v;
    if (boundary == 0) {
        // sponge layer rule: axis d contributes its layer value at distance l from the
        // outer edge when every other axis is at least l cells from its edges
        int x[3] = {i, j, k};
        int dist[3];
        for (int d = 0; d < 3; ++d) dist[d] = g.act[d] ? min(x[d], g.n[d] - 1 - x[d]) : 0x7fffffff;
        float s = 0.f;
        for (int d = 0; d < 3; ++d) {
            if (!g.act[d] || g.nabs[d] == 0) continue;
            int other = 0x7fffffff;
            for (int e2 = 0; e2 < 3; ++e2)
                if (e2 != d) other = min(other, dist[e2]);
            int l0 = x[d], l1 = g.n[d] - 1 - x[d];
            if (l0 < g.nabs[d] && other >= l0) s += tables[d * tlen + l0];
            if (l1 < g.nabs[d] && other >= l1) s += tables[d * tlen + l1];
        }
        float D = (7.f * (s * v)) * dt;
        a += dt * v2 * D;
        e = dt * dt * v2 * D * D;
    }
    if (att_power == -1) W = alpha[p];
    float ia = 1.f / (1.f + a);
    A[p] = W * (2.f - e) * ia;
    B[p] = -W * (1.f - a) * ia;
    float cq = W * (dt * dt) * ia;
    G[p] = cq * m;
    M[p] = m;
    if (Cbeta) Cbeta[p] = cq * 2.f * alpha[p] * v2 * v / dt;
    if (Cpml) Cpml[p] = cq * v2;
}

// out = K(u * uscale) for the Laplacian form
__global__ void k_laplace(Geo g, const float *__restrict__ u, const float *__restrict__ uscale,
                          float *__restrict__ out) {
    DOMAIN_INDEX
    if (!uscale) {
        out[p] = sa_lap(g, u, p);
        return;
    }
    float acc = 0.f;
    for (int a = 0; a < 3; ++a) {
        if (!g.act[a]) continue;
        long long s = sa_stride(g, a);
        float v = g.w2[a][0] * u[p] * uscale[p];
        for (int r = 1; r <= 5; ++r)
            v += g.w2[a][r] * (u[p + r * s] * uscale[p + r * s] + u[p - r * s] * uscale[p - r * s]);
        acc += v;
    }
    out[p] = acc;
}

// st = bbar * D+ (u * uscale) over the domain widened by 5 cells along axis a
__global__ void k_stagger(Geo g, int a, const float *__restrict__ u, const float *__restrict__ uscale,
                          const float *__restrict__ b, float *__restrict__ st) {
    int lo[3] = {0, 0, 0};
    lo[a] = -5;
    int ext[3] = {g.n[0], g.n[1], g.n[2]};
    ext[a] += 10;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int i = blockIdx.z;
    if (k >= ext[2] || j >= ext[1] || i >= ext[0]) return;
    long long q = sa_pidx(g, i + lo[0], j + lo[1], k + lo[2]);
    long long s = sa_stride(g, a);
    float v;
    if (!uscale) {
        v = sa_dplus(g, u, q, a);
    } else {
        v = 0.f;
        for (int r = 0; r < 5; ++r)
            v += g.ws[a][r] * (u[q + (r + 1) * s] * uscale[q + (r + 1) * s] - u[q - r * s] * uscale[q - r * s]);
    }
    if (b) v *= 0.5f * (b[q] + b[q + s]);
    st[q] = v;
}

__global__ void k_divergence(Geo g, const float *__restrict__ st0, const float *__restrict__ st1,
                             const float *__restrict__ st2, float *__restrict__ out) {
    DOMAIN_INDEX
    float acc = 0.f;
    if (g.act[0]) acc += sa_dminus(g, st0, p, 0);
    if (g.act[1]) acc += sa_dminus(g, st1, p, 1);
    if (g.act[2]) acc += sa_dminus(g, st2, p, 2);
    out[p] = acc;
}

struct UpdateArgs {
    const float *u, *other;
    float *out;
    const float *A, *B, *G, *Ku, *KmKu, *Cbeta, *Cpml;
    const float *lapother;  // Laplacian of `other`, taken before this launch overwrites it
    const float *u1[3];
    float c4;  // dt^2 / 12 (OT4)
    float *save;
    int save_mode;  // 0 none, 1 dt2, 2 raw current level
    Box box;
    float idt2;
};

__global__ void k_update(Geo g, UpdateArgs x) {
    DOMAIN_INDEX
    float u = x.u[p], o = x.other[p];
    float q = x.Ku[p];
    if (x.KmKu) q += x.c4 * x.KmKu[p];
    float acc = x.G[p] * q;
    if (x.Cbeta) acc += x.Cbeta[p] * (sa_lap(g, x.u, p) - x.lapother[p]);
    if (x.Cpml) {
        float s1 = 0.f;
        for (int a = 0; a < 3; ++a)
            if (x.u1[a]) s1 += x.u1[a][p];
        acc -= x.Cpml[p] * s1;
    }
    float out = x.A[p] * u + x.B[p] * o + acc;
    x.out[p] = out;
    if (x.save && sa_inbox(x.box, i, j, k)) {
        long long b = sa_boxidx(x.box, i, j, k);
        x.save[b] = x.save_mode == 1 ? (out - 2.f * u + o) * x.idt2 : u;
    }
}

// CFS-PML auxiliary update of axis a at cell p, x its coordinate along a, inside a slab of a
// (boundaries/devito.py:133-171)
__device__ __forceinline__ void pml_cell(const Geo &g, int a, const float *__restrict__ F, long long p, int x,
                                         float *u1, float *u2, float *u3, const float *__restrict__ sig,
                                         const float *__restrict__ alp, const float *__restrict__ dsig,
                                         const float *__restrict__ dalp, float dt, int backward) {
    float s = sig[x], al = alp[x], ds = dsig[x], da = dalp[x];
    float dF = sa_d1(g, F, p, a), d2F = sa_d2(g, F, p, a);
    float v3, v2, v1;
    if (!backward) {
        v3 = u3[p] - dt * ((al + s) * u3[p] - s * s * (ds + da) * dF);
        v2 = u2[p] - dt * ((al + s) * u2[p] - s * (2.f * ds * dF + da * dF + s * d2F) + v3);
        v1 = u1[p] - dt * ((al + s) * u1[p] - (ds * dF + 2.f * s * d2F) + v2);
    } else {
        v3 = u3[p] - dt * ((al + s) * u3[p] + s * s * (ds + da) * dF);
        v2 = u2[p] - dt * ((al + s) * u2[p] + s * (2.f * ds * dF + da * dF) - s * s * d2F + v3);
        v1 = u1[p] - dt * ((al + s) * u1[p] + ds * dF - 2.f * s * d2F + v2);
    }
    u3[p] = v3;
    u2[p] = v2;
    u1[p] = v1;
}

// CFS-PML auxiliary update on the two slabs of axis a
__global__ void k_pml(Geo g, int a, const float *__restrict__ F, float *u1, float *u2, float *u3,
                      const float *__restrict__ sig, const float *__restrict__ alp, const float *__restrict__ dsig,
                      const float *__restrict__ dalp, float dt, int backward) {
    DOMAIN_INDEX
    int x = a == 0 ? i : (a == 1 ? j : k);
    if (!(x < g.nabs[a] || x >= g.n[a] - g.nabs[a])) return;
    pml_cell(g, a, F, p, x, u1, u2, u3, sig, alp, dsig, dalp, dt, backward);
}

// The CFS-PML fields of every active axis (3D fused updates)
struct PmlSet {
    float *u1[3], *u2[3], *u3[3];
    const float *prof[3][4];  // sigma, alpha, dsigma, dalpha per axis
    const float *Cpml;
};

__device__ __forceinline__ bool sa_in_slab(const Geo &g, int a, int x) {
    return x < g.nabs[a] || x >= g.n[a] - g.nabs[a];
}

// Cells of the absorbing layers (the union of the slabs), one wave per row (i, j) of the domain
// (kSlabRows rows per workgroup): every cell of a row inside an axis-0 or axis-1 slab, else the
// cells of the two axis-2 slabs (cell e of the row's list is k = slab_k(r, e)).
constexpr int kSlabRows = 4;
struct SlabRow {
    int i, j, cnt, nz, na;
    bool whole;
};

__device__ __forceinline__ bool slab_row(const Geo &g, SlabRow &r) {
    const int row = blockIdx.x * kSlabRows + threadIdx.y;
    r.i = row / g.n[1];
    r.j = row - r.i * g.n[1];
    if (r.i >= g.n[0]) return false;
    r.nz = g.n[2];
    r.na = g.nabs[2];
    r.whole = sa_in_slab(g, 0, r.i) || sa_in_slab(g, 1, r.j) || 2 * r.na >= r.nz;
    r.cnt = r.whole ? r.nz : 2 * r.na;
    return true;
}

__device__ __forceinline__ int slab_k(const SlabRow &r, int e) {
    return r.whole ? e : (e < r.na ? e : r.nz - 2 * r.na + e);
}

// every active axis's CFS-PML update on its slabs (the separate per-axis k_pml launches, in one)
__global__ void k_pml_slabs(Geo g, const float *__restrict__ F, PmlSet ps, float dt, int backward) {
    SlabRow r;
    if (!slab_row(g, r)) return;
    for (int e = threadIdx.x; e < r.cnt; e += blockDim.x) {
        const int k = slab_k(r, e);
        const long long p = sa_pidx(g, r.i, r.j, k);
        const int xs[3] = {r.i, r.j, k};
        for (int a = 0; a < 3; ++a)
            if (g.act[a] && sa_in_slab(g, a, xs[a]))
                pml_cell(g, a, F, p, xs[a], ps.u1[a], ps.u2[a], ps.u3[a], ps.prof[a][0], ps.prof[a][1],
                         ps.prof[a][2], ps.prof[a][3], dt, backward);
    }
}

// the PML term of (1) after a fused update that left it out: out -= Cpml sum_d u1_d on the slabs
// (zero elsewhere: u1_d is only written on the slabs of d)
__global__ void k_pml_correct(Geo g, PmlSet ps, float *out) {
    SlabRow r;
    if (!slab_row(g, r)) return;
    for (int e = threadIdx.x; e < r.cnt; e += blockDim.x) {
        const long long p = sa_pidx(g, r.i, r.j, slab_k(r, e));
        float s1 = 0.f;
        for (int a = 0; a < 3; ++a)
            if (ps.u1[a]) s1 += ps.u1[a][p];
        out[p] -= ps.Cpml[p] * s1;
    }
}

// padded field -> unpadded C-order array of the domain
__global__ void k_unpad(Geo g, const float *__restrict__ src, float *__restrict__ dst) {
    DOMAIN_INDEX
    dst[((long long)i * g.n[1] + j) * g.n[2] + k] = src[p];
}

__global__ void k_scale_copy(Geo g, const float *__restrict__ src, const float *__restrict__ scale,
                             float *__restrict__ dst) {
    DOMAIN_INDEX
    dst[p] = scale ? scale[p] * src[p] : src[p];
}

// sparse injection, deterministic: one thread per destination cell sums its contributions in
// a fixed order. dest: padded index of each destination cell, ijk: its domain coordinates
// (3 ints); coef per contribution; values[point].
// Fields of S interleaved shots hold shot sh of cell p at p * S + sh (S = 1: one shot).
__global__ void k_inject(int ndest, const long long *__restrict__ dest, const int *__restrict__ ijk,
                         const int *__restrict__ off, const int *__restrict__ point, const float *__restrict__ coef,
                         const float *__restrict__ values, float sign, float *field, float *save, Box box,
                         float save_scale, int S, int sh) {
    int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ndest) return;
    float acc = 0.f;
    for (int c = off[d]; c < off[d + 1]; ++c) acc += coef[c] * values[point[c]];
    acc *= sign;
    field[dest[d] * S + sh] += acc;
    if (save) {
        const int i = ijk[3 * d], j = ijk[3 * d + 1], k = ijk[3 * d + 2];
        if (sa_inbox(box, i, j, k)) save[sa_boxidx(box, i, j, k) * S + sh] += acc * save_scale;
    }
}

// rec[r] = sign * sum_q w * field (receiver interpolation, wavelet gradient); idx: padded
// index per tap, negative for none
__global__ void k_gather(int npt, int ntap, const long long *__restrict__ idx, const float *__restrict__ w,
                         const float *__restrict__ field, float sign, float *__restrict__ out, int S, int sh) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= npt) return;
    float acc = 0.f;
    for (int q = 0; q < ntap; ++q) {
        const long long p = idx[(long long)r * ntap + q];
        if (p < 0) continue;
        acc += w[(long long)r * ntap + q] * field[p * S + sh];
    }
    out[r] = sign * acc;
}

// both shots of a two-shot pass in one launch (blockIdx.y = shot; interleaved fields, S = 2)
struct InjectSet {
    int ndest;
    const long long *dest;
    const int *ijk, *off, *point;
    const float *coef, *values;
};
struct GatherSet {
    int npt, ntap;
    const long long *idx;
    const float *w;
    float *out;
};

__global__ void k_inject2(InjectSet s0, InjectSet s1, float sign, float *field, float *save, Box box,
                          float save_scale) {
    const int sh = blockIdx.y;
    const InjectSet q = sh ? s1 : s0;
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= q.ndest) return;
    float acc = 0.f;
    for (int c = q.off[d]; c < q.off[d + 1]; ++c) acc += q.coef[c] * q.values[q.point[c]];
    acc *= sign;
    field[q.dest[d] * 2 + sh] += acc;
    if (save) {
        const int i = q.ijk[3 * d], j = q.ijk[3 * d + 1], k = q.ijk[3 * d + 2];
        if (sa_inbox(box, i, j, k)) save[sa_boxidx(box, i, j, k) * 2 + sh] += acc * save_scale;
    }
}

__global__ void k_gather2(GatherSet s0, GatherSet s1, const float *__restrict__ field, float sign) {
    const int sh = blockIdx.y;
    const GatherSet q = sh ? s1 : s0;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= q.npt) return;
    float acc = 0.f;
    for (int t = 0; t < q.ntap; ++t) {
        const long long p = q.idx[(long long)r * q.ntap + t];
        if (p < 0) continue;
        acc += q.w[(long long)r * q.ntap + t] * field[p * 2 + sh];
    }
    q.out[r] = sign * acc;
}

// the multi-shot schedule: one shot's injection and interpolation, its pair and the half of the
// pair's interleaved cells it owns. Sets are ordered longest shot first, so the shots still
// running at a step are a prefix (blockIdx.y).
struct MInject {
    int ndest, nsrc;
    const long long *dest;
    const int *off, *point;
    const float *coef;
    const float *values;  // (nt, nsrc) samples of this shot
    int pair, half;
};
struct MGather {
    int npt, ntap;
    const long long *idx;
    const float *w;
    float *out;  // (nt, npt) traces of this shot
    int pair, half;
};

// the level written at step t (the new one): parity = t % 2 reads lev[parity], writes lev[1 - parity]
__device__ __forceinline__ float *written_level(const PairLevels &pl, int parity) {
    return parity ? pl.lev0 : pl.lev1;
}

__global__ void k_inject_multi(const MInject *__restrict__ sets, const PairLevels *__restrict__ pairs, int t,
                               int parity) {
    const MInject q = sets[blockIdx.y];
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= q.ndest) return;
    const float *v = q.values + (size_t)t * q.nsrc;
    float acc = 0.f;
    for (int c = q.off[d]; c < q.off[d + 1]; ++c) acc += q.coef[c] * v[q.point[c]];
    written_level(pairs[q.pair], parity)[q.dest[d] * 2 + q.half] += acc;
}

__global__ void k_gather_multi(const MGather *__restrict__ sets, const PairLevels *__restrict__ pairs, int t,
                               int parity) {
    const MGather q = sets[blockIdx.y];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= q.npt) return;
    const float *field = written_level(pairs[q.pair], parity);
    float acc = 0.f;
    for (int k = 0; k < q.ntap; ++k) {
        const long long p = q.idx[(long long)r * q.ntap + k];
        if (p < 0) continue;
        acc += q.w[(long long)r * q.ntap + k] * field[p * 2 + q.half];
    }
    q.out[(size_t)t * q.npt + r] = acc;
}

// padded copy of an unpadded extended-grid array, edges replicated into the halo (the device
// form of the host padding in sa_set_medium); lo: leading halo per internal axis, P: padded sizes
__global__ void k_pad_edge(Geo g, const float *__restrict__ src, float *__restrict__ dst, int lo0, int lo1, int lo2,
                           int P1, int P2) {
    const int a2 = blockIdx.x * blockDim.x + threadIdx.x;
    const int a1 = blockIdx.y * blockDim.y + threadIdx.y;
    const int a0 = blockIdx.z;
    if (a2 >= P2 || a1 >= P1) return;
    const int i = min(max(a0 - lo0, 0), g.n[0] - 1);
    const int j = min(max(a1 - lo1, 0), g.n[1] - 1);
    const int k = min(max(a2 - lo2, 0), g.n[2] - 1);
    dst[(long long)a0 * g.s0 + (long long)a1 * g.s1 + a2] = src[((long long)i * g.n[1] + j) * g.n[2] + k];
}

__global__ void k_reciprocal(const float *__restrict__ x, float *__restrict__ y, long long n) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) y[q] = 1.f / x[q];
}

// complex arithmetic in double precision for the complex-step derivative of the step
// coefficients (every operation is analytic, so Im f(x + i h) / h is f'(x) to rounding)
struct cplx {
    double r, i;
};
__device__ __forceinline__ cplx operator+(cplx a, cplx b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cplx operator-(cplx a, cplx b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cplx operator*(cplx a, cplx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cplx operator*(double a, cplx b) { return {a * b.r, a * b.i}; }
__device__ __forceinline__ cplx inv(cplx a) {
    const double d = a.r * a.r + a.i * a.i;
    return {a.r / d, -a.i / d};
}

// A, B, G, m, b of equation (1) for one point (the same formulas as k_coefficients)
__device__ void step_coefficients(cplx v, cplx rho, cplx alpha, double prof, int boundary, int att_power,
                                  int rho_mode, double dt, cplx out[5]) {
    const cplx one = {1.0, 0.0}, zero = {0.0, 0.0};
    cplx m = v * v;
    if (rho_mode == 2) m = m * rho;
    cplx a = zero, e = zero, W = one;
    if (att_power == 0) a = a + dt * (alpha * v);
    if (boundary == 0) {
        const cplx D = (7.0 * dt * prof) * v;
        a = a + dt * (v * v * D);
        e = (dt * dt) * (v * v * D * D);
    }
    if (att_power == -1) W = alpha;
    const cplx ia = inv(one + a);
    out[0] = W * ((cplx{2.0, 0.0} - e) * ia);
    out[1] = -1.0 * (W * ((one - a) * ia));
    out[2] = (dt * dt) * (W * ia * m);
    out[3] = m;
    out[4] = rho_mode == 2 ? inv(rho) : zero;
}

// dJ/dvp, dJ/drho, dJ/dalpha from the sensitivities to A, B, G, m, b (unpadded extended arrays,
// gm / gb may be null) and the medium (padded); alpha_scale converts the alpha derivative to
// the caller's units. Any output may be null.
__global__ void k_exact_chain(Geo g, const float *__restrict__ gA, const float *__restrict__ gB,
                              const float *__restrict__ gG, const float *__restrict__ gm, const float *__restrict__ gb,
                              const float *__restrict__ vp, const float *__restrict__ rho,
                              const float *__restrict__ alpha, const float *__restrict__ tables, int tlen,
                              int boundary, int att_power, int rho_mode, double dt, double alpha_scale,
                              float *__restrict__ o_vp, float *__restrict__ o_rho, float *__restrict__ o_alpha) {
    DOMAIN_INDEX
    const long long q = ((long long)i * g.n[1] + j) * g.n[2] + k;
    double prof = 0.0;
    if (boundary == 0) {  // the sponge layer profile without the velocity factor (k_coefficients' rule)
        int x[3] = {i, j, k};
        int dist[3];
        for (int d = 0; d < 3; ++d) dist[d] = g.act[d] ? min(x[d], g.n[d] - 1 - x[d]) : 0x7fffffff;
        float sum = 0.f;
        for (int d = 0; d < 3; ++d) {
            if (!g.act[d] || g.nabs[d] == 0) continue;
            int other = 0x7fffffff;
            for (int e2 = 0; e2 < 3; ++e2)
                if (e2 != d) other = min(other, dist[e2]);
            int l0 = x[d], l1 = g.n[d] - 1 - x[d];
            if (l0 < g.nabs[d] && other >= l0) sum += tables[d * tlen + l0];
            if (l1 < g.nabs[d] && other >= l1) sum += tables[d * tlen + l1];
        }
        prof = sum;
    }
    const double h = 1e-30;
    const double v0 = vp[p], r0 = rho_mode == 2 ? rho[p] : 1.0, a0 = att_power != -2 ? alpha[p] : 0.0;
    const double w[5] = {gA[q], gB[q], gG[q], gm ? gm[q] : 0.0, gb ? gb[q] : 0.0};
    cplx c[5];
    for (int which = 0; which < 3; ++which) {
        float *out = which == 0 ? o_vp : (which == 1 ? o_rho : o_alpha);
        if (!out) continue;
        const cplx v = {v0, which == 0 ? h : 0.0}, r = {r0, which == 1 ? h : 0.0}, al = {a0, which == 2 ? h : 0.0};
        step_coefficients(v, r, al, prof, boundary, att_power, rho_mode, dt, c);
        double t = 0.0;
        for (int n = 0; n < 5; ++n) t += w[n] * (c[n].i / h);
        out[q] = (float)(which == 2 ? t * alpha_scale : t);
    }
}

// source taps: dJ/dvp at each tap cell (the injection coefficient w scale vp[cell]) and
// dJ/dsrc (nt, nsrc); one thread, taps and time in order (deterministic; taps of different
// points may share a cell). lam (nt, nsrc, ntap): the adjoint field at each tap; src (nt, nsrc).
__global__ void k_source_grads(int nt, int nsrc, int ntap, const float *__restrict__ lam,
                               const float *__restrict__ src, const long long *__restrict__ pidx,
                               const long long *__restrict__ lin, const float *__restrict__ w,
                               const float *__restrict__ vp, double scale, float *__restrict__ o_vp,
                               float *__restrict__ o_src) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int sp = 0; sp < nsrc; ++sp)
        for (int q = 0; q < ntap; ++q) {
            const int c = sp * ntap + q;
            if (pidx[c] < 0) continue;
            double acc = 0.0;
            for (int t = 0; t < nt; ++t) acc += (double)lam[((long long)t * nsrc + sp) * ntap + q] * src[(long long)t * nsrc + sp];
            if (o_vp) o_vp[lin[c]] += (float)(acc * w[c] * scale);
        }
    if (!o_src) return;
    for (int t = 0; t < nt; ++t)
        for (int sp = 0; sp < nsrc; ++sp) {
            double acc = 0.0;
            for (int q = 0; q < ntap; ++q) {
                const int c = sp * ntap + q;
                if (pidx[c] < 0) continue;
                acc += (double)lam[((long long)t * nsrc + sp) * ntap + q] * w[c] * scale * vp[pidx[c]];
            }
            o_src[(long long)t * nsrc + sp] = (float)acc;
        }
}

// per-tap values field / G (exact adjoint: lambda at each source tap); idx padded
__global__ void k_gather_taps(int nq, const long long *__restrict__ idx, const float *__restrict__ field,
                              const float *__restrict__ G, float *__restrict__ out) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const long long p = idx[q];
    out[q] = p < 0 ? 0.f : field[p] / G[p];
}

// injection coefficients per contribution: 0 source (w 2 dt^k vp / h), 1 compat adjoint
// (w dt^2 vp^2), 2 exact adjoint (w G); pidx: padded index per contribution
__global__ void k_tap_coef(int nnz, const long long *__restrict__ pidx, const float *__restrict__ w,
                           const float *__restrict__ vp, const float *__restrict__ G, int kind, float scale,
                           float dt2, float *__restrict__ out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nnz) return;
    const long long p = pidx[c];
    float v = vp[p];
    if (kind == 0) out[c] = w[c] * scale * v;
    else if (kind == 1) out[c] = w[c] * dt2 * v * v;
    else out[c] = w[c] * G[p];
}

// snapshot (box layout) -> padded work field (zero outside the box)
__global__ void k_unpack(Geo g, Box box, const float *__restrict__ snap, float *__restrict__ dst) {
    DOMAIN_INDEX
    dst[p] = sa_inbox(box, i, j, k) ? snap[sa_boxidx(box, i, j, k)] : 0.f;
}

// padded field -> snapshot (box layout)
__global__ void k_pack(Geo g, Box box, const float *__restrict__ src, float *__restrict__ snap) {
    DOMAIN_INDEX
    if (sa_inbox(box, i, j, k)) snap[sa_boxidx(box, i, j, k)] = src[p];
}

// snapshot (box layout) <-> contiguous array over the region r (grid coordinates inside the box)
__global__ void k_region(Box box, Box r, float *snap, float *flat, int to_flat) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int i = blockIdx.z;
    if (k >= r.n[2] || j >= r.n[1] || i >= r.n[0]) return;
    long long q = ((long long)i * r.n[1] + j) * r.n[2] + k;
    long long b = sa_boxidx(box, r.lo[0] + i, r.lo[1] + j, r.lo[2] + k);
    if (to_flat)
        flat[q] = snap[b];
    else
        snap[b] = flat[q];
}

// compat vp gradient on the interior (prepare_grad_vp, devito.py:985-1024)
__global__ void k_grad_vp(Geo g, Box box, const float *__restrict__ ps, const float *__restrict__ psm,
                          const float *__restrict__ psp, float fdt2, const float *__restrict__ pa, float *grad,
                          float *prec) {
    DOMAIN_INDEX
    if (i < g.nabs[0] || i >= g.n[0] - g.nabs[0] || j < g.nabs[1] || j >= g.n[1] - g.nabs[1] || k < g.nabs[2] ||
        k >= g.n[2] - g.nabs[2])
        return;
    if (!sa_inbox(box, i, j, k)) return;
    long long b = sa_boxidx(box, i, j, k);
    float pdt = psm ? (psp[b] - 2.f * ps[b] + psm[b]) * fdt2 : ps[b];
    grad[p] += pdt * pa[p];
    prec[p] += pdt * pdt;
}

// compat rho gradient (prepare_grad_rho, devito.py:1115-1150): work = padded raw snapshot
__global__ void k_grad_rho(Geo g, const float *__restrict__ work, const float *__restrict__ b,
                           const float *__restrict__ pa, float *grad, float *prec) {
    DOMAIN_INDEX
    if (i < g.nabs[0] || i >= g.n[0] - g.nabs[0] || j < g.nabs[1] || j >= g.n[1] - g.nabs[1] || k < g.nabs[2] ||
        k >= g.n[2] - g.nabs[2])
        return;
    float acc = 0.f;
    for (int a = 0; a < 3; ++a)
        if (g.act[a]) acc += sa_dminus(g, b, p, a) * sa_dminus(g, work, p, a);
    float f = -acc - b[p] * sa_lap(g, work, p);
    grad[p] += f * pa[p];
    prec[p] += f * f;
}

// exact adjoint accumulation for equation t: lambda = mu_{t+1} / G
__global__ void k_exact_acc_m(Geo g, const float *__restrict__ Kmu, const float *__restrict__ Ku, float c4,
                              float *gm) {
    DOMAIN_INDEX
    gm[p] += c4 * Kmu[p] * Ku[p];
}

__global__ void k_axpy_domain(Geo g, float alpha, const float *__restrict__ x, float *y) {
    DOMAIN_INDEX
    y[p] += alpha * x[p];
}

// The exact accumulations of one adjoint step in one pass: gA += lam u, gB += lam um,
// gG += lam Ku (lam = mu / G) and, with gb, d/db through the half-point buoyancy,
// gb -= 1/2 sum_a (D+ mu D+ u)(p) + (D+ mu D+ u)(p - s_a), the staggered first derivatives formed
// here from u and mu (sa_dplus, the staggered derivative of k_stagger).
__global__ void k_exact_acc_fused(Geo g, const float *__restrict__ mu, const float *__restrict__ G,
                                  const float *__restrict__ u, const float *__restrict__ um,
                                  const float *__restrict__ Ku, float *gA, float *gB, float *gG, float *gb) {
    DOMAIN_INDEX
    float lam = mu[p] / G[p];
    gA[p] += lam * u[p];
    gB[p] += lam * um[p];
    gG[p] += lam * Ku[p];
    if (gb) {
        float acc = 0.f;
        for (int a = 0; a < 3; ++a) {
            if (!g.act[a]) continue;
            const long long s = sa_stride(g, a);
            const float up = sa_dplus(g, u, p, a), ul = sa_dplus(g, u, p - s, a);
            const float mp = sa_dplus(g, mu, p, a), ml = sa_dplus(g, mu, p - s, a);
            acc += mp * up + ml * ul;
        }
        gb[p] += -0.5f * acc;
    }
}

// ------------------------------------------------------------------------------------------
// host side

struct DevBuf {
    float *p = nullptr;
    size_t n = 0;
};

struct Sparse {
    int npoint = 0, ntap = 0;
    std::vector<long long> idx;  // host copy (extended-grid linear indices, -1: none)
    std::vector<float> w;
    long long *d_idx = nullptr;  // padded indices per tap (-1: none)
    float *d_w = nullptr;
    // deterministic scatter: destinations sorted, CSR of contributions
    int ndest = 0, nnz = 0;
    long long *d_dest = nullptr, *d_lin = nullptr;  // padded index per destination / contribution
    int *d_ijk = nullptr;                           // domain coordinates per destination
    int *d_off = nullptr, *d_point = nullptr;
    float *d_cw = nullptr;    // raw weights per contribution
    float *d_coef = nullptr;  // scaled coefficients per contribution
    int *d_map = nullptr;     // per padded cell: its destination index, -1 for none (fused 2D update)
    int *d_toff = nullptr, *d_tpt = nullptr;  // points per 2D update tile (CSR by first tap)
};

// gradient accumulation the fused 2D update performs on a guarded adjoint step
struct GradSpec {
    const float *ps, *psm, *psp;  // snapshot of this step; raw mode: its neighbours (null: zero)
    float fdt2;
    int raw;
    float *gv, *pv, *gr, *pr;
};

// an interpolation the fused 2D update performs from its u (k_gather's sum, same order)
struct GatherSpec {
    Sparse *sp;
    float sign;
    float *out;
};

// an injection the fused 2D update applies to the cells it writes (k_inject's sum, same order)
struct InjectSpec {
    Sparse *sp;
    const float *values;
    float sign;
    bool into_save;  // also add (scaled by idt2) into the p.dt2 snapshot
};

struct sa_solver {
    sa_grid grid;
    sa_scheme sch;
    bool have_scheme = false;
    int device = 0;
    hipStream_t stream = nullptr;
    Geo geo;
    int P[3];
    long long total = 0, npts = 0;
    float dt = 0.f;
    double hmax = 0;
    // medium (padded)
    float *vp = nullptr, *rho = nullptr, *b = nullptr, *alpha = nullptr;
    bool medium_set = false;
    // coefficients
    float *A = nullptr, *B = nullptr, *G = nullptr, *M = nullptr, *Cbeta = nullptr, *Cpml = nullptr;
    bool coef_dirty = true;
    // boundary
    float *tables = nullptr;
    int tlen = 0;
    float *pml[3][4] = {};
    float *u1[3] = {}, *u2[3] = {}, *u3[3] = {};
    // wavefields and work
    float *buf[2] = {};
    float *Ku = nullptr, *L2 = nullptr, *KmKu = nullptr, *lapo = nullptr;
    float *lapother = nullptr;  // Laplacian of the other level (power-2 attenuation)
    // fused 2D update: it also writes the Laplacian of u into lapnext, which is the Laplacian of
    // the next step's other level; lap_of is the level lapother holds the Laplacian of (null:
    // none, e.g. at the start of a run)
    float *lapnext = nullptr;
    const float *lap_of = nullptr;
    float *st[3] = {};
    float *work0 = nullptr, *work1 = nullptr;
    int cur = 0;  // index of the last computed level in buf
    // sparse
    Sparse src, rec;
    // saved wavefield
    float *saved = nullptr;
    size_t saved_bytes = 0;
    int nsave = 0;
    Box box;
    int saved_mode = 0;
    int saved_f = 0;
    // checkpointed exact adjoint (save_mode 4): pairs (p[jK+1], p[jK]) in `saved`, a K+1 level
    // segment buffer, forward work buffers and the source samples kept for the recomputation
    float *segment = nullptr;
    size_t segment_bytes = 0;
    float *fb[2] = {};
    float *src_keep = nullptr;
    size_t src_keep_n = 0;
    float *medium_stage = nullptr;  // unpadded medium array on its way to the padded layout
    float *kseg = nullptr;  // save mode 4 with the fused density update: K u of the segment's levels
    size_t kseg_bytes = 0;
    bool kseg_valid = false;  // the current segment's K u slots were written by its recomputation
    size_t medium_stage_n = 0;
    // checkpointed raw saves (save_mode 5): `saved` holds the states (p[T_j], p[T_j - 1]),
    // T_j = (jK - 1) f, of segments j >= 1; `segment` is a ring of K + 2 snapshots (snapshot i in
    // slot i % (K + 2)) holding snapshots jK - 1 .. (j + 1)K of segment `ck_seg`; snapshots outside
    // [ck_first, ck_last] read as zero, as the slots mode 2 never writes
    int ck_K = 0, ck_J = 0, ck_first = 0, ck_last = -1, ck_seg = -1;
    // snapshots moved to host memory (cache_forward): the layout above is kept, `saved` is null
    bool saved_on_host = false;
    float *xfer = nullptr;  // one region of one snapshot, staged for host copies
    size_t xfer_n = 0;
    // wavefield dumps (dump_forward_wavefield / dump_adjoint_wavefield): raw level every
    // dump_f[w] steps over the whole domain, w = 0 forward, 1 adjoint
    int dump_f[2] = {};
    int dump_n[2] = {};
    float *dump[2] = {};
    size_t dump_bytes[2] = {};
    // host staging
    float *d_series = nullptr;
    size_t series_n = 0;
    float *d_out = nullptr;
    size_t out_n = 0;
    // timing
    bool timing = false;
    int timing_stride = 1;     // every timing_stride-th update launch is timed
    long long timing_seq = 0;  // update launches since timing was enabled
    bool timing_this = false;  // the launch between timed_begin and timed_end is timed
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    double kernel_ms = 0;
    long long launches = 0;
    // gradient accumulators (padded): grad_vp, prec_vp, grad_rho, prec_rho
    float *gacc[4] = {};
    // fast 3D path
    Fast3D fast;
    Rho3DTune rho3d;  // fused div(b grad) update (3D, OT2, no PML, no power-2 attenuation)
    Rho3DTune rho3d_b2 = {1, 60, 4, true};  // its two-shot form (sa_batch_*): rows 1 only
    // two shots per pass (sa_batch_*): interleaved levels, per-shot taps, interleaved snapshots
    int nshot = 0;
    float *bbuf[2] = {};
    Sparse bsrc[2], brec[2];
    float *bsaved = nullptr;
    size_t bsaved_bytes = 0;
    // lockstep 2D shots (sa_shots2d_*, as the first solver of a group): the per-step, per-shot
    // argument table (pinned host copy and device copy) and the event after its upload
    Flat2DStep *tab_h = nullptr, *tab_d = nullptr;
    size_t tab_cap = 0;
    hipEvent_t tab_ev = nullptr;
    int bnsave = 0, bf = 0;
    Box bbox;
    // many shots of uneven duration (sa_multi_*): shots paired longest first, pair j holding
    // sorted shots 2j and 2j + 1 in two level buffers; device tables of the pairs' levels and
    // of the shots' injections and interpolations; a stream the schedule is captured on
    struct MShot {
        int nt = 0;
        bool set = false;
        Sparse src, rec;
    };
    std::vector<MShot> mshot;
    std::vector<int> morder;  // sorted position -> shot
    std::vector<float *> mbuf;
    PairLevels *d_pairs = nullptr;
    MInject *d_minj = nullptr;
    MGather *d_mgat = nullptr;
    hipStream_t mstream = nullptr;
    hipGraph_t mgraph = nullptr;      // the last captured schedule and its executable form, kept
    hipGraphExec_t mexec = nullptr;   // until the next schedule or sa_destroy (after a wait)
    float graph_ms = 0.f;  // capture + instantiation time of the last graph
};

static void release_graph(sa_solver *s) {
    if (s->mexec) (void)hipGraphExecDestroy(s->mexec);
    if (s->mgraph) (void)hipGraphDestroy(s->mgraph);
    s->mexec = nullptr;
    s->mgraph = nullptr;
}

static int dalloc(float **p, size_t n) {
    if (*p) return SA_OK;
    if (hipMalloc((void **)p, n * sizeof(float)) != hipSuccess) {
        *p = nullptr;
        return fail(SA_E_NOMEM, "hipMalloc of %zu floats failed", n);
    }
    // hipMemset on device memory may return before the fill has run, and the solver's kernels run
    // on non-blocking streams that do not wait for the null stream: wait for it here
    if (hipMemset(*p, 0, n * sizeof(float)) != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess)
        return fail(SA_E_HIP, "hipMemset failed");
    return SA_OK;
}

static void dfree(void *p) {
    if (p) (void)hipFree(p);
}

// Host-to-device upload on the null stream, waited for: the solver's kernels run on streams that
// may be non-blocking (sa_stream_create), which do not wait for null-stream work.
static hipError_t h2d(void *dst, const void *src, size_t bytes) {
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    return e != hipSuccess ? e : hipStreamSynchronize(nullptr);
}

#define TRY(x)                  \
    do {                        \
        int rc_ = (x);          \
        if (rc_ != SA_OK) return rc_; \
    } while (0)

// Setters copy into (or free and reallocate) buffers that kernels already queued on the
// solver's stream may still read: the "_dev" calls return at once, and the solver's stream
// may be non-blocking (sa_stream_create), so the synchronous copies on the null stream would
// not wait for it. Each setter first waits for the solver's stream.
static int quiesce(sa_solver *s) {
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SA_OK;
}

static dim3 domain_grid(const Geo &g, dim3 blk) {
    return dim3((g.n[2] + blk.x - 1) / blk.x, (g.n[1] + blk.y - 1) / blk.y, g.n[0]);
}

static const dim3 kBlk(64, 4, 1);

static int zero_field(sa_solver *s, float *p) {
    HIP_TRY(hipMemsetAsync(p, 0, s->total * sizeof(float), s->stream));
    return SA_OK;
}

static int zero_field2(sa_solver *s, float *p) {  // two interleaved shots
    HIP_TRY(hipMemsetAsync(p, 0, 2 * s->total * sizeof(float), s->stream));
    return SA_OK;
}

// unpadded extended-grid host copy of a padded device array
static int download_unpadded(sa_solver *s, float *host, const float *dev) {
    const Geo &g = s->geo;
    std::vector<float> tmp(s->total);
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(tmp.data(), dev, s->total * sizeof(float), hipMemcpyDeviceToHost));
    for (int i = 0; i < g.n[0]; ++i)
        for (int j = 0; j < g.n[1]; ++j)
            for (int k = 0; k < g.n[2]; ++k)
                host[((long long)i * g.n[1] + j) * g.n[2] + k] = tmp[g.base + (long long)i * g.s0 + (long long)j * g.s1 + k];
    return SA_OK;
}

extern "C" const char *sa_last_error(void) { return g_err.c_str(); }
extern "C" const char *sa_version(void) { return "stride_amd 0.1 gfx950"; }

extern "C" int sa_create(const sa_grid *grid, sa_solver **out) {
    if (!grid || !out) return fail(SA_E_ARG, "null argument");
    if (grid->dim != 2 && grid->dim != 3) return fail(SA_E_ARG, "dim must be 2 or 3, got %d", grid->dim);
    if (grid->nt < 3) return fail(SA_E_ARG, "nt must be >= 3");
    for (int d = 0; d < grid->dim; ++d) {
        if (grid->shape[d] < 1) return fail(SA_E_ARG, "shape[%d] = %d", d, grid->shape[d]);
        if (grid->absorbing[d] < 0 || 2 * grid->absorbing[d] > grid->shape[d])
            return fail(SA_E_ARG, "absorbing[%d] = %d does not fit shape %d", d, grid->absorbing[d], grid->shape[d]);
        if (!(grid->spacing[d] > 0)) return fail(SA_E_ARG, "spacing[%d] must be positive", d);
    }
    if (!(grid->dt > 0)) return fail(SA_E_ARG, "dt must be positive");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (grid->device < 0 || grid->device >= ndev) return fail(SA_E_ARG, "device %d not present (%d devices)", grid->device, ndev);
    HIP_TRY(hipSetDevice(grid->device));
    sa_solver *s = new sa_solver();
    s->grid = *grid;
    s->device = grid->device;
    Geo &g = s->geo;
    memset(&g, 0, sizeof(g));
    int off = 3 - grid->dim;
    int lead[3];
    for (int a = 0; a < 3; ++a) {
        int d = a - off;
        g.act[a] = d >= 0;
        g.n[a] = d >= 0 ? grid->shape[d] : 1;
        g.nabs[a] = d >= 0 ? grid->absorbing[d] : 0;
        lead[a] = g.act[a] ? (a == 2 ? SA_LEAD : SA_HALO) : 0;
        s->P[a] = lead[a] + g.n[a] + (g.act[a] ? SA_HALO : 0);
        double h = d >= 0 ? grid->spacing[d] : 1.0;
        for (int r = 0; r < 6; ++r) {
            g.w2[a][r] = g.act[a] ? (float)(kW2[r] / (h * h)) : 0.f;
            g.w1[a][r] = g.act[a] ? (float)(kW1[r] / h) : 0.f;
        }
        for (int r = 0; r < 5; ++r) g.ws[a][r] = g.act[a] ? (float)(kWS[r] / h) : 0.f;
        if (d >= 0) s->hmax = std::max(s->hmax, h);
    }
    s->P[2] = (s->P[2] + 31) / 32 * 32;
    g.s1 = s->P[2];
    g.s0 = (long long)s->P[1] * s->P[2];
    g.base = (long long)lead[0] * g.s0 + (long long)lead[1] * g.s1 + lead[2];
    g.total = (long long)s->P[0] * g.s0;
    s->total = g.total;
    s->npts = (long long)g.n[0] * g.n[1] * g.n[2];
    s->dt = (float)grid->dt;
    int rc = dalloc(&s->buf[0], s->total);
    if (rc == SA_OK) rc = dalloc(&s->buf[1], s->total);
    if (rc == SA_OK) rc = dalloc(&s->Ku, s->total);
    if (rc != SA_OK) {
        sa_destroy(s);
        return rc;
    }
    *out = s;
    return SA_OK;
}

static void free_sparse(Sparse &sp) {
    dfree(sp.d_map);
    dfree(sp.d_toff);
    dfree(sp.d_tpt);
    dfree(sp.d_idx); dfree(sp.d_w); dfree(sp.d_dest); dfree(sp.d_lin); dfree(sp.d_off); dfree(sp.d_point);
    dfree(sp.d_ijk);
    dfree(sp.d_cw); dfree(sp.d_coef);
    sp = Sparse();
}

extern "C" int sa_destroy(sa_solver *s) {
    if (!s) return SA_OK;
    (void)hipSetDevice(s->device);
    (void)hipDeviceSynchronize();
    float *fs[] = {s->vp, s->rho, s->b, s->alpha, s->A, s->B, s->G, s->M, s->Cbeta, s->Cpml, s->tables,
                   s->buf[0], s->buf[1], s->Ku, s->L2, s->KmKu, s->lapo, s->lapother, s->lapnext, s->work0, s->work1, s->saved,
                   s->d_series, s->d_out};
    for (float *p : fs) dfree(p);
    for (float *p : s->gacc) dfree(p);
    dfree(s->segment);
    dfree(s->fb[0]);
    dfree(s->fb[1]);
    dfree(s->src_keep);
    dfree(s->medium_stage);
    dfree(s->kseg);
    dfree(s->xfer);
    dfree(s->dump[0]);
    dfree(s->dump[1]);
    for (int a = 0; a < 3; ++a) {
        dfree(s->u1[a]); dfree(s->u2[a]); dfree(s->u3[a]); dfree(s->st[a]);
        for (int r = 0; r < 4; ++r) dfree(s->pml[a][r]);
    }
    free_sparse(s->src);
    free_sparse(s->rec);
    for (int k = 0; k < 2; ++k) {
        free_sparse(s->bsrc[k]);
        free_sparse(s->brec[k]);
        dfree(s->bbuf[k]);
    }
    dfree(s->bsaved);
    if (s->tab_h) (void)hipHostFree(s->tab_h);
    if (s->tab_d) (void)hipFree(s->tab_d);
    if (s->tab_ev) (void)hipEventDestroy(s->tab_ev);
    for (auto &m : s->mshot) {
        free_sparse(m.src);
        free_sparse(m.rec);
    }
    for (float *p : s->mbuf) dfree(p);
    dfree(s->d_pairs);
    dfree(s->d_minj);
    dfree(s->d_mgat);
    release_graph(s);
    if (s->mstream) (void)hipStreamDestroy(s->mstream);
    fast3d_free(s->fast);
    for (auto e : s->ev) (void)hipEventDestroy(e);
    delete s;
    return SA_OK;
}

extern "C" int sa_set_stream(sa_solver *s, void *stream) {
    if (!s) return fail(SA_E_ARG, "null solver");
    s->stream = (hipStream_t)stream;
    return SA_OK;
}

extern "C" int sa_set_scheme(sa_solver *s, const sa_scheme *sch) {
    if (!s || !sch) return fail(SA_E_ARG, "null argument");
    if (sch->kernel != 2 && sch->kernel != 4) return fail(SA_E_ARG, "Unrecognized kernel %d", sch->kernel);
    if (sch->rho_mode < 0 || sch->rho_mode > 2) return fail(SA_E_ARG, "rho_mode %d", sch->rho_mode);
    if (!(sch->att_power == -2 || sch->att_power == -1 || sch->att_power == 0 || sch->att_power == 2))
        return fail(SA_E_ARG, "The \"attenuation_exponent\" can only take values (0, 2).");
    if (sch->boundary < -1 || sch->boundary > 1) return fail(SA_E_ARG, "boundary %d", sch->boundary);
    if (sch->undersampling < 1) return fail(SA_E_ARG, "undersampling must be >= 1");
    if (sch->tb0 < 0 || sch->tb1 > s->grid.nt || sch->tb1 <= sch->tb0 + 2)
        return fail(SA_E_ARG, "time bounds (%d, %d) invalid for nt %d", sch->tb0, sch->tb1, s->grid.nt);
    if (sch->save_mode < 0 || sch->save_mode > 5) return fail(SA_E_ARG, "save_mode %d", sch->save_mode);
    if (sch->save_mode == 5) {
        if (sch->checkpoint < 1) return fail(SA_E_ARG, "save_mode 5 needs checkpoint >= 1 snapshots per segment");
        // the recomputation restarts from the two time levels alone
        if (sch->boundary == 1 || sch->att_power == 2)
            return fail(SA_E_UNSUPPORTED, "checkpointed raw saves (save_mode 5) support the sponge or no boundary "
                                          "and attenuation none/0/None");
    }
    if (!s->have_scheme || memcmp(&s->sch, sch, sizeof(*sch)) != 0) s->coef_dirty = true;
    s->sch = *sch;
    s->have_scheme = true;
    return SA_OK;
}

static int upload_padded_dev(sa_solver *s, float **dev, const float *src);
static int ensure_staging(sa_solver *s, float **p, size_t *have, size_t n);

// A host medium array goes to the device unpadded (one copy of the extended grid) and is padded
// there (edge halo, k_pad_edge); the staging buffer is reused by the next array once the padding
// launch has run.
static int upload_medium(sa_solver *s, float **dev, const float *host) {
    TRY(ensure_staging(s, &s->medium_stage, &s->medium_stage_n, (size_t)s->npts));
    HIP_TRY(h2d(s->medium_stage, host, (size_t)s->npts * sizeof(float)));
    TRY(upload_padded_dev(s, dev, s->medium_stage));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SA_OK;
}

extern "C" int sa_set_medium(sa_solver *s, const float *vp, const float *rho, const float *alpha) {
    if (!s || !vp) return fail(SA_E_ARG, "vp is required");
    TRY(quiesce(s));
    if (rho)
        for (long long q = 0; q < s->npts; ++q)
            if (!(rho[q] > 0)) return fail(SA_E_ARG, "rho must be positive");
    TRY(upload_medium(s, &s->vp, vp));
    if (rho) {
        TRY(upload_medium(s, &s->rho, rho));
        // buoyancy 1 / rho over the whole padded array (the edge halo of rho is positive)
        TRY(dalloc(&s->b, s->total));
        k_reciprocal<<<(unsigned)((s->total + 255) / 256), 256, 0, s->stream>>>(s->rho, s->b, s->total);
        CHECK_LAUNCH();
    }
    if (alpha) TRY(upload_medium(s, &s->alpha, alpha));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->medium_set = true;
    s->coef_dirty = true;
    return SA_OK;
}

static int upload_padded_dev(sa_solver *s, float **dev, const float *src) {
    TRY(dalloc(dev, s->total));
    const Geo &g = s->geo;
    int lo[3];
    for (int a = 0; a < 3; ++a) lo[a] = g.act[a] ? (a == 2 ? SA_LEAD : SA_HALO) : 0;
    const dim3 grid((s->P[2] + kBlk.x - 1) / kBlk.x, (s->P[1] + kBlk.y - 1) / kBlk.y, s->P[0]);
    k_pad_edge<<<grid, kBlk, 0, s->stream>>>(g, src, *dev, lo[0], lo[1], lo[2], s->P[1], s->P[2]);
    CHECK_LAUNCH();
    return SA_OK;
}

extern "C" int sa_set_medium_dev(sa_solver *s, const float *vp, const float *rho, const float *alpha) {
    if (!s || !vp) return fail(SA_E_ARG, "vp is required");
    TRY(quiesce(s));
    TRY(upload_padded_dev(s, &s->vp, vp));
    if (rho) {
        TRY(upload_padded_dev(s, &s->rho, rho));
        TRY(dalloc(&s->b, s->total));
        k_reciprocal<<<(unsigned)((s->total + 255) / 256), 256, 0, s->stream>>>(s->rho, s->b, s->total);
        CHECK_LAUNCH();
    }
    if (alpha) TRY(upload_padded_dev(s, &s->alpha, alpha));
    s->medium_set = true;
    s->coef_dirty = true;
    return SA_OK;
}

extern "C" int sa_set_sponge(sa_solver *s, const float *tables, int len) {
    if (!s || !tables || len < 1) return fail(SA_E_ARG, "sponge tables required");
    for (int d = 0; d < s->grid.dim; ++d)
        if (s->grid.absorbing[d] > len) return fail(SA_E_ARG, "sponge table shorter than the layer");
    TRY(quiesce(s));
    dfree(s->tables);
    s->tables = nullptr;
    // re-index tables to internal axes (2D: axis 0 unused)
    int off = 3 - s->grid.dim;
    std::vector<float> t(3 * (size_t)len, 0.f);
    for (int d = 0; d < s->grid.dim; ++d)
        for (int l = 0; l < len; ++l) t[(size_t)(d + off) * len + l] = tables[(size_t)d * len + l];
    TRY(dalloc(&s->tables, t.size()));
    HIP_TRY(h2d(s->tables, t.data(), t.size() * sizeof(float)));
    s->tlen = len;
    s->coef_dirty = true;
    return SA_OK;
}

extern "C" int sa_set_pml(sa_solver *s, const float *profiles, int stride) {
    if (!s || !profiles) return fail(SA_E_ARG, "pml profiles required");
    TRY(quiesce(s));
    int off = 3 - s->grid.dim;
    for (int d = 0; d < s->grid.dim; ++d) {
        int a = d + off;
        if (stride < s->geo.n[a]) return fail(SA_E_ARG, "pml stride too short");
        for (int r = 0; r < 4; ++r) {
            TRY(dalloc(&s->pml[a][r], s->geo.n[a]));
            HIP_TRY(h2d(s->pml[a][r], profiles + ((size_t)d * 4 + r) * stride, s->geo.n[a] * sizeof(float)));
        }
    }
    return SA_OK;
}

static int set_sparse(sa_solver *s, Sparse &sp, int npoint, int ntap, const int64_t *index, const float *weight) {
    if (npoint < 0 || ntap < 1 || (npoint > 0 && (!index || !weight)))
        return fail(SA_E_ARG, "sparse points need index and weight arrays");
    TRY(quiesce(s));
    free_sparse(sp);
    sp.npoint = npoint;
    sp.ntap = ntap;
    size_t nq = (size_t)npoint * ntap;
    sp.idx.assign(index, index + nq);
    sp.w.assign(weight, weight + nq);
    for (size_t q = 0; q < nq; ++q)
        if (sp.idx[q] >= s->npts) return fail(SA_E_ARG, "tap index %lld outside the grid", (long long)sp.idx[q]);
    if (nq == 0) return SA_OK;
    // padded indices, computed once here so that the per-step kernels do no 64-bit division
    const Geo &g = s->geo;
    auto padded = [&](long long lin) -> long long {
        if (lin < 0) return -1;
        const long long k = lin % g.n[2], r = lin / g.n[2];
        return g.base + (r / g.n[1]) * g.s0 + (r % g.n[1]) * g.s1 + k;
    };
    std::vector<long long> pidx(nq);
    for (size_t q = 0; q < nq; ++q) pidx[q] = padded(sp.idx[q]);
    HIP_TRY(hipMalloc((void **)&sp.d_idx, nq * sizeof(long long)));
    HIP_TRY(hipMalloc((void **)&sp.d_w, nq * sizeof(float)));
    HIP_TRY(h2d(sp.d_idx, pidx.data(), nq * sizeof(long long)));
    HIP_TRY(h2d(sp.d_w, sp.w.data(), nq * sizeof(float)));
    // destination-sorted CSR for the deterministic scatter (stable order of contributions)
    std::vector<size_t> ord;
    for (size_t q = 0; q < nq; ++q)
        if (sp.idx[q] >= 0) ord.push_back(q);
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return sp.idx[a] < sp.idx[b]; });
    std::vector<long long> dest, lin;
    std::vector<int> offv, point, ijk;
    std::vector<float> cw;
    for (size_t c = 0; c < ord.size(); ++c) {
        size_t q = ord[c];
        if (c == 0 || sp.idx[q] != sp.idx[ord[c - 1]]) {
            const long long l = sp.idx[q];
            dest.push_back(pidx[q]);
            ijk.push_back((int)(l / ((long long)g.n[1] * g.n[2])));
            ijk.push_back((int)((l / g.n[2]) % g.n[1]));
            ijk.push_back((int)(l % g.n[2]));
            offv.push_back((int)c);
        }
        point.push_back((int)(q / ntap));
        lin.push_back(pidx[q]);
        cw.push_back(sp.w[q]);
    }
    offv.push_back((int)ord.size());
    sp.ndest = (int)dest.size();
    sp.nnz = (int)ord.size();
    if (sp.nnz == 0) return SA_OK;
    HIP_TRY(hipMalloc((void **)&sp.d_dest, dest.size() * sizeof(long long)));
    HIP_TRY(hipMalloc((void **)&sp.d_lin, lin.size() * sizeof(long long)));
    HIP_TRY(hipMalloc((void **)&sp.d_off, offv.size() * sizeof(int)));
    HIP_TRY(hipMalloc((void **)&sp.d_point, point.size() * sizeof(int)));
    HIP_TRY(hipMalloc((void **)&sp.d_cw, cw.size() * sizeof(float)));
    HIP_TRY(hipMalloc((void **)&sp.d_coef, cw.size() * sizeof(float)));
    HIP_TRY(hipMalloc((void **)&sp.d_ijk, ijk.size() * sizeof(int)));
    HIP_TRY(h2d(sp.d_ijk, ijk.data(), ijk.size() * sizeof(int)));
    HIP_TRY(h2d(sp.d_dest, dest.data(), dest.size() * sizeof(long long)));
    HIP_TRY(h2d(sp.d_lin, lin.data(), lin.size() * sizeof(long long)));
    HIP_TRY(h2d(sp.d_off, offv.data(), offv.size() * sizeof(int)));
    HIP_TRY(h2d(sp.d_point, point.data(), point.size() * sizeof(int)));
    HIP_TRY(h2d(sp.d_cw, cw.data(), cw.size() * sizeof(float)));
    return SA_OK;
}

extern "C" int sa_set_sources(sa_solver *s, int npoint, int ntap, const int64_t *index, const float *weight) {
    if (!s) return fail(SA_E_ARG, "null solver");
    return set_sparse(s, s->src, npoint, ntap, index, weight);
}

extern "C" int sa_set_receivers(sa_solver *s, int npoint, int ntap, const int64_t *index, const float *weight) {
    if (!s) return fail(SA_E_ARG, "null solver");
    return set_sparse(s, s->rec, npoint, ntap, index, weight);
}

static int lap_other(sa_solver *s, const float *other);
static void lap_advance(sa_solver *s, const float *u);

// out = the centred Laplacian of u over the domain: in 3D the specialised update's Laplacian-only
// pass, else the general kernel
static int laplacian(sa_solver *s, const float *u, float *out) {
    const Geo &g = s->geo;
    if (s->grid.dim == 3 && fast3d_enabled() && fast3d_supported(g)) {
        if (fast3d_laplacian(s->fast, g, u, out, s->stream) != 0)
            return fail(SA_E_HIP, "Laplacian launch: %s", hipGetErrorString(hipGetLastError()));
        return SA_OK;
    }
    k_laplace<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, u, nullptr, out);
    CHECK_LAUNCH();
    return SA_OK;
}

// the fused 3D density update (with CFS-PML: the auxiliary fields by k_pml_slabs before it and
// their term by k_pml_correct after it)
// (OT4: two passes, no PML, no power-2 attenuation)
static bool use_rho3d(const sa_solver *s) {
    const sa_scheme &c = s->sch;
    const bool order = c.kernel == 2 || (c.kernel == 4 && c.boundary != 1 && c.att_power != 2);
    return s->grid.dim == 3 && order && c.rho_mode != 0 && s->rho3d.enabled && rho3d_supported(s->geo);
}

static bool use_flat2d(const sa_solver *s) {
    return s->grid.dim == 2 && s->sch.kernel == 2 && s->rho3d.enabled && flat2d_supported(s->geo);
}

// the specialised 3D update: no density term; attenuation of any form with one shot per pass;
// sponge, no boundary or CFS-PML (as for the density update)
// (OT4: two passes, attenuation off / power 0 / None, no PML); every save mode: p.dt2 or raw
// snapshots fused, and the checkpoint recomputations (modes 4, 5) and the exact adjoint's steps run
// the same update
static bool use_fast(const sa_solver *s) {
    const sa_scheme &c = s->sch;
    const bool order = c.kernel == 2 || (c.kernel == 4 && c.boundary != 1 && c.att_power != 2);
    return s->grid.dim == 3 && order && c.rho_mode == 0 && fast3d_enabled() && fast3d_supported(s->geo);
}

// the two-shot and multi-shot forms of the specialised update: no attenuation, no PML
static bool use_fast_pairs(const sa_solver *s) {
    return use_fast(s) && s->sch.kernel == 2 && s->sch.att_power == -2 && s->sch.boundary != 1;
}

// the two-shot form of the fused density update: no power-2 attenuation, no PML
static bool use_rho3d_pairs(const sa_solver *s) {
    return use_rho3d(s) && s->sch.kernel == 2 && s->sch.att_power != 2 && s->sch.boundary != 1 &&
           rho3d_supported(s->geo, 2);
}

static PmlSet pml_set(const sa_solver *s) {
    PmlSet ps;
    memset(&ps, 0, sizeof(ps));
    for (int a = 0; a < 3; ++a) {
        if (!s->geo.act[a]) continue;
        ps.u1[a] = s->u1[a];
        ps.u2[a] = s->u2[a];
        ps.u3[a] = s->u3[a];
        for (int r = 0; r < 4; ++r) ps.prof[a][r] = s->pml[a][r];
    }
    ps.Cpml = s->Cpml;
    return ps;
}

// CFS-PML around a fused 3D update (OT2, F = u): the auxiliary fields of every axis before it ...
static int pml_before(sa_solver *s, const float *u, int backward) {
    const Geo &g = s->geo;
    const unsigned rows = (unsigned)((g.n[0] * g.n[1] + kSlabRows - 1) / kSlabRows);
    k_pml_slabs<<<rows, dim3(64, kSlabRows), 0, s->stream>>>(g, u, pml_set(s), s->dt, backward);
    CHECK_LAUNCH();
    return SA_OK;
}

// ... and their term of (1): applied by the specialised update itself (its operands added to
// `at`, which becomes the attenuation argument when there was none) ...
static void pml_term(const sa_solver *s, FastAtt &at, FastAtt *&pat) {
    if (!pat) {
        at.power = -2;
        pat = &at;
    }
    at.cpml = s->Cpml;
    for (int a = 0; a < 3; ++a) at.u1[a] = s->u1[a];
}

// OT4 on the specialised update: a first pass forms K u and w = m K u (m = vp^2, the operand of
// the general path's K(m K u)); the update then stages w, so its Laplacian is K(m K u), and reads
// u and K u as own-cell operands. Returns the field the update stages.
static int ot4_first(sa_solver *s, const float *u, FastAtt &at, FastAtt *&pat, const float *&staged) {
    if (fast3d_ot4_first(s->fast, s->geo, u, s->Ku, s->L2, s->stream) != 0)
        return fail(SA_E_HIP, "OT4 first pass launch: %s", hipGetErrorString(hipGetLastError()));
    if (!pat) {
        at.power = -2;
        pat = &at;
    }
    at.ot4_u = u;
    at.ot4_ku = s->Ku;
    at.c4 = s->dt * s->dt / 12.f;
    staged = s->L2;
    return SA_OK;
}

// ... or, after the fused density update, by a launch of its own
static int pml_after(sa_solver *s, float *out) {
    const Geo &g = s->geo;
    const unsigned rows = (unsigned)((g.n[0] * g.n[1] + kSlabRows - 1) / kSlabRows);
    k_pml_correct<<<rows, dim3(64, kSlabRows), 0, s->stream>>>(g, pml_set(s), out);
    CHECK_LAUNCH();
    return SA_OK;
}

// build per-point coefficients and boundary state for the current scheme + medium
static int prepare(sa_solver *s) {
    if (!s->have_scheme) return fail(SA_E_STATE, "sa_set_scheme must be called first");
    if (!s->medium_set) return fail(SA_E_STATE, "sa_set_medium must be called first");
    const sa_scheme &c = s->sch;
    if (c.rho_mode == 2 && !s->rho) return fail(SA_E_ARG, "variable density needs rho");
    if (c.att_power != -2 && !s->alpha) return fail(SA_E_ARG, "attenuation needs alpha");
    if (c.boundary == 0 && !s->tables) return fail(SA_E_ARG, "sponge boundary needs sa_set_sponge");
    if (c.boundary == 1)
        for (int a = 0; a < 3; ++a)
            if (s->geo.act[a] && !s->pml[a][0]) return fail(SA_E_ARG, "PML boundary needs sa_set_pml");
    HIP_TRY(hipSetDevice(s->device));
    if (!s->coef_dirty) return SA_OK;
    TRY(dalloc(&s->A, s->total));
    TRY(dalloc(&s->B, s->total));
    TRY(dalloc(&s->G, s->total));
    TRY(dalloc(&s->M, s->total));
    if (c.att_power == 2) TRY(dalloc(&s->Cbeta, s->total));
    if (c.att_power == 2) TRY(dalloc(&s->lapother, s->total));
    // the Laplacian of u for the next step (the fused updates); allocated for every power-2 scheme,
    // since which update runs can change after this (sa_set_tile_density) without a new prepare
    if (c.att_power == 2) TRY(dalloc(&s->lapnext, s->total));
    if (c.boundary == 1) TRY(dalloc(&s->Cpml, s->total));
    const Geo &g = s->geo;
    k_coefficients<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(
        g, s->vp, s->rho, s->alpha, s->tables, s->tlen, c.boundary, c.att_power, c.rho_mode, s->dt, s->A, s->B,
        s->G, s->M, c.att_power == 2 ? s->Cbeta : nullptr, c.boundary == 1 ? s->Cpml : nullptr);
    CHECK_LAUNCH();
    if (c.rho_mode != 0)
        for (int a = 0; a < 3; ++a)
            if (g.act[a]) TRY(dalloc(&s->st[a], s->total));
    if (c.kernel == 4) {
        TRY(dalloc(&s->L2, s->total));
        TRY(dalloc(&s->KmKu, s->total));
    }
    if (c.boundary == 1) {
        for (int a = 0; a < 3; ++a)
            if (g.act[a]) {
                TRY(dalloc(&s->u1[a], s->total));
                TRY(dalloc(&s->u2[a], s->total));
                TRY(dalloc(&s->u3[a], s->total));
            }
        if (c.kernel == 4) TRY(dalloc(&s->lapo, s->total));
    }
    // with CFS-PML the specialised update runs without a boundary term (pml_before / pml_after)
    if (use_fast(s))
        TRY(fast3d_prepare(s->fast, g, s->vp, s->tables, s->tlen, c.boundary == 1 ? -1 : c.boundary, s->dt,
                           s->stream));
    s->coef_dirty = false;
    return SA_OK;
}

static int timed_begin(sa_solver *s, hipEvent_t *e0) {
    if (!s->timing) return SA_OK;
    s->timing_this = s->timing_seq++ % s->timing_stride == 0;
    if (!s->timing_this) return SA_OK;
    if (s->ev_used + 2 > s->ev.size()) {
        for (int q = 0; q < 256; ++q) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            s->ev.push_back(e);
        }
    }
    *e0 = s->ev[s->ev_used];
    HIP_TRY(hipEventRecord(*e0, s->stream));
    return SA_OK;
}

static int timed_end(sa_solver *s) {
    if (!s->timing || !s->timing_this) return SA_OK;
    hipEvent_t e1 = s->ev[s->ev_used + 1];
    HIP_TRY(hipEventRecord(e1, s->stream));
    s->ev_used += 2;
    s->launches += 1;
    if (s->ev_used + 2 > 4096) {  // drain the pool
        HIP_TRY(hipEventSynchronize(s->ev[s->ev_used - 1]));
        for (size_t q = 0; q < s->ev_used; q += 2) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, s->ev[q], s->ev[q + 1]));
            s->kernel_ms += ms;
        }
        s->ev_used = 0;
    }
    return SA_OK;
}

static int drain_timing(sa_solver *s) {
    if (!s->timing || s->ev_used == 0) return SA_OK;
    HIP_TRY(hipEventSynchronize(s->ev[s->ev_used - 1]));
    for (size_t q = 0; q < s->ev_used; q += 2) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev[q], s->ev[q + 1]));
        s->kernel_ms += ms;
    }
    s->ev_used = 0;
    return SA_OK;
}

// K(src * scale) -> dst (domain)
static int apply_K(sa_solver *s, const float *src, const float *scale, float *dst) {
    const Geo &g = s->geo;
    if (s->sch.rho_mode == 0) {
        if (!scale) return laplacian(s, src, dst);
        k_laplace<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, src, scale, dst);
        CHECK_LAUNCH();
        return SA_OK;
    }
    const float *b = s->sch.rho_mode == 2 ? s->b : nullptr;
    if (!scale && s->grid.dim == 3 && s->rho3d.enabled && rho3d_supported(g)) {
        // K u of the density forms in one pass (the fused density update's K-only form)
        Rho3DStep st;
        st.u = src;
        st.other = nullptr;
        st.out = dst;
        st.b = b;
        st.A = st.B = st.G = nullptr;
        st.save = nullptr;
        st.save_mode = 0;
        st.box = s->box;
        st.idt2 = 0.f;
        st.konly = 1;
        if (rho3d_step(s->rho3d, g, st, s->stream) != 0)
            return fail(SA_E_HIP, "K-only density launch: %s", hipGetErrorString(hipGetLastError()));
        return SA_OK;
    }
    for (int a = 0; a < 3; ++a) {
        if (!g.act[a]) continue;
        int ext[3] = {g.n[0], g.n[1], g.n[2]};
        ext[a] += 10;
        dim3 grid((ext[2] + kBlk.x - 1) / kBlk.x, (ext[1] + kBlk.y - 1) / kBlk.y, ext[0]);
        k_stagger<<<grid, kBlk, 0, s->stream>>>(g, a, src, scale, b, s->st[a]);
        CHECK_LAUNCH();
    }
    k_divergence<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->st[0], s->st[1], s->st[2], dst);
    CHECK_LAUNCH();
    return SA_OK;
}

// one time step of (1): out = step(u, other); optional fused save
__global__ void k_dest_map(int ndest, const long long *__restrict__ dest, int *map) {
    int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < ndest) map[dest[d]] = d;
}

static int dest_map(sa_solver *s, Sparse &sp) {
    if (sp.d_map || sp.ndest == 0) return SA_OK;
    if (hipMalloc((void **)&sp.d_map, s->total * sizeof(int)) != hipSuccess) {
        sp.d_map = nullptr;
        return fail(SA_E_NOMEM, "hipMalloc of the destination map failed");
    }
    HIP_TRY(hipMemsetAsync(sp.d_map, 0xff, s->total * sizeof(int), s->stream));
    k_dest_map<<<(sp.ndest + 255) / 256, 256, 0, s->stream>>>(sp.ndest, sp.d_dest, sp.d_map);
    CHECK_LAUNCH();
    return SA_OK;
}

// points of `sp` listed per tile of the fused 2D update, by the tile of their first tap
static int tile_points(sa_solver *s, Sparse &sp) {
    if (sp.d_toff) return SA_OK;
    const Geo &g = s->geo;
    const int ntz = (g.n[2] + kFlat2DTileZ - 1) / kFlat2DTileZ, nty = (g.n[1] + kFlat2DTileY - 1) / kFlat2DTileY;
    std::vector<int> tile(sp.npoint, -1), off(ntz * nty + 1, 0), pt;
    for (int r = 0; r < sp.npoint; ++r)
        for (int q = 0; q < sp.ntap; ++q) {
            const long long lin = sp.idx[(size_t)r * sp.ntap + q];
            if (lin < 0) continue;
            const int j = (int)(lin / g.n[2]), k = (int)(lin % g.n[2]);
            tile[r] = (j / kFlat2DTileY) * ntz + k / kFlat2DTileZ;
            break;
        }
    for (int r = 0; r < sp.npoint; ++r)
        if (tile[r] >= 0) ++off[tile[r] + 1];
    for (size_t i = 1; i < off.size(); ++i) off[i] += off[i - 1];
    pt.assign(std::max(off.back(), 1), 0);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int r = 0; r < sp.npoint; ++r)
        if (tile[r] >= 0) pt[fill[tile[r]]++] = r;
    HIP_TRY(hipMalloc((void **)&sp.d_toff, off.size() * sizeof(int)));
    HIP_TRY(hipMalloc((void **)&sp.d_tpt, pt.size() * sizeof(int)));
    HIP_TRY(h2d(sp.d_toff, off.data(), off.size() * sizeof(int)));
    HIP_TRY(h2d(sp.d_tpt, pt.data(), pt.size() * sizeof(int)));
    return SA_OK;
}

// The fused 2D update's arguments for one step (the one-shot loops and the lockstep shots build
// them alike). With power-2 attenuation the launch reads lap(other) from lapother (the caller
// forms it at the first step of a run) and writes lap(u) into lapnext; flat2d_lap_advance then
// makes it the next step's lapother.
static int flat2d_args(sa_solver *s, const float *u, const float *other, float *out, int backward, float *save,
                       int save_mode, const InjectSpec *inj, const GatherSpec *gat, const GradSpec *grd,
                       Flat2DStep &x) {
    const sa_scheme &c = s->sch;
    memset(&x, 0, sizeof(x));
    x.u = u;
    x.other = other;
    x.out = out;
    x.b = c.rho_mode == 2 ? s->b : nullptr;
    x.A = s->A;
    x.B = s->B;
    x.G = s->G;
    if (c.att_power == 2) {
        x.Cbeta = s->Cbeta;
        x.lapother = s->lapother;
        x.lapout = s->lapnext;
    }
    if (c.boundary == 1) {
        x.Cpml = s->Cpml;
        for (int a = 0; a < 3; ++a) {
            x.u1[a] = s->u1[a];
            x.u2[a] = s->u2[a];
            x.u3[a] = s->u3[a];
            for (int r = 0; r < 4; ++r) x.pml[a][r] = s->pml[a][r];
        }
    }
    x.dt = s->dt;
    x.backward = backward;
    x.save = save;
    x.save_mode = save_mode;
    x.box = s->box;
    x.idt2 = 1.f / (s->dt * s->dt);
    if (inj && inj->sp->ndest > 0) {
        TRY(dest_map(s, *inj->sp));
        x.imap = inj->sp->d_map;
        x.ioff = inj->sp->d_off;
        x.ipoint = inj->sp->d_point;
        x.icoef = inj->sp->d_coef;
        x.ivalues = inj->values;
        x.isign = inj->sign;
        x.isave = inj->into_save;
    }
    if (gat && gat->sp->npoint > 0) {
        TRY(tile_points(s, *gat->sp));
        x.goff = gat->sp->d_toff;
        x.gpt = gat->sp->d_tpt;
        x.gidx = gat->sp->d_idx;
        x.gw = gat->sp->d_w;
        x.gntap = gat->sp->ntap;
        x.gsign = gat->sign;
        x.gout = gat->out;
    }
    if (grd) {
        x.gps = grd->ps;
        x.gpsm = grd->psm;
        x.gpsp = grd->psp;
        x.gfdt2 = grd->fdt2;
        x.graw = grd->raw;
        x.ggv = grd->gv;
        x.gpv = grd->pv;
        x.ggr = grd->gr;
        x.gpr = grd->pr;
    }
    return SA_OK;
}

static void flat2d_lap_advance(sa_solver *s, const Flat2DStep &x, const float *u) {
    if (x.lapout) {
        std::swap(s->lapother, s->lapnext);
        s->lap_of = u;
    } else {
        s->lap_of = nullptr;
    }
}

// One time step. With `inj` given and the fused 2D update in use, the injection is applied by
// the update itself and *injected is set; otherwise the caller injects.
// kout (the fused density update only): K u of every node is also stored there; *kout_done tells
// whether it was.
static int step_general(sa_solver *s, const float *u, const float *other, float *out, int backward, float *save,
                        int save_mode, const InjectSpec *inj = nullptr, bool *injected = nullptr,
                        const GatherSpec *gat = nullptr, const GradSpec *grd = nullptr, float *kout = nullptr,
                        bool *kout_done = nullptr) {
    if (injected) *injected = false;
    if (kout_done) *kout_done = false;
    const Geo &g = s->geo;
    const sa_scheme &c = s->sch;
    hipEvent_t e0;
    TRY(timed_begin(s, &e0));
    if (use_rho3d(s)) {
        Rho3DStep st;
        st.u = u;
        st.other = other;
        st.out = out;
        st.b = c.rho_mode == 2 ? s->b : nullptr;
        st.A = s->A;
        st.B = s->B;
        st.G = s->G;
        st.save = save;
        st.save_mode = save_mode;
        st.box = s->box;
        st.idt2 = 1.f / (s->dt * s->dt);
        if (c.att_power == 2) {
            // the Laplacians of the two levels: the other level's kept from the previous step (or
            // taken here at the first step of a run), u's by a launch of its own (the next step's)
            if (!s->lapother || !s->lapnext || !s->Cbeta)
                return fail(SA_E_STATE, "power-2 attenuation buffers missing (prepare the scheme first)");
            TRY(lap_other(s, other));
            TRY(laplacian(s, u, s->lapnext));
            st.Cbeta = s->Cbeta;
            st.lapother = s->lapother;
            st.lapu = s->lapnext;
        }
        if (kout && c.att_power != 2 && c.kernel == 2) st.kout = kout;
        if (c.kernel == 4) {
            // OT4: pass 1 forms K u and w = m K u; the update (pass 2) stages w, so its K is
            // K(m K u), and reads u and K u as operands
            Rho3DStep p1 = st;
            p1.save = nullptr;
            p1.out = s->Ku;
            p1.ot4 = 1;
            p1.M = s->M;
            p1.wout = s->L2;
            if (rho3d_step(s->rho3d, g, p1, s->stream) != 0)
                return fail(SA_E_HIP, "fused density OT4 first pass launch: %s",
                            hipGetErrorString(hipGetLastError()));
            st.u = s->L2;
            st.ot4 = 2;
            st.uown = u;
            st.kuin = s->Ku;
            st.c4 = s->dt * s->dt / 12.f;
        }
        // CFS-PML: the auxiliary fields first; their term inside the update (or, with power-2
        // attenuation, by a launch of its own after it)
        const bool pml_in = c.boundary == 1 && c.att_power != 2 && c.kernel == 2 && !st.kout;
        if (c.boundary == 1) TRY(pml_before(s, u, backward));
        if (pml_in) {
            st.Cpml = s->Cpml;
            for (int a = 0; a < 3; ++a) st.pu1[a] = s->u1[a];
        }
        if (rho3d_step(s->rho3d, g, st, s->stream) != 0)
            return fail(SA_E_HIP, "fused density update launch: %s", hipGetErrorString(hipGetLastError()));
        if (c.boundary == 1 && !pml_in) TRY(pml_after(s, out));
        if (kout_done) *kout_done = st.kout != nullptr;
        if (c.att_power == 2) lap_advance(s, u);
        TRY(timed_end(s));
        return SA_OK;
    }
    if (use_flat2d(s)) {
        if (c.att_power == 2 && (s->lap_of != other || !s->lapnext)) {
            // `other` is overwritten in place, so its Laplacian must come from before: the
            // previous step's launch wrote it (as the Laplacian of that step's u), unless this is
            // the first step of a run
            k_laplace<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, other, nullptr, s->lapother);
            CHECK_LAUNCH();
        }
        Flat2DStep x;
        TRY(flat2d_args(s, u, other, out, backward, save, save_mode, inj, gat, grd, x));
        if (flat2d_step(g, c.rho_mode, x, s->stream) != 0)
            return fail(SA_E_HIP, "fused 2D update launch: %s", hipGetErrorString(hipGetLastError()));
        flat2d_lap_advance(s, x, u);
        if (injected && inj) *injected = true;
        TRY(timed_end(s));
        return SA_OK;
    }
    TRY(apply_K(s, u, nullptr, s->Ku));
    const float *F = u;
    if (c.kernel == 4) {
        TRY(apply_K(s, s->Ku, s->M, s->KmKu));
        k_scale_copy<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->Ku, s->M, s->L2);
        CHECK_LAUNCH();
        F = s->L2;
    }
    if (c.boundary == 1) {
        const float *Fp = c.kernel == 4 ? s->lapo : F;
        for (int a = 0; a < 3; ++a) {
            if (!g.act[a]) continue;
            k_pml<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, a, Fp, s->u1[a], s->u2[a], s->u3[a], s->pml[a][0],
                                                                s->pml[a][1], s->pml[a][2], s->pml[a][3], s->dt,
                                                                backward);
            CHECK_LAUNCH();
        }
    }
    UpdateArgs x;
    memset(&x, 0, sizeof(x));
    x.u = u;
    x.other = other;
    x.out = out;
    x.A = s->A;
    x.B = s->B;
    x.G = s->G;
    x.Ku = s->Ku;
    x.KmKu = c.kernel == 4 ? s->KmKu : nullptr;
    x.c4 = s->dt * s->dt / 12.f;
    x.Cbeta = c.att_power == 2 ? s->Cbeta : nullptr;
    if (c.att_power == 2) {
        // out may alias other (the new level overwrites the oldest in place): its Laplacian
        // is taken by a launch of its own before the update reads neighbours of it
        k_laplace<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, other, nullptr, s->lapother);
        CHECK_LAUNCH();
        x.lapother = s->lapother;
    }
    x.Cpml = c.boundary == 1 ? s->Cpml : nullptr;
    for (int a = 0; a < 3; ++a) x.u1[a] = c.boundary == 1 ? s->u1[a] : nullptr;
    x.save = save;
    x.save_mode = save_mode;
    x.box = s->box;
    x.idt2 = 1.f / (s->dt * s->dt);
    k_update<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, x);
    CHECK_LAUNCH();
    if (c.kernel == 4 && c.boundary == 1)
        HIP_TRY(hipMemcpyAsync(s->lapo, s->L2, s->total * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
    TRY(timed_end(s));
    return SA_OK;
}

static int coef_sparse(sa_solver *s, Sparse &sp, int kind) {
    if (sp.nnz == 0) return SA_OK;
    float scale = s->sch.diff_source ? (float)(2.0 * s->grid.dt * s->grid.dt / s->hmax)
                                     : (float)(2.0 * s->grid.dt / s->hmax);
    k_tap_coef<<<(sp.nnz + 255) / 256, 256, 0, s->stream>>>(sp.nnz, sp.d_lin, sp.d_cw, s->vp, s->G, kind,
                                                            scale, s->dt * s->dt, sp.d_coef);
    CHECK_LAUNCH();
    return SA_OK;
}

static int inject(sa_solver *s, Sparse &sp, const float *values, float sign, float *field, float *save,
                  float save_scale, const Box &box, int S = 1, int sh = 0) {
    if (sp.ndest == 0) return SA_OK;
    k_inject<<<(sp.ndest + 127) / 128, 128, 0, s->stream>>>(sp.ndest, sp.d_dest, sp.d_ijk, sp.d_off, sp.d_point,
                                                            sp.d_coef, values, sign, field, save, box, save_scale, S,
                                                            sh);
    CHECK_LAUNCH();
    return SA_OK;
}

static int gather(sa_solver *s, Sparse &sp, const float *field, float sign, float *out, int S = 1, int sh = 0) {
    if (sp.npoint == 0) return SA_OK;
    k_gather<<<(sp.npoint + 127) / 128, 128, 0, s->stream>>>(sp.npoint, sp.ntap, sp.d_idx, sp.d_w, field,
                                                             sign, out, S, sh);
    CHECK_LAUNCH();
    return SA_OK;
}

// injections of both shots of a two-shot pass (values[k]: shot k's samples of this step)
static int inject2(sa_solver *s, Sparse (&sp)[2], const float *const (&values)[2], float sign, float *field,
                   float *save, float save_scale, const Box &box) {
    const int n = std::max(sp[0].ndest, sp[1].ndest);
    if (n == 0) return SA_OK;
    InjectSet q[2];
    for (int k = 0; k < 2; ++k)
        q[k] = InjectSet{sp[k].ndest, sp[k].d_dest, sp[k].d_ijk, sp[k].d_off, sp[k].d_point, sp[k].d_coef, values[k]};
    k_inject2<<<dim3((n + 127) / 128, 2), 128, 0, s->stream>>>(q[0], q[1], sign, field, save, box, save_scale);
    CHECK_LAUNCH();
    return SA_OK;
}

static int gather2(sa_solver *s, Sparse (&sp)[2], const float *field, float sign, float *const (&out)[2]) {
    const int n = std::max(sp[0].npoint, sp[1].npoint);
    if (n == 0) return SA_OK;
    GatherSet q[2];
    for (int k = 0; k < 2; ++k) q[k] = GatherSet{sp[k].npoint, sp[k].ntap, sp[k].d_idx, sp[k].d_w, out[k]};
    k_gather2<<<dim3((n + 127) / 128, 2), 128, 0, s->stream>>>(q[0], q[1], field, sign);
    CHECK_LAUNCH();
    return SA_OK;
}

static int ensure_staging(sa_solver *s, float **p, size_t *have, size_t n) {
    if (*have >= n && *p) return SA_OK;
    dfree(*p);
    *p = nullptr;
    *have = 0;
    if (hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(float)) != hipSuccess)
        return fail(SA_E_NOMEM, "staging allocation failed");
    *have = n;
    return SA_OK;
}

static int alloc_saved(sa_solver *s) {
    const sa_scheme &c = s->sch;
    const Geo &g = s->geo;
    int f = c.save_mode == 3 ? 1 : c.undersampling;
    // vp-only gradient needs the interior only; everything else keeps the domain
    Box b;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = 0;
        b.n[a] = g.n[a];
    }
    if (c.save_mode == 1) {
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = g.nabs[a];
            b.n[a] = g.n[a] - 2 * g.nabs[a];
        }
    }
    b.off2 = b.lo[2] % 32;
    b.p2 = (b.off2 + b.n[2] + 31) / 32 * 32;
    int nsave = c.tb1 / f + 1;
    if (c.save_mode == 4) nsave = 2 * ((c.tb1 - 2) / f + 1);  // checkpoint pairs
    size_t per = (size_t)b.n[0] * b.n[1] * b.p2;
    size_t slots = nsave;
    if (c.save_mode == 5) {
        // snapshots first..last as mode 2 writes them; segments of K snapshots; states of segments >= 1
        s->ck_first = (std::max(1, c.tb0) + f - 1) / f;
        s->ck_last = std::min((c.tb1 - 2) / f, nsave - 1);
        s->ck_K = std::min(c.checkpoint, std::max(1, s->ck_last + 1));  // one segment at most
        s->ck_J = s->ck_last >= 0 ? s->ck_last / s->ck_K + 1 : 1;
        s->ck_seg = -1;
        slots = (size_t)std::max(1, 2 * (s->ck_J - 1));
    }
    size_t bytes = per * slots * sizeof(float);
    if (!s->saved || s->saved_bytes < bytes) {
        dfree(s->saved);
        s->saved = nullptr;
        s->saved_bytes = 0;
        if (hipMalloc((void **)&s->saved, bytes) != hipSuccess) {
            s->saved = nullptr;
            return fail(SA_E_NOMEM, "cannot allocate %.2f GB for the saved wavefield", bytes / 1e9);
        }
        s->saved_bytes = bytes;
    }
    // Slots the forward run never writes read as zero (as the reference's zero-initialised save
    // buffer). Every-f saves (modes 1, 2) write every box point of the slots t / f for
    // t % f == 0, max(1, tb0) <= t <= tb1 - 2, so only the slots outside that run are cleared;
    // the other modes clear the whole buffer.
    if (c.save_mode == 5) {
        // every state slot the adjoint reads is written by the forward run
    } else if (c.save_mode == 1 || c.save_mode == 2) {
        const int first = (std::max(1, c.tb0) + f - 1) / f;
        const int last = std::min((c.tb1 - 2) / f, nsave - 1);
        if (first > last) {
            HIP_TRY(hipMemsetAsync(s->saved, 0, bytes, s->stream));
        } else {
            if (first > 0) HIP_TRY(hipMemsetAsync(s->saved, 0, per * first * sizeof(float), s->stream));
            if (last + 1 < nsave)
                HIP_TRY(hipMemsetAsync(s->saved + per * (size_t)(last + 1), 0, per * (nsave - 1 - last) * sizeof(float),
                                       s->stream));
        }
    } else {
        HIP_TRY(hipMemsetAsync(s->saved, 0, bytes, s->stream));
    }
    s->box = b;
    s->nsave = nsave;
    s->saved_mode = c.save_mode;
    s->saved_on_host = false;
    s->saved_f = f;
    if (c.save_mode == 4 || c.save_mode == 5) {
        // save mode 4: the K + 1 recomputed levels of a segment in the padded layout (the adjoint
        // reads them with their halo); mode 5: K + 2 raw snapshots in the box layout
        size_t sb = c.save_mode == 4 ? s->total * (size_t)(f + 1) * sizeof(float)
                                     : per * (size_t)(s->ck_K + 2) * sizeof(float);
        if (!s->segment || s->segment_bytes < sb) {
            dfree(s->segment);
            s->segment = nullptr;
            if (hipMalloc((void **)&s->segment, sb) != hipSuccess) {
                s->segment = nullptr;
                return fail(SA_E_NOMEM, "cannot allocate %.2f GB for the recomputation segment", sb / 1e9);
            }
            s->segment_bytes = sb;
        }
        // the recomputed levels of save mode 4 are read with their halo, which no step writes
        if (c.save_mode == 4) HIP_TRY(hipMemsetAsync(s->segment, 0, sb, s->stream));
        // save mode 4 with the fused density update: a K u slot per recomputed level, so the
        // exact adjoint reads K u_t instead of forming it again
        s->kseg_valid = false;
        const char *noku = getenv("STRIDE_AMD_NO_KU_REUSE");
        if (c.save_mode == 4 && use_rho3d(s) && c.att_power != 2 && !(noku && noku[0] == '1')) {
            if (!s->kseg || s->kseg_bytes < sb) {
                dfree(s->kseg);
                s->kseg = nullptr;
                s->kseg_bytes = 0;
                if (hipMalloc((void **)&s->kseg, sb) == hipSuccess) s->kseg_bytes = sb;
                else s->kseg = nullptr;  // not enough memory: the adjoint forms K u itself
                (void)hipGetLastError();
            }
        }
        TRY(dalloc(&s->fb[0], s->total));
        TRY(dalloc(&s->fb[1], s->total));
    }
    return SA_OK;
}

// floats per snapshot slot (rows padded to the pitch)
static size_t box_points(const Box &b) { return (size_t)b.n[0] * b.n[1] * b.p2; }

static int dump_alloc(sa_solver *s, int w);
static int dump_level(sa_solver *s, int w, int t, const float *u);

// Power-2 attenuation on the specialised 3D update: the update reads the Laplacian of the level it
// overwrites (kept from the previous step, or taken here at the first step of a run) and writes
// the Laplacian of u, which becomes the next step's.
static int lap_other(sa_solver *s, const float *other) {
    if (!s->lapother) return fail(SA_E_STATE, "power-2 attenuation buffers missing");
    if (s->lap_of != other) TRY(laplacian(s, other, s->lapother));
    return SA_OK;
}

static void lap_advance(sa_solver *s, const float *u) {
    std::swap(s->lapother, s->lapnext);
    s->lap_of = u;
}

// One step of the specialised 3D update (use_fast): out = step(u, o) (out may alias o) with the
// scheme's attenuation, CFS-PML and OT4 forms; forward: an optional fused save into `save` (mode 1
// p.dt2, 2 raw); backward (the adjoints' steps): an optional fused vp gradient from the p.dt2
// snapshot ps.
static int fast_step(sa_solver *s, const float *u, const float *o, float *out, int backward, float *save, int mode,
                     const float *ps, float *gv, float *pv) {
    const sa_scheme &c = s->sch;
    FastAtt at{}, *pat = nullptr;
    if (c.att_power != -2) {
        if (c.att_power == 2) TRY(lap_other(s, o));
        at = {c.att_power, s->alpha, s->lapother, s->lapnext};
        pat = &at;
    }
    if (c.boundary == 1) {
        // the auxiliary fields first; the update applies their term
        TRY(pml_before(s, u, backward));
        pml_term(s, at, pat);
    }
    const float *staged = u;
    if (c.kernel == 4) TRY(ot4_first(s, u, at, pat, staged));
    hipEvent_t e0;
    TRY(timed_begin(s, &e0));
    const int rc = backward ? fast3d_step_adjoint(s->fast, s->geo, staged, o, out, ps, s->box, gv, pv, s->stream, pat)
                            : fast3d_step(s->fast, s->geo, staged, o, out, save, s->box, 1.f / (s->dt * s->dt),
                                          s->stream, pat, mode == 2 ? 2 : 1);
    if (rc != 0) return fail(SA_E_HIP, "specialised 3D update launch: %s", hipGetErrorString(hipGetLastError()));
    TRY(timed_end(s));
    if (c.att_power == 2) lap_advance(s, u);
    return SA_OK;
}

// One forward step t: p[t+1] into o from u = p[t] and o = p[t-1], then the source injection.
// The forward run and the checkpointed recomputation issue the same launches through it, so the
// recomputed levels equal the forward run's bit for bit.
static int forward_step(sa_solver *s, int t, float *u, float *o, const float *src, float *save, int mode,
                        const GatherSpec *gat) {
    const float idt2 = 1.f / (s->dt * s->dt);
    bool injected = false;
    if (use_fast(s)) {
        TRY(fast_step(s, u, o, o, 0, save, mode, nullptr, nullptr, nullptr));
    } else {
        const InjectSpec spec{&s->src, src + (size_t)t * s->src.npoint, 1.f, mode == 1};
        TRY(step_general(s, u, o, o, 0, save, mode, &spec, &injected, gat));
    }
    if (!injected)
        TRY(inject(s, s->src, src + (size_t)t * s->src.npoint, 1.f, o, mode == 1 ? save : nullptr, idt2, s->box));
    return SA_OK;
}

extern "C" int sa_forward_dev(sa_solver *s, const float *src, float *rec) {
    if (!s) return fail(SA_E_ARG, "null solver");
    TRY(prepare(s));
    s->lap_of = nullptr;
    const sa_scheme &c = s->sch;
    TRY(coef_sparse(s, s->src, 0));
    TRY(zero_field(s, s->buf[0]));
    TRY(zero_field(s, s->buf[1]));
    if (c.boundary == 1)
        for (int a = 0; a < 3; ++a)
            if (s->geo.act[a]) {
                TRY(zero_field(s, s->u1[a]));
                TRY(zero_field(s, s->u2[a]));
                TRY(zero_field(s, s->u3[a]));
            }
    if (s->lapo) TRY(zero_field(s, s->lapo));
    if (s->rec.npoint > 0) HIP_TRY(hipMemsetAsync(rec, 0, sizeof(float) * (size_t)s->grid.nt * s->rec.npoint, s->stream));
    if (c.save_mode != 0) TRY(alloc_saved(s));
    TRY(dump_alloc(s, 0));
    const int f = c.save_mode == 3 ? 1 : c.undersampling;
    const size_t per = box_points(s->box);
    const bool fast = use_fast(s);
    const bool ckpt = c.save_mode == 4, raw_ckpt = c.save_mode == 5;
    if (ckpt || raw_ckpt) {
        size_t ns = (size_t)s->grid.nt * s->src.npoint;
        TRY(ensure_staging(s, &s->src_keep, &s->src_keep_n, ns));
        if (ns) HIP_TRY(hipMemcpyAsync(s->src_keep, src, ns * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
    }
    // levels: buf[t % 2] = p[t]; step t overwrites p[t-1] with p[t+1]. With the fused 2D update,
    // the receivers of step t are interpolated by step t + 1's launch (its u is that level).
    const bool defer = !fast && use_flat2d(s) && s->rec.npoint > 0;
    float *pending = nullptr;  // receiver row of the previous step, not yet interpolated
    int tlast = c.tb1 - 2;
    for (int t = 1; t <= tlast; ++t) {
        float *u = s->buf[t % 2], *o = s->buf[(t + 1) % 2];
        float *save = nullptr;
        int mode = 0;
        if (c.save_mode != 0 && !ckpt && !raw_ckpt && t % f == 0 && t >= c.tb0 && t <= c.tb1 && t / f < s->nsave) {
            save = s->saved + (size_t)(t / f) * per;
            mode = c.save_mode == 1 ? 1 : 2;
        }
        TRY(dump_level(s, 0, t, u));
        const GatherSpec gspec{&s->rec, 1.f, pending};
        TRY(forward_step(s, t, u, o, src, save, mode, pending ? &gspec : nullptr));
        if (defer)
            pending = rec + (size_t)t * s->rec.npoint;
        else
            TRY(gather(s, s->rec, o, 1.f, rec + (size_t)t * s->rec.npoint));
        if (ckpt && t % f == 0) {
            // checkpoint j = t / K holds (p[jK + 1], p[jK])
            float *slot = s->saved + (size_t)(2 * (t / f)) * per;
            k_pack<<<domain_grid(s->geo, kBlk), kBlk, 0, s->stream>>>(s->geo, s->box, o, slot);
            k_pack<<<domain_grid(s->geo, kBlk), kBlk, 0, s->stream>>>(s->geo, s->box, u, slot + per);
            CHECK_LAUNCH();
        }
        if (raw_ckpt && (t + 1) % f == 0 && t + 1 >= 2 && ((t + 1) / f + 1) % s->ck_K == 0) {
            // the state of segment j, (p[T_j], p[T_j - 1]) with T_j = (jK - 1) f = t + 1
            const int j = ((t + 1) / f + 1) / s->ck_K;
            if (j >= 1 && j < s->ck_J) {
                float *slot = s->saved + (size_t)(2 * (j - 1)) * per;
                k_pack<<<domain_grid(s->geo, kBlk), kBlk, 0, s->stream>>>(s->geo, s->box, o, slot);
                k_pack<<<domain_grid(s->geo, kBlk), kBlk, 0, s->stream>>>(s->geo, s->box, u, slot + per);
                CHECK_LAUNCH();
            }
        }
    }
    if (pending) TRY(gather(s, s->rec, s->buf[(tlast + 1) % 2], 1.f, pending));
    s->cur = (tlast + 1) % 2;
    return SA_OK;
}

// Slot of level t in the recomputed segment j (save mode 4): padded fields, zero halo.
static float *segment_level(sa_solver *s, int j, int t) {
    return s->segment + (size_t)(t - j * s->saved_f) * s->total;
}

// Slot of K u_t in segment j (save mode 4 with the fused density update: the recomputation's
// update stores the K u it forms, which the exact adjoint's accumulation at step t reads).
static float *segment_ku(sa_solver *s, int j, int t) {
    return s->kseg + (size_t)(t - j * s->saved_f) * s->total;
}

// Recompute forward levels u_{jK} .. u_{min((j+1)K, nt-1)} into the segment buffer from checkpoint j.
// Every step writes its new level into its own slot (the levels stay in the padded layout, so the
// adjoint reads them without unpacking).
static int recompute_segment(sa_solver *s, int j) {
    s->lap_of = nullptr;
    const int K = s->saved_f;
    const int nt = s->grid.nt;
    const size_t per = box_points(s->box);
    const Geo &g = s->geo;
    const float *pair = s->saved + (size_t)(2 * j) * per;  // (p[jK+1], p[jK])
    int t0 = j * K;
    int tend = std::min((j + 1) * K, nt - 1);
    k_unpack<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->box, pair + per, segment_level(s, j, t0));
    if (tend >= t0 + 1)
        k_unpack<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->box, pair, segment_level(s, j, t0 + 1));
    CHECK_LAUNCH();
    // K u of levels t0 + 1 .. tend - 1 (formed by their steps) when every step stored it
    bool ku = s->kseg != nullptr;
    for (int t = t0 + 1; t < tend; ++t) {
        float *out = segment_level(s, j, t + 1);
        bool done = false;
        if (use_fast(s))  // the forward run's own steps, so the recomputed levels equal its bits
            TRY(fast_step(s, segment_level(s, j, t), segment_level(s, j, t - 1), out, 0, nullptr, 0, nullptr, nullptr,
                          nullptr));
        else
            TRY(step_general(s, segment_level(s, j, t), segment_level(s, j, t - 1), out, 0, nullptr, 0, nullptr,
                             nullptr, nullptr, nullptr, s->kseg ? segment_ku(s, j, t) : nullptr, &done));
        ku = ku && done;
        TRY(inject(s, s->src, s->src_keep + (size_t)t * s->src.npoint, 1.f, out, nullptr, 0.f, s->box));
    }
    s->kseg_valid = ku;
    return SA_OK;
}

// Checkpointed raw saves: recompute snapshots jK - 1 .. (j + 1)K of segment j (p[i f] for
// snapshot i) into the ring, from the state (p[T_j], p[T_j - 1]), T_j = (jK - 1) f, or from the
// zero state before step 1. Snapshots outside [ck_first, ck_last] are zero.
static int recompute_raw_segment(sa_solver *s, int j) {
    s->lap_of = nullptr;
    const int f = s->saved_f, K = s->ck_K, R = K + 2;
    const size_t per = box_points(s->box);
    const Geo &g = s->geo;
    auto slot = [&](int i) { return s->segment + (size_t)(i % R) * per; };
    const int ilo = std::max(j * K - 1, 0), ihi = (j + 1) * K;
    for (int i = ilo; i <= ihi; ++i)
        if (i < s->ck_first || i > s->ck_last) HIP_TRY(hipMemsetAsync(slot(i), 0, per * sizeof(float), s->stream));
    const int vlo = std::max(ilo, s->ck_first), vhi = std::min(ihi, s->ck_last);
    s->ck_seg = j;
    if (vlo > vhi) return SA_OK;
    const int T = (j * K - 1) * f;
    int t0 = 1;
    if (j >= 1 && T >= 2) {
        const float *state = s->saved + (size_t)(2 * (j - 1)) * per;  // (p[T], p[T - 1])
        k_unpack<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->box, state, s->fb[T % 2]);
        k_unpack<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->box, state + per, s->fb[(T - 1) % 2]);
        CHECK_LAUNCH();
        t0 = T;
    } else {
        TRY(zero_field(s, s->fb[0]));
        TRY(zero_field(s, s->fb[1]));
    }
    const int tend = vhi * f;
    for (int t = t0; t < tend; ++t) {
        float *u = s->fb[t % 2], *o = s->fb[(t + 1) % 2];
        if (t % f == 0 && t / f >= vlo) {
            k_pack<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->box, u, slot(t / f));
            CHECK_LAUNCH();
        }
        TRY(forward_step(s, t, u, o, s->src_keep, nullptr, 0, nullptr));
    }
    k_pack<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->box, s->fb[tend % 2], slot(vhi));
    CHECK_LAUNCH();
    s->lap_of = nullptr;
    return SA_OK;
}

extern "C" int sa_forward(sa_solver *s, const float *src, float *rec) {
    if (!s || !src || !rec) return fail(SA_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(s->device));
    size_t ns = (size_t)s->grid.nt * s->src.npoint, nr = (size_t)s->grid.nt * s->rec.npoint;
    TRY(ensure_staging(s, &s->d_series, &s->series_n, ns));
    TRY(ensure_staging(s, &s->d_out, &s->out_n, nr));
    if (ns) HIP_TRY(hipMemcpyAsync(s->d_series, src, ns * sizeof(float), hipMemcpyHostToDevice, s->stream));
    TRY(sa_forward_dev(s, s->d_series, s->d_out));
    if (nr) HIP_TRY(hipMemcpyAsync(rec, s->d_out, nr * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    TRY(drain_timing(s));
    return SA_OK;
}

static int zero_aux(sa_solver *s) {
    s->lap_of = nullptr;
    TRY(zero_field(s, s->buf[0]));
    TRY(zero_field(s, s->buf[1]));
    if (s->sch.boundary == 1)
        for (int a = 0; a < 3; ++a)
            if (s->geo.act[a]) {
                TRY(zero_field(s, s->u1[a]));
                TRY(zero_field(s, s->u2[a]));
                TRY(zero_field(s, s->u3[a]));
            }
    if (s->lapo) TRY(zero_field(s, s->lapo));
    return SA_OK;
}

// compat adjoint with device-resident outputs (padded accumulators gv, pv, gr, pr; gsrc (nt, nsrc))
static int adjoint_loop(sa_solver *s, const float *res, float *gv, float *pv, float *gr, float *pr, float *gsrc) {
    const sa_scheme &c = s->sch;
    const Geo &g = s->geo;
    if (s->saved_on_host) return fail(SA_E_STATE, "the saved wavefield is in host memory: sa_saved_from_host first");
    if (!s->saved || s->saved_mode == 0 || s->saved_mode == 3 || s->saved_mode == 4)
        return fail(SA_E_STATE, "the adjoint needs a forward run that saved the wavefield (save_mode 1, 2 or 5)");
    const bool raw = s->saved_mode == 2 || s->saved_mode == 5, raw_ckpt = s->saved_mode == 5;
    if (gr && !raw) return fail(SA_E_STATE, "the density gradient needs the raw wavefield (save_mode 2 or 5)");
    TRY(coef_sparse(s, s->rec, 1));
    if (raw_ckpt) s->ck_seg = -1;
    TRY(zero_aux(s));
    TRY(dump_alloc(s, 1));
    const int f = s->saved_f;
    const size_t per = box_points(s->box);
    const float fdt2 = (float)(1.0 / ((double)f * s->grid.dt * (double)f * s->grid.dt));
    if (gr) {
        TRY(dalloc(&s->work0, s->total));
    }
    const bool fast = use_fast(s);
    // the fused 2D update takes the gradients itself: vp (any density form) and rho (variable
    // density, interior at least 5 cells from the edges, so the snapshot stencils stay inside)
    bool grad2d = !fast && use_flat2d(s) && (s->saved_mode == 1 || raw);
    if (gr) grad2d = grad2d && c.rho_mode == 2 && raw && g.nabs[1] >= 5 && g.nabs[2] >= 5;
    // snapshot i: slot i of `saved`, or slot i % (K + 2) of the recomputed ring
    auto snap = [&](int i) -> const float * {
        return raw_ckpt ? s->segment + (size_t)(i % (s->ck_K + 2)) * per : s->saved + (size_t)i * per;
    };
    int tstart = c.tb1 - 1 - f, tend = c.tb0 + 1;
    // buf[t % 2] = p_a[t]; step t overwrites p_a[t+1] with p_a[t-1]
    for (int t = tstart; t >= tend; --t) {
        float *u = s->buf[t % 2], *o = s->buf[(t + 1) % 2];
        bool injected = false, grads_done = false;
        bool guard = t % f == 0 && t >= c.tb0 && t <= c.tb1 && t / f < s->nsave && (gv || gr);
        TRY(dump_level(s, 1, t, u));
        if (guard && raw_ckpt && (t / f) / s->ck_K != s->ck_seg) TRY(recompute_raw_segment(s, (t / f) / s->ck_K));
        const float *ps = guard ? snap(t / f) : nullptr;
        if (fast) {
            TRY(fast_step(s, u, o, o, 1, nullptr, 0, s->saved_mode == 1 ? ps : nullptr, gv, pv));
        } else {
            const InjectSpec spec{&s->rec, res + (size_t)t * s->rec.npoint, -1.f, false};
            // the fused 2D update also accumulates this step's gradients (it reads u and the
            // snapshots only, as the separate passes after the injection do)
            GradSpec gspec{};
            if (guard && grad2d) {
                gspec.ps = ps;
                gspec.raw = raw;
                if (gspec.raw) {
                    gspec.psm = t / f >= 1 ? snap(t / f - 1) : nullptr;
                    gspec.psp = t / f + 1 < s->nsave ? snap(t / f + 1) : nullptr;
                }
                gspec.fdt2 = fdt2;
                gspec.gv = gv;
                gspec.pv = pv;
                gspec.gr = gr;
                gspec.pr = pr;
            }
            TRY(step_general(s, u, o, o, 1, nullptr, 0, &spec, &injected, nullptr,
                             guard && grad2d ? &gspec : nullptr));
            grads_done = guard && grad2d;
        }
        if (!injected) TRY(inject(s, s->rec, res + (size_t)t * s->rec.npoint, -1.f, o, nullptr, 0.f, s->box));
        if (gsrc) TRY(gather(s, s->src, u, -1.f, gsrc + (size_t)t * s->src.npoint));
        if (!guard || grads_done) continue;
        if (gr && raw && s->b && use_rho3d(s) && rho3d_grad_supported(g, s->box)) {
            // vp and rho gradients in one pass over the raw snapshots
            const float *psm = t / f >= 1 ? snap(t / f - 1) : nullptr;
            const float *psp = t / f + 1 < s->nsave ? snap(t / f + 1) : nullptr;
            if (rho3d_grad(g, s->box, ps, psm, psp, fdt2, s->b, u, gv, pv, gr, pr, s->stream) != 0)
                return fail(SA_E_HIP, "density gradient launch: %s", hipGetErrorString(hipGetLastError()));
            continue;
        }
        if (gv && !(fast && s->saved_mode == 1)) {
            const float *psm = nullptr, *psp = nullptr;
            if (raw) {
                psm = t / f >= 1 ? snap(t / f - 1) : nullptr;
                psp = t / f + 1 < s->nsave ? snap(t / f + 1) : nullptr;
                if (!psm || !psp) {
                    // out-of-range neighbours read as zero: stage a zero snapshot
                    TRY(dalloc(&s->work1, std::max<long long>(s->total, (long long)per)));
                    HIP_TRY(hipMemsetAsync(s->work1, 0, per * sizeof(float), s->stream));
                    if (!psm) psm = s->work1;
                    if (!psp) psp = s->work1;
                }
            }
            k_grad_vp<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->box, ps, psm, psp, fdt2, u, gv, pv);
            CHECK_LAUNCH();
        }
        if (gr) {
            k_unpack<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->box, ps, s->work0);
            CHECK_LAUNCH();
            k_grad_rho<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->work0, s->b, u, gr, pr);
            CHECK_LAUNCH();
        }
    }
    s->cur = (tend - 1 + 2) % 2;
    return SA_OK;
}

static int grad_outputs(sa_solver *s, int n, float **dev) {
    for (int q = 0; q < n; ++q) TRY(dalloc(&dev[q], s->total));
    for (int q = 0; q < n; ++q) TRY(zero_field(s, dev[q]));
    return SA_OK;
}

extern "C" int sa_adjoint_dev(sa_solver *s, const float *res, int want, float *grad_src) {
    if (!s) return fail(SA_E_ARG, "null solver");
    HIP_TRY(hipSetDevice(s->device));
    TRY(prepare(s));
    bool want_v = want & SA_WANT_VP, want_r = want & SA_WANT_RHO;
    if (want_v) TRY(grad_outputs(s, 2, s->gacc));
    if (want_r) TRY(grad_outputs(s, 2, s->gacc + 2));
    if (grad_src && s->src.npoint > 0)
        HIP_TRY(hipMemsetAsync(grad_src, 0, sizeof(float) * (size_t)s->grid.nt * s->src.npoint, s->stream));
    return adjoint_loop(s, res, want_v ? s->gacc[0] : nullptr, want_v ? s->gacc[1] : nullptr,
                        want_r ? s->gacc[2] : nullptr, want_r ? s->gacc[3] : nullptr, grad_src);
}

extern "C" int sa_read_gradient(sa_solver *s, int which, float *out) {
    if (!s || !out || which < 0 || which > 3) return fail(SA_E_ARG, "bad arguments");
    if (!s->gacc[which]) return fail(SA_E_STATE, "gradient %d was not computed", which);
    HIP_TRY(hipSetDevice(s->device));
    return download_unpadded(s, out, s->gacc[which]);
}

extern "C" int sa_adjoint(sa_solver *s, const float *res, float *grad_vp, float *prec_vp, float *grad_rho,
                          float *prec_rho, float *grad_src) {
    if (!s || !res) return fail(SA_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(s->device));
    TRY(prepare(s));
    size_t nr = (size_t)s->grid.nt * s->rec.npoint, ns = (size_t)s->grid.nt * s->src.npoint;
    TRY(ensure_staging(s, &s->d_series, &s->series_n, nr));
    TRY(ensure_staging(s, &s->d_out, &s->out_n, ns));
    if (nr) HIP_TRY(hipMemcpyAsync(s->d_series, res, nr * sizeof(float), hipMemcpyHostToDevice, s->stream));
    int want = ((grad_vp || prec_vp) ? SA_WANT_VP : 0) | ((grad_rho || prec_rho) ? SA_WANT_RHO : 0);
    TRY(sa_adjoint_dev(s, s->d_series, want, grad_src ? s->d_out : nullptr));
    float *outs[4] = {grad_vp, prec_vp, grad_rho, prec_rho};
    for (int q = 0; q < 4; ++q)
        if (outs[q]) TRY(download_unpadded(s, outs[q], s->gacc[q]));
    if (grad_src && ns) HIP_TRY(hipMemcpyAsync(grad_src, s->d_out, ns * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return drain_timing(s);
}

// the exact adjoint with the residual res_dev (nt, nrec) and the per-tap field lam_dev
// (nt, nsrc * ntap) on the device; acc[5] receives the padded accumulators (gA, gB, gG, and gm /
// gb when the scheme has them), which the caller frees
static int adjoint_exact_core(sa_solver *s, const float *res_dev, float *lam_dev, float *(&acc)[5]) {
    HIP_TRY(hipSetDevice(s->device));
    TRY(prepare(s));
    const sa_scheme &c = s->sch;
    const Geo &g = s->geo;
    if (c.boundary == 1 || c.att_power == 2 || (c.kernel == 4 && c.rho_mode == 2))
        return fail(SA_E_UNSUPPORTED, "exact adjoint supports sponge/no boundary, attenuation none/0/None, "
                                      "and OT4 only without variable density");
    if (!s->saved || (s->saved_mode != 3 && s->saved_mode != 4))
        return fail(SA_E_STATE, "the exact adjoint needs a forward run with save_mode 3 or 4");
    const bool ckpt = s->saved_mode == 4;
    const int nt = s->grid.nt;
    size_t nq = (size_t)s->src.npoint * s->src.ntap;
    if (nq) HIP_TRY(hipMemsetAsync(lam_dev, 0, (size_t)nt * nq * sizeof(float), s->stream));
    TRY(coef_sparse(s, s->rec, 2));
    TRY(zero_aux(s));
    for (auto &a : acc) a = nullptr;
    TRY(grad_outputs(s, 3, acc));
    if (c.kernel == 4) TRY(grad_outputs(s, 1, acc + 3));
    if (c.rho_mode == 2) TRY(grad_outputs(s, 1, acc + 4));
    TRY(dalloc(&s->work0, s->total));
    TRY(dalloc(&s->work1, s->total));
    float *KuT = nullptr, *Kmu = nullptr;
    TRY(dalloc(&KuT, s->total));
    if (c.kernel == 4) TRY(dalloc(&Kmu, s->total));
    const size_t per = box_points(s->box);
    const float c4 = s->dt * s->dt / 12.f;
    int rc = SA_OK;
    int seg = -1;
    // buf[lv % 2] = mu_lv
    for (int lv = nt - 1; lv >= 1 && rc == SA_OK; --lv) {
        float *m1 = s->buf[(lv + 1) % 2], *m2 = s->buf[lv % 2];  // mu_{lv+1}, mu_{lv+2}
        if (lv >= 2) {
            // mu_lv over mu_{lv+2}
            rc = use_fast(s) ? fast_step(s, m1, m2, m2, 1, nullptr, 0, nullptr, nullptr, nullptr)
                             : step_general(s, m1, m2, m2, 1, nullptr, 0);
            if (rc == SA_OK && lv - 1 >= 1 && lv - 1 <= nt - 2)
                rc = inject(s, s->rec, res_dev + (size_t)(lv - 1) * s->rec.npoint, 1.f, m2, nullptr, 0.f, s->box);
        }
        int t = lv;
        if (rc != SA_OK || t < 1 || t > nt - 2) continue;
        const float *ut = nullptr, *utm = nullptr;
        if (ckpt) {
            const int K = s->saved_f;
            const int j = (t - 1) / K;
            if (j != seg) {
                // the recomputation uses the forward coefficients of the source taps
                rc = coef_sparse(s, s->src, 0);
                if (rc == SA_OK) rc = recompute_segment(s, j);
                if (rc != SA_OK) break;
                seg = j;
            }
            ut = segment_level(s, j, t);
            utm = segment_level(s, j, t - 1);
        } else {
            ut = s->saved + (size_t)t * per;
            utm = s->saved + (size_t)(t - 1) * per;
        }
        if (nq) {
            k_gather_taps<<<((int)nq + 127) / 128, 128, 0, s->stream>>>((int)nq, s->src.d_idx, m1, s->G,
                                                                         lam_dev + (size_t)t * nq);
        }
        // u_t and u_{t-1} in the padded layout: the recomputed slots as they are, stored levels unpacked
        const float *u0 = ut, *u1 = utm;
        if (!ckpt) {
            k_unpack<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->box, ut, s->work0);
            k_unpack<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->box, utm, s->work1);
            u0 = s->work0;
            u1 = s->work1;
        }
        if ((rc = (hipGetLastError() == hipSuccess ? SA_OK : fail(SA_E_HIP, "launch failed"))) != SA_OK) break;
        // K u_t: stored by the segment's recomputation for its levels t0 + 1 .. tend - 1, else formed
        const float *Ku = KuT;
        if (ckpt && s->kseg && s->kseg_valid && c.kernel == 2) {
            const int K = s->saved_f, j = (t - 1) / K;
            const int tend = std::min((j + 1) * K, nt - 1);
            if (t >= j * K + 1 && t <= tend - 1) Ku = segment_ku(s, j, t);
        }
        if (Ku == KuT && (rc = apply_K(s, u0, nullptr, KuT)) != SA_OK) break;
        if (c.kernel == 4) {
            // gm += dt^2/12 (K mu)(K u); Ku += dt^2/12 K(m Ku)
            if ((rc = apply_K(s, m1, nullptr, Kmu)) != SA_OK) break;
            k_exact_acc_m<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, Kmu, KuT, c4, acc[3]);
            if ((rc = apply_K(s, KuT, s->M, Kmu)) != SA_OK) break;
            k_axpy_domain<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, c4, Kmu, KuT);
        }
        // gA, gB, gG and (variable density) gb, with the staggered derivatives of u_t and
        // mu_{t+1} (m1) formed in the same pass (3D: a march with LDS tiles)
        float *gbp = c.rho_mode == 2 ? acc[4] : nullptr;
        if (s->grid.dim == 3 && rho3d_supported(g)) {
            if (rho3d_exact_acc(g, u0, m1, s->G, u1, Ku, acc[0], acc[1], acc[2], gbp, s->stream) != 0)
                rc = fail(SA_E_HIP, "exact accumulation launch: %s", hipGetErrorString(hipGetLastError()));
        } else {
            k_exact_acc_fused<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, m1, s->G, u0, u1, Ku, acc[0],
                                                                              acc[1], acc[2], gbp);
        }
        if (hipGetLastError() != hipSuccess) rc = fail(SA_E_HIP, "launch failed");
    }
    dfree(KuT);
    dfree(Kmu);
    s->cur = 2 % 2;
    return rc;
}

extern "C" int sa_adjoint_exact(sa_solver *s, const float *res, float *gA, float *gB, float *gG, float *gm,
                                float *gb, float *lam_src) {
    if (!s || !res) return fail(SA_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(s->device));
    const int nt = s->grid.nt;
    size_t nr = (size_t)nt * s->rec.npoint;
    size_t nq = (size_t)s->src.npoint * s->src.ntap;
    TRY(ensure_staging(s, &s->d_series, &s->series_n, nr));
    TRY(ensure_staging(s, &s->d_out, &s->out_n, (size_t)nt * nq));
    if (nr) HIP_TRY(hipMemcpyAsync(s->d_series, res, nr * sizeof(float), hipMemcpyHostToDevice, s->stream));
    float *acc[5] = {};
    int rc = adjoint_exact_core(s, s->d_series, s->d_out, acc);
    float *outs[5] = {gA, gB, gG, gm, gb};
    for (int q = 0; q < 5 && rc == SA_OK; ++q)
        if (outs[q] && acc[q]) rc = download_unpadded(s, outs[q], acc[q]);
    if (rc == SA_OK && lam_src && nq) {
        if (hipStreamSynchronize(s->stream) != hipSuccess ||
            hipMemcpy(lam_src, s->d_out, (size_t)nt * nq * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(SA_E_HIP, "copy of lam_src failed");
    }
    for (float *p : acc) dfree(p);
    if (rc == SA_OK) rc = drain_timing(s);
    return rc;
}

// device-resident form: res (nt, nrec), lam_src (nt, nsrc * ntap) and the accumulators
// (unpadded extended grid; any may be NULL) are device pointers; queued on the solver's stream
extern "C" int sa_adjoint_exact_dev(sa_solver *s, const float *res, float *gA, float *gB, float *gG, float *gm,
                                    float *gb, float *lam_src) {
    if (!s || !res) return fail(SA_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(s->device));
    const int nt = s->grid.nt;
    size_t nq = (size_t)s->src.npoint * s->src.ntap;
    float *lam = lam_src;
    if (!lam) {
        TRY(ensure_staging(s, &s->d_out, &s->out_n, (size_t)nt * nq));
        lam = s->d_out;
    }
    float *acc[5] = {};
    int rc = adjoint_exact_core(s, res, lam, acc);
    float *outs[5] = {gA, gB, gG, gm, gb};
    for (int q = 0; q < 5 && rc == SA_OK; ++q)
        if (outs[q] && acc[q]) {
            k_unpad<<<domain_grid(s->geo, kBlk), kBlk, 0, s->stream>>>(s->geo, acc[q], outs[q]);
            if (hipGetLastError() != hipSuccess) rc = fail(SA_E_HIP, "launch failed");
        }
    for (float *p : acc) dfree(p);  // a device free waits for the queued work
    return rc;
}

// The chain rule of the exact adjoint on the device: dJ/dvp, dJ/drho, dJ/dalpha (alpha_scale:
// nepers per caller unit) from the accumulators of sa_adjoint_exact_dev (unpadded; gm / gb may be
// NULL) and the solver's medium, boundary and source taps; with lam_src and src (device,
// (nt, nsrc * ntap) and (nt, nsrc)) the source taps' term is added to dJ/dvp and dJ/dsrc is
// written to g_src (nt, nsrc). Outputs unpadded extended grid, device; any may be NULL.
extern "C" int sa_exact_gradients_dev(sa_solver *s, const float *gA, const float *gB, const float *gG,
                                      const float *gm, const float *gb, const float *lam_src, const float *src,
                                      float *g_vp, float *g_rho, float *g_alpha, float *g_src, double alpha_scale) {
    if (!s || !gA || !gB || !gG) return fail(SA_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(s->device));
    TRY(prepare(s));
    const sa_scheme &c = s->sch;
    if (g_rho && c.rho_mode != 2) return fail(SA_E_ARG, "a density gradient needs variable density");
    if (g_alpha && c.att_power == -2) return fail(SA_E_ARG, "an alpha gradient needs attenuation");
    k_exact_chain<<<domain_grid(s->geo, kBlk), kBlk, 0, s->stream>>>(
        s->geo, gA, gB, gG, gm, gb, s->vp, s->rho, s->alpha, s->tables, s->tlen, c.boundary, c.att_power,
        c.rho_mode, (double)s->grid.dt, alpha_scale, g_vp, g_rho, g_alpha);
    CHECK_LAUNCH();
    if (lam_src && src && s->src.npoint > 0) {
        const double scale = c.diff_source ? 2.0 * s->grid.dt * s->grid.dt / s->hmax : 2.0 * s->grid.dt / s->hmax;
        // unpadded linear index of each tap cell (the outputs' layout)
        std::vector<long long> lin(s->src.idx.begin(), s->src.idx.end());
        long long *d_lin = nullptr;
        HIP_TRY(hipMalloc((void **)&d_lin, std::max<size_t>(lin.size(), 1) * sizeof(long long)));
        int rc = SA_OK;
        if (h2d(d_lin, lin.data(), lin.size() * sizeof(long long)) != hipSuccess)
            rc = fail(SA_E_HIP, "copy of the tap indices failed");
        if (rc == SA_OK) {
            k_source_grads<<<1, 64, 0, s->stream>>>(s->grid.nt, s->src.npoint, s->src.ntap, lam_src, src, s->src.d_idx,
                                                     d_lin, s->src.d_w, s->vp, scale, g_vp, g_src);
            if (hipGetLastError() != hipSuccess) rc = fail(SA_E_HIP, "launch failed");
        }
        dfree(d_lin);  // waits for the launch
        return rc;
    }
    return SA_OK;
}

extern "C" int sa_get_wavefield(sa_solver *s, int which, float *out) {
    if (!s || !out) return fail(SA_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(s->device));
    int idx = which == 0 ? s->cur : 1 - s->cur;
    return download_unpadded(s, out, s->buf[idx]);
}

extern "C" int sa_free_saved(sa_solver *s) {
    if (!s) return fail(SA_E_ARG, "null solver");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s->saved);
    s->saved = nullptr;
    s->saved_bytes = 0;
    s->saved_mode = 0;
    s->saved_on_host = false;
    dfree(s->segment);
    s->segment = nullptr;
    s->segment_bytes = 0;
    dfree(s->kseg);
    s->kseg = nullptr;
    s->kseg_bytes = 0;
    s->kseg_valid = false;
    return SA_OK;
}

// the whole domain in snapshot layout
static Box whole_box(const Geo &g) {
    Box b;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = 0;
        b.n[a] = g.n[a];
    }
    b.off2 = 0;
    b.p2 = (b.n[2] + 31) / 32 * 32;
    return b;
}

// the part of the box at least margin[a] cells from the domain edges on each axis (margin null:
// the whole box); margins follow the caller's axis order (dim of them), not the padded one
static Box region_of(const sa_solver *s, const Box &box, const int *margin) {
    Box r = box;
    if (margin)
        for (int a = 0; a < 3; ++a) {
            const int src = a - (3 - s->grid.dim);  // 2D: axis 0 is the unit axis
            const int m = src >= 0 ? margin[src] : 0;
            int lo = std::max(box.lo[a], m);
            int hi = std::min(box.lo[a] + box.n[a], s->geo.n[a] - m);
            r.lo[a] = lo;
            r.n[a] = std::max(0, hi - lo);
        }
    return r;
}

// copy nsnap snapshots of a store between device (box layout) and host (region, contiguous)
static int store_copy(sa_solver *s, float *store, int nsnap, const Box &box, const Box &r, float *host, bool to_host) {
    const size_t m = (size_t)r.n[0] * r.n[1] * r.n[2];
    const size_t per = box_points(box);
    if (m == 0 || nsnap == 0) return SA_OK;
    TRY(ensure_staging(s, &s->xfer, &s->xfer_n, m));
    const dim3 grid((r.n[2] + kBlk.x - 1) / kBlk.x, (r.n[1] + kBlk.y - 1) / kBlk.y, r.n[0]);
    for (int i = 0; i < nsnap; ++i) {
        float *slot = store + (size_t)i * per;
        if (to_host) {
            k_region<<<grid, kBlk, 0, s->stream>>>(box, r, slot, s->xfer, 1);
            CHECK_LAUNCH();
            HIP_TRY(hipMemcpyAsync(host + (size_t)i * m, s->xfer, m * sizeof(float), hipMemcpyDeviceToHost, s->stream));
        } else {
            HIP_TRY(hipMemcpyAsync(s->xfer, host + (size_t)i * m, m * sizeof(float), hipMemcpyHostToDevice, s->stream));
            k_region<<<grid, kBlk, 0, s->stream>>>(box, r, slot, s->xfer, 0);
            CHECK_LAUNCH();
        }
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SA_OK;
}

static bool stored_snapshots(const sa_solver *s) { return s->saved_mode == 1 || s->saved_mode == 2; }

extern "C" int sa_saved_count(sa_solver *s, const int *margin, long long *n) {
    if (!s || !n) return fail(SA_E_ARG, "null argument");
    if (!stored_snapshots(s) || (!s->saved && !s->saved_on_host))
        return fail(SA_E_STATE, "no stored snapshots (a forward run with save_mode 1 or 2)");
    for (int a = 0; margin && a < s->grid.dim; ++a)
        if (margin[a] < 0) return fail(SA_E_ARG, "margins must be >= 0");
    const Box r = region_of(s, s->box, margin);
    *n = (long long)s->nsave * r.n[0] * r.n[1] * r.n[2];
    return SA_OK;
}

extern "C" int sa_saved_to_host(sa_solver *s, float *host, long long n, const int *margin, int release) {
    long long want = 0;
    TRY(sa_saved_count(s, margin, &want));
    if (!host || n != want) return fail(SA_E_ARG, "host buffer of %lld floats, the snapshots hold %lld", n, want);
    if (!s->saved) return fail(SA_E_STATE, "the snapshots are already in host memory");
    HIP_TRY(hipSetDevice(s->device));
    TRY(store_copy(s, s->saved, s->nsave, s->box, region_of(s, s->box, margin), host, true));
    if (release) {
        dfree(s->saved);
        s->saved = nullptr;
        s->saved_bytes = 0;
        s->saved_on_host = true;
    }
    return SA_OK;
}

extern "C" int sa_saved_from_host(sa_solver *s, const float *host, long long n, const int *margin) {
    long long want = 0;
    TRY(sa_saved_count(s, margin, &want));
    if (!host || n != want) return fail(SA_E_ARG, "host buffer of %lld floats, the snapshots hold %lld", n, want);
    HIP_TRY(hipSetDevice(s->device));
    const size_t bytes = box_points(s->box) * (size_t)s->nsave * sizeof(float);
    if (!s->saved) {
        if (hipMalloc((void **)&s->saved, bytes) != hipSuccess) {
            s->saved = nullptr;
            return fail(SA_E_NOMEM, "cannot allocate %.2f GB for the saved wavefield", bytes / 1e9);
        }
        s->saved_bytes = bytes;
    }
    // what the host copy does not cover reads as zero
    HIP_TRY(hipMemsetAsync(s->saved, 0, bytes, s->stream));
    TRY(store_copy(s, s->saved, s->nsave, s->box, region_of(s, s->box, margin), const_cast<float *>(host), false));
    s->saved_on_host = false;
    return SA_OK;
}

extern "C" int sa_set_dump(sa_solver *s, int forward_factor, int adjoint_factor) {
    if (!s) return fail(SA_E_ARG, "null solver");
    if (forward_factor < 0 || adjoint_factor < 0) return fail(SA_E_ARG, "dump factors must be >= 0");
    s->dump_f[0] = forward_factor;
    s->dump_f[1] = adjoint_factor;
    return SA_OK;
}

// (tb1 - tb0 + f) / f + 1 slots, slot t / f (common/devito.py:552-578)
static int dump_alloc(sa_solver *s, int w) {
    const int f = s->dump_f[w];
    s->dump_n[w] = 0;
    if (f <= 0) return SA_OK;
    const int n = (s->sch.tb1 - s->sch.tb0 + f) / f + 1;
    const size_t bytes = box_points(whole_box(s->geo)) * (size_t)n * sizeof(float);
    if (!s->dump[w] || s->dump_bytes[w] < bytes) {
        dfree(s->dump[w]);
        s->dump[w] = nullptr;
        s->dump_bytes[w] = 0;
        if (hipMalloc((void **)&s->dump[w], bytes) != hipSuccess) {
            s->dump[w] = nullptr;
            return fail(SA_E_NOMEM, "cannot allocate %.2f GB for the wavefield dump", bytes / 1e9);
        }
        s->dump_bytes[w] = bytes;
    }
    HIP_TRY(hipMemsetAsync(s->dump[w], 0, bytes, s->stream));
    s->dump_n[w] = n;
    return SA_OK;
}

// level `u` of time t into the dump of pass w when t is a dump step
static int dump_level(sa_solver *s, int w, int t, const float *u) {
    const int f = s->dump_f[w];
    if (f <= 0 || !s->dump[w] || t % f != 0 || t < s->sch.tb0 || t > s->sch.tb1 || t / f >= s->dump_n[w])
        return SA_OK;
    const Box b = whole_box(s->geo);
    k_pack<<<domain_grid(s->geo, kBlk), kBlk, 0, s->stream>>>(s->geo, b, u, s->dump[w] + (size_t)(t / f) * box_points(b));
    CHECK_LAUNCH();
    return SA_OK;
}

extern "C" int sa_dump_count(sa_solver *s, int which, long long *n) {
    if (!s || !n || which < 0 || which > 1) return fail(SA_E_ARG, "bad argument");
    *n = (long long)s->dump_n[which] * s->geo.n[0] * s->geo.n[1] * s->geo.n[2];
    return SA_OK;
}

extern "C" int sa_dump_to_host(sa_solver *s, int which, float *host, long long n) {
    long long want = 0;
    TRY(sa_dump_count(s, which, &want));
    if (want == 0) return fail(SA_E_STATE, "no %s wavefield dump: sa_set_dump before the run", which ? "adjoint" : "forward");
    if (!host || n != want) return fail(SA_E_ARG, "host buffer of %lld floats, the dump holds %lld", n, want);
    HIP_TRY(hipSetDevice(s->device));
    const Box b = whole_box(s->geo);
    return store_copy(s, s->dump[which], s->dump_n[which], b, b, host, true);
}

extern "C" int sa_set_tile(sa_solver *s, int rows, int chunk, int version) {
    if (!s) return fail(SA_E_ARG, "null solver");
    s->fast.prepared = true;  // an explicit tile overrides the default
    if (fast3d_set_tune(s->fast, rows, chunk, version) != 0)
        return fail(SA_E_ARG, "tile: chunk >= 1 and version 7 with rows 2 or version 8 with rows 1/2");
    return SA_OK;
}

extern "C" int sa_set_tile_density(sa_solver *s, int rows, int chunk, int waves) {
    if (!s) return fail(SA_E_ARG, "null solver");
    if (rows == 0) {
        s->rho3d.enabled = false;
        return SA_OK;
    }
    if (!rho3d_valid_tune(rows, chunk, waves))
        return fail(SA_E_ARG, "density tile: chunk >= 1, rows 1 or 2 with 4 or 8 waves, rows 1 with 16 waves");
    s->rho3d.rows = rows;
    s->rho3d.chunk = chunk;
    s->rho3d.waves = waves;
    s->rho3d.enabled = true;
    return SA_OK;
}

extern "C" int sa_batch_set_tile_density(sa_solver *s, int rows, int chunk, int waves) {
    if (!s) return fail(SA_E_ARG, "null solver");
    if (rows != 1 || !rho3d_valid_tune(rows, chunk, waves) || waves == 16)
        return fail(SA_E_ARG, "two-shot density tile: rows 1, 4 or 8 waves, chunk >= 1");
    s->rho3d_b2.rows = rows;
    s->rho3d_b2.chunk = chunk;
    s->rho3d_b2.waves = waves;
    return SA_OK;
}

extern "C" int sa_timing_enable(sa_solver *s, int enable) {
    if (!s) return fail(SA_E_ARG, "null solver");
    TRY(drain_timing(s));
    s->timing = enable != 0;
    s->timing_stride = enable > 1 ? enable : 1;
    s->timing_seq = 0;
    s->timing_this = false;
    s->kernel_ms = 0;
    s->launches = 0;
    return SA_OK;
}

extern "C" int sa_timing_read(sa_solver *s, double *kernel_ms, int64_t *launches) {
    if (!s) return fail(SA_E_ARG, "null solver");
    TRY(drain_timing(s));
    if (kernel_ms) *kernel_ms = s->kernel_ms;
    if (launches) *launches = s->launches;
    return SA_OK;
}

extern "C" int sa_device_alloc(sa_solver *s, int64_t bytes, void **out) {
    if (!s || !out || bytes < 0) return fail(SA_E_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(s->device));
    if (hipMalloc(out, std::max<int64_t>(bytes, 1)) != hipSuccess) return fail(SA_E_NOMEM, "hipMalloc failed");
    return SA_OK;
}

extern "C" int sa_device_free(sa_solver *s, void *ptr) {
    if (!s) return fail(SA_E_ARG, "null solver");
    HIP_TRY(hipSetDevice(s->device));
    dfree(ptr);
    return SA_OK;
}

extern "C" int sa_memcpy(sa_solver *s, void *dst, const void *src, int64_t bytes, int kind) {
    if (!s) return fail(SA_E_ARG, "null solver");
    HIP_TRY(hipSetDevice(s->device));
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, k, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SA_OK;
}

extern "C" int sa_stream_create(int device, void **stream) {
    if (!stream) return fail(SA_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(device));
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));  // default priority
    *stream = (void *)st;
    return SA_OK;
}

extern "C" int sa_stream_destroy(void *stream) {
    if (!stream) return SA_OK;
    HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return SA_OK;
}

extern "C" int sa_synchronize(sa_solver *s) {
    if (!s) return fail(SA_E_ARG, "null solver");
    HIP_TRY(hipStreamSynchronize(s->stream));
    return drain_timing(s);
}

// ---------------------------------------------------------------------------------------
// Two shots per pass. The shots share the grid, scheme, medium and boundary; each has its own
// source / receiver taps, source samples, traces and residual. Their wavefields are stored
// interleaved per cell so that one pass of the specialised 3D update advances both with one
// read of vp, and the adjoint accumulates the sum of the two shots' vp gradients and
// preconditioners (the quantity an FWI iteration uses) with one read-modify-write of each.

// Two-shot runs: the specialised update (no density; save mode 0 or 1, the interior p.dt2) or
// the fused density update (save mode 0 or 2, the raw wavefield of the whole domain).
static bool batch_density(const sa_solver *s) { return s->sch.rho_mode != 0; }

static bool batch_ok(const sa_solver *s) {
    const sa_scheme &c = s->sch;
    if (s->nshot != 2) return false;
    if (batch_density(s)) return use_rho3d_pairs(s) && (c.save_mode == 0 || c.save_mode == 2);
    return use_fast_pairs(s) && (c.save_mode == 0 || c.save_mode == 1);
}

extern "C" int sa_batch_create(sa_solver *s, int nshot) {
    if (!s) return fail(SA_E_ARG, "null solver");
    if (nshot != 2) return fail(SA_E_ARG, "batches hold 2 shots");
    if (s->total * 8 >= (1ll << 31)) return fail(SA_E_UNSUPPORTED, "grid too large for the two-shot kernel");
    HIP_TRY(hipSetDevice(s->device));
    for (int k = 0; k < 2; ++k)
        if (!s->bbuf[k]) TRY(dalloc(&s->bbuf[k], 2 * s->total));
    s->nshot = 2;
    return SA_OK;
}

extern "C" int sa_batch_set_taps(sa_solver *s, int shot, int nsrc, int src_ntap, const int64_t *src_index,
                                 const float *src_weight, int nrec, int rec_ntap, const int64_t *rec_index,
                                 const float *rec_weight) {
    if (!s) return fail(SA_E_ARG, "null solver");
    if (s->nshot != 2 || shot < 0 || shot > 1) return fail(SA_E_ARG, "shot must be 0 or 1 of a batch");
    TRY(set_sparse(s, s->bsrc[shot], nsrc, src_ntap, src_index, src_weight));
    return set_sparse(s, s->brec[shot], nrec, rec_ntap, rec_index, rec_weight);
}

extern "C" int sa_batch_set_tile(sa_solver *s, int rows, int chunk, int waves) {
    if (!s) return fail(SA_E_ARG, "null solver");
    if (fast3d_set_tune_b2(s->fast, rows, chunk, waves) != 0)
        return fail(SA_E_ARG, "two-shot tile: chunk >= 1, waves 8 with rows 1 or 2, or waves 4 with rows 2");
    return SA_OK;
}

// Snapshot t / f of a batch starts at bsaved + (t / f) * 2 * box_points(bbox). Save mode 1: the
// interior p.dt2 of the two shots interleaved per cell; save mode 2: the raw wavefield of the
// whole domain, shot 0's snapshot followed by shot 1's.
static int batch_alloc_saved(sa_solver *s) {
    const sa_scheme &c = s->sch;
    const Geo &g = s->geo;
    const int f = c.undersampling;
    Box b;
    for (int a = 0; a < 3; ++a) {
        const bool interior = c.save_mode == 1;
        b.lo[a] = interior ? g.nabs[a] : 0;
        b.n[a] = interior ? g.n[a] - 2 * g.nabs[a] : g.n[a];
    }
    b.off2 = b.lo[2] % 32;
    b.p2 = (b.off2 + b.n[2] + 31) / 32 * 32;
    const int nsave = c.tb1 / f + 1;
    const size_t per = 2 * box_points(b);  // two shots per cell
    const size_t bytes = per * nsave * sizeof(float);
    if (!s->bsaved || s->bsaved_bytes < bytes) {
        dfree(s->bsaved);
        s->bsaved = nullptr;
        s->bsaved_bytes = 0;
        if (hipMalloc((void **)&s->bsaved, bytes) != hipSuccess) {
            s->bsaved = nullptr;
            return fail(SA_E_NOMEM, "cannot allocate %.2f GB for the saved wavefields", bytes / 1e9);
        }
        s->bsaved_bytes = bytes;
    }
    // slots the forward run never writes read as zero (see alloc_saved)
    const int first = (std::max(1, c.tb0) + f - 1) / f;
    const int last = std::min((c.tb1 - 2) / f, nsave - 1);
    if (first > last) {
        HIP_TRY(hipMemsetAsync(s->bsaved, 0, bytes, s->stream));
    } else {
        if (first > 0) HIP_TRY(hipMemsetAsync(s->bsaved, 0, per * first * sizeof(float), s->stream));
        if (last + 1 < nsave)
            HIP_TRY(hipMemsetAsync(s->bsaved + per * (size_t)(last + 1), 0, per * (nsave - 1 - last) * sizeof(float),
                                   s->stream));
    }
    s->bbox = b;
    s->bnsave = nsave;
    s->bf = f;
    return SA_OK;
}

// src[k]: device (nt, nsrc_k) samples of shot k; rec[k]: device (nt, nrec_k) traces of shot k
extern "C" int sa_batch_forward_dev(sa_solver *s, const float *const *src, float *const *rec) {
    if (!s || !src || !rec) return fail(SA_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(s->device));
    TRY(prepare(s));
    if (!batch_ok(s))
        return fail(SA_E_UNSUPPORTED, "two-shot runs need sa_batch_create and a 3D OT2 scheme with sponge or no "
                                      "boundary: without density, no attenuation and save mode 0 or 1; with "
                                      "density, no power-2 attenuation and save mode 0 or 2");
    const sa_scheme &c = s->sch;
    const bool dens = batch_density(s);
    for (int k = 0; k < 2; ++k) {
        TRY(coef_sparse(s, s->bsrc[k], 0));
        TRY(zero_field2(s, s->bbuf[k]));
        if (s->brec[k].npoint > 0)
            HIP_TRY(hipMemsetAsync(rec[k], 0, sizeof(float) * (size_t)s->grid.nt * s->brec[k].npoint, s->stream));
    }
    if (c.save_mode != 0) TRY(batch_alloc_saved(s));
    const int f = c.undersampling;
    const size_t per = 2 * box_points(s->bbox);
    const float idt2 = 1.f / (s->dt * s->dt);
    const int tlast = c.tb1 - 2;
    for (int t = 1; t <= tlast; ++t) {
        float *u = s->bbuf[t % 2], *o = s->bbuf[(t + 1) % 2];
        float *save = nullptr;
        if (c.save_mode != 0 && t % f == 0 && t >= c.tb0 && t <= c.tb1 && t / f < s->bnsave)
            save = s->bsaved + (size_t)(t / f) * per;
        hipEvent_t e0;
        TRY(timed_begin(s, &e0));
        if (dens) {
            Rho3DStep st;
            st.u = u;
            st.other = o;
            st.out = o;
            st.b = c.rho_mode == 2 ? s->b : nullptr;
            st.A = s->A;
            st.B = s->B;
            st.G = s->G;
            st.save = save;
            st.save1 = save ? save + box_points(s->bbox) : nullptr;
            st.save_mode = 2;
            st.box = s->bbox;
            st.idt2 = idt2;
            st.nshot = 2;
            if (rho3d_step(s->rho3d_b2, s->geo, st, s->stream) != 0)
                return fail(SA_E_HIP, "two-shot density update launch: %s", hipGetErrorString(hipGetLastError()));
        } else {
            TRY(fast3d_step_b2(s->fast, s->geo, u, o, o, save, s->bbox, idt2, s->stream));
        }
        TRY(timed_end(s));
        const float *const sv[2] = {src[0] + (size_t)t * s->bsrc[0].npoint, src[1] + (size_t)t * s->bsrc[1].npoint};
        float *const rv[2] = {rec[0] + (size_t)t * s->brec[0].npoint, rec[1] + (size_t)t * s->brec[1].npoint};
        // the p.dt2 save takes the injected term too; the raw save holds u, before this injection
        TRY(inject2(s, s->bsrc, sv, 1.f, o, dens ? nullptr : save, idt2, s->bbox));
        TRY(gather2(s, s->brec, o, 1.f, rv));
    }
    s->cur = (tlast + 1) % 2;
    return SA_OK;
}

// res[k]: device (nt, nrec_k) adjoint source of shot k (windowed as for sa_adjoint_dev); the
// vp gradient and preconditioner of both shots are summed (sa_read_gradient 0 / 1)
extern "C" int sa_batch_adjoint_dev(sa_solver *s, const float *const *res) {
    if (!s || !res) return fail(SA_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(s->device));
    TRY(prepare(s));
    if (!batch_ok(s)) return fail(SA_E_UNSUPPORTED, "two-shot runs need sa_batch_create and a two-shot scheme");
    const sa_scheme &c = s->sch;
    const bool dens = batch_density(s);
    if (!s->bsaved || (c.save_mode != 1 && c.save_mode != 2))
        return fail(SA_E_STATE, "the two-shot adjoint needs a two-shot forward run with save mode 1 (no density) "
                                "or 2 (density)");
    if (dens && !(c.rho_mode == 2 && s->b && rho3d_grad_supported(s->geo, s->bbox)))
        return fail(SA_E_UNSUPPORTED, "the two-shot density adjoint needs variable density and absorbing layers "
                                      "of at least 8 cells");
    // vp gradient and preconditioner (both schemes), rho gradient and preconditioner (density)
    TRY(grad_outputs(s, dens ? 4 : 2, s->gacc));
    for (int k = 0; k < 2; ++k) {
        TRY(coef_sparse(s, s->brec[k], 1));
        TRY(zero_field2(s, s->bbuf[k]));
    }
    const int f = s->bf;
    const size_t per = 2 * box_points(s->bbox), one = box_points(s->bbox);
    const float fdt2 = (float)(1.0 / ((double)f * s->grid.dt * (double)f * s->grid.dt));
    const int tstart = c.tb1 - 1 - f, tend = c.tb0 + 1;
    for (int t = tstart; t >= tend; --t) {
        float *u = s->bbuf[t % 2], *o = s->bbuf[(t + 1) % 2];
        const bool guard = t % f == 0 && t >= c.tb0 && t <= c.tb1 && t / f < s->bnsave;
        const float *ps = guard ? s->bsaved + (size_t)(t / f) * per : nullptr;
        hipEvent_t e0;
        TRY(timed_begin(s, &e0));
        if (dens) {
            Rho3DStep st;
            st.u = u;
            st.other = o;
            st.out = o;
            st.b = s->b;
            st.A = s->A;
            st.B = s->B;
            st.G = s->G;
            st.save = nullptr;
            st.save_mode = 0;
            st.box = s->bbox;
            st.idt2 = 0.f;
            st.nshot = 2;
            if (rho3d_step(s->rho3d_b2, s->geo, st, s->stream) != 0)
                return fail(SA_E_HIP, "two-shot density update launch: %s", hipGetErrorString(hipGetLastError()));
        } else {
            TRY(fast3d_step_adjoint_b2(s->fast, s->geo, u, o, o, ps, s->bbox, s->gacc[0], s->gacc[1], s->stream));
        }
        TRY(timed_end(s));
        const float *const rv[2] = {res[0] + (size_t)t * s->brec[0].npoint, res[1] + (size_t)t * s->brec[1].npoint};
        TRY(inject2(s, s->brec, rv, -1.f, o, nullptr, 0.f, s->bbox));
        if (dens && guard) {
            // each shot's vp and rho terms from its raw snapshots t - f, t, t + f and its own
            // adjoint level (shot k of u), summed into the shared accumulators
            const int i = t / f;
            for (int k = 0; k < 2; ++k) {
                const float *psk = ps + k * one;
                const float *psm = i >= 1 ? s->bsaved + (size_t)(i - 1) * per + k * one : nullptr;
                const float *psp = i + 1 < s->bnsave ? s->bsaved + (size_t)(i + 1) * per + k * one : nullptr;
                if (rho3d_grad(s->geo, s->bbox, psk, psm, psp, fdt2, s->b, u, s->gacc[0], s->gacc[1], s->gacc[2],
                               s->gacc[3], s->stream, 2, k) != 0)
                    return fail(SA_E_HIP, "density gradient launch: %s", hipGetErrorString(hipGetLastError()));
            }
        }
    }
    s->cur = (tend - 1 + 2) % 2;
    return SA_OK;
}

// wavefield `which` (0: last level, 1: the one before) of shot `shot` of a batch, unpadded
extern "C" int sa_batch_get_wavefield(sa_solver *s, int shot, int which, float *out) {
    if (!s || !out || shot < 0 || shot > 1 || s->nshot != 2) return fail(SA_E_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(s->device));
    const Geo &g = s->geo;
    std::vector<float> tmp(2 * (size_t)s->total);
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(tmp.data(), s->bbuf[which == 0 ? s->cur : 1 - s->cur], tmp.size() * sizeof(float),
                      hipMemcpyDeviceToHost));
    for (int i = 0; i < g.n[0]; ++i)
        for (int j = 0; j < g.n[1]; ++j)
            for (int k = 0; k < g.n[2]; ++k)
                out[((long long)i * g.n[1] + j) * g.n[2] + k] =
                    tmp[2 * (g.base + (long long)i * g.s0 + (long long)j * g.s1 + k) + shot];
    return SA_OK;
}

// ---- several 2D shots per launch (lockstep) -------------------------------------------------
// Shots on one 2D grid and scheme, each held by its own solver (medium, taps, levels, saved
// wavefield, accumulators), advance together: every time step is ONE launch of the fused 2D
// update over the tiles of all shots (grid axis z = shot). Each shot's arguments for each step
// are those its one-shot run passes (flat2d_args), laid out in a table that goes to the device
// once per run, so every shot's traces and gradients equal its one-shot run's. All work is issued
// on the first solver's stream, after the work already queued on each solver's stream; each
// solver's stream then waits for the group's work.

namespace {
// issues the group's work on one stream and restores every solver's own stream on exit
struct GroupStream {
    sa_solver *const *v;
    int n;
    std::vector<hipStream_t> own;
    std::vector<hipEvent_t> evs;
    hipStream_t st;
    GroupStream(sa_solver *const *v_, int n_) : v(v_), n(n_), own(n_), st(v_[0]->stream) {
        for (int k = 0; k < n; ++k) own[k] = v[k]->stream;
    }
    int begin() {
        for (int k = 0; k < n; ++k) {
            if (own[k] != st) {
                hipEvent_t e;
                HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                evs.push_back(e);
                HIP_TRY(hipEventRecord(e, own[k]));
                HIP_TRY(hipStreamWaitEvent(st, e, 0));
            }
            v[k]->stream = st;
        }
        return SA_OK;
    }
    ~GroupStream() {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess && hipEventRecord(e, st) == hipSuccess)
            for (int k = 0; k < n; ++k)
                if (own[k] != st) (void)hipStreamWaitEvent(own[k], e, 0);
        if (e) (void)hipEventDestroy(e);
        for (int k = 0; k < n; ++k) v[k]->stream = own[k];
        for (auto x : evs) (void)hipEventDestroy(x);
    }
};
}  // namespace

static int shots2d_check(sa_solver *const *v, int n) {
    if (!v || n < 1 || n > 1024) return fail(SA_E_ARG, "1 .. 1024 solvers");
    for (int k = 0; k < n; ++k)
        if (!v[k]) return fail(SA_E_ARG, "null solver");
    HIP_TRY(hipSetDevice(v[0]->device));
    const sa_scheme &c0 = v[0]->sch;
    for (int k = 0; k < n; ++k) {
        sa_solver *s = v[k];
        if (s->device != v[0]->device) return fail(SA_E_ARG, "the solvers of a group share one device");
        TRY(prepare(s));
        if (!use_flat2d(s)) return fail(SA_E_UNSUPPORTED, "lockstep shots run the fused 2D update (2D, OT2)");
        const sa_scheme &c = s->sch;
        if (c.kernel != c0.kernel || c.rho_mode != c0.rho_mode || c.att_power != c0.att_power ||
            c.boundary != c0.boundary || c.diff_source != c0.diff_source || c.undersampling != c0.undersampling ||
            c.save_mode != c0.save_mode || c.tb0 != c0.tb0 || c.tb1 != c0.tb1 || s->grid.nt != v[0]->grid.nt)
            return fail(SA_E_ARG, "the solvers of a group need the same scheme and number of steps");
        for (int a = 0; a < 3; ++a)
            if (s->geo.n[a] != v[0]->geo.n[a] || s->geo.nabs[a] != v[0]->geo.nabs[a] || s->P[a] != v[0]->P[a])
                return fail(SA_E_ARG, "the solvers of a group need the same grid");
        if (c.save_mode > 2) return fail(SA_E_UNSUPPORTED, "lockstep shots: save mode 0, 1 or 2");
        if (s->dump_f[0] > 0 || s->dump_f[1] > 0)
            return fail(SA_E_UNSUPPORTED, "wavefield dumps: run the shots one by one");
    }
    return SA_OK;
}

static int shots2d_table(sa_solver *s0, size_t n) {
    if (s0->tab_ev) HIP_TRY(hipEventSynchronize(s0->tab_ev));  // the previous upload has read the table
    if (n <= s0->tab_cap) return SA_OK;
    if (s0->tab_h) HIP_TRY(hipHostFree(s0->tab_h));
    if (s0->tab_d) HIP_TRY(hipFree(s0->tab_d));
    s0->tab_h = nullptr;
    s0->tab_d = nullptr;
    s0->tab_cap = 0;
    HIP_TRY(hipHostMalloc((void **)&s0->tab_h, n * sizeof(Flat2DStep), hipHostMallocDefault));
    if (hipMalloc((void **)&s0->tab_d, n * sizeof(Flat2DStep)) != hipSuccess) {
        s0->tab_d = nullptr;
        return fail(SA_E_NOMEM, "cannot allocate the lockstep argument table");
    }
    s0->tab_cap = n;
    return SA_OK;
}

// launches the steps of the table (nstep rows of n shots) on the group's stream
static int shots2d_run(sa_solver *const *v, int n, int nstep) {
    sa_solver *s0 = v[0];
    const size_t bytes = (size_t)nstep * n * sizeof(Flat2DStep);
    if (nstep <= 0) return SA_OK;
    HIP_TRY(hipMemcpyAsync(s0->tab_d, s0->tab_h, bytes, hipMemcpyHostToDevice, s0->stream));
    if (!s0->tab_ev) HIP_TRY(hipEventCreateWithFlags(&s0->tab_ev, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(s0->tab_ev, s0->stream));
    for (int i = 0; i < nstep; ++i) {
        hipEvent_t e0;
        TRY(timed_begin(s0, &e0));
        if (flat2d_step_shots(s0->geo, s0->sch.rho_mode, s0->tab_h[(size_t)i * n], s0->tab_d + (size_t)i * n, n,
                              s0->stream) != 0)
            return fail(SA_E_HIP, "lockstep 2D update launch: %s", hipGetErrorString(hipGetLastError()));
        TRY(timed_end(s0));
    }
    return SA_OK;
}

// src[k], rec[k]: device (nt, nsrc_k) samples and (nt, nrec_k) traces of shot k (solver v[k])
extern "C" int sa_shots2d_forward_dev(sa_solver *const *v, int n, const float *const *src, float *const *rec) {
    if (!src || !rec) return fail(SA_E_ARG, "null argument");
    TRY(shots2d_check(v, n));
    GroupStream gs(v, n);
    TRY(gs.begin());
    const sa_scheme &c = v[0]->sch;
    const int f = c.undersampling, tlast = c.tb1 - 2, nstep = std::max(0, tlast);
    TRY(shots2d_table(v[0], (size_t)std::max(1, nstep) * n));
    std::vector<float *> pend(n, nullptr);
    for (int k = 0; k < n; ++k) {
        sa_solver *s = v[k];
        const Geo &g = s->geo;
        s->lap_of = nullptr;
        TRY(coef_sparse(s, s->src, 0));
        TRY(zero_field(s, s->buf[0]));
        TRY(zero_field(s, s->buf[1]));
        if (c.boundary == 1)
            for (int a = 0; a < 3; ++a)
                if (g.act[a]) {
                    TRY(zero_field(s, s->u1[a]));
                    TRY(zero_field(s, s->u2[a]));
                    TRY(zero_field(s, s->u3[a]));
                }
        if (s->lapo) TRY(zero_field(s, s->lapo));
        if (s->rec.npoint > 0)
            HIP_TRY(hipMemsetAsync(rec[k], 0, sizeof(float) * (size_t)s->grid.nt * s->rec.npoint, s->stream));
        if (c.save_mode != 0) TRY(alloc_saved(s));
        if (c.att_power == 2 && nstep > 0) {
            // the first step's other level (buf[0]): its Laplacian
            if (!s->lapother || !s->lapnext) return fail(SA_E_STATE, "power-2 attenuation buffers missing");
            k_laplace<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->buf[0], nullptr, s->lapother);
            CHECK_LAUNCH();
        }
        const size_t per = box_points(s->box);
        float *pending = nullptr;  // the previous step's receiver row, interpolated by this step's launch
        for (int t = 1; t <= tlast; ++t) {
            float *u = s->buf[t % 2], *o = s->buf[(t + 1) % 2];
            float *save = nullptr;
            int mode = 0;
            if (c.save_mode != 0 && t % f == 0 && t >= c.tb0 && t <= c.tb1 && t / f < s->nsave) {
                save = s->saved + (size_t)(t / f) * per;
                mode = c.save_mode == 1 ? 1 : 2;
            }
            const InjectSpec spec{&s->src, src[k] + (size_t)t * s->src.npoint, 1.f, mode == 1};
            const GatherSpec gspec{&s->rec, 1.f, pending};
            Flat2DStep &x = v[0]->tab_h[(size_t)(t - 1) * n + k];
            TRY(flat2d_args(s, u, o, o, 0, save, mode, &spec, pending ? &gspec : nullptr, nullptr, x));
            flat2d_lap_advance(s, x, u);
            pending = s->rec.npoint > 0 ? rec[k] + (size_t)t * s->rec.npoint : nullptr;
        }
        pend[k] = pending;
        s->cur = (tlast + 1) % 2;
    }
    TRY(shots2d_run(v, n, nstep));
    for (int k = 0; k < n; ++k)
        if (pend[k]) TRY(gather(v[k], v[k]->rec, v[k]->buf[(tlast + 1) % 2], 1.f, pend[k]));
    return SA_OK;
}

// res[k]: device (nt, nrec_k) adjoint source of shot k; want: SA_WANT_VP / SA_WANT_RHO. Each
// solver's gradients are its own shot's (sa_read_gradient / sa_read_gradient_dev on it).
extern "C" int sa_shots2d_adjoint_dev(sa_solver *const *v, int n, const float *const *res, int want) {
    if (!res) return fail(SA_E_ARG, "null argument");
    TRY(shots2d_check(v, n));
    const sa_scheme &c = v[0]->sch;
    const bool want_v = want & SA_WANT_VP, want_r = want & SA_WANT_RHO;
    for (int k = 0; k < n; ++k) {
        sa_solver *s = v[k];
        if (s->saved_on_host) return fail(SA_E_STATE, "a saved wavefield is in host memory: sa_saved_from_host first");
        if (!s->saved || (s->saved_mode != 1 && s->saved_mode != 2))
            return fail(SA_E_STATE, "the lockstep adjoint needs forward runs that saved the wavefield (save mode 1 or 2)");
        const bool raw = s->saved_mode == 2;
        if (want_r && !(raw && c.rho_mode == 2 && s->geo.nabs[1] >= 5 && s->geo.nabs[2] >= 5))
            return fail(SA_E_UNSUPPORTED, "the lockstep rho gradient needs variable density, the raw wavefield and "
                                          "absorbing layers of at least 5 cells");
    }
    GroupStream gs(v, n);
    TRY(gs.begin());
    const int tstart = c.tb1 - 1 - v[0]->saved_f, tend = c.tb0 + 1;
    const int nstep = std::max(0, tstart - tend + 1);
    TRY(shots2d_table(v[0], (size_t)std::max(1, nstep) * n));
    for (int k = 0; k < n; ++k) {
        sa_solver *s = v[k];
        const Geo &g = s->geo;
        if (want_v) TRY(grad_outputs(s, 2, s->gacc));
        if (want_r) TRY(grad_outputs(s, 2, s->gacc + 2));
        float *gv = want_v ? s->gacc[0] : nullptr, *pv = want_v ? s->gacc[1] : nullptr;
        float *gr = want_r ? s->gacc[2] : nullptr, *pr = want_r ? s->gacc[3] : nullptr;
        TRY(coef_sparse(s, s->rec, 1));
        TRY(zero_aux(s));
        if (c.att_power == 2 && nstep > 0) {
            if (!s->lapother || !s->lapnext) return fail(SA_E_STATE, "power-2 attenuation buffers missing");
            k_laplace<<<domain_grid(g, kBlk), kBlk, 0, s->stream>>>(g, s->buf[(tstart + 1) % 2], nullptr,
                                                                     s->lapother);
            CHECK_LAUNCH();
        }
        const int f = s->saved_f;
        const bool raw = s->saved_mode == 2;
        const size_t per = box_points(s->box);
        const float fdt2 = (float)(1.0 / ((double)f * s->grid.dt * (double)f * s->grid.dt));
        for (int t = tstart; t >= tend; --t) {
            float *u = s->buf[t % 2], *o = s->buf[(t + 1) % 2];
            const bool guard = t % f == 0 && t >= c.tb0 && t <= c.tb1 && t / f < s->nsave && (gv || gr);
            GradSpec gspec{};
            if (guard) {
                gspec.ps = s->saved + (size_t)(t / f) * per;
                gspec.raw = raw;
                if (raw) {
                    gspec.psm = t / f >= 1 ? s->saved + (size_t)(t / f - 1) * per : nullptr;
                    gspec.psp = t / f + 1 < s->nsave ? s->saved + (size_t)(t / f + 1) * per : nullptr;
                }
                gspec.fdt2 = fdt2;
                gspec.gv = gv;
                gspec.pv = pv;
                gspec.gr = gr;
                gspec.pr = pr;
            }
            const InjectSpec spec{&s->rec, res[k] + (size_t)t * s->rec.npoint, -1.f, false};
            Flat2DStep &x = v[0]->tab_h[(size_t)(tstart - t) * n + k];
            TRY(flat2d_args(s, u, o, o, 1, nullptr, 0, &spec, nullptr, guard ? &gspec : nullptr, x));
            flat2d_lap_advance(s, x, u);
        }
        s->cur = (tend - 1 + 2) % 2;
    }
    return shots2d_run(v, n, nstep);
}

// ---- many shots of uneven duration and receiver count (forward) -----------------------------

extern "C" int sa_multi_create(sa_solver *s, int nshot) {
    if (!s) return fail(SA_E_ARG, "null solver");
    if (nshot < 1 || nshot > 4096) return fail(SA_E_ARG, "nshot must be between 1 and 4096, got %d", nshot);
    if (s->total * 8 >= (1ll << 31)) return fail(SA_E_UNSUPPORTED, "grid too large for the shot-pair kernel");
    TRY(quiesce(s));
    release_graph(s);
    for (auto &m : s->mshot) {
        free_sparse(m.src);
        free_sparse(m.rec);
    }
    for (float *p : s->mbuf) dfree(p);
    s->mbuf.clear();
    s->mshot.assign(nshot, sa_solver::MShot());
    const int npair = (nshot + 1) / 2;
    s->mbuf.assign(2 * npair, nullptr);
    for (auto &p : s->mbuf) TRY(dalloc(&p, 2 * s->total));
    dfree(s->d_pairs);
    dfree(s->d_minj);
    dfree(s->d_mgat);
    s->d_pairs = nullptr;
    s->d_minj = nullptr;
    s->d_mgat = nullptr;
    HIP_TRY(hipMalloc((void **)&s->d_pairs, npair * sizeof(PairLevels)));
    HIP_TRY(hipMalloc((void **)&s->d_minj, nshot * sizeof(MInject)));
    HIP_TRY(hipMalloc((void **)&s->d_mgat, nshot * sizeof(MGather)));
    if (!s->mstream) HIP_TRY(hipStreamCreateWithFlags(&s->mstream, hipStreamNonBlocking));
    return SA_OK;
}

extern "C" int sa_multi_set_shot(sa_solver *s, int shot, int nt, int nsrc, int src_ntap, const int64_t *src_index,
                                 const float *src_weight, int nrec, int rec_ntap, const int64_t *rec_index,
                                 const float *rec_weight) {
    if (!s) return fail(SA_E_ARG, "null solver");
    if (shot < 0 || shot >= (int)s->mshot.size()) return fail(SA_E_ARG, "shot %d outside the schedule", shot);
    if (nt < 3 || nt > s->grid.nt) return fail(SA_E_ARG, "shot nt %d must be within 3 .. %d", nt, s->grid.nt);
    auto &m = s->mshot[shot];
    TRY(set_sparse(s, m.src, nsrc, src_ntap, src_index, src_weight));
    TRY(set_sparse(s, m.rec, nrec, rec_ntap, rec_index, rec_weight));
    m.nt = nt;
    m.set = true;
    return SA_OK;
}

// src[k]: device (nt_k, nsrc_k) samples of shot k; rec[k]: device (nt_k, nrec_k) traces of shot k.
// Every shot runs its own steps 1 .. nt_k - 2 (run_forward with time bounds (0, nt_k)). With
// graph != 0 the whole schedule is captured into one HIP graph and launched once.
extern "C" int sa_multi_forward_dev(sa_solver *s, const float *const *src, float *const *rec, int graph) {
    if (!s || !src || !rec) return fail(SA_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(s->device));
    const int nshot = (int)s->mshot.size();
    if (nshot == 0) return fail(SA_E_STATE, "sa_multi_create must be called first");
    for (int k = 0; k < nshot; ++k)
        if (!s->mshot[k].set) return fail(SA_E_STATE, "shot %d of the schedule has no taps", k);
    TRY(prepare(s));
    if (!use_fast_pairs(s) || s->sch.save_mode != 0)
        return fail(SA_E_UNSUPPORTED, "the multi-shot schedule runs the specialised 3D forward (OT2, no density, "
                                      "sponge or no boundary, no attenuation, save mode 0)");
    TRY(quiesce(s));  // the device tables below may still be read by a previous schedule
    release_graph(s);
    // longest shot first: the shots (and pairs) still running at any step are a prefix
    std::vector<int> order(nshot);
    for (int k = 0; k < nshot; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return s->mshot[a].nt > s->mshot[b].nt; });
    s->morder = order;
    const int npair = (nshot + 1) / 2;
    std::vector<PairLevels> pl(npair);
    for (int j = 0; j < npair; ++j) pl[j] = PairLevels{s->mbuf[2 * j], s->mbuf[2 * j + 1]};
    std::vector<MInject> inj(nshot);
    std::vector<MGather> gat(nshot);
    int max_dest = 0, max_rec = 0;
    for (int i = 0; i < nshot; ++i) {
        auto &m = s->mshot[order[i]];
        TRY(coef_sparse(s, m.src, 0));
        inj[i] = MInject{m.src.ndest, m.src.npoint, m.src.d_dest, m.src.d_off, m.src.d_point, m.src.d_coef,
                         src[order[i]], i / 2, i % 2};
        gat[i] = MGather{m.rec.npoint, m.rec.ntap, m.rec.d_idx, m.rec.d_w, rec[order[i]], i / 2, i % 2};
        max_dest = std::max(max_dest, m.src.ndest);
        max_rec = std::max(max_rec, m.rec.npoint);
        if (m.rec.npoint > 0)
            HIP_TRY(hipMemsetAsync(rec[order[i]], 0, sizeof(float) * (size_t)m.nt * m.rec.npoint, s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));  // coefficients computed, traces cleared
    HIP_TRY(h2d(s->d_pairs, pl.data(), pl.size() * sizeof(PairLevels)));
    HIP_TRY(h2d(s->d_minj, inj.data(), inj.size() * sizeof(MInject)));
    HIP_TRY(h2d(s->d_mgat, gat.data(), gat.size() * sizeof(MGather)));
    // clear the levels on the solver's stream, outside any graph
    for (float *p : s->mbuf) HIP_TRY(hipMemsetAsync(p, 0, 2 * s->total * sizeof(float), s->stream));
    const int tlast = s->mshot[order[0]].nt - 2;
    hipStream_t st = graph ? s->mstream : s->stream;
    if (graph) {
        HIP_TRY(hipStreamSynchronize(s->stream));
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    }
    auto t0 = std::chrono::steady_clock::now();
    int rc = SA_OK;
    int nact = nshot;
    for (int t = 1; t <= tlast && rc == SA_OK; ++t) {
        while (nact > 0 && s->mshot[order[nact - 1]].nt - 2 < t) --nact;  // shots still running
        const int pact = (nact + 1) / 2;                                    // their pairs
        hipEvent_t e0;
        if (!graph && (rc = timed_begin(s, &e0)) != SA_OK) break;
        if (fast3d_step_multi(s->fast, s->geo, s->d_pairs, pact, t % 2, st) != 0) {
            rc = fail(SA_E_HIP, "multi-shot update launch failed");
            break;
        }
        if (!graph && (rc = timed_end(s)) != SA_OK) break;
        if (max_dest > 0)
            k_inject_multi<<<dim3((max_dest + 127) / 128, nact), 128, 0, st>>>(s->d_minj, s->d_pairs, t, t % 2);
        if (max_rec > 0)
            k_gather_multi<<<dim3((max_rec + 127) / 128, nact), 128, 0, st>>>(s->d_mgat, s->d_pairs, t, t % 2);
        if (hipGetLastError() != hipSuccess) rc = fail(SA_E_HIP, "multi-shot sparse launch failed");
    }
    if (graph) {
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc != SA_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess || !g) return fail(SA_E_HIP, "capture of the multi-shot schedule failed");
        hipGraphExec_t exec = nullptr;
        if (hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(g);
            return fail(SA_E_HIP, "instantiation of the multi-shot graph failed");
        }
        s->graph_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        s->mgraph = g;
        s->mexec = exec;
        if (hipGraphLaunch(exec, s->stream) != hipSuccess) return fail(SA_E_HIP, "launch of the multi-shot graph failed");
    }
    return rc;
}

// wavefield `which` (0: last level, 1: the one before) of the pair holding shot `shot`, its half;
// for the shorter shot of a pair the pair ran on to its partner's last step
extern "C" int sa_multi_get_wavefield(sa_solver *s, int shot, int which, float *out) {
    if (!s || !out || shot < 0 || shot >= (int)s->mshot.size() || s->morder.empty())
        return fail(SA_E_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(s->device));
    int pos = 0;
    while (s->morder[pos] != shot) ++pos;
    const int j = pos / 2, half = pos % 2;
    const int tlast = s->mshot[s->morder[2 * j]].nt - 2;
    const int last = (tlast + 1) % 2;
    const Geo &g = s->geo;
    std::vector<float> tmp(2 * (size_t)s->total);
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(tmp.data(), s->mbuf[2 * j + (which == 0 ? last : 1 - last)], tmp.size() * sizeof(float),
                      hipMemcpyDeviceToHost));
    for (int i = 0; i < g.n[0]; ++i)
        for (int jj = 0; jj < g.n[1]; ++jj)
            for (int k = 0; k < g.n[2]; ++k)
                out[((long long)i * g.n[1] + jj) * g.n[2] + k] =
                    tmp[2 * (g.base + (long long)i * g.s0 + (long long)jj * g.s1 + k) + half];
    return SA_OK;
}

// capture + instantiation time of the last graph-launched schedule (milliseconds)
extern "C" int sa_multi_graph_ms(sa_solver *s, double *ms) {
    if (!s || !ms) return fail(SA_E_ARG, "null argument");
    *ms = s->graph_ms;
    return SA_OK;
}

extern "C" int sa_read_gradient_dev(sa_solver *s, int which, float *dev_out) {
    if (!s || !dev_out || which < 0 || which > 3) return fail(SA_E_ARG, "bad arguments");
    if (!s->gacc[which]) return fail(SA_E_STATE, "gradient %d was not computed", which);
    HIP_TRY(hipSetDevice(s->device));
    k_unpad<<<domain_grid(s->geo, kBlk), kBlk, 0, s->stream>>>(s->geo, s->gacc[which], dev_out);
    CHECK_LAUNCH();
    return SA_OK;
}

// accessors for the processing kernels (csrc/process.hip)
hipStream_t sa_solver_stream(sa_solver *s) { return s->stream; }
int sa_solver_device(sa_solver *s) { return s->device; }
int sa_fail(int code, const char *msg) { return fail(code, "%s", msg); }
