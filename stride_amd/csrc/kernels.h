// Device-side geometry and finite-difference helpers shared by the solver kernels.
//
// Wavefields and per-point coefficient arrays live in one padded layout: three
// internal axes (a 2D grid runs with axis 0 of size 1), C order with axis 2
// fastest, a zero halo of 10 cells on axes 0/1 and the trailing side of axis 2,
// 32 leading cells on axis 2 so that every interior row starts 128-byte
// aligned, and the row pitch rounded up to 32 floats.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SA_HALO 10
#define SA_LEAD 32
#define SA_RADIUS 5

struct Geo {
    int n[3];        // domain size per internal axis
    int act[3];      // axis active (2D: axis 0 inactive)
    int nabs[3];     // absorbing layer per axis
    long long s0, s1;  // strides of axis 0 and 1 (axis 2 stride 1)
    long long base;    // padded offset of domain point (0, 0, 0)
    long long total;   // padded element count
    float w2[3][6];  // SO10 second-derivative weights / h^2 (zero on inactive axes)
    float w1[3][6];  // SO10 centred first-derivative weights / h
    float ws[3][5];  // SO10 staggered first-derivative weights / h
};

__device__ __forceinline__ long long sa_stride(const Geo &g, int a) {
    return a == 0 ? g.s0 : (a == 1 ? g.s1 : 1);
}

__device__ __forceinline__ long long sa_pidx(const Geo &g, int i, int j, int k) {
    return g.base + (long long)i * g.s0 + (long long)j * g.s1 + k;
}

// unpadded linear index (extended-grid C order) -> padded index
__device__ __forceinline__ long long sa_lin2p(const Geo &g, long long lin) {
    int k = (int)(lin % g.n[2]);
    long long r = lin / g.n[2];
    int j = (int)(r % g.n[1]);
    int i = (int)(r / g.n[1]);
    return sa_pidx(g, i, j, k);
}

__device__ __forceinline__ float sa_lap(const Geo &g, const float *__restrict__ u, long long p) {
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!g.act[a]) continue;
        long long s = sa_stride(g, a);
        float v = g.w2[a][0] * u[p];
#pragma unroll
        for (int r = 1; r <= 5; ++r) v += g.w2[a][r] * (u[p + r * s] + u[p - r * s]);
        acc += v;
    }
    return acc;
}

__device__ __forceinline__ float sa_d1(const Geo &g, const float *__restrict__ u, long long p, int a) {
    long long s = sa_stride(g, a);
    float v = 0.f;
#pragma unroll
    for (int r = 1; r <= 5; ++r) v += g.w1[a][r] * (u[p + r * s] - u[p - r * s]);
    return v;
}

__device__ __forceinline__ float sa_d2(const Geo &g, const float *__restrict__ u, long long p, int a) {
    long long s = sa_stride(g, a);
    float v = g.w2[a][0] * u[p];
#pragma unroll
    for (int r = 1; r <= 5; ++r) v += g.w2[a][r] * (u[p + r * s] + u[p - r * s]);
    return v;
}

// staggered derivative at q + 1/2 along a
__device__ __forceinline__ float sa_dplus(const Geo &g, const float *__restrict__ u, long long q, int a) {
    long long s = sa_stride(g, a);
    float v = 0.f;
#pragma unroll
    for (int r = 0; r < 5; ++r) v += g.ws[a][r] * (u[q + (r + 1) * s] - u[q - r * s]);
    return v;
}

// staggered derivative at node p of a half-point array (h[q] holds the value at q + 1/2);
// with h = u it is the derivative evaluated at p - 1/2 (grad(f, shift=-0.5) on nodes)
__device__ __forceinline__ float sa_dminus(const Geo &g, const float *__restrict__ h, long long p, int a) {
    long long s = sa_stride(g, a);
    float v = 0.f;
#pragma unroll
    for (int r = 0; r < 5; ++r) v += g.ws[a][r] * (h[p + r * s] - h[p - (r + 1) * s]);
    return v;
}

// Sub-box of the domain, C order; snapshots are stored in this layout. Rows (axis 2) have a
// pitch p2 that is a multiple of 32 floats and start off2 = lo[2] % 32 cells in, so that a
// cell sits at the same 128-byte line offset as in the padded field layout and a 64-cell
// segment of a row is two whole cache lines in both.
struct Box {
    int lo[3], n[3];
    int p2, off2;
};

__device__ __forceinline__ long long sa_boxidx(const Box &b, int i, int j, int k) {
    return ((long long)(i - b.lo[0]) * b.n[1] + (j - b.lo[1])) * b.p2 + b.off2 + (k - b.lo[2]);
}

__device__ __forceinline__ bool sa_inbox(const Box &b, int i, int j, int k) {
    return i >= b.lo[0] && i < b.lo[0] + b.n[0] && j >= b.lo[1] && j < b.lo[1] + b.n[1] && k >= b.lo[2] &&
           k < b.lo[2] + b.n[2];
}
