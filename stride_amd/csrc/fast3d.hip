// Specialised 3D wavefield update (see fast3d.h).
//
// Work decomposition: a workgroup of 4 waves owns a tile of 64 (axis 2, the
// contiguous axis) x 4*RPT (axis 1) cells and marches along axis 0 over a chunk of
// planes. Each lane keeps an 11-deep register queue of u along axis 0 for each of
// its RPT cells, so the axis-0 second derivative costs no memory traffic; the
// current plane is staged in LDS with a 5-cell halo for the axis-1 / axis-2
// derivatives. The halo cells are re-read from global memory, where the
// neighbouring tiles loaded them as their own cells five planes earlier: the
// block -> tile mapping puts consecutive tiles on the same XCD (blocks are
// dispatched round-robin over the 8 XCDs), so those re-reads hit that XCD's L2.
//
// Per cell and step the kernel reads u (4 B), the other level (4 B), vp (4 B) and
// writes the new level (4 B). The sponge coefficients are evaluated on the fly
// from vp and the per-axis layer tables, only in tiles that touch the layer.

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/stride_amd.h"
#include "fast3d.h"

namespace {

constexpr int kTZ = 64;
constexpr int kWaves = 4;

struct FastArgs {
    const float *u;
    const float *o;
    float *out;
    const float *vp;
    const float *tables;
    int tlen;
    const float *axis;  // per-axis layer tables [3][alen] (version 5)
    int alen;
    float dt, dt2, idt2;
    float *save;
    const float *ps;
    Box box;
    float *grad;
    float *prec;
    int ntile_y, ntile_z, chunk;
    int nblocks, per_xcd;
    // several shot pairs per launch (the multi-shot schedule): per-pair level buffers, active
    // pairs, which level is u (parity of the step) and the tiles per group of the mapping
    const PairLevels *pairs;
    int npair, parity, group;
};

// MODE 0: plain step, 1: fused save of p.dt2, 2: fused gradient accumulation
template <int RPT, int MODE, bool SPONGE>
__global__ __launch_bounds__(256) void k_fast3d(Geo g, FastArgs a) {
    constexpr int TY = kWaves * RPT;
    constexpr int LY = TY + 10, LZ = kTZ + 10;
    __shared__ float s[LY][LZ];

    // XCD-aware mapping: hardware block b runs on XCD b % 8; give each XCD a contiguous
    // range of logical tiles
    const int b = blockIdx.x;
    const int L = (b & 7) * a.per_xcd + (b >> 3);
    if (L >= a.nblocks) return;
    const int tz = L % a.ntile_z;
    const int rest = L / a.ntile_z;
    const int ty = rest % a.ntile_y;
    const int ch = rest / a.ntile_y;
    const int z0 = tz * kTZ, y0 = ty * TY;
    const int x0 = ch * a.chunk;
    const int x1 = min(x0 + a.chunk, g.n[0]);

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int z = z0 + lane;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const bool zload = z < nz + SA_RADIUS;
    const bool zin = z < nz;
    const long long s0 = g.s0, s1 = g.s1;

    const float w2x[6] = {g.w2[0][0], g.w2[0][1], g.w2[0][2], g.w2[0][3], g.w2[0][4], g.w2[0][5]};
    const float w2y[6] = {g.w2[1][0], g.w2[1][1], g.w2[1][2], g.w2[1][3], g.w2[1][4], g.w2[1][5]};
    const float w2z[6] = {g.w2[2][0], g.w2[2][1], g.w2[2][2], g.w2[2][3], g.w2[2][4], g.w2[2][5]};

    int yr[RPT];
    bool rin[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        yr[r] = y0 + w + kWaves * r;
        rin[r] = yr[r] < ny && zin;
    }
    const bool tile_yz_interior = SPONGE && y0 >= g.nabs[1] && y0 + TY <= ny - g.nabs[1] && z0 >= g.nabs[2] &&
                                  z0 + kTZ <= nz - g.nabs[2];

    // register queue: q[r][k] = u at plane x - 5 + k
    float q[RPT][11];
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        const int xx = x0 - SA_RADIUS + k;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const long long p = g.base + xx * s0 + yr[r] * s1 + z;
            q[r][k] = (yr[r] < ny + SA_RADIUS && zload && xx < nx + SA_RADIUS) ? a.u[p] : 0.f;
        }
    }

    for (int x = x0; x < x1; ++x) {
        const long long px = g.base + (long long)x * s0;
        // front plane x + 5
#pragma unroll
        for (int r = 0; r < RPT; ++r)
            q[r][10] = (yr[r] < ny + SA_RADIUS && zload && x + SA_RADIUS < nx + SA_RADIUS)
                           ? a.u[px + SA_RADIUS * s0 + yr[r] * s1 + z]
                           : 0.f;
        // the pointwise operands of this plane
        float ov[RPT], vv[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const long long p = px + yr[r] * s1 + z;
            ov[r] = rin[r] ? a.o[p] : 0.f;
            vv[r] = rin[r] ? a.vp[p] : 0.f;
        }
        // stage plane x: own cells from the queue, halo rows / columns from memory
#pragma unroll
        for (int r = 0; r < RPT; ++r) s[w + kWaves * r + 5][lane + 5] = q[r][5];
        for (int hy = w; hy < 10; hy += kWaves) {
            const int yy = hy < 5 ? y0 - 5 + hy : y0 + TY + hy - 5;
            const int ly = hy < 5 ? hy : TY + hy;
            s[ly][lane + 5] = (yy < ny + SA_RADIUS && zload) ? a.u[px + yy * s1 + z] : 0.f;
        }
        for (int i = threadIdx.x; i < 10 * TY; i += 256) {
            const int row = i / 10, c = i % 10;
            const int zz = c < 5 ? z0 - 5 + c : z0 + kTZ + c - 5;
            const int yy = y0 + row;
            const int lz = c < 5 ? c : kTZ + c;
            s[row + 5][lz] = (yy < ny + SA_RADIUS && zz < nz + SA_RADIUS) ? a.u[px + yy * s1 + zz] : 0.f;
        }
        __syncthreads();

        const bool plane_interior = tile_yz_interior && x >= g.nabs[0] && x < nx - g.nabs[0];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            if (!rin[r]) continue;
            const int ly = w + kWaves * r + 5, lz = lane + 5;
            const float c = q[r][5];
            float vx = w2x[0] * c, vy = w2y[0] * c, vz = w2z[0] * c;
#pragma unroll
            for (int k = 1; k <= 5; ++k) {
                vx += w2x[k] * (q[r][5 + k] + q[r][5 - k]);
                vy += w2y[k] * (s[ly + k][lz] + s[ly - k][lz]);
                vz += w2z[k] * (s[ly][lz + k] + s[ly][lz - k]);
            }
            const float lap = vx + vy + vz;
            const float v = vv[r];
            const float v2 = v * v;
            float A = 2.f, B = -1.f, G = a.dt2 * v2;
            if (SPONGE && !plane_interior) {
                // layer value from the corner table, indexed by the clipped distance to the edge per axis
                const int M = a.tlen;
                const int d0 = min(min(x, nx - 1 - x), M);
                const int d1 = min(min(yr[r], ny - 1 - yr[r]), M);
                const int d2 = min(min(z, nz - 1 - z), M);
                const float sm = a.tables[(d0 * (M + 1) + d1) * (M + 1) + d2];
                if (sm != 0.f) {
                    const float D = (7.f * (sm * v)) * a.dt;
                    const float aa = a.dt * v2 * D;
                    const float e = a.dt * a.dt * v2 * D * D;
                    const float ia = __builtin_amdgcn_rcpf(1.f + aa);  // v_rcp_f32 (1 ulp), no divide sequence
                    A = (2.f - e) * ia;
                    B = -(1.f - aa) * ia;
                    G = (a.dt2 * ia) * v2;
                }
            }
            const float val = A * c + B * ov[r] + G * lap;
            const long long p = px + yr[r] * s1 + z;
            a.out[p] = val;
            if (MODE == 1) {
                if (sa_inbox(a.box, x, yr[r], z)) a.save[sa_boxidx(a.box, x, yr[r], z)] = (val - 2.f * c + ov[r]) * a.idt2;
            } else if (MODE == 2) {
                if (sa_inbox(a.box, x, yr[r], z)) {
                    const float psv = a.ps[sa_boxidx(a.box, x, yr[r], z)];
                    a.grad[p] += psv * c;
                    a.prec[p] += psv * psv;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RPT; ++r)
#pragma unroll
            for (int k = 0; k < 10; ++k) q[r][k] = q[r][k + 1];
    }
}


// Second-generation march (version 2): ~40 VALU instructions per cell and plane instead of ~100.
//  - u, the other level, vp, out and the gradient arrays are read / written with raw buffer
//    instructions: one scalar resource per array, a constant per-lane byte offset computed
//    once, and the plane offset in a scalar register, so no address arithmetic runs per load;
//  - out-of-range lanes are clamped onto zero halo cells instead of predicated;
//  - the register queue rotates statically (the plane loop is unrolled by 11);
//  - plane x + 1 is loaded before plane x is computed, and the LDS plane is double-buffered,
//    so there is one barrier per plane.
// Requires the 3D padded layout (leads 10, 10, 32 cells) and arrays below 2 GiB.
typedef __amdgpu_buffer_rsrc_t rsrc_t;

__device__ __forceinline__ rsrc_t make_rsrc(const void *p, long long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float bload(rsrc_t r, unsigned voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// Streaming (non-temporal) variants, cache-policy bit "nt": for arrays touched once per step
// (the other level, vp, the new level, snapshots and gradient accumulators), so that the L2
// keeps the u planes that neighbouring tiles read as halo.
constexpr int kNT = 2;

__device__ __forceinline__ float bload_nt(rsrc_t r, unsigned voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, kNT));
}

__device__ __forceinline__ void bstore_nt(float v, rsrc_t r, unsigned voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, kNT);
}

__device__ __forceinline__ void bstore(float v, rsrc_t r, unsigned voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}

template <int RPT>
struct Lanes {
    static constexpr int TY = kWaves * RPT;
    static constexpr int LY = TY + 10, LZ = kTZ + 10;
    static constexpr int NHZ = (10 * TY + 255) / 256;  // z-halo cells per lane (upper bound)
    unsigned row[RPT];   // byte offset of each own cell within a padded plane (clamped)
    unsigned hy[3];      // y-halo cells
    unsigned hz[NHZ];    // z-halo cells
    int hy_lds[3];       // LDS row of each y-halo cell (-1: none)
    int hz_lds[NHZ];     // flattened LDS index of each z-halo cell (-1: none)
    unsigned brow[RPT];  // byte offset of each own cell within a box plane (MODE 1/2)
    bool bin[RPT];       // own cell inside the box along axes 1/2
};

struct Rsrc {
    rsrc_t u, o, vp, out, save, ps, grad, prec;
    int plane_bytes, box_plane_bytes;
};

// operands of one plane, loaded ahead of use: front cell of the queue, other level, vp, halo cells
template <int RPT>
struct Stage {
    float q[RPT], o[RPT], v[RPT], hy[3], hz[Lanes<RPT>::NHZ];
};

template <int RPT>
__device__ __forceinline__ void load_stage(Stage<RPT> &st, const Rsrc &rs, const Lanes<RPT> &ln, int xf, int xp,
                                           int nx) {
    const int sf = (min(xf, nx + SA_RADIUS - 1) + SA_HALO) * rs.plane_bytes;
    const int sp = (min(xp, nx - 1) + SA_HALO) * rs.plane_bytes;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        st.q[r] = bload(rs.u, ln.row[r], sf);
        st.o[r] = bload(rs.o, ln.row[r], sp);
        st.v[r] = bload(rs.vp, ln.row[r], sp);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) st.hy[k] = bload(rs.u, ln.hy[k], sp);
#pragma unroll
    for (int k = 0; k < Lanes<RPT>::NHZ; ++k) st.hz[k] = bload(rs.u, ln.hz[k], sp);
}

// One plane of the march. PD = prefetch distance in planes: with PD = 1 the operands of plane
// x + 1 are loaded at the top of plane x; with PD = 2 those of plane x + 2 are (and `s1` holds
// plane x + 1's, loaded one plane earlier).
template <int RPT, int MODE, bool SPONGE, int PD, int J>
__device__ __forceinline__ void march2_plane(const Geo &g, const FastArgs &a, const Rsrc &rs, float *__restrict__ lds,
                                             int &cur, float (&q)[RPT][11], float (&ov)[RPT], float (&vv)[RPT],
                                             Stage<RPT> &s1, const int x, const int x1, const Lanes<RPT> &ln,
                                             const int (&yr)[RPT], const bool (&rin)[RPT], const int z,
                                             const int lane, const int w, const bool tile_yz_interior) {
    using L = Lanes<RPT>;
    constexpr int LZ = L::LZ;
    constexpr int PLANE = L::LY * L::LZ;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const bool more = x + 1 < x1;
    // 1. loads PD planes ahead
    Stage<RPT> sn;
    load_stage<RPT>(sn, rs, ln, x + 5 + PD, x + PD, nx);
    // 2. compute plane x (LDS buffer `cur`; queue slot (J + k) % 11 holds plane x - 5 + k)
    const float *__restrict__ sb = lds + cur * PLANE;
    const bool plane_interior = tile_yz_interior && x >= g.nabs[0] && x < nx - g.nabs[0];
    const int sx = (x + SA_HALO) * rs.plane_bytes;
    const bool xin_box = MODE != 0 && x >= a.box.lo[0] && x < a.box.lo[0] + a.box.n[0];
    const int sbx = MODE != 0 ? (x - a.box.lo[0]) * rs.box_plane_bytes : 0;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int li = (w + kWaves * r + 5) * LZ + lane + 5;
        const float c = q[r][(J + 5) % 11];
        float vx = g.w2[0][0] * c, vy = g.w2[1][0] * c, vz = g.w2[2][0] * c;
#pragma unroll
        for (int k = 1; k <= 5; ++k) {
            vx += g.w2[0][k] * (q[r][(J + 5 + k) % 11] + q[r][(J + 5 - k) % 11]);
            vy += g.w2[1][k] * (sb[li + k * LZ] + sb[li - k * LZ]);
            vz += g.w2[2][k] * (sb[li + k] + sb[li - k]);
        }
        const float lap = vx + vy + vz;
        const float v = vv[r];
        const float v2 = v * v;
        float A = 2.f, B = -1.f, G = a.dt2 * v2;
        if (SPONGE && !plane_interior) {
            const int M = a.tlen;
            const int d0 = min(min(x, nx - 1 - x), M);
            const int d1 = min(min(yr[r], ny - 1 - yr[r]), M);
            const int d2 = min(min(z, nz - 1 - z), M);
            const float sm = a.tables[(d0 * (M + 1) + d1) * (M + 1) + d2];
            if (sm != 0.f) {
                const float D = (7.f * (sm * v)) * a.dt;
                const float aa = a.dt * v2 * D;
                const float e = a.dt * a.dt * v2 * D * D;
                const float ia = __builtin_amdgcn_rcpf(1.f + aa);  // v_rcp_f32 (1 ulp), no divide sequence
                A = (2.f - e) * ia;
                B = -(1.f - aa) * ia;
                G = (a.dt2 * ia) * v2;
            }
        }
        const float val = A * c + B * ov[r] + G * lap;
        if (rin[r]) {
            bstore(val, rs.out, ln.row[r], sx);
            if (MODE != 0 && xin_box && ln.bin[r]) {
                if (MODE == 1) {
                    bstore((val - 2.f * c + ov[r]) * a.idt2, rs.save, ln.brow[r], sbx);
                } else {
                    const float psv = bload(rs.ps, ln.brow[r], sbx);
                    bstore(bload(rs.grad, ln.row[r], sx) + psv * c, rs.grad, ln.row[r], sx);
                    bstore(bload(rs.prec, ln.row[r], sx) + psv * psv, rs.prec, ln.row[r], sx);
                }
            }
        }
    }
    // 3. stage plane x + 1 in the other buffer, rotate the queue (plane x + 6 replaces x - 5)
    const Stage<RPT> &nx1 = PD == 1 ? sn : s1;
    if (more) {
        float *__restrict__ sw = lds + (cur ^ 1) * PLANE;
#pragma unroll
        for (int r = 0; r < RPT; ++r) sw[(w + kWaves * r + 5) * LZ + lane + 5] = q[r][(J + 6) % 11];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (ln.hy_lds[k] >= 0) sw[ln.hy_lds[k] * LZ + lane + 5] = nx1.hy[k];
#pragma unroll
        for (int k = 0; k < L::NHZ; ++k)
            if (ln.hz_lds[k] >= 0) sw[ln.hz_lds[k]] = nx1.hz[k];
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        q[r][J % 11] = nx1.q[r];
        ov[r] = nx1.o[r];
        vv[r] = nx1.v[r];
    }
    if (PD == 2) s1 = sn;
    __syncthreads();
    cur ^= 1;
}

template <int RPT, int MODE, bool SPONGE, int PD>
__global__ __launch_bounds__(256) void k_fast3d_v2(Geo g, FastArgs a) {
    using L = Lanes<RPT>;
    constexpr int TY = L::TY;
    constexpr int LZ = L::LZ;
    __shared__ float lds[2 * L::LY * L::LZ];
    const int b = blockIdx.x;
    const int Lg = (b & 7) * a.per_xcd + (b >> 3);
    if (Lg >= a.nblocks) return;
    const int tz = Lg % a.ntile_z;
    const int rest = Lg / a.ntile_z;
    const int ty = rest % a.ntile_y;
    const int ch = rest / a.ntile_y;
    const int z0 = tz * kTZ, y0 = ty * TY;
    const int x0 = ch * a.chunk;
    const int x1 = min(x0 + a.chunk, g.n[0]);
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int z = z0 + lane;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const unsigned s1 = (unsigned)g.s1;
    Rsrc rs;
    const long long bytes = g.total * 4;
    rs.u = make_rsrc(a.u, bytes);
    rs.o = make_rsrc(a.o, bytes);
    rs.vp = make_rsrc(a.vp, bytes);
    rs.out = make_rsrc(a.out, bytes);
    rs.plane_bytes = (int)(g.s0 * 4);
    rs.box_plane_bytes = a.box.n[1] * a.box.p2 * 4;
    const long long box_bytes = (long long)a.box.n[0] * rs.box_plane_bytes;
    if (MODE == 1) rs.save = make_rsrc(a.save, box_bytes);
    if (MODE == 2) {
        rs.ps = make_rsrc(a.ps, box_bytes);
        rs.grad = make_rsrc(a.grad, bytes);
        rs.prec = make_rsrc(a.prec, bytes);
    }
    // lane offsets (padded coordinates; clamped onto zero halo cells when out of range)
    const int zc = min(z, nz + SA_RADIUS - 1) + SA_LEAD;
    L ln;
    int yr[RPT];
    bool rin[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        yr[r] = y0 + w + kWaves * r;
        rin[r] = yr[r] < ny && z < nz;
        ln.row[r] = ((unsigned)(min(yr[r], ny + SA_RADIUS - 1) + SA_HALO) * s1 + (unsigned)zc) * 4u;
        ln.bin[r] = yr[r] >= a.box.lo[1] && yr[r] < a.box.lo[1] + a.box.n[1] && z >= a.box.lo[2] &&
                    z < a.box.lo[2] + a.box.n[2];
        ln.brow[r] = ln.bin[r] ? (unsigned)((yr[r] - a.box.lo[1]) * a.box.p2 + a.box.off2 + (z - a.box.lo[2])) * 4u : 0u;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int hy = w + kWaves * k;
        ln.hy_lds[k] = -1;
        ln.hy[k] = 0u;
        if (hy < 10) {
            const int yy = hy < 5 ? y0 - 5 + hy : y0 + TY + hy - 5;
            ln.hy_lds[k] = hy < 5 ? hy : TY + hy;
            ln.hy[k] = ((unsigned)(min(yy, ny + SA_RADIUS - 1) + SA_HALO) * s1 + (unsigned)zc) * 4u;
        }
    }
#pragma unroll
    for (int k = 0; k < L::NHZ; ++k) {
        const int i = threadIdx.x + 256 * k;
        ln.hz_lds[k] = -1;
        ln.hz[k] = 0u;
        if (i < 10 * TY) {
            const int row = i / 10, c = i % 10;
            const int zz = c < 5 ? z0 - 5 + c : z0 + kTZ + c - 5;
            const int yy = y0 + row;
            ln.hz_lds[k] = (row + 5) * LZ + (c < 5 ? c : kTZ + c);
            ln.hz[k] = ((unsigned)(min(yy, ny + SA_RADIUS - 1) + SA_HALO) * s1 +
                        (unsigned)(min(zz, nz + SA_RADIUS - 1) + SA_LEAD)) * 4u;
        }
    }
    const bool tile_yz_interior = SPONGE && y0 >= g.nabs[1] && y0 + TY <= ny - g.nabs[1] && z0 >= g.nabs[2] &&
                                  z0 + kTZ <= nz - g.nabs[2];
    // prologue: queue planes x0 - 5 .. x0 + 5, operands and LDS plane of x0 (and plane x0 + 1's
    // operands for PD = 2)
    float q[RPT][11], ov[RPT], vv[RPT];
#pragma unroll
    for (int k = 0; k < 11; ++k) {
        const int sk = (min(x0 - SA_RADIUS + k, nx + SA_RADIUS - 1) + SA_HALO) * rs.plane_bytes;
#pragma unroll
        for (int r = 0; r < RPT; ++r) q[r][k] = bload(rs.u, ln.row[r], sk);
    }
    Stage<RPT> s1st;
    {
        Stage<RPT> s0st;
        load_stage<RPT>(s0st, rs, ln, x0 + 5, x0, nx);
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            ov[r] = s0st.o[r];
            vv[r] = s0st.v[r];
            lds[(w + kWaves * r + 5) * LZ + lane + 5] = q[r][5];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (ln.hy_lds[k] >= 0) lds[ln.hy_lds[k] * LZ + lane + 5] = s0st.hy[k];
#pragma unroll
        for (int k = 0; k < L::NHZ; ++k)
            if (ln.hz_lds[k] >= 0) lds[ln.hz_lds[k]] = s0st.hz[k];
        if (PD == 2) load_stage<RPT>(s1st, rs, ln, x0 + 6, x0 + 1, nx);
    }
    __syncthreads();
    int cur = 0;
    for (int xb = x0; xb < x1; xb += 11) {
#define SA_PLANE(J)                                                                                            \
    if (xb + J < x1)                                                                                           \
        march2_plane<RPT, MODE, SPONGE, PD, J>(g, a, rs, lds, cur, q, ov, vv, s1st, xb + J, x1, ln, yr, rin, z, \
                                               lane, w, tile_yz_interior);
        SA_PLANE(0) SA_PLANE(1) SA_PLANE(2) SA_PLANE(3) SA_PLANE(4) SA_PLANE(5)
        SA_PLANE(6) SA_PLANE(7) SA_PLANE(8) SA_PLANE(9) SA_PLANE(10)
#undef SA_PLANE
    }
}

// Fifth-generation march (versions 5 and 6): the plane tile with its halo is staged in a ring
// of three LDS slots by direct-to-LDS buffer loads (16 B per lane, no VGPR staging and no LDS
// store instructions), issued two planes ahead; the own-column operands (u five planes ahead
// for the register queue, the other level, vp and, when fused, the gradient operands) are
// register loads issued two to three planes ahead. Every load and store is unconditional:
// lanes outside the domain get a byte offset past the end of the buffer, which the buffer
// range check turns into a zero read or a dropped write, so the plane loop has no divergent
// branches. The sponge coefficient is evaluated from the per-axis layer tables (two per-lane
// constants and one scalar per plane) instead of a table gather.
//
// Synchronisation per plane: each wave waits for its own direct-to-LDS loads into the slot of
// the plane (a dummy read of the slot makes the compiler emit exactly that wait), then one
// workgroup barrier; the slot written next is the one every wave finished reading before that
// barrier.
constexpr unsigned kOOB = 0x80000000u;  // byte offset past every buffer (< 2 GiB)

template <int TY>
struct Ring {
    static constexpr int ROWS = TY + 10;     // tile rows and a 5-row halo on each side
    static constexpr int CPR = 20;           // 16-byte chunks per row: axis-2 cells z0 - 8 .. z0 + 71
    static constexpr int PITCH = 4 * CPR;    // floats per LDS row
    static constexpr int NCH = ROWS * CPR;   // chunks per plane
    static constexpr int NI = (NCH + 63) / 64;  // 64-lane load instructions per plane
};

// every wave issues the same number of direct-to-LDS loads per plane (the surplus ones read
// past the end of the buffer and fill slot padding with zeros), so that all waves run one
// straight-line instruction stream and the compiler's wait counts stay exact
template <int TY, int NW>
struct RingW : Ring<TY> {
    static constexpr int KI = (Ring<TY>::NI + NW - 1) / NW;  // loads per wave per plane
    static constexpr int FLOATS = KI * NW * 256;             // slot size in floats
};

typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ void dma16(rsrc_t r, float *lds, unsigned voff, int soff) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void *)lds, 16, voff, soff, 0, 0);
}

template <int RPT, int MODE>
struct Stage5 {
    float o[RPT], v[RPT];
    float ps[MODE == 2 ? RPT : 1], gr[MODE == 2 ? RPT : 1], pr[MODE == 2 ? RPT : 1];
};

template <int RPT, int MODE>
__device__ __forceinline__ void load_stage5(Stage5<RPT, MODE> &st, const Rsrc &rs, const unsigned (&row)[RPT],
                                            const unsigned (&grow)[RPT], const unsigned (&brow)[RPT], int x,
                                            const FastArgs &a) {
    const int sp = (x + SA_HALO) * rs.plane_bytes;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        st.o[r] = bload_nt(rs.o, row[r], sp);
        st.v[r] = bload_nt(rs.vp, row[r], sp);
    }
    if (MODE == 2) {
        // box planes outside [lo, lo + n) read as zero (offset past the end)
        const bool xin = (x >= a.box.lo[0]) & (x < a.box.lo[0] + a.box.n[0]);
        const int sbx = xin ? (x - a.box.lo[0]) * rs.box_plane_bytes : 0;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            st.ps[r] = bload_nt(rs.ps, xin ? brow[r] : kOOB, sbx);
            st.gr[r] = bload_nt(rs.grad, grow[r], sp);
            st.pr[r] = bload_nt(rs.prec, grow[r], sp);
        }
    }
}

template <int RPT, int NW, int MODE, bool SPONGE, int NS, int J>
__device__ __forceinline__ void march5_plane(const Geo &g, const FastArgs &a, const Rsrc &rs, float *const (&sl)[6],
                                             float (&q)[RPT][12],
                                             Stage5<RPT, MODE> (&st)[3], const int x, const int x1, const int lane,
                                             const int w, const unsigned (&row)[RPT], const unsigned (&orow)[RPT],
                                             const unsigned (&grow)[RPT], const unsigned (&brow)[RPT],
                                             const unsigned (&dv)[RingW<NW * RPT, NW>::KI], const int (&d1)[RPT],
                                             const float (&t1v)[RPT], const int d2, const float t2v,
                                             const bool tile_yz_interior) {
    using R = RingW<NW * RPT, NW>;
    constexpr int KI = R::KI;
    float *__restrict__ cur = sl[J % NS];              // slot of plane x
    float *__restrict__ nxt = sl[(J + NS - 1) % NS];   // slot of plane x + NS - 1
    const int nx = g.n[0];
    // 1. this wave's loads into the slot of plane x have landed; then every wave's have. The
    // probe read is pinned here (memory clobbers on both sides): hoisted, its wait would
    // shorten the prefetch distance.
    asm volatile("" ::: "memory");
    {
        const float probe = cur[lane];
        asm volatile("" ::"v"(probe));
    }
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    // 2. plane x + NS - 1 into the slot every wave finished reading before the barrier (plane
    // indices past the last halo plane are clamped: those planes are never stored)
    const int xlast = nx + SA_HALO - 1;
    {
        const int soff = (min(x + NS - 1, xlast) + SA_HALO) * rs.plane_bytes;
#pragma unroll
        for (int k = 0; k < KI; ++k) dma16(rs.u, nxt + (w + NW * k) * 256, dv[k], soff);
    }
    // 3. compute plane x; queue slot (J + 5 + k) % 12 holds plane x + k
    Stage5<RPT, MODE> &sx = st[J % 3];
    const int sxo = (x + SA_HALO) * rs.plane_bytes;
    // planes past the chunk end (the last 12-plane period) are computed but not stored
    const bool xin = x < x1;
    // bitwise, not short-circuit: a branch here splits the plane's schedule
    const bool xin_box = MODE != 0 && (xin & (x >= a.box.lo[0]) & (x < a.box.lo[0] + a.box.n[0]));
    const int sbx = MODE == 1 && xin_box ? (x - a.box.lo[0]) * rs.box_plane_bytes : 0;
    const bool plane_interior = tile_yz_interior && x >= g.nabs[0] && x < nx - g.nabs[0];
    int d0 = 0;
    float t0v = 0.f;
    if (SPONGE && !plane_interior) {
        d0 = max(min(x, nx - 1 - x), 0);
        t0v = d0 < g.nabs[0] ? a.axis[d0] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        // the lane offset is made opaque per plane so that the LDS addresses (one base per
        // 1 KiB window of ds_read2 offsets) are formed here instead of being hoisted out of
        // the march for every slot, which would hold tens of VGPRs
        int li = (w + NW * r + 5) * R::PITCH + lane + 8;
        asm volatile("" : "+v"(li));
        const float c = q[r][(J + 5) % 12];
        float vx = g.w2[0][0] * c, vy = g.w2[1][0] * c, vz = g.w2[2][0] * c;
#pragma unroll
        for (int k = 1; k <= 5; ++k) {
            vx += g.w2[0][k] * (q[r][(J + 5 + k) % 12] + q[r][(J + 17 - k) % 12]);
            vy += g.w2[1][k] * (cur[li + k * R::PITCH] + cur[li - k * R::PITCH]);
            vz += g.w2[2][k] * (cur[li + k] + cur[li - k]);
        }
        const float lap = vx + vy + vz;
        const float v = sx.v[r];
        const float v2 = v * v;
        float A = 2.f, B = -1.f, G = a.dt2 * v2;
        if (SPONGE && !plane_interior) {
            // layer rule: axis d contributes its table value at distance d_d from the edge
            // when no other axis is closer to its own edge
            float sm = 0.f;
            if (d0 <= min(d1[r], d2)) sm += t0v;
            if (d1[r] <= min(d0, d2)) sm += t1v[r];
            if (d2 <= min(d0, d1[r])) sm += t2v;
            if (sm != 0.f) {
                const float D = (7.f * (sm * v)) * a.dt;
                const float aa = a.dt * v2 * D;
                const float e = a.dt * a.dt * v2 * D * D;
                const float ia = __builtin_amdgcn_rcpf(1.f + aa);  // v_rcp_f32 (1 ulp), no divide sequence
                A = (2.f - e) * ia;
                B = -(1.f - aa) * ia;
                G = (a.dt2 * ia) * v2;
            }
        }
        const float val = A * c + B * sx.o[r] + G * lap;
        bstore_nt(val, rs.out, xin ? orow[r] : kOOB, sxo);
        if (MODE == 1) bstore_nt((val - 2.f * c + sx.o[r]) * a.idt2, rs.save, xin_box ? brow[r] : kOOB, sbx);
        if (MODE == 2) {
            const float psv = sx.ps[r];
            const unsigned go = xin_box ? grow[r] : kOOB;
            bstore_nt(sx.gr[r] + psv * c, rs.grad, go, sxo);
            bstore_nt(sx.pr[r] + psv * psv, rs.prec, go, sxo);
        }
    }
    // 4. refill: u of plane x + 7 into the queue slot of plane x - 5, operands of plane x + 3
    {
        const int sq = (min(x + 7, xlast) + SA_HALO) * rs.plane_bytes;
#pragma unroll
        for (int r = 0; r < RPT; ++r) q[r][J % 12] = bload(rs.u, row[r], sq);
    }
    load_stage5<RPT, MODE>(sx, rs, row, grow, brow, min(x + 3, xlast), a);
}

// LDS ring of NS plane slots (one array each: the compiler tracks direct-to-LDS loads per
// array, so a read of one slot waits only for the loads into that slot); plane x + NS - 1 is
// requested while plane x is computed
#define SA_RING_SLOTS(NS, FLOATS)                                                   \
    __shared__ __attribute__((aligned(16))) float s0[FLOATS];                       \
    __shared__ __attribute__((aligned(16))) float s1[FLOATS];                       \
    __shared__ __attribute__((aligned(16))) float s2[NS > 2 ? FLOATS : 4];          \
    __shared__ __attribute__((aligned(16))) float s3[NS > 3 ? FLOATS : 4];          \
    __shared__ __attribute__((aligned(16))) float s4[NS > 4 ? FLOATS : 4];          \
    __shared__ __attribute__((aligned(16))) float s5[NS > 5 ? FLOATS : 4];          \
    float *const sl[6] = {s0, s1, s2, s3, s4, s5};

template <int RPT, int NW, int MODE, bool SPONGE, int NS>
__global__ __launch_bounds__(NW * 64) void k_fast3d_v5(Geo g, FastArgs a) {
    constexpr int TY = NW * RPT;
    using R = RingW<TY, NW>;
    constexpr int KI = R::KI;
    static_assert(NS >= 3 && NS <= 6 && 12 % NS == 0, "ring depth: 3, 4 or 6 slots");
    SA_RING_SLOTS(NS, R::FLOATS)
    const int b = blockIdx.x;
    const int Lg = (b & 7) * a.per_xcd + (b >> 3);
    if (Lg >= a.nblocks) return;
    const int tz = Lg % a.ntile_z;
    const int rest = Lg / a.ntile_z;
    const int ty = rest % a.ntile_y;
    const int ch = rest / a.ntile_y;
    const int z0 = tz * kTZ, y0 = ty * TY;
    const int x0 = ch * a.chunk;
    const int x1 = min(x0 + a.chunk, g.n[0]);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int z = z0 + lane;
    const int ny = g.n[1], nz = g.n[2];
    const unsigned ps1 = (unsigned)g.s1;
    Rsrc rs;
    const long long bytes = g.total * 4;
    rs.u = make_rsrc(a.u, bytes);
    rs.o = make_rsrc(a.o, bytes);
    rs.vp = make_rsrc(a.vp, bytes);
    rs.out = make_rsrc(a.out, bytes);
    rs.plane_bytes = (int)(g.s0 * 4);
    rs.box_plane_bytes = a.box.n[1] * a.box.p2 * 4;
    const long long box_bytes = (long long)a.box.n[0] * rs.box_plane_bytes;
    if (MODE == 1) rs.save = make_rsrc(a.save, box_bytes);
    if (MODE == 2) {
        rs.ps = make_rsrc(a.ps, box_bytes);
        rs.grad = make_rsrc(a.grad, bytes);
        rs.prec = make_rsrc(a.prec, bytes);
    }
    // per-lane byte offsets within a plane: loads clamp onto halo cells, stores of lanes
    // outside the domain (or the box) go past the end of the buffer
    const unsigned zc = (unsigned)(min(z, nz + SA_RADIUS - 1) + SA_LEAD);
    unsigned row[RPT], orow[RPT], grow[RPT], brow[RPT];
    int d1[RPT];
    float t1v[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int yr = y0 + w + NW * r;
        const bool rin = yr < ny && z < nz;
        row[r] = ((unsigned)(min(yr, ny + SA_RADIUS - 1) + SA_HALO) * ps1 + zc) * 4u;
        orow[r] = rin ? row[r] : kOOB;
        const bool bin = rin && yr >= a.box.lo[1] && yr < a.box.lo[1] + a.box.n[1] && z >= a.box.lo[2] &&
                         z < a.box.lo[2] + a.box.n[2];
        brow[r] = bin ? (unsigned)((yr - a.box.lo[1]) * a.box.p2 + a.box.off2 + (z - a.box.lo[2])) * 4u : kOOB;
        grow[r] = bin ? row[r] : kOOB;
        d1[r] = min(yr, ny - 1 - yr);
        t1v[r] = 0.f;
        if (SPONGE && d1[r] >= 0 && d1[r] < g.nabs[1]) t1v[r] = a.axis[a.alen + d1[r]];
    }
    const int d2 = min(z, nz - 1 - z);
    const float t2v = SPONGE && d2 >= 0 && d2 < g.nabs[2] ? a.axis[2 * a.alen + d2] : 0.f;
    // direct-to-LDS chunk offsets: instruction w + NW k, lane -> chunk (w + NW k) * 64 + lane
    unsigned dv[KI];
#pragma unroll
    for (int k = 0; k < KI; ++k) {
        const int c = (w + NW * k) * 64 + lane;
        dv[k] = kOOB;
        if (c < R::NCH) {
            const int rr = c / R::CPR, cc = c % R::CPR;
            const int yy = min(y0 - 5 + rr, ny + SA_HALO - 1);
            const int zz = min(z0 + SA_LEAD - 8 + 4 * cc, (int)ps1 - 4);
            dv[k] = ((unsigned)(yy + SA_HALO) * ps1 + (unsigned)zz) * 4u;
        }
    }
    const bool tile_yz_interior = SPONGE && y0 >= g.nabs[1] && y0 + TY <= ny - g.nabs[1] && z0 >= g.nabs[2] &&
                                  z0 + kTZ <= nz - g.nabs[2];
    // prologue: LDS slots of planes x0, x0 + 1; queue planes x0 - 5 .. x0 + 6; operands of
    // planes x0 .. x0 + 2
    const int xlast = g.n[0] + SA_HALO - 1;
#pragma unroll
    for (int p = 0; p < NS - 1; ++p)
#pragma unroll
        for (int k = 0; k < KI; ++k)
            dma16(rs.u, sl[p] + (w + NW * k) * 256, dv[k], (min(x0 + p, xlast) + SA_HALO) * rs.plane_bytes);
    float q[RPT][12];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        const int sk = (min(x0 - SA_RADIUS + k, xlast) + SA_HALO) * rs.plane_bytes;
#pragma unroll
        for (int r = 0; r < RPT; ++r) q[r][k] = bload(rs.u, row[r], sk);
    }
    Stage5<RPT, MODE> st[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) load_stage5<RPT, MODE>(st[k], rs, row, grow, brow, min(x0 + k, xlast), a);
    // whole 12-plane periods, straight-line (no per-plane branch)
    const int nper = (x1 - x0 + 11) / 12;
    for (int it = 0; it < nper; ++it) {
        const int xb = x0 + 12 * it;
#define SA_PLANE5(J)                                                                                               \
    march5_plane<RPT, NW, MODE, SPONGE, NS, J>(g, a, rs, sl, q, st, xb + J, x1, lane, w, row, orow, grow, brow, \
                                           dv, d1, t1v, d2, t2v, tile_yz_interior);
        SA_PLANE5(0) SA_PLANE5(1) SA_PLANE5(2) SA_PLANE5(3) SA_PLANE5(4) SA_PLANE5(5)
        SA_PLANE5(6) SA_PLANE5(7) SA_PLANE5(8) SA_PLANE5(9) SA_PLANE5(10) SA_PLANE5(11)
#undef SA_PLANE5
    }
    // no direct-to-LDS load may outlive the workgroup's LDS allocation
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
}

// Seventh-generation march (versions 7 and 8): like version 5 (LDS plane ring filled by
// direct-to-LDS loads, branch-free planes), but u is read from memory once. Version 5 keeps
// a register window of u along axis 0, filled by separate loads five planes ahead of the
// plane tile that brings the same cells into LDS; with hundreds of tiles in flight per XCD
// the L2 cannot hold those five planes, so u is fetched from HBM twice. Here the own cell of
// plane p is read from the LDS tile when plane p arrives, and its axis-0 second-derivative
// contributions are pushed into a ring of partial sums for planes p - 5 .. p + 5; plane
// p - 5, now complete, is updated and stored. The first five planes of a chunk get their
// contributions from own-column loads before the march, the last five are finished after it
// with own-column loads of the five planes past the chunk.
template <int RPT, int MODE, bool SPONGE>
__device__ __forceinline__ void finish7(const Geo &g, const FastArgs &a, const Rsrc &rs, const int x, const int x0,
                                        const int x1, const float (&lap)[RPT], const float (&c)[RPT],
                                        const Stage5<RPT, MODE> &sx, const unsigned (&orow)[RPT],
                                        const unsigned (&grow)[RPT], const unsigned (&brow)[RPT],
                                        const int (&d1)[RPT], const float (&t1v)[RPT], const int d2, const float t2v) {
    const int nx = g.n[0];
    const int sxo = (x + SA_HALO) * rs.plane_bytes;
    // planes outside the chunk (before it, or past its end in the last 12-plane period) are
    // computed but not stored
    const bool xin = (x >= x0) & (x < x1);
    // bitwise, not short-circuit: a branch here splits the plane's schedule
    const bool xin_box = MODE != 0 && (xin & (x >= a.box.lo[0]) & (x < a.box.lo[0] + a.box.n[0]));
    const int sbx = MODE == 1 && xin_box ? (x - a.box.lo[0]) * rs.box_plane_bytes : 0;
    int d0 = 0;
    float t0v = 0.f;
    if (SPONGE) {
        d0 = max(min(x, nx - 1 - x), 0);
        const float tv = a.axis[min(d0, a.alen - 1)];
        t0v = d0 < g.nabs[0] ? tv : 0.f;
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const float v = sx.v[r];
        const float v2 = v * v;
        float A = 2.f, B = -1.f, G = a.dt2 * v2;
        if (SPONGE) {
            // branch-free layer rule (axis d contributes its table value at distance d_d from
            // the edge when no other axis is closer to its own edge); with sm = 0 every
            // expression reduces exactly to the interior values
            float sm = d0 <= min(d1[r], d2) ? t0v : 0.f;
            sm += d1[r] <= min(d0, d2) ? t1v[r] : 0.f;
            sm += d2 <= min(d0, d1[r]) ? t2v : 0.f;
            const float D = (7.f * (sm * v)) * a.dt;
            const float aa = a.dt * v2 * D;
            const float e = a.dt * a.dt * v2 * D * D;
            const float ia = __builtin_amdgcn_rcpf(1.f + aa);  // v_rcp_f32 (1 ulp), no divide sequence
            A = (2.f - e) * ia;
            B = -(1.f - aa) * ia;
            G = (a.dt2 * ia) * v2;
        }
        const float val = A * c[r] + B * sx.o[r] + G * lap[r];
        bstore_nt(val, rs.out, xin ? orow[r] : kOOB, sxo);
        if (MODE == 1) bstore_nt((val - 2.f * c[r] + sx.o[r]) * a.idt2, rs.save, xin_box ? brow[r] : kOOB, sbx);
        if (MODE == 2) {
            const float psv = sx.ps[r];
            const unsigned go = xin_box ? grow[r] : kOOB;
            bstore_nt(sx.gr[r] + psv * c[r], rs.grad, go, sxo);
            bstore_nt(sx.pr[r] + psv * psv, rs.prec, go, sxo);
        }
    }
}

// One step of the march: plane p (LDS slot J % 3, J = p - x0 mod 12). Partial-sum slot
// (J + k) % 12 holds plane p + k (k = -5 .. 5); centre slot J % 6 holds u of plane p; operand
// stage (J + 1) % 3 holds plane p - 5.
template <int RPT, int NW, int MODE, bool SPONGE, int NS, int J>
__device__ __forceinline__ void march7_step(const Geo &g, const FastArgs &a, const Rsrc &rs, float *const (&sl)[6],
                                            float (&acc)[RPT][12],
                                            float (&cq)[RPT][6], Stage5<RPT, MODE> (&st)[3], const int p,
                                            const int x0, const int x1, const int lane, const int w,
                                            const unsigned (&row)[RPT], const unsigned (&orow)[RPT],
                                            const unsigned (&grow)[RPT], const unsigned (&brow)[RPT],
                                            const unsigned (&dv)[RingW<NW * RPT, NW>::KI], const int (&d1)[RPT],
                                            const float (&t1v)[RPT], const int d2, const float t2v) {
    using R = RingW<NW * RPT, NW>;
    constexpr int KI = R::KI;
    float *__restrict__ cur = sl[J % NS];              // slot of plane p
    float *__restrict__ nxt = sl[(J + NS - 1) % NS];   // slot of plane p + NS - 1
    const int xlast = g.n[0] + SA_HALO - 1;
    // 1. this wave's loads into the slot of plane p have landed; then every wave's have (the
    // probe read is pinned between memory clobbers so that its wait is not hoisted)
    asm volatile("" ::: "memory");
    {
        const float probe = cur[lane];
        asm volatile("" ::"v"(probe));
    }
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    // 2. plane p + NS - 1 into the slot every wave finished reading before the barrier
    {
        const int soff = (min(p + NS - 1, xlast) + SA_HALO) * rs.plane_bytes;
#pragma unroll
        for (int k = 0; k < KI; ++k) dma16(rs.u, nxt + (w + NW * k) * 256, dv[k], soff);
    }
    // 3. plane p: axis-1/2 terms, own centre, pushes along axis 0
    float lap[RPT], c5[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        // the lane offset is made opaque per plane so that the LDS addresses (one base per
        // 1 KiB window of ds_read2 offsets) are formed here instead of being hoisted out of
        // the march for every slot, which would hold tens of VGPRs
        int li = (w + NW * r + 5) * R::PITCH + lane + 8;
        asm volatile("" : "+v"(li));
        const float c = cur[li];
        float vy = g.w2[1][0] * c, vz = g.w2[2][0] * c;
#pragma unroll
        for (int k = 1; k <= 5; ++k) {
            vy += g.w2[1][k] * (cur[li + k * R::PITCH] + cur[li - k * R::PITCH]);
            vz += g.w2[2][k] * (cur[li + k] + cur[li - k]);
        }
        acc[r][(J + 5) % 12] = g.w2[0][5] * c;  // first contribution to plane p + 5
#pragma unroll
        for (int k = 1; k <= 4; ++k) acc[r][(J + k) % 12] += g.w2[0][k] * c;
        acc[r][J % 12] += g.w2[0][0] * c + (vy + vz);
#pragma unroll
        for (int k = 1; k <= 5; ++k) acc[r][(J + 12 - k) % 12] += g.w2[0][k] * c;
        lap[r] = acc[r][(J + 7) % 12];  // plane p - 5 is complete
        c5[r] = cq[r][(J + 1) % 6];
        cq[r][J % 6] = c;
    }
    // 4. update and store plane p - 5; operands of plane p - 2 into its stage
    Stage5<RPT, MODE> &sx = st[(J + 1) % 3];
    finish7<RPT, MODE, SPONGE>(g, a, rs, p - 5, x0, x1, lap, c5, sx, orow, grow, brow, d1, t1v, d2, t2v);
    load_stage5<RPT, MODE>(st[(J + 1) % 3], rs, row, grow, brow, min(max(p - 2, 0), xlast), a);
}

// Epilogue step: plane p past the march (own column from memory), finishing plane p - 5.
template <int RPT, int NW, int MODE, bool SPONGE, int J>
__device__ __forceinline__ void tail7_step(const Geo &g, const FastArgs &a, const Rsrc &rs, float (&acc)[RPT][12],
                                           float (&cq)[RPT][6], Stage5<RPT, MODE> (&st)[3], const float (&up)[RPT],
                                           const int p, const int x0, const int x1, const unsigned (&row)[RPT],
                                           const unsigned (&orow)[RPT], const unsigned (&grow)[RPT],
                                           const unsigned (&brow)[RPT], const int (&d1)[RPT],
                                           const float (&t1v)[RPT], const int d2, const float t2v) {
    const int xlast = g.n[0] + SA_HALO - 1;
    float lap[RPT], c5[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        // only planes p - 5 .. p - 1 are still open; J = 0 at the first tail step
#pragma unroll
        for (int k = J + 1; k <= 5; ++k) acc[r][(J + 12 - k) % 12] += g.w2[0][k] * up[r];
        lap[r] = acc[r][(J + 7) % 12];
        c5[r] = cq[r][(J + 1) % 6];
    }
    finish7<RPT, MODE, SPONGE>(g, a, rs, p - 5, x0, x1, lap, c5, st[(J + 1) % 3], orow, grow, brow, d1, t1v, d2,
                                   t2v);
    if (J < 2) load_stage5<RPT, MODE>(st[(J + 1) % 3], rs, row, grow, brow, min(p - 2, xlast), a);
}

// The march of one chunk; SPONGE: the layer coefficients are evaluated on every cell.
template <int RPT, int NW, int MODE, bool SPONGE, int NS>
__device__ __forceinline__ void march7(const Geo &g, const FastArgs &a, const Rsrc &rs, float *const (&sl)[6],
                                       const int x0, const int x1,
                                       const int lane, const int w, const unsigned (&row)[RPT],
                                       const unsigned (&orow)[RPT], const unsigned (&grow)[RPT],
                                       const unsigned (&brow)[RPT], const unsigned (&dv)[RingW<NW * RPT, NW>::KI],
                                       const int (&d1)[RPT], const float (&t1v)[RPT], const int d2,
                                       const float t2v) {
    constexpr int KI = RingW<NW * RPT, NW>::KI;
    const int xlast = g.n[0] + SA_HALO - 1;
    // prologue: LDS slots of planes x0 .. x0 + NS - 2; partial sums of planes x0 .. x0 + 4
    // from the own column of planes x0 - 5 .. x0 - 1
#pragma unroll
    for (int p = 0; p < NS - 1; ++p)
#pragma unroll
        for (int k = 0; k < KI; ++k)
            dma16(rs.u, sl[p] + (w + NW * k) * 256, dv[k], (min(x0 + p, xlast) + SA_HALO) * rs.plane_bytes);
    float acc[RPT][12], cq[RPT][6];
    Stage5<RPT, MODE> st[3] = {};  // filled inside the march before their first stored use
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[r][k] = 0.f;
#pragma unroll
        for (int k = 0; k < 6; ++k) cq[r][k] = 0.f;
#pragma unroll
        for (int i = 1; i <= 5; ++i) {
            const float ub = bload(rs.u, row[r], (x0 - i + SA_HALO) * rs.plane_bytes);
#pragma unroll
            for (int j = 0; j + i <= 5; ++j) acc[r][j] += g.w2[0][i + j] * ub;
        }
    }
    // whole 12-plane periods, straight-line (no per-plane branch)
    const int nper = (x1 - x0 + 11) / 12;
    const int pend = x0 + 12 * nper;
    for (int it = 0; it < nper; ++it) {
        const int xb = x0 + 12 * it;
#define SA_STEP7(J)                                                                                              \
    march7_step<RPT, NW, MODE, SPONGE, NS, J>(g, a, rs, sl, acc, cq, st, xb + J, x0, x1, lane, w, row, orow, \
                                          grow, brow, dv, d1, t1v, d2, t2v);
        SA_STEP7(0) SA_STEP7(1) SA_STEP7(2) SA_STEP7(3) SA_STEP7(4) SA_STEP7(5)
        SA_STEP7(6) SA_STEP7(7) SA_STEP7(8) SA_STEP7(9) SA_STEP7(10) SA_STEP7(11)
#undef SA_STEP7
    }
    // epilogue: planes pend - 5 .. pend - 1 from the own column of planes pend .. pend + 4
    float up[5][RPT];
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
        for (int r = 0; r < RPT; ++r) up[j][r] = bload(rs.u, row[r], (min(pend + j, xlast) + SA_HALO) * rs.plane_bytes);
#define SA_TAIL7(J)                                                                                                 \
    tail7_step<RPT, NW, MODE, SPONGE, J>(g, a, rs, acc, cq, st, up[J], pend + J, x0, x1, row, orow, grow, brow, d1, \
                                         t1v, d2, t2v);
    SA_TAIL7(0) SA_TAIL7(1) SA_TAIL7(2) SA_TAIL7(3) SA_TAIL7(4)
#undef SA_TAIL7
    // no direct-to-LDS load may outlive the workgroup's LDS allocation
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
}

// register budget with one row per lane: 80 VGPRs (6 waves per SIMD);
// without it the scheduler may trade occupancy for latency hiding that this
// bandwidth-bound loop does not need
template <int RPT, int NW, int MODE, bool SPONGE, int NS>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(RPT == 1 ? 6 : 1))) void k_fast3d_v7(
    Geo g, FastArgs a) {
    constexpr int TY = NW * RPT;
    using R = RingW<TY, NW>;
    constexpr int KI = R::KI;
    static_assert(NS >= 3 && NS <= 6 && 12 % NS == 0, "ring depth: 3, 4 or 6 slots");
    SA_RING_SLOTS(NS, R::FLOATS)
    const int b = blockIdx.x;
    const int Lg = (b & 7) * a.per_xcd + (b >> 3);
    if (Lg >= a.nblocks) return;
    const int tz = Lg % a.ntile_z;
    const int rest = Lg / a.ntile_z;
    const int ty = rest % a.ntile_y;
    const int ch = rest / a.ntile_y;
    const int z0 = tz * kTZ, y0 = ty * TY;
    const int x0 = ch * a.chunk;
    const int x1 = min(x0 + a.chunk, g.n[0]);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int z = z0 + lane;
    const int ny = g.n[1], nz = g.n[2];
    const unsigned ps1 = (unsigned)g.s1;
    Rsrc rs;
    const long long bytes = g.total * 4;
    rs.u = make_rsrc(a.u, bytes);
    rs.o = make_rsrc(a.o, bytes);
    rs.vp = make_rsrc(a.vp, bytes);
    rs.out = make_rsrc(a.out, bytes);
    rs.plane_bytes = (int)(g.s0 * 4);
    rs.box_plane_bytes = a.box.n[1] * a.box.p2 * 4;
    const long long box_bytes = (long long)a.box.n[0] * rs.box_plane_bytes;
    if (MODE == 1) rs.save = make_rsrc(a.save, box_bytes);
    if (MODE == 2) {
        rs.ps = make_rsrc(a.ps, box_bytes);
        rs.grad = make_rsrc(a.grad, bytes);
        rs.prec = make_rsrc(a.prec, bytes);
    }
    const unsigned zc = (unsigned)(min(z, nz + SA_RADIUS - 1) + SA_LEAD);
    unsigned row[RPT], orow[RPT], grow[RPT], brow[RPT];
    int d1[RPT];
    float t1v[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int yr = y0 + w + NW * r;
        const bool rin = yr < ny && z < nz;
        row[r] = ((unsigned)(min(yr, ny + SA_RADIUS - 1) + SA_HALO) * ps1 + zc) * 4u;
        orow[r] = rin ? row[r] : kOOB;
        const bool bin = rin && yr >= a.box.lo[1] && yr < a.box.lo[1] + a.box.n[1] && z >= a.box.lo[2] &&
                         z < a.box.lo[2] + a.box.n[2];
        brow[r] = bin ? (unsigned)((yr - a.box.lo[1]) * a.box.p2 + a.box.off2 + (z - a.box.lo[2])) * 4u : kOOB;
        grow[r] = bin ? row[r] : kOOB;
        d1[r] = min(yr, ny - 1 - yr);
        t1v[r] = 0.f;
        if (SPONGE && d1[r] >= 0 && d1[r] < g.nabs[1]) t1v[r] = a.axis[a.alen + d1[r]];
    }
    const int d2 = min(z, nz - 1 - z);
    const float t2v = SPONGE && d2 >= 0 && d2 < g.nabs[2] ? a.axis[2 * a.alen + d2] : 0.f;
    unsigned dv[KI];
#pragma unroll
    for (int k = 0; k < KI; ++k) {
        const int c = (w + NW * k) * 64 + lane;
        dv[k] = kOOB;
        if (c < R::NCH) {
            const int rr = c / R::CPR, cc = c % R::CPR;
            const int yy = min(y0 - 5 + rr, ny + SA_HALO - 1);
            const int zz = min(z0 + SA_LEAD - 8 + 4 * cc, (int)ps1 - 4);
            dv[k] = ((unsigned)(yy + SA_HALO) * ps1 + (unsigned)zz) * 4u;
        }
    }
    // tiles whose cells (axes 1/2) and chunk planes all lie inside the layers take the plain
    // update; the others evaluate the layer coefficients on every cell, branch-free
    const bool interior = y0 >= g.nabs[1] && y0 + TY <= ny - g.nabs[1] && z0 >= g.nabs[2] &&
                          z0 + kTZ <= nz - g.nabs[2] && x0 >= g.nabs[0] && x1 <= g.n[0] - g.nabs[0];
    if (SPONGE && !interior)
        march7<RPT, NW, MODE, true, NS>(g, a, rs, sl, x0, x1, lane, w, row, orow, grow, brow, dv, d1, t1v, d2, t2v);
    else
        march7<RPT, NW, MODE, false, NS>(g, a, rs, sl, x0, x1, lane, w, row, orow, grow, brow, dv, d1, t1v, d2, t2v);
}

// ---------------------------------------------------------------------------------------
// Two shots per pass (the batched entry points): the two shots' wavefields are interleaved
// per cell (8 bytes per padded cell, shot-minor), vp and the gradient / preconditioner are
// shared. Structure as version 7 (LDS plane ring by direct-to-LDS loads, u read once, axis-0
// terms pushed into partial sums), with every per-cell quantity a pair. Per update this
// moves 14 B instead of 16 (the vp read is shared), and on gradient steps the 16 B of
// gradient and preconditioner read-modify-write are shared by the two shots.
typedef float f2 __attribute__((ext_vector_type(2)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ f2 bload2(rsrc_t r, unsigned voff, int soff) {
    return __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

__device__ __forceinline__ f2 bload2_nt(rsrc_t r, unsigned voff, int soff) {
    return __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, kNT));
}

__device__ __forceinline__ void bstore2_nt(f2 v, rsrc_t r, unsigned voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2v, v), r, voff, soff, kNT);
}

template <int TY>
struct Ring2 {                               // LDS plane slot of pairs
    static constexpr int ROWS = TY + 10;
    static constexpr int CPR = 40;           // 16-byte chunks per row: cells z0 - 8 .. z0 + 71
    static constexpr int PITCH = 80;         // cells (pairs) per LDS row
    static constexpr int NCH = ROWS * CPR;
    static constexpr int NI = (NCH + 63) / 64;
};

template <int TY, int NW>
struct Ring2W : Ring2<TY> {
    static constexpr int KI = (Ring2<TY>::NI + NW - 1) / NW;
    static constexpr int FLOATS = KI * NW * 256;
};

template <int RPT, int MODE>
struct StageB {
    f2 o[RPT];
    float v[RPT];
    f2 ps[MODE == 2 ? RPT : 1];
    float gr[MODE == 2 ? RPT : 1], pr[MODE == 2 ? RPT : 1];
};

template <int RPT, int MODE>
__device__ __forceinline__ void load_stageB(StageB<RPT, MODE> &st, const Rsrc &rs, const unsigned (&row)[RPT],
                                            const unsigned (&grow)[RPT], const unsigned (&brow)[RPT], int x,
                                            const FastArgs &a) {
    const int sp = (x + SA_HALO) * rs.plane_bytes;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        st.o[r] = bload2_nt(rs.o, 2 * row[r], 2 * sp);
        st.v[r] = bload_nt(rs.vp, row[r], sp);
    }
    if (MODE == 2) {
        const bool xin = (x >= a.box.lo[0]) & (x < a.box.lo[0] + a.box.n[0]);
        const int sbx = xin ? (x - a.box.lo[0]) * rs.box_plane_bytes : 0;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            st.ps[r] = bload2_nt(rs.ps, xin ? 2 * brow[r] : kOOB, 2 * sbx);
            st.gr[r] = bload_nt(rs.grad, grow[r], sp);
            st.pr[r] = bload_nt(rs.prec, grow[r], sp);
        }
    }
}

template <int RPT, int MODE, bool SPONGE>
__device__ __forceinline__ void finishB(const Geo &g, const FastArgs &a, const Rsrc &rs, const int x, const int x0,
                                        const int x1, const f2 (&lap)[RPT], const f2 (&c)[RPT],
                                        const StageB<RPT, MODE> &sx, const unsigned (&orow)[RPT],
                                        const unsigned (&grow)[RPT], const unsigned (&brow)[RPT],
                                        const int (&d1)[RPT], const float (&t1v)[RPT], const int d2, const float t2v) {
    const int nx = g.n[0];
    const int sxo = (x + SA_HALO) * rs.plane_bytes;
    const bool xin = (x >= x0) & (x < x1);
    const bool xin_box = MODE != 0 && (xin & (x >= a.box.lo[0]) & (x < a.box.lo[0] + a.box.n[0]));
    const int sbx = MODE == 1 && xin_box ? (x - a.box.lo[0]) * rs.box_plane_bytes : 0;
    int d0 = 0;
    float t0v = 0.f;
    if (SPONGE) {
        d0 = max(min(x, nx - 1 - x), 0);
        const float tv = a.axis[min(d0, a.alen - 1)];
        t0v = d0 < g.nabs[0] ? tv : 0.f;
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const float v = sx.v[r];
        const float v2 = v * v;
        float A = 2.f, B = -1.f, G = a.dt2 * v2;
        if (SPONGE) {
            float sm = d0 <= min(d1[r], d2) ? t0v : 0.f;
            sm += d1[r] <= min(d0, d2) ? t1v[r] : 0.f;
            sm += d2 <= min(d0, d1[r]) ? t2v : 0.f;
            const float D = (7.f * (sm * v)) * a.dt;
            const float aa = a.dt * v2 * D;
            const float e = a.dt * a.dt * v2 * D * D;
            const float ia = __builtin_amdgcn_rcpf(1.f + aa);  // v_rcp_f32 (1 ulp), no divide sequence
            A = (2.f - e) * ia;
            B = -(1.f - aa) * ia;
            G = (a.dt2 * ia) * v2;
        }
        f2 val;
        val.x = A * c[r].x + B * sx.o[r].x + G * lap[r].x;
        val.y = A * c[r].y + B * sx.o[r].y + G * lap[r].y;
        bstore2_nt(val, rs.out, xin ? 2 * orow[r] : kOOB, 2 * sxo);
        if (MODE == 1) {
            f2 ps;
            ps.x = (val.x - 2.f * c[r].x + sx.o[r].x) * a.idt2;
            ps.y = (val.y - 2.f * c[r].y + sx.o[r].y) * a.idt2;
            bstore2_nt(ps, rs.save, xin_box ? 2 * brow[r] : kOOB, 2 * sbx);
        }
        if (MODE == 2) {
            const f2 psv = sx.ps[r];
            const unsigned go = xin_box ? grow[r] : kOOB;
            bstore_nt((sx.gr[r] + psv.x * c[r].x) + psv.y * c[r].y, rs.grad, go, sxo);
            bstore_nt((sx.pr[r] + psv.x * psv.x) + psv.y * psv.y, rs.prec, go, sxo);
        }
    }
}

template <int RPT, int NW, int MODE, bool SPONGE, bool ISO, int J>
__device__ __forceinline__ void marchB_step(const Geo &g, const FastArgs &a, const Rsrc &rs, float *const (&sl)[6],
                                            f2 (&acc)[RPT][12], f2 (&cq)[RPT][6], StageB<RPT, MODE> (&st)[3],
                                            const int p, const int x0, const int x1, const int lane, const int w,
                                            const unsigned (&row)[RPT], const unsigned (&orow)[RPT],
                                            const unsigned (&grow)[RPT], const unsigned (&brow)[RPT],
                                            const unsigned (&dv)[Ring2W<NW * RPT, NW>::KI], const int (&d1)[RPT],
                                            const float (&t1v)[RPT], const int d2, const float t2v) {
    using R = Ring2W<NW * RPT, NW>;
    constexpr int KI = R::KI;
    float *__restrict__ cur = sl[J % 3];
    float *__restrict__ nxt = sl[(J + 2) % 3];
    const int xlast = g.n[0] + SA_HALO - 1;
    asm volatile("" ::: "memory");
    {
        const float probe = cur[lane];
        asm volatile("" ::"v"(probe));
    }
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    {
        const int soff = 2 * (min(p + 2, xlast) + SA_HALO) * rs.plane_bytes;
#pragma unroll
        for (int k = 0; k < KI; ++k) dma16(rs.u, nxt + (w + NW * k) * 256, dv[k], soff);
    }
    const f2 *__restrict__ pc = reinterpret_cast<const f2 *>(cur);
    f2 lap[RPT], c5[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        int li = (w + NW * r + 5) * R::PITCH + lane + 8;  // in pairs; opaque per plane (see version 5)
        asm volatile("" : "+v"(li));
        const f2 c = pc[li];
        f2 vy, vz;
        if (ISO) {  // one spacing on every axis: the y and z neighbours share the weights of x
            vy = (2.f * g.w2[0][0]) * c;
            vz = f2{0.f, 0.f};
#pragma unroll
            for (int k = 1; k <= 5; ++k) {
                const f2 syz = (pc[li + k * R::PITCH] + pc[li - k * R::PITCH]) + (pc[li + k] + pc[li - k]);
                vy.x += g.w2[0][k] * syz.x;
                vy.y += g.w2[0][k] * syz.y;
            }
        } else {
            vy = g.w2[1][0] * c;
            vz = g.w2[2][0] * c;
#pragma unroll
            for (int k = 1; k <= 5; ++k) {
                const f2 sy = pc[li + k * R::PITCH] + pc[li - k * R::PITCH];
                const f2 sz = pc[li + k] + pc[li - k];
                vy.x += g.w2[1][k] * sy.x;
                vy.y += g.w2[1][k] * sy.y;
                vz.x += g.w2[2][k] * sz.x;
                vz.y += g.w2[2][k] * sz.y;
            }
        }
        acc[r][(J + 5) % 12] = g.w2[0][5] * c;
#pragma unroll
        for (int k = 1; k <= 4; ++k) {
            acc[r][(J + k) % 12].x += g.w2[0][k] * c.x;
            acc[r][(J + k) % 12].y += g.w2[0][k] * c.y;
        }
        acc[r][J % 12].x += g.w2[0][0] * c.x + (vy.x + vz.x);
        acc[r][J % 12].y += g.w2[0][0] * c.y + (vy.y + vz.y);
#pragma unroll
        for (int k = 1; k <= 5; ++k) {
            acc[r][(J + 12 - k) % 12].x += g.w2[0][k] * c.x;
            acc[r][(J + 12 - k) % 12].y += g.w2[0][k] * c.y;
        }
        lap[r] = acc[r][(J + 7) % 12];
        c5[r] = cq[r][(J + 1) % 6];
        cq[r][J % 6] = c;
    }
    finishB<RPT, MODE, SPONGE>(g, a, rs, p - 5, x0, x1, lap, c5, st[(J + 1) % 3], orow, grow, brow, d1, t1v, d2, t2v);
    load_stageB<RPT, MODE>(st[(J + 1) % 3], rs, row, grow, brow, min(max(p - 2, 0), xlast), a);
}

template <int RPT, int NW, int MODE, bool SPONGE, int J>
__device__ __forceinline__ void tailB_step(const Geo &g, const FastArgs &a, const Rsrc &rs, f2 (&acc)[RPT][12],
                                           f2 (&cq)[RPT][6], StageB<RPT, MODE> (&st)[3], const f2 (&up)[RPT],
                                           const int p, const int x0, const int x1, const unsigned (&row)[RPT],
                                           const unsigned (&orow)[RPT], const unsigned (&grow)[RPT],
                                           const unsigned (&brow)[RPT], const int (&d1)[RPT],
                                           const float (&t1v)[RPT], const int d2, const float t2v) {
    const int xlast = g.n[0] + SA_HALO - 1;
    f2 lap[RPT], c5[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int k = J + 1; k <= 5; ++k) {
            acc[r][(J + 12 - k) % 12].x += g.w2[0][k] * up[r].x;
            acc[r][(J + 12 - k) % 12].y += g.w2[0][k] * up[r].y;
        }
        lap[r] = acc[r][(J + 7) % 12];
        c5[r] = cq[r][(J + 1) % 6];
    }
    finishB<RPT, MODE, SPONGE>(g, a, rs, p - 5, x0, x1, lap, c5, st[(J + 1) % 3], orow, grow, brow, d1, t1v, d2, t2v);
    if (J < 2) load_stageB<RPT, MODE>(st[(J + 1) % 3], rs, row, grow, brow, min(p - 2, xlast), a);
}

template <int RPT, int NW, int MODE, bool SPONGE, bool ISO>
__device__ __forceinline__ void marchB(const Geo &g, const FastArgs &a, const Rsrc &rs, float *const (&sl)[6],
                                       const int x0, const int x1, const int lane, const int w,
                                       const unsigned (&row)[RPT], const unsigned (&orow)[RPT],
                                       const unsigned (&grow)[RPT], const unsigned (&brow)[RPT],
                                       const unsigned (&dv)[Ring2W<NW * RPT, NW>::KI], const int (&d1)[RPT],
                                       const float (&t1v)[RPT], const int d2, const float t2v) {
    constexpr int KI = Ring2W<NW * RPT, NW>::KI;
    const int xlast = g.n[0] + SA_HALO - 1;
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int k = 0; k < KI; ++k)
            dma16(rs.u, sl[q] + (w + NW * k) * 256, dv[k], 2 * (min(x0 + q, xlast) + SA_HALO) * rs.plane_bytes);
    f2 acc[RPT][12], cq[RPT][6];
    StageB<RPT, MODE> st[3] = {};
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[r][k] = f2{0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 6; ++k) cq[r][k] = f2{0.f, 0.f};
#pragma unroll
        for (int i = 1; i <= 5; ++i) {
            const f2 ub = bload2(rs.u, 2 * row[r], 2 * (x0 - i + SA_HALO) * rs.plane_bytes);
#pragma unroll
            for (int j = 0; j + i <= 5; ++j) {
                acc[r][j].x += g.w2[0][i + j] * ub.x;
                acc[r][j].y += g.w2[0][i + j] * ub.y;
            }
        }
    }
    const int nper = (x1 - x0 + 11) / 12;
    const int pend = x0 + 12 * nper;
    for (int it = 0; it < nper; ++it) {
        const int xb = x0 + 12 * it;
#define SA_STEPB(J)                                                                                            \
    marchB_step<RPT, NW, MODE, SPONGE, ISO, J>(g, a, rs, sl, acc, cq, st, xb + J, x0, x1, lane, w, row, orow, grow, \
                                          brow, dv, d1, t1v, d2, t2v);
        SA_STEPB(0) SA_STEPB(1) SA_STEPB(2) SA_STEPB(3) SA_STEPB(4) SA_STEPB(5)
        SA_STEPB(6) SA_STEPB(7) SA_STEPB(8) SA_STEPB(9) SA_STEPB(10) SA_STEPB(11)
#undef SA_STEPB
    }
    f2 up[5][RPT];
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
        for (int r = 0; r < RPT; ++r)
            up[j][r] = bload2(rs.u, 2 * row[r], 2 * (min(pend + j, xlast) + SA_HALO) * rs.plane_bytes);
#define SA_TAILB(J)                                                                                                \
    tailB_step<RPT, NW, MODE, SPONGE, J>(g, a, rs, acc, cq, st, up[J], pend + J, x0, x1, row, orow, grow, brow, d1, \
                                         t1v, d2, t2v);
    SA_TAILB(0) SA_TAILB(1) SA_TAILB(2) SA_TAILB(3) SA_TAILB(4)
#undef SA_TAILB
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): no direct-to-LDS load outlives the workgroup
}

// MULTI: one launch advances a.npair shot pairs (the multi-shot schedule). The logical blocks
// of all pairs are ordered tile group major, pair next, tile within the group last. With a
// group of a few tiles the pairs' blocks for one group run close together on one XCD and share
// its vp lines in the L2; with a whole pair per group (the default) the order is pair-major,
// which keeps the neighbouring tiles of one pair together (their halo rows hit in the L2) and
// measured faster (301 against 345 us per pair and step on the 360^3 grid, 16 pairs).
template <int RPT, int NW, int MODE, bool SPONGE, bool ISO, bool MULTI>
__global__ __launch_bounds__(NW * 64) void k_fast3d_b2(Geo g, FastArgs a) {
    constexpr int TY = NW * RPT;
    using R = Ring2W<TY, NW>;
    constexpr int KI = R::KI;
    SA_RING_SLOTS(3, R::FLOATS)
    const int b = blockIdx.x;
    int Lg = (b & 7) * a.per_xcd + (b >> 3);
    const float *pu = a.u, *po = a.o;
    float *pout = a.out;
    if (MULTI) {
        if (Lg >= a.nblocks * a.npair) return;
        const int G = a.group;
        const int gi = Lg / (G * a.npair);
        const int gsz = min(G, a.nblocks - gi * G);  // the last group may hold fewer tiles
        const int r = Lg - gi * G * a.npair;
        const int pair = r / gsz;
        Lg = gi * G + (r - pair * gsz);
        const PairLevels pl = a.pairs[pair];
        pu = a.parity ? pl.lev1 : pl.lev0;
        pout = a.parity ? pl.lev0 : pl.lev1;
        po = pout;
    } else if (Lg >= a.nblocks) {
        return;
    }
    const int tz = Lg % a.ntile_z;
    const int rest = Lg / a.ntile_z;
    const int ty = rest % a.ntile_y;
    const int ch = rest / a.ntile_y;
    const int z0 = tz * kTZ, y0 = ty * TY;
    const int x0 = ch * a.chunk;
    const int x1 = min(x0 + a.chunk, g.n[0]);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int z = z0 + lane;
    const int ny = g.n[1], nz = g.n[2];
    const unsigned ps1 = (unsigned)g.s1;
    Rsrc rs;
    const long long bytes = g.total * 4;
    rs.u = make_rsrc(pu, 2 * bytes);
    rs.o = make_rsrc(po, 2 * bytes);
    rs.vp = make_rsrc(a.vp, bytes);
    rs.out = make_rsrc(pout, 2 * bytes);
    rs.plane_bytes = (int)(g.s0 * 4);  // of a one-shot field; pair fields use twice the offsets
    rs.box_plane_bytes = a.box.n[1] * a.box.p2 * 4;
    const long long box_bytes = (long long)a.box.n[0] * rs.box_plane_bytes;
    if (MODE == 1) rs.save = make_rsrc(a.save, 2 * box_bytes);
    if (MODE == 2) {
        rs.ps = make_rsrc(a.ps, 2 * box_bytes);
        rs.grad = make_rsrc(a.grad, bytes);
        rs.prec = make_rsrc(a.prec, bytes);
    }
    // offsets are in one-shot bytes (pair arrays double them at the use)
    const unsigned zc = (unsigned)(min(z, nz + SA_RADIUS - 1) + SA_LEAD);
    unsigned row[RPT], orow[RPT], grow[RPT], brow[RPT];
    int d1[RPT];
    float t1v[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int yr = y0 + w + NW * r;
        const bool rin = yr < ny && z < nz;
        row[r] = ((unsigned)(min(yr, ny + SA_RADIUS - 1) + SA_HALO) * ps1 + zc) * 4u;
        orow[r] = rin ? row[r] : kOOB / 2;
        const bool bin = rin && yr >= a.box.lo[1] && yr < a.box.lo[1] + a.box.n[1] && z >= a.box.lo[2] &&
                         z < a.box.lo[2] + a.box.n[2];
        brow[r] = bin ? (unsigned)((yr - a.box.lo[1]) * a.box.p2 + a.box.off2 + (z - a.box.lo[2])) * 4u : kOOB / 2;
        grow[r] = bin ? row[r] : kOOB;
        d1[r] = min(yr, ny - 1 - yr);
        t1v[r] = 0.f;
        if (SPONGE && d1[r] >= 0 && d1[r] < g.nabs[1]) t1v[r] = a.axis[a.alen + d1[r]];
    }
    const int d2 = min(z, nz - 1 - z);
    const float t2v = SPONGE && d2 >= 0 && d2 < g.nabs[2] ? a.axis[2 * a.alen + d2] : 0.f;
    // direct-to-LDS chunk offsets (pair bytes): chunk c -> row c / 40, cells 2 (c % 40) .. + 1
    unsigned dv[KI];
#pragma unroll
    for (int k = 0; k < KI; ++k) {
        const int c = (w + NW * k) * 64 + lane;
        dv[k] = kOOB;
        if (c < R::NCH) {
            const int rr = c / R::CPR, cc = c % R::CPR;
            const int yy = min(y0 - 5 + rr, ny + SA_HALO - 1);
            const int zz = min(z0 + SA_LEAD - 8 + 2 * cc, (int)ps1 - 2);
            dv[k] = ((unsigned)(yy + SA_HALO) * ps1 + (unsigned)zz) * 8u;
        }
    }
    const bool interior = y0 >= g.nabs[1] && y0 + TY <= ny - g.nabs[1] && z0 >= g.nabs[2] &&
                          z0 + kTZ <= nz - g.nabs[2] && x0 >= g.nabs[0] && x1 <= g.n[0] - g.nabs[0];
    if (SPONGE && !interior)
        marchB<RPT, NW, MODE, true, ISO>(g, a, rs, sl, x0, x1, lane, w, row, orow, grow, brow, dv, d1, t1v, d2, t2v);
    else
        marchB<RPT, NW, MODE, false, ISO>(g, a, rs, sl, x0, x1, lane, w, row, orow, grow, brow, dv, d1, t1v, d2, t2v);
}

template <int RPT, int MODE>
void launch_rpt(int version, bool sponge, dim3 grid, hipStream_t st, const Geo &g, const FastArgs &a) {
    if (version == 2) {
        if (sponge)
            k_fast3d_v2<RPT, MODE, true, 1><<<grid, 256, 0, st>>>(g, a);
        else
            k_fast3d_v2<RPT, MODE, false, 1><<<grid, 256, 0, st>>>(g, a);
    } else if (version == 3) {
        if (sponge)
            k_fast3d_v2<RPT, MODE, true, 2><<<grid, 256, 0, st>>>(g, a);
        else
            k_fast3d_v2<RPT, MODE, false, 2><<<grid, 256, 0, st>>>(g, a);
    } else if (sponge) {
        k_fast3d<RPT, MODE, true><<<grid, 256, 0, st>>>(g, a);
    } else {
        k_fast3d<RPT, MODE, false><<<grid, 256, 0, st>>>(g, a);
    }
}

template <int RPT, int NW, int MODE, int NS>
void launch_v5(bool sponge, dim3 grid, hipStream_t st, const Geo &g, const FastArgs &a) {
    if (sponge)
        k_fast3d_v5<RPT, NW, MODE, true, NS><<<grid, NW * 64, 0, st>>>(g, a);
    else
        k_fast3d_v5<RPT, NW, MODE, false, NS><<<grid, NW * 64, 0, st>>>(g, a);
}

template <int RPT, int NW, int MODE, int NS>
void launch_v7(bool sponge, dim3 grid, hipStream_t st, const Geo &g, const FastArgs &a) {
    if (sponge)
        k_fast3d_v7<RPT, NW, MODE, true, NS><<<grid, NW * 64, 0, st>>>(g, a);
    else
        k_fast3d_v7<RPT, NW, MODE, false, NS><<<grid, NW * 64, 0, st>>>(g, a);
}

template <int MODE>
int launch(const Fast3D &f, const Geo &g, FastArgs &a, hipStream_t st) {
    int rows = f.tune.rows;
    int version = f.tune.version;
    if (g.total * 4 >= (1ll << 31)) {  // buffer offsets are 32-bit: general-addressing version
        version = 1;
        rows = std::max(rows, 2);
    }
    const int waves = version == 6 || version == 8 ? 8 : kWaves;
    const int TY = waves * rows;
    a.ntile_y = (g.n[1] + TY - 1) / TY;
    a.ntile_z = (g.n[2] + kTZ - 1) / kTZ;
    a.chunk = f.tune.chunk;
    const int nchunk = (g.n[0] + a.chunk - 1) / a.chunk;
    a.nblocks = a.ntile_y * a.ntile_z * nchunk;
    a.per_xcd = (a.nblocks + 7) / 8;
    dim3 grid(a.per_xcd * 8);
    const bool sponge = f.boundary == 0;
    if (version == 5) {
        switch (rows) {
            case 2: launch_v5<2, 4, MODE, 3>(sponge, grid, st, g, a); break;
            case 4: launch_v5<4, 4, MODE, 3>(sponge, grid, st, g, a); break;
            default: return SA_E_ARG;
        }
    } else if (version == 6) {
        switch (rows) {
            case 1: launch_v5<1, 8, MODE, 3>(sponge, grid, st, g, a); break;
            case 2: launch_v5<2, 8, MODE, 3>(sponge, grid, st, g, a); break;
            default: return SA_E_ARG;
        }
    } else if (version == 7) {
        switch (rows) {
            case 2: launch_v7<2, 4, MODE, 3>(sponge, grid, st, g, a); break;
            default: return SA_E_ARG;
        }
    } else if (version == 8) {
        switch (rows) {
            case 1: launch_v7<1, 8, MODE, 3>(sponge, grid, st, g, a); break;
            case 2: launch_v7<2, 8, MODE, 3>(sponge, grid, st, g, a); break;
            default: return SA_E_ARG;
        }
    } else {
        switch (rows) {
            case 2: launch_rpt<2, MODE>(version, sponge, grid, st, g, a); break;
            case 4: launch_rpt<4, MODE>(version, sponge, grid, st, g, a); break;
            case 8: launch_rpt<8, MODE>(version, sponge, grid, st, g, a); break;
            default: return SA_E_ARG;
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : SA_E_HIP;
}

template <int RPT, int NW, int MODE, bool MULTI>
void launch_b2_rpt(bool sponge, dim3 grid, hipStream_t st, const Geo &g, const FastArgs &a) {
    bool iso = true;  // the same weights on every axis (equal spacing)
    for (int k = 0; k < 6; ++k) iso = iso && g.w2[1][k] == g.w2[0][k] && g.w2[2][k] == g.w2[0][k];
    if (sponge && iso)
        k_fast3d_b2<RPT, NW, MODE, true, true, MULTI><<<grid, NW * 64, 0, st>>>(g, a);
    else if (sponge)
        k_fast3d_b2<RPT, NW, MODE, true, false, MULTI><<<grid, NW * 64, 0, st>>>(g, a);
    else if (iso)
        k_fast3d_b2<RPT, NW, MODE, false, true, MULTI><<<grid, NW * 64, 0, st>>>(g, a);
    else
        k_fast3d_b2<RPT, NW, MODE, false, false, MULTI><<<grid, NW * 64, 0, st>>>(g, a);
}

// npair > 0: the multi-shot form (a.pairs, a.parity set by the caller), MODE 0 only
template <int MODE>
int launch_b2(const Fast3D &f, const Geo &g, FastArgs &a, hipStream_t st, int npair = 0) {
    if (g.total * 8 >= (1ll << 31)) return SA_E_UNSUPPORTED;  // pair buffers need 32-bit offsets
    const int rows = f.tune2.rows, waves = f.tune2.version;
    const int TY = waves * rows;
    a.ntile_y = (g.n[1] + TY - 1) / TY;
    a.ntile_z = (g.n[2] + kTZ - 1) / kTZ;
    a.chunk = f.tune2.chunk;
    const int nchunk = (g.n[0] + a.chunk - 1) / a.chunk;
    a.nblocks = a.ntile_y * a.ntile_z * nchunk;
    a.npair = npair;
    // tiles per group of the multi-shot mapping (STRIDE_AMD_MULTI_GROUP overrides it for tuning)
    a.group = f.multi_group == 1   ? a.ntile_z * a.ntile_y   // one chunk slab
              : f.multi_group == 2 ? a.nblocks               // a whole pair: pair-major order
                                   : a.ntile_z;
    if (const char *e = getenv("STRIDE_AMD_MULTI_GROUP"))
        a.group = e[0] == '1' ? a.ntile_z * a.ntile_y : (e[0] == '2' ? a.nblocks : a.ntile_z);
    const long long total_blocks = (long long)a.nblocks * std::max(npair, 1);
    if (total_blocks >= (1ll << 31) / 8) return SA_E_UNSUPPORTED;
    a.per_xcd = (int)((total_blocks + 7) / 8);
    dim3 grid(a.per_xcd * 8);
    const bool sponge = f.boundary == 0;
    const bool multi = npair > 0;
    if (multi && MODE != 0) return SA_E_UNSUPPORTED;
#define SA_B2(R, W)                                               \
    if constexpr (MODE == 0) {                                    \
        if (multi)                                                \
            launch_b2_rpt<R, W, 0, true>(sponge, grid, st, g, a); \
        else                                                      \
            launch_b2_rpt<R, W, 0, false>(sponge, grid, st, g, a); \
    } else {                                                      \
        launch_b2_rpt<R, W, MODE, false>(sponge, grid, st, g, a); \
    }
    if (waves == 8 && rows == 1) {
        SA_B2(1, 8)
    } else if (waves == 8 && rows == 2) {
        SA_B2(2, 8)
    } else if (waves == 4 && rows == 2) {
        SA_B2(2, 4)
    } else {
        return SA_E_ARG;
    }
#undef SA_B2
    return hipGetLastError() == hipSuccess ? 0 : SA_E_HIP;
}

FastArgs base_args(const Fast3D &f, const float *u, const float *o, float *out) {
    FastArgs a = {};
    a.u = u;
    a.o = o;
    a.out = out;
    a.vp = f.vp;
    a.tables = f.tables;
    a.tlen = f.tlen;
    a.axis = f.axis;
    a.alen = f.alen;
    a.dt = f.dt;
    a.dt2 = f.dt * f.dt;
    return a;
}

}  // namespace

bool fast3d_enabled() {
    const char *e = getenv("STRIDE_AMD_GENERAL_ONLY");
    return !(e && e[0] == '1');
}

int fast3d_set_tune(Fast3D &f, int rows, int chunk, int version) {
    if (chunk < 1) return SA_E_ARG;
    const bool ok = version >= 1 && version <= 3   ? (rows == 2 || rows == 4 || rows == 8)
                    : version == 5               ? (rows == 2 || rows == 4)
                    : version == 7               ? rows == 2
                    : version == 6 || version == 8 ? (rows == 1 || rows == 2)
                                                 : false;
    if (!ok) return SA_E_ARG;
    f.tune.rows = rows;
    f.tune.chunk = chunk;
    f.tune.version = version;
    return 0;
}

int fast3d_prepare(Fast3D &f, const Geo &g, const float *vp, const float *tables, int tlen, int boundary, float dt,
                   hipStream_t stream) {
    f.boundary = boundary;
    f.vp = vp;
    f.dt = dt;
    f.axis = tables;
    f.alen = tlen;
    if (boundary == 0) {
        // corner table: layer sum for distances to the edge (d0, d1, d2) clipped at the widest
        // layer M, built with the rule of the general kernels (axis d contributes its value at
        // distance l < nabs[d] when every other axis is at least l from its edges)
        std::vector<float> t1(3 * (size_t)tlen);
        if (hipMemcpy(t1.data(), tables, t1.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            return SA_E_HIP;
        const int M = std::max(g.nabs[0], std::max(g.nabs[1], g.nabs[2]));
        const int m0 = M + 1, m1 = M + 1, m2 = M + 1;
        std::vector<float> corner((size_t)m0 * m1 * m2, 0.f);
        for (int d0 = 0; d0 < m0; ++d0)
            for (int d1 = 0; d1 < m1; ++d1)
                for (int d2 = 0; d2 < m2; ++d2) {
                    const int dist[3] = {d0, d1, d2};
                    float s = 0.f;
                    for (int d = 0; d < 3; ++d) {
                        if (dist[d] >= g.nabs[d]) continue;
                        int other = 0x7fffffff;
                        for (int e = 0; e < 3; ++e)
                            if (e != d) other = std::min(other, dist[e]);
                        if (other >= dist[d]) s += t1[(size_t)d * tlen + dist[d]];
                    }
                    corner[((size_t)d0 * m1 + d1) * m2 + d2] = s;
                }
        if (f.corner) (void)hipFree(f.corner);
        f.corner = nullptr;
        if (hipMalloc((void **)&f.corner, corner.size() * sizeof(float)) != hipSuccess) return SA_E_NOMEM;
        if (hipMemcpyAsync(f.corner, corner.data(), corner.size() * sizeof(float), hipMemcpyHostToDevice, stream) !=
            hipSuccess)
            return SA_E_HIP;
        if (hipStreamSynchronize(stream) != hipSuccess) return SA_E_HIP;
    }
    f.tables = f.corner;
    f.tlen = std::max(g.nabs[0], std::max(g.nabs[1], g.nabs[2]));  // clip distance of the table
    if (!f.prepared) {
        // default: version 7, tile 64 x 8 cells, 24-plane chunks (the autotuner refines it)
        f.tune.rows = 2;
        f.tune.chunk = 24;
        f.tune.version = 7;
        const char *e = getenv("STRIDE_AMD_TILE");
        if (e) {
            int r = 0, c = 0, v = 2;
            if (sscanf(e, "%d,%d,%d", &r, &c, &v) >= 2) fast3d_set_tune(f, r, c, v);
        }
        f.prepared = true;
    }
    return 0;
}

int fast3d_step(Fast3D &f, const Geo &g, const float *u, const float *o, float *out, float *save, const Box &box,
                float idt2, hipStream_t stream) {
    FastArgs a = base_args(f, u, o, out);
    if (save) {
        a.save = save;
        a.box = box;
        a.idt2 = idt2;
        return launch<1>(f, g, a, stream);
    }
    return launch<0>(f, g, a, stream);
}

int fast3d_step_adjoint(Fast3D &f, const Geo &g, const float *u, const float *o, float *out, const float *ps,
                        const Box &box, float *grad, float *prec, hipStream_t stream) {
    FastArgs a = base_args(f, u, o, out);
    if (ps && grad && prec) {
        a.ps = ps;
        a.box = box;
        a.grad = grad;
        a.prec = prec;
        return launch<2>(f, g, a, stream);
    }
    return launch<0>(f, g, a, stream);
}

int fast3d_step_b2(Fast3D &f, const Geo &g, const float *u, const float *o, float *out, float *save, const Box &box,
                   float idt2, hipStream_t stream) {
    FastArgs a = base_args(f, u, o, out);
    if (save) {
        a.save = save;
        a.box = box;
        a.idt2 = idt2;
        return launch_b2<1>(f, g, a, stream);
    }
    return launch_b2<0>(f, g, a, stream);
}

int fast3d_step_adjoint_b2(Fast3D &f, const Geo &g, const float *u, const float *o, float *out, const float *ps,
                           const Box &box, float *grad, float *prec, hipStream_t stream) {
    FastArgs a = base_args(f, u, o, out);
    if (ps && grad && prec) {
        a.ps = ps;
        a.box = box;
        a.grad = grad;
        a.prec = prec;
        return launch_b2<2>(f, g, a, stream);
    }
    return launch_b2<0>(f, g, a, stream);
}

int fast3d_step_multi(Fast3D &f, const Geo &g, const PairLevels *pairs, int npair, int parity, hipStream_t stream) {
    if (npair < 1 || !pairs) return SA_E_ARG;
    FastArgs a = base_args(f, nullptr, nullptr, nullptr);
    a.pairs = pairs;
    a.parity = parity & 1;
    return launch_b2<0>(f, g, a, stream, npair);
}

int fast3d_set_tune_b2(Fast3D &f, int rows, int chunk, int waves) {
    if (chunk < 1 || !((waves == 8 && (rows == 1 || rows == 2)) || (waves == 4 && rows == 2))) return SA_E_ARG;
    f.tune2.rows = rows;
    f.tune2.chunk = chunk;
    f.tune2.version = waves;
    return 0;
}

void fast3d_free(Fast3D &f) {
    if (f.corner) (void)hipFree(f.corner);
    f.corner = nullptr;
    f.tables = nullptr;
    f.prepared = false;
}
