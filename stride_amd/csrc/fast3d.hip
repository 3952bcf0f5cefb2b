// Specialised 3D wavefield update (see fast3d.h).
//
// One pass per time step over the padded 3D layout: a workgroup of NW waves owns a tile of 64
// (axis 2, the contiguous axis) x NW*RPT (axis 1) cells and marches along axis 0 over a chunk
// of planes. Each plane tile with its 5-cell halo is brought into a ring of three LDS slots by
// direct-to-LDS buffer loads two planes ahead; the own cell of each plane feeds a register ring
// of axis-0 partial sums, so u is read from memory once. Loads and stores are unconditional
// buffer instructions whose range check turns out-of-domain lanes into zero reads and dropped
// writes. The sponge coefficients are evaluated on the fly from vp and the per-axis layer
// tables, only in tiles that touch the layer.
//
// Per cell and step the one-shot kernel reads u (4 B), the other level (4 B), vp (4 B) and
// writes the new level (4 B); the two-shot kernel moves 28 B per cell for two updates (vp is
// shared). Saves of p.dt2 (forward) and the vp gradient / preconditioner accumulation
// (adjoint) are fused into the same pass. Buffer offsets are 32-bit: fields must stay below
// 2 GiB (fast3d_supported); larger grids take the general kernels.

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
    const float *axis;  // per-axis layer tables [3][alen]
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
    // attenuation (one-shot kernel): form (0 none, 1 power 0, 2 power 2, 3 power None), alpha
    // in Np, and for power 2 the Laplacian of the other level (read) and of u (written: the next
    // step's other level)
    int att;
    const float *alpha, *lapo;
    float *lapn;
    // CFS-PML (att bit 4): Cpml and the u1 auxiliary field of each axis (read on the cells of
    // the absorbing layers only), the layer widths and the length of axis 0
    const float *cpml, *pu1[3];
    int nabs[3], n0;
    // OT4 (att bit 8: the second pass, whose staged field is w = m K u): the own u and K u, and
    // c4 = dt^2 / 12; MODE 4 (the first pass) writes K u into out and w into wout
    const float *uown, *kuin;
    float *wout;
    float c4;
    int raw_save;  // MODE 1: save the level itself (raw snapshots) instead of p.dt2
};

// MODE 0: plain step, 1: fused save of p.dt2, 2: fused gradient accumulation
// Raw buffer access: one 128-bit resource per array, a constant per-lane byte offset and the
// plane offset in a scalar register; arrays below 2 GiB.
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

struct Rsrc {
    rsrc_t u, o, vp, out, save, ps, grad, prec, alpha, lapo, lapn, cp, q1, q2, q3, uo, ku, wo;
    int plane_bytes, box_plane_bytes;
};


// The plane tile with its halo is staged in a ring of three LDS slots by direct-to-LDS buffer
// loads (16 B per lane, no VGPR staging and no LDS store instructions), issued two planes
// ahead; the own-column operands (the other level, vp and, when fused, the gradient operands)
// are register loads issued three planes ahead. Synchronisation per plane: each wave waits for
// its own direct-to-LDS loads into the slot of the plane (a dummy read of the slot makes the
// compiler emit exactly that wait), then one workgroup barrier; the slot written next is the
// one every wave finished reading before that barrier.
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

// ATT: attenuation form in bits 0-1 (0 none, 1 power 0, 2 power 2, 3 power None), bit 2 the
// CFS-PML term, bit 3 the OT4 second pass. alpha is an own-cell operand with attenuation, and
// with power 2 also the other level's Laplacian; with the PML, Cpml and the three u1 fields (zero
// off the layers); with OT4, u and K u (the staged field is then w = m K u).
template <int RPT, int MODE, int ATT>
struct Stage5 {
    static constexpr int AF = ATT & 3;
    static constexpr bool PM = (ATT & 4) != 0;
    static constexpr bool O4 = (ATT & 8) != 0;
    float uo[O4 ? RPT : 1], ku[O4 ? RPT : 1];
    float o[RPT], v[RPT];
    float ps[MODE == 2 ? RPT : 1], gr[MODE == 2 ? RPT : 1], pr[MODE == 2 ? RPT : 1];
    float al[AF ? RPT : 1], lo[AF == 2 ? RPT : 1];
    float cp[PM ? RPT : 1], q1[PM ? RPT : 1], q2[PM ? RPT : 1], q3[PM ? RPT : 1];
};

template <int RPT, int MODE, int ATT>
__device__ __forceinline__ void load_stage5(Stage5<RPT, MODE, ATT> &st, const Rsrc &rs, const unsigned (&row)[RPT],
                                            const unsigned (&grow)[RPT], const unsigned (&brow)[RPT], int x,
                                            const FastArgs &a, const int (&d1)[RPT], const int d2) {
    using S5 = Stage5<RPT, MODE, ATT>;
    const int sp = (x + SA_HALO) * rs.plane_bytes;
    if (MODE == 3) return;  // the Laplacian alone: no own operands
    if (MODE == 4) {        // OT4 first pass: vp only (m = vp^2)
#pragma unroll
        for (int r = 0; r < RPT; ++r) st.v[r] = bload_nt(rs.vp, row[r], sp);
        return;
    }
    // PML operands: read on the layer cells (plane, row or column within a layer of its axis;
    // planes past the domain too, whose results are never stored), zero elsewhere
    const bool xlay = S5::PM && min(x, a.n0 - 1 - x) < a.nabs[0];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        st.o[r] = bload_nt(rs.o, row[r], sp);
        st.v[r] = bload_nt(rs.vp, row[r], sp);
        if (S5::AF) st.al[r] = bload_nt(rs.alpha, row[r], sp);
        if (S5::AF == 2) st.lo[r] = bload_nt(rs.lapo, row[r], sp);
        if (S5::O4) {
            st.uo[r] = bload_nt(rs.uo, row[r], sp);
            st.ku[r] = bload_nt(rs.ku, row[r], sp);
        }
        if (S5::PM) {
            const unsigned po = (xlay | (d1[r] < a.nabs[1]) | (d2 < a.nabs[2])) ? row[r] : kOOB;
            st.cp[r] = bload_nt(rs.cp, po, sp);
            st.q1[r] = bload_nt(rs.q1, po, sp);
            st.q2[r] = bload_nt(rs.q2, po, sp);
            st.q3[r] = bload_nt(rs.q3, po, sp);
        }
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


// The one-shot march (versions 7: 4 waves x 2 rows, and 8: 8 waves x 1 or 2 rows). u is read
// from memory once: a register window of u along axis 0 filled by separate loads would fetch
// u from HBM twice (with hundreds of tiles in flight per XCD the L2 cannot hold the planes
// between the two reads). Instead the own cell of
// plane p is read from the LDS tile when plane p arrives, and its axis-0 second-derivative
// contributions are pushed into a ring of partial sums for planes p - 5 .. p + 5; plane
// p - 5, now complete, is updated and stored. The first five planes of a chunk get their
// contributions from own-column loads before the march, the last five are finished after it
// with own-column loads of the five planes past the chunk.
template <int RPT, int MODE, bool SPONGE, int ATT>
__device__ __forceinline__ void finish7(const Geo &g, const FastArgs &a, const Rsrc &rs, const int x, const int x0,
                                        const int x1, const float (&lap)[RPT], const float (&c)[RPT],
                                        const Stage5<RPT, MODE, ATT> &sx, const unsigned (&orow)[RPT],
                                        const unsigned (&grow)[RPT], const unsigned (&brow)[RPT],
                                        const int (&d1)[RPT], const float (&t1v)[RPT], const int d2, const float t2v) {
    const int nx = g.n[0];
    const int sxo = (x + SA_HALO) * rs.plane_bytes;
    // planes outside the chunk (before it, or past its end in the last 12-plane period) are
    // computed but not stored
    const bool xin = (x >= x0) & (x < x1);
    // bitwise, not short-circuit: a branch here splits the plane's schedule
    if (MODE == 3) {  // the centred Laplacian of u
#pragma unroll
        for (int r = 0; r < RPT; ++r) bstore_nt(lap[r], rs.out, xin ? orow[r] : kOOB, sxo);
        return;
    }
    if (MODE == 4) {  // OT4 first pass: K u, and w = m K u with m = vp^2 (the general K(m K u) operand)
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const float v = sx.v[r];
            bstore_nt(lap[r], rs.out, xin ? orow[r] : kOOB, sxo);
            bstore_nt((v * v) * lap[r], rs.wo, xin ? orow[r] : kOOB, sxo);
        }
        return;
    }
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
        float A = 2.f, B = -1.f, G = a.dt2 * v2, cq = a.dt2;
        constexpr int AF = ATT & 3;
        if (SPONGE || AF == 1 || AF == 3) {
            // the coefficients of the general update (k_coefficients): power 0 adds dt alpha vp
            // to the damping, power None scales the step by W = alpha
            float aa = 0.f, e = 0.f;
            if (AF == 1) aa += a.dt * sx.al[r] * v;
            if (SPONGE) {
                // branch-free layer rule (axis d contributes its table value at distance d_d from
                // the edge when no other axis is closer to its own edge); with sm = 0 every
                // expression reduces exactly to the interior values
                float sm = d0 <= min(d1[r], d2) ? t0v : 0.f;
                sm += d1[r] <= min(d0, d2) ? t1v[r] : 0.f;
                sm += d2 <= min(d0, d1[r]) ? t2v : 0.f;
                const float D = (7.f * (sm * v)) * a.dt;
                aa += a.dt * v2 * D;
                e = a.dt * a.dt * v2 * D * D;
            }
            const float W = AF == 3 ? sx.al[r] : 1.f;
            const float ia = __builtin_amdgcn_rcpf(1.f + aa);  // v_rcp_f32 (1 ulp), no divide sequence
            A = W * (2.f - e) * ia;
            B = -W * (1.f - aa) * ia;
            cq = W * a.dt2 * ia;
            G = cq * v2;
        }
        // OT4: the staged field is w, so lap = K w = K(m K u); own u and K u are operands
        constexpr bool O4 = (ATT & 8) != 0;
        const float cc = O4 ? sx.uo[r] : c[r];
        float acc = O4 ? G * (sx.ku[r] + a.c4 * lap[r]) : G * lap[r];
        if (AF == 2) {
            // power-2 attenuation (the general update's Cbeta term): Cbeta (lap u - lap other),
            // Cbeta = Cq 2 alpha vp^3 / dt; the Laplacian of u is the next step's lap other
            const float cb = cq * 2.f * sx.al[r] * v2 * v / a.dt;
            acc += cb * (lap[r] - sx.lo[r]);
            bstore_nt(lap[r], rs.lapn, xin ? orow[r] : kOOB, sxo);
        }
        // CFS-PML: - Cpml sum_d u1_d (the general update's order: the axes summed from zero)
        if (ATT & 4) acc -= sx.cp[r] * ((sx.q1[r] + sx.q2[r]) + sx.q3[r]);
        const float val = A * cc + B * sx.o[r] + acc;
        bstore_nt(val, rs.out, xin ? orow[r] : kOOB, sxo);
        if (MODE == 1)
            bstore_nt(a.raw_save ? cc : (val - 2.f * cc + sx.o[r]) * a.idt2, rs.save, xin_box ? brow[r] : kOOB, sbx);
        if (MODE == 2) {
            const float psv = sx.ps[r];
            const unsigned go = xin_box ? grow[r] : kOOB;
            bstore_nt(sx.gr[r] + psv * cc, rs.grad, go, sxo);
            bstore_nt(sx.pr[r] + psv * psv, rs.prec, go, sxo);
        }
    }
}

// One step of the march: plane p (LDS slot J % 3, J = p - x0 mod 12). Partial-sum slot
// (J + k) % 12 holds plane p + k (k = -5 .. 5); centre slot J % 6 holds u of plane p; operand
// stage (J + 1) % 3 holds plane p - 5.
template <int RPT, int NW, int MODE, bool SPONGE, int ATT, int NS, int J>
__device__ __forceinline__ void march7_step(const Geo &g, const FastArgs &a, const Rsrc &rs, float *const (&sl)[6],
                                            float (&acc)[RPT][12],
                                            float (&cq)[RPT][6], Stage5<RPT, MODE, ATT> (&st)[3], const int p,
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
    Stage5<RPT, MODE, ATT> &sx = st[(J + 1) % 3];
    finish7<RPT, MODE, SPONGE, ATT>(g, a, rs, p - 5, x0, x1, lap, c5, sx, orow, grow, brow, d1, t1v, d2, t2v);
    load_stage5<RPT, MODE, ATT>(st[(J + 1) % 3], rs, row, grow, brow, min(max(p - 2, 0), xlast), a, d1, d2);
}

// Epilogue step: plane p past the march (own column from memory), finishing plane p - 5.
template <int RPT, int NW, int MODE, bool SPONGE, int ATT, int J>
__device__ __forceinline__ void tail7_step(const Geo &g, const FastArgs &a, const Rsrc &rs, float (&acc)[RPT][12],
                                           float (&cq)[RPT][6], Stage5<RPT, MODE, ATT> (&st)[3],
                                           const float (&up)[RPT],
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
    finish7<RPT, MODE, SPONGE, ATT>(g, a, rs, p - 5, x0, x1, lap, c5, st[(J + 1) % 3], orow, grow, brow, d1, t1v,
                                     d2, t2v);
    if (J < 2) load_stage5<RPT, MODE, ATT>(st[(J + 1) % 3], rs, row, grow, brow, min(p - 2, xlast), a, d1, d2);
}

// The march of one chunk; SPONGE: the layer coefficients are evaluated on every cell.
template <int RPT, int NW, int MODE, bool SPONGE, int ATT, int NS>
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
    Stage5<RPT, MODE, ATT> st[3] = {};  // filled inside the march before their first stored use
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
    march7_step<RPT, NW, MODE, SPONGE, ATT, NS, J>(g, a, rs, sl, acc, cq, st, xb + J, x0, x1, lane, w, row, orow, \
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
    tail7_step<RPT, NW, MODE, SPONGE, ATT, J>(g, a, rs, acc, cq, st, up[J], pend + J, x0, x1, row, orow, grow, brow, \
                                               d1, t1v, d2, t2v);
    SA_TAIL7(0) SA_TAIL7(1) SA_TAIL7(2) SA_TAIL7(3) SA_TAIL7(4)
#undef SA_TAIL7
    // no direct-to-LDS load may outlive the workgroup's LDS allocation
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
}

// register budget with one row per lane: 80 VGPRs (6 waves per SIMD);
// without it the scheduler may trade occupancy for latency hiding that this
// bandwidth-bound loop does not need
template <int RPT, int NW, int MODE, bool SPONGE, int ATT, int NS>
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
    if (ATT & 3) rs.alpha = make_rsrc(a.alpha, bytes);
    if ((ATT & 3) == 2) {
        rs.lapo = make_rsrc(a.lapo, bytes);
        rs.lapn = make_rsrc(a.lapn, bytes);
    }
    if (MODE == 4) rs.wo = make_rsrc(a.wout, bytes);
    if (ATT & 8) {
        rs.uo = make_rsrc(a.uown, bytes);
        rs.ku = make_rsrc(a.kuin, bytes);
    }
    if (ATT & 4) {
        rs.cp = make_rsrc(a.cpml, bytes);
        rs.q1 = make_rsrc(a.pu1[0], bytes);
        rs.q2 = make_rsrc(a.pu1[1], bytes);
        rs.q3 = make_rsrc(a.pu1[2], bytes);
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
        march7<RPT, NW, MODE, true, ATT, NS>(g, a, rs, sl, x0, x1, lane, w, row, orow, grow, brow, dv, d1, t1v, d2,
                                              t2v);
    else
        march7<RPT, NW, MODE, false, ATT, NS>(g, a, rs, sl, x0, x1, lane, w, row, orow, grow, brow, dv, d1, t1v, d2,
                                               t2v);
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
        int li = (w + NW * r + 5) * R::PITCH + lane + 8;  // in pairs; opaque per plane (as in march7_step)
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

template <int RPT, int NW, int MODE, int ATT, int NS>
void launch_v7_att(bool sponge, dim3 grid, hipStream_t st, const Geo &g, const FastArgs &a) {
    if (sponge)
        k_fast3d_v7<RPT, NW, MODE, true, ATT, NS><<<grid, NW * 64, 0, st>>>(g, a);
    else
        k_fast3d_v7<RPT, NW, MODE, false, ATT, NS><<<grid, NW * 64, 0, st>>>(g, a);
}

// attenuation forms 0 - 3 on every tile
template <int RPT, int NW, int MODE, int NS>
void launch_v7(bool sponge, dim3 grid, hipStream_t st, const Geo &g, const FastArgs &a) {
    if constexpr (MODE >= 3) {
        // the Laplacian-only and OT4 first passes: no own-operand forms
        launch_v7_att<RPT, NW, MODE, 0, NS>(sponge, grid, st, g, a);
    } else {
        switch (a.att) {
            case 1: launch_v7_att<RPT, NW, MODE, 1, NS>(sponge, grid, st, g, a); break;
            case 2: launch_v7_att<RPT, NW, MODE, 2, NS>(sponge, grid, st, g, a); break;
            case 3: launch_v7_att<RPT, NW, MODE, 3, NS>(sponge, grid, st, g, a); break;
            default: launch_v7_att<RPT, NW, MODE, 0, NS>(sponge, grid, st, g, a); break;
        }
    }
}

// the CFS-PML (4 - 7) and OT4 second-pass (8, 9, 11) forms, on the version-7 tile only
template <int MODE>
bool launch_v7_forms(bool sponge, dim3 grid, hipStream_t st, const Geo &g, const FastArgs &a) {
    if constexpr (MODE >= 3) {
        return false;
    } else {
        switch (a.att) {
            case 4: k_fast3d_v7<2, 4, MODE, false, 4, 3><<<grid, 4 * 64, 0, st>>>(g, a); break;
            case 5: k_fast3d_v7<2, 4, MODE, false, 5, 3><<<grid, 4 * 64, 0, st>>>(g, a); break;
            case 6: k_fast3d_v7<2, 4, MODE, false, 6, 3><<<grid, 4 * 64, 0, st>>>(g, a); break;
            case 7: k_fast3d_v7<2, 4, MODE, false, 7, 3><<<grid, 4 * 64, 0, st>>>(g, a); break;
            case 8: launch_v7_att<2, 4, MODE, 8, 3>(sponge, grid, st, g, a); break;
            case 9: launch_v7_att<2, 4, MODE, 9, 3>(sponge, grid, st, g, a); break;
            case 11: launch_v7_att<2, 4, MODE, 11, 3>(sponge, grid, st, g, a); break;
            default: return false;
        }
        return true;
    }
}

template <int MODE>
int launch(const Fast3D &f, const Geo &g, FastArgs &a, hipStream_t st) {
    if (!fast3d_supported(g)) return SA_E_UNSUPPORTED;  // buffer offsets are 32-bit
    // the PML and OT4 forms run on the version-7 tile (4 waves x 2 rows) whatever the tune
    const bool forms = MODE < 3 && a.att >= 4;
    const int rows = forms ? 2 : f.tune.rows;
    const int version = forms ? 7 : f.tune.version;
    const int waves = version == 8 ? 8 : kWaves;
    const int TY = waves * rows;
    a.ntile_y = (g.n[1] + TY - 1) / TY;
    a.ntile_z = (g.n[2] + kTZ - 1) / kTZ;
    a.chunk = f.tune.chunk;
    const int nchunk = (g.n[0] + a.chunk - 1) / a.chunk;
    a.nblocks = a.ntile_y * a.ntile_z * nchunk;
    a.per_xcd = (a.nblocks + 7) / 8;
    dim3 grid(a.per_xcd * 8);
    const bool sponge = f.boundary == 0;
    if (forms) {
        if (!launch_v7_forms<MODE>(sponge, grid, st, g, a)) return SA_E_ARG;
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
        return SA_E_ARG;
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

bool fast3d_supported(const Geo &g) {
    // two-shot pair fields double the offsets; one-shot fields need half of that
    return g.total * 4 < (1ll << 31);
}

int fast3d_set_tune(Fast3D &f, int rows, int chunk, int version) {
    if (chunk < 1) return SA_E_ARG;
    const bool ok = version == 7 ? rows == 2 : version == 8 ? (rows == 1 || rows == 2) : false;
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
    if (!f.prepared) {
        // default: version 7, tile 64 x 8 cells, 24-plane chunks (the autotuner refines it)
        f.tune.rows = 2;
        f.tune.chunk = 24;
        f.tune.version = 7;
        const char *e = getenv("STRIDE_AMD_TILE");
        if (e) {
            int r = 0, c = 0, v = 7;
            if (sscanf(e, "%d,%d,%d", &r, &c, &v) >= 2) fast3d_set_tune(f, r, c, v);
        }
        f.prepared = true;
    }
    return 0;
}

int fast3d_step(Fast3D &f, const Geo &g, const float *u, const float *o, float *out, float *save, const Box &box,
                float idt2, hipStream_t stream, const FastAtt *att, int save_mode) {
    FastArgs a = base_args(f, u, o, out);
    if (att) {
        if (att->power == 2 && (!att->alpha || !att->lapo || !att->lapn)) return SA_E_ARG;
        if (att->power != -2 && !att->alpha) return SA_E_ARG;
        a.att = att->power == 0 ? 1 : att->power == 2 ? 2 : att->power == -1 ? 3 : 0;
        a.alpha = att->alpha;
        a.lapo = att->lapo;
        a.lapn = att->lapn;
        if (att->cpml) {
            if (!att->u1[0] || !att->u1[1] || !att->u1[2] || f.boundary != -1) return SA_E_ARG;
            a.att |= 4;
            a.cpml = att->cpml;
            for (int d = 0; d < 3; ++d) {
                a.pu1[d] = att->u1[d];
                a.nabs[d] = g.nabs[d];
            }
            a.n0 = g.n[0];
        }
        if (att->ot4_u) {
            if (!att->ot4_ku || att->power == 2 || att->cpml) return SA_E_ARG;
            a.att |= 8;
            a.uown = att->ot4_u;
            a.kuin = att->ot4_ku;
            a.c4 = att->c4;
        }
    }
    if (save) {
        if (save_mode != 1 && save_mode != 2) return SA_E_ARG;
        a.save = save;
        a.box = box;
        a.idt2 = idt2;
        a.raw_save = save_mode == 2;
        return launch<1>(f, g, a, stream);
    }
    return launch<0>(f, g, a, stream);
}

int fast3d_step_adjoint(Fast3D &f, const Geo &g, const float *u, const float *o, float *out, const float *ps,
                        const Box &box, float *grad, float *prec, hipStream_t stream, const FastAtt *att) {
    FastArgs a = base_args(f, u, o, out);
    if (att) {
        if (att->power == 2 && (!att->alpha || !att->lapo || !att->lapn)) return SA_E_ARG;
        if (att->power != -2 && !att->alpha) return SA_E_ARG;
        a.att = att->power == 0 ? 1 : att->power == 2 ? 2 : att->power == -1 ? 3 : 0;
        a.alpha = att->alpha;
        a.lapo = att->lapo;
        a.lapn = att->lapn;
        if (att->cpml) {
            if (!att->u1[0] || !att->u1[1] || !att->u1[2] || f.boundary != -1) return SA_E_ARG;
            a.att |= 4;
            a.cpml = att->cpml;
            for (int d = 0; d < 3; ++d) {
                a.pu1[d] = att->u1[d];
                a.nabs[d] = g.nabs[d];
            }
            a.n0 = g.n[0];
        }
        if (att->ot4_u) {
            if (!att->ot4_ku || att->power == 2 || att->cpml) return SA_E_ARG;
            a.att |= 8;
            a.uown = att->ot4_u;
            a.kuin = att->ot4_ku;
            a.c4 = att->c4;
        }
    }
    if (ps && grad && prec) {
        a.ps = ps;
        a.box = box;
        a.grad = grad;
        a.prec = prec;
        return launch<2>(f, g, a, stream);
    }
    return launch<0>(f, g, a, stream);
}

int fast3d_ot4_first(Fast3D &f, const Geo &g, const float *u, float *ku, float *w, hipStream_t stream) {
    if (!u || !ku || !w) return SA_E_ARG;
    FastArgs a = base_args(f, u, nullptr, ku);
    a.wout = w;
    const int boundary = f.boundary;
    f.boundary = -1;  // no layer coefficients in this pass
    const int rc = launch<4>(f, g, a, stream);
    f.boundary = boundary;
    return rc;
}

int fast3d_laplacian(Fast3D &f, const Geo &g, const float *u, float *out, hipStream_t stream) {
    if (!u || !out) return SA_E_ARG;
    FastArgs a = base_args(f, u, nullptr, out);
    const int boundary = f.boundary;
    f.boundary = -1;  // no layer coefficients: the Laplacian alone
    const int rc = launch<3>(f, g, a, stream);
    f.boundary = boundary;
    return rc;
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
    f.prepared = false;
}
