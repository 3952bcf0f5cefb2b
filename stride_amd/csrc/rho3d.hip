// Fused 3D update for the div(b grad) operator (see rho3d.h).
//
// The operator is K u = sum_d D-_d (bbar_d D+_d u): D+_d the SO10 staggered derivative from
// nodes to the half points i + 1/2, bbar_d the mean of the buoyancy at i and i + 1 (1 for
// constant density), D-_d the staggered derivative from half points back to nodes
// (stride/physics/iso_acoustic/devito.py:1646-1666). The step is
//   out = A u + B other + G K u
// with the per-point coefficients A, B, G of the general kernels (sponge and attenuation
// folded in), i.e. the same update as k_stagger x 3 + k_divergence + k_update in one pass.
//
// Work decomposition: a workgroup of NW waves owns a tile of 64 (axis 2) x NW*RPT (axis 1)
// cells and marches along axis 0 over a chunk of planes. For every input plane m:
//  - the plane's u with a 9-cell halo (rows y0-9 .. y0+TY+8, columns z0-12 .. z0+75) and b
//    with a 5-cell halo are staged in LDS (requested two planes ahead into registers);
//  - the in-plane fluxes bbar D+ u along axes 1 and 2 are formed once per half point into LDS
//    tiles, and their D- divergence is added to the node accumulator of plane m;
//  - the own cell of plane m enters a register ring of u (planes m-9 .. m), from which the
//    axis-0 flux of half plane m-5 is formed, scaled by its bbar and pushed into the node
//    accumulators of planes m-9 .. m; node plane m-9 is then complete, is updated (its other
//    level and coefficients were requested one plane earlier) and stored.
// So u, b, the other level and the coefficients are each read once per step (plus the
// 18 planes of lead-in and lead-out per chunk), and the new level is written once.
// Every load is unconditional: cells beyond the padded rows read zero through the buffer
// range check.

#include "rho3d.h"
#include "bufio.h"

namespace {

using namespace bufio;

constexpr int kTZ = 64;

struct RhoArgs {
    const float *u, *o, *b, *A, *B, *G;
    float *out, *save;
    float *save1;  // two shots: shot 1's save (box layout); `save` holds shot 0's
    long long bytes;   // one-shot field
    long long sbytes;  // field of the shots of a pass (u, o, out: NS floats per cell)
    int plane_bytes, splane_bytes;
    int chunk, ntz, nty, ntiles, per_xcd;
    int save_mode;
    float idt2;
    Box box;
    int konly;  // out = K u (no update, no save)
    float *kout;  // update form: K u of every node also stored here (null: not)
    // power-2 attenuation (Cbeta non-null): Cbeta and the centred Laplacians of the other level
    // and of u
    const float *Cbeta, *lapo, *lapu;
    // OT4: pass 1 writes K u (out) and w = m K u (wout, m from M); pass 2 stages w (the u field)
    // and reads the level itself (uown) and its K u (kuin): out = A u + B other + G (K u + c4 K w)
    int ot4;
    const float *M, *uown, *kuin;
    float *wout;
    float c4;
    // CFS-PML term (form 5): Cpml and the u1 field of each axis, read on the cells of the
    // absorbing layers only (layer widths nabs, axis-0 length n0)
    const float *Cpml, *pu1[3];
    int nabs[3], n0;
};

// NS shots per pass: u, o and out hold NS floats per cell (shot-minor); b and the coefficients
// are read once for all shots. The LDS u tile and the flux tiles hold NS floats per cell.
template <int RPT, int NW, bool VARB, int NS>
struct Tile {
    static constexpr int NT = NW * 64, TY = NW * RPT;
    static constexpr int UY = TY + 18, UZ = 88 * NS, UC = UZ / 4;  // u: rows y0-9.., columns z0-12.. (floats)
    static constexpr int BY = TY + 10, BZ = 80, BC = BZ / 4;       // b: rows y0-5.., columns z0-8..
    static constexpr int FY = TY + 9;                              // axis-1 fluxes, half rows y0-5 .. y0+TY+3
    static constexpr int FZN = 73, FZP = 76;                       // axis-2 fluxes, half columns z0-5 .. z0+67
    static constexpr int NLU = (UY * UC + NT - 1) / NT;
    static constexpr int NLB = VARB ? (BY * BC + NT - 1) / NT : 1;
};

template <int RPT, int NW, bool VARB, int NS>
struct State {
    using T = Tile<RPT, NW, VARB, NS>;
    float nd[NS][RPT][10];  // node accumulators (K u); slot of plane p: (p - mbeg) % 10
    float uq[NS][RPT][10];  // own u; slot of plane p
    float hb[RPT][5];       // axis-0 bbar of half plane q; slot (q - mbeg) % 5
    float bprev[RPT];       // own b of the previous plane
    float4 su[2][T::NLU];   // u tiles of the next two planes, staged in registers (slot: plane % 2)
    float4 sb[2][T::NLB];
    float ov[2][NS][RPT], Av[2][RPT], Bv[2][RPT], Gv[2][RPT];  // own operands of the next node plane
    // power-2 attenuation: own Cbeta and the centred Laplacians of the other level and of u of the
    // next node plane
    float Cb[2][RPT], Lo[2][RPT], Lu[2][RPT];
    float Mv[2][RPT], Uo[2][RPT], Kv[2][RPT];  // OT4: m (pass 1); u and K u (pass 2)
    float Cp[2][RPT], Q1[2][RPT], Q2[2][RPT], Q3[2][RPT];  // CFS-PML: Cpml, u1 of each axis
    bool lay[RPT];  // own cell's row or column inside a layer of axis 1 or 2
    unsigned uoff[T::NLU], boff[T::NLB];  // in-plane byte offsets of the tile chunks (kOOB: none)
    int ulds[T::NLU], blds[T::NLB];       // their LDS float offsets (the dump slot: none)
    unsigned own[RPT];                    // own cells' in-plane byte offsets (kOOB outside the domain)
    unsigned owns[RPT];                   // the same in the NS-shot fields (NS > 1)
};

// Workgroup barrier that orders LDS accesses only: a __syncthreads() would also wait for every
// outstanding global load and store (the prefetched planes) before the barrier.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

struct Res {
    rsrc_t u, o, b, A, B, G, out, Cb, Lo, Lu, Ko, M, Wo, Uo, Ki, Cp, Q1, Q2, Q3;
};

// plane pitch in bytes of the NS-shot fields (the one-shot pitch, known equal, for NS = 1)
#define SPB(a) (NS == 1 ? (a).plane_bytes : (a).splane_bytes)

// NS consecutive floats of LDS (8-byte aligned for NS = 2)
template <int NS>
__device__ __forceinline__ void lds_get(const float *p, float (&v)[NS]) {
    if (NS == 1) {
        v[0] = p[0];
    } else {
        const float2 q = *reinterpret_cast<const float2 *>(p);
        v[0] = q.x;
        v[NS - 1] = q.y;
    }
}

template <int NS>
__device__ __forceinline__ void lds_put(float *p, const float (&v)[NS]) {
    if (NS == 1) *p = v[0];
    else *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[NS - 1]);
}

template <int S, int RPT, int NW, bool VARB, int NS>
__device__ __forceinline__ void load_tile(State<RPT, NW, VARB, NS> &st, const Res &rs, unsigned poff, unsigned spoff) {
    using T = Tile<RPT, NW, VARB, NS>;
#pragma unroll
    for (int i = 0; i < T::NLU; ++i) st.su[S][i] = load4(rs.u, st.uoff[i] + spoff);
    if (VARB) {
#pragma unroll
        for (int i = 0; i < T::NLB; ++i) st.sb[S][i] = load4(rs.b, st.boff[i] + poff);
    }
}

template <int S, int RPT, int NW, bool VARB, int FORM, int NS>
__device__ __forceinline__ void load_own(State<RPT, NW, VARB, NS> &st, const Res &rs, const RhoArgs &a, int x,
                                         unsigned poff, unsigned spoff) {
    if (FORM == 1) return;  // out = K u: no own operands
    if (FORM == 3) {        // OT4 pass 1: m
#pragma unroll
        for (int r = 0; r < RPT; ++r) st.Mv[S][r] = load_nt(rs.M, st.own[r] + poff);
        return;
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        if (FORM == 4) {  // OT4 pass 2: the level and its K u
            st.Uo[S][r] = load_nt(rs.Uo, st.own[r] + poff);
            st.Kv[S][r] = load_nt(rs.Ki, st.own[r] + poff);
        }
        if (NS == 1) {
            st.ov[S][0][r] = load_nt(rs.o, st.own[r] + spoff);
        } else {
            const float2 o2 = load2_nt(rs.o, st.owns[r] + spoff);
            st.ov[S][0][r] = o2.x;
            st.ov[S][NS - 1][r] = o2.y;
        }
        st.Av[S][r] = load_nt(rs.A, st.own[r] + poff);
        st.Bv[S][r] = load_nt(rs.B, st.own[r] + poff);
        st.Gv[S][r] = load_nt(rs.G, st.own[r] + poff);
        if (FORM == 2) {
            st.Cb[S][r] = load_nt(rs.Cb, st.own[r] + poff);
            st.Lo[S][r] = load_nt(rs.Lo, st.own[r] + poff);
            st.Lu[S][r] = load_nt(rs.Lu, st.own[r] + poff);
        }
        if (FORM == 5) {
            // node plane x (planes past the domain read too; their results are never stored)
            const bool xl = min(x, a.n0 - 1 - x) < a.nabs[0];
            const unsigned po = ((xl | st.lay[r]) ? st.own[r] : kOOB) + poff;
            st.Cp[S][r] = load_nt(rs.Cp, po);
            st.Q1[S][r] = load_nt(rs.Q1, po);
            st.Q2[S][r] = load_nt(rs.Q2, po);
            st.Q3[S][r] = load_nt(rs.Q3, po);
        }
    }
}

template <int RPT, int NW, bool VARB, bool SAVE, int FORM, int NS, int J>
__device__ __forceinline__ void plane(const Geo &g, const RhoArgs &a, const Res &rs, State<RPT, NW, VARB, NS> &st,
                                      float *sU, float *sB, float *sFy, float *sFz, const int m, const int x0,
                                      const int x1, const int y0, const int z0, const int lane, const int w) {
    using T = Tile<RPT, NW, VARB, NS>;
    // slots: plane m -> J, m - 9 -> (J + 1) % 10; staging slot of plane m: J % 2 (m - mbeg = J mod 10)
    constexpr int S = J % 2;
    lds_barrier();  // every wave is done with the previous plane's LDS tiles
#pragma unroll
    for (int i = 0; i < T::NLU; ++i) *reinterpret_cast<float4 *>(sU + st.ulds[i]) = st.su[S][i];
    if (VARB) {
#pragma unroll
        for (int i = 0; i < T::NLB; ++i) *reinterpret_cast<float4 *>(sB + st.blds[i]) = st.sb[S][i];
    }
    // requests two planes ahead for the tiles, one node plane ahead for the own operands; all
    // unconditional (planes outside the chunk are read but never stored; past the array they
    // read zero), so the wait counts before each use are exact
    load_tile<S>(st, rs, (unsigned)((m + 2) * a.plane_bytes), (unsigned)((m + 2) * SPB(a)));
    const int p = m - 9;
    const bool emit = p >= x0 && p < x1;
    load_own<1 - S, RPT, NW, VARB, FORM, NS>(st, rs, a, p + 1, (unsigned)((p + 1) * a.plane_bytes),
                                             (unsigned)((p + 1) * SPB(a)));
    lds_barrier();  // plane m's tiles are in LDS
    const bool inplane = m >= x0 && m < x1;
    if (inplane) {
        // one index space over both flux tiles, so that every wave gets the same share: items
        // [0, FY*64) are axis-1 fluxes (whole waves, since FY*64 is a multiple of 64), the rest
        // axis-2 fluxes
        for (int e = w * 64 + lane; e < T::FY * kTZ + T::TY * T::FZN; e += T::NT) {
            if (e < T::FY * kTZ) {
                // axis-1 flux at half row q = y0 - 5 + fr, column z0 + lane
                const int fr = e >> 6;
                const float *c = sU + (fr + 4) * T::UZ + (lane + 12) * NS;  // u at row q
                float d[NS] = {};
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    float hi[NS], lo[NS];
                    lds_get<NS>(c + (k + 1) * T::UZ, hi);
                    lds_get<NS>(c - k * T::UZ, lo);
#pragma unroll
                    for (int s = 0; s < NS; ++s) d[s] += g.ws[1][k] * (hi[s] - lo[s]);
                }
                if (VARB) {
                    const float bb = 0.5f * (sB[fr * T::BZ + lane + 8] + sB[(fr + 1) * T::BZ + lane + 8]);
#pragma unroll
                    for (int s = 0; s < NS; ++s) d[s] *= bb;
                }
                lds_put<NS>(sFy + (fr * kTZ + lane) * NS, d);
            } else {
                // axis-2 flux at half column q = z0 - 5 + fz, row y0 + row
                const int ez = e - T::FY * kTZ;
                const int row = ez / T::FZN, fz = ez - row * T::FZN;
                const float *c = sU + (row + 9) * T::UZ + (fz + 7) * NS;  // u at column q
                float d[NS] = {};
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    float hi[NS], lo[NS];
                    lds_get<NS>(c + (k + 1) * NS, hi);
                    lds_get<NS>(c - k * NS, lo);
#pragma unroll
                    for (int s = 0; s < NS; ++s) d[s] += g.ws[2][k] * (hi[s] - lo[s]);
                }
                if (VARB) {
                    const float bb = 0.5f * (sB[(row + 5) * T::BZ + fz + 3] + sB[(row + 5) * T::BZ + fz + 4]);
#pragma unroll
                    for (int s = 0; s < NS; ++s) d[s] *= bb;
                }
                lds_put<NS>(sFz + (row * T::FZP + fz) * NS, d);
            }
        }
    }
    // axis 0: own cell of plane m into the u ring (planes m-9 .. m)
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int ry = w * RPT + r;
        float uc[NS];
        lds_get<NS>(sU + (ry + 9) * T::UZ + (lane + 12) * NS, uc);
#pragma unroll
        for (int s = 0; s < NS; ++s) st.uq[s][r][J] = uc[s];
        if (VARB) {
            const float bc = sB[(ry + 5) * T::BZ + lane + 8];
            st.hb[r][(J + 4) % 5] = 0.5f * (st.bprev[r] + bc);  // half plane m - 1
            st.bprev[r] = bc;
        }
    }
    if (inplane) {
        lds_barrier();  // fluxes are in LDS
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int ry = w * RPT + r;
            const float *fy = sFy + ((ry + 5) * kTZ + lane) * NS;  // flux at half row y
            const float *fz = sFz + (ry * T::FZP + lane + 5) * NS;  // flux at half column z
            float d[NS] = {};
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                float hi[NS], lo[NS];
                lds_get<NS>(fy + k * kTZ * NS, hi);
                lds_get<NS>(fy - (k + 1) * kTZ * NS, lo);
#pragma unroll
                for (int s = 0; s < NS; ++s) d[s] += g.ws[1][k] * (hi[s] - lo[s]);
            }
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                float hi[NS], lo[NS];
                lds_get<NS>(fz + k * NS, hi);
                lds_get<NS>(fz - (k + 1) * NS, lo);
#pragma unroll
                for (int s = 0; s < NS; ++s) d[s] += g.ws[2][k] * (hi[s] - lo[s]);
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) st.nd[s][r][J] += d[s];
        }
    }
    // axis-0 flux at half plane m - 5 (u of planes m-9 .. m) goes to the nodes m-9 .. m
#pragma unroll
    for (int s = 0; s < NS; ++s) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            float f = 0.f;
#pragma unroll
            for (int k = 0; k < 5; ++k)
                f += g.ws[0][k] * (st.uq[s][r][(J + 6 + k) % 10] - st.uq[s][r][(J + 15 - k) % 10]);
            if (VARB) f *= st.hb[r][J % 5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                st.nd[s][r][(J + 5 + 10 - k) % 10] += g.ws[0][k] * f;  // node m - 5 - k
                st.nd[s][r][(J + 6 + k) % 10] -= g.ws[0][k] * f;       // node m - 4 + k
            }
        }
    }
    // node plane m - 9 is complete (stored only inside the chunk)
    {
        const unsigned po = (unsigned)(p * a.plane_bytes), spo = (unsigned)(p * SPB(a));
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            float out[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                // OT4 pass 2: the staged field is w, so nd = K w; the level is an operand
                const float u = FORM == 4 ? st.Uo[S][r] : st.uq[s][r][(J + 1) % 10];
                const float kn = st.nd[s][r][(J + 1) % 10];
                float acc = FORM == 4 ? st.Gv[S][r] * (st.Kv[S][r] + a.c4 * kn) : st.Gv[S][r] * kn;
                if (FORM == 2) acc += st.Cb[S][r] * (st.Lu[S][r] - st.Lo[S][r]);  // power 2: Cbeta (lap u - lap other)
                if (FORM == 5) acc -= st.Cp[S][r] * ((st.Q1[S][r] + st.Q2[S][r]) + st.Q3[S][r]);  // CFS-PML
                out[s] = (FORM == 1 || FORM == 3) ? kn : st.Av[S][r] * u + st.Bv[S][r] * st.ov[S][s][r] + acc;
            }
            if (NS == 1) store_nt(out[0], rs.out, emit ? st.own[r] + spo : kOOB);
            else store2_nt(out[0], out[NS - 1], rs.out, emit ? st.owns[r] + spo : kOOB);
            if (FORM == 0 && NS == 1 && a.kout) store_nt(st.nd[0][r][(J + 1) % 10], rs.Ko, emit ? st.own[r] + po : kOOB);
            if (FORM == 3) store_nt(st.Mv[S][r] * out[0], rs.Wo, emit ? st.own[r] + po : kOOB);  // w = m K u
            if (SAVE && emit) {
                const int y = y0 + w * RPT + r, z = z0 + lane;
                if (sa_inbox(a.box, p, y, z)) {
                    const long long bi = sa_boxidx(a.box, p, y, z);
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        const float u = FORM == 4 ? st.Uo[S][r] : st.uq[s][r][(J + 1) % 10];
                        (s == 0 ? a.save : a.save1)[bi] =
                            a.save_mode == 1 ? (out[s] - 2.f * u + st.ov[S][s][r]) * a.idt2 : u;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) st.nd[s][r][(J + 1) % 10] = 0.f;
    }
}

template <int RPT, int NW, bool VARB, bool SAVE, int FORM, int NS>
__global__ __launch_bounds__(NW * 64) void k_rho3d(Geo g, RhoArgs a) {
    using T = Tile<RPT, NW, VARB, NS>;
    __shared__ float4 sU4[T::UY * T::UZ / 4 + 1];  // + a dump slot for threads without a chunk
    __shared__ float4 sB4[VARB ? T::BY * T::BZ / 4 + 1 : 1];
    __shared__ float2 sFy2[T::FY * kTZ * NS / 2 + 1];
    __shared__ float2 sFz2[T::TY * T::FZP * NS / 2 + 1];
    float *sU = reinterpret_cast<float *>(sU4), *sB = reinterpret_cast<float *>(sB4);
    float *sFy = reinterpret_cast<float *>(sFy2), *sFz = reinterpret_cast<float *>(sFz2);

    // XCD-aware tile order: hardware block b runs on XCD b % 8; each XCD takes a contiguous
    // run of tiles (axis 2 fastest, then axis 1, then chunks) so halo re-reads hit its L2
    const int L = (blockIdx.x & 7) * a.per_xcd + (blockIdx.x >> 3);
    if (L >= a.ntiles) return;
    const int tz = L % a.ntz, rest = L / a.ntz;
    const int ty = rest % a.nty, ch = rest / a.nty;
    const int z0 = tz * kTZ, y0 = ty * T::TY, x0 = ch * a.chunk;
    const int x1 = min(x0 + a.chunk, g.n[0]);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;

    Res rs;
    rs.u = make(a.u, NS == 1 ? a.bytes : a.sbytes);
    rs.o = make(a.o, NS == 1 ? a.bytes : a.sbytes);
    rs.b = make(VARB ? a.b : a.A, a.bytes);
    rs.A = make(a.A, a.bytes);
    rs.B = make(a.B, a.bytes);
    rs.G = make(a.G, a.bytes);
    rs.out = make(a.out, NS == 1 ? a.bytes : a.sbytes);
    rs.Ko = make(a.kout ? a.kout : a.out, a.kout ? a.bytes : 0);
    if (FORM == 5) {
        rs.Cp = make(a.Cpml, a.bytes);
        rs.Q1 = make(a.pu1[0], a.bytes);
        rs.Q2 = make(a.pu1[1], a.bytes);
        rs.Q3 = make(a.pu1[2], a.bytes);
    }
    if (FORM == 3) {
        rs.M = make(a.M, a.bytes);
        rs.Wo = make(a.wout, a.bytes);
    }
    if (FORM == 4) {
        rs.Uo = make(a.uown, a.bytes);
        rs.Ki = make(a.kuin, a.bytes);
    }
    if (FORM == 2) {
        rs.Cb = make(a.Cbeta, a.bytes);
        rs.Lo = make(a.lapo, a.bytes);
        rs.Lu = make(a.lapu, a.bytes);
    }

    State<RPT, NW, VARB, NS> st;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int k = 0; k < 10; ++k) {
#pragma unroll
            for (int s = 0; s < NS; ++s) st.nd[s][r][k] = st.uq[s][r][k] = 0.f;
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) st.hb[r][k] = 0.f;
        st.bprev[r] = 0.f;
        const int y = y0 + w * RPT + r, z = z0 + lane;
        const bool in = y < g.n[1] && z < g.n[2];
        const long long cell = g.base + (long long)y * g.s1 + z;
        st.own[r] = in ? (unsigned)(cell * 4) : kOOB;
        st.lay[r] = min(y, g.n[1] - 1 - y) < g.nabs[1] || min(z, g.n[2] - 1 - z) < g.nabs[2];
        if (NS > 1) st.owns[r] = in ? (unsigned)(cell * 4 * NS) : kOOB;
    }
    // rows past the padded plane (y > n1 + 9) would read the next plane: they read zero. A u
    // chunk is 4 floats: 4 / NS cells of NS shots.
#pragma unroll
    for (int i = 0; i < T::NLU; ++i) {
        const int e = t + T::NT * i;
        const int row = e / T::UC, c = e - row * T::UC;
        const int y = y0 - 9 + row, z = z0 - 12 + 4 * c / NS;
        const bool ok = e < T::UY * T::UC;
        st.ulds[i] = ok ? row * T::UZ + 4 * c : T::UY * T::UZ;
        st.uoff[i] = ok && y <= g.n[1] + 9 ? (unsigned)((g.base + (long long)y * g.s1 + z) * 4 * NS) : kOOB;
    }
#pragma unroll
    for (int i = 0; i < T::NLB; ++i) {
        const int e = t + T::NT * i;
        const int row = e / T::BC, c = e - row * T::BC;
        const int y = y0 - 5 + row, z = z0 - 8 + 4 * c;
        const bool ok = VARB && e < T::BY * T::BC;
        st.blds[i] = ok ? row * T::BZ + 4 * c : T::BY * T::BZ;
        st.boff[i] = ok && y <= g.n[1] + 9 ? (unsigned)((g.base + (long long)y * g.s1 + z) * 4) : kOOB;
    }

    const int mbeg = x0 - 9;
    load_tile<0>(st, rs, (unsigned)(mbeg * a.plane_bytes), (unsigned)(mbeg * SPB(a)));
    load_tile<1>(st, rs, (unsigned)((mbeg + 1) * a.plane_bytes), (unsigned)((mbeg + 1) * SPB(a)));
    // whole periods of 10 planes (compile-time ring slots); planes past x1 + 8 store nothing
    const int nper = (x1 - x0 + 18 + 9) / 10;
    for (int q = 0; q < nper; ++q) {
        const int m = mbeg + 10 * q;
#define SA_RHO_PLANE(J) \
    plane<RPT, NW, VARB, SAVE, FORM, NS, J>(g, a, rs, st, sU, sB, sFy, sFz, m + J, x0, x1, y0, z0, lane, w)
        SA_RHO_PLANE(0);
        SA_RHO_PLANE(1);
        SA_RHO_PLANE(2);
        SA_RHO_PLANE(3);
        SA_RHO_PLANE(4);
        SA_RHO_PLANE(5);
        SA_RHO_PLANE(6);
        SA_RHO_PLANE(7);
        SA_RHO_PLANE(8);
        SA_RHO_PLANE(9);
#undef SA_RHO_PLANE
    }
}

template <int RPT, int NW>
int launch(const Geo &g, RhoArgs a, bool varb, bool save, int nshot, hipStream_t stream) {
    constexpr int TY = RPT * NW;
    a.nty = (g.n[1] + TY - 1) / TY;
    a.ntiles = a.ntz * a.nty * ((g.n[0] + a.chunk - 1) / a.chunk);
    a.per_xcd = (a.ntiles + 7) / 8;
    const dim3 grid(8 * a.per_xcd), block(NW * 64);
    if (nshot == 2) {
        // two shots: the update form only (no K u output, no power-2 term)
        if (a.konly || a.kout || a.Cbeta) return -1;
        if (varb && save) k_rho3d<RPT, NW, true, true, 0, 2><<<grid, block, 0, stream>>>(g, a);
        else if (varb) k_rho3d<RPT, NW, true, false, 0, 2><<<grid, block, 0, stream>>>(g, a);
        else if (save) k_rho3d<RPT, NW, false, true, 0, 2><<<grid, block, 0, stream>>>(g, a);
        else k_rho3d<RPT, NW, false, false, 0, 2><<<grid, block, 0, stream>>>(g, a);
    } else if (a.Cpml && varb && save) k_rho3d<RPT, NW, true, true, 5, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.Cpml && varb) k_rho3d<RPT, NW, true, false, 5, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.Cpml && save) k_rho3d<RPT, NW, false, true, 5, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.Cpml) k_rho3d<RPT, NW, false, false, 5, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.ot4 == 1 && varb) k_rho3d<RPT, NW, true, false, 3, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.ot4 == 1) k_rho3d<RPT, NW, false, false, 3, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.ot4 == 2 && varb && save) k_rho3d<RPT, NW, true, true, 4, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.ot4 == 2 && varb) k_rho3d<RPT, NW, true, false, 4, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.ot4 == 2 && save) k_rho3d<RPT, NW, false, true, 4, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.ot4 == 2) k_rho3d<RPT, NW, false, false, 4, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.konly && varb) k_rho3d<RPT, NW, true, false, 1, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.konly) k_rho3d<RPT, NW, false, false, 1, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.Cbeta && varb && save) k_rho3d<RPT, NW, true, true, 2, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.Cbeta && varb) k_rho3d<RPT, NW, true, false, 2, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.Cbeta && save) k_rho3d<RPT, NW, false, true, 2, 1><<<grid, block, 0, stream>>>(g, a);
    else if (a.Cbeta) k_rho3d<RPT, NW, false, false, 2, 1><<<grid, block, 0, stream>>>(g, a);
    else if (varb && save) k_rho3d<RPT, NW, true, true, 0, 1><<<grid, block, 0, stream>>>(g, a);
    else if (varb) k_rho3d<RPT, NW, true, false, 0, 1><<<grid, block, 0, stream>>>(g, a);
    else if (save) k_rho3d<RPT, NW, false, true, 0, 1><<<grid, block, 0, stream>>>(g, a);
    else k_rho3d<RPT, NW, false, false, 0, 1><<<grid, block, 0, stream>>>(g, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}


// ------------------------------------------------------------------------------------------
// Density gradient step (as-written adjoint, every f-th step): on the interior,
//   grad_rho += f pa,  prec_rho += f^2,  f = -sum_d D-_d(b) D-_d(p) - b lap(p)
// (stride/physics/iso_acoustic/devito.py:1115-1150) with p the raw saved wavefield, and
//   grad_vp += pdt pa,  prec_vp += pdt^2,  pdt = (p[t+f] - 2 p[t] + p[t-f]) / (f dt)^2
// (common/devito.py:541-550), the two read from the snapshots directly (whole-domain box
// layout), so no unpacking pass. Same march as the update: the plane tiles of p and b with a
// 5-cell halo in LDS, register rings of the own column of p and b along axis 0 (planes p-5 ..
// p+5) and of the in-plane partial terms; node plane m-5 is finished when plane m arrives.

struct GradArgs {
    const float *ps, *psm, *psp;  // snapshots t, t - f, t + f (psm / psp null: zero)
    const float *b, *pa;          // buoyancy, adjoint level (padded layout)
    float *gv, *pv, *gr, *pr;     // accumulators (padded layout); gv / pv null: no vp gradient
    long long bytes, sbytes;
    int plane_bytes, splane_bytes, srow;  // srow: snapshot row pitch in floats
    int pa_nshot, pa_shot;                // pa holds pa_nshot interleaved shots; this is shot pa_shot
    float fdt2;
    int lo[3], hi[3];  // interior [lo, hi)
    int chunk, ntz, nty, ntiles, per_xcd;
};

template <int NW>
struct GTile {
    static constexpr int NT = NW * 64, TY = NW;
    static constexpr int LY = TY + 10, LZ = 80, LC = LZ / 4;  // rows y0-5 .., columns z0-8 ..
    static constexpr int NL = (LY * LC + NT - 1) / NT;
};

template <int NW>
struct GState {
    using T = GTile<NW>;
    float pr_[12], br_[12], ir_[12], lr_[12];  // own p, own b, in-plane products, in-plane Laplacian
    float4 sp[2][T::NL], sb[2][T::NL];
    unsigned poff[T::NL], boff[T::NL];
    int lds[T::NL];
    unsigned own, sown;  // own cell: padded-layout and snapshot byte offsets within a plane
    unsigned aown;       // own cell's byte offset in pa (its shot of the interleaved shots)
    float pa[2], pm[2], pp[2], gv[2], pv[2], gr[2], prr[2];
};

struct GRes {
    rsrc_t ps, psm, psp, b, pa, gv, pv, gr, pr;
};

template <int S, int NW>
__device__ __forceinline__ void gload_tile(GState<NW> &st, const GRes &rs, const GradArgs &a, int m) {
    using T = GTile<NW>;
    const unsigned sp = (unsigned)(m * a.splane_bytes), pp = (unsigned)(m * a.plane_bytes);
#pragma unroll
    for (int i = 0; i < T::NL; ++i) {
        st.sp[S][i] = load4(rs.ps, st.poff[i] + sp);
        st.sb[S][i] = load4(rs.b, st.boff[i] + pp);
    }
}

template <int S, int NW>
__device__ __forceinline__ void gload_own(GState<NW> &st, const GRes &rs, const GradArgs &a, int p) {
    const unsigned po = st.own + (unsigned)(p * a.plane_bytes), so = st.sown + (unsigned)(p * a.splane_bytes);
    st.pa[S] = load_nt(rs.pa, st.aown + (unsigned)(p * a.plane_bytes * a.pa_nshot));
    st.pm[S] = load_nt(rs.psm, so);
    st.pp[S] = load_nt(rs.psp, so);
    st.gr[S] = load_nt(rs.gr, po);
    st.prr[S] = load_nt(rs.pr, po);
    st.gv[S] = load_nt(rs.gv, po);
    st.pv[S] = load_nt(rs.pv, po);
}

template <int NW, int J>
__device__ __forceinline__ void gplane(const Geo &g, const GradArgs &a, const GRes &rs, GState<NW> &st, float *sP,
                                       float *sB, const int m, const int x0, const int x1, const int lane,
                                       const int w) {
    using T = GTile<NW>;
    constexpr int S = J % 2;
    lds_barrier();
#pragma unroll
    for (int i = 0; i < T::NL; ++i) {
        *reinterpret_cast<float4 *>(sP + st.lds[i]) = st.sp[S][i];
        *reinterpret_cast<float4 *>(sB + st.lds[i]) = st.sb[S][i];
    }
    gload_tile<S>(st, rs, a, m + 2);
    const int p = m - 5;
    gload_own<1 - S>(st, rs, a, p + 1);
    lds_barrier();
    const int c = (w + 5) * T::LZ + lane + 8;  // own cell in the tiles
    st.pr_[J] = sP[c];
    st.br_[J] = sB[c];
    if (m >= x0 && m < x1) {
        float prod = 0.f, lap = 0.f;
        // axis 1 (rows)
        {
            float db = 0.f, dp = 0.f, l = g.w2[1][0] * sP[c];
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                db += g.ws[1][r] * (sB[c + r * T::LZ] - sB[c - (r + 1) * T::LZ]);
                dp += g.ws[1][r] * (sP[c + r * T::LZ] - sP[c - (r + 1) * T::LZ]);
            }
#pragma unroll
            for (int r = 1; r <= 5; ++r) l += g.w2[1][r] * (sP[c + r * T::LZ] + sP[c - r * T::LZ]);
            prod += db * dp;
            lap += l;
        }
        // axis 2 (columns)
        {
            float db = 0.f, dp = 0.f, l = g.w2[2][0] * sP[c];
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                db += g.ws[2][r] * (sB[c + r] - sB[c - (r + 1)]);
                dp += g.ws[2][r] * (sP[c + r] - sP[c - (r + 1)]);
            }
#pragma unroll
            for (int r = 1; r <= 5; ++r) l += g.w2[2][r] * (sP[c + r] + sP[c - r]);
            prod += db * dp;
            lap += l;
        }
        st.ir_[J] = prod;
        st.lr_[J] = lap;
    }
    // node plane p = m - 5 (ring slot (J + 7) % 12): axis 0 from the rings of planes p-5 .. p+5
    {
        constexpr int C = (J + 7) % 12;
        float db = 0.f, dp = 0.f, l = g.w2[0][0] * st.pr_[C];
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            db += g.ws[0][r] * (st.br_[(J + 7 + r) % 12] - st.br_[(J + 6 - r) % 12]);
            dp += g.ws[0][r] * (st.pr_[(J + 7 + r) % 12] - st.pr_[(J + 6 - r) % 12]);
        }
#pragma unroll
        for (int r = 1; r <= 5; ++r) l += g.w2[0][r] * (st.pr_[(J + 7 + r) % 12] + st.pr_[(J + 7 - r) % 12]);
        const float f = -(st.ir_[C] + db * dp) - st.br_[C] * (st.lr_[C] + l);
        const bool emit = p >= x0 && p < x1;
        const unsigned po = emit ? st.own + (unsigned)(p * a.plane_bytes) : kOOB;
        const float pav = st.pa[S];
        store_nt(st.gr[S] + f * pav, rs.gr, po);
        store_nt(st.prr[S] + f * f, rs.pr, po);
        if (a.gv) {
            const float pdt = (st.pp[S] - 2.f * st.pr_[C] + st.pm[S]) * a.fdt2;
            store_nt(st.gv[S] + pdt * pav, rs.gv, po);
            store_nt(st.pv[S] + pdt * pdt, rs.pv, po);
        }
    }
}

template <int NW>
__global__ __launch_bounds__(NW * 64) void k_grad_density(Geo g, GradArgs a) {
    using T = GTile<NW>;
    __shared__ float4 sP4[T::LY * T::LZ / 4 + 1];  // + a dump slot for threads without a chunk
    __shared__ float4 sB4[T::LY * T::LZ / 4 + 1];
    float *sP = reinterpret_cast<float *>(sP4), *sB = reinterpret_cast<float *>(sB4);
    const int L = (blockIdx.x & 7) * a.per_xcd + (blockIdx.x >> 3);
    if (L >= a.ntiles) return;
    const int tz = L % a.ntz, rest = L / a.ntz;
    const int ty = rest % a.nty, ch = rest / a.nty;
    const int z0 = a.lo[2] + tz * kTZ, y0 = a.lo[1] + ty * T::TY, x0 = a.lo[0] + ch * a.chunk;
    const int x1 = min(x0 + a.chunk, a.hi[0]);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    GRes rs;
    rs.ps = make(a.ps, a.sbytes);
    rs.psm = make(a.psm ? a.psm : a.ps, a.psm ? a.sbytes : 0);  // null: every read is zero
    rs.psp = make(a.psp ? a.psp : a.ps, a.psp ? a.sbytes : 0);
    rs.b = make(a.b, a.bytes);
    rs.pa = make(a.pa, a.bytes * a.pa_nshot);
    rs.gr = make(a.gr, a.bytes);
    rs.pr = make(a.pr, a.bytes);
    rs.gv = make(a.gv ? a.gv : a.gr, a.gv ? a.bytes : 0);
    rs.pv = make(a.pv ? a.pv : a.pr, a.pv ? a.bytes : 0);
    GState<NW> st;
#pragma unroll
    for (int k = 0; k < 12; ++k) st.pr_[k] = st.br_[k] = st.ir_[k] = st.lr_[k] = 0.f;
    {
        const int y = y0 + w, z = z0 + lane;
        const bool ok = y < a.hi[1] && z < a.hi[2];
        const long long cell = g.base + (long long)y * g.s1 + z;
        st.own = ok ? (unsigned)(cell * 4) : kOOB;
        st.aown = ok ? (unsigned)((cell * a.pa_nshot + a.pa_shot) * 4) : kOOB;
        st.sown = ok ? (unsigned)(((long long)y * a.srow + z) * 4) : kOOB;
    }
#pragma unroll
    for (int i = 0; i < T::NL; ++i) {
        const int e = t + T::NT * i;
        const int row = e / T::LC, c = e - row * T::LC;
        const int y = y0 - 5 + row, z = z0 - 8 + 4 * c;
        const bool ok = e < T::LY * T::LC && y < g.n[1];
        st.lds[i] = e < T::LY * T::LC ? row * T::LZ + 4 * c : T::LY * T::LZ;
        st.poff[i] = ok ? (unsigned)(((long long)y * a.srow + z) * 4) : kOOB;
        st.boff[i] = ok ? (unsigned)((g.base + (long long)y * g.s1 + z) * 4) : kOOB;
    }
    const int mbeg = x0 - 5;
    gload_tile<0>(st, rs, a, mbeg);
    gload_tile<1>(st, rs, a, mbeg + 1);
    gload_own<0>(st, rs, a, mbeg - 5);
    const int nper = (x1 - x0 + 10 + 11) / 12;
    for (int q = 0; q < nper; ++q) {
        const int m = mbeg + 12 * q;
#define SA_GRAD_PLANE(J) gplane<NW, J>(g, a, rs, st, sP, sB, m + J, x0, x1, lane, w)
        SA_GRAD_PLANE(0);
        SA_GRAD_PLANE(1);
        SA_GRAD_PLANE(2);
        SA_GRAD_PLANE(3);
        SA_GRAD_PLANE(4);
        SA_GRAD_PLANE(5);
        SA_GRAD_PLANE(6);
        SA_GRAD_PLANE(7);
        SA_GRAD_PLANE(8);
        SA_GRAD_PLANE(9);
        SA_GRAD_PLANE(10);
        SA_GRAD_PLANE(11);
#undef SA_GRAD_PLANE
    }
}

// ------------------------------------------------------------------------------------------
// Exact-adjoint accumulations of one step (every node of the domain), in one march:
//   lam = mu / G,  gA += lam u,  gB += lam um,  gG += lam Ku,
//   gb -= 1/2 sum_a (D+ mu D+ u)(p) + (D+ mu D+ u)(p - s_a)   (d/db through the half-point buoyancy)
// with u = u_t and mu = mu_{t+1} (padded layout, zero halo). The plane tiles of u and mu with a
// 5-cell halo sit in LDS for the axis-1/2 staggered derivatives; register rings of the own column
// of u and mu (planes p-5 .. p+5) give the axis-0 ones and the in-plane products of each node
// wait in a ring until the node's axis-0 term is complete (plane p + 5).

struct XAccArgs {
    const float *u, *mu, *G, *um, *Ku;
    float *gA, *gB, *gG, *gb;
    long long bytes;
    int plane_bytes;
    int chunk, ntz, nty, ntiles, per_xcd;
};

template <int NW>
struct XState {
    using T = GTile<NW>;
    float ur[12], mr[12], ir[12];  // own u, own mu, in-plane products of plane p (slot (p - mbeg) % 12)
    float4 su[2][T::NL], sm[2][T::NL];
    unsigned off[T::NL];
    int lds[T::NL];
    unsigned own;
    float G[2], um[2], Ku[2], gA[2], gB[2], gG[2], gb[2];
};

struct XRes {
    rsrc_t u, mu, G, um, Ku, gA, gB, gG, gb;
};

template <int S, int NW>
__device__ __forceinline__ void xload_tile(XState<NW> &st, const XRes &rs, const XAccArgs &a, int m) {
    using T = GTile<NW>;
    const unsigned pp = (unsigned)(m * a.plane_bytes);
#pragma unroll
    for (int i = 0; i < T::NL; ++i) {
        st.su[S][i] = load4(rs.u, st.off[i] + pp);
        st.sm[S][i] = load4(rs.mu, st.off[i] + pp);
    }
}

template <int S, int NW>
__device__ __forceinline__ void xload_own(XState<NW> &st, const XRes &rs, const XAccArgs &a, int p) {
    const unsigned po = st.own + (unsigned)(p * a.plane_bytes);
    st.G[S] = load_nt(rs.G, po);
    st.um[S] = load_nt(rs.um, po);
    st.Ku[S] = load_nt(rs.Ku, po);
    st.gA[S] = load_nt(rs.gA, po);
    st.gB[S] = load_nt(rs.gB, po);
    st.gG[S] = load_nt(rs.gG, po);
    if (a.gb) st.gb[S] = load_nt(rs.gb, po);
}

template <int NW, int J>
__device__ __forceinline__ void xplane(const Geo &g, const XAccArgs &a, const XRes &rs, XState<NW> &st, float *sU,
                                       float *sM, const int m, const int x0, const int x1, const int lane,
                                       const int w) {
    using T = GTile<NW>;
    constexpr int S = J % 2;
    lds_barrier();
#pragma unroll
    for (int i = 0; i < T::NL; ++i) {
        *reinterpret_cast<float4 *>(sU + st.lds[i]) = st.su[S][i];
        *reinterpret_cast<float4 *>(sM + st.lds[i]) = st.sm[S][i];
    }
    xload_tile<S>(st, rs, a, m + 2);
    const int p = m - 5;
    xload_own<1 - S>(st, rs, a, p + 1);
    lds_barrier();
    const int c = (w + 5) * T::LZ + lane + 8;  // own cell in the tiles
    st.ur[J] = sU[c];
    st.mr[J] = sM[c];
    if (a.gb && m >= x0 && m < x1) {
        // axis 1 (rows) then axis 2 (columns): products at the half points +1/2 and -1/2
        float prod = 0.f;
        {
            float up = 0.f, ul = 0.f, mp = 0.f, ml = 0.f;
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                up += g.ws[1][r] * (sU[c + (r + 1) * T::LZ] - sU[c - r * T::LZ]);
                ul += g.ws[1][r] * (sU[c + r * T::LZ] - sU[c - (r + 1) * T::LZ]);
                mp += g.ws[1][r] * (sM[c + (r + 1) * T::LZ] - sM[c - r * T::LZ]);
                ml += g.ws[1][r] * (sM[c + r * T::LZ] - sM[c - (r + 1) * T::LZ]);
            }
            prod += mp * up + ml * ul;
        }
        {
            float up = 0.f, ul = 0.f, mp = 0.f, ml = 0.f;
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                up += g.ws[2][r] * (sU[c + r + 1] - sU[c - r]);
                ul += g.ws[2][r] * (sU[c + r] - sU[c - (r + 1)]);
                mp += g.ws[2][r] * (sM[c + r + 1] - sM[c - r]);
                ml += g.ws[2][r] * (sM[c + r] - sM[c - (r + 1)]);
            }
            prod += mp * up + ml * ul;
        }
        st.ir[J] = prod;
    }
    // node plane p = m - 5 (ring slot (J + 7) % 12): axis 0 from the rings of planes p-5 .. p+5
    {
        constexpr int C = (J + 7) % 12;
        const bool emit = p >= x0 && p < x1;
        const unsigned po = emit ? st.own + (unsigned)(p * a.plane_bytes) : kOOB;
        const float lam = st.mr[C] / st.G[S];
        store_nt(st.gA[S] + lam * st.ur[C], rs.gA, po);
        store_nt(st.gB[S] + lam * st.um[S], rs.gB, po);
        store_nt(st.gG[S] + lam * st.Ku[S], rs.gG, po);
        if (a.gb) {
            float up = 0.f, ul = 0.f, mp = 0.f, ml = 0.f;
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                up += g.ws[0][r] * (st.ur[(J + 8 + r) % 12] - st.ur[(J + 7 - r) % 12]);
                ul += g.ws[0][r] * (st.ur[(J + 7 + r) % 12] - st.ur[(J + 6 - r) % 12]);
                mp += g.ws[0][r] * (st.mr[(J + 8 + r) % 12] - st.mr[(J + 7 - r) % 12]);
                ml += g.ws[0][r] * (st.mr[(J + 7 + r) % 12] - st.mr[(J + 6 - r) % 12]);
            }
            const float acc = (mp * up + ml * ul) + st.ir[C];
            store_nt(st.gb[S] + -0.5f * acc, rs.gb, po);
        }
    }
}

template <int NW>
__global__ __launch_bounds__(NW * 64) void k_exact_acc_density(Geo g, XAccArgs a) {
    using T = GTile<NW>;
    __shared__ float4 sU4[T::LY * T::LZ / 4 + 1];  // + a dump slot for threads without a chunk
    __shared__ float4 sM4[T::LY * T::LZ / 4 + 1];
    float *sU = reinterpret_cast<float *>(sU4), *sM = reinterpret_cast<float *>(sM4);
    const int L = (blockIdx.x & 7) * a.per_xcd + (blockIdx.x >> 3);
    if (L >= a.ntiles) return;
    const int tz = L % a.ntz, rest = L / a.ntz;
    const int ty = rest % a.nty, ch = rest / a.nty;
    const int z0 = tz * kTZ, y0 = ty * T::TY, x0 = ch * a.chunk;
    const int x1 = min(x0 + a.chunk, g.n[0]);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    XRes rs;
    rs.u = make(a.u, a.bytes);
    rs.mu = make(a.mu, a.bytes);
    rs.G = make(a.G, a.bytes);
    rs.um = make(a.um, a.bytes);
    rs.Ku = make(a.Ku, a.bytes);
    rs.gA = make(a.gA, a.bytes);
    rs.gB = make(a.gB, a.bytes);
    rs.gG = make(a.gG, a.bytes);
    rs.gb = make(a.gb ? a.gb : a.gA, a.gb ? a.bytes : 0);  // null: every read is zero
    XState<NW> st;
#pragma unroll
    for (int k = 0; k < 12; ++k) st.ur[k] = st.mr[k] = st.ir[k] = 0.f;
    {
        const int y = y0 + w, z = z0 + lane;
        st.own = (y < g.n[1] && z < g.n[2]) ? (unsigned)((g.base + (long long)y * g.s1 + z) * 4) : kOOB;
    }
    // tile rows y0-5 .. y0+TY+4, columns z0-8 .. z0+71; rows past the domain read zero (their
    // cells are halo or beyond it)
#pragma unroll
    for (int i = 0; i < T::NL; ++i) {
        const int e = t + T::NT * i;
        const int row = e / T::LC, c = e - row * T::LC;
        const int y = y0 - 5 + row, z = z0 - 8 + 4 * c;
        const bool ok = e < T::LY * T::LC && y < g.n[1];
        st.lds[i] = e < T::LY * T::LC ? row * T::LZ + 4 * c : T::LY * T::LZ;
        st.off[i] = ok ? (unsigned)((g.base + (long long)y * g.s1 + z) * 4) : kOOB;
    }
    const int mbeg = x0 - 5;
    xload_tile<0>(st, rs, a, mbeg);
    xload_tile<1>(st, rs, a, mbeg + 1);
    xload_own<0>(st, rs, a, mbeg - 5);
    const int nper = (x1 - x0 + 10 + 11) / 12;
    for (int q = 0; q < nper; ++q) {
        const int m = mbeg + 12 * q;
#define SA_XACC_PLANE(J) xplane<NW, J>(g, a, rs, st, sU, sM, m + J, x0, x1, lane, w)
        SA_XACC_PLANE(0);
        SA_XACC_PLANE(1);
        SA_XACC_PLANE(2);
        SA_XACC_PLANE(3);
        SA_XACC_PLANE(4);
        SA_XACC_PLANE(5);
        SA_XACC_PLANE(6);
        SA_XACC_PLANE(7);
        SA_XACC_PLANE(8);
        SA_XACC_PLANE(9);
        SA_XACC_PLANE(10);
        SA_XACC_PLANE(11);
#undef SA_XACC_PLANE
    }
}

}  // namespace

bool rho3d_supported(const Geo &g, int nshot) {
    if (!(g.act[0] && g.act[1] && g.act[2])) return false;
    // every byte offset (plus the plane offsets of the lead-in/out planes) stays below kOOB
    return (g.total * 4 + 32 * g.s0 * 4) * nshot < (long long)bufio::kOOB;
}

bool rho3d_valid_tune(int rows, int chunk, int waves) {
    if (chunk < 1) return false;
    return ((waves == 4 || waves == 8) && (rows == 1 || rows == 2)) || (waves == 16 && rows == 1);
}

int rho3d_step(const Rho3DTune &t, const Geo &g, const Rho3DStep &st, hipStream_t stream) {
    RhoArgs a;
    a.u = st.u;
    a.o = st.other;
    a.b = st.b;
    a.A = st.A;
    a.B = st.B;
    a.G = st.G;
    a.out = st.out;
    a.save = st.save;
    a.save1 = st.save1;
    if (st.nshot != 1 && st.nshot != 2) return -1;
    if (st.nshot == 2 && st.save && !st.save1) return -1;
    a.bytes = g.total * 4;
    a.sbytes = a.bytes * st.nshot;
    a.plane_bytes = (int)(g.s0 * 4);
    a.splane_bytes = a.plane_bytes * st.nshot;
    a.chunk = t.chunk;
    a.ntz = (g.n[2] + kTZ - 1) / kTZ;
    a.save_mode = st.save_mode;
    a.idt2 = st.idt2;
    a.box = st.box;
    a.konly = st.konly;
    a.kout = st.kout;
    if (st.kout && (st.konly || st.Cbeta)) return -1;
    a.Cbeta = st.Cbeta;
    a.lapo = st.lapother;
    a.lapu = st.lapu;
    if (st.konly && (st.save || st.Cbeta)) return -1;
    if (st.Cbeta && (!st.lapother || !st.lapu)) return -1;
    a.ot4 = st.ot4;
    a.M = st.M;
    a.uown = st.uown;
    a.kuin = st.kuin;
    a.wout = st.wout;
    a.c4 = st.c4;
    if (st.ot4 < 0 || st.ot4 > 2) return -1;
    if (st.ot4 && (st.nshot != 1 || st.konly || st.kout || st.Cbeta)) return -1;
    if (st.ot4 == 1 && (!st.M || !st.wout || st.save)) return -1;
    if (st.ot4 == 2 && (!st.uown || !st.kuin)) return -1;
    a.Cpml = st.Cpml;
    for (int d = 0; d < 3; ++d) {
        a.pu1[d] = st.pu1[d];
        a.nabs[d] = g.nabs[d];
    }
    a.n0 = g.n[0];
    if (st.Cpml && (st.nshot != 1 || st.konly || st.kout || st.Cbeta || st.ot4 || !st.pu1[0] || !st.pu1[1] ||
                    !st.pu1[2]))
        return -1;
    const bool varb = st.b != nullptr, save = st.save != nullptr;
    const int ns = st.nshot;
    if (t.waves == 4 && t.rows == 1) return launch<1, 4>(g, a, varb, save, ns, stream);
    if (t.waves == 4 && t.rows == 2) return launch<2, 4>(g, a, varb, save, ns, stream);
    if (t.waves == 8 && t.rows == 1) return launch<1, 8>(g, a, varb, save, ns, stream);
    if (t.waves == 8 && t.rows == 2) return launch<2, 8>(g, a, varb, save, ns, stream);
    if (t.waves == 16 && t.rows == 1) return launch<1, 16>(g, a, varb, save, ns, stream);
    return -1;
}

bool rho3d_grad_supported(const Geo &g, const Box &box) {
    if (!rho3d_supported(g)) return false;
    for (int d = 0; d < 3; ++d)
        if (g.nabs[d] < 8 || 2 * g.nabs[d] >= g.n[d] || box.lo[d] != 0 || box.n[d] != g.n[d]) return false;
    const long long sbytes = (long long)box.n[0] * box.n[1] * box.p2 * 4;
    return box.off2 == 0 && sbytes + 32LL * box.n[1] * box.p2 * 4 < (long long)bufio::kOOB;
}

int rho3d_grad(const Geo &g, const Box &box, const float *ps, const float *psm, const float *psp, float fdt2,
               const float *b, const float *pa, float *gv, float *pv, float *gr, float *pr, hipStream_t stream,
               int pa_nshot, int pa_shot) {
    constexpr int NW = 8;
    if (pa_nshot < 1 || pa_nshot > 2 || pa_shot < 0 || pa_shot >= pa_nshot || !rho3d_supported(g, pa_nshot))
        return -1;
    GradArgs a;
    a.pa_nshot = pa_nshot;
    a.pa_shot = pa_shot;
    a.ps = ps;
    a.psm = psm;
    a.psp = psp;
    a.b = b;
    a.pa = pa;
    a.gv = gv;
    a.pv = pv;
    a.gr = gr;
    a.pr = pr;
    a.bytes = g.total * 4;
    a.sbytes = (long long)box.n[0] * box.n[1] * box.p2 * 4;
    a.plane_bytes = (int)(g.s0 * 4);
    a.splane_bytes = box.n[1] * box.p2 * 4;
    a.srow = box.p2;
    a.fdt2 = fdt2;
    for (int d = 0; d < 3; ++d) {
        a.lo[d] = g.nabs[d];
        a.hi[d] = g.n[d] - g.nabs[d];
    }
    a.chunk = 60;  // of 24 / 40 / 60 / 90 on the 360^3 grid, the fastest
    a.ntz = (a.hi[2] - a.lo[2] + kTZ - 1) / kTZ;
    a.nty = (a.hi[1] - a.lo[1] + NW - 1) / NW;
    a.ntiles = a.ntz * a.nty * ((a.hi[0] - a.lo[0] + a.chunk - 1) / a.chunk);
    a.per_xcd = (a.ntiles + 7) / 8;
    k_grad_density<NW><<<dim3(8 * a.per_xcd), dim3(NW * 64), 0, stream>>>(g, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int rho3d_exact_acc(const Geo &g, const float *u, const float *mu, const float *G, const float *um, const float *Ku,
                    float *gA, float *gB, float *gG, float *gb, hipStream_t stream) {
    if (!rho3d_supported(g)) return -1;
    constexpr int NW = 8;
    XAccArgs a;
    a.u = u;
    a.mu = mu;
    a.G = G;
    a.um = um;
    a.Ku = Ku;
    a.gA = gA;
    a.gB = gB;
    a.gG = gG;
    a.gb = gb;
    a.bytes = g.total * 4;
    a.plane_bytes = (int)(g.s0 * 4);
    a.chunk = 60;
    a.ntz = (g.n[2] + kTZ - 1) / kTZ;
    a.nty = (g.n[1] + NW - 1) / NW;
    a.ntiles = a.ntz * a.nty * ((g.n[0] + a.chunk - 1) / a.chunk);
    a.per_xcd = (a.ntiles + 7) / 8;
    k_exact_acc_density<NW><<<dim3(8 * a.per_xcd), dim3(NW * 64), 0, stream>>>(g, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
