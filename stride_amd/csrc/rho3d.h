// Fused 3D wavefield update for the div(b grad) forms of the operator (constant or variable
// density), OT2, sponge or no boundary, attenuation power 0 / 2 / None / off: one pass per time
// step in place of three staggered-derivative passes, a divergence pass and the update pass.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

struct Rho3DTune {
    int rows = 1;    // rows (axis 1) per thread; the tile is 64 (axis 2) x waves*rows (axis 1)
    int chunk = 60;  // planes (axis 0) per workgroup
    int waves = 8;   // 4 or 8
    bool enabled = true;
};

struct Rho3DStep {
    const float *u, *other;  // current level and the level it replaces
    float *out;              // new level (may alias other)
    const float *b;          // buoyancy (variable density) or null (constant density: b = 1)
    const float *A, *B, *G;  // per-point update coefficients
    float *save;             // optional fused save into `box` (save_mode 1: p.dt2, 2: raw u)
    int save_mode;
    Box box;
    float idt2;
    int konly = 0;  // out = K u only (A, B, G, other and save unused)
    float *kout = nullptr;  // update without attenuation power 2: K u of every node also stored here
    // power-2 attenuation: out += Cbeta (lap u - lap other) with the centred SO10 Laplacians of
    // the other level and of u read from lapother and lapu
    const float *Cbeta = nullptr, *lapother = nullptr, *lapu = nullptr;
    // two shots per pass (update form only: no konly, kout or power-2 term): u, other and out
    // hold the two shots interleaved per cell (8 bytes per padded cell, shot-minor); b and the
    // coefficients are shared; `save` takes shot 0's save and save1 shot 1's (same box layout)
    int nshot = 1;
    float *save1 = nullptr;
    // OT4 (one shot, no konly / kout / power 2): pass 1 (ot4 = 1) writes K u into out and
    // w = M K u into wout; pass 2 (ot4 = 2) takes w as u and reads the level from uown and its K u
    // from kuin: out = A uown + B other + G (kuin + c4 K w)
    int ot4 = 0;
    const float *M = nullptr, *uown = nullptr, *kuin = nullptr;
    float *wout = nullptr;
    float c4 = 0.f;
    // CFS-PML term (one shot, OT2, no power 2, no konly / kout): out -= Cpml sum_d u1_d, the u1
    // fields already advanced for this step (read on the absorbing-layer cells only)
    const float *Cpml = nullptr;
    const float *pu1[3] = {nullptr, nullptr, nullptr};
};

// true when the layout of `nshot` interleaved shots fits the kernel's 32-bit buffer offsets
bool rho3d_supported(const Geo &g, int nshot = 1);
int rho3d_step(const Rho3DTune &t, const Geo &g, const Rho3DStep &st, hipStream_t stream);
// rows 1/2 with waves 4/8, or rows 1 with waves 16; chunk >= 1
bool rho3d_valid_tune(int rows, int chunk, int waves);

// Density gradient step of the as-written adjoint on the interior, read from the raw snapshots
// (whole-domain box): grad_rho / prec_rho and, when gv / pv are given, grad_vp / prec_vp from the
// second time difference of the snapshots t - f, t, t + f (psm / psp null: zero).
bool rho3d_grad_supported(const Geo &g, const Box &box);
// pa may hold pa_nshot (1 or 2) interleaved shots, of which shot pa_shot is read.
int rho3d_grad(const Geo &g, const Box &box, const float *ps, const float *psm, const float *psp, float fdt2,
               const float *b, const float *pa, float *gv, float *pv, float *gr, float *pr, hipStream_t stream,
               int pa_nshot = 1, int pa_shot = 0);

// Exact-adjoint accumulations of one step over the whole domain (3D): lam = mu / G, gA += lam u,
// gB += lam um, gG += lam Ku and, with gb, gb -= 1/2 sum_a (D+ mu D+ u)(p) + (D+ mu D+ u)(p - s_a)
// (all arrays padded layout, zero halo).
int rho3d_exact_acc(const Geo &g, const float *u, const float *mu, const float *G, const float *um, const float *Ku,
                    float *gA, float *gB, float *gG, float *gb, hipStream_t stream);
