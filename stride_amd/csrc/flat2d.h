// Fused 2D wavefield update: OT2, any density form, sponge or CFS-PML, attenuation off / power
// 0 / power None / power 2, in one launch per step instead of the general kernels' separate
// passes.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

// tile of one workgroup: 64 cells along axis 2 x 8 along axis 1 (of 2 / 4 / 8 rows, the fastest)
constexpr int kFlat2DTileZ = 64, kFlat2DTileY = 8;

struct Flat2DStep {
    const float *u, *other;  // current level and the level it replaces
    float *out;              // new level (may alias other)
    const float *b;          // buoyancy (variable density) or null
    const float *A, *B, *G;
    const float *Cbeta, *lapother;  // power-2 attenuation (null: off); lapother = lap(other)
    float *lapout;                  // optional: receives lap(u) (the next step's lapother)
    const float *Cpml;              // CFS-PML (null: off)
    float *u1[3], *u2[3], *u3[3];   // PML auxiliary fields per axis
    const float *pml[3][4];         // PML profiles per axis: sigma, alpha, dsigma, dalpha
    float dt;
    int backward;
    // optional injection into the new level after the update (the deterministic scatter's sum,
    // in its order): imap[p] = destination index or -1; contributions ioff[d] .. ioff[d + 1]
    const int *imap, *ioff, *ipoint;
    const float *icoef, *ivalues;
    float isign;
    bool isave;   // also add it, times idt2, to the p.dt2 snapshot
    // optional interpolation of u at sparse points (the previous step's receivers: u is the
    // level that step wrote): the workgroup of tile (ty, tz) computes the points listed in
    // gpt[goff[tile] .. goff[tile + 1]), tile = ty * tiles along axis 2 + tz, as k_gather does
    const int *goff, *gpt;
    const long long *gidx;
    const float *gw;
    int gntap;
    float gsign;
    float *gout;
    // optional gradient accumulation of the as-written adjoint on the interior (u is the
    // adjoint level): grad_vp += pdt u, prec_vp += pdt^2 with pdt the saved p.dt2 (graw 0) or the
    // second difference of the raw snapshots (graw 1: gps -/+ one slot, null: zero); with grho,
    // grad_rho += f u, prec_rho += f^2, f = -sum_d D-_d(b) D-_d(p) - b lap(p) of the raw snapshot
    const float *gps, *gpsm, *gpsp;
    float gfdt2;
    int graw;
    float *ggv, *gpv, *ggr, *gpr;
    float *save;  // optional fused save into `box` (save_mode 1: p.dt2, 2: raw u)
    int save_mode;
    Box box;
    float idt2;
};

// true for the 2D padded layout within the kernel's 32-bit buffer offsets
bool flat2d_supported(const Geo &g);
int flat2d_step(const Geo &g, int rho_mode, const Flat2DStep &x, hipStream_t stream);
// Several shots on one grid per launch (workgroups of shot k along grid axis z): shot k's
// arguments are tab[k] (device memory, read by the kernel); `first` is the host copy of tab[0],
// whose options (PML, power 2, save) select the kernel and must be those of every shot.
int flat2d_step_shots(const Geo &g, int rho_mode, const Flat2DStep &first, const Flat2DStep *tab, int nshot,
                      hipStream_t stream);
