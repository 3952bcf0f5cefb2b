// Specialised 3D wavefield update for the headline schedule: OT2, no density
// term, sponge or no boundary, no attenuation (one shot per pass: also attenuation
// power 0, 2 or None). One pass per time step reads u (with halo), the other level,
// vp, and writes the new level in place of the other level: 16 algorithmic bytes per
// grid-point update (+ 4 B alpha with attenuation; power 2 also lap other and lap u:
// 28 B). The every-f-th-step
// save of p.dt2 (forward) and the vp gradient / preconditioner accumulation
// (adjoint) are fused into the same pass.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

struct PairLevels {
    float *lev0, *lev1;
};

struct FastTune {
    int rows;     // rows (axis 1) per thread; the tile is 64 (axis 2) x waves*rows (axis 1)
    int chunk;    // planes (axis 0) marched per workgroup
    int version;  // 7: 4 waves (rows 2), 8: 8 waves (rows 1 or 2)
};

struct Fast3D {
    const float *axis = nullptr;  // per-axis layer tables [3][alen] (device, owned by the solver)
    int alen = 0;
    int boundary = -1;
    const float *vp = nullptr;
    float dt = 0.f;
    FastTune tune = {2, 24, 7};
    FastTune tune2 = {2, 24, 4};  // two-shot kernel: rows, chunk, waves (4 or 8)
    // multi-shot mapping, tiles per group: 0 a row of tiles along axis 2, 1 a chunk slab, 2 a whole
    // pair (pair-major order; the fastest: one launch's tail is shared by all its pairs)
    int multi_group = 2;
    bool prepared = false;
};

bool fast3d_enabled();
// true when the grid's fields fit the kernels' 32-bit buffer offsets (else: general kernels)
bool fast3d_supported(const Geo &g);
int fast3d_prepare(Fast3D &f, const Geo &g, const float *vp, const float *tables, int tlen, int boundary, float dt,
                   hipStream_t stream);
// Attenuation operands of the one-shot update: the power (0, 2, or -1 for None; -2 none) and
// alpha (Np); for power 2 also the Laplacian of the level `o` (read; out may overwrite o, so it
// was kept from the previous step) and the Laplacian of u (written: the next step's). The
// coefficients are the general update's: power 0 adds dt alpha vp to the damping, power None
// scales the step by alpha, power 2 adds Cbeta (lap u - lap o), Cbeta = dt^2 ia 2 alpha vp^3 / dt.
struct FastAtt {
    int power;
    const float *alpha, *lapo;
    float *lapn;
    // CFS-PML (non-null cpml; the update prepared without a boundary): the term - Cpml sum_d u1_d
    // of the general update, with the u1 fields already advanced for this step
    const float *cpml = nullptr;
    const float *u1[3] = {nullptr, nullptr, nullptr};
    // OT4 second pass (non-null ot4_u; attenuation off / power 0 / None, no PML): the update's u
    // argument is w = m K u from fast3d_ot4_first, ot4_u the level itself and ot4_ku its K u;
    // out = A u + B other + G (K u + c4 K w), c4 = dt^2 / 12
    const float *ot4_u = nullptr, *ot4_ku = nullptr;
    float c4 = 0.f;
};
// OT4 first pass: ku = K u and w = vp^2 K u over the domain (one march)
int fast3d_ot4_first(Fast3D &f, const Geo &g, const float *u, float *ku, float *w, hipStream_t stream);
// forward step: out = step(u, o) (out may alias o); optional fused save into `save` (box layout):
// p.dt2 (save_mode 1) or the level u itself (2: raw snapshots)
int fast3d_step(Fast3D &f, const Geo &g, const float *u, const float *o, float *out, float *save, const Box &box,
                float idt2, hipStream_t stream, const FastAtt *att = nullptr, int save_mode = 1);
// adjoint step with optional fused gradient: grad += ps * u, prec += ps^2 on the interior (box layout ps)
int fast3d_step_adjoint(Fast3D &f, const Geo &g, const float *u, const float *o, float *out, const float *ps,
                        const Box &box, float *grad, float *prec, hipStream_t stream,
                        const FastAtt *att = nullptr);
// out = the centred SO10 Laplacian of u over the domain (the update's own stencil, one pass)
int fast3d_laplacian(Fast3D &f, const Geo &g, const float *u, float *out, hipStream_t stream);
void fast3d_free(Fast3D &f);
// tile shape used by later launches (autotuner)
int fast3d_set_tune(Fast3D &f, int rows, int chunk, int version);
// Two shots per pass: u, o, out, save and ps hold two interleaved shots per cell (8 bytes per
// padded cell / box cell); vp, grad and prec are shared (the gradient sums both shots).
int fast3d_step_b2(Fast3D &f, const Geo &g, const float *u, const float *o, float *out, float *save, const Box &box,
                   float idt2, hipStream_t stream);
int fast3d_step_adjoint_b2(Fast3D &f, const Geo &g, const float *u, const float *o, float *out, const float *ps,
                           const Box &box, float *grad, float *prec, hipStream_t stream);
int fast3d_set_tune_b2(Fast3D &f, int rows, int chunk, int waves);
// Several shot pairs per launch, plain step (the multi-shot schedule): pair k's two level
// buffers (two interleaved shots per cell each) are pairs[k] (a device table); this step reads
// u = parity ? lev1 : lev0 and writes the new level over the other one. vp is shared.
int fast3d_step_multi(Fast3D &f, const Geo &g, const PairLevels *pairs, int npair, int parity, hipStream_t stream);
