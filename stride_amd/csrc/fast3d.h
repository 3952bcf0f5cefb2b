// Specialised 3D wavefield-update kernels (declarations).
#pragma once
#include <hip/hip_runtime.h>
#include "kernels.h"

struct Fast3D {
    float *coef = nullptr;  // placeholder until the specialised path lands
};

bool fast3d_enabled();
int fast3d_prepare(Fast3D &f, const Geo &g, const float *vp, const float *tables, int tlen, int boundary, float dt,
                   hipStream_t stream);
int fast3d_step(Fast3D &f, const Geo &g, const float *u, const float *o, float *out, float *save, const Box &box,
                float idt2, hipStream_t stream);
int fast3d_step_adjoint(Fast3D &f, const Geo &g, const float *u, const float *o, float *out, const float *ps,
                        const Box &box, float *grad, float *prec, hipStream_t stream);
void fast3d_free(Fast3D &f);
