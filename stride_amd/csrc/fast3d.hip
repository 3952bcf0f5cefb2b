#include "fast3d.h"
bool fast3d_enabled() { return false; }
int fast3d_prepare(Fast3D &, const Geo &, const float *, const float *, int, int, float, hipStream_t) { return 0; }
int fast3d_step(Fast3D &, const Geo &, const float *, const float *, float *, float *, const Box &, float, hipStream_t) { return -5; }
int fast3d_step_adjoint(Fast3D &, const Geo &, const float *, const float *, float *, const float *, const Box &, float *, float *, hipStream_t) { return -5; }
void fast3d_free(Fast3D &) {}
