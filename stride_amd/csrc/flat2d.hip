// Fused 2D wavefield update (see flat2d.h).
//
// One launch per time step computes, per cell,
//   out = A u + B other + G K u + Cbeta (lap u - lap other) - Cpml sum_d u1_d
// -- equation (1) of the general kernels -- where K is the SO10 Laplacian (no density) or
// sum_d D-_d (bbar_d D+_d u) (constant / variable density), and, on the CFS-PML slabs, first
// advances the auxiliary fields u3 -> u2 -> u1 of each axis from the SO10 first and second
// derivatives of u (stride/physics/boundaries/devito.py:133-171). The general path runs this as
// up to seven launches (two staggered-derivative passes, a divergence pass, two PML passes,
// the update); on the 2D grids (240^2 - 572^2) each of those is dominated by its fixed cost.
// The source (forward) or residual (adjoint) injection into the new level is applied here as
// well, by the thread that writes the destination cell, and the previous step's receiver
// interpolation (which reads this step's u) by the workgroup whose tile holds a receiver's first
// tap. The Laplacian of `other` (power-2
// attenuation) must be taken before the update overwrites
// `other` in place: each launch also writes the Laplacian of its u, which is the next step's
// other level, so a separate Laplacian launch runs only on the first step of a run.
//
// Work decomposition: a workgroup of 8 waves owns 64 (axis 2) x 8 (axis 1) cells, one per
// thread. u is staged in LDS with a 9-cell halo (rows y0-9 .. y0+16, columns z0-12 .. z0+75),
// b with a 5-cell halo; with density the fluxes bbar D+ u are formed once per half point into
// LDS tiles and differenced back to the nodes.

#include "flat2d.h"
#include "bufio.h"

namespace {

using namespace bufio;

constexpr int kTZ = kFlat2DTileZ, kNW = kFlat2DTileY, kTY = kNW;
constexpr int UY = kTY + 18, UZ = 88, UC = UZ / 4;
constexpr int BY = kTY + 10, BZ = 80, BC = BZ / 4;
constexpr int FY = kTY + 9, FZN = 73, FZP = 76;
constexpr int NT = kNW * 64;

__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// RHO: 0 Laplacian, 1 constant density (bbar = 1), 2 variable density
template <int RHO, bool PML, bool BETA, bool SAVE>
__global__ __launch_bounds__(NT) void k_flat2d(Geo g, Flat2DStep x) {
#include "flat2d_tile.h"
}

// several shots per launch: shot blockIdx.z's arguments from a device table
template <int RHO, bool PML, bool BETA, bool SAVE>
__global__ __launch_bounds__(NT) void k_flat2d_shots(Geo g, const Flat2DStep *__restrict__ tab) {
    const Flat2DStep &x = tab[blockIdx.z];
#include "flat2d_tile.h"
}

template <int RHO, bool PML, bool BETA>
void launch_rho(const Geo &g, const Flat2DStep &x, const Flat2DStep *tab, dim3 grid, hipStream_t stream) {
    if (tab && x.save)
        k_flat2d_shots<RHO, PML, BETA, true><<<grid, NT, 0, stream>>>(g, tab);
    else if (tab)
        k_flat2d_shots<RHO, PML, BETA, false><<<grid, NT, 0, stream>>>(g, tab);
    else if (x.save)
        k_flat2d<RHO, PML, BETA, true><<<grid, NT, 0, stream>>>(g, x);
    else
        k_flat2d<RHO, PML, BETA, false><<<grid, NT, 0, stream>>>(g, x);
}

template <int RHO>
void launch_flags(const Geo &g, const Flat2DStep &x, const Flat2DStep *tab, dim3 grid, hipStream_t stream) {
    const bool pml = x.Cpml != nullptr, beta = x.Cbeta != nullptr;
    if (pml && beta) launch_rho<RHO, true, true>(g, x, tab, grid, stream);
    else if (pml) launch_rho<RHO, true, false>(g, x, tab, grid, stream);
    else if (beta) launch_rho<RHO, false, true>(g, x, tab, grid, stream);
    else launch_rho<RHO, false, false>(g, x, tab, grid, stream);
}

}  // namespace

bool flat2d_supported(const Geo &g) {
    if (g.act[0] || !g.act[1] || !g.act[2] || g.n[0] != 1) return false;
    return g.total * 4 < (long long)bufio::kOOB;
}

int flat2d_step(const Geo &g, int rho_mode, const Flat2DStep &x, hipStream_t stream) {
    const dim3 grid((g.n[2] + kTZ - 1) / kTZ, (g.n[1] + kTY - 1) / kTY);
    if (rho_mode == 0) launch_flags<0>(g, x, nullptr, grid, stream);
    else if (rho_mode == 1) launch_flags<1>(g, x, nullptr, grid, stream);
    else launch_flags<2>(g, x, nullptr, grid, stream);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int flat2d_step_shots(const Geo &g, int rho_mode, const Flat2DStep &first, const Flat2DStep *tab, int nshot,
                      hipStream_t stream) {
    if (!tab || nshot < 1 || nshot > 65535) return -1;
    const dim3 grid((g.n[2] + kTZ - 1) / kTZ, (g.n[1] + kTY - 1) / kTY, nshot);
    if (rho_mode == 0) launch_flags<0>(g, first, tab, grid, stream);
    else if (rho_mode == 1) launch_flags<1>(g, first, tab, grid, stream);
    else launch_flags<2>(g, first, tab, grid, stream);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
