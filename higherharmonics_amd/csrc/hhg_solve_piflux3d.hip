// hhg_solve_piflux3d.hip — solve, trajectory and cost-probe kernels of the
// reference's 3D pi-flux lattice (field along z, relaxation toward the
// instantaneous equilibrium; templates in hhg_solve.hpp), one translation unit
// per model so the instantiations compile in parallel.
#include "hhg_solve.hpp"

HHG_INSTANTIATE_MODEL(hhg::MODEL_PIFLUX3D)

namespace hhg {

// PiFlux::compute_occupation_numbers' per-k evolution; x.s_out receives
// [n_sel][N+1][2] occupations
hipError_t launch_occupations(int pulse, const TrajArgs& x) {
    const int blocks = (int)((x.n_sel + 63) / 64);
#define HHG_OCC(P)                                                                                                \
    hipLaunchKernelGGL((occupation_kernel<P>), dim3(blocks), dim3(64), 0, x.st, x.jobs, x.kx, x.ky, x.kz, x.sel, \
                       x.n_sel, x.sp, x.s_out, x.ctr);                                                           \
    break
    switch (pulse) {
    case PULSE_COS2: HHG_OCC(PULSE_COS2);
    case PULSE_GAUSS: HHG_OCC(PULSE_GAUSS);
    case PULSE_DOUBLE_GAUSS: HHG_OCC(PULSE_DOUBLE_GAUSS);
    case PULSE_DOUBLE_COS2: HHG_OCC(PULSE_DOUBLE_COS2);
    case PULSE_CONTINUOUS: HHG_OCC(PULSE_CONTINUOUS);
    case PULSE_SPLINE: HHG_OCC(PULSE_SPLINE);
    default: return hipErrorInvalidValue;
    }
#undef HHG_OCC
    return hipGetLastError();
}

}  // namespace hhg
