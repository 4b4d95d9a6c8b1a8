// hhg_solve_piflux3d.hip — solve, trajectory and cost-probe kernels of the
// reference's 3D pi-flux lattice (field along z, relaxation toward the
// instantaneous equilibrium; templates in hhg_solve.hpp), one translation unit
// per model so the instantiations compile in parallel.
#include "hhg_solve.hpp"

HHG_INSTANTIATE_MODEL(hhg::MODEL_PIFLUX3D)
