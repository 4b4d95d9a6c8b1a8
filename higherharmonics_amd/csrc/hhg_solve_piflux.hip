// hhg_solve_piflux.hip — solve, trajectory and cost-probe kernels of the piflux model
// (templates in hhg_solve.hpp), one translation unit per model so the
// instantiations compile in parallel.
#include "hhg_solve.hpp"

HHG_INSTANTIATE_MODEL(hhg::MODEL_PIFLUX2D)
