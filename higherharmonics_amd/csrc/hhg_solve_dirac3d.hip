// hhg_solve_dirac3d.hip — solve, trajectory and cost-probe kernels of the
// reference's 3D Dirac cone (band frame, field along z; templates in
// hhg_solve.hpp), one translation unit per model so the instantiations compile
// in parallel.
#include "hhg_solve.hpp"

HHG_INSTANTIATE_MODEL(hhg::MODEL_DIRAC3D)
