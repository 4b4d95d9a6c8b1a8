// hhg_solve_fused.hip — fused RKF78 launches over the items of two
// (model, pulse, relaxation) groups (solve_kernel_fused in hhg_solve.hpp).
// Instantiated for the mixed sweep's pair: honeycomb under double cos^2
// pulses with PiFlux-2D under double Gaussian pulses, no relaxation.
#include "hhg_solve.hpp"

namespace hhg {

namespace {

using HoneyDoubleCos2 = GroupKey<MODEL_HONEYCOMB, PULSE_DOUBLE_COS2, RELAX_NONE>;
using PiFluxDoubleGauss = GroupKey<MODEL_PIFLUX2D, PULSE_DOUBLE_GAUSS, RELAX_NONE>;

template <class G>
bool is_key(const int* k) {
    return k[0] == G::M && k[1] == G::P && k[2] == G::R;
}

template <class G1, class G2>
bool is_pair(const int* k1, const int* k2) {
    return (is_key<G1>(k1) && is_key<G2>(k2)) || (is_key<G2>(k1) && is_key<G1>(k2));
}

template <class G1, class G2, int L>
hipError_t launch_pair(const KernelArgs& a, int blocks, int threads, size_t smem, hipStream_t st, const void** fn) {
    if (fn) {
        *fn = reinterpret_cast<const void*>(&solve_kernel_fused<G1, G2, L>);
        return hipSuccess;
    }
    hipLaunchKernelGGL((solve_kernel_fused<G1, G2, L>), dim3(blocks), dim3(threads), smem, st, a);
    return hipGetLastError();
}

hipError_t fused(const int* k1, const int* k2, int learned, const KernelArgs* a, int blocks, int threads, size_t smem,
                 hipStream_t st, const void** fn) {
    if (is_pair<PiFluxDoubleGauss, HoneyDoubleCos2>(k1, k2)) {
        KernelArgs none{};
        const KernelArgs& args = a ? *a : none;
        switch (learned) {
        case LEARN_NONE:
            return launch_pair<PiFluxDoubleGauss, HoneyDoubleCos2, LEARN_NONE>(args, blocks, threads, smem, st, fn);
        case LEARN_SCHEDULED:
            return launch_pair<PiFluxDoubleGauss, HoneyDoubleCos2, LEARN_SCHEDULED>(args, blocks, threads, smem, st,
                                                                                   fn);
        case LEARN_NONE | LOG_STEPS:
            return launch_pair<PiFluxDoubleGauss, HoneyDoubleCos2, LEARN_NONE | LOG_STEPS>(args, blocks, threads, smem,
                                                                                         st, fn);
        case LEARN_SCHEDULED | LOG_STEPS:
            return launch_pair<PiFluxDoubleGauss, HoneyDoubleCos2, LEARN_SCHEDULED | LOG_STEPS>(args, blocks, threads,
                                                                                              smem, st, fn);
        }
    }
    return hipErrorInvalidValue;
}

}  // namespace

const void* solve_fused_fn(const int key1[3], const int key2[3], int learned) {
    const void* fn = nullptr;
    return fused(key1, key2, learned, nullptr, 0, 0, 0, nullptr, &fn) == hipSuccess ? fn : nullptr;
}

hipError_t launch_solve_fused(const int key1[3], const int key2[3], int learned, const KernelArgs& a, int blocks,
                              int threads, size_t smem, hipStream_t st) {
    return fused(key1, key2, learned, &a, blocks, threads, smem, st, nullptr);
}

}  // namespace hhg
