// Mixed-precision kernels: FP32 traces, fluxes, Rusanov and derivative
// contractions; the state, the RK stage sums and the metrics stay FP64.
#include "kernels.hpp"

namespace chorus {

hipError_t upload_operators_mixed() { return upload_tables(); }

template <int N> static int nblocks(int ncell) { return (ncell + Shape<N>::CPB - 1) / Shape<N>::CPB; }

hipError_t launch_stage_ideal_mixed(int N, const StageArgs &a, hipStream_t s) {
#define L_(n)                                                                                     \
    case n:                                                                                       \
        hipLaunchKernelGGL((k_stage_ideal<n, float>), dim3(nblocks<n>(a.ncell)), dim3(Shape<n>::THREADS), 0, s, a); \
        return hipGetLastError();
    switch (N) {
        L_(2) L_(3) L_(4) L_(5) L_(6) L_(7) L_(8)
    default: return hipErrorInvalidValue;
    }
#undef L_
}

}  // namespace chorus
