// Mixed-precision kernels: FP32 traces, fluxes, Rusanov and derivative
// contractions in per-cell delta form; the state, the RK stage sums, the
// metrics, the AV indicator and the CT update stay FP64.
#include "kernels_mixed.hpp"

namespace chorus {

hipError_t upload_operators_mixed() {
    hipError_t e = upload_tables();
    return e != hipSuccess ? e : upload_av_tables();
}

template <class S> static int blocks_of(int ncell) { return (ncell + S::CPB - 1) / S::CPB; }
static int fblocks(int N, int nface) { return (nface * N + 255) / 256; }

hipError_t launch_stage_ideal_mixed(int N, const StageArgs &a, const FaceArgs &f, hipStream_t s) {
#define L_(n)                                                                                     \
    case n:                                                                                       \
        hipLaunchKernelGGL((k_face_inv_mx<n>), dim3(fblocks(n, f.nface)), dim3(256), 0, s, f);    \
        hipLaunchKernelGGL((k_stage_ideal_mx<n>), dim3(blocks_of<IShape<n>>(a.ncell)), dim3(IShape<n>::THREADS), 0, s, a); \
        return hipGetLastError();
    switch (N) {
        L_(2) L_(3) L_(4) L_(5) L_(6) L_(7) L_(8)
    default: return hipErrorInvalidValue;
    }
#undef L_
}

template <int n, int VM>
static hipError_t visc_stage(const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s) {
    hipLaunchKernelGGL((k_face_inv_mx<n>), dim3(fblocks(n, f.nface)), dim3(256), 0, s, f);
    hipLaunchKernelGGL((k_grad_mx<n>), dim3(blocks_of<VShape<n>>(a.ncell)), dim3(V2Shape<n>::THREADS), 0, s, v);
    hipLaunchKernelGGL((k_face_visc_mx<n, VM>), dim3(fblocks(n, f.nface)), dim3(256), 0, s, f);
    hipLaunchKernelGGL((k_stage_visc_mx<n, VM>), dim3(blocks_of<VShape<n>>(a.ncell)), dim3(MxVShape<n>::THREADS), 0, s, a, v);
    return hipGetLastError();
}

template <int n>
static hipError_t visc_stage_vm(const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s) {
    const int vm = ((v.mu != 0.0 || v.eta0 != 0.0) ? 1 : 0) | (v.artiv ? 2 : 0);
    switch (vm) {
    case 1: return visc_stage<n, 1>(a, v, f, s);
    case 2: return visc_stage<n, 2>(a, v, f, s);
    case 3: return visc_stage<n, 3>(a, v, f, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_stage_visc_mixed(int N, const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s) {
    switch (N) {
    case 2: return visc_stage_vm<2>(a, v, f, s);
    case 3: return visc_stage_vm<3>(a, v, f, s);
    case 4: return visc_stage_vm<4>(a, v, f, s);
    case 5: return visc_stage_vm<5>(a, v, f, s);
    case 6: return visc_stage_vm<6>(a, v, f, s);
    case 7: return visc_stage_vm<7>(a, v, f, s);
    case 8: return visc_stage_vm<8>(a, v, f, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace chorus
