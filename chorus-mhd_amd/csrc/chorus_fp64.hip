// FP64 drop-in kernels.  Built with -ffp-contract=off: no FMA is formed, so each
// product and sum rounds separately, exactly like the reference build.
#include "kernels_visc.hpp"

namespace chorus {

hipError_t upload_operators_fp64() {
    hipError_t e = upload_tables();
    return e != hipSuccess ? e : upload_av_tables();
}

template <int N> static int nblocks(int ncell) { return (ncell + Shape<N>::CPB - 1) / Shape<N>::CPB; }
template <class S> static int blocks_of(int ncell) { return (ncell + S::CPB - 1) / S::CPB; }

#define CHORUS_DISPATCH_N(N, CALL) \
    switch (N) {                   \
    case 2: CALL(2);               \
    case 3: CALL(3);               \
    case 4: CALL(4);               \
    case 5: CALL(5);               \
    case 6: CALL(6);               \
    case 7: CALL(7);               \
    case 8: CALL(8);               \
    default: return hipErrorInvalidValue; \
    }

static int fblocks(int N, int nface) { return (nface * N + 255) / 256; }

hipError_t launch_stage_ideal_fp64(int N, const StageArgs &a, const FaceArgs &f, hipStream_t s) {
#define L_(n)                                                                                        \
    hipLaunchKernelGGL((k_face_inv<n, double>), dim3(fblocks(n, f.nface)), dim3(256), 0, s, f);      \
    hipLaunchKernelGGL((k_stage_ideal<n, double>), dim3(blocks_of<IShape<n>>(a.ncell)), dim3(IShape<n>::THREADS), 0, s, a); \
    return hipGetLastError();
    CHORUS_DISPATCH_N(N, L_)
#undef L_
}

hipError_t launch_trace(int N, int ncell, const double *Q, double *TQ, double *sm, hipStream_t s) {
#define L_(n) \
    hipLaunchKernelGGL((k_trace<n>), dim3(nblocks<n>(ncell)), dim3(Shape<n>::THREADS), 0, s, ncell, Q, TQ, sm); \
    return hipGetLastError();
    CHORUS_DISPATCH_N(N, L_)
#undef L_
}

hipError_t launch_ct(int N, const CTArgs &a, hipStream_t s) {
#define L_(n)                                                                                            \
    if (a.which == 0)                                                                                    \
        hipLaunchKernelGGL((k_ct<n, 0>), dim3(nblocks<n>(a.ncell - a.c_first)), dim3(Shape<n>::THREADS), 0, s, a); \
    else                                                                                                 \
        hipLaunchKernelGGL((k_ct<n, 1>), dim3(nblocks<n>(a.ncell - a.c_first)), dim3(Shape<n>::THREADS), 0, s, a); \
    return hipGetLastError();
    CHORUS_DISPATCH_N(N, L_)
#undef L_
}

int stage_blocks(int N, int ncell) {
    switch (N) {
    case 2: return nblocks<2>(ncell);
    case 3: return nblocks<3>(ncell);
    case 4: return nblocks<4>(ncell);
    case 5: return nblocks<5>(ncell);
    case 6: return nblocks<6>(ncell);
    case 7: return nblocks<7>(ncell);
    case 8: return nblocks<8>(ncell);
    default: return 0;
    }
}
int ct_blocks(int N, int ncell) { return stage_blocks(N, ncell); }

template <int N> static int vblocks(int ncell) { return (ncell + VShape<N>::CPB - 1) / VShape<N>::CPB; }

template <int n, int VM>
static hipError_t visc_stage(const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s) {
    hipLaunchKernelGGL((k_face_inv<n, double>), dim3(fblocks(n, f.nface)), dim3(256), 0, s, f);
    hipLaunchKernelGGL((k_grad<n, double>), dim3(vblocks<n>(a.ncell)), dim3(V2Shape<n>::THREADS), 0, s, v);
    hipLaunchKernelGGL((k_face_visc<n, double, VM>), dim3(fblocks(n, f.nface)), dim3(256), 0, s, f);
    hipLaunchKernelGGL((k_stage_visc<n, double, VM>), dim3(vblocks<n>(a.ncell)), dim3(V2Shape<n>::THREADS), 0, s, a, v);
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

hipError_t launch_stage_visc_fp64(int N, const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s) {
#define V_(n) return visc_stage_vm<n>(a, v, f, s);
    CHORUS_DISPATCH_N(N, V_)
#undef V_
}

}  // namespace chorus
