// Mixed-precision kernels: FP32 traces, fluxes, Rusanov and derivative
// contractions in per-cell delta form; the state, the RK stage sums, the
// metrics, the AV indicator and the CT update stay FP64.
#include "kernels_mixed.hpp"

namespace chorus {

hipError_t upload_operators_mixed() {
    hipError_t e = upload_tables();
    if (e == hipSuccess) e = upload_av_tables();
    return e != hipSuccess ? e : upload_av_roundtrip();
}

template <class S> static int blocks_of(int ncell) { return (ncell + S::CPB - 1) / S::CPB; }

template <int n> static hipError_t trace_mx(int ncell, const double *Q, const MxTraces &mt, double *sm, hipStream_t s) {
    hipLaunchKernelGGL((k_trace_mx<n>), dim3(blocks_of<Shape<n>>(ncell)), dim3(Shape<n>::THREADS), 0, s, ncell, Q, mt,
                       sm);
    return hipGetLastError();
}

hipError_t launch_trace_mixed(int N, int ncell, const double *Q, const MxTraces &mt, double *sm, hipStream_t s) {
    switch (N) {
    case 2: return trace_mx<2>(ncell, Q, mt, sm, s);
    case 3: return trace_mx<3>(ncell, Q, mt, sm, s);
    case 4: return trace_mx<4>(ncell, Q, mt, sm, s);
    case 5: return trace_mx<5>(ncell, Q, mt, sm, s);
    case 6: return trace_mx<6>(ncell, Q, mt, sm, s);
    case 7: return trace_mx<7>(ncell, Q, mt, sm, s);
    case 8: return trace_mx<8>(ncell, Q, mt, sm, s);
    default: return hipErrorInvalidValue;
    }
}
static int fblocks(int N, int nface) { return (nface * N + 255) / 256; }

// tile shapes of the mixed cell kernels, chosen per order by the autotuner.
// A shape is a threads-per-block target of the ideal / viscous kernels:
//   shape 0: 256 / 128, shape 1: 128 / 64, shape 2: 512 / 256.
// Tile t gives the ideal or viscous stage kernel shape kStage[t] and the
// gradient kernel shape kGrad[t]: tiles 0-2 one shape for both, 3-8 the other
// pairs (the two kernels of an order need not prefer the same shape).
namespace {
constexpr int kStage[9] = {0, 1, 2, 0, 0, 1, 1, 2, 2};
constexpr int kGrad[9] = {0, 1, 2, 1, 2, 0, 2, 0, 1};
}
int tile_count() { return 9; }

template <int n, int TBI, int NG>
static hipError_t ideal_stage(const StageArgs &a, const FaceArgs &f, hipStream_t s) {
    using S = FluxCellShape<n, TBI>;
    hipLaunchKernelGGL((k_face_inv_mx<n, NG>), dim3(fblocks(n, f.nface)), dim3(256), 0, s, f);
    hipLaunchKernelGGL((k_stage_ideal_mx<n, TBI, NG>), dim3(blocks_of<S>(a.ncell)), dim3(S::THREADS), 0, s, a);
    return hipGetLastError();
}

template <int n, int NG>
static hipError_t ideal_stage_ng(const StageArgs &a, const FaceArgs &f, hipStream_t s, int tile) {
    switch (kStage[tile]) {
    case 1: return ideal_stage<n, 128, NG>(a, f, s);
    case 2: return ideal_stage<n, 512, NG>(a, f, s);
    default: return ideal_stage<n, 256, NG>(a, f, s);
    }
}

// cells with 8 geometry nodes (curved edges) take the NG = 8 instantiations
template <int n>
static hipError_t ideal_stage_tile(const StageArgs &a, const FaceArgs &f, hipStream_t s, int tile) {
    if ((a.ng == 8) != (f.ng == 8)) return hipErrorInvalidValue;
    return a.ng == 8 ? ideal_stage_ng<n, 8>(a, f, s, tile) : ideal_stage_ng<n, 4>(a, f, s, tile);
}

hipError_t launch_stage_ideal_mixed(int N, const StageArgs &a, const FaceArgs &f, hipStream_t s, int tile) {
    switch (N) {
    case 2: return ideal_stage_tile<2>(a, f, s, tile);
    case 3: return ideal_stage_tile<3>(a, f, s, tile);
    case 4: return ideal_stage_tile<4>(a, f, s, tile);
    case 5: return ideal_stage_tile<5>(a, f, s, tile);
    case 6: return ideal_stage_tile<6>(a, f, s, tile);
    case 7: return ideal_stage_tile<7>(a, f, s, tile);
    case 8: return ideal_stage_tile<8>(a, f, s, tile);
    default: return hipErrorInvalidValue;
    }
}

template <int n, int TBG, int TBV, int VM, int NG>
static hipError_t visc_stage(const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s) {
    using S = CellShape<n, TBV>;
    using SG = CellShape<n, TBG>;
    // the gradient kernel reads the other cells' traces itself; only physical
    // viscosity with AV needs the wave-speed maximum before it (with AV alone
    // the gradient kernel forms it)
    if constexpr (VM == 3)
        hipLaunchKernelGGL((k_face_inv_mx<n, NG, false>), dim3(fblocks(n, f.nface)), dim3(256), 0, s, f);
    hipLaunchKernelGGL((k_grad_mx<n, TBG, VM, NG>), dim3(blocks_of<SG>(a.ncell)), dim3(V2ShapeT<n, TBG>::THREADS), 0, s,
                       v);
    hipLaunchKernelGGL((k_face_visc_mx<n, VM, NG>), dim3(fblocks(n, f.nface)), dim3(256), 0, s, f);
    hipLaunchKernelGGL((k_stage_visc_mx<n, TBV, VM, NG>), dim3(blocks_of<S>(a.ncell)), dim3(MxVShape<n, TBV>::THREADS),
                       0, s, a, v);
    return hipGetLastError();
}

template <int n, int TBG, int TBV, int NG>
static hipError_t visc_stage_vm_ng(const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s) {
    const int vm = ((v.mu != 0.0 || v.eta0 != 0.0) ? 1 : 0) | (v.artiv ? 2 : 0);
    switch (vm) {
    case 1: return visc_stage<n, TBG, TBV, 1, NG>(a, v, f, s);
    case 2: return visc_stage<n, TBG, TBV, 2, NG>(a, v, f, s);
    case 3: return visc_stage<n, TBG, TBV, 3, NG>(a, v, f, s);
    default: return hipErrorInvalidValue;
    }
}

template <int n, int TBG, int TBV>
static hipError_t visc_stage_vm(const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s) {
    if ((a.ng == 8) != (f.ng == 8)) return hipErrorInvalidValue;
    return a.ng == 8 ? visc_stage_vm_ng<n, TBG, TBV, 8>(a, v, f, s) : visc_stage_vm_ng<n, TBG, TBV, 4>(a, v, f, s);
}

// the viscous kernels' TB of each shape (shape 0: 128, 1: 64, 2: 256)
template <int n, int TBV>
static hipError_t visc_stage_grad(const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s, int g) {
    switch (g) {
    case 1: return visc_stage_vm<n, 64, TBV>(a, v, f, s);
    case 2: return visc_stage_vm<n, 256, TBV>(a, v, f, s);
    default: return visc_stage_vm<n, 128, TBV>(a, v, f, s);
    }
}

template <int n>
static hipError_t visc_stage_tile(const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s, int tile) {
    switch (kStage[tile]) {
    case 1: return visc_stage_grad<n, 64>(a, v, f, s, kGrad[tile]);
    case 2: return visc_stage_grad<n, 256>(a, v, f, s, kGrad[tile]);
    default: return visc_stage_grad<n, 128>(a, v, f, s, kGrad[tile]);
    }
}

hipError_t launch_stage_visc_mixed(int N, const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s,
                                   int tile) {
    switch (N) {
    case 2: return visc_stage_tile<2>(a, v, f, s, tile);
    case 3: return visc_stage_tile<3>(a, v, f, s, tile);
    case 4: return visc_stage_tile<4>(a, v, f, s, tile);
    case 5: return visc_stage_tile<5>(a, v, f, s, tile);
    case 6: return visc_stage_tile<6>(a, v, f, s, tile);
    case 7: return visc_stage_tile<7>(a, v, f, s, tile);
    case 8: return visc_stage_tile<8>(a, v, f, s, tile);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace chorus
