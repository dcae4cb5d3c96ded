// FP64 drop-in kernels.  Built with -ffp-contract=off: no FMA is formed, so each
// product and sum rounds separately, exactly like the reference build.
#include <mutex>

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

template <int N, int NG> static hipError_t ideal_stage(const StageArgs &a, const FaceArgs &f, hipStream_t s) {
    hipLaunchKernelGGL((k_face_inv<N, double, NG>), dim3(fblocks(N, f.nface)), dim3(256), 0, s, f);
    hipLaunchKernelGGL((k_stage_ideal<N, double, NG>), dim3(blocks_of<IShape<N>>(a.ncell)), dim3(IShape<N>::THREADS),
                       0, s, a);
    return hipGetLastError();
}

// cells with 8 geometry nodes (curved edges) take the NG = 8 instantiations
hipError_t launch_stage_ideal_fp64(int N, const StageArgs &a, const FaceArgs &f, hipStream_t s) {
    if ((a.ng == 8) != (f.ng == 8)) return hipErrorInvalidValue;
#define L_(n) return a.ng == 8 ? ideal_stage<n, 8>(a, f, s) : ideal_stage<n, 4>(a, f, s);
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

template <int N, int NG> static hipError_t ct_pass(const CTArgs &a, hipStream_t s) {
    if (a.which == 0)
        hipLaunchKernelGGL((k_ct<N, 0, NG>), dim3(nblocks<N>(a.ncell - a.c_first)), dim3(Shape<N>::THREADS), 0, s, a);
    else
        hipLaunchKernelGGL((k_ct<N, 1, NG>), dim3(nblocks<N>(a.ncell - a.c_first)), dim3(Shape<N>::THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ct(int N, const CTArgs &a, hipStream_t s) {
#define L_(n) return a.ng == 8 ? ct_pass<n, 8>(a, s) : ct_pass<n, 4>(a, s);
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

// ---------------------------------------------------------------------------
// Scratch of chorus_diag_noalloc: a fixed device array per GPU (held until the
// process ends) instead of memory the call allocates.  Layout for a context of
// nblk cell blocks:
//   [0, 2 nblk)                      k_ct<N,1> partials {sum divB^2, max |divB|} per block
//   [2 nblk, 2 nblk + kDiagResBlocks) k_sumsq partial sums of resid^2
//   the next 3 doubles               k_diag_final: sum divB^2, max |divB|, sum resid^2
// In host code the name `scratch` only stands in for the array: kernels get
// its device address from diag_scratch().  Every call on one GPU shares the
// array; calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kDiagScratchDoubles];

hipError_t diag_scratch(int device, double **ptr) {
    static std::mutex mu;
    static double *addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;   // resolved with that device current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = static_cast<double *>(p);
    }
    *ptr = addr[device];
    return hipSuccess;
}

// partial sums of R[i]^2 over a grid-stride of the array, one per block
__global__ void __launch_bounds__(256) k_sumsq(const double *R, size_t n, double *part) {
    __shared__ double red[256];
    const int tid = threadIdx.x;
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + tid; i < n; i += (size_t)gridDim.x * 256) s = s + R[i] * R[i];
    red[tid] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (tid < k) red[tid] = red[tid] + red[tid + k];
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x] = red[0];
}

// one thread: the div B block partials added and maxed in block order (the
// order chorus_diag adds them on the host, so the two agree bit for bit), and
// the residual partials added
__global__ void k_diag_final(const double *part, int nblk, const double *rpart, int nres, double *out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s2 = 0.0, mx = 0.0, r2 = 0.0;
    for (int b = 0; b < nblk; b++) {
        s2 += part[2 * b];
        const double m = part[2 * b + 1];
        mx = (mx < m) ? m : mx;
    }
    for (int b = 0; b < nres; b++) r2 += rpart[b];
    out[0] = s2;
    out[1] = mx;
    out[2] = r2;
}

hipError_t launch_diag_reduce(const double *R, size_t n, int nblk, double *scr, hipStream_t s) {
    double *rpart = scr + 2 * (size_t)nblk;
    int nres = 0;
    if (R) {
        const size_t want = (n + 255) / 256;
        nres = (int)(want < (size_t)kDiagResBlocks ? want : (size_t)kDiagResBlocks);
        if (nres < 1) nres = 1;
        hipLaunchKernelGGL(k_sumsq, dim3(nres), dim3(256), 0, s, R, n, rpart);
    }
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s, scr, nblk, rpart, nres, rpart + kDiagResBlocks);
    return hipGetLastError();
}

template <int N> static int vblocks(int ncell) { return (ncell + VShape<N>::CPB - 1) / VShape<N>::CPB; }

template <int n, int VM, int NG>
static hipError_t visc_stage(const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s) {
    hipLaunchKernelGGL((k_face_inv<n, double, NG>), dim3(fblocks(n, f.nface)), dim3(256), 0, s, f);
    hipLaunchKernelGGL((k_grad<n, double, VM, NG>), dim3(vblocks<n>(a.ncell)), dim3(V2Shape<n>::THREADS), 0, s, v);
    hipLaunchKernelGGL((k_face_visc<n, double, VM, NG>), dim3(fblocks(n, f.nface)), dim3(256), 0, s, f);
    hipLaunchKernelGGL((k_stage_visc<n, double, VM, NG>), dim3(vblocks<n>(a.ncell)), dim3(V2Shape<n>::THREADS), 0, s,
                       a, v);
    return hipGetLastError();
}

template <int n, int NG>
static hipError_t visc_stage_vm(const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s) {
    const int vm = ((v.mu != 0.0 || v.eta0 != 0.0) ? 1 : 0) | (v.artiv ? 2 : 0);
    switch (vm) {
    case 1: return visc_stage<n, 1, NG>(a, v, f, s);
    case 2: return visc_stage<n, 2, NG>(a, v, f, s);
    case 3: return visc_stage<n, 3, NG>(a, v, f, s);
    default: return hipErrorInvalidValue;
    }
}

// cells with 8 geometry nodes (curved edges) take the NG = 8 instantiations
hipError_t launch_stage_visc_fp64(int N, const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s) {
    if ((a.ng == 8) != (f.ng == 8)) return hipErrorInvalidValue;
#define V_(n) return a.ng == 8 ? visc_stage_vm<n, 8>(a, v, f, s) : visc_stage_vm<n, 4>(a, v, f, s);
    CHORUS_DISPATCH_N(N, V_)
#undef V_
}

// ---------------------------------------------------------------------------
// Plot-node values of the Tecplot output (tecplotter2d, tecplotter.f90:19-58)
// in two passes with the reference's bits:
//   k_plot_cells: one thread per (cell, variable) holds the variable's N^2
//     values in registers and forms the cell's contribution to every plot
//     point, Qs = sum_js sum_is Q(is,js)*h(is,xsi)*h(js,eta) (js outer, is
//     inner, as the reference), into V[cell][point][var];
//   k_plot_nodes: one thread per plot node adds its contributions in the order
//     the reference does (cells ascending), divides by their count and forms
//     the primitive variables; muAV(1) and muAV(6) likewise.
// ---------------------------------------------------------------------------
template <int N> __global__ void __launch_bounds__(256) k_plot_cells(PlotArgs p) {
    constexpr int NP = N * N;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.ncell * NV) return;
    const int c = t / NV, v = t % NV, W = p.nre + 1;
    const double *qp = p.Q + (size_t)c * NV * NP + v * NP;
    double q[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) q[k] = qp[k];
    double *out = p.V + (size_t)c * W * W * NV + v;
    for (int jj = 0; jj < W; jj++) {
        const double *hj = p.H + jj * N;
        for (int ii = 0; ii < W; ii++) {
            const double *hi = p.H + ii * N;
            double qs = 0.0;
#pragma unroll
            for (int js = 0; js < N; js++) {
                const double hv = hj[js];
#pragma unroll
                for (int is = 0; is < N; is++) qs = qs + q[js * N + is] * hi[is] * hv;
            }
            out[(size_t)(jj * W + ii) * NV] = qs;
        }
    }
}

__global__ void __launch_bounds__(256) k_plot_nodes(PlotArgs p) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= p.nnode) return;
    const int W = p.nre + 1, N = p.N;
    double sum[NV], a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int v = 0; v < NV; v++) sum[v] = 0.0;
    const double eps0 = p.artiv ? p.eig[0] * p.h_msh / (double)(N - 1) * 1.0 : 0.0;
    for (int k = p.cstart[n]; k < p.cstart[n + 1]; k++) {
        const int c = p.contrib[2 * k], ij = p.contrib[2 * k + 1];
        const double *qs = p.V + ((size_t)c * W * W + ij) * NV;
#pragma unroll
        for (int v = 0; v < NV; v++) sum[v] = sum[v] + qs[v];
        if (p.artiv) {
            a1 = a1 + p.mucell[c];
            a2 = a2 + mu_av(p.sm6[c], eps0, p.pi);
        }
    }
    const double f = double(p.factor[n + 1]);
    double qs[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) qs[v] = sum[v] / f;
    double o[NV];
    o[0] = qs[0];
    o[1] = qs[1] / qs[0];
    o[2] = qs[2] / qs[0];
    o[3] = qs[3] / qs[0];
    o[5] = qs[5]; o[6] = qs[6]; o[7] = qs[7]; o[8] = qs[8];
    const double nu = o[1] * o[1] + o[2] * o[2] + o[3] * o[3];
    const double nb = o[5] * o[5] + o[6] * o[6] + o[7] * o[7];
    o[4] = (qs[4] - 0.5 * qs[0] * nu - 0.5 * nb) * (p.gam - 1.0);
#pragma unroll
    for (int v = 0; v < NV; v++) p.q[(size_t)v * p.nnode + n] = o[v];
    p.av[n] = a1 / f;
    p.av[(size_t)p.nnode + n] = a2 / f;
}

hipError_t launch_plot_nodes(const PlotArgs &p, hipStream_t s) {
    const int nt = p.ncell * NV;
#define L_(n)                                                                                         \
    if (nt) hipLaunchKernelGGL(k_plot_cells<n>, dim3((nt + 255) / 256), dim3(256), 0, s, p); \
    break;
    CHORUS_DISPATCH_N(p.N, L_)
#undef L_
    if (p.nnode) hipLaunchKernelGGL(k_plot_nodes, dim3((p.nnode + 255) / 256), dim3(256), 0, s, p);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Restart dump image (write_data.f90:44-74): per cell one record of an int32
// global cell number followed by Q(1:9,1:N,1:N) of the cell in Fortran order,
// records back to back (4-byte aligned, so each FP64 value is stored as two
// 32-bit words).  One thread per value: it reads Q[c][v][js][is] of the device
// layout and writes the value's two words at its place in the record; thread 0
// of the cell also writes the cell number c+1.  The 12-byte header (iter,
// ctime) is the host's.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_restart_image(int ncell, int N, const double *__restrict__ Q,
                                                       unsigned *__restrict__ img) {
    const int NP = N * N, NQ = 9 * NP;
    const size_t rec = 1 + 2 * (size_t)NQ;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)ncell * NQ) return;
    const int c = (int)(t / NQ), d = (int)(t % NQ);
    const int v = d % 9, p = d / 9;                      // p = is + N*js
    const unsigned long long bits = (unsigned long long)__double_as_longlong(Q[(size_t)c * NQ + v * NP + p]);
    unsigned *r = img + (size_t)c * rec;
    r[1 + 2 * d] = (unsigned)bits;
    r[2 + 2 * d] = (unsigned)(bits >> 32);
    if (d == 0) r[0] = (unsigned)(c + 1);
}

hipError_t launch_restart_image(int ncell, int N, const double *Q, unsigned *img, hipStream_t s) {
    const size_t n = (size_t)ncell * 9 * N * N;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_restart_image, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ncell, N, Q, img);
    return hipGetLastError();
}

}  // namespace chorus
