// Internal interfaces shared by the kernel translation units and the C-ABI layer.
#pragma once

#include <string>
#include <vector>

#include <hip/hip_runtime.h>
#include <cstddef>

namespace chorus {

constexpr int kMaxN = 8;       // solution points per direction supported by the kernels

// minimum waves per SIMD requested from the register allocator (launch bounds);
// FP64 viscous kernels: 4 (at most 128 VGPRs) except at N=3, where k_grad
// would spill 24 registers
#ifndef CHORUS_VISC_MINW
#define CHORUS_VISC_MINW(N) ((N) == 3 ? 3 : 4)
#endif
#ifndef CHORUS_IDEAL_MINW
#define CHORUS_IDEAL_MINW 3
#endif
// unroll factor of the contraction index in FP64 traces / gradient traces /
// derivative sums: full unrolling hoists every LDS load and exhausts the
// register file in the FP64 kernels (3 measured fastest at N=5 against 1, 2
// and full); FP32 kernels unroll fully
#ifndef CHORUS_F64_UNROLL
#define CHORUS_F64_UNROLL 3
#endif
template <class R, int N> constexpr int contraction_unroll() { return sizeof(R) == 8 ? CHORUS_F64_UNROLL : N + 1; }
// mixed form: unroll of the solution-point index in the delta traces (FP64
// loads) and in the gradient traces; full unrolling hoists every load and
// exhausts the register file of the viscous stage kernel
#ifndef CHORUS_MX_TRACE_UNROLL
#define CHORUS_MX_TRACE_UNROLL 2
#endif
#ifndef CHORUS_MX_GRAD_UNROLL
#define CHORUS_MX_GRAD_UNROLL 2
#endif

// stage combination kinds (FP64 in both precision forms; the association of
// each expression is the reference's, iterations.f90:92-217)
enum StageKind : int {
    KIND_RESID = 0,   // write the residual only (compresid)
    KIND_AXPY = 1,    // Q + c0*R                           c0 = b*dt   (RK_S5P3 s1,s2; FiveStageRk s1)
    KIND_3A = 2,      // (c1*Qa + c2*Q) + c0*R              c0 = b*dt   (RK_S5P3 s3..s5; FiveStageRk s3)
    KIND_3B = 3,      // (c1*Qa + c2*Q) + (c3*R)*c4         c4 = dt     (FiveStageRk s2)
    KIND_S54_4 = 4,   // KIND_3A and Rout = R                           (FiveStageRk s4, keeps resid0)
    KIND_S54_5 = 5,   // ((((c1*Qa + c2*Qb) + c3*Qc) + c4*Q) + c5*Rprev) + c0*R  (FiveStageRk s5)
};

// Mixed form: the traces of the state at the face points in the compact form
// the face passes read (the FP64 TQ is then written only for the div B
// diagnostic).  Per cell face point (TQ order): dev, the FP32 difference of
// variables 1-8 from the cell's FP64 reference state ref (any state of the
// cell's own scale: the jump of two traces is formed as the rounded
// difference of the two references plus the difference of the deviations,
// so it keeps the precision of a difference of FP64 traces to FP32 rounding
// of terms the size of a cell's variation), and az, the FP64 trace of A_z.
//   dev[cell][lfi][nfp][8] (float), az[cell][lfi][nfp] (double), ref[cell][8] (double)
struct MxTraces {
    float *dev;
    double *az;
    double *ref;
};

struct StageArgs {
    int ncell;
    const double *cellxy;
    const int2 *nbr;
    const double *Qin;
    double *Qout;
    const double *Qa, *Qb, *Qc, *Rprev;
    double *Rout;
    int kind;
    double c0, c1, c2, c3, c4, c5;
    double gam;
    const void *FF;     // common face fluxes of each cell's face points (k_face_inv)
    const double *FA;   // mixed form: the FP64 upwind A_z of each cell's face points
    const void *FFv;    // viscous face fluxes (k_face_visc)
    double *TQ;         // traces of Qout at the face points, written by the stage kernel
    MxTraces mt;        // mixed form: the same traces in the compact form
    double *sm;         // AV: indicator of Qout per cell, written by the viscous stage kernel
    int ng;             // geometry nodes per cell in cellxy: 4 (0 taken as 4) or 8
};

// Face-ordered pass: one thread per face point.
//   faces[f] = {cellL, lfiL, cellR, lfiR}: the reference's left and right cell
//   of face f with their local face index 0..3; cellR = -1 on a FREE face.
//   Face point k of face f is point k of the left cell's face numbering and
//   point N-1-k of the right cell's.
// Per cell, local face lfi and own face point nfp (face2fpconnec order):
//   TQ [cell][lfi][nfp][9]   trace of Q (FP64)
//   FF [cell][lfi][nfp][9]   this cell's common inviscid flux (contravariant,
//                            signed as the reference hands it to the cell) + A_z
//   FFv[cell][lfi][nfp][9]   this cell's common viscous + AV flux
//   gtr[cell][lfi][nfp][16]  own trace of nablaQs (k_grad)
struct FaceArgs {
    int nface;
    const int4 *faces;
    const double *cellxy;
    const double *TQ;
    MxTraces mt;             // mixed form: the traces in the compact form (TQ unused)
    void *FF;
    double *FA;              // mixed form: FP64 upwind A_z per cell face point
    void *FFv;
    const void *gtr;
    const double *sm;        // AV indicator per cell
    const double *mucell;    // k_face_visc: muAV(1) per cell of this stage's input (k_grad)
    double *eig;             // AV: the stage maximum of the Rusanov wave speed, per instance (k_face_inv
                             // forms it by atomic max; the mixed viscous face kernel reads it)
    const int *inst;         // batched instances: instance of each cell (null: one instance)
    const double *hinst;     // batched instances: AV length scale per instance
    int artiv;
    double gam, mu, eta0, h_msh, pi;
    double qin[8];           // INLE faces: the prescribed state (chorus_bc)
    int ng;                  // geometry nodes per cell in cellxy: 4 (0 taken as 4) or 8
};

// Boundary kinds of the faces without a right cell (faces[f].z < 0), held in
// faces[f].w: the right state of the inviscid flux at the face point
enum BoundaryKind : int {
    BK_ZEROGRAD = 0,   // FREE, OUTL: Q_R = Q_L
    BK_MIRROR = 1,     // SYMP, WALL: normal momentum and normal B reversed
    BK_INLET = 2,      // INLE: Q_R(1:8) = qin, A_z from the cell
};

// viscous / AV path (k_grad + k_stage_visc)
struct ViscArgs {
    int ncell;
    const double *cellxy;
    const int2 *nbr;
    const double *Qin;
    void *gtr;          // own nablaQs traces [cell][lf][nfp][16]  (Real)
    void *fvi;          // written by the gradient kernel: FP64 form, the interior viscous/AV fluxes
                        // [cell][fam][9][N*(N-1)]; mixed form, their divergence part [cell][point][8] (FP32)
    const int *inst;    // batched instances: instance of each cell (null: one instance)
    const double *hinst;   // batched instances: AV length scale per instance
    double *sm;         // log10 smoothness indicator of density per cell
    double *eig;        // this stage's maximum Rusanov wave speed (atomic max; with AV alone the
                        // mixed gradient kernel forms it, otherwise a face pass before it)
    double *eig_next;   // the next stage's slots (one per instance), zeroed by k_stage_visc
    int ninst;          // instances sharing the launch
    const double *TQ;   // traces of Qin at the face points
    MxTraces mt;        // mixed form: the same traces in the compact form
    double *mucell;     // AV: muAV(1) per cell of this stage's input, written by k_grad
                        // (mixed form with AV alone: for a unit wave speed until the stage kernel
                        // scales it by the stage maximum)
    double *sm6;        // when set (Tecplot output): log10 indicator of By of the stage's input
    int artiv;
    double mu, eta0, gam, h_msh, pi;
    double qin[8];      // INLE faces: the prescribed state (chorus_bc)
    template <class R> __host__ __device__ R *gtr_r() const { return (R *)gtr; }
    template <class R> __host__ __device__ R *fvi_r() const { return (R *)fvi; }
};

struct CTArgs {
    int ncell;          // one past the last cell processed
    int c_first;        // first cell processed (block b starts at c_first + b*CPB)
    const double *cellxy;
    const int2 *nbr;
    double *Q;
    int which;          // 0: correct B from curl A_z; 1: div B diagnostic
    double *partial;    // which=1: per block {sum divB^2, max |divB|}
    const double *TQ;   // traces of A_z (which=0) / Bx, By (which=1) at the face points
    double *TQout;      // which=0: the traces of the corrected Bx, By are written here
    MxTraces mt;        // which=0, mixed form: A_z traces read from mt.az and the corrected
                        // Bx, By traces written to mt.dev (TQ, TQout unused)
    int ng;             // geometry nodes per cell in cellxy: 4 (0 taken as 4) or 8
};

// SSP stage combination at one solution point; the new state also replaces
// the point's entries of qs (the block's LDS copy, read by no other thread).
// qa: a register copy of the point's Qa (prefetched) or null.
template <class R>
__device__ __forceinline__ void stage_combine(const StageArgs &a, size_t base, int NP, double *qs,
                                              const R *res, const double *qa = nullptr) {
    // without a prefetched Qa, the nine loads are issued together before any
    // combination (one memory latency instead of nine)
    double qv[9];
    const bool need_a = a.kind != KIND_RESID && a.kind != KIND_AXPY;
#pragma unroll
    for (int v = 0; v < 9; v++) qv[v] = qa ? qa[v] : (need_a ? a.Qa[base + (size_t)v * NP] : 0.0);
#pragma unroll
    for (int v = 0; v < 9; v++) {
        const double q = qs[v * NP];
        const double r = double(res[v]);
        const size_t o = base + (size_t)v * NP;
        double qn;
        switch (a.kind) {
        case KIND_RESID: a.Rout[o] = r; continue;
        case KIND_AXPY: qn = q + a.c0 * r; break;
        case KIND_3A: qn = a.c1 * qv[v] + a.c2 * q + a.c0 * r; break;
        case KIND_3B: qn = a.c1 * qv[v] + a.c2 * q + a.c3 * r * a.c4; break;
        case KIND_S54_4: qn = a.c1 * qv[v] + a.c2 * q + a.c0 * r; a.Rout[o] = r; break;
        default: qn = a.c1 * qv[v] + a.c2 * a.Qb[o] + a.c3 * a.Qc[o] + a.c4 * q + a.c5 * a.Rprev[o] + a.c0 * r; break;
        }
        a.Qout[o] = qn;
        qs[v * NP] = qn;
    }
}

// combination of the variables v0 .. v0+NVG-1 of a point (stage_combine for a
// subset): qs the point's FP64 state in LDS, qa its prefetched Qa, res the
// residuals of the subset, r8 the A_z residual when the subset holds var 8
template <int NVG, class R>
__device__ __forceinline__ void combine_vars(const StageArgs &a, size_t base, int NP, double *qs, int v0,
                                             const double *qa, const R *res, double r8 = 0.0) {
#pragma unroll
    for (int k = 0; k < NVG; k++) {
        const int v = v0 + k;
        const double r = v == 8 ? r8 : double(res[k]);
        const size_t o = base + (size_t)v * NP;
        const double q = qs[v * NP];
        double qn;
        switch (a.kind) {
        case KIND_RESID: a.Rout[o] = r; continue;
        case KIND_AXPY: qn = q + a.c0 * r; break;
        case KIND_3A: qn = a.c1 * qa[k] + a.c2 * q + a.c0 * r; break;
        case KIND_3B: qn = a.c1 * qa[k] + a.c2 * q + a.c3 * r * a.c4; break;
        case KIND_S54_4: qn = a.c1 * qa[k] + a.c2 * q + a.c0 * r; a.Rout[o] = r; break;
        default: qn = a.c1 * qa[k] + a.c2 * a.Qb[o] + a.c3 * a.Qc[o] + a.c4 * q + a.c5 * a.Rprev[o] + a.c0 * r; break;
        }
        a.Qout[o] = qn;
        qs[v * NP] = qn;
    }
}

// register prefetch of a point's Qa for the stage kinds that combine with it
__device__ __forceinline__ void prefetch_qa(const StageArgs &a, size_t base, int NP, double *qa) {
    const bool need = a.kind != KIND_RESID && a.kind != KIND_AXPY;
#pragma unroll
    for (int v = 0; v < 9; v++) qa[v] = need ? a.Qa[base + (size_t)v * NP] : 0.0;
}

// launchers (defined in the kernel translation units); return hipError_t
hipError_t launch_stage_ideal_fp64(int N, const StageArgs &a, const FaceArgs &f, hipStream_t s);
hipError_t launch_trace(int N, int ncell, const double *Q, double *TQ, double *sm, hipStream_t s);
hipError_t launch_trace_mixed(int N, int ncell, const double *Q, const MxTraces &mt, double *sm, hipStream_t s);
hipError_t launch_stage_ideal_mixed(int N, const StageArgs &a, const FaceArgs &f, hipStream_t s, int tile);
int tile_count();
hipError_t launch_ct(int N, const CTArgs &a, hipStream_t s);
hipError_t launch_stage_visc_fp64(int N, const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s);
hipError_t launch_stage_visc_mixed(int N, const StageArgs &a, const ViscArgs &v, const FaceArgs &f, hipStream_t s,
                                   int tile);
int ct_blocks(int N, int ncell);
int stage_blocks(int N, int ncell);

// chorus_diag_noalloc: a fixed device array per GPU holds the partials of the
// diagnostics (see chorus_fp64.hip for its layout).  8 MiB: the div B partials
// of up to about 520k cell blocks (5.2M cells at N=5) fit.
constexpr size_t kDiagScratchDoubles = size_t(1) << 20;
constexpr int kDiagResBlocks = 1024;      // partial sums of the residual norm
constexpr int kMaxDevices = 64;
// device address of the array on `device`, which must be the current device
// on the first call for it (hipGetSymbolAddress, once per device)
hipError_t diag_scratch(int device, double **ptr);
// residual norm partials of R[0, n) (when R is set) and the final sums of the
// div B partials already in scr[0, 2 nblk)
hipError_t launch_diag_reduce(const double *R, size_t n, int nblk, double *scr, hipStream_t s);
hipError_t upload_operators_fp64();
hipError_t upload_operators_mixed();

// plot-node topology of the Tecplot output (tecplotter2dsetup.f90:1-164)
struct PlotTopology {
    int nre = 0, nnode = 0, ncellp = 0;
    std::vector<double> px, py;     // node coordinates, 1-based [1..nnode]
    std::vector<int> factor;        // cells sharing each node, 1-based
    std::vector<int> gnum;          // node of (ii, jj) of each cell: [c][jj-1][ii-1], 1-based ids
    std::vector<int> c2n;           // 4 nodes (1-based) per plot cell
    std::vector<int> cstart;        // contributions of node n: [cstart[n], cstart[n+1])
    std::vector<int> contrib;       // per contribution: cell (0-based), (jj-1)*(nre+1) + (ii-1)
};
int build_plot_topology(int ncell, int nvert, int nface, const double *xv, const double *yv, const int *ivcell,
                        const int *if2c, int nre, PlotTopology &t, std::string &err);
int write_tecplot_file(const char *path, double ctime, const PlotTopology &t, const double *q, const double *av,
                       std::string &err);
// plot-node values (tecplotter.f90:19-58) on the device: 9 primitive plot
// variables q[v*nnode+n] and muAV(1), muAV(6) in av[k*nnode+n]
struct PlotArgs {
    int nnode, N, nre, artiv, ncell;
    const int *cstart, *contrib, *factor;   // factor 1-based
    const double *Q, *H;                    // state [cell][var][js][is]; H[ii*N+is] = hval(is, ii/nre)
    const double *mucell, *sm6, *eig;       // muAV(1) per cell, By indicator per cell, stage maximum
    double h_msh, gam, pi;                  // epsilon0 = eig * h_msh / (N-1) (AV.f90:147-148)
    double *V;                              // each cell's contribution per plot point [cell][point][var]
    double *q, *av;
};
hipError_t launch_plot_nodes(const PlotArgs &p, hipStream_t s);
// restart dump records (write_data.f90:49-74) of Q[cell][var][js][is]: per cell
// the 32-bit words {c+1, Q(1:9,1:N,1:N) as FP64 in Fortran order}
hipError_t launch_restart_image(int ncell, int N, const double *Q, unsigned *img, hipStream_t s);
// hval of the N solution points at the NRE+1 equally spaced plot coordinates
// (tecplotter.f90:64-82): H[ii*N+is]
void host_plot_basis(int N, int nre, double *H);

// Xs, Xf, Lmat, Mmat of the reference for N points (setsandpoints,
// computeLandM, initsetup.f90:167-315), computed on the host in the
// reference's operation order; L(ifp,is) at [ifp*N+is], M(r,is) at [r*N+is].
void host_operators(int N, double *Xs, double *Xf, double *L, double *M);
// Xs for N points (used for N-1 by the AV sensor, initsetup.f90:44-45)
void host_solution_points(int N, double *Xs);
// AV interpolation tables (AV.f90:19-58): H1[isl*N+is] = hhval(is, Xs_low(isl)),
// H2[is*(N-1)+isl] = hhval_low(isl, Xs(is))
void host_av_tables(int N, double *H1, double *H2);

}  // namespace chorus
