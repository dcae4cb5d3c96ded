// Viscous / artificial-viscosity path (vismode=1 with ARTIV=1 or mu, eta0 != 0).
//
// The reference stage (compflux.f90:1-31) has three neighbour-coupled phases
// and one global reduction; on the device they are four launches per stage:
//
//   k_face_inv   per face point: Rusanov flux and the stage maximum of its wave
//                speed (getrusanov.f90:1-109, 105-107)
//   k_grad       per cell: central Qv at every flux point (face values average
//                both stored traces, interfaceflux_all.f90:79-91), nablaQs at
//                the solution points (nablaQsp, compflux.f90:107-157), its own
//                traces at the four faces, and the viscous and AV fluxes of the
//                interior flux points (comp_viscous_flux, compflux.f90:205-256;
//                FP64 form)
//   k_face_visc  per face point: the viscous and AV flux from the face-averaged
//                Qv and gradient (interfaceVISflux, interfaceAVflux, BCVISFLUX,
//                comp_viscous_flux) for both cells, each in its own metric
//   k_stage_visc per cell: interior inviscid fluxes, the stored interior
//                viscous and face fluxes, the residual (compresid.f90:1-67),
//                the SSP stage combination, the traces of the new state and its
//                AV smoothness indicator of density (AV.f90:15-82)
//
// Per cell, local face lf and face point nfp (own numbering, face2fpconnec order):
//   gtr[cell][lf][nfp][16]  this cell's trace of nablaQs (vars 1..8, x then y)
//   fvi[cell][fam][9][N(N-1)]  FP64 form: viscous + AV flux (contravariant) of
//                           the interior flux points, point k = line*(N-1) + r-1
//   sm[cell]                log10 smoothness indicator of density (SM_IND(1,:))
// Only muAV(1,:) reaches the flux (mu_art = epsilon_loc(1), getrusanov.f90:279),
// so only density's indicator is evaluated.
#pragma once

#include "kernels.hpp"

namespace chorus {

static hipError_t upload_av_tables() {
    static double H1[kMaxN + 1][kMaxN * kMaxN], H2[kMaxN + 1][kMaxN * kMaxN];
    for (int n = 2; n <= kMaxN; n++) host_av_tables(n, H1[n], H2[n]);
    hipError_t e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_H1), H1, sizeof(H1))) != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_H2), H2, sizeof(H2));
}

#ifndef CHORUS_VISC_TB
#define CHORUS_VISC_TB 128
#endif
template <int N> using VShape = CellShape<N, CHORUS_VISC_TB>;

// Two threads per solution point for the viscous kernels (k_grad,
// k_stage_visc): in the flux
// phase each thread takes one task of the family (inviscid interior, viscous
// interior or face point; the inviscid and viscous segments start on wave
// boundaries and the face points fill the gaps), in the derivative phase group
// 0 sums the inviscid and A_z terms and group 1 the viscous ones.
template <int N, int TB> struct V2ShapeT {
    using S = CellShape<N, TB>;
    static constexpr int CPB = S::CPB, NP = S::NP;
    static constexpr int NIT = N * (N - 1);                  // interior points per family and cell
    static constexpr int NI = CPB * NIT;                     // inviscid (and viscous) tasks
    static constexpr int S1 = ((NI + 63) / 64) * 64;         // first viscous task
    static constexpr int NFC = CPB * 2 * N;                  // face tasks
    static constexpr int FB = NFC > (S1 - NI) ? NFC - (S1 - NI) : 0;   // face tasks after the viscous ones
    static constexpr int BT = S1 + NI + FB;                  // flux-phase task slots
    static constexpr int G1 = ((CPB * NP + 63) / 64) * 64;   // first thread of derivative group 1
    static constexpr int CT = G1 + CPB * NP;
    static constexpr int THREADS = (((BT > CT ? BT : CT) + 63) / 64) * 64;
};
template <int N> using V2Shape = V2ShapeT<N, CHORUS_VISC_TB>;


// local-face index (0..3) and own face-point number of flux point (fam, r, line)
template <int N> __device__ __forceinline__ void face_of(int fam, int r, int line, int &lfi, int &nfp) {
    if (fam == 0) { lfi = (r == 0) ? 3 : 1; nfp = (r == 0) ? (N - 1 - line) : line; }
    else          { lfi = (r == 0) ? 0 : 2; nfp = (r == 0) ? line : (N - 1 - line); }
}

// partner's own face-point number of the same physical point
template <int N> __device__ __forceinline__ int partner_nfp(int nfp) { return N - 1 - nfp; }

// AV smoothness indicator of density (AV.f90:15-82, FP64 in both forms: the
// N -> N-1 -> N interpolation round trip and log10 of the relative energy).
// Called by all threads of the block.
// rho[c * CS + p]: the block's densities (cell stride CS); CPB cells per block
// ops: the block's LDS copy of the FP64 tables (OpTables64 layout)
template <int N, int CPB, int CS = NV * N * N>
__device__ __forceinline__ void av_sensor(double *sm, const double *rho, double *sAV, int ncl, int c0,
                                          const double *ops) {
    constexpr int NP = N * N, NQ = CS, NL = N - 1;
    const double *sQ = rho;
    const double *H1 = ops + Ops64Layout<N>::H1, *H2 = ops + Ops64Layout<N>::H2;
    const int tid = threadIdx.x;
    const int pc = tid / NP, pk = tid % NP, pjs = pk / N, pis = pk % N;
    const bool pact = (tid < CPB * NP) && (pc < ncl);
    double *low = sAV;   // [c][jsl*NL+isl] uses the first NL*NL entries of each cell
    for (int it = tid; it < ncl * NL * NL; it += blockDim.x) {
        const int c = it / (NL * NL), k = it % (NL * NL), jsl = k / NL, isl = k % NL;
        double acc = 0.0;
#pragma unroll
        for (int js = 0; js < N; js++)
#pragma unroll
            for (int is = 0; is < N; is++)
                acc = acc + sQ[c * NQ + js * N + is] * H1[isl * N + is] * H1[jsl * N + js];
        low[c * NP + k] = acc;
    }
    __syncthreads();
    double dev = 0.0;
    if (pact) {
        double acc = 0.0;
#pragma unroll
        for (int jsl = 0; jsl < NL; jsl++)
#pragma unroll
            for (int isl = 0; isl < NL; isl++)
                acc = acc + low[pc * NP + jsl * NL + isl] * H2[pis * NL + isl] * H2[pjs * NL + jsl];
        const double q = sQ[pc * NQ + pk];
        dev = (acc - q) * (acc - q) / (q * q);
    }
    __syncthreads();
    if (pact) sAV[pc * NP + pk] = dev;
    __syncthreads();
    if (tid < ncl) {
        double se = 0.0;
        for (int p = 0; p < NP; p++) se = se + sAV[tid * NP + p];
        sm[c0 + tid] = log10(se);
    }
}

// traces of a state (after set-up, or for a state not produced by a stage) and,
// with sm, its AV indicator
template <int N>
__global__ void __launch_bounds__(256) k_trace(int ncell, const double *Q, double *TQ, double *sm) {
    using S = Shape<N>;
    constexpr int CPB = S::CPB, NP = S::NP, NQ = S::NQ;
    __shared__ double sQ[CPB * NQ];
    __shared__ double sAV[CPB * NP];
    __shared__ double sOps[Ops64Layout<N>::SIZE];   // FP64 operator tables (OpTables64 layout)
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = blk * CPB;
    const int ncl = min(CPB, ncell - c0);
    {
        constexpr int PQ = per_thread<CPB * NQ, S::THREADS>();
        double hq[PQ];
        gather_regs<PQ, S::THREADS>(hq, Q + (size_t)c0 * NQ, ncl * NQ);
        OpTables64<N, S::THREADS> ops;
        ops.gather();
        scatter_lds<PQ, S::THREADS>(sQ, hq, ncl * NQ);
        ops.scatter(sOps);
    }
    __syncthreads();
    store_traces<N>(sQ, ncl, c0, TQ, sOps);
    if (sm) av_sensor<N, CPB>(sm, sQ, sAV, ncl, c0, sOps);
}

// the same in the mixed form's compact form, each cell's reference state its
// central solution point (the stage kernels' choice)
template <int N>
__global__ void __launch_bounds__(256) k_trace_mx(int ncell, const double *Q, MxTraces mt, double *sm) {
    using S = Shape<N>;
    constexpr int CPB = S::CPB, NP = S::NP, NQ = S::NQ, PC = (N / 2) * N + N / 2;
    __shared__ double sQ[CPB * NQ];
    __shared__ double sAV[CPB * NP];
    __shared__ double sOps[Ops64Layout<N>::SIZE];   // FP64 operator tables (OpTables64 layout)
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = blk * CPB;
    const int ncl = min(CPB, ncell - c0);
    {
        constexpr int PQ = per_thread<CPB * NQ, S::THREADS>();
        double hq[PQ];
        gather_regs<PQ, S::THREADS>(hq, Q + (size_t)c0 * NQ, ncl * NQ);
        OpTables64<N, S::THREADS> ops;
        ops.gather();
        scatter_lds<PQ, S::THREADS>(sQ, hq, ncl * NQ);
        ops.scatter(sOps);
    }
    __syncthreads();
    store_traces_mx<N>(sQ, ncl, c0, mt, sQ + PC, NQ, NP, sOps);
    if (sm) av_sensor<N, CPB>(sm, sQ, sAV, ncl, c0, sOps);
}

// muAV(1,:) of one cell from its indicator and epsilon0 (AV.f90:120-172)
__device__ __forceinline__ double mu_av(double sm, double eps0, double pi) {
    const double se_low = -6.0, se_high = -1.0;
    const double k_range = (se_high - se_low) * 0.5, s0 = (se_high + se_low) * 0.5;
    if (sm < se_low) return 0.0;
    if (sm > se_high) return eps0;
    return eps0 / 2.0 * (1.0 + sin(pi * (sm - s0) / 2.0 / k_range));
}

// gradient of the 8 variables at flux point (ifp,jfp) of family fam: own trace
// of nablaQs (nablaQ_interior_FP, compflux.f90:159-196)
template <int N, class R>
__device__ __forceinline__ void grad_trace(const R *gq, int fam, int ifp, int jfp, R *g, const R *Lt = nullptr) {
    constexpr int NP = N * N;
#pragma unroll
    for (int d = 0; d < 2; d++) {
#pragma unroll
        for (int w = 0; w < 8; w++) g[d * NV + w] = R(0);
        g[d * NV + 8] = R(0);
    }
    // contraction index outermost: 16 loads in flight per step
    constexpr int U = sizeof(R) == 8 ? contraction_unroll<R, N>() : CHORUS_MX_GRAD_UNROLL;
#pragma unroll U
    for (int sp = 0; sp < N; sp++) {
        const int row = fam == 0 ? ifp : jfp;
        const R l = Lt ? Lt[row * N + sp] : Lop<R>(N, row, sp);
        const int idx = fam == 0 ? jfp * N + sp : sp * N + ifp;
#pragma unroll
        for (int d = 0; d < 2; d++)
#pragma unroll
            for (int w = 0; w < 8; w++) g[d * NV + w] = g[d * NV + w] + gq[(d * 8 + w) * NP + idx] * l;
    }
}

// 8 consecutive values of LDS at p (16-byte aligned) with 16-byte accesses
__device__ __forceinline__ void lds_load8(const float *p, float *v) {
    const float4 a = reinterpret_cast<const float4 *>(p)[0], b = reinterpret_cast<const float4 *>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void lds_load8(const double *p, double *v) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double2 a = reinterpret_cast<const double2 *>(p)[k];
        v[2 * k] = a.x;
        v[2 * k + 1] = a.y;
    }
}
__device__ __forceinline__ void lds_store8(float *p, const float *v) {
    reinterpret_cast<float4 *>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
    reinterpret_cast<float4 *>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
}
__device__ __forceinline__ void lds_store8(double *p, const double *v) {
#pragma unroll
    for (int k = 0; k < 4; k++) reinterpret_cast<double2 *>(p)[k] = make_double2(v[2 * k], v[2 * k + 1]);
}

// the same with the cell's nablaQs point-major (mixed form: 16-byte LDS reads)
template <int N, class R, int PS = 16>
__device__ __forceinline__ void grad_trace_pm(const R *gc, int fam, int ifp, int jfp, R *g, const R *Lt = nullptr) {
    // gc: the cell's nablaQs, point-major: [p][d*8+v] (16 values per point, PS apart)
#pragma unroll
    for (int d = 0; d < 2; d++) {
#pragma unroll
        for (int w = 0; w < 8; w++) g[d * NV + w] = R(0);
        g[d * NV + 8] = R(0);
    }
    // contraction index outermost: 16 loads in flight per step
    constexpr int U = sizeof(R) == 8 ? contraction_unroll<R, N>() : CHORUS_MX_GRAD_UNROLL;
#pragma unroll U
    for (int sp = 0; sp < N; sp++) {
        const int row = fam == 0 ? ifp : jfp;
        const R l = Lt ? Lt[row * N + sp] : Lop<R>(N, row, sp);
        const int idx = fam == 0 ? jfp * N + sp : sp * N + ifp;
        R v[16];
        lds_load8(gc + idx * PS, v);
        lds_load8(gc + idx * PS + 8, v + 8);
#pragma unroll
        for (int d = 0; d < 2; d++)
#pragma unroll
            for (int w = 0; w < 8; w++) g[d * NV + w] = g[d * NV + w] + v[d * 8 + w] * l;
    }
}

// viscous + AV contravariant flux at one flux point (comp_viscous_flux, compflux.f90:205-256)
// VM: bit 0 physical viscosity/resistivity (mu or eta0 nonzero), bit 1 AV
// (xf: the LDS copy of the flux-point coordinates Xf)
template <int N, class R, int VM>
__device__ __forceinline__ void visc_point(const ViscArgs &v, const double *mu4, int fam, int ifp, int jfp, R s1,
                                           R s2, const R *qv, const R *g, R *fv, const double *xf) {
    const R gam = R(v.gam), lam = R(-2.0 / 3.0), prandt = R(0.72);
    R Fv[NV], Gv[NV];
    if constexpr ((VM & 1) != 0) {
        vflux<R>(gam, lam, prandt, R(v.mu), R(v.eta0), qv, g, Fv, Gv);
#pragma unroll
        for (int w = 0; w < NV; w++) fv[w] = Fv[w] * s1 + Gv[w] * s2;
    } else {
#pragma unroll
        for (int w = 0; w < NV; w++) fv[w] = R(0);
    }
    if constexpr ((VM & 2) != 0) {
        const int lo = fam == 0 ? 3 : 0, hi = fam == 0 ? 1 : 2;   // sides 1 and 2 of the family
        const double x = fam == 0 ? xf[ifp] : xf[jfp];
        const R mu_art = R(mu4[lo] * (1.0 - x) + mu4[hi] * x);
        vflux<R>(gam, lam, prandt, mu_art, R(0), qv, g, Fv, Gv);
#pragma unroll
        for (int w = 0; w < NV; w++) fv[w] = fv[w] + Fv[w] * s1 + Gv[w] * s2;
    }
}

// ---------------------------------------------------------------------------
// Two threads per solution point (V2Shape): the flux phase has one task per
// flux point, the derivative phase splits the x (group 0) and y (group 1)
// components of nablaQs.
template <int N, class R, int VM, int NG = 4>
__global__ void __launch_bounds__(V2Shape<N>::THREADS) __attribute__((amdgpu_waves_per_eu(CHORUS_VISC_MINW(N)))) k_grad(ViscArgs a) {
    using S = VShape<N>;
    using T = V2Shape<N>;
    constexpr int CPB = S::CPB, NP = S::NP, NF = S::NF, NQ = S::NQ;
    constexpr int NIT = T::NIT;
    constexpr int VST = 10 * NF + 1;     // per-cell stride: Qv (8 planes) + metric rows (2 planes)
    __shared__ double sQ[CPB * NQ];
    __shared__ double sX[CPB * 2 * NG];
    __shared__ R sV[CPB * VST];          // [c][v | s1 | s2][line][r], one family at a time
    __shared__ R sG[CPB * 16 * NP];      // nablaQs [c][d*8+v][p]
    __shared__ int2 sNbr[CPB * 4];       // face neighbours of the block's cells
    __shared__ double sMu[CPB * 4];      // muAVfi/fj(1,side) per local face
    __shared__ double sOps[Ops64Layout<N>::SIZE];   // FP64 operator tables (OpTables64 layout)
    constexpr int OM = Ops64Layout<N>::M, OXF = Ops64Layout<N>::XF, OXS = Ops64Layout<N>::XS;
    const int tid = threadIdx.x;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = blk * CPB;
    const int ncl = min(CPB, a.ncell - c0);
    {
        constexpr int PQ = per_thread<CPB * NQ, T::THREADS>(), PX = per_thread<CPB * 2 * NG, T::THREADS>();
        constexpr int PN = per_thread<CPB * 4, T::THREADS>();
        double hq[PQ], hx[PX];
        int2 hn[PN];
        gather_regs<PQ, T::THREADS>(hq, a.Qin + (size_t)c0 * NQ, ncl * NQ);
        gather_regs<PX, T::THREADS>(hx, a.cellxy + (size_t)c0 * 2 * NG, ncl * 2 * NG);
        gather_regs<PN, T::THREADS>(hn, a.nbr + (size_t)c0 * 4, ncl * 4);
        OpTables64<N, T::THREADS> ops;
        ops.gather();
        scatter_lds<PQ, T::THREADS>(sQ, hq, ncl * NQ);
        scatter_lds<PX, T::THREADS>(sX, hx, ncl * 2 * NG);
        scatter_lds<PN, T::THREADS>(sNbr, hn, ncl * 4);
        ops.scatter(sOps);
    }
    __syncthreads();
    // face averages of muAV(1) of every local face of the block's cells
    // (compAVflux / interfaceAVflux / cyclocAVflux, AV.f90:199-303, 494-564,
    // with epsilon0 of AV.f90:145-148 from this stage's wave-speed maximum);
    // k_face_visc forms the same average at the face points
    if constexpr ((VM & 2) != 0) {
        for (int i = tid; i < ncl * 4; i += blockDim.x) {
            const int cell = c0 + i / 4;
            const int2 info = sNbr[i];
            const int in = a.inst ? a.inst[cell] : 0;
            const double eps0 = a.eig[in] * (a.inst ? a.hinst[in] : a.h_msh) / (double)(N - 1) * 1.0;
            const double mo = mu_av(a.sm[cell], eps0, a.pi);
            sMu[i] = ((info.y >> 4) & 1) ? mo : 0.5 * (mo + mu_av(a.sm[info.x], eps0, a.pi));
            if ((i & 3) == 0) a.mucell[cell] = mo;   // for the viscous face kernel (and the output)
        }
    }
    const int grp = tid >= T::G1 ? 1 : 0;
    const int pt = tid - grp * T::G1;
    const int pc = pt / NP, pk = pt % NP, pjs = pk / N, pis = pk % N;
    const bool pact = (pt < CPB * NP) && (pc < ncl);
    R gp[8];
#pragma unroll 1
    for (int fam = 0; fam < 2; fam++) {
        for (int t = tid; t < T::NI + T::NFC; t += blockDim.x) {
            if (t < T::NI) {
                // interior flux point: Qv = own trace
                const int c = t / NIT, k = t % NIT;
                if (c >= ncl) continue;
                const int line = k / (N - 1), r = 1 + k % (N - 1);
                const int ifp = fam == 0 ? r : line, jfp = fam == 0 ? line : r;
                double s1d, s2d;
                metric_row_g<N, NG>(sX + c * 2 * NG, fam, ifp, jfp, s1d, s2d, sOps + OXF, sOps + OXS);
                R Qown[NV];
                trace9<N, R>(sQ + c * NQ, fam, ifp, jfp, Qown, sOps);
                R *q = sV + c * VST + line * (N + 1) + r;
#pragma unroll
                for (int v = 0; v < 8; v++) q[v * NF] = Qown[v];
                q[8 * NF] = R(s1d);
                q[9 * NF] = R(s2d);
            } else {
                // face point: central average of both stored traces (a FREE
                // face averages the own trace with itself, bcflux.f90:55-61)
                const int item = t - T::NI;
                const int c = item / (2 * N), k = item % (2 * N);
                if (c >= ncl) continue;
                const int line = k >> 1, r = (k & 1) ? N : 0;
                const int cell = c0 + c;
                const int ifp = fam == 0 ? r : line, jfp = fam == 0 ? line : r;
                double s1d, s2d;
                metric_row_g<N, NG>(sX + c * 2 * NG, fam, ifp, jfp, s1d, s2d, sOps + OXF, sOps + OXS);
                int lfi, nfp;
                face_of<N>(fam, r, line, lfi, nfp);
                const int2 info = sNbr[c * 4 + lfi];
                const double *own = a.TQ + (size_t)((cell * 4 + lfi) * N + nfp) * NV;
                const double *oth = ((info.y >> 4) & 1) ? own
                                    : a.TQ + (size_t)((info.x * 4 + (info.y & 7) - 1) * N + partner_nfp<N>(nfp)) * NV;
                R *q = sV + c * VST + line * (N + 1) + r;
#pragma unroll
                for (int v = 0; v < 8; v++) q[v * NF] = R(0.5) * R(own[v]) + R(0.5) * R(oth[v]);
                q[8 * NF] = R(s1d);
                q[9 * NF] = R(s2d);
            }
        }
        __syncthreads();
        // nablaQs (nablaQsp, compflux.f90:107-157); group 0: x, group 1: y
        if (pact) {
            const int line = fam == 0 ? pjs : pis, pos = fam == 0 ? pis : pjs;
            const R *q = sV + pc * VST + line * (N + 1);
            const R *mq = q + (8 + grp) * NF;
            R acc[8];
#pragma unroll
            for (int v = 0; v < 8; v++) acc[v] = R(0);
            constexpr int U = contraction_unroll<R, N>();
#pragma unroll U
            for (int r = 0; r <= N; r++) {
                const R m = R(sOps[OM + r * N + pos]);
                const R mm = mq[r];
#pragma unroll
                for (int v = 0; v < 8; v++) {
                    const R t = q[v * NF + r] * m;
                    acc[v] = acc[v] + t * mm;
                }
            }
            if (fam == 0) {
#pragma unroll
                for (int v = 0; v < 8; v++) gp[v] = acc[v];
            } else {
                const R J = R(jacobian_g<N, NG>(sX + pc * 2 * NG, pis, pjs, sOps + OXS));
#pragma unroll
                for (int v = 0; v < 8; v++) sG[(pc * 16 + grp * 8 + v) * NP + pk] = (gp[v] + acc[v]) / J;
            }
        }
        __syncthreads();
    }
    // own traces of nablaQs at the four faces (interfaceVISflux, interfaceflux_all.f90:499-536):
    // a thread per face point and half of the 16 components, the L row in registers
    for (int it = tid; it < ncl * 4 * N * 2; it += blockDim.x) {
        const int c = it / (8 * N), k = it % (8 * N), fpi = k >> 1, h = k & 1;
        const int lfi = fpi / N, nfp = fpi % N;
        int ifp, jfp;
        face2fp<N>(lfi + 1, nfp, ifp, jfp);
        const bool fam0 = lf_family(lfi + 1) == 0;
        const int lrow = fam0 ? ifp : jfp, base = fam0 ? jfp * N : ifp, stride = fam0 ? 1 : N;
        R Lr[N];
#pragma unroll
        for (int sp = 0; sp < N; sp++) Lr[sp] = R(sOps[lrow * N + sp]);
        R *dst = a.gtr_r<R>() + (size_t)((c0 + c) * 4 * N + fpi) * 16 + h * 8;
#pragma unroll
        for (int gg = 0; gg < 8; gg++) {
            const R *gq = sG + (c * 16 + h * 8 + gg) * NP + base;
            R acc = R(0);
#pragma unroll
            for (int sp = 0; sp < N; sp++) acc = acc + gq[sp * stride] * Lr[sp];
            dst[gg] = acc;
        }
    }
    // viscous / AV flux of the own traces at the interior flux points of both
    // families (comp_viscous_flux, compflux.f90:205-256), for the stage kernel
    // (tasks family-major, so that a wave's tasks share a family: those of
    // family 1 read the traces and metric rows its flux phase left in sV)
    R *fvi = a.fvi_r<R>();
    for (int it = tid; it < 2 * CPB * NIT; it += blockDim.x) {
        const int fam = it / (CPB * NIT), k1 = it % (CPB * NIT), c = k1 / NIT, k = k1 % NIT;
        if (c >= ncl) continue;
        const int line = k / (N - 1), r = 1 + k % (N - 1);
        const int ifp = fam == 0 ? r : line, jfp = fam == 0 ? line : r;
        R Qown[NV], g[2 * NV], fv[NV], s1, s2;
        if (fam == 1) {
            const R *q = sV + c * VST + line * (N + 1) + r;
#pragma unroll
            for (int v = 0; v < 8; v++) Qown[v] = q[v * NF];
            Qown[8] = R(0);   // (A_z does not enter the viscous flux)
            s1 = q[8 * NF];
            s2 = q[9 * NF];
        } else {
            double s1d, s2d;
            metric_row_g<N, NG>(sX + c * 2 * NG, fam, ifp, jfp, s1d, s2d, sOps + OXF, sOps + OXS);
            s1 = R(s1d);
            s2 = R(s2d);
            trace9<N, R>(sQ + c * NQ, fam, ifp, jfp, Qown, sOps);
        }
        grad_trace<N, R>(sG + c * 16 * NP, fam, ifp, jfp, g, sOps);
        visc_point<N, R, VM>(a, sMu + c * 4, fam, ifp, jfp, s1, s2, Qown, g, fv, sOps + OXF);
        R *dst = fvi + (size_t)((c0 + c) * 2 + fam) * (NV * NIT) + k;
#pragma unroll
        for (int w = 0; w < NV; w++) dst[w * NIT] = fv[w];
    }
}

// ---------------------------------------------------------------------------
// Viscous and AV flux of every face point, once per face (interfaceVISflux /
// cyclocVISflux / BCVISFLUX with interfaceAVflux, compflux.f90:198-258,
// AV.f90:199-303, 494-564): Qv and nablaQ averaged over both stored traces
// (a FREE face averages its own), muAV(1) averaged over both cells, the
// Cartesian fluxes of getVfluxvectors / getVfluxvectors_AV evaluated once and
// contracted with each cell's own metric row in visc_point's order.  At a face
// point the flux-point coordinate of mu_art's interpolation is exactly 0 or 1,
// so mu_art is the face average itself.
// ---------------------------------------------------------------------------
template <int N, class R, int VM, int NG = 4>
__global__ void __launch_bounds__(256) k_face_visc(FaceArgs a) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= a.nface * N) return;
    const int f = t / N, k = t % N;
    const int4 fc = a.faces[f];
    const bool free_face = fc.z < 0;
    const int kr = N - 1 - k;
    // own metric rows of both cells at the point
    R sl1, sl2, sr1 = R(0), sr2 = R(0);
    {
        int ifp, jfp;
        double xy[2 * NG], m1, m2;
        face2fp<N>(fc.y + 1, k, ifp, jfp);
#pragma unroll
        for (int q = 0; q < 2 * NG; q++) xy[q] = a.cellxy[(size_t)fc.x * (2 * NG) + q];
        metric_row_g<N, NG>(xy, lf_family(fc.y + 1), ifp, jfp, m1, m2);
        sl1 = R(m1); sl2 = R(m2);
        if (!free_face) {
            face2fp<N>(fc.w + 1, kr, ifp, jfp);
#pragma unroll
            for (int q = 0; q < 2 * NG; q++) xy[q] = a.cellxy[(size_t)fc.z * (2 * NG) + q];
            metric_row_g<N, NG>(xy, lf_family(fc.w + 1), ifp, jfp, m1, m2);
            sr1 = R(m1); sr2 = R(m2);
        }
    }
    const size_t pl = (size_t)((fc.x * 4 + fc.y) * N + k);
    const size_t pr = free_face ? pl : (size_t)((fc.z * 4 + fc.w) * N + kr);
    R qv[NV], g[2 * NV];
#pragma unroll
    for (int v = 0; v < 8; v++) qv[v] = R(0.5) * R(a.TQ[pl * NV + v]) + R(0.5) * R(a.TQ[pr * NV + v]);
    qv[8] = R(0);
    const R *gl = (const R *)a.gtr + pl * 16, *gr = (const R *)a.gtr + pr * 16;
#pragma unroll
    for (int d = 0; d < 2; d++) {
#pragma unroll
        for (int w = 0; w < 8; w++) g[d * NV + w] = R(0.5) * gl[d * 8 + w] + R(0.5) * gr[d * 8 + w];
        g[d * NV + 8] = R(0);
    }
    const R gam = R(a.gam), lam = R(-2.0 / 3.0), prandt = R(0.72);
    R fl[NV], fr[NV], Fv[NV], Gv[NV];
    if constexpr ((VM & 1) != 0) {
        vflux<R>(gam, lam, prandt, R(a.mu), R(a.eta0), qv, g, Fv, Gv);
#pragma unroll
        for (int w = 0; w < NV; w++) { fl[w] = Fv[w] * sl1 + Gv[w] * sl2; fr[w] = Fv[w] * sr1 + Gv[w] * sr2; }
    } else {
#pragma unroll
        for (int w = 0; w < NV; w++) { fl[w] = R(0); fr[w] = R(0); }
    }
    if constexpr ((VM & 2) != 0) {
        // muAV(1) of both cells (mu_av of this stage's epsilon0, kept by k_grad)
        const double ml = a.mucell[fc.x];
        const double mr = free_face ? ml : a.mucell[fc.z];
        const double mf = free_face ? ml : 0.5 * (ml + mr);
        vflux<R>(gam, lam, prandt, R(mf), R(0), qv, g, Fv, Gv);
#pragma unroll
        for (int w = 0; w < NV; w++) {
            fl[w] = fl[w] + Fv[w] * sl1 + Gv[w] * sl2;
            fr[w] = fr[w] + Fv[w] * sr1 + Gv[w] * sr2;
        }
    }
    R *FFv = (R *)a.FFv;
#pragma unroll
    for (int w = 0; w < NV; w++) FFv[pl * NV + w] = fl[w];
    if (!free_face) {
#pragma unroll
        for (int w = 0; w < NV; w++) FFv[pr * NV + w] = fr[w];
    }
}

// ---------------------------------------------------------------------------
template <int N, class R, int VM, int NG = 4>
__global__ void __launch_bounds__(V2Shape<N>::THREADS) __attribute__((amdgpu_waves_per_eu(CHORUS_VISC_MINW(N)))) k_stage_visc(StageArgs a, ViscArgs v) {
    using S = VShape<N>;
    using T = V2Shape<N>;
    constexpr int CPB = S::CPB, NP = S::NP, NF = S::NF, NQ = S::NQ;
    constexpr int NIT = T::NIT;
    constexpr int UST = 20 * NF + 1;     // per-cell stride: F (8) + Fv (9) + A_z + metric rows (2)
    __shared__ double sQ[CPB * NQ];
    __shared__ double sX[CPB * 2 * NG];
    __shared__ R sU[CPB * UST];          // [c][plane][line][r], one family at a time
    __shared__ double sAV[CPB * NP];     // AV indicator work
    __shared__ double sOps[Ops64Layout<N>::SIZE];   // FP64 operator tables (OpTables64 layout)
    constexpr int OM = Ops64Layout<N>::M, OXF = Ops64Layout<N>::XF, OXS = Ops64Layout<N>::XS;
    static_assert(9 * CPB * NP <= CPB * UST, "exchange area");
    const int tid = threadIdx.x;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = blk * CPB;
    const int ncl = min(CPB, a.ncell - c0);
    {
        constexpr int PQ = per_thread<CPB * NQ, T::THREADS>(), PX = per_thread<CPB * 2 * NG, T::THREADS>();
        double hq[PQ], hx[PX];
        gather_regs<PQ, T::THREADS>(hq, a.Qin + (size_t)c0 * NQ, ncl * NQ);
        gather_regs<PX, T::THREADS>(hx, a.cellxy + (size_t)c0 * 2 * NG, ncl * 2 * NG);
        OpTables64<N, T::THREADS> ops;
        ops.gather();
        scatter_lds<PQ, T::THREADS>(sQ, hq, ncl * NQ);
        scatter_lds<PX, T::THREADS>(sX, hx, ncl * 2 * NG);
        ops.scatter(sOps);
    }
    if constexpr ((VM & 2) != 0) {
        if (blockIdx.x == 0)
            for (int i = tid; i < v.ninst; i += blockDim.x) v.eig_next[i] = 0.0;
    }
    __syncthreads();
    // the By indicator of the input state, for the Tecplot output's muAV(6)
    // (AV.f90:60-80 for variable 6)
    if constexpr ((VM & 2) != 0) {
        if (v.sm6) av_sensor<N, CPB>(v.sm6, sQ + 5 * NP, sAV, ncl, c0, sOps);
    }
    // derivative-phase role of this thread
    const int grp = tid >= T::G1 ? 1 : 0;
    const int pt = tid - grp * T::G1;
    const int pc = pt / NP, pk = pt % NP, pjs = pk / N, pis = pk % N;
    const bool pact = (pt < CPB * NP) && (pc < ncl);
    const R *FF = (const R *)a.FF, *FFv = (const R *)a.FFv, *FVI = v.fvi_r<R>();

    R acc8[8], accv[NV], dA1 = R(0), dA2 = R(0);   // group 0: acc8 = dF; group 1: accv = dFv
    const size_t base = (size_t)(c0 + pc) * NQ + pk;
#pragma unroll 1
    for (int fam = 0; fam < 2; fam++) {
        for (int t = tid; t < T::BT; t += blockDim.x) {
            int kind, item;
            if (t < T::NI) { kind = 0; item = t; }
            else if (t < T::S1) { kind = 2; item = t - T::NI; }
            else if (t < T::S1 + T::NI) { kind = 1; item = t - T::S1; }
            else { kind = 2; item = (T::S1 - T::NI) + (t - T::S1 - T::NI); }
            if (kind == 0) {
                // interior flux point: inviscid flux of the own trace
                const int c = item / NIT, k = item % NIT;
                if (c >= ncl) continue;
                const int line = k / (N - 1), r = 1 + k % (N - 1);
                const int ifp = fam == 0 ? r : line, jfp = fam == 0 ? line : r;
                double s1d, s2d;
                metric_row_g<N, NG>(sX + c * 2 * NG, fam, ifp, jfp, s1d, s2d, sOps + OXF, sOps + OXS);
                R Qown[NV];
                trace9<N, R>(sQ + c * NQ, fam, ifp, jfp, Qown, sOps);
                R *u = sU + c * UST + line * (N + 1) + r;
                const R s1 = R(s1d), s2 = R(s2d);
                R F[8], G[8];
                iflux<R>(R(a.gam), Qown, F, G);
#pragma unroll
                for (int w = 0; w < 8; w++) u[w * NF] = F[w] * s1 + G[w] * s2;
                u[17 * NF] = Qown[8];
                u[18 * NF] = s1;
                u[19 * NF] = s2;
            } else if (kind == 1) {
                // interior flux point: the viscous / AV flux formed by k_grad
                const int c = item / NIT, k = item % NIT;
                if (c >= ncl) continue;
                const int line = k / (N - 1), r = 1 + k % (N - 1);
                const R *src = FVI + (size_t)((c0 + c) * 2 + fam) * (NV * NIT) + k;
                R *u = sU + c * UST + line * (N + 1) + r;
#pragma unroll
                for (int w = 0; w < NV; w++) u[(8 + w) * NF] = src[w * NIT];
            } else {
                // face point: the stored common inviscid, A_z and viscous fluxes
                const int c = item / (2 * N), k = item % (2 * N);
                if (item >= T::NFC || c >= ncl) continue;
                const int line = k >> 1, r = (k & 1) ? N : 0;
                const int ifp = fam == 0 ? r : line, jfp = fam == 0 ? line : r;
                int lfi, nfp;
                face_of<N>(fam, r, line, lfi, nfp);
                const size_t o = (size_t)(((c0 + c) * 4 + lfi) * N + nfp) * NV;
                double s1d, s2d;
                metric_row_g<N, NG>(sX + c * 2 * NG, fam, ifp, jfp, s1d, s2d, sOps + OXF, sOps + OXS);
                R *u = sU + c * UST + line * (N + 1) + r;
#pragma unroll
                for (int w = 0; w < 8; w++) u[w * NF] = FF[o + w];
                u[17 * NF] = FF[o + 8];
#pragma unroll
                for (int w = 0; w < NV; w++) u[(8 + w) * NF] = FFv[o + w];
                u[18 * NF] = R(s1d);
                u[19 * NF] = R(s2d);
            }
        }
        __syncthreads();
        if (pact) {
            const int line = fam == 0 ? pjs : pis, pos = fam == 0 ? pis : pjs;
            const R *f = sU + pc * UST + line * (N + 1);
            constexpr int U = contraction_unroll<R, N>();
            if (grp == 0) {
                R d[8], e1 = R(0), e2 = R(0);
#pragma unroll
                for (int w = 0; w < 8; w++) d[w] = R(0);
#pragma unroll U
                for (int r = 0; r <= N; r++) {
                    const R m = R(sOps[OM + r * N + pos]);
#pragma unroll
                    for (int w = 0; w < 8; w++) d[w] = d[w] + f[w * NF + r] * m;
                }
#pragma unroll U
                for (int r = 0; r <= N; r++) e1 = e1 + f[17 * NF + r] * R(sOps[OM + r * N + pos]) * f[18 * NF + r];
#pragma unroll U
                for (int r = 0; r <= N; r++) e2 = e2 + f[17 * NF + r] * R(sOps[OM + r * N + pos]) * f[19 * NF + r];
                if (fam == 0) {
#pragma unroll
                    for (int w = 0; w < 8; w++) acc8[w] = d[w];
                    dA1 = e1; dA2 = e2;
                } else {
#pragma unroll
                    for (int w = 0; w < 8; w++) acc8[w] = -(acc8[w] + d[w]);
                    dA1 = dA1 + e1; dA2 = dA2 + e2;
                }
            } else {
                R dv[NV];
#pragma unroll
                for (int w = 0; w < NV; w++) dv[w] = R(0);
#pragma unroll U
                for (int r = 0; r <= N; r++) {
                    const R m = R(sOps[OM + r * N + pos]);
#pragma unroll
                    for (int w = 0; w < NV; w++) dv[w] = dv[w] + f[(8 + w) * NF + r] * m;
                }
#pragma unroll
                for (int w = 0; w < NV; w++) accv[w] = fam == 0 ? dv[w] : accv[w] + dv[w];
            }
        }
        __syncthreads();
    }
    // the residual and the stage combination are split over both groups
    // (group 0: variables 0-3 and A_z, group 1: variables 4-7); each hands the
    // sums of the other group's variables through the (now free) flux area
    R *xv = sU;
    if (pact) {
        if (grp == 0) {
#pragma unroll
            for (int w = 4; w < 8; w++) xv[w * CPB * NP + pt] = acc8[w];
        } else {
#pragma unroll
            for (int w = 0; w < 4; w++) xv[w * CPB * NP + pt] = accv[w];
            xv[8 * CPB * NP + pt] = accv[8];
        }
    }
    // the point's Qa, loaded across the barrier
    double qa[5];
    if (pact) {
        const bool need = a.kind != KIND_RESID && a.kind != KIND_AXPY;
        const int vg = 4 * grp;
#pragma unroll
        for (int k = 0; k < 4; k++) qa[k] = need ? a.Qa[base + (size_t)(vg + k) * NP] : 0.0;
        qa[4] = need && grp == 0 ? a.Qa[base + (size_t)8 * NP] : 0.0;
    }
    __syncthreads();
    if (pact) {
        // residual at the solution point (compresid.f90:12-65)
        const R J = R(jacobian_g<N, NG>(sX + pc * 2 * NG, pis, pjs, sOps + OXS));
        double *qs = sQ + pc * NQ + pk;
        if (grp == 0) {
            const R q0 = R(qs[0]), q1 = R(qs[NP]), q2 = R(qs[2 * NP]);
            const R au = q1 / q0, av = q2 / q0;
            const R nA1 = dA1 / J, nA2 = dA2 / J;
            R res[4];
#pragma unroll
            for (int w = 0; w < 4; w++) res[w] = (acc8[w] + xv[w * CPB * NP + pt]) / J;
            R r8 = -(au * nA1) - av * nA2;
            r8 = r8 * J;
            r8 = (r8 + xv[8 * CPB * NP + pt]) / J;
            combine_vars<4>(a, base, NP, qs, 0, qa, res);
            combine_vars<1>(a, base, NP, qs, 8, qa + 4, res, r8);
        } else {
            R res[4];
#pragma unroll
            for (int w = 0; w < 4; w++) res[w] = (xv[(4 + w) * CPB * NP + pt] + accv[4 + w]) / J;
            combine_vars<4>(a, base, NP, qs, 4, qa, res);
        }
    }
    if (a.kind == KIND_RESID) return;
    __syncthreads();
    // traces of the new state for the next stage's face passes, and its AV
    // indicator (AV.f90:15-82) for the next stage
    // (the flux planes are free now: they stage the trace stores)
    constexpr bool STG = trace_stage_doubles<N>(CPB) * sizeof(double) <= CPB * UST * sizeof(R);
    store_traces<N>(sQ, ncl, c0, a.TQ, sOps, STG ? reinterpret_cast<double *>(sU) : nullptr);
    if constexpr ((VM & 2) != 0) av_sensor<N, CPB>(a.sm, sQ, sAV, ncl, c0, sOps);
}

}  // namespace chorus
