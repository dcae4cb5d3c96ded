// Stage, constrained-transport and diagnostic kernels for gfx950.
//
// Included by two translation units: chorus_fp64.hip (Real = double, built
// with -ffp-contract=off so every operation rounds as in the reference) and
// chorus_mixed.hip (Real = float for the flux/contraction arithmetic).
//
// Data layout in HBM (one cell = one contiguous block):
//   Q[cell][var][js][is]       FP64 state, 9*N*N doubles per cell
//   cellxy[cell][8]            vertex coordinates x1..x4, y1..y4 (IVCELL order)
//   nbr[cell][lf] (int2)       .x neighbour cell across local face lf (-1: FREE)
//                              .y bits 0-2 neighbour's local face (1..4),
//                                 bit 3 this cell is the reference's RIGHT cell,
//                                 bit 4 FREE boundary face
// A workgroup owns CPB whole cells: it stages their Q in LDS with contiguous
// loads, evaluates every flux point of its cells (face points read the
// neighbour's trace through L2), then every solution point.
#pragma once

#include "physics.hpp"
#include "chorus_internal.hpp"

namespace chorus {

// operator tables per N, one private copy per translation unit (no
// relocatable device code); filled by upload_tables() from host_operators()
static __constant__ double c_L[kMaxN + 1][(kMaxN + 1) * kMaxN];   // L(ifp,is) at [ifp*N+is]
static __constant__ double c_M[kMaxN + 1][(kMaxN + 1) * kMaxN];   // M(r,is) at [r*N+is]
static __constant__ double c_Xs[kMaxN + 1][kMaxN];
static __constant__ double c_Xf[kMaxN + 1][kMaxN + 1];
static __constant__ float c_Lf[kMaxN + 1][(kMaxN + 1) * kMaxN];
static __constant__ float c_Mf[kMaxN + 1][(kMaxN + 1) * kMaxN];
// AV interpolation tables per N: H1[isl*N+is] = hhval(is, Xs_low(isl)),
// H2[is*(N-1)+isl] = hhval_low(isl, Xs(is))  (AV.f90:19-58)
static __constant__ double c_H1[kMaxN + 1][kMaxN * kMaxN];
static __constant__ double c_H2[kMaxN + 1][kMaxN * kMaxN];

static hipError_t upload_tables() {
    static double L[kMaxN + 1][(kMaxN + 1) * kMaxN], M[kMaxN + 1][(kMaxN + 1) * kMaxN];
    static double Xs[kMaxN + 1][kMaxN], Xf[kMaxN + 1][kMaxN + 1];
    static float Lf[kMaxN + 1][(kMaxN + 1) * kMaxN], Mf[kMaxN + 1][(kMaxN + 1) * kMaxN];
    for (int n = 2; n <= kMaxN; n++) {
        host_operators(n, Xs[n], Xf[n], L[n], M[n]);
        for (int i = 0; i < (n + 1) * n; i++) { Lf[n][i] = float(L[n][i]); Mf[n][i] = float(M[n][i]); }
    }
    hipError_t e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_L), L, sizeof(L))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_M), M, sizeof(M))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_Xs), Xs, sizeof(Xs))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_Xf), Xf, sizeof(Xf))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_Lf), Lf, sizeof(Lf))) != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_Mf), Mf, sizeof(Mf));
}

// a workgroup of about TB threads owns CPB whole cells, one thread per solution point
template <int N, int TB> struct CellShape {
    static constexpr int NP = N * N;          // solution points per cell
    static constexpr int NF = N * (N + 1);    // flux points per family
    static constexpr int NQ = NV * NP;        // doubles of state per cell
    static constexpr int CPB = (TB / NP) > 0 ? (TB / NP) : 1;   // cells per block
    static constexpr int THREADS = ((CPB * NP + 63) / 64) * 64;
};
template <int N> using Shape = CellShape<N, 256>;
// a block of whole cells with one thread per flux point of a family (N(N+1)
// per cell): a kernel whose flux phase evaluates every flux point runs it in
// one round; its derivative phase uses the first CPB*N^2 threads
template <int N, int TB> struct FluxCellShape {
    static constexpr int NP = N * N, NF = N * (N + 1), NQ = NV * NP;
    static constexpr int CPB = (TB / NF) > 0 ? (TB / NF) : 1;
    static constexpr int THREADS = ((CPB * NF + 63) / 64) * 64;
};
#ifndef CHORUS_IDEAL_TB
#define CHORUS_IDEAL_TB 256
#endif
template <int N> using IShape = CellShape<N, CHORUS_IDEAL_TB>;

// XCD-aware block order: the hardware deals consecutive workgroups round-robin
// over the 8 XCDs; give each XCD a contiguous run of cell blocks so the
// neighbour traces a block reads were staged by a block on the same L2.
__device__ __forceinline__ int xcd_block(int b, int nblk) {
    int q = nblk >> 3, r = nblk & 7;
    int x = b & 7, k = b >> 3;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Block-cooperative copy of n <= PER*THREADS consecutive elements from global
// memory: gather_regs issues every load of the thread into registers before
// any use, scatter_lds then writes them to LDS.  A strided copy loop would
// wait for each element's load before the next one is issued (one HBM
// latency per element a thread copies); this costs one.  blockDim.x must be
// THREADS.
template <int PER, int THREADS, class T>
__device__ __forceinline__ void gather_regs(T *v, const T *src, int n) {
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = (int)threadIdx.x + k * THREADS;
        v[k] = i < n ? src[i] : T{};
    }
}
template <int PER, int THREADS, class T>
__device__ __forceinline__ void scatter_lds(T *dst, const T *v, int n) {
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = (int)threadIdx.x + k * THREADS;
        if (i < n) dst[i] = v[k];
    }
}
template <int NMAX, int THREADS> constexpr int per_thread() { return (NMAX + THREADS - 1) / THREADS; }

// FP64 extrapolation / derivative operators and point tables of order N staged
// in LDS (per-thread indices would otherwise make every coefficient a vector
// load from the constant tables): gather before the block's first barrier,
// scatter after it.  Layout of the LDS copy: L[(N+1)N], M[(N+1)N], Xf[N+1], Xs[N],
// and the AV interpolation tables H1[(N-1)N], H2[N(N-1)].
template <int N> struct Ops64Layout {
    static constexpr int K = (N + 1) * N, KH = (N - 1) * N;
    static constexpr int L = 0, M = K, XF = 2 * K, XS = 2 * K + N + 1, H1 = XS + N, H2 = H1 + KH, SIZE = H2 + KH;
};
template <int N, int THREADS> struct OpTables64 {
    using Y = Ops64Layout<N>;
    static constexpr int PO = (Y::SIZE + THREADS - 1) / THREADS;
    double v[PO];
    // one load per element from a selected address (a chain of conditional
    // loads into one register would wait for each load before the next)
    __device__ __forceinline__ void gather() {
#pragma unroll
        for (int k = 0; k < PO; k++) {
            const int i = (int)threadIdx.x + k * THREADS;
            const double *src = i < Y::M ? &c_L[N][i] : i < Y::XF ? &c_M[N][i - Y::M]
                              : i < Y::XS ? &c_Xf[N][i - Y::XF] : i < Y::H1 ? &c_Xs[N][i - Y::XS]
                              : i < Y::H2 ? &c_H1[N][i - Y::H1] : &c_H2[N][min(i, Y::SIZE - 1) - Y::H2];
            v[k] = *src;
        }
    }
    __device__ __forceinline__ void scatter(double *sOps) const { scatter_lds<PO, THREADS>(sOps, v, Y::SIZE); }
};

// local face -> flux-point index of face point nfp (face2fpconnec, initsetup.f90:343-379)
template <int N> __device__ __forceinline__ void face2fp(int lf, int nfp, int &ifp, int &jfp) {
    switch (lf) {
    case 1: ifp = nfp; jfp = 0; break;
    case 2: ifp = N; jfp = nfp; break;
    case 3: ifp = N - 1 - nfp; jfp = N; break;
    default: ifp = 0; jfp = N - 1 - nfp; break;
    }
}

// face family of a local face (interfaceflux_all.f90:26-27): 0 = xi (F1), 1 = eta (G2)
__device__ __forceinline__ int lf_family(int lf) { return (lf & 1) ? 1 : 0; }

// contravariant metric row used by the flux at a point of family fam
// (calcjacob.f90:139-170): fam 0 -> (S1(1,1), S1(1,2)) = (yeta, -xeta) at
// (Xf(ifp), Xs(jfp)); fam 1 -> (S2(2,1), S2(2,2)) = (-ypsi, xpsi) at (Xs(ifp), Xf(jfp))
// Geometry of a cell: NG = 4 corner vertices (x1..x4, y1..y4: the bilinear map
// of mapder.f90:1-76) or NG = 8 nodes (x1..x8, y1..y8: corners then edge
// nodes, the 8-node map of MAPDER_8_NODE, mapder.f90:117-253)
template <int NG>
__device__ __forceinline__ void map_derivs_g(const double *xy, double psi, double eta, double &xpsi, double &ypsi,
                                             double &xeta, double &yeta) {
    if constexpr (NG == 8) map_derivs8(xy, xy + 8, psi, eta, xpsi, ypsi, xeta, yeta);
    else map_derivs(xy, xy + 4, psi, eta, xpsi, ypsi, xeta, yeta);
}

template <int N, int NG>
__device__ __forceinline__ void metric_row_g(const double *xy, int fam, int ifp, int jfp, double &s1, double &s2,
                                             const double *Xf = nullptr, const double *Xs = nullptr) {
    double xpsi, ypsi, xeta, yeta;
    if (fam == 0) {
        map_derivs_g<NG>(xy, Xf ? Xf[ifp] : c_Xf[N][ifp], Xs ? Xs[jfp] : c_Xs[N][jfp], xpsi, ypsi, xeta, yeta);
        s1 = yeta; s2 = -xeta;
    } else {
        map_derivs_g<NG>(xy, Xs ? Xs[ifp] : c_Xs[N][ifp], Xf ? Xf[jfp] : c_Xf[N][jfp], xpsi, ypsi, xeta, yeta);
        s1 = -ypsi; s2 = xpsi;
    }
}

template <int N, int NG>
__device__ __forceinline__ double jacobian_g(const double *xy, int is, int js, const double *Xs = nullptr) {
    double xpsi, ypsi, xeta, yeta;
    map_derivs_g<NG>(xy, Xs ? Xs[is] : c_Xs[N][is], Xs ? Xs[js] : c_Xs[N][js], xpsi, ypsi, xeta, yeta);
    return xpsi * yeta - xeta * ypsi;
}

template <int N>
// (Xf, Xs: LDS copies of the point tables, or null for the constant tables)
__device__ __forceinline__ void metric_row(const double *xy, int fam, int ifp, int jfp, double &s1,
                                           double &s2, const double *Xf = nullptr, const double *Xs = nullptr) {
    metric_row_g<N, 4>(xy, fam, ifp, jfp, s1, s2, Xf, Xs);
}

template <int N>
__device__ __forceinline__ double jacobian(const double *xy, int is, int js, const double *Xs = nullptr) {
    return jacobian_g<N, 4>(xy, is, js, Xs);
}

template <class R> __device__ __forceinline__ R Lop(int N, int ifp, int is);
template <> __device__ __forceinline__ double Lop<double>(int N, int ifp, int is) { return c_L[N][ifp * N + is]; }
template <> __device__ __forceinline__ float Lop<float>(int N, int ifp, int is) { return c_Lf[N][ifp * N + is]; }
template <class R> __device__ __forceinline__ R Mop(int N, int r, int is);
template <> __device__ __forceinline__ double Mop<double>(int N, int r, int is) { return c_M[N][r * N + is]; }
template <> __device__ __forceinline__ float Mop<float>(int N, int r, int is) { return c_Mf[N][r * N + is]; }

// trace of all 9 variables at flux point (ifp,jfp) of family fam of one cell
// whose state block q[var][js][is] is given (inviscid_flux_interior,
// compflux.f90:59-94 and the trace loops of interfaceflux_all.f90:40-73)
// (Lt: LDS copy of the FP64 extrapolation table, or null for the constant one)
template <int N, class R>
__device__ __forceinline__ void trace9(const double *q, int fam, int ifp, int jfp, R *Qf, const double *Lt = nullptr) {
    constexpr int NP = N * N;
#pragma unroll
    for (int v = 0; v < NV; v++) Qf[v] = R(0);
    constexpr int U = contraction_unroll<R, N>();
#pragma unroll U
    for (int sp = 0; sp < N; sp++) {
        const int row = fam == 0 ? ifp : jfp;
        const R l = Lt ? R(Lt[row * N + sp]) : Lop<R>(N, row, sp);
        const int idx = fam == 0 ? jfp * N + sp : sp * N + ifp;
#pragma unroll
        for (int v = 0; v < NV; v++) Qf[v] = Qf[v] + R(q[v * NP + idx]) * l;
    }
}

template <int N>
__device__ __forceinline__ double trace1(const double *q1, int fam, int ifp, int jfp, const double *Lt = nullptr) {
    const double *L = Lt ? Lt : c_L[N];
    double a = 0.0;
    if (fam == 0) {
#pragma unroll
        for (int sp = 0; sp < N; sp++) a = a + q1[jfp * N + sp] * L[ifp * N + sp];
    } else {
#pragma unroll
        for (int sp = 0; sp < N; sp++) a = a + q1[sp * N + ifp] * L[jfp * N + sp];
    }
    return a;
}

// reference sign of a cell's contribution at its local face lfi (0..3):
// -1 on faces 1 and 4 (flux point index 0 of the family), +1 on faces 2 and 3
// (interfaceflux_all.f90:28-37)
__device__ __forceinline__ int side_sign(int lfi) { return (lfi == 0 || lfi == 3) ? -1 : 1; }

// traces of all 9 variables of the block's cells at their 4N face points, from
// the LDS copy q[c][var][js][is]; same summation order as trace9.  Called by
// all threads of the block: one thread per (cell, face point) forms the nine
// variables of its record (the point's index arithmetic and its row of L are
// shared by the nine sums).  TQ[cell][lfi][nfp][var].
// With `stage` (LDS, trace_stage_doubles(CPB) doubles, free for the call), a
// wave's 64 records are exchanged through its own part of it, so that each
// of its nine global stores writes 64 consecutive doubles; without it, each
// thread stores its own 72-byte record.
template <int N> constexpr int trace_stage_doubles(int cpb) { return (cpb * 4 * N + 63) / 64 * 64 * NV; }

template <int N>
__device__ __forceinline__ void store_traces(const double *q, int ncl, int c0, double *TQ, const double *Lt = nullptr,
                                             double *stage = nullptr) {
    constexpr int NP = N * N, NQ = NV * NP, FP = 4 * N, PT = FP * NV;
    const double *L = Lt ? Lt : c_L[N];
    const int nrec = ncl * FP;
    for (int base = 0; base < nrec; base += blockDim.x) {
        const int it = base + (int)threadIdx.x;
        double acc[NV];
        if (it < nrec) {
            const int c = it / FP, fpi = it % FP, lfi = fpi / N, nfp = fpi % N;
            int ifp, jfp;
            face2fp<N>(lfi + 1, nfp, ifp, jfp);
            const bool fam0 = lf_family(lfi + 1) == 0;
            // family 0: the line js = jfp, extrapolated to ifp; family 1: the
            // column is = ifp, extrapolated to jfp
            const double *qp = q + c * NQ + (fam0 ? jfp * N : ifp);
            const int stride = fam0 ? 1 : N;
            const double *Lrow = L + (fam0 ? ifp : jfp) * N;
            double Lr[N];
#pragma unroll
            for (int sp = 0; sp < N; sp++) Lr[sp] = Lrow[sp];
#pragma unroll
            for (int v = 0; v < NV; v++) acc[v] = 0.0;
#pragma unroll
            for (int sp = 0; sp < N; sp++) {
                const double *qs = qp + sp * stride;
#pragma unroll
                for (int v = 0; v < NV; v++) acc[v] = acc[v] + qs[v * NP] * Lr[sp];
            }
            if (!stage) {
                double *dst = TQ + (size_t)c0 * PT + (size_t)it * NV;
#pragma unroll
                for (int v = 0; v < NV; v++) dst[v] = acc[v];
            }
        }
        if (stage) {
            // the wave's records [w0, w0 + nw) leave as nw*9 consecutive doubles
            const int lane = (int)threadIdx.x & 63, w0 = base + ((int)threadIdx.x & ~63);
            const int nw = min(64, nrec - w0);
            if (nw > 0) {
                double *st = stage + (((int)threadIdx.x >> 6) << 6) * NV;
                if (it < nrec) {
#pragma unroll
                    for (int v = 0; v < NV; v++) st[lane * NV + v] = acc[v];
                }
                // the wave reads what its own lanes wrote: LDS operations of a
                // wave complete in order; keep the compiler from moving them
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                double *dst = TQ + (size_t)c0 * PT + (size_t)w0 * NV;
#pragma unroll
                for (int j = 0; j < NV; j++) {
                    const int k = j * 64 + lane;
                    if (k < nw * NV) dst[k] = st[k];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
    }
}

// The same traces in the mixed form's compact form (MxTraces): the FP64 sums
// of store_traces, less the cell's reference state ref(c, v) = ref[c*rs + v*vs]
// for variables 1-8 and rounded to FP32, A_z in FP64; the reference states of
// the block's cells are stored beside them.  One thread per record: its 32
// deviation bytes leave as two 16-byte stores, consecutive threads on
// consecutive records.  Lt: the FP64 extrapolation table (c_L layout).
template <int N>
__device__ __forceinline__ void store_traces_mx(const double *q, int ncl, int c0, const MxTraces &mt,
                                                const double *ref, int rs, int vs, const double *Lt) {
    constexpr int NP = N * N, NQ = NV * NP, FP = 4 * N;
    const int nrec = ncl * FP;
    for (int it = (int)threadIdx.x; it < nrec; it += blockDim.x) {
        const int c = it / FP, fpi = it % FP, lfi = fpi / N, nfp = fpi % N;
        int ifp, jfp;
        face2fp<N>(lfi + 1, nfp, ifp, jfp);
        const bool fam0 = lf_family(lfi + 1) == 0;
        const double *qp = q + c * NQ + (fam0 ? jfp * N : ifp);
        const int stride = fam0 ? 1 : N;
        const double *Lrow = Lt + (fam0 ? ifp : jfp) * N;
        double Lr[N], acc[NV];
#pragma unroll
        for (int sp = 0; sp < N; sp++) Lr[sp] = Lrow[sp];
#pragma unroll
        for (int v = 0; v < NV; v++) acc[v] = 0.0;
#pragma unroll
        for (int sp = 0; sp < N; sp++) {
            const double *qs = qp + sp * stride;
#pragma unroll
            for (int v = 0; v < NV; v++) acc[v] = acc[v] + qs[v * NP] * Lr[sp];
        }
        float dv[8];
#pragma unroll
        for (int v = 0; v < 8; v++) dv[v] = float(acc[v] - ref[c * rs + v * vs]);
        const size_t rec = (size_t)c0 * FP + it;
        float4 *d = reinterpret_cast<float4 *>(mt.dev + rec * 8);
        d[0] = make_float4(dv[0], dv[1], dv[2], dv[3]);
        d[1] = make_float4(dv[4], dv[5], dv[6], dv[7]);
        mt.az[rec] = acc[8];
    }
    for (int i = (int)threadIdx.x; i < ncl * 8; i += blockDim.x)
        mt.ref[(size_t)c0 * 8 + i] = ref[(i >> 3) * rs + (i & 7) * vs];
}

template <class R>
__device__ __forceinline__ void block_max_atomic(double v, double *red, double *dst) {
    const int tid = threadIdx.x, n = blockDim.x;
    red[tid] = v;
    __syncthreads();
    int p = 1;
    while (p < n) p <<= 1;
    for (int s = p >> 1; s > 0; s >>= 1) {
        if (tid < s && tid + s < n) red[tid] = fmax(red[tid], red[tid + s]);
        __syncthreads();
    }
    if (tid == 0) {
        // non-negative doubles order like their bit patterns
        atomicMax(reinterpret_cast<unsigned long long *>(dst), (unsigned long long)__double_as_longlong(red[0]));
    }
}

// maximum per instance: one atomic per wave when the wave's face points belong
// to one instance (the usual case: instances are contiguous cell ranges),
// otherwise one per lane
__device__ __forceinline__ void instance_max_atomic(double v, int inst, double *dst) {
    const int first = __builtin_amdgcn_readfirstlane(inst);
    if (__all(inst == first)) {
        for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
        if ((threadIdx.x & 63) == 0)
            atomicMax(reinterpret_cast<unsigned long long *>(dst + first), (unsigned long long)__double_as_longlong(v));
    } else {
        atomicMax(reinterpret_cast<unsigned long long *>(dst + inst), (unsigned long long)__double_as_longlong(v));
    }
}

// the same for a kernel with many waves per instance: a wave whose maximum does
// not exceed the value already stored issues no atomic (the stored value only
// grows during the kernel, so a stale read costs an atomic, never a result)
__device__ __forceinline__ void instance_max_atomic_filtered(double v, int inst, double *dst) {
    const int first = __builtin_amdgcn_readfirstlane(inst);
    const bool uniform = __all(inst == first);
    if (uniform)
        for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    if (uniform && (threadIdx.x & 63) != 0) return;
    // (non-negative doubles order like their bit patterns)
    unsigned long long *p = reinterpret_cast<unsigned long long *>(dst + (uniform ? first : inst));
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    if (b > __atomic_load_n(p, __ATOMIC_RELAXED)) atomicMax(p, b);
}

// ---------------------------------------------------------------------------
// Rusanov flux of every face point, once per face (getrusanov.f90:1-109 called
// from interfaceflux_all.f90:16-114, cyclocflux_all.f90:15-114, bcflux.f90:19-73):
// the left cell's trace and metric, the right cell's trace (a FREE face
// mirrors the left one).  The left cell receives Fnl, the right one
// (sign_l*sign_r)*Fnl; both receive the upwind A_z.  With ARTIV the block
// maximum of the wave speeds goes to the stage maximum (getrusanov.f90:105-107).
// ---------------------------------------------------------------------------
// The block's 256 face points move through LDS: their 9-value traces (left,
// right) are gathered with consecutive threads on consecutive values of a
// record, and the fluxes leave the same way, so every global access of a wave
// covers a few contiguous runs instead of 64 records 72 bytes apart.
// (record numbers fit 32 bits: ncell * 4 * N < 2^31 for every mesh a context
// can hold; only byte offsets need 64 bits)
template <int N> __device__ __forceinline__ unsigned face_rec(const int4 &fc, int k, bool right_side) {
    return (right_side && fc.z >= 0) ? unsigned((fc.z * 4 + fc.w) * N + (N - 1 - k)) : unsigned((fc.x * 4 + fc.y) * N + k);
}

// Slot of a thread in the block-wide copies of 2*NV values per face point
// (left record, right record): a fixed value w and the points pb, pb+PSTEP, ...
// so consecutive threads touch consecutive values of a record.
template <int N, int FPB> struct FaceSlots {
    static constexpr int RW = 2 * NV, PSTEP = 256 / RW, LANES = PSTEP * RW, NIT = (FPB + PSTEP - 1) / PSTEP;
    int w, pb;
    bool on, side;
    __device__ FaceSlots() : w(threadIdx.x % RW), pb(threadIdx.x / RW), on(threadIdx.x < LANES), side(w >= NV) {}
    __device__ int wv() const { return side ? w - NV : w; }
};

// sT[p][w] = TQ value w of face point p0+p (w < NV left record, else right;
// a FREE face repeats its left record); every load of the thread issued first
template <int N, int FPB>
__device__ __forceinline__ void gather_face_traces(const double *__restrict__ TQ, const int4 *sF, int f0, int p0, int npt,
                                                   double *sT) {
    using SL = FaceSlots<N, FPB>;
    const SL sl;
    if (!sl.on) return;
    double val[SL::NIT];
#pragma unroll
    for (int j = 0; j < SL::NIT; j++) {
        const int p = sl.pb + j * SL::PSTEP;
        if (p < npt) {
            const int gp = p0 + p, f = gp / N;
            val[j] = TQ[(size_t)face_rec<N>(sF[f - f0], gp - f * N, sl.side) * NV + sl.wv()];
        }
    }
#pragma unroll
    for (int j = 0; j < SL::NIT; j++) {
        const int p = sl.pb + j * SL::PSTEP;
        if (p < npt) sT[p * SL::RW + sl.w] = val[j];
    }
}

// out[record value] = sT[p][w] for the slots keep(fc, w) selects
template <int N, int FPB, class T, class S, class Keep>
__device__ __forceinline__ void scatter_face_values(T *__restrict__ out, const int4 *sF, int f0, int p0, int npt,
                                                    const S *sT, Keep keep) {
    using SL = FaceSlots<N, FPB>;
    const SL sl;
    if (!sl.on) return;
#pragma unroll
    for (int j = 0; j < SL::NIT; j++) {
        const int p = sl.pb + j * SL::PSTEP;
        if (p < npt) {
            const int gp = p0 + p, f = gp / N;
            const int4 fc = sF[f - f0];
            if (keep(fc, sl.w)) out[(size_t)face_rec<N>(fc, gp - f * N, sl.side) * NV + sl.wv()] = T(sT[p * SL::RW + sl.w]);
        }
    }
}

template <int N, class R, int NG = 4>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_face_inv(FaceArgs a) {
    constexpr int FPB = 256, RW = 2 * NV;          // face points per block, values per point (left, right)
    constexpr int NFB = FPB / N + 2;               // faces a block's points can touch
    __shared__ double sT[FPB * RW];               // [p][left 9 | right 9]: traces, then fluxes
    __shared__ int4 sF[NFB];
    __shared__ double red[256];
    const int tid = threadIdx.x;
    const int p0 = xcd_block(blockIdx.x, gridDim.x) * FPB;
    const int npt = min(FPB, a.nface * N - p0);
    const int f0 = p0 / N, nf = (p0 + npt - 1) / N - f0 + 1;
    if (tid < nf) sF[tid] = a.faces[f0 + tid];
    __syncthreads();
    gather_face_traces<N, FPB>(a.TQ, sF, f0, p0, npt, sT);
    __syncthreads();
    double emax = 0.0;
    R Fn[8], az, slr = R(1);
    bool free_face = true;
    if (tid < npt) {
        const int gp = p0 + tid, k = gp % N;
        const int4 fc = sF[gp / N - f0];
        free_face = fc.z < 0;
        int ifp, jfp;
        face2fp<N>(fc.y + 1, k, ifp, jfp);
        double xy[2 * NG];
#pragma unroll
        for (int q = 0; q < 2 * NG; q++) xy[q] = a.cellxy[(size_t)fc.x * (2 * NG) + q];
        double s1, s2;
        metric_row_g<N, NG>(xy, lf_family(fc.y + 1), ifp, jfp, s1, s2);
        R Ql[NV], Qr[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) { Ql[v] = R(sT[tid * RW + v]); Qr[v] = R(sT[tid * RW + NV + v]); }
        if (free_face && fc.w != BK_ZEROGRAD) {
            // INLE / SYMP / WALL: the right state of the boundary kind
            double qr[NV];
            boundary_state(fc.w, sT + tid * RW, s1, s2, a.qin, qr);
#pragma unroll
            for (int v = 0; v < NV; v++) Qr[v] = R(qr[v]);
        }
        const int sl = side_sign(fc.y);
        R eig;
        rusanov<R>(R(a.gam), Ql, Qr, R(s1), R(s2), sl, Fn, az, eig);
        if (!free_face) slr = R(sl * -side_sign(fc.w));
        emax = double(eig);
    }
    __syncthreads();
    if (tid < npt) {
#pragma unroll
        for (int v = 0; v < 8; v++) sT[tid * RW + v] = double(Fn[v]);
        sT[tid * RW + 8] = double(az);
        if (!free_face) {
#pragma unroll
            for (int v = 0; v < 8; v++) sT[tid * RW + NV + v] = double(slr * Fn[v]);
            sT[tid * RW + NV + 8] = double(az);
        }
    }
    __syncthreads();
    // a FREE face has no right cell
    scatter_face_values<N, FPB>((R *)a.FF, sF, f0, p0, npt, sT,
                                [](const int4 &fc, int w) { return w < NV || fc.z >= 0; });
    if (a.artiv) {
        if (a.inst) {
            const int tt = min(p0 + tid, a.nface * N - 1);
            instance_max_atomic(emax, a.inst[a.faces[tt / N].x], a.eig);
        } else {
            block_max_atomic<R>(emax, red, a.eig);
        }
    }
}

// ---------------------------------------------------------------------------
// Stage kernel of the ideal path (vismode=1, mu=eta0=0, ARTIV=0): compresid
// (compresid.f90:1-67 via compflux.f90:1-104 and magnetic_potential.f90:1-42)
// with the face fluxes of k_face_inv, fused with one SSP stage combination
// (iterations.f90:92-217); the traces of the new state are stored for the next
// stage's face pass.  With mu=eta0=0 and no AV the viscous terms of compresid
// are exact zeros, so they are not evaluated.
// ---------------------------------------------------------------------------
template <int N, class R, int NG = 4>
__global__ void __launch_bounds__(IShape<N>::THREADS) __attribute__((amdgpu_waves_per_eu(CHORUS_IDEAL_MINW))) k_stage_ideal(StageArgs a) {
    using S = IShape<N>;
    constexpr int CPB = S::CPB, NP = S::NP, NF = S::NF, NQ = S::NQ;
    constexpr int FST = 8 * NF + 1;      // per-cell stride of sF (odd: cells land on different banks)
    constexpr int AST = 3 * NF + 1;      // per-cell stride of sA
    __shared__ double sQ[CPB * NQ];
    __shared__ double sX[CPB * 2 * NG];
    __shared__ R sF[CPB * FST];          // [c][v][line][r] contravariant flux, one family at a time
    __shared__ R sA[CPB * AST];          // [c][A_z | metric 1 | metric 2][line][r]
    __shared__ double sOps[Ops64Layout<N>::SIZE];   // FP64 operator tables (OpTables64 layout)
    constexpr int OM = Ops64Layout<N>::M, OXF = Ops64Layout<N>::XF, OXS = Ops64Layout<N>::XS;
    const int tid = threadIdx.x;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = blk * CPB;
    const int ncl = min(CPB, a.ncell - c0);

    // A: stage the block's cells (contiguous) and their vertex coordinates
    {
        constexpr int PQ = per_thread<CPB * NQ, S::THREADS>(), PX = per_thread<CPB * 2 * NG, S::THREADS>();
        double hq[PQ], hx[PX];
        gather_regs<PQ, S::THREADS>(hq, a.Qin + (size_t)c0 * NQ, ncl * NQ);
        gather_regs<PX, S::THREADS>(hx, a.cellxy + (size_t)c0 * 2 * NG, ncl * 2 * NG);
        OpTables64<N, S::THREADS> ops;
        ops.gather();
        scatter_lds<PQ, S::THREADS>(sQ, hq, ncl * NQ);
        scatter_lds<PX, S::THREADS>(sX, hx, ncl * 2 * NG);
        ops.scatter(sOps);
    }
    __syncthreads();

    // thread <-> solution point for the C phases
    const int pc = tid / NP, pk = tid % NP;
    const int pjs = pk / N, pis = pk % N;
    const bool pact = (tid < CPB * NP) && (pc < ncl);
    const R *FF = (const R *)a.FF;
    R dF[8], dA1 = R(0), dA2 = R(0);
    double qa[NV];

#pragma unroll 1
    for (int fam = 0; fam < 2; fam++) {
        // B1: interior flux points (inviscid_flux_interior, compflux.f90:59-102)
        for (int it = tid; it < ncl * N * (N - 1); it += blockDim.x) {
            const int c = it / (N * (N - 1)), k = it % (N * (N - 1));
            const int line = k / (N - 1), r = 1 + k % (N - 1);
            const int ifp = fam == 0 ? r : line, jfp = fam == 0 ? line : r;
            double s1d, s2d;
            metric_row_g<N, NG>(sX + c * 2 * NG, fam, ifp, jfp, s1d, s2d, sOps + OXF, sOps + OXS);
            R Qown[NV], F[8], G[8];
            trace9<N, R>(sQ + c * NQ, fam, ifp, jfp, Qown, sOps);
            iflux<R>(R(a.gam), Qown, F, G);
            const R s1 = R(s1d), s2 = R(s2d);
            R *f = sF + c * FST + line * (N + 1) + r;
#pragma unroll
            for (int v = 0; v < 8; v++) f[v * NF] = F[v] * s1 + G[v] * s2;
            R *aa = sA + c * AST + line * (N + 1) + r;
            aa[0] = Qown[8];
            aa[NF] = s1;
            aa[2 * NF] = s2;
        }
        // B2: face points: the common flux and A_z of k_face_inv
        for (int it = tid; it < ncl * 2 * N; it += blockDim.x) {
            const int c = it / (2 * N), k = it % (2 * N);
            const int line = k >> 1, r = (k & 1) ? N : 0;
            const int ifp = fam == 0 ? r : line, jfp = fam == 0 ? line : r;
            int lfi, nfp;
            if (fam == 0) { lfi = (r == 0) ? 3 : 1; nfp = (r == 0) ? (N - 1 - line) : line; }
            else          { lfi = (r == 0) ? 0 : 2; nfp = (r == 0) ? line : (N - 1 - line); }
            const R *src = FF + (size_t)(((c0 + c) * 4 + lfi) * N + nfp) * NV;
            double s1d, s2d;
            metric_row_g<N, NG>(sX + c * 2 * NG, fam, ifp, jfp, s1d, s2d, sOps + OXF, sOps + OXS);
            R *f = sF + c * FST + line * (N + 1) + r;
#pragma unroll
            for (int v = 0; v < 8; v++) f[v * NF] = src[v];
            R *aa = sA + c * AST + line * (N + 1) + r;
            aa[0] = src[8];
            aa[NF] = R(s1d);
            aa[2 * NF] = R(s2d);
        }
        __syncthreads();
        // the stage's Qa at the point, loaded while the derivative is summed
        if (fam == 1 && pact) prefetch_qa(a, (size_t)(c0 + pc) * NQ + pk, NP, qa);
        // C: derivative along this family at the thread's solution point
        if (pact) {
            const int line = fam == 0 ? pjs : pis;    // the line through the point
            const int pos = fam == 0 ? pis : pjs;     // position along the line
            const R *f = sF + pc * FST + line * (N + 1);
            const R *aa = sA + pc * AST + line * (N + 1);
            R d[8];
#pragma unroll
            for (int v = 0; v < 8; v++) d[v] = R(0);
            R e1 = R(0), e2 = R(0);
            constexpr int U = contraction_unroll<R, N>();
#pragma unroll U
            for (int r = 0; r <= N; r++) {
                R m = R(sOps[OM + r * N + pos]);
#pragma unroll
                for (int v = 0; v < 8; v++) d[v] = d[v] + f[v * NF + r] * m;
            }
#pragma unroll U
            for (int r = 0; r <= N; r++) e1 = e1 + aa[r] * R(sOps[OM + r * N + pos]) * aa[NF + r];
#pragma unroll U
            for (int r = 0; r <= N; r++) e2 = e2 + aa[r] * R(sOps[OM + r * N + pos]) * aa[2 * NF + r];
            if (fam == 0) {
#pragma unroll
                for (int v = 0; v < 8; v++) dF[v] = d[v];
                dA1 = e1; dA2 = e2;
            } else {
                // residual at the solution point (compresid.f90:12-65)
                const double Jd = jacobian_g<N, NG>(sX + pc * 2 * NG, pis, pjs, sOps + OXS);
                const R J = R(Jd);
                double *qs = sQ + pc * NQ + pk;
                const R q0 = R(qs[0]), q1 = R(qs[NP]), q2 = R(qs[2 * NP]);
                const R au = q1 / q0, av = q2 / q0;
                const R nA1 = (dA1 + e1) / J, nA2 = (dA2 + e2) / J;
                R res[NV];
#pragma unroll
                for (int v = 0; v < 8; v++) res[v] = -(dF[v] + d[v]);
                res[8] = -(au * nA1) - av * nA2;
                res[8] = res[8] * J;
#pragma unroll
                for (int v = 0; v < NV; v++) res[v] = res[v] / J;
                const size_t base = (size_t)(c0 + pc) * NQ + pk;
                stage_combine(a, base, NP, qs, res, qa);
            }
        }
        __syncthreads();
    }
    // D: traces of the new state for the next stage's face pass (the flux
    // planes, free now, stage the stores)
    constexpr bool STG = trace_stage_doubles<N>(CPB) * sizeof(double) <= CPB * FST * sizeof(R);
    if (a.kind != KIND_RESID) store_traces<N>(sQ, ncl, c0, a.TQ, sOps, STG ? reinterpret_cast<double *>(sF) : nullptr);
}

// ---------------------------------------------------------------------------
// correct_B_via_curl_Az (setinitialcond.f90:82-92): central A_z at every flux
// point (interior: L extrapolation, compflux.f90:69-75/88-94; faces: average of
// both traces, interfaceflux_all.f90:79-91), nablaQsp(9,9) (compflux.f90:107-157)
// and update_B (magnetic_potential.f90:44-67).  FP64 in both precision forms.
// In place: reads var 9, writes vars 6 and 7 and their face traces.
// WHICH=1 instead evaluates div B = d(Bx)/dx + d(By)/dy at the solution points
// (iterations.f90:35-48) and writes per-block sum of squares and max |div B|
// (no state change).  The partner's trace comes from the stored traces.
// ---------------------------------------------------------------------------
template <int N, int WHICH, int NG = 4>
__global__ void __launch_bounds__(256) k_ct(CTArgs a) {
    using S = Shape<N>;
    constexpr int CPB = S::CPB, NP = S::NP, NF = S::NF, NQ = S::NQ;
    constexpr int NVAR = WHICH == 0 ? 1 : 2, VAR0 = WHICH == 0 ? 8 : 5;
    __shared__ double sV[2][CPB * NP];   // variable planes staged (A_z, or Bx and By)
    __shared__ double sX[CPB * 2 * NG];
    __shared__ double sF[2][CPB * NF];   // central flux-point values, one family at a time
    __shared__ double sS[2][CPB * NF];   // metric rows, same layout
    __shared__ double red[2][256];
    __shared__ int2 sNbr[CPB * 4];
    __shared__ double sP[NVAR][CPB * 4 * N];   // the partner's trace at each face point [c][lfi][nfp]
    __shared__ double sOps[Ops64Layout<N>::SIZE];   // FP64 operator tables (OpTables64 layout)
    constexpr int OM = Ops64Layout<N>::M, OXF = Ops64Layout<N>::XF, OXS = Ops64Layout<N>::XS;
    const int tid = threadIdx.x;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = a.c_first + blk * CPB;
    const int ncl = min(CPB, a.ncell - c0);
    {
        constexpr int PV = per_thread<CPB * NP, S::THREADS>(), PX = per_thread<CPB * 2 * NG, S::THREADS>();
        double hv[NVAR][PV], hx[PX];
#pragma unroll
        for (int k = 0; k < PV; k++) {
            const int i = tid + k * S::THREADS, c = i / NP, p = i % NP;
#pragma unroll
            for (int w = 0; w < NVAR; w++) hv[w][k] = i < ncl * NP ? a.Q[(size_t)(c0 + c) * NQ + (VAR0 + w) * NP + p] : 0.0;
        }
        constexpr int PN = per_thread<CPB * 4, S::THREADS>();
        int2 hn[PN];
        gather_regs<PX, S::THREADS>(hx, a.cellxy + (size_t)c0 * 2 * NG, ncl * 2 * NG);
        gather_regs<PN, S::THREADS>(hn, a.nbr + (size_t)c0 * 4, ncl * 4);
        OpTables64<N, S::THREADS> ops;
        ops.gather();
        // the partners' stored traces at the block's face points, fetched with the
        // staging loads (a FREE face has none: its own trace is used instead)
        constexpr int PP = per_thread<CPB * 4 * N, S::THREADS>();
        double hp[NVAR][PP];
#pragma unroll
        for (int k = 0; k < PP; k++) {
            const int i = tid + k * S::THREADS, c = i / (4 * N), lfi = (i / N) % 4, nfp = i % N;
#pragma unroll
            for (int w = 0; w < NVAR; w++) hp[w][k] = 0.0;
            if (i < ncl * 4 * N) {
                const int2 info = a.nbr[(size_t)(c0 + c) * 4 + lfi];
                if (!((info.y >> 4) & 1)) {
                    const size_t prec = (size_t)((info.x * 4 + (info.y & 7) - 1) * N + (N - 1 - nfp));
                    if (WHICH == 0 && a.mt.az) {
                        hp[0][k] = a.mt.az[prec];
                    } else {
                        const double *t = a.TQ + prec * NV + VAR0;
#pragma unroll
                        for (int w = 0; w < NVAR; w++) hp[w][k] = t[w];
                    }
                }
            }
        }
#pragma unroll
        for (int w = 0; w < NVAR; w++) scatter_lds<PP, S::THREADS>(sP[w], hp[w], ncl * 4 * N);
#pragma unroll
        for (int w = 0; w < NVAR; w++) scatter_lds<PV, S::THREADS>(sV[w], hv[w], ncl * NP);
        scatter_lds<PX, S::THREADS>(sX, hx, ncl * 2 * NG);
        scatter_lds<PN, S::THREADS>(sNbr, hn, ncl * 4);
        ops.scatter(sOps);
    }
    __syncthreads();
    const int pc = tid / NP, pk = tid % NP, pjs = pk / N, pis = pk % N;
    const bool pact = (tid < CPB * NP) && (pc < ncl);
    double dpsi[NVAR][2], out[NVAR][2];   // [w][component]
#pragma unroll 1
    for (int fam = 0; fam < 2; fam++) {
        for (int it = tid; it < ncl * NF; it += blockDim.x) {
            const int c = it / NF, k = it % NF, line = k / (N + 1), r = k % (N + 1);
            const int ifp = fam == 0 ? r : line, jfp = fam == 0 ? line : r;
            double s1, s2;
            metric_row_g<N, NG>(sX + c * 2 * NG, fam, ifp, jfp, s1, s2, sOps + OXF, sOps + OXS);
            int2 info = make_int2(-1, 1 << 4);
            int lfi = 0, nfp = 0;
            if (r == 0 || r == N) {
                if (fam == 0) { lfi = (r == 0) ? 3 : 1; nfp = (r == 0) ? (N - 1 - line) : line; }
                else          { lfi = (r == 0) ? 0 : 2; nfp = (r == 0) ? line : (N - 1 - line); }
                info = sNbr[c * 4 + lfi];
            }
#pragma unroll
            for (int w = 0; w < NVAR; w++) {
                double own = trace1<N>(sV[w] + c * NP, fam, ifp, jfp, sOps);
                double val = own;
                if (r == 0 || r == N) {
                    const double other = ((info.y >> 4) & 1) ? own : sP[w][(c * 4 + lfi) * N + nfp];
                    val = 0.5 * own + 0.5 * other;
                }
                sF[w][c * NF + line * (N + 1) + r] = val;
            }
            sS[0][c * NF + line * (N + 1) + r] = s1;
            sS[1][c * NF + line * (N + 1) + r] = s2;
        }
        __syncthreads();
        if (pact) {
            const int line = fam == 0 ? pjs : pis, pos = fam == 0 ? pis : pjs;
#pragma unroll
            for (int w = 0; w < NVAR; w++) {
                const double *f = sF[w] + pc * NF + line * (N + 1);
#pragma unroll
                for (int d = 0; d < 2; d++) {
                    const double *ss = sS[d] + pc * NF + line * (N + 1);
                    double acc = 0.0;
#pragma unroll
                    for (int r = 0; r <= N; r++) acc = acc + f[r] * sOps[OM + r * N + pos] * ss[r];
                    if (fam == 0) dpsi[w][d] = acc;
                    else out[w][d] = acc;
                }
            }
        }
        __syncthreads();
    }
    if (WHICH == 0) {
        // update_B, then the traces of the new Bx, By for the next face pass
        if (pact) {
            const double J = jacobian_g<N, NG>(sX + pc * 2 * NG, pis, pjs, sOps + OXS);
            const size_t base = (size_t)(c0 + pc) * NQ + pk;
            const double n1 = (dpsi[0][0] + out[0][0]) / J;   // nablaQs(9,1)
            const double n2 = (dpsi[0][1] + out[0][1]) / J;   // nablaQs(9,2)
            a.Q[base + 5 * NP] = n2;
            a.Q[base + 6 * NP] = -n1;
            sV[0][pc * NP + pk] = n2;
            sV[1][pc * NP + pk] = -n1;
        }
        __syncthreads();
        for (int it = tid; it < ncl * 4 * N * 2; it += blockDim.x) {
            const int c = it / (8 * N), k = it % (8 * N), fpi = k / 2, w = k % 2;
            const int lfi = fpi / N, nfp = fpi % N;
            int ifp, jfp;
            face2fp<N>(lfi + 1, nfp, ifp, jfp);
            const double v = trace1<N>(sV[w] + c * NP, lf_family(lfi + 1), ifp, jfp, sOps);
            const size_t rec = (size_t)(((c0 + c) * 4 + lfi) * N + nfp);
            if (a.mt.dev) a.mt.dev[rec * 8 + 5 + w] = float(v - a.mt.ref[(size_t)(c0 + c) * 8 + 5 + w]);
            else a.TQout[rec * NV + 5 + w] = v;
        }
    } else {
        double sumsq = 0.0, dmax = 0.0;
        if (pact) {
            const double J = jacobian_g<N, NG>(sX + pc * 2 * NG, pis, pjs, sOps + OXS);
            const double dbx = (dpsi[0][0] + out[0][0]) / J;  // nablaQs(6,1)
            const double dby = (dpsi[NVAR - 1][1] + out[NVAR - 1][1]) / J;  // nablaQs(7,2)
            const double db = dbx + dby;
            sumsq = db * db;
            dmax = __builtin_fabs(db);
        }
        red[0][tid] = sumsq;
        red[1][tid] = dmax;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (tid < s) {
                red[0][tid] += red[0][tid + s];
                red[1][tid] = fmax(red[1][tid], red[1][tid + s]);
            }
            __syncthreads();
        }
        if (tid == 0) {
            a.partial[blk * 2 + 0] = red[0][0];
            a.partial[blk * 2 + 1] = red[1][0];
        }
    }
}

}  // namespace chorus
