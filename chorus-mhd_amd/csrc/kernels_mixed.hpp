// Mixed-precision kernels: FP32 flux arithmetic in delta form on the
// face-ordered pipeline of the FP64 path.
//
// Plain FP32 flux values of size |F| lose log10(N^2/h) digits when the
// divergence differences them over a cell.  Here every FP32 quantity on a flux
// point is a deviation from a FP64 reference state Qbar of the cell (its state
// at the central solution point, so deviations are O(h)):
//   dQf   = sum_sp (Q(sp) - Qbar) L(r,sp)                 (FP64 difference, FP32 MAC)
//   dF    = F(Qbar + dQf) - F(Qbar)                       (exact difference formulas)
//   dF1   = dF*S11 + dG*S12                               (contravariant)
// and the divergence is taken of dF1 only: the constant part F(Qbar).S has a
// vanishing discrete divergence on straight-sided bilinear cells (the metric
// identity behind free-stream preservation: d/dpsi(y_eta) - d/deta(y_psi) = 0,
// exact for the SD derivative of the linear metric terms), so dropping it
// changes the residual only at FP64 round-off.  The same holds for A_z and for
// the central Qv of the gradient.  FP32 errors are then relative to O(h)
// deviations, i.e. eps32*N^2 relative to the residual, independent of h.
//
// Every flux point, interior or face, evaluates dF1 of its own trace; a face
// point then adds the jump term of the common flux, computed once per face
// point by k_face_inv_mx from the two FP64 traces of that point:
//   Fnl - F(Q_L).S = 1/2 (F(Q_R) - F(Q_L)).S - 1/2 sign_l |S| lambda (Q_R - Q_L)
//   for the right cell (sign_l sign_r)(Fnl - F(Q_R).S)
// where F(Q_R) - F(Q_L) is again an exact difference about Q_L, so only jumps
// and deviations are ever rounded to FP32.
// The state, the RK sums, the traces, the metrics, the AV indicator and the
// CT update stay FP64, and so does the whole vector-potential equation (A_z at
// the flux points, its upwind face value, its derivative and residual): the
// SD div B of B = curl A_z amplifies grid-scale noise in A_z by ~(N^2/h)^2.
#pragma once

#include "kernels_visc.hpp"

// prototype: the derivative phase of k_grad_mx on the matrix cores
#ifndef CHORUS_GRAD_MFMA
#define CHORUS_GRAD_MFMA 0
#endif
typedef float v4f __attribute__((ext_vector_type(4)));
#ifndef CHORUS_MX_CUNROLL
#define CHORUS_MX_CUNROLL 2
#endif

namespace chorus {

// round trip of the AV sensor's interpolation per direction, N -> N-1 -> N
// points (AV.f90:15-58): R = H2 H1 (N x N, FP32; the 2-D round trip is R (x) R)
static __constant__ float c_Rav[kMaxN + 1][kMaxN * kMaxN];

static hipError_t upload_av_roundtrip() {
    static float R[kMaxN + 1][kMaxN * kMaxN];
    static double H1[kMaxN * kMaxN], H2[kMaxN * kMaxN];
    for (int n = 2; n <= kMaxN; n++) {
        host_av_tables(n, H1, H2);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double acc = 0.0;
                for (int l = 0; l < n - 1; l++) acc += H2[i * (n - 1) + l] * H1[l * n + j];
                R[n][i * n + j] = float(acc);
            }
    }
    return hipMemcpyToSymbol(HIP_SYMBOL(c_Rav), R, sizeof(R));
}

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// reference state of a cell (or a face point) and the derived quantities the
// delta fluxes need
struct Ref {
    float q[NV];          // Qbar
    float u[3];           // m/rho
    float b2, p, H, ub;   // |B|^2, pressure, E + p + |B|^2/2, u.B
};

// (the derived quantities only need FP32 consistency with Qbar: an error in
// them multiplies a deviation in the exact-difference fluxes)
__device__ __forceinline__ void ref_from(const double *qbar, double gam, Ref &s) {
#pragma unroll
    for (int v = 0; v < NV; v++) s.q[v] = float(qbar[v]);
    const double ir = 1.0 / qbar[0];
    const double ux = qbar[1] * ir, uy = qbar[2] * ir, uz = qbar[3] * ir;
    const double b2 = qbar[5] * qbar[5] + qbar[6] * qbar[6] + qbar[7] * qbar[7];
    const double p = (gam - 1.0) * (qbar[4] - 0.5 * (qbar[1] * ux + qbar[2] * uy + qbar[3] * uz) - 0.5 * b2);
    s.u[0] = float(ux); s.u[1] = float(uy); s.u[2] = float(uz);
    s.b2 = float(b2);
    s.p = float(p);
    s.H = float(qbar[4] + p + 0.5 * b2);
    s.ub = float(ux * qbar[5] + uy * qbar[6] + uz * qbar[7]);
}

// F(Qbar+d) - F(Qbar), G(Qbar+d) - G(Qbar) for the ideal-MHD fluxes of
// getIfluxvectors (getrusanov.f90:111-153), written as exact differences
__device__ __forceinline__ void dflux(const Ref &s, const float *d, float gm1, float *dF, float *dG) {
    const float rho = s.q[0] + d[0];
    const float ir = frcp(rho);
    const float du = (d[1] - s.u[0] * d[0]) * ir;
    const float dv = (d[2] - s.u[1] * d[0]) * ir;
    const float dw = (d[3] - s.u[2] * d[0]) * ir;
    const float ux = s.u[0] + du, uy = s.u[1] + dv, uz = s.u[2] + dw;
    const float bx = s.q[5] + d[5], by = s.q[6] + d[6], bz = s.q[7] + d[7];
    const float dB2 = (2.f * s.q[5] + d[5]) * d[5] + (2.f * s.q[6] + d[6]) * d[6] + (2.f * s.q[7] + d[7]) * d[7];
    const float dmu = s.q[1] * du + d[1] * ux + s.q[2] * dv + d[2] * uy + s.q[3] * dw + d[3] * uz;
    const float dp = gm1 * (d[4] - 0.5f * dmu - 0.5f * dB2);
    const float dH = d[4] + dp + 0.5f * dB2;
    const float H = s.H + dH;
    const float dub = s.u[0] * d[5] + du * bx + s.u[1] * d[6] + dv * by + s.u[2] * d[7] + dw * bz;
    const float pm = dp + 0.5f * dB2;
    dF[0] = d[1];
    dF[1] = s.q[1] * du + d[1] * ux + pm - (2.f * s.q[5] + d[5]) * d[5];
    dF[2] = s.q[1] * dv + d[1] * uy - (s.q[5] * d[6] + d[5] * by);
    dF[3] = s.q[1] * dw + d[1] * uz - (s.q[5] * d[7] + d[5] * bz);
    dF[4] = s.u[0] * dH + du * H - (s.ub * d[5] + dub * bx);
    dF[5] = 0.f;
    dF[6] = s.u[0] * d[6] + du * by - (s.u[1] * d[5] + dv * bx);
    dF[7] = s.u[0] * d[7] + du * bz - (s.u[2] * d[5] + dw * bx);
    dG[0] = d[2];
    dG[1] = dF[2];
    dG[2] = s.q[2] * dv + d[2] * uy + pm - (2.f * s.q[6] + d[6]) * d[6];
    dG[3] = s.q[2] * dw + d[2] * uz - (s.q[6] * d[7] + d[6] * bz);
    dG[4] = s.u[1] * dH + dv * H - (s.ub * d[6] + dub * by);
    dG[5] = -dF[6];
    dG[6] = 0.f;
    dG[7] = s.u[1] * d[7] + dv * bz - (s.u[2] * d[6] + dw * by);
}

// fast-magnetosonic Rusanov wave speed of getrusanov.f90:74-87 from the full
// FP32 states; nx, ny is a unit normal (orientation irrelevant)
__device__ __forceinline__ float wave_speed(const float *Ql, const float *Qr, float nx, float ny, float gam) {
    const float il = frcp(Ql[0]), ir = frcp(Qr[0]);
    const float ul = Ql[1] * il, vl = Ql[2] * il, wl = Ql[3] * il;
    const float ur = Qr[1] * ir, vr = Qr[2] * ir, wr = Qr[3] * ir;
    const float b2l = Ql[5] * Ql[5] + Ql[6] * Ql[6] + Ql[7] * Ql[7];
    const float b2r = Qr[5] * Qr[5] + Qr[6] * Qr[6] + Qr[7] * Qr[7];
    const float pl = (gam - 1.f) * (Ql[4] - 0.5f * Ql[0] * (ul * ul + vl * vl + wl * wl) - 0.5f * b2l);
    const float pr = (gam - 1.f) * (Qr[4] - 0.5f * Qr[0] * (ur * ur + vr * vr + wr * wr) - 0.5f * b2r);
    const float Bx = 0.5f * (Ql[5] + Qr[5]), By = 0.5f * (Ql[6] + Qr[6]), Bz = 0.5f * (Ql[7] + Qr[7]);
    const float irs = frcp(Ql[0] + Qr[0]);
    const float c_a = gam * (pl + pr) * irs;
    const float c_d = (Bx * Bx + By * By + Bz * Bz) * 2.f * irs;
    const float bn = Bx * nx + By * ny;
    const float c_n = bn * bn * 2.f * irs;
    const float c_tmp = fsqrt(fmaxf((c_a + c_d) * (c_a + c_d) - 4.f * c_a * c_n, 0.f));
    const float c_f = fsqrt(0.5f * (c_a + c_d + c_tmp));
    return 0.5f * fabsf((ul + ur) * nx + (vl + vr) * ny) + c_f;
}

// Rusanov jump terms of one face point for its left (ol) and right (orr)
// cells from the two FP64 traces (see the header); sl = sign_l, slr =
// sign_l sign_r; returns the FP32 wave speed.  d: the FP32 trace jump qr - ql.
__device__ __forceinline__ float rusanov_jumps(const double *ql, const float *d, float s1, float s2, float sl,
                                               float slr, double gam, float *ol, float *orr) {
    Ref s;
    ref_from(ql, gam, s);
    const float gamf = float(gam);
    float dF[8], dG[8];
    dflux(s, d, gamf - 1.f, dF, dG);
    const float mag = fsqrt(s1 * s1 + s2 * s2);
    const float im = frcp(mag);
    float Qr[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) Qr[v] = s.q[v] + d[v];
    const float eig = wave_speed(s.q, Qr, s1 * im, s2 * im, gamf);
    const float dis = 0.5f * sl * mag * eig;
#pragma unroll
    for (int v = 0; v < 8; v++) {
        const float h = 0.5f * (dF[v] * s1 + dG[v] * s2);
        const float e = dis * d[v];
        ol[v] = h - e;
        orr[v] = slr * (-h - e);
    }
    return eig;
}

// FP32 metric of a straight-sided cell for the delta fluxes.  The bilinear
// map's metric rows (mapder.f90:1-76) are linear along a line of flux points:
//   fam 0: (y_eta, -x_eta) = (1 - psi) A + psi B,   psi = Xf(ifp)
//   fam 1: (-y_psi, x_psi) from (1 - eta) C + eta D, eta = Xf(jfp)
// with the edge vectors A = v4 - v1, B = v3 - v2, C = v2 - v1, D = v3 - v4,
// differenced in FP64 and rounded once: e[8] = {Ax, Ay, Bx, By, Cx, Cy, Dx, Dy}.
// The block's edge vectors are staged with the state (ncl*8 values from the
// FP64 vertex coordinates of cells c0.., 8 doubles per cell).
// gather() issues the coordinate loads with the rest of the staging; scatter()
// writes the rounded differences to LDS before the block's staging barrier
template <int CPB, int T> struct EdgeStage {
    static constexpr int PER = (CPB * 8 + T - 1) / T;
    double p[PER], m[PER];
    __device__ __forceinline__ void gather(const double *xy, int ncl) {
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int i = (int)threadIdx.x + j * T, c = i >> 3, k = i & 7;
            // (plus, minus) coordinate of component k: x1..x4 at 0..3, y1..y4 at 4..7
            const int pl = (0x62516273 >> (4 * k)) & 15, mi = (0x73405140 >> (4 * k)) & 15;
            p[j] = i < ncl * 8 ? xy[c * 8 + pl] : 0.0;
            m[j] = i < ncl * 8 ? xy[c * 8 + mi] : 0.0;
        }
    }
    __device__ __forceinline__ void scatter(float *sE, int ncl) const {
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int i = (int)threadIdx.x + j * T;
            if (i < ncl * 8) sE[i] = float(p[j] - m[j]);
        }
    }
};
// The other cells' compact traces at the face points of the block's cells
// (cells c0.., ncl of them), staged with the rest of the block: per own face
// point the other cell's 8 FP32 deviations (in the own point order), per own
// local face the difference of the other cell's FP64 reference state and the
// own one (sBar), rounded to FP32; zeros at boundary faces.  gather_nbr()
// loads the neighbour entries first, gather() the traces (after the other
// staging loads are issued), scatter() writes them to LDS once sBar is there.
// point stride of the gradient kernel's point-major nablaQs in LDS: 16 values
// padded to 20 floats (80 bytes, still 16-byte aligned), so that the 16-byte
// accesses of neighbouring points start on different banks
#ifndef CHORUS_MX_GS
#define CHORUS_MX_GS 20
#endif
template <int N, int CPB, int T> struct PartnerStage {
    static constexpr int ND = CPB * 8 * N, NR = CPB * 16;   // float4 and double2 items
    static constexpr int PD = (ND + T - 1) / T, PR = (NR + T - 1) / T;
    int2 nd[PD], nr[PR];
    float4 d[PD];
    double2 r[PR];
    __device__ __forceinline__ void gather_nbr(const int2 *nbr, int c0, int ncl) {
#pragma unroll
        for (int k = 0; k < PD; k++) {
            const int i = (int)threadIdx.x + k * T;
            nd[k] = i < ncl * 8 * N ? nbr[(c0 + i / (8 * N)) * 4 + (i % (8 * N)) / (2 * N)] : make_int2(-1, 1 << 4);
        }
#pragma unroll
        for (int k = 0; k < PR; k++) {
            const int i = (int)threadIdx.x + k * T;
            nr[k] = i < ncl * 16 ? nbr[(c0 + i / 16) * 4 + (i / 4) % 4] : make_int2(-1, 1 << 4);
        }
    }
    __device__ __forceinline__ void gather(const MxTraces &mt) {
#pragma unroll
        for (int k = 0; k < PD; k++) {
            const int i = (int)threadIdx.x + k * T, rem = i % (8 * N), nfp = (rem >> 1) % N;
            const bool other = !((nd[k].y >> 4) & 1);
            const size_t rec = (size_t)((nd[k].x * 4 + (nd[k].y & 7) - 1) * N + (N - 1 - nfp));
            d[k] = other ? reinterpret_cast<const float4 *>(mt.dev)[rec * 2 + (rem & 1)] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < PR; k++) {
            const int i = (int)threadIdx.x + k * T;
            const bool other = !((nr[k].y >> 4) & 1);
            r[k] = other ? reinterpret_cast<const double2 *>(mt.ref)[(size_t)nr[k].x * 4 + i % 4] : make_double2(0.0, 0.0);
        }
    }
    __device__ __forceinline__ void scatter(float *sPD, float *sPR, const double *sBar, int ncl) const {
#pragma unroll
        for (int k = 0; k < PD; k++) {
            const int i = (int)threadIdx.x + k * T;
            if (i < ncl * 8 * N) reinterpret_cast<float4 *>(sPD)[i] = d[k];
        }
#pragma unroll
        for (int k = 0; k < PR; k++) {
            const int i = (int)threadIdx.x + k * T;
            if (i < ncl * 16) {
                const bool other = !((nr[k].y >> 4) & 1);
                const double *b = sBar + (i / 16) * NV + 2 * (i % 4);
                reinterpret_cast<float2 *>(sPR)[i] =
                    other ? make_float2(float(r[k].x - b[0]), float(r[k].y - b[1])) : make_float2(0.f, 0.f);
            }
        }
    }
};

__device__ __forceinline__ void metric_row_f(const float *e, int fam, float x, float &s1, float &s2) {
    if (fam == 0) {
        s1 = (1.f - x) * e[1] + x * e[3];
        s2 = -((1.f - x) * e[0] + x * e[2]);
    } else {
        s1 = -((1.f - x) * e[5] + x * e[7]);
        s2 = (1.f - x) * e[4] + x * e[6];
    }
}
// metric row of flux point (fam, r along the family, ifp, jfp) of block cell c
// in FP32: from the staged edge vectors, or the FP64 8-node map (NG = 8)
template <int N, int NG>
__device__ __forceinline__ void metric_row_mx(const float *sE, const double *sX, const double *ops, int c, int fam,
                                              int r, int ifp, int jfp, float &s1, float &s2) {
    if constexpr (NG == 4) {
        metric_row_f(sE + c * 8, fam, float(ops[Ops64Layout<N>::XF + r]), s1, s2);
    } else {
        double s1d, s2d;
        metric_row_g<N, NG>(sX + c * 2 * NG, fam, ifp, jfp, s1d, s2d, ops + Ops64Layout<N>::XF,
                            ops + Ops64Layout<N>::XS);
        s1 = float(s1d);
        s2 = float(s2d);
    }
}
// 1/J at solution point (is, js) of block cell c in FP32 (J = x_psi y_eta - x_eta y_psi)
template <int N, int NG>
__device__ __forceinline__ float inv_jacobian_mx(const float *sE, const double *sX, const double *ops, int c, int is,
                                                 int js) {
    if constexpr (NG == 4) {
        const float *e = sE + c * 8;
        const float psi = float(ops[Ops64Layout<N>::XS + is]), eta = float(ops[Ops64Layout<N>::XS + js]);
        const float xpsi = (1.f - eta) * e[4] + eta * e[6], ypsi = (1.f - eta) * e[5] + eta * e[7];
        const float xeta = (1.f - psi) * e[0] + psi * e[2], yeta = (1.f - psi) * e[1] + psi * e[3];
        return frcp(xpsi * yeta - xeta * ypsi);
    } else {
        return float(1.0 / jacobian_g<N, NG>(sX + c * 2 * NG, is, js, ops + Ops64Layout<N>::XS));
    }
}

// Phase A of the mixed cell kernels: the block's cells (contiguous in HBM)
// become FP32 deviations from their reference states Qbar (the FP64 state at
// the central solution point).  Each thread keeps the values it loaded in
// registers; the threads holding a central point publish it as Qbar, and q64,
// if given, receives the FP64 state itself (the combination, the FP64 A_z
// derivative and the traces read it from there).
// Called by all THREADS threads of the block.
template <int N, int CPB, int THREADS>
__device__ __forceinline__ void stage_deviations(const double *Qg, int ncl, double gam, double *sBar, Ref *sRef,
                                                 float *dq, double *q64 = nullptr) {
    constexpr int NP = N * N, NQ = NV * NP, PER = (CPB * NQ + THREADS - 1) / THREADS;
    constexpr int PC = (N / 2) * N + N / 2;   // central solution point
    const int tid = threadIdx.x, n = ncl * NQ;
    double held[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = tid + k * THREADS;
        held[k] = i < n ? Qg[i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = tid + k * THREADS;
        const int c = i / NQ, r = i % NQ, v = r / NP, p = r % NP;
        if (i < n && p == PC) sBar[c * NV + v] = held[k];
        if (q64 && i < n) q64[i] = held[k];
    }
    __syncthreads();
    for (int c = tid; c < ncl; c += THREADS) ref_from(sBar + c * NV, gam, sRef[c]);
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = tid + k * THREADS;
        if (i < n) dq[i] = float(held[k] - sBar[(i / NQ) * NV + (i % NQ) / NP]);
    }
    __syncthreads();
}

// The face-point records of the block's cells (cells c0.., ncl of them) read
// by the mixed stage kernels: the 8-float records of FF and the FP64 upwind
// A_z of FA, staged into LDS with the state, so that the flux tasks wait for
// no global load
template <int N, int CPB, int T> struct FaceRecStage {
    static constexpr int N4 = CPB * 8 * N, NA = CPB * 4 * N;   // float4 items, doubles
    static constexpr int P4 = (N4 + T - 1) / T, PA = (NA + T - 1) / T;
    float4 f[P4];
    double az[PA];
    __device__ __forceinline__ void gather(const float *FF, const double *FA, int c0, int ncl) {
        const float4 *src = reinterpret_cast<const float4 *>(FF) + (size_t)c0 * 8 * N;
#pragma unroll
        for (int k = 0; k < P4; k++) {
            const int i = (int)threadIdx.x + k * T;
            f[k] = i < ncl * 8 * N ? src[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < PA; k++) {
            const int i = (int)threadIdx.x + k * T;
            az[k] = i < ncl * 4 * N ? FA[(size_t)c0 * 4 * N + i] : 0.0;
        }
    }
    __device__ __forceinline__ void scatter(float *sFF, double *sFA, int ncl) const {
#pragma unroll
        for (int k = 0; k < P4; k++) {
            const int i = (int)threadIdx.x + k * T;
            if (i < ncl * 8 * N) reinterpret_cast<float4 *>(sFF)[i] = f[k];
        }
#pragma unroll
        for (int k = 0; k < PA; k++) {
            const int i = (int)threadIdx.x + k * T;
            if (i < ncl * 4 * N) sFA[i] = az[k];
        }
    }
};

// stage_deviations in two parts, so that loads which depend on other staged
// values can be issued between them: load() issues the state loads, finish()
// does the rest (with its two barriers)
template <int N, int CPB, int THREADS> struct DevStage {
    static constexpr int NP = N * N, NQ = NV * NP, PER = (CPB * NQ + THREADS - 1) / THREADS;
    double held[PER];
    __device__ __forceinline__ void load(const double *Qg, int ncl) {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = (int)threadIdx.x + k * THREADS;
            held[k] = i < ncl * NQ ? Qg[i] : 0.0;
        }
    }
    __device__ __forceinline__ void finish(int ncl, double gam, double *sBar, Ref *sRef, float *dq) const {
        constexpr int PC = (N / 2) * N + N / 2;   // central solution point
        const int tid = threadIdx.x, n = ncl * NQ;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = tid + k * THREADS;
            const int c = i / NQ, r = i % NQ, v = r / NP, p = r % NP;
            if (i < n && p == PC) sBar[c * NV + v] = held[k];
        }
        __syncthreads();
        for (int c = tid; c < ncl; c += THREADS) ref_from(sBar + c * NV, gam, sRef[c]);
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = tid + k * THREADS;
            if (i < n) dq[i] = float(held[k] - sBar[(i / NQ) * NV + (i % NQ) / NP]);
        }
        __syncthreads();
    }
};

// FP32 extrapolation and derivative operators of order N staged in LDS
// (the per-thread indices of the contractions would otherwise make every
// coefficient a vector load from the constant tables): gather before the
// block's first barrier, scatter after it
template <int N, int THREADS> struct OpTables {
    static constexpr int K = (N + 1) * N, PO = (K + THREADS - 1) / THREADS;
    float l[PO], m[PO];
    __device__ __forceinline__ void gather() {
        gather_regs<PO, THREADS>(l, &c_Lf[N][0], K);
        gather_regs<PO, THREADS>(m, &c_Mf[N][0], K);
    }
    __device__ __forceinline__ void scatter(float *sL, float *sM) const {
        scatter_lds<PO, THREADS>(sL, l, K);
        scatter_lds<PO, THREADS>(sM, m, K);
    }
};

// AV smoothness indicator of density, mixed form: the same quantity as
// av_sensor (AV.f90:15-82), log10 of sum_p (Qsm - Q)^2 / Q^2 with Qsm the
// N -> N-1 -> N interpolation round trip R (x) R of the cell's density, here in
// FP32 on the deviation d = rho - rhobar from a constant of the cell (the
// round trip reproduces constants, so Qsm - Q = (R (x) R - I) d): one pass per
// direction, then the per-cell sum and a FP32 logarithm.
// rho[c * CS + p]: the block's FP64 densities; rbar[c * NV]: the constant;
// work: 2 * CPB * N^2 floats of LDS; Rt: the LDS copy of c_Rav[N].
// Called by all threads of the block.
template <int N, int CPB, int CS = NV * N * N>
__device__ __forceinline__ void av_sensor_mx(double *sm, const double *rho, const double *rbar, float *work,
                                             int ncl, int c0, const float *Rt) {
    constexpr int NP = N * N;
    const int tid = threadIdx.x;
    const int pc = tid / NP, pk = tid % NP, pjs = pk / N, pis = pk % N;
    const bool pact = (tid < CPB * NP) && (pc < ncl);
    float *d = work, *t = work + CPB * NP;
    float dq = 0.f, r = 1.f;
    if (pact) {
        const double q = rho[pc * CS + pk];
        dq = float(q - rbar[pc * NV]);
        r = float(q);
        d[tid] = dq;
    }
    __syncthreads();
    if (pact) {
        // along is: t[js][is] = sum_k R[is][k] d[js][k]
        const float *dr = d + pc * NP + pjs * N;
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < N; k++) acc += Rt[pis * N + k] * dr[k];
        t[tid] = acc;
    }
    __syncthreads();
    if (pact) {
        // along js, minus the point's own deviation
        const float *tc = t + pc * NP + pis;
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < N; k++) acc += Rt[pjs * N + k] * tc[k * N];
        const float dev = acc - dq;
        d[tid] = dev * dev * frcp(r * r);
    }
    __syncthreads();
    if (tid < ncl) {
        float se = 0.f;
#pragma unroll
        for (int p = 0; p < NP; p++) se += d[tid * NP + p];
        sm[c0 + tid] = double(__builtin_amdgcn_logf(se) * 0.30102999566398120f);
    }
}

// SSP stage combination at one solution point: the point's FP64 state qs[v*NP]
// (the block's LDS copy) and a register copy of Qa (when the stage needs it);
// the new state replaces the LDS copy, for the traces
__device__ __forceinline__ void combine_point(const StageArgs &a, size_t base, int NP, double *qs,
                                              const double *qa, const float *res, double res8) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const double r = v == 8 ? res8 : double(res[v]);
        const size_t o = base + (size_t)v * NP;
        const double q = qs[v * NP];
        double qn;
        switch (a.kind) {
        case KIND_RESID: a.Rout[o] = r; continue;
        case KIND_AXPY: qn = q + a.c0 * r; break;
        case KIND_3A: qn = a.c1 * qa[v] + a.c2 * q + a.c0 * r; break;
        case KIND_3B: qn = a.c1 * qa[v] + a.c2 * q + a.c3 * r * a.c4; break;
        case KIND_S54_4: qn = a.c1 * qa[v] + a.c2 * q + a.c0 * r; a.Rout[o] = r; break;
        default: qn = a.c1 * qa[v] + a.c2 * a.Qb[o] + a.c3 * a.Qc[o] + a.c4 * q + a.c5 * a.Rprev[o] + a.c0 * r; break;
        }
        a.Qout[o] = qn;
        qs[v * NP] = qn;
    }
}

// delta trace: sum_sp dq(sp) L(r,sp) of all 9 variables
template <int N>
__device__ __forceinline__ void dtrace9(const float *dq, int fam, int ifp, int jfp, float *d, const float *Lf) {
    constexpr int NP = N * N;
#pragma unroll
    for (int v = 0; v < NV; v++) d[v] = 0.f;
#pragma unroll CHORUS_MX_TRACE_UNROLL
    for (int sp = 0; sp < N; sp++) {
        const float l = fam == 0 ? Lf[ifp * N + sp] : Lf[jfp * N + sp];
        const int idx = fam == 0 ? jfp * N + sp : sp * N + ifp;
#pragma unroll
        for (int v = 0; v < NV; v++) d[v] += dq[v * NP + idx] * l;
    }
}

// FP64 A_z derivative term of one family at position pos of a line: the
// contravariant A_z flux A*s along the line, differentiated, with A(r) the
// FP64 A_z at the N+1 flux points of the line.  The metric row s(r) of a
// bilinear cell is linear along the line and independent of the line, with
// end values (flux points 0 and N sit at 0 and 1) that are vertex differences:
//   fam 0: s = (y_eta, -x_eta),  s(0) = (y4-y1, x1-x4), s(N) = (y3-y2, x2-x3)
//   fam 1: s = (-y_psi, x_psi),  s(0) = (y1-y2, x2-x1), s(N) = (y4-y3, x3-x4)
// so sum_r A(r) M(r,pos) s(r) = s(0) T0 + s(N) T1 with
// T0 = sum_r A(r) M(r,pos) (1 - Xf(r)), T1 = sum_r A(r) M(r,pos) Xf(r).
// The metric rows of an 8-node cell (NG = 8) are quadratic along a line: there
// the products are summed point by point, with s(r) of the 8-node map.
template <int N, int NG = 4>
__device__ __forceinline__ void az_derivative(const double *A, const double *xy, int fam, int pos, int line,
                                              double &e1, double &e2, const double *ops) {
    if constexpr (NG == 8) {
        e1 = 0.0;
        e2 = 0.0;
#pragma unroll
        for (int r = 0; r <= N; r++) {
            const double am = A[r] * ops[Ops64Layout<N>::M + r * N + pos];
            double s1, s2;
            metric_row_g<N, 8>(xy, fam, fam == 0 ? r : line, fam == 0 ? line : r, s1, s2,
                               ops + Ops64Layout<N>::XF, ops + Ops64Layout<N>::XS);
            e1 += am * s1;
            e2 += am * s2;
        }
        return;
    }
    double T0 = 0.0, T1 = 0.0;
#pragma unroll
    for (int r = 0; r <= N; r++) {
        const double am = A[r] * ops[Ops64Layout<N>::M + r * N + pos];
        const double xf = c_Xf[N][r];   // uniform index: a scalar load
        T1 += am * xf;
        T0 += am - am * xf;
    }
    const double *x = xy, *y = xy + 4;
    if (fam == 0) {
        e1 = (y[3] - y[0]) * T0 + (y[2] - y[1]) * T1;
        e2 = (x[0] - x[3]) * T0 + (x[1] - x[2]) * T1;
    } else {
        e1 = (y[0] - y[1]) * T0 + (y[3] - y[2]) * T1;
        e2 = (x[1] - x[0]) * T0 + (x[2] - x[3]) * T1;
    }
}

// -(u J dA/dx + v J dA/dy) at solution point (pis, pjs) from the FP64 A_z
// flux-point values A[fam * FAMST + line * (N + 1) + r] of the point's cell
// (magnetic_potential.f90:1-42, compresid.f90:40-47); q is the point's state
template <int N, int FAMST, int NG = 4>
__device__ __forceinline__ double az_residual(const double *A, const double *xy, int pis, int pjs, const double *q,
                                              int qstride, const double *ops) {
    double e1a, e2a, e1b, e2b;
    az_derivative<N, NG>(A + pjs * (N + 1), xy, 0, pis, pjs, e1a, e2a, ops);
    az_derivative<N, NG>(A + FAMST + pis * (N + 1), xy, 1, pjs, pis, e1b, e2b, ops);
    const double au = q[qstride] / q[0], av = q[2 * qstride] / q[0];
    return -(au * (e1a + e1b) + av * (e2a + e2b));
}

// FP64 A_z at flux point (fam, r, line) of a cell: the upwind face value of a
// face pass at face points (FA: the records of the cell `cell`), the trace of
// the cell's FP64 A_z plane (az: N^2 values) elsewhere
template <int N>
__device__ __forceinline__ double az_at(const double *FA, const double *az, int cell, int fam, int r, int line,
                                        const double *ops) {
    if (r == 0 || r == N) {
        int lfi, nfp;
        face_of<N>(fam, r, line, lfi, nfp);
        return FA[(size_t)((cell * 4 + lfi) * N + nfp)];
    }
    return trace1<N>(az, fam, fam == 0 ? r : line, fam == 0 ? line : r, ops);
}

// a cell face point's trace in the compact form: the FP32 deviations of
// variables 1-8 and the cell's FP64 reference state
__device__ __forceinline__ void mx_trace(const MxTraces &mt, size_t rec, int cell, float *dv, double *ref) {
    const float4 *d = reinterpret_cast<const float4 *>(mt.dev + rec * 8);
    const float4 d0 = d[0], d1 = d[1];
    dv[0] = d0.x; dv[1] = d0.y; dv[2] = d0.z; dv[3] = d0.w;
    dv[4] = d1.x; dv[5] = d1.y; dv[6] = d1.z; dv[7] = d1.w;
    const double2 *r = reinterpret_cast<const double2 *>(mt.ref + (size_t)cell * 8);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const double2 x = r[j];
        ref[2 * j] = x.x;
        ref[2 * j + 1] = x.y;
    }
}

// upwind side of A_z at a face point: true for the left cell's trace, by the
// left cell's outward normal velocity, evaluated in FP64 as the FP64 form does
// (getrusanov.f90:61-66, 98-103): near-zero normal velocities then pick the
// same side, so A_z keeps the FP64 form's face values.  ql: the left trace
// (FP64), (m1, m2): its metric row, lfl: its local face.
__device__ __forceinline__ bool upwind_az(const double *ql, double m1, double m2, int lfl) {
    const double rhoi = 1.0 / ql[0];
    const double mn = sqrt(m1 * m1 + m2 * m2);
    const double ux = m1 / mn * double(side_sign(lfl)), uy = m2 / mn * double(side_sign(lfl));
    return rhoi * ql[1] * ux + rhoi * ql[2] * uy > 0.0;
}

// boundary_state in FP32 (for the wave speed of a boundary face point)
__device__ __forceinline__ void boundary_state_f(int kind, const float *ql, float s1, float s2, const double *qin,
                                                 float *qr) {
#pragma unroll
    for (int v = 0; v < NV; v++) qr[v] = ql[v];
    if (kind == BK_MIRROR) {
        const float im = frcp(fsqrt(s1 * s1 + s2 * s2));
        const float nx = s1 * im, ny = s2 * im;
        const float mn = ql[1] * nx + ql[2] * ny, bn = ql[5] * nx + ql[6] * ny;
        qr[1] = ql[1] - 2.f * mn * nx;
        qr[2] = ql[2] - 2.f * mn * ny;
        qr[5] = ql[5] - 2.f * bn * nx;
        qr[6] = ql[6] - 2.f * bn * ny;
    } else if (kind == BK_INLET) {
#pragma unroll
        for (int v = 0; v < 8; v++) qr[v] = float(qin[v]);
    }
}

// ---------------------------------------------------------------------------
// Face pass, mixed form: the jump terms of the Rusanov flux for both cells of
// each face point (see the header) and the FP64 upwind A_z; the stage maximum
// of the FP32 wave speed with ARTIV.  A FREE / OUTL face has no jump (zeros).
// FF[cell][lfi][nfp][8] (float): the 8 flux jump terms of the cell's side
// (the viscous face kernel subtracts its flux from them in place).
// ---------------------------------------------------------------------------
// JUMPS = false (viscous path with physical viscosity and AV together): only
// the wave-speed maximum, which the gradient kernel's AV viscosity needs
// before its interior fluxes.  The viscous face kernel forms the jump terms
// and the upwind A_z, and with AV alone the gradient kernel forms the maximum
// itself (the AV fluxes are linear in it).
template <int N, int NG = 4, bool JUMPS = true>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_face_inv_mx(FaceArgs a) {
    // one thread per face point: both sides' compact traces are loaded
    // directly (the face's points share the cells' reference states through
    // the cache), and the 32-byte output records leave directly
    __shared__ double red[256];
    const int tid = threadIdx.x;
    const int gp = xcd_block(blockIdx.x, gridDim.x) * 256 + tid;
    double emax = 0.0;
    if (gp < a.nface * N) {
        const int k = gp % N;
        const int4 fc = a.faces[gp / N];
        const bool free_face = fc.z < 0;
        int ifp, jfp;
        face2fp<N>(fc.y + 1, k, ifp, jfp);
        double xy[2 * NG];
#pragma unroll
        for (int q = 0; q < 2 * NG; q++) xy[q] = a.cellxy[(size_t)fc.x * (2 * NG) + q];
        double m1, m2;
        metric_row_g<N, NG>(xy, lf_family(fc.y + 1), ifp, jfp, m1, m2);
        const float s1 = float(m1), s2 = float(m2);
        const unsigned rl = face_rec<N>(fc, k, false);
        const unsigned rr = free_face ? rl : face_rec<N>(fc, k, true);
        // the own trace in FP64 (reference plus deviation) and the jump to the
        // right state: the other cell's trace (the rounded difference of the
        // two references plus that of the deviations); at an INLE / SYMP /
        // WALL face the boundary kind's state; none at a FREE / OUTL face
        float dl[8];
        double bl[8], ql[NV];
        mx_trace(a.mt, rl, fc.x, dl, bl);
#pragma unroll
        for (int v = 0; v < 8; v++) ql[v] = bl[v] + double(dl[v]);
        ql[8] = a.mt.az[rl];
        float d[NV];
        d[8] = 0.f;
        if (!free_face) {
            float dr[8];
            double br[8];
            mx_trace(a.mt, rr, fc.z, dr, br);
#pragma unroll
            for (int v = 0; v < 8; v++) d[v] = float(br[v] - bl[v]) + (dr[v] - dl[v]);
        } else if (fc.w != BK_ZEROGRAD) {
            double qr[NV];
            boundary_state(fc.w, ql, m1, m2, a.qin, qr);
#pragma unroll
            for (int v = 0; v < 8; v++) d[v] = float(qr[v] - ql[v]);
        } else {
#pragma unroll
            for (int v = 0; v < 8; v++) d[v] = 0.f;
        }
        const float sl = float(side_sign(fc.y));
        const float slr = free_face ? 0.f : sl * float(-side_sign(fc.w));
        float ol[8], orr[8], eig;
        if constexpr (JUMPS) {
            eig = rusanov_jumps(ql, d, s1, s2, sl, slr, a.gam, ol, orr);
        } else {
            float Ql[NV], Qr[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                Ql[v] = float(ql[v]);
                Qr[v] = Ql[v] + d[v];
            }
            const float im = frcp(fsqrt(s1 * s1 + s2 * s2));
            eig = wave_speed(Ql, Qr, s1 * im, s2 * im, float(a.gam));
        }
        if constexpr (JUMPS) {
            float4 *fl = reinterpret_cast<float4 *>((float *)a.FF + (size_t)rl * 8);
            if (free_face) {
                // the left cell's jump term against the boundary state (none
                // at a FREE / OUTL face)
                if (fc.w == BK_ZEROGRAD) {
#pragma unroll
                    for (int v = 0; v < 8; v++) ol[v] = 0.f;
                }
                fl[0] = make_float4(ol[0], ol[1], ol[2], ol[3]);
                fl[1] = make_float4(ol[4], ol[5], ol[6], ol[7]);
                // A_z of the boundary state is the cell's own: the upwind value is too
                a.FA[rl] = ql[8];
            } else {
                float4 *fr = reinterpret_cast<float4 *>((float *)a.FF + (size_t)rr * 8);
                fl[0] = make_float4(ol[0], ol[1], ol[2], ol[3]);
                fl[1] = make_float4(ol[4], ol[5], ol[6], ol[7]);
                fr[0] = make_float4(orr[0], orr[1], orr[2], orr[3]);
                fr[1] = make_float4(orr[4], orr[5], orr[6], orr[7]);
                const double az = upwind_az(ql, m1, m2, fc.y) ? ql[8] : a.mt.az[rr];
                a.FA[rl] = az;
                a.FA[rr] = az;
            }
        }
        emax = double(eig);
    }
    if (a.artiv) {
        if (a.inst) {
            const int tt = min(gp, a.nface * N - 1);
            instance_max_atomic(emax, a.inst[a.faces[tt / N].x], a.eig);
        } else {
            block_max_atomic<float>(emax, red, a.eig);
        }
    }
}

// ---------------------------------------------------------------------------
// Ideal-path stage kernel, mixed form: every flux point evaluates the delta
// flux of its own trace, face points add the jump terms of k_face_inv_mx; the
// residual is formed in FP32 (A_z in FP64) and combined into the FP64 state.
// One thread per flux point of a family (FluxCellShape), so the flux phase is
// one round.  LDS holds the FP32 deviations and flux planes (one union area,
// reused for the new state at the end) and the FP64 A_z values.
// ---------------------------------------------------------------------------
// 4 waves per SIMD (<= 128 VGPRs) where that needs no spills
#ifndef CHORUS_MX_IDEAL_MINW
#define CHORUS_MX_IDEAL_MINW(N) ((N) <= 5 ? 4 : 3)
#endif
template <int N, int TB, int NG = 4>
__global__ void __launch_bounds__((FluxCellShape<N, TB>::THREADS)) __attribute__((amdgpu_waves_per_eu(CHORUS_MX_IDEAL_MINW(N)))) k_stage_ideal_mx(StageArgs a) {
    using S = FluxCellShape<N, TB>;
    constexpr int CPB = S::CPB, NP = S::NP, NF = S::NF, NQ = S::NQ, T = S::THREADS;
    constexpr int FST = 8 * NF + 1;      // per-cell stride of the dF1 planes
    constexpr int AST = NF + 1;          // per-cell stride of the FP64 A_z flux-point values
    constexpr int UF = CPB * NQ + CPB * FST;
    constexpr int UD = (UF + 1) / 2;
    __shared__ double U[UD];
    __shared__ double sQ64[CPB * NQ];    // FP64 state (combined in place into the new state)
    __shared__ double sX[CPB * 2 * NG];
    __shared__ float sE[NG == 4 ? CPB * 8 : 1];   // FP32 edge vectors of the block's cells
    __shared__ double sBar[CPB * NV];
    __shared__ Ref sRef[CPB];
    __shared__ double sA[2 * CPB * AST]; // FP64 A_z at the flux points of both families
    __shared__ float sLf[(N + 1) * N], sMf[(N + 1) * N];
    __shared__ double sOps[Ops64Layout<N>::SIZE];   // FP64 operator tables (OpTables64 layout)
    float *sD = (float *)U, *sF = sD + CPB * NQ;
    const int tid = threadIdx.x;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = blk * CPB;
    const int ncl = min(CPB, a.ncell - c0);
    {
        // the small per-cell arrays are loaded with the state (one memory latency)
        constexpr int PX = per_thread<CPB * 2 * NG, T>();
        double hx[PX];
        OpTables<N, T> ops;
        gather_regs<PX, T>(hx, a.cellxy + (size_t)c0 * 2 * NG, ncl * 2 * NG);
        ops.gather();
        OpTables64<N, T> ops64;
        ops64.gather();
        EdgeStage<CPB, T> edges;
        if constexpr (NG == 4) edges.gather(a.cellxy + (size_t)c0 * 2 * NG, ncl);
        stage_deviations<N, CPB, T>(a.Qin + (size_t)c0 * NQ, ncl, a.gam, sBar, sRef, sD, sQ64);
        scatter_lds<PX, T>(sX, hx, ncl * 2 * NG);
        if constexpr (NG == 4) edges.scatter(sE, ncl);
        ops.scatter(sLf, sMf);
        ops64.scatter(sOps);
        __syncthreads();
    }
    const float gam = float(a.gam);
    const int pc = tid / NP, pk = tid % NP, pjs = pk / N, pis = pk % N;
    const bool pact = (tid < CPB * NP) && (pc < ncl);
    const size_t base = (size_t)(c0 + pc) * NQ + pk;
    const float *FF = (const float *)a.FF;
    float dF[8], res[NV];
    double qa[NV];
#pragma unroll 1
    for (int fam = 0; fam < 2; fam++) {
        for (int it = tid; it < ncl * NF; it += T) {
            const int c = it / NF, k = it % NF;
            const int line = k / (N + 1), r = k % (N + 1);
            const int ifp = fam == 0 ? r : line, jfp = fam == 0 ? line : r;
            float jump[8];
#pragma unroll
            for (int v = 0; v < 8; v++) jump[v] = 0.f;
            if (r == 0 || r == N) {
                int lfi, nfp;
                face_of<N>(fam, r, line, lfi, nfp);
                const float4 *src = reinterpret_cast<const float4 *>(FF + (size_t)(((c0 + c) * 4 + lfi) * N + nfp) * 8);
                const float4 j0 = src[0], j1 = src[1];
                jump[0] = j0.x; jump[1] = j0.y; jump[2] = j0.z; jump[3] = j0.w;
                jump[4] = j1.x; jump[5] = j1.y; jump[6] = j1.z; jump[7] = j1.w;
            }
            float s1, s2;
            metric_row_mx<N, NG>(sE, sX, sOps, c, fam, r, ifp, jfp, s1, s2);
            float d[NV], F[8], G[8];
            dtrace9<N>(sD + c * NQ, fam, ifp, jfp, d, sLf);
            dflux(sRef[c], d, gam - 1.f, F, G);
            float *f = sF + c * FST + line * (N + 1) + r;
#pragma unroll
            for (int v = 0; v < 8; v++) f[v * NF] = F[v] * s1 + G[v] * s2 + jump[v];
            // (the FP64 A_z plane is the state's, combined only after the flux phase)
            sA[(fam * CPB + c) * AST + line * (N + 1) + r] = az_at<N>(a.FA, sQ64 + c * NQ + 8 * NP, c0 + c, fam, r, line, sOps);
        }
        __syncthreads();
        // the point's Qa for the combination, loaded while the derivative is summed
        if (fam == 1 && pact) prefetch_qa(a, base, NP, qa);
        if (pact) {
            const int line = fam == 0 ? pjs : pis, pos = fam == 0 ? pis : pjs;
            const float *f = sF + pc * FST + line * (N + 1);
            float d[8];
#pragma unroll
            for (int v = 0; v < 8; v++) d[v] = 0.f;
#pragma unroll CHORUS_MX_CUNROLL
            for (int r = 0; r <= N; r++) {
                const float m = sMf[r * N + pos];
#pragma unroll
                for (int v = 0; v < 8; v++) d[v] += f[v * NF + r] * m;
            }
#pragma unroll
            for (int v = 0; v < 8; v++) dF[v] = fam == 0 ? d[v] : -(dF[v] + d[v]);
        }
        __syncthreads();
    }
    // A_z residual in FP64 from the flux-point values of both families
    double res8 = 0.0;
    if (pact) {
        const double iJd = 1.0 / jacobian_g<N, NG>(sX + pc * 2 * NG, pis, pjs, sOps + Ops64Layout<N>::XS);
        const float iJ = float(iJd);
#pragma unroll
        for (int v = 0; v < 8; v++) res[v] = dF[v] * iJ;
        res[8] = 0.f;
        res8 = az_residual<N, CPB * AST, NG>(sA + pc * AST, sX + pc * 2 * NG, pis, pjs, sQ64 + pc * NQ + pk, NP, sOps) * iJd;
    }
    if (pact) combine_point(a, base, NP, sQ64 + pc * NQ + pk, qa, res, res8);
    if (a.kind == KIND_RESID) return;
    __syncthreads();
    store_traces_mx<N>(sQ64, ncl, c0, a.mt, sBar, NV, 1, sOps);
}

// ---------------------------------------------------------------------------
// Viscous / AV path, mixed form: two threads per solution point as in the
// FP64 path, and no inviscid face pass before the gradient kernel (only with
// physical viscosity and AV together, for the wave-speed maximum).  The
// central Qv of the gradient pass is a deviation from the cell's reference
// state (at face points the mean of the own delta trace and the other cell's
// compact trace); nablaQs, its traces and the viscous fluxes are plain FP32
// (they carry no large constant part); the viscous face flux is evaluated once
// per face point from the trace average, beside the Rusanov jump terms and the
// upwind A_z.  k_grad_mx also forms the viscous and AV fluxes of the interior
// flux points and their part of the divergence at the solution points (fvi),
// which the stage kernel adds.  With AV alone, k_grad_mx forms the stage
// maximum of the wave speed from its face points and the AV viscosity for a
// unit wave speed; the viscous face and stage kernels scale by the maximum.
// ---------------------------------------------------------------------------
template <int N, int TB, int VM, int NG = 4>
__global__ void __launch_bounds__((V2ShapeT<N, TB>::THREADS)) k_grad_mx(ViscArgs a) {
    using S = CellShape<N, TB>;
    using TS = V2ShapeT<N, TB>;
    constexpr int CPB = S::CPB, NP = S::NP, NF = S::NF, NQ = S::NQ, T = TS::THREADS, NIT = TS::NIT;
    constexpr int VST = 10 * NF + 1;
    constexpr int GS = CHORUS_MX_GS;                          // sG point stride (floats)
    constexpr int GO = (CPB * NQ + CPB * VST + 3) / 4 * 4;   // sG offset (16-byte aligned)
    constexpr int UF = GO + CPB * GS * NP;
    constexpr int UD = (UF + 1) / 2 > CPB * NQ ? (UF + 1) / 2 : CPB * NQ;
    __shared__ __align__(16) double U[UD];
    __shared__ double sX[CPB * 2 * NG];
    __shared__ float sE[NG == 4 ? CPB * 8 : 1];   // FP32 edge vectors of the block's cells
    __shared__ double sBar[CPB * NV];
    __shared__ Ref sRef[CPB];
    __shared__ int2 sNbr[CPB * 4];
    __shared__ double sMu[CPB * 4];      // muAVfi/fj(1,side) per local face
    __shared__ float sLf[(N + 1) * N], sMf[(N + 1) * N];
    __shared__ double sOps[Ops64Layout<N>::SIZE];   // FP64 operator tables (OpTables64 layout)
    float *sD = (float *)U, *sV = sD + CPB * NQ, *sG = sD + GO;   // sG: nablaQs, point-major [c][p][GS]
    // the other cells' traces (PartnerStage): their deviations [c][face point][8]
    // and reference differences [c][lf][8]; they live in sG, which is written
    // only by the second family's derivative, where N allows
#if CHORUS_GRAD_MFMA
    constexpr bool IN_G = false;
#else
    constexpr bool IN_G = 32 * (N + 1) <= GS * N * N;
#endif
    __shared__ __align__(16) float sPX[IN_G ? 1 : CPB * 32 * (N + 1)];
    float *sPD = IN_G ? sG : sPX, *sPR = sPD + CPB * 4 * N * 8;
    const int tid = threadIdx.x;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = blk * CPB;
    const int ncl = min(CPB, a.ncell - c0);
    {
        constexpr int PX = per_thread<CPB * 2 * NG, T>(), PN = per_thread<CPB * 4, T>();
        double hx[PX];
        int2 hn[PN];
        OpTables<N, T> ops;
        PartnerStage<N, CPB, T> partners;
        partners.gather_nbr(a.nbr, c0, ncl);
        gather_regs<PX, T>(hx, a.cellxy + (size_t)c0 * 2 * NG, ncl * 2 * NG);
        gather_regs<PN, T>(hn, a.nbr + (size_t)c0 * 4, ncl * 4);
        ops.gather();
        OpTables64<N, T> ops64;
        ops64.gather();
        EdgeStage<CPB, T> edges;
        if constexpr (NG == 4) edges.gather(a.cellxy + (size_t)c0 * 2 * NG, ncl);
        DevStage<N, CPB, T> dev;
        dev.load(a.Qin + (size_t)c0 * NQ, ncl);
        // (the neighbour entries were the first loads issued)
        partners.gather(a.mt);
        dev.finish(ncl, a.gam, sBar, sRef, sD);
        scatter_lds<PX, T>(sX, hx, ncl * 2 * NG);
        if constexpr (NG == 4) edges.scatter(sE, ncl);
        scatter_lds<PN, T>(sNbr, hn, ncl * 4);
        partners.scatter(sPD, sPR, sBar, ncl);
        ops.scatter(sLf, sMf);
        ops64.scatter(sOps);
        __syncthreads();
    }
    // face averages of muAV(1) of the block's local faces, as in k_grad
    if constexpr ((VM & 2) != 0) {
        for (int i = tid; i < ncl * 4; i += T) {
            const int cell = c0 + i / 4;
            const int2 info = sNbr[i];
            const int in = a.inst ? a.inst[cell] : 0;
            // with AV alone, for a unit wave speed: the AV fluxes are linear in
            // it, and the viscous face and stage kernels scale them by the stage
            // maximum that this kernel forms
            const double ev = VM == 2 ? 1.0 : a.eig[in];
            const double eps0 = ev * (a.inst ? a.hinst[in] : a.h_msh) / (double)(N - 1) * 1.0;
            const double mo = mu_av(a.sm[cell], eps0, a.pi);
            sMu[i] = ((info.y >> 4) & 1) ? mo : 0.5 * (mo + mu_av(a.sm[info.x], eps0, a.pi));
            if ((i & 3) == 0) a.mucell[cell] = mo;   // for the viscous face kernel (and the output)
        }
    }
    const int grp = tid >= TS::G1 ? 1 : 0;
    const int pt = tid - grp * TS::G1;
    const int pc = pt / NP, pk = pt % NP, pjs = pk / N, pis = pk % N;
    const bool pact = (pt < CPB * NP) && (pc < ncl);
    float gp[8];
#if CHORUS_GRAD_MFMA
    // 1/J of every point of the block, for the MFMA derivative phase
    __shared__ float sIJ[CPB * NP];
    for (int i = tid; i < ncl * NP; i += T) {
        const int k = i % NP;
        sIJ[i] = float(1.0 / jacobian_g<N, NG>(sX + (i / NP) * 2 * NG, k % N, k / N, sOps + Ops64Layout<N>::XS));
    }
#endif
    // with AV alone the stage maximum of the Rusanov wave speed is formed here,
    // from the face tasks (each of the cell's face points is one of them, in
    // one family or the other; the speed is symmetric in the two sides)
    float emax = 0.f;
    const float gamf = float(a.gam);
#pragma unroll 1
    for (int fam = 0; fam < 2; fam++) {
        for (int t = tid; t < TS::NI + TS::NFC; t += T) {
            int c, line, r;
            const bool face_task = t >= TS::NI;
            bool bnd = true;
            int kind = BK_ZEROGRAD, pfp = 0;
            if (!face_task) {
                c = t / NIT;
                const int k = t % NIT;
                line = k / (N - 1); r = 1 + k % (N - 1);
            } else {
                const int item = t - TS::NI, k = item % (2 * N);
                c = item / (2 * N);
                line = k >> 1; r = (k & 1) ? N : 0;
                if (c < ncl) {
                    int lfi, nfp;
                    face_of<N>(fam, r, line, lfi, nfp);
                    const int2 info = sNbr[c * 4 + lfi];
                    bnd = (info.y >> 4) & 1;
                    kind = (info.y >> 5) & 3;
                    pfp = lfi * N + nfp;
                }
            }
            if (c >= ncl) continue;
            const int ifp = fam == 0 ? r : line, jfp = fam == 0 ? line : r;
            float s1d, s2d;
            metric_row_mx<N, NG>(sE, sX, sOps, c, fam, r, ifp, jfp, s1d, s2d);
            float d[NV];
            dtrace9<N>(sD + c * NQ, fam, ifp, jfp, d, sLf);
            float *q = sV + c * VST + line * (N + 1) + r;
            // the other cell's trace as a deviation from this cell's reference
            // state (the own one at a boundary face)
            float dj[8];
            if (face_task && !bnd) {
                const float *pd = sPD + (c * 4 * N + pfp) * 8, *pr = sPR + (c * 4 + pfp / N) * 8;
#pragma unroll
                for (int v = 0; v < 8; v++) dj[v] = pr[v] + pd[v];
                // the central value (own + other)/2
#pragma unroll
                for (int v = 0; v < 8; v++) q[v * NF] = 0.5f * (d[v] + dj[v]);
            } else {
#pragma unroll
                for (int v = 0; v < 8; v++) {
                    dj[v] = d[v];
                    q[v * NF] = d[v];
                }
            }
            q[8 * NF] = float(s1d);
            q[9 * NF] = float(s2d);
            if constexpr (VM == 2) {
                if (face_task) {
                    float Ql[NV], Qr[NV];
#pragma unroll
                    for (int v = 0; v < 8; v++) {
                        Ql[v] = sRef[c].q[v] + d[v];
                        Qr[v] = sRef[c].q[v] + dj[v];
                    }
                    Ql[8] = Qr[8] = 0.f;
                    if (bnd && kind != BK_ZEROGRAD) boundary_state_f(kind, Ql, s1d, s2d, a.qin, Qr);
                    const float im = frcp(fsqrt(s1d * s1d + s2d * s2d));
                    emax = fmaxf(emax, wave_speed(Ql, Qr, s1d * im, s2d * im, gamf));
                }
            }
        }
        __syncthreads();
#if CHORUS_GRAD_MFMA
        // derivative phase on the matrix cores: per (cell, line) one 16x16 tile
        // D[pos][col] = sum_r M(r,pos) X[r][col], col = (metric row grp, variable v),
        // X = q_v(r) s_grp(r), with v_mfma_f32_16x16x4f32 over k-blocks of 4 flux
        // points (rows pos >= N and flux points r > N are zero operands)
        {
            constexpr int KB = (N + 4) / 4;
            const int lane = tid & 63, wv = tid >> 6, kq = lane >> 4, col = lane & 15;
            const int cgrp = col >> 3, cv = col & 7;
            float am[KB];
#pragma unroll
            for (int kb = 0; kb < KB; kb++) {
                const int r = kb * 4 + kq;
                am[kb] = (col < N && r <= N) ? sMf[r * N + col] : 0.f;
            }
            for (int tile = wv; tile < ncl * N; tile += T / 64) {
                const int c = tile / N, line = tile % N;
                const float *q = sV + c * VST + line * (N + 1);
                v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kb = 0; kb < KB; kb++) {
                    const int r = kb * 4 + kq;
                    const float b = r <= N ? q[cv * NF + r] * q[(8 + cgrp) * NF + r] : 0.f;
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(am[kb], b, acc, 0, 0, 0);
                }
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int pos = 4 * kq + j;
                    if (pos < N) {
                        const int k = fam == 0 ? line * N + pos : pos * N + line;
                        float *g = sG + (c * NP + k) * GS + cgrp * 8 + cv;
                        *g = fam == 0 ? acc[j] : (*g + acc[j]) * sIJ[c * NP + k];
                    }
                }
            }
        }
        (void)gp;
#else
        if (pact) {
            const int line = fam == 0 ? pjs : pis, pos = fam == 0 ? pis : pjs;
            const float *q = sV + pc * VST + line * (N + 1);
            const float *mq = q + (8 + grp) * NF;
            float acc[8];
#pragma unroll
            for (int v = 0; v < 8; v++) acc[v] = 0.f;
#pragma unroll CHORUS_MX_CUNROLL
            for (int r = 0; r <= N; r++) {
                const float mm = sMf[r * N + pos] * mq[r];
#pragma unroll
                for (int v = 0; v < 8; v++) acc[v] += q[v * NF + r] * mm;
            }
            if (fam == 0) {
#pragma unroll
                for (int v = 0; v < 8; v++) gp[v] = acc[v];
            } else {
                const float iJ = inv_jacobian_mx<N, NG>(sE, sX, sOps, pc, pis, pjs);
#pragma unroll
                for (int v = 0; v < 8; v++) gp[v] = (gp[v] + acc[v]) * iJ;
                lds_store8(sG + (pc * NP + pk) * GS + grp * 8, gp);
            }
        }
#endif
        __syncthreads();
    }
    if constexpr (VM == 2) {
        // (every thread: a wave-wide maximum per instance)
        const int tc = tid >= TS::NI ? min((tid - TS::NI) / (2 * N), ncl - 1) : 0;
        instance_max_atomic_filtered(double(emax), a.inst ? a.inst[c0 + tc] : 0, a.eig);
    }
    // own face traces of nablaQs: a thread per face point and half of the 16
    // components, the L row in registers
    for (int it = tid; it < ncl * 4 * N * 2; it += T) {
        const int c = it / (8 * N), k = it % (8 * N), fpi = k >> 1, h = k & 1;
        const int lfi = fpi / N, nfp = fpi % N;
        int ifp, jfp;
        face2fp<N>(lfi + 1, nfp, ifp, jfp);
        const bool fam0 = lf_family(lfi + 1) == 0;
        const int lrow = fam0 ? ifp : jfp, base = fam0 ? jfp * N : ifp, stride = fam0 ? 1 : N;
        float Lr[N];
#pragma unroll
        for (int sp = 0; sp < N; sp++) Lr[sp] = sLf[lrow * N + sp];
        float *dst = a.gtr_r<float>() + (size_t)((c0 + c) * 4 * N + fpi) * 16 + h * 8;
        float acc[8];
#pragma unroll
        for (int gg = 0; gg < 8; gg++) acc[gg] = 0.f;
#pragma unroll
        for (int sp = 0; sp < N; sp++) {
            float v[8];
            lds_load8(sG + (c * NP + base + sp * stride) * GS + h * 8, v);
#pragma unroll
            for (int gg = 0; gg < 8; gg++) acc[gg] += v[gg] * Lr[sp];
        }
        reinterpret_cast<float4 *>(dst)[0] = make_float4(acc[0], acc[1], acc[2], acc[3]);
        reinterpret_cast<float4 *>(dst)[1] = make_float4(acc[4], acc[5], acc[6], acc[7]);
    }
    // viscous / AV flux of the own traces at the interior flux points of both
    // families (comp_viscous_flux, compflux.f90:205-256), one task per thread
    // (tasks family-major, so that a wave's tasks share a family: those of
    // family 1 read the delta traces and metric rows its flux phase left in sV);
    // then their part of the divergence at the solution points, for the stage
    // kernel: fvi[cell][point][8] = sum over interior r of M(r, pos) Fv(r), both
    // families (the face points' viscous flux joins the stage kernel's sum
    // through the face records)
    static_assert(2 * CPB * NIT <= T, "one interior flux task per thread");
    float fvr[8];
    int tfam = -1, tc = 0, tk = 0;
    {
        const int it = tid;
        const int fam = it / (CPB * NIT), k1 = it % (CPB * NIT), c = k1 / NIT, k = k1 % NIT;
        if (it < 2 * CPB * NIT && c < ncl) {
        const int line = k / (N - 1), r = 1 + k % (N - 1);
        const int ifp = fam == 0 ? r : line, jfp = fam == 0 ? line : r;
        float d[NV], qv[NV], g[2 * NV], fv[NV], s1, s2;
        if (fam == 1) {
            const float *q = sV + c * VST + line * (N + 1) + r;
#pragma unroll
            for (int w = 0; w < 8; w++) d[w] = q[w * NF];
            d[8] = 0.f;   // (A_z does not enter the viscous flux)
            s1 = q[8 * NF];
            s2 = q[9 * NF];
        } else {
            metric_row_mx<N, NG>(sE, sX, sOps, c, fam, r, ifp, jfp, s1, s2);
            dtrace9<N>(sD + c * NQ, fam, ifp, jfp, d, sLf);
        }
#pragma unroll
        for (int w = 0; w < NV; w++) qv[w] = sRef[c].q[w] + d[w];
        grad_trace_pm<N, float, GS>(sG + c * NP * GS, fam, ifp, jfp, g, sLf);
        visc_point<N, float, VM>(a, sMu + c * 4, fam, ifp, jfp, s1, s2, qv, g, fv, sOps + Ops64Layout<N>::XF);
        // (8 components: the viscous flux of A_z is zero)
#pragma unroll
        for (int w = 0; w < 8; w++) fvr[w] = fv[w];
        tfam = fam;
        tc = c;
        tk = k;
        }
    }
    __syncthreads();   // the deviations, flux planes and gradient in LDS are free now
    float *sFv = sD;   // [fam][c][w][line][r-1]
    if (tfam >= 0) {
#pragma unroll
        for (int w = 0; w < 8; w++) sFv[((tfam * CPB + tc) * 8 + w) * NIT + tk] = fvr[w];
    }
    __syncthreads();
    if (pact) {
        const int vg = 4 * grp;
        float p4[4] = {0.f, 0.f, 0.f, 0.f};
        const float *f0 = sFv + ((0 * CPB + pc) * 8 + vg) * NIT + pjs * (N - 1);
        const float *f1 = sFv + ((1 * CPB + pc) * 8 + vg) * NIT + pis * (N - 1);
#pragma unroll
        for (int r = 1; r < N; r++) {
            const float m0 = sMf[r * N + pis], m1 = sMf[r * N + pjs];
#pragma unroll
            for (int w = 0; w < 4; w++) p4[w] += f0[w * NIT + r - 1] * m0 + f1[w * NIT + r - 1] * m1;
        }
        float *dst = a.fvi_r<float>() + ((size_t)(c0 + pc) * NP + pk) * 8 + vg;
        reinterpret_cast<float4 *>(dst)[0] = make_float4(p4[0], p4[1], p4[2], p4[3]);
    }
}

// viscous and AV face flux, mixed form: k_face_visc in FP32 from the FP64
// trace average, together with the Rusanov jump terms of the face point (the
// viscous path's k_face_inv_mx leaves them to this pass): FF[cell side] =
// jump - viscous flux
// (5 waves per SIMD: at most 96 VGPRs, which it reaches without spilling)
template <int N, int VM, int NG = 4>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_face_visc_mx(FaceArgs a) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= a.nface * N) return;
    const int f = t / N, k = t % N;
    const int4 fc = a.faces[f];
    const bool free_face = fc.z < 0;
    const int kr = N - 1 - k;
    float sl1, sl2, sr1 = 0.f, sr2 = 0.f;
    double ml1, ml2;
    {
        int ifp, jfp;
        double xy[2 * NG], m1, m2;
        face2fp<N>(fc.y + 1, k, ifp, jfp);
#pragma unroll
        for (int q = 0; q < 2 * NG; q++) xy[q] = a.cellxy[(size_t)fc.x * (2 * NG) + q];
        metric_row_g<N, NG>(xy, lf_family(fc.y + 1), ifp, jfp, ml1, ml2);
        sl1 = float(ml1); sl2 = float(ml2);
        if (!free_face) {
            face2fp<N>(fc.w + 1, kr, ifp, jfp);
#pragma unroll
            for (int q = 0; q < 2 * NG; q++) xy[q] = a.cellxy[(size_t)fc.z * (2 * NG) + q];
            metric_row_g<N, NG>(xy, lf_family(fc.w + 1), ifp, jfp, m1, m2);
            sr1 = float(m1); sr2 = float(m2);
        }
    }
    const size_t pl = (size_t)((fc.x * 4 + fc.y) * N + k);
    const size_t pr = free_face ? pl : (size_t)((fc.z * 4 + fc.w) * N + kr);
    // the own trace in FP64 and the jump to the other cell's trace (the
    // rounded difference of the two references plus that of the deviations;
    // none at a boundary face): the central value own + jump/2 of the viscous
    // flux, and the Rusanov jump terms of the inviscid flux
    double ql[NV];
    float dj[8];
    {
        float dl[8];
        double bl[8];
        mx_trace(a.mt, pl, fc.x, dl, bl);
#pragma unroll
        for (int v = 0; v < 8; v++) ql[v] = bl[v] + double(dl[v]);
        if (!free_face) {
            float dr[8];
            double br[8];
            mx_trace(a.mt, pr, fc.z, dr, br);
#pragma unroll
            for (int v = 0; v < 8; v++) dj[v] = float(br[v] - bl[v]) + (dr[v] - dl[v]);
        } else {
#pragma unroll
            for (int v = 0; v < 8; v++) dj[v] = 0.f;
        }
    }
    // the FP64 upwind A_z of the face point for both cells (the own one at a
    // boundary face: the boundary state keeps the cell's A_z), stored last
    const double az = a.mt.az[free_face || upwind_az(ql, ml1, ml2, fc.y) ? pl : pr];
    ql[8] = 0.0;
    float qv[NV], g[2 * NV];
#pragma unroll
    for (int v = 0; v < 8; v++) qv[v] = float(ql[v]) + 0.5f * dj[v];
    qv[8] = 0.f;
    float jl[8], jr[8];
    {
        // the jump to the right state of the inviscid flux: to the other
        // cell's trace; at an INLE / SYMP / WALL face to the boundary kind's
        // state; none at a FREE / OUTL face
        float d[NV];
#pragma unroll
        for (int v = 0; v < 8; v++) d[v] = dj[v];
        d[8] = 0.f;
        if (free_face && fc.w != BK_ZEROGRAD) {
            double qr[NV];
            boundary_state(fc.w, ql, ml1, ml2, a.qin, qr);
#pragma unroll
            for (int v = 0; v < NV; v++) d[v] = float(qr[v] - ql[v]);
        }
        const float sl = float(side_sign(fc.y));
        const float slr = free_face ? 0.f : sl * float(-side_sign(fc.w));
        rusanov_jumps(ql, d, sl1, sl2, sl, slr, a.gam, jl, jr);
        if (free_face && fc.w == BK_ZEROGRAD) {
#pragma unroll
            for (int v = 0; v < 8; v++) jl[v] = 0.f;
        }
    }
    const float *gl = (const float *)a.gtr + pl * 16, *gr = (const float *)a.gtr + pr * 16;
#pragma unroll
    for (int d = 0; d < 2; d++) {
#pragma unroll
        for (int w = 0; w < 8; w++) g[d * NV + w] = 0.5f * (gl[d * 8 + w] + gr[d * 8 + w]);
        g[d * NV + 8] = 0.f;
    }
    const float gam = float(a.gam), lam = -2.f / 3.f, prandt = 0.72f;
    float fl[NV], fr[NV], Fv[NV], Gv[NV];
#pragma unroll
    for (int w = 0; w < NV; w++) { fl[w] = 0.f; fr[w] = 0.f; }
    if constexpr ((VM & 1) != 0) {
        vflux<float>(gam, lam, prandt, float(a.mu), float(a.eta0), qv, g, Fv, Gv);
#pragma unroll
        for (int w = 0; w < NV; w++) { fl[w] = Fv[w] * sl1 + Gv[w] * sl2; fr[w] = Fv[w] * sr1 + Gv[w] * sr2; }
    }
    if constexpr ((VM & 2) != 0) {
        // muAV(1) of both cells (mu_av of this stage's epsilon0, kept by k_grad_mx)
        const double ml = a.mucell[fc.x];
        const double mr = free_face ? ml : a.mucell[fc.z];
        double mf = free_face ? ml : 0.5 * (ml + mr);
        // with AV alone the gradient kernel formed muAV for a unit wave speed
        // and the stage maximum beside it
        if constexpr (VM == 2) mf = mf * a.eig[a.inst ? a.inst[fc.x] : 0];
        vflux<float>(gam, lam, prandt, float(mf), 0.f, qv, g, Fv, Gv);
#pragma unroll
        for (int w = 0; w < NV; w++) {
            fl[w] += Fv[w] * sl1 + Gv[w] * sl2;
            fr[w] += Fv[w] * sr1 + Gv[w] * sr2;
        }
    }
    // each cell side's record: jump - viscous flux (components 0-7; the
    // viscous flux of A_z is zero)
    float4 *rl = reinterpret_cast<float4 *>((float *)a.FF + pl * 8);
    rl[0] = make_float4(jl[0] - fl[0], jl[1] - fl[1], jl[2] - fl[2], jl[3] - fl[3]);
    rl[1] = make_float4(jl[4] - fl[4], jl[5] - fl[5], jl[6] - fl[6], jl[7] - fl[7]);
    a.FA[pl] = az;
    if (!free_face) {
        float4 *rr = reinterpret_cast<float4 *>((float *)a.FF + pr * 8);
        rr[0] = make_float4(jr[0] - fr[0], jr[1] - fr[1], jr[2] - fr[2], jr[3] - fr[3]);
        rr[1] = make_float4(jr[4] - fr[4], jr[5] - fr[5], jr[6] - fr[6], jr[7] - fr[7]);
        a.FA[pr] = az;
    }
}

// task layout of the mixed viscous stage kernel: one flux task per flux point
// of a family, then two threads per solution point for the derivative and
// the stage combination (group 0: variables 0-3 and A_z, group 1: 4-7)
template <int N, int TB> struct MxVShape {
    using S = CellShape<N, TB>;
    static constexpr int CPB = S::CPB, NP = S::NP, NF = S::NF;
    static constexpr int NIT = N * (N - 1);
    static constexpr int NA = CPB * NF;                      // flux tasks (all flux points of a family)
    static constexpr int G1 = ((CPB * NP + 63) / 64) * 64;   // first thread of group 1
    static constexpr int CT = G1 + CPB * NP;
    static constexpr int THREADS = (((NA > CT ? NA : CT) + 63) / 64) * 64;
};


// Viscous / AV stage kernel, mixed form.  The mixed form keeps no reference
// association, so the inviscid and viscous contributions of a flux point are
// combined before the derivative: every flux task forms
//   u = dF1(own delta trace) + jump - Fv        (8 components)
// with the jump terms of k_face_inv_mx and the viscous face flux of
// k_face_visc_mx at face points; the interior viscous fluxes enter as their
// divergence at the solution points, formed by k_grad_mx (fvi).  The viscous
// flux of A_z is zero (getrusanov.f90:157-361), so A_z keeps its FP64
// upwind/trace values alone.  Then one sum per component.
// waves per SIMD requested from the register allocator: 5 (at most 96 VGPRs)
// up to N = 5, where it spills at most a few words; N = 6 would spill
// (measured per order and tile shape)
#ifndef CHORUS_MX_VISC_MINW
#define CHORUS_MX_VISC_MINW(N) ((N) <= 5 ? 5 : (N) <= 6 ? 4 : 3)
#endif
template <int N, int TB, int VM, int NG = 4>
__global__ void __launch_bounds__((MxVShape<N, TB>::THREADS)) __attribute__((amdgpu_waves_per_eu(CHORUS_MX_VISC_MINW(N)))) k_stage_visc_mx(StageArgs a, ViscArgs v) {
    using S = CellShape<N, TB>;
    using TS = MxVShape<N, TB>;
    constexpr int CPB = S::CPB, NP = S::NP, NF = S::NF, NQ = S::NQ, T = TS::THREADS;
    constexpr int UST = 8 * NF + 1;      // per-cell stride of the combined flux planes
    constexpr int AST = NF + 1;          // per-cell stride of the FP64 A_z flux-point values
    // the flux planes of both families side by side at N >= 6 (two 512-thread
    // blocks per CU either way there), of one family at a time below (at N = 5
    // both side by side measured slower even where the LDS kept five blocks
    // per CU)
    constexpr bool MERGE = N >= 6;
    constexpr int UF = CPB * NQ + (MERGE ? 2 : 1) * CPB * UST;   // deviations + flux planes
    constexpr int UD = (UF + 1) / 2;
    __shared__ double U[UD];
    __shared__ double sQ64[CPB * NQ];    // FP64 state (combined in place into the new state)
    __shared__ double sX[CPB * 2 * NG];
    __shared__ float sE[NG == 4 ? CPB * 8 : 1];   // FP32 edge vectors of the block's cells
    __shared__ double sBar[CPB * NV];
    __shared__ Ref sRef[CPB];
    __shared__ double sA[2 * CPB * AST]; // FP64 A_z at the flux points of both families
    __shared__ float sLf[(N + 1) * N], sMf[(N + 1) * N];
    __shared__ double sOps[Ops64Layout<N>::SIZE];   // FP64 operator tables (OpTables64 layout)
    __shared__ float sRav[N * N];        // AV interpolation round trip (c_Rav)
    __shared__ __align__(16) float sFF[CPB * 4 * N * 8];   // the cells' face records (FaceRecStage)
    __shared__ double sFA[CPB * 4 * N];
    float *sD = (float *)U, *sU = sD + CPB * NQ;
    if constexpr ((VM & 2) != 0) {
        if (threadIdx.x < N * N) sRav[threadIdx.x] = c_Rav[N][threadIdx.x];
    }
    const int tid = threadIdx.x;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = blk * CPB;
    const int ncl = min(CPB, a.ncell - c0);
    if constexpr ((VM & 2) != 0) {
        if (blockIdx.x == 0)
            for (int i = tid; i < v.ninst; i += T) v.eig_next[i] = 0.0;
    }
    {
        // the per-cell arrays are loaded with the state (one memory latency)
        constexpr int PX = per_thread<CPB * 2 * NG, T>();
        double hx[PX];
        OpTables<N, T> ops;
        ops.gather();
        OpTables64<N, T> ops64;
        ops64.gather();
        gather_regs<PX, T>(hx, a.cellxy + (size_t)c0 * 2 * NG, ncl * 2 * NG);
        EdgeStage<CPB, T> edges;
        if constexpr (NG == 4) edges.gather(a.cellxy + (size_t)c0 * 2 * NG, ncl);
        FaceRecStage<N, CPB, T> recs;
        recs.gather((const float *)a.FF, a.FA, c0, ncl);
        stage_deviations<N, CPB, T>(a.Qin + (size_t)c0 * NQ, ncl, a.gam, sBar, sRef, sD, sQ64);
        scatter_lds<PX, T>(sX, hx, ncl * 2 * NG);
        if constexpr (NG == 4) edges.scatter(sE, ncl);
        recs.scatter(sFF, sFA, ncl);
        ops.scatter(sLf, sMf);
        ops64.scatter(sOps);
    }
    __syncthreads();
    // the By indicator of the input state, for the Tecplot output's muAV(6)
    if constexpr ((VM & 2) != 0) {
        if (v.sm6) {
            // (its work area: the flux planes, unused until the flux phase)
            av_sensor<N, CPB>(v.sm6, sQ64 + 5 * NP, U + (CPB * NQ + 1) / 2, ncl, c0, sOps);
            __syncthreads();
        }
    }
    const float gam = float(a.gam);
    const int grp = tid >= TS::G1 ? 1 : 0;
    const int pt = tid - grp * TS::G1;
    const int pc = pt / NP, pk = pt % NP, pjs = pk / N, pis = pk % N;
    const bool pact = (pt < CPB * NP) && (pc < ncl);
    const size_t base = (size_t)(c0 + pc) * NQ + pk;
    const int vg = 4 * grp;              // first variable of the thread's group
    const float *FVI = v.fvi_r<float>();
    float acc[4];
    double qa[5];
    float4 pv = make_float4(0.f, 0.f, 0.f, 0.f);   // the interior viscous fluxes' divergence (k_grad_mx)
    double eg = 0.0;                                 // AV alone: the stage maximum of the wave speed
    // one flux task per flux point: the own delta flux plus, at a face point,
    // the jump - viscous face flux record (the interior viscous flux enters
    // through k_grad_mx's partial divergence); and the FP64 A_z of the point
    auto flux_task = [&](int fam, int c, int k) {
        const int line = k / (N + 1), r = k % (N + 1);
        const int ifp = fam == 0 ? r : line, jfp = fam == 0 ? line : r;
        float *u = sU + ((MERGE ? fam : 0) * CPB + c) * UST + line * (N + 1) + r;
        float extra[8];
        if (r == 0 || r == N) {
            int lfi, nfp;
            face_of<N>(fam, r, line, lfi, nfp);
            const float4 *src = reinterpret_cast<const float4 *>(sFF + ((c * 4 + lfi) * N + nfp) * 8);
            const float4 e0 = src[0], e1 = src[1];
            extra[0] = e0.x; extra[1] = e0.y; extra[2] = e0.z; extra[3] = e0.w;
            extra[4] = e1.x; extra[5] = e1.y; extra[6] = e1.z; extra[7] = e1.w;
        } else {
#pragma unroll
            for (int w = 0; w < 8; w++) extra[w] = 0.f;
        }
        float s1, s2;
        metric_row_mx<N, NG>(sE, sX, sOps, c, fam, r, ifp, jfp, s1, s2);
        float d[NV], F[8], G[8];
        dtrace9<N>(sD + c * NQ, fam, ifp, jfp, d, sLf);
        dflux(sRef[c], d, gam - 1.f, F, G);
#pragma unroll
        for (int w = 0; w < 8; w++) u[w * NF] = F[w] * s1 + G[w] * s2 + extra[w];
        // (the FP64 A_z plane is the state's, combined only after the flux phase)
        sA[(fam * CPB + c) * AST + line * (N + 1) + r] = az_at<N>(sFA, sQ64 + c * NQ + 8 * NP, c, fam, r, line, sOps);
    };
    // the point's Qa for the combination, loaded while the derivative is summed
    auto load_point = [&]() {
        const bool need = a.kind != KIND_RESID && a.kind != KIND_AXPY;
#pragma unroll
        for (int k = 0; k < 4; k++) qa[k] = need ? a.Qa[base + (size_t)(vg + k) * NP] : 0.0;
        qa[4] = need && grp == 0 ? a.Qa[base + (size_t)8 * NP] : 0.0;
        pv = reinterpret_cast<const float4 *>(FVI + ((size_t)(c0 + pc) * NP + pk) * 8 + vg)[0];
        if constexpr (VM == 2) eg = v.eig[v.inst ? v.inst[c0 + pc] : 0];
    };
    // the derivative of one family's planes (the family-1 sum added to the family-0 one)
    auto derivative = [&](int fam) {
        const int line = fam == 0 ? pjs : pis, pos = fam == 0 ? pis : pjs;
        const float *f = sU + ((MERGE ? fam : 0) * CPB + pc) * UST + line * (N + 1) + vg * NF;
        float d4[4];
#pragma unroll
        for (int w = 0; w < 4; w++) d4[w] = 0.f;
#pragma unroll CHORUS_MX_CUNROLL
        for (int r = 0; r <= N; r++) {
            const float m = sMf[r * N + pos];
#pragma unroll
            for (int w = 0; w < 4; w++) d4[w] += f[w * NF + r] * m;
        }
#pragma unroll
        for (int w = 0; w < 4; w++) acc[w] = fam == 0 ? d4[w] : acc[w] + d4[w];
    };
    if constexpr (MERGE) {
        // both families' flux tasks in one phase (their planes side by side):
        // they fill the threads more evenly than a family per round, and two
        // barriers go
        for (int t = tid; t < 2 * TS::NA; t += T) {
            const int fam = t >= TS::NA ? 1 : 0, tf = t - fam * TS::NA;
            if (tf / NF < ncl) flux_task(fam, tf / NF, tf % NF);
        }
        __syncthreads();
        if (pact) {
            load_point();
            derivative(0);
            derivative(1);
        }
    } else {
#pragma unroll 1
        for (int fam = 0; fam < 2; fam++) {
            for (int t = tid; t < TS::NA; t += T)
                if (t / NF < ncl) flux_task(fam, t / NF, t % NF);
            __syncthreads();
            if (pact) {
                if (fam == 1) load_point();
                derivative(fam);
            }
            __syncthreads();
        }
    }
    if (pact) {
        // 1/J: FP64 for the A_z residual (group 0), FP32 for the rest
        const double iJd = grp == 0 ? 1.0 / jacobian_g<N, NG>(sX + pc * 2 * NG, pis, pjs, sOps + Ops64Layout<N>::XS) : 0.0;
        const float iJ = grp == 0 ? float(iJd) : inv_jacobian_mx<N, NG>(sE, sX, sOps, pc, pis, pjs);
        if constexpr (VM == 2) {
            // AV alone: the gradient kernel formed the interior AV fluxes for a
            // unit wave speed; the stage maximum scales them (and muAV, kept
            // for the output)
            const float ef = float(eg);
            pv.x *= ef; pv.y *= ef; pv.z *= ef; pv.w *= ef;
            if (pk == 0 && grp == 0) v.mucell[c0 + pc] *= eg;
        }
        float res[4];
        res[0] = (pv.x - acc[0]) * iJ;
        res[1] = (pv.y - acc[1]) * iJ;
        res[2] = (pv.z - acc[2]) * iJ;
        res[3] = (pv.w - acc[3]) * iJ;
        double *qs = sQ64 + pc * NQ + pk;
        if (grp == 0) {
            // A_z residual in FP64 from the flux-point values of both families
            // (before the combination overwrites the point's density and momenta)
            const double r8 = az_residual<N, CPB * AST, NG>(sA + pc * AST, sX + pc * 2 * NG, pis, pjs, qs, NP, sOps) * iJd;
            combine_vars<4>(a, base, NP, qs, 0, qa, res);
            combine_vars<1>(a, base, NP, qs, 8, qa + 4, res, r8);
        } else {
            combine_vars<4>(a, base, NP, qs, 4, qa, res);
        }
    }
    if (a.kind == KIND_RESID) return;
    __syncthreads();
    store_traces_mx<N>(sQ64, ncl, c0, a.mt, sBar, NV, 1, sOps);
    // (its work area: the deviations and flux planes, free since the derivative)
    if constexpr ((VM & 2) != 0) av_sensor_mx<N, CPB>(a.sm, sQ64, sBar, reinterpret_cast<float *>(U), ncl, c0, sRav);
}

}  // namespace chorus
