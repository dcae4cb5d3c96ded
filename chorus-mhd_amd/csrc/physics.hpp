// Pointwise ideal-MHD physics of the reference, as device functions templated
// on the arithmetic type.  In the FP64 translation unit (built with
// -ffp-contract=off) every expression keeps the reference's association, so
// results are bitwise those of the reference build, which emits no FMA.
#pragma once

#include <hip/hip_runtime.h>

namespace chorus {

constexpr int NV = 9;

template <class R> __device__ __forceinline__ R sq(R x) { return x * x; }
__device__ __forceinline__ double dsqrt_(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float dsqrt_(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double dabs_(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float dabs_(float x) { return __builtin_fabsf(x); }
// quotient: IEEE division in FP64; reciprocal times numerator in FP32 (mixed form)
__device__ __forceinline__ double div_(double a, double b) { return a / b; }
__device__ __forceinline__ float div_(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// getIfluxvectors (getrusanov.f90:111-153): ideal-MHD fluxes F, G of Qi(1:8)
template <class R>
__device__ __forceinline__ void iflux(R gam, const R *Qi, R *Fi, R *Gi) {
    R rho = Qi[0], rhou = Qi[1], rhov = Qi[2], rhow = Qi[3], rhoE = Qi[4];
    R Bx = Qi[5], By = Qi[6], Bz = Qi[7];
    R NORM_B = sq(Bx) + sq(By) + sq(Bz);
    R pr = (gam - R(1)) * (rhoE - R(0.5) * (sq(rhou) + sq(rhov) + sq(rhow)) / rho - R(0.5) * NORM_B);
    Fi[0] = rhou;
    Fi[1] = sq(rhou) / rho + pr + (R(0.5) * NORM_B - sq(Bx));
    Fi[2] = rhou * rhov / rho - Bx * By;
    Fi[3] = rhou * rhow / rho - Bx * Bz;
    Fi[4] = (rhou / rho) * (rhoE + pr + R(0.5) * NORM_B) -
            (rhou / rho * Bx + rhov / rho * By + rhow / rho * Bz) * Bx;
    Fi[5] = R(0);
    Fi[6] = rhou / rho * By - rhov / rho * Bx;
    Fi[7] = rhou / rho * Bz - rhow / rho * Bx;
    Gi[0] = rhov;
    Gi[1] = rhou * rhov / rho - Bx * By;
    Gi[2] = sq(rhov) / rho + pr + (R(0.5) * NORM_B - sq(By));
    Gi[3] = rhow * rhov / rho - By * Bz;
    Gi[4] = (rhov / rho) * (rhoE + pr + R(0.5) * NORM_B) -
            (rhou / rho * Bx + rhov / rho * By + rhow / rho * Bz) * By;
    Gi[5] = rhov / rho * Bx - rhou / rho * By;
    Gi[6] = R(0);
    Gi[7] = rhov / rho * Bz - rhow / rho * By;
}

// getrusanovflux (getrusanov.f90:1-109).  Returns the left-cell flux Fnl
// (already scaled by |S|*sign_l, getrusanov.f90:93), the A_z upwind value
// (getrusanov.f90:99-103) and the wave speed eigv (getrusanov.f90:87).  The
// right cell receives (sign_l*sign_r)*Fnl (getrusanov.f90:96).
template <class R>
__device__ __forceinline__ void rusanov(R gam, const R *Qfl, const R *Qfr, R nx, R ny, int sign_l,
                                        R *Fnl, R &Az, R &eigv) {
    R magnorm = dsqrt_(sq(nx) + sq(ny));
    R unit_x = nx / magnorm * R(sign_l);
    R unit_y = ny / magnorm * R(sign_l);
    R rhoi = R(1) / Qfl[0];
    R ul = rhoi * Qfl[1], vl = rhoi * Qfl[2], wl = rhoi * Qfl[3];
    R Bxl = Qfl[5], Byl = Qfl[6], Bzl = Qfl[7];
    R NORM_Bl = sq(Bxl) + sq(Byl) + sq(Bzl);
    R pl = (gam - R(1)) * (Qfl[4] - R(0.5) * Qfl[0] * (sq(ul) + sq(vl) + sq(wl)) - NORM_Bl * R(0.5));
    pl = pl + R(0.5) * NORM_Bl;
    R Vn_l = ul * unit_x + vl * unit_y;
    R Bn_l = Bxl * unit_x + Byl * unit_y;
    rhoi = R(1) / Qfr[0];
    R ur = rhoi * Qfr[1], vr = rhoi * Qfr[2], wr = rhoi * Qfr[3];
    R Bxr = Qfr[5], Byr = Qfr[6], Bzr = Qfr[7];
    R NORM_Br = sq(Bxr) + sq(Byr) + sq(Bzr);
    R pr = (gam - R(1)) * (Qfr[4] - R(0.5) * Qfr[0] * (sq(ur) + sq(vr) + sq(wr)) - NORM_Br * R(0.5));
    pr = pr + NORM_Br * R(0.5);
    R Vn_r = ur * unit_x + vr * unit_y;
    R Bn_r = Bxr * unit_x + Byr * unit_y;
    R fl[8], fr[8];
    fl[0] = Qfl[0] * Vn_l;
    fl[1] = Qfl[1] * Vn_l + pl * unit_x - Bxl * Bn_l;
    fl[2] = Qfl[2] * Vn_l + pl * unit_y - Byl * Bn_l;
    fl[3] = Qfl[3] * Vn_l - Bzl * Bn_l;
    fl[4] = (Qfl[4] + pl) * Vn_l - Bn_l * (ul * Bxl + vl * Byl + wl * Bzl);
    fl[5] = (vl * Bxl - ul * Byl) * unit_y;
    fl[6] = (ul * Byl - vl * Bxl) * unit_x;
    fl[7] = Bzl * Vn_l - wl * Bn_l;
    fr[0] = Qfr[0] * Vn_r;
    fr[1] = Qfr[1] * Vn_r + pr * unit_x - Bxr * Bn_r;
    fr[2] = Qfr[2] * Vn_r + pr * unit_y - Byr * Bn_r;
    fr[3] = Qfr[3] * Vn_r - Bzr * Bn_r;
    fr[4] = (Qfr[4] + pr) * Vn_r - Bn_r * (ur * Bxr + vr * Byr + wr * Bzr);
    fr[5] = (vr * Bxr - ur * Byr) * unit_y;
    fr[6] = (ur * Byr - vr * Bxr) * unit_x;
    fr[7] = Bzr * Vn_r - wr * Bn_r;
    pl = pl - R(0.5) * NORM_Bl;
    pr = pr - R(0.5) * NORM_Br;
    R Bx = R(0.5) * (Bxl + Bxr), By = R(0.5) * (Byl + Byr), Bz = R(0.5) * (Bzl + Bzr);
    R c_a = gam * (pl + pr) / (Qfl[0] + Qfr[0]);
    R c_d = (sq(Bx) + sq(By) + sq(Bz)) / (R(0.5) * (Qfl[0] + Qfr[0]));
    R c_n = sq(Bx * unit_x + By * unit_y) / (R(0.5) * (Qfl[0] + Qfr[0]));
    R c_tmp = dsqrt_(sq(c_a + c_d) - R(4) * c_a * c_n);
    R c_f = dsqrt_(R(0.5) * (c_a + c_d + c_tmp));
    eigv = R(0.5) * dabs_(Vn_l + Vn_r) + c_f;
    R sl = magnorm * R(sign_l);
#pragma unroll
    for (int k = 0; k < 8; k++) Fnl[k] = sl * (R(0.5) * (fl[k] + fr[k] - eigv * (Qfr[k] - Qfl[k])));
    Az = (Vn_l > R(0)) ? Qfl[8] : Qfr[8];
}

// Right state of the inviscid flux at a boundary face point of kind `kind`
// (BoundaryKind) from the cell's own trace ql and the metric row (s1, s2) of
// the face point (any orientation: the mirror is symmetric in n).  FP64.
__device__ __forceinline__ void boundary_state(int kind, const double *ql, double s1, double s2, const double *qin,
                                               double *qr) {
#pragma unroll
    for (int v = 0; v < 9; v++) qr[v] = ql[v];
    if (kind == 1) {
        const double mag = sqrt(s1 * s1 + s2 * s2);
        const double nx = s1 / mag, ny = s2 / mag;
        const double mn = ql[1] * nx + ql[2] * ny;
        const double bn = ql[5] * nx + ql[6] * ny;
        qr[1] = ql[1] - 2.0 * mn * nx;
        qr[2] = ql[2] - 2.0 * mn * ny;
        qr[5] = ql[5] - 2.0 * bn * nx;
        qr[6] = ql[6] - 2.0 * bn * ny;
    } else if (kind == 2) {
#pragma unroll
        for (int v = 0; v < 8; v++) qr[v] = qin[v];
    }
}

// getVfluxvectors (getrusanov.f90:157-249) with viscosity m and resistivity e.
// getVfluxvectors_AV (getrusanov.f90:251-361) is the same body with m = mu_art
// and e = 0.  nQ[d*9+v] = nablaQ(v,d).  Entries 6..9 of the outputs are the
// resistive terms; entry 9 is zero.
template <class R>
__device__ __forceinline__ void vflux(R gam, R lambda, R prandt, R m, R e, const R *Qv, const R *nQ,
                                      R *Fv, R *Gv) {
    R rho = Qv[0], rhou = Qv[1], rhov = Qv[2], rhow = Qv[3], rhoE = Qv[4];
    R Bx = Qv[5], By = Qv[6], Bz = Qv[7];
    R u = div_(rhou, rho), v = div_(rhov, rho), w = div_(rhow, rho);
    R NORM_U = sq(u) + sq(v) + sq(w);
    R NORM_B = sq(Bx) + sq(By) + sq(Bz);
    R inte = div_(rhoE - R(0.5) * NORM_U * rho - R(0.5) * NORM_B, rho);
    R drho[2], du[2], dv[2], dw[2], dBx[2], dBy[2], dBz[2], de[2];
#pragma unroll
    for (int d = 0; d < 2; d++) {
        const R *g = nQ + d * NV;
        drho[d] = g[0];
        du[d] = div_(g[1] - drho[d] * u, rho);
        dv[d] = div_(g[2] - drho[d] * v, rho);
        dw[d] = div_(g[3] - drho[d] * w, rho);
        dBx[d] = g[5]; dBy[d] = g[6]; dBz[d] = g[7];
    }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        R dKE = R(0.5) * NORM_U * drho[d] + rho * (u * du[d] + v * dv[d] + w * dw[d]);
        R dpre_B = Bx * dBx[d] + By * dBy[d] + Bz * dBz[d];
        de[d] = div_(nQ[d * NV + 4] - dKE - dpre_B - drho[d] * inte, rho);
    }
    R ux = du[0], uy = du[1], vx = dv[0], vy = dv[1], wx = dw[0], wy = dw[1];
    R ex = de[0], ey = de[1];
    R Bxy = dBx[1], Byx = dBy[0], Bzx = dBz[0], Bzy = dBz[1];
    Fv[0] = R(0);
    Fv[1] = m * (R(2) * ux + lambda * (ux + vy));
    Fv[2] = m * (vx + uy);
    Fv[3] = m * wx;
    Fv[4] = u * Fv[1] + v * Fv[2] + w * Fv[3] + div_(m * gam * ex, prandt);
    Gv[0] = R(0);
    Gv[1] = m * (vx + uy);
    Gv[2] = m * (R(2) * vy + lambda * (ux + vy));
    Gv[3] = m * wy;
    Gv[4] = Gv[1] * u + Gv[2] * v + Gv[3] * w + div_(m * gam * ey, prandt);
    R Jx = Bzy - R(0), Jy = R(0) - Bzx, Jz = Byx - Bxy;
    Fv[5] = R(0);
    Fv[6] = e * Jz;
    Fv[7] = -e * Jy;
    Fv[4] = Fv[4] + Fv[6] * By + Fv[7] * Bz;
    Gv[5] = -e * Jz;
    Gv[6] = R(0);
    Gv[7] = e * Jx;
    Gv[4] = Gv[4] + Gv[5] * Bx + Gv[7] * Bz;
    Fv[8] = R(0);
    Gv[8] = R(0);
}

// Derivatives of the bilinear cell map (mapder.f90:14-26) contracted with the
// vertex coordinates in calcjacob.f90:110-119 order.  X[k], Y[k]: vertex k.
__device__ __forceinline__ void map_derivs(const double *X, const double *Y, double psi, double eta,
                                           double &xpsi, double &ypsi, double &xeta, double &yeta) {
    const double da[4] = {eta - 1.0, 1.0 - eta, eta, -eta};
    const double db[4] = {psi - 1.0, -psi, psi, 1.0 - psi};
    xpsi = 0.0; ypsi = 0.0; xeta = 0.0; yeta = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        xpsi = xpsi + da[k] * X[k];
        ypsi = ypsi + da[k] * Y[k];
        xeta = xeta + db[k] * X[k];
        yeta = yeta + db[k] * Y[k];
    }
}

// The same derivatives for an 8-node cell: nodes 1-4 the corners, 5-8 the
// edge nodes (5 on edge 1-2, 6 on 2-3, 7 on 3-4, 8 on 4-1), with the
// derivatives of the 8-node shape functions of MAPDER_8_NODE
// (mapder.f90:117-253), summed in calcjacob's order (calcjacob.f90:20-28).
__device__ __forceinline__ void map_derivs8(const double *X, const double *Y, double psi, double eta,
                                            double &xpsi, double &ypsi, double &xeta, double &yeta) {
    const double da[8] = {(1.0 - eta) * (4.0 * psi + 2.0 * eta - 3.0), (1.0 - eta) * (4.0 * psi - 2.0 * eta - 1.0),
                          eta * (4.0 * psi + 2.0 * eta - 3.0),        eta * (4.0 * psi - 2.0 * eta - 1.0),
                          4.0 * (1.0 - eta) * (1.0 - 2.0 * psi),       4.0 * eta * (1.0 - eta),
                          4.0 * eta * (1.0 - 2.0 * psi),               4.0 * eta * (eta - 1.0)};
    const double db[8] = {(1.0 - psi) * (4.0 * eta + 2.0 * psi - 3.0), psi * (4.0 * eta - 2.0 * psi - 1.0),
                          psi * (4.0 * eta + 2.0 * psi - 3.0),        (1.0 - psi) * (4.0 * eta - 2.0 * psi - 1.0),
                          4.0 * psi * (psi - 1.0),                     4.0 * psi * (1.0 - 2.0 * eta),
                          4.0 * psi * (1.0 - psi),                     4.0 * (1.0 - psi) * (1.0 - 2.0 * eta)};
    xpsi = 0.0; ypsi = 0.0; xeta = 0.0; yeta = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        xpsi = xpsi + da[k] * X[k];
        ypsi = ypsi + da[k] * Y[k];
        xeta = xeta + db[k] * X[k];
        yeta = yeta + db[k] * Y[k];
    }
}

}  // namespace chorus
