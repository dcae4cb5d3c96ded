/* Oracle: plain-C FP64 restatement of the CHORUS-MHD hot path.
 *
 * TEST INFRASTRUCTURE ONLY -- the checker, never the thing measured or shipped.
 *
 * Each function names the reference routine (file:line) it restates.  Loop
 * nests, summation order and the association of every expression follow the
 * reference so that, built without FMA contraction, results match the
 * compiled reference bit for bit.  Arrays use the reference's column-major
 * layouts; the index macros below spell them out.
 *
 * Pinning (tests/test_oracle.py): bitwise against fixtures written by the
 * compiled, unmodified reference (tests/golden/), plus closed-form checks of
 * the operators, metrics, fluxes and RK coefficients.
 */
#define _GNU_SOURCE
#include "chorus_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NV 9

struct orc_ctx {
    orc_params p;
    int N, ncell, nvert, nface, ninter, npair, nfree;
    double gam, lambda, prandt, pi;
    int *ivcell, *if2c, *ibfinter, *cyc2fpair, *ibfree;
    /* every face without a right cell (IBFREE, IBFOUT, IBFSYMP, IBFWAL, IBFINL)
     * and its kind: 0 zero gradient, 1 mirror, 2 prescribed inlet state */
    int nbf, *ibf, *bkind;
    double *xv, *yv;
    double *Xs, *Xf, *Lmat, *Mmat, *Xs_low, *Xf_low;
    double *dmdxs, *dmdxf1, *dmdxf2;
    double *Jac, *S1, *S2, *XXsolu;
    int iface2fp[4 * 16], jface2fp[4 * 16];
    double *Q, *F1, *G2, *Fv1, *Gv2, *resid, *Qvfi, *Qvfj;
    double *nablaQs, *nablaQvfi, *nablaQvfj, *Azfi, *Azfj, *nablaAs;
    double *muAV, *muAVfi, *muAVfj, *SM_IND;
    double eigvmax, eigvmax_g, epsilon0;
    /* 8-node cells (orc_set_nodes8): mid-edge nodes 5..8 per cell, XM(NCELL,4) */
    double *xm8, *ym8;
};

/* ---- layouts (0-based arguments) --------------------------------------- */
#define IDX_Q(c, v, is, js, ic)   ((((size_t)(ic) * (c)->N + (js)) * (c)->N + (is)) * NV + (v))
#define IDX_FI(c, nv, v, ifp, jfp, ic) ((((size_t)(ic) * (c)->N + (jfp)) * ((c)->N + 1) + (ifp)) * (nv) + (v))
#define IDX_FJ(c, nv, v, ifp, jfp, ic) ((((size_t)(ic) * ((c)->N + 1) + (jfp)) * (c)->N + (ifp)) * (nv) + (v))
#define IDX_NQS(c, v, d, is, js, ic) (((((size_t)(ic) * (c)->N + (js)) * (c)->N + (is)) * 2 + (d)) * NV + (v))
#define IDX_NQFI(c, v, d, ifp, jfp, ic) (((((size_t)(ic) * (c)->N + (jfp)) * ((c)->N + 1) + (ifp)) * 2 + (d)) * NV + (v))
#define IDX_NQFJ(c, v, d, ifp, jfp, ic) (((((size_t)(ic) * ((c)->N + 1) + (jfp)) * (c)->N + (ifp)) * 2 + (d)) * NV + (v))
#define IDX_AFI(c, ifp, jfp, ic)  (((size_t)(ic) * (c)->N + (jfp)) * ((c)->N + 1) + (ifp))
#define IDX_AFJ(c, ifp, jfp, ic)  (((size_t)(ic) * ((c)->N + 1) + (jfp)) * (c)->N + (ifp))
#define IDX_NAS(c, d, is, js, ic) ((((size_t)(ic) * (c)->N + (js)) * (c)->N + (is)) * 2 + (d))
#define IDX_JAC(c, is, js, ic)    (((size_t)(ic) * (c)->N + (js)) * (c)->N + (is))
/* S1(a,b,ifp,jfp,ic) dims (2,2,N+1,N,NC); S2 dims (2,2,N,N+1,NC) */
#define IDX_S1(c, a, b, ifp, jfp, ic) (((((size_t)(ic) * (c)->N + (jfp)) * ((c)->N + 1) + (ifp)) * 2 + (b)) * 2 + (a))
#define IDX_S2(c, a, b, ifp, jfp, ic) (((((size_t)(ic) * ((c)->N + 1) + (jfp)) * (c)->N + (ifp)) * 2 + (b)) * 2 + (a))
#define LM(c, M, ifp, is) ((M)[(size_t)(is) * ((c)->N + 1) + (ifp)])
#define IF2C(c, f, k)   ((c)->if2c[(size_t)(k) * (c)->nface + (f)])
#define IVCELL(c, ic, k) ((c)->ivcell[(size_t)(k) * (c)->ncell + (ic)])
#define MUV(c, arr, v, side, ic) ((arr)[((size_t)(ic) * 2 + (side)) * NV + (v)])

static double sq(double x) { return x * x; }

/* ======================================================================== */
/* topology: connectivity.f90:1-233, connect_bdry.f90:1-354, map_cyc_loc.f90 */
/* ======================================================================== */
int orc_topology(int ncell, int nvert, const int *ivcell, const double *xv, const double *yv,
                 int nbnd, const int *bnd, const int *btype, double dxcycl, double dycycl,
                 orc_topo *out)
{
    const int IVFACE[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    const double tolCYC = 1e-5;
#define IVC(ic, k) ivcell[(size_t)(k) * ncell + (ic)]
    memset(out, 0, sizeof(*out));
    /* vertex -> cell CSR (connectivity.f90:29-70) */
    int *ifill = calloc((size_t)nvert, sizeof(int));
    int *icvsta = malloc(sizeof(int) * ((size_t)nvert + 1));
    for (int ic = 0; ic < ncell; ic++)
        for (int k = 0; k < 4; k++) ifill[IVC(ic, k) - 1]++;
    for (int iv = 0; iv < nvert; iv++)
        if (ifill[iv] == 0) { free(ifill); free(icvsta); return -1; }
    int kk = 0;
    int *idummy = malloc(sizeof(int) * (size_t)nvert);
    for (int iv = 0; iv < nvert; iv++) { icvsta[iv] = kk; idummy[iv] = kk; kk += ifill[iv]; }
    icvsta[nvert] = kk;
    int *icvert = malloc(sizeof(int) * (size_t)kk);
    for (int ic = 0; ic < ncell; ic++)
        for (int k = 0; k < 4; k++) { int iv = IVC(ic, k) - 1; icvert[idummy[iv]++] = ic; }
    /* face discovery (connectivity.f90:96-151) */
    int maxface = ncell * 4;
    int *ic2f = calloc((size_t)ncell * 4, sizeof(int));      /* 1-based face ids, 0 = unset */
    int *f2c = calloc((size_t)maxface * 2, sizeof(int));     /* 1-based cells, 0 = none */
    int *f2v = calloc((size_t)maxface * 2, sizeof(int));
    int iface = 0;
    for (int ic = 0; ic < ncell; ic++) {
        for (int ifa = 0; ifa < 4; ifa++) {
            int ifaceold = iface;
            if (ic2f[ic * 4 + ifa] != 0) continue;
            int ivbase = IVC(ic, IVFACE[ifa][0]);
            int iv1 = IVC(ic, IVFACE[ifa][1]);
            int found = 0;
            for (int k = icvsta[ivbase - 1]; k < icvsta[ivbase] && !found; k++) {
                int icc = icvert[k];
                if (icc == ic) continue;
                for (int iffa = 0; iffa < 4; iffa++) {
                    int ivv1 = IVC(icc, IVFACE[iffa][0]);
                    int ivv2 = IVC(icc, IVFACE[iffa][1]);
                    if ((ivbase == ivv1 && iv1 == ivv2) || (ivbase == ivv2 && iv1 == ivv1)) {
                        iface++;
                        ic2f[icc * 4 + iffa] = iface;
                        f2v[(iface - 1) * 2 + 0] = ivbase;
                        f2v[(iface - 1) * 2 + 1] = iv1;
                        f2c[(iface - 1) * 2 + 0] = ic + 1;
                        f2c[(iface - 1) * 2 + 1] = icc + 1;
                        found = 1;
                        break;
                    }
                }
            }
            if (!found) {
                if (f2c[ifaceold * 2 + 0] == 0) {
                    f2c[ifaceold * 2 + 0] = ic + 1;
                    f2c[ifaceold * 2 + 1] = 0;
                    iface = ifaceold + 1;
                    f2v[(iface - 1) * 2 + 0] = ivbase;
                    f2v[(iface - 1) * 2 + 1] = iv1;
                }
            }
            ic2f[ic * 4 + ifa] = iface;
        }
    }
    int nface = iface;
    int *if2c = calloc((size_t)nface * 4, sizeof(int));
    int *if2v = calloc((size_t)nface * 2, sizeof(int));
    for (int f = 0; f < nface; f++) {
        if2c[f] = f2c[f * 2]; if2c[nface + f] = f2c[f * 2 + 1];
        if2v[f] = f2v[f * 2]; if2v[nface + f] = f2v[f * 2 + 1];
    }
    /* local face ids of left/right cells (connectivity.f90:166-215) */
    for (int f = 0; f < nface; f++) {
        int c1 = if2c[f] - 1, c2 = if2c[nface + f] - 1;
        int v1 = if2v[f], v2 = if2v[nface + f];
        int a = IVC(c1, 0), b = IVC(c1, 1), cc = IVC(c1, 2), d = IVC(c1, 3);
        int lf;
        if (a == v1 && b == v2) lf = 1;
        else if (b == v1 && cc == v2) lf = 2;
        else if (cc == v1 && d == v2) lf = 3;
        else if (d == v1 && a == v2) lf = 4;
        else { lf = -1; }
        if (lf < 0) { free(if2c); free(if2v); goto fail; }
        if2c[2 * nface + f] = lf;
        if (c2 >= 0) {
            a = IVC(c2, 0); b = IVC(c2, 1); cc = IVC(c2, 2); d = IVC(c2, 3);
            int rf = 0;
            if (b == v1 && a == v2) rf = 1;
            else if (cc == v1 && b == v2) rf = 2;
            else if (d == v1 && cc == v2) rf = 3;
            else if (a == v1 && d == v2) rf = 4;
            if2c[3 * nface + f] = rf;
        }
    }
    /* boundary attachment (connect_bdry.f90:126-223) */
    int *ibfcyc = malloc(sizeof(int) * ((size_t)nbnd + 1));
    int *ibfree = malloc(sizeof(int) * ((size_t)nbnd + 1));
    int ncycl = 0, nfree = 0;
    int *lists[4], nlist[4] = {0, 0, 0, 0};   /* INLE, OUTL, SYMP, WALL */
    for (int k = 0; k < 4; k++) lists[k] = malloc(sizeof(int) * ((size_t)nbnd + 1));
    for (int ib = 0; ib < nbnd; ib++) {
        int ivb1 = bnd[ib], ivb3 = bnd[nbnd + ib];
        int icc = -1;
        for (int k1 = icvsta[ivb1 - 1]; k1 < icvsta[ivb1]; k1++)
            for (int k3 = icvsta[ivb3 - 1]; k3 < icvsta[ivb3]; k3++)
                if (icvert[k1] == icvert[k3]) icc = icvert[k1];
        if (icc < 0) continue;
        int ifa;
        for (ifa = 0; ifa < 4; ifa++) {
            int v1 = IVC(icc, IVFACE[ifa][0]), v2 = IVC(icc, IVFACE[ifa][1]);
            if ((v1 == ivb1 || v2 == ivb1) && (v1 == ivb3 || v2 == ivb3)) break;
        }
        if (ifa == 4) continue;
        int k = ic2f[icc * 4 + ifa];
        if (btype[ib] == 1) ibfcyc[ncycl++] = k;
        else if (btype[ib] == 2) ibfree[nfree++] = k;
        else if (btype[ib] >= 3 && btype[ib] <= 6) lists[btype[ib] - 3][nlist[btype[ib] - 3]++] = k;
    }
    out->ninl = nlist[0]; out->ibfinl = lists[0];
    out->nout = nlist[1]; out->ibfout = lists[1];
    out->nsymp = nlist[2]; out->ibfsymp = lists[2];
    out->nwal = nlist[3]; out->ibfwal = lists[3];
    int ninter = 0;
    for (int f = 0; f < nface; f++) if (if2c[nface + f] != 0) ninter++;
    int *ibfinter = malloc(sizeof(int) * ((size_t)ninter + 1));
    ninter = 0;
    for (int f = 0; f < nface; f++) if (if2c[nface + f] != 0) ibfinter[ninter++] = f + 1;
    /* local periodic faces (connect_bdry.f90:281-308): every CYCL face must have a partner */
    int *ibfcycloc = malloc(sizeof(int) * ((size_t)ncycl + 1));
    int ncycloc = 0;
    for (int ib = 0; ib < ncycl; ib++) {
        int K = ibfcyc[ib] - 1;
        double xfin = (xv[if2v[K] - 1] + xv[if2v[nface + K] - 1]) / 2;
        double yfin = (yv[if2v[K] - 1] + yv[if2v[nface + K] - 1]) / 2;
        int hit = 0;
        for (int ib2 = 0; ib2 < ncycl && !hit; ib2++) {
            int K2 = ibfcyc[ib2] - 1;
            double xfou = (xv[if2v[K2] - 1] + xv[if2v[nface + K2] - 1]) / 2;
            double yfou = (yv[if2v[K2] - 1] + yv[if2v[nface + K2] - 1]) / 2;
            double epsX = fabs(fabs(xfin - xfou) - dxcycl);
            double epsY = fabs(fabs(yfin - yfou) - dycycl);
            if ((fabs(xfin - xfou) < tolCYC && epsY < tolCYC) ||
                (fabs(yfin - yfou) < tolCYC && epsX < tolCYC)) hit = 1;
        }
        if (!hit) {
            free(ibfcycloc); free(ibfinter); free(if2c); free(if2v); free(ibfcyc); free(ibfree);
            for (int k = 0; k < 4; k++) free(lists[k]);
            goto fail;
        }
        ibfcycloc[ncycloc++] = ibfcyc[ib];
    }
    /* MAPCYCLOC (map_cyc_loc.f90:12-40) */
    int *pair = malloc(sizeof(int) * ((size_t)ncycloc + 2));
    int npair = 0;
    for (int nf = 0; nf < ncycloc; nf++) {
        int K = ibfcycloc[nf] - 1;
        double xfin = (xv[if2v[K] - 1] + xv[if2v[nface + K] - 1]) / 2;
        double yfin = (yv[if2v[K] - 1] + yv[if2v[nface + K] - 1]) / 2;
        for (int nf2 = nf + 1; nf2 < ncycloc; nf2++) {
            int K2 = ibfcycloc[nf2] - 1;
            double xfou = (xv[if2v[K2] - 1] + xv[if2v[nface + K2] - 1]) / 2;
            double yfou = (yv[if2v[K2] - 1] + yv[if2v[nface + K2] - 1]) / 2;
            double epsX = fabs(fabs(xfin - xfou) - dxcycl);
            double epsY = fabs(fabs(yfin - yfou) - dycycl);
            if ((epsX < tolCYC && fabs(yfin - yfou) < tolCYC) ||
                (epsY < tolCYC && fabs(xfin - xfou) < tolCYC)) {
                if (npair >= ncycloc / 2) { npair = -1; break; }
                pair[npair * 2 + 0] = K + 1;
                pair[npair * 2 + 1] = K2 + 1;
                npair++;
            }
        }
        if (npair < 0) break;
    }
    if (npair != ncycloc / 2) {
        free(pair); free(ibfcycloc); free(ibfinter); free(if2c); free(if2v); free(ibfcyc); free(ibfree);
        for (int k = 0; k < 4; k++) free(lists[k]);
        goto fail;
    }
    out->nface = nface; out->ninter = ninter; out->npair = npair; out->nfree = nfree;
    out->if2c = if2c; out->if2v = if2v; out->ibfinter = ibfinter; out->ibfree = ibfree;
    out->cyc2fpair = malloc(sizeof(int) * ((size_t)npair * 2 + 1));
    for (int p = 0; p < npair; p++) {
        out->cyc2fpair[p] = pair[p * 2];
        out->cyc2fpair[npair + p] = pair[p * 2 + 1];
    }
    free(pair); free(ibfcycloc); free(ibfcyc);
    free(ifill); free(icvsta); free(idummy); free(icvert); free(ic2f); free(f2c); free(f2v);
    return 0;
fail:
    free(ifill); free(icvsta); free(idummy); free(icvert); free(ic2f); free(f2c); free(f2v);
    return -2;
#undef IVC
}

void orc_topo_free(orc_topo *t)
{
    free(t->if2c); free(t->if2v); free(t->ibfinter); free(t->cyc2fpair); free(t->ibfree);
    free(t->ibfinl); free(t->ibfout); free(t->ibfsymp); free(t->ibfwal);
    memset(t, 0, sizeof(*t));
}

/* ======================================================================== */
/* set-up: initsetup.f90, mapder.f90, calcjacob.f90                          */
/* ======================================================================== */

/* setsandpoints (initsetup.f90:167-255) */
static void setsandpoints(int np, double pi, double *xs, double *xf)
{
    for (int i = 1; i <= np; i++)
        xs[i - 1] = 0.5 * (1.0 - cos((double)(2 * i - 1) * pi / (2 * (double)np)));
    double *f = xf;
    switch (np) {
    case 1: f[0] = 0.0; f[1] = 1.0; break;
    case 2: f[0] = 0.0; f[1] = 0.5; f[2] = 1.0; break;
    case 3:
        f[0] = 0.0; f[1] = (1.0 - sqrt(3.0) / 3.0) / 2.0; f[2] = (1.0 + sqrt(3.0) / 3.0) / 2.0; f[3] = 1.0;
        break;
    case 4:
        f[0] = 0.0; f[1] = (1.0 - sqrt(15.0) / 5.0) / 2.0; f[2] = 0.5;
        f[3] = (1.0 + sqrt(15.0) / 5.0) / 2.0; f[4] = 1.0;
        break;
    case 5:
        f[0] = 0.0;
        f[1] = (1.0 - sqrt(525.0 + 70.0 * sqrt(30.0)) / 35.0) / 2.0;
        f[2] = (1.0 - sqrt(525.0 - 70.0 * sqrt(30.0)) / 35.0) / 2.0;
        f[3] = (1.0 + sqrt(525.0 - 70.0 * sqrt(30.0)) / 35.0) / 2.0;
        f[4] = (1.0 + sqrt(525.0 + 70.0 * sqrt(30.0)) / 35.0) / 2.0;
        f[5] = 1.0;
        break;
    case 6:
        f[0] = 0.0;
        f[1] = (1.0 - sqrt(245.0 + 14.0 * sqrt(70.0)) / 21.0) / 2.0;
        f[2] = (1.0 - sqrt(245.0 - 14.0 * sqrt(70.0)) / 21.0) / 2.0;
        f[3] = 0.5;
        f[4] = (1.0 + sqrt(245.0 - 14.0 * sqrt(70.0)) / 21.0) / 2.0;
        f[5] = (1.0 + sqrt(245.0 + 14.0 * sqrt(70.0)) / 21.0) / 2.0;
        f[6] = 1.0;
        break;
    default: {
        /* 12-digit Gauss-Legendre roots for np = 7..10 (initsetup.f90:211-252) */
        static const double r7[3] = {0.932469514203, 0.661209386466, 0.238619186083};
        static const double r8[4] = {0.949107912343, 0.741531185599, 0.405845151377, 0.0};
        static const double r9[4] = {0.960289856498, 0.796666477414, 0.525532409916, 0.183434642496};
        static const double r10[5] = {0.968160239508, 0.836031107327, 0.613371432701, 0.324253423404, 0.0};
        const double *r = np == 7 ? r7 : np == 8 ? r8 : np == 9 ? r9 : r10;
        int h = (np - 1) / 2;  /* roots on each side of the middle */
        f[0] = 0.0;
        for (int k = 0; k < h; k++) f[1 + k] = (1.0 - r[k]) / 2.0;
        if ((np - 1) % 2 == 1) f[1 + h] = 0.5;
        for (int k = 0; k < h; k++) f[np - 1 - k] = (1.0 + r[k]) / 2.0;
        f[np] = 1.0;
        break;
    }
    }
}

/* hhval (initsetup.f90:318-340) / hhval_low (AV.f90:176-197) */
static double hhval(int np, const double *xs, int i, double xval)
{
    double hvaln = 1.0, hvald = 1.0;
    for (int s = 0; s < np; s++)
        if (s != i) { hvaln = hvaln * (xval - xs[s]); hvald = hvald * (xs[i] - xs[s]); }
    return hvaln / hvald;
}

/* computeLandM (initsetup.f90:258-315) */
static void compute_LM(orc_ctx *c)
{
    int np = c->N;
    for (int ifp = 0; ifp <= np; ifp++)
        for (int is = 0; is < np; is++) LM(c, c->Lmat, ifp, is) = hhval(np, c->Xs, is, c->Xf[ifp]);
    for (int ifp = 0; ifp <= np; ifp++)
        for (int is = 0; is < np; is++) {
            double xval = c->Xs[is], sum = 0.0;
            for (int k = 0; k <= np; k++) {
                if (k == ifp) continue;
                double num = 1.0, den = 1.0;
                for (int s = 0; s <= np; s++) {
                    if (s != ifp && s != k) num = num * (xval - c->Xf[s]);
                    if (s != ifp) den = den * (c->Xf[ifp] - c->Xf[s]);
                }
                sum = sum + num / den;
            }
            LM(c, c->Mmat, ifp, is) = sum;
        }
}

/* shape-function derivatives of the bilinear map (mapder.f90:14-26) */
static void dmdx4(double psi, double eta, double d[4][2])
{
    d[0][0] = eta - 1.0; d[0][1] = psi - 1.0;
    d[1][0] = 1.0 - eta; d[1][1] = -psi;
    d[2][0] = eta;       d[2][1] = psi;
    d[3][0] = -eta;      d[3][1] = 1.0 - psi;
}

/* MAPDER_8_NODE (mapder.f90:117-253): derivatives d/dpsi, d/deta of the
 * 8-node serendipity shape functions at (psi, eta); the reference's REAL
 * literals (4., 2., 3., 1.) are exact in double precision */
static void dmdx8(double psi, double eta, double d[8][2])
{
    d[0][0] = (1.0 - eta) * (4.0 * psi + 2.0 * eta - 3.0);
    d[0][1] = (1.0 - psi) * (4.0 * eta + 2.0 * psi - 3.0);
    d[1][0] = (1.0 - eta) * (4.0 * psi - 2.0 * eta - 1.0);
    d[1][1] = psi * (4.0 * eta - 2.0 * psi - 1.0);
    d[2][0] = eta * (4.0 * psi + 2.0 * eta - 3.0);
    d[2][1] = psi * (4.0 * eta + 2.0 * psi - 3.0);
    d[3][0] = eta * (4.0 * psi - 2.0 * eta - 1.0);
    d[3][1] = (1.0 - psi) * (4.0 * eta - 2.0 * psi - 1.0);
    d[4][0] = 4.0 * (1.0 - eta) * (1.0 - 2.0 * psi);
    d[4][1] = 4.0 * psi * (psi - 1.0);
    d[5][0] = 4.0 * eta * (1.0 - eta);
    d[5][1] = 4.0 * psi * (1.0 - 2.0 * eta);
    d[6][0] = 4.0 * eta * (1.0 - 2.0 * psi);
    d[6][1] = 4.0 * psi * (1.0 - psi);
    d[7][0] = 4.0 * eta * (eta - 1.0);
    d[7][1] = 4.0 * (1.0 - psi) * (1.0 - 2.0 * eta);
}

/* MapBaseFunc8 (mapder.f90:92-104) */
static void base8(double xi, double eta, double f[8])
{
    f[0] = (1.0 - xi) * (1.0 - eta) * (-2.0 * xi - 2.0 * eta + 1.0);
    f[1] = xi * (1.0 - eta) * (2.0 * xi - 2.0 * eta - 1.0);
    f[2] = xi * eta * (2.0 * xi + 2.0 * eta - 3.0);
    f[3] = (1.0 - xi) * eta * (-2.0 * xi + 2.0 * eta - 1.0);
    f[4] = 4.0 * (1.0 - eta) * (1.0 - xi) * xi;
    f[5] = 4.0 * eta * (1.0 - eta) * xi;
    f[6] = 4.0 * (1.0 - xi) * xi * eta;
    f[7] = 4.0 * (1.0 - eta) * (1.0 - xi) * eta;
}

void orc_mapder8(int N, double *dmdxs, double *dmdxf1, double *dmdxf2)
{
    double pi = atan(1.0) * 4.0, Xs[16], Xf[17], d[8][2];
    setsandpoints(N, pi, Xs, Xf);
    for (int js = 0; js < N; js++)
        for (int is = 0; is < N; is++) {
            dmdx8(Xs[is], Xs[js], d);
            for (int k = 0; k < 8; k++)
                for (int a = 0; a < 2; a++) dmdxs[(((size_t)js * N + is) * 2 + a) * 8 + k] = d[k][a];
        }
    for (int jfp = 0; jfp < N; jfp++)
        for (int ifp = 0; ifp <= N; ifp++) {
            dmdx8(Xf[ifp], Xs[jfp], d);
            for (int k = 0; k < 8; k++)
                for (int a = 0; a < 2; a++) dmdxf1[(((size_t)jfp * (N + 1) + ifp) * 2 + a) * 8 + k] = d[k][a];
        }
    for (int jfp = 0; jfp <= N; jfp++)
        for (int ifp = 0; ifp < N; ifp++) {
            dmdx8(Xs[ifp], Xf[jfp], d);
            for (int k = 0; k < 8; k++)
                for (int a = 0; a < 2; a++) dmdxf2[(((size_t)jfp * N + ifp) * 2 + a) * 8 + k] = d[k][a];
        }
}

/* map derivatives at (psi, eta) from the cell's nn nodes X, Y in calcjacob's
 * order (calcjacob.f90:20-28); nn = 8 takes MAPDER_8_NODE's derivatives */
static void map_derivs(int nn, const double *X, const double *Y, double psi, double eta, double *xpsi,
                       double *ypsi, double *xeta, double *yeta)
{
    double d[8][2];
    if (nn == 8) dmdx8(psi, eta, d);
    else dmdx4(psi, eta, (double (*)[2])d);
    *xpsi = 0.0; *xeta = 0.0; *ypsi = 0.0; *yeta = 0.0;
    for (int k = 0; k < nn; k++) {
        *xpsi = *xpsi + d[k][0] * X[k];
        *ypsi = *ypsi + d[k][0] * Y[k];
        *xeta = *xeta + d[k][1] * X[k];
        *yeta = *yeta + d[k][1] * Y[k];
    }
}

/* calcjacob (calcjacob.f90:1-86) + xyCoor_atGps (mapder.f90:255-273); with
 * mid-edge nodes (orc_set_nodes8) the same sums run over the 8 nodes, which
 * the reference prepares (MAPDER_8_NODE) but never reaches (initsetup.f90:71-78) */
static void metrics(orc_ctx *c)
{
    int N = c->N, nn = c->xm8 ? 8 : 4;
    for (int ic = 0; ic < c->ncell; ic++) {
        double X[8], Y[8];
        for (int k = 0; k < 4; k++) { X[k] = c->xv[IVCELL(c, ic, k) - 1]; Y[k] = c->yv[IVCELL(c, ic, k) - 1]; }
        if (nn == 8)
            for (int k = 0; k < 4; k++) {
                X[4 + k] = c->xm8[(size_t)k * c->ncell + ic];
                Y[4 + k] = c->ym8[(size_t)k * c->ncell + ic];
            }
        double xpsi, ypsi, xeta, yeta;
        for (int js = 0; js < N; js++)
            for (int is = 0; is < N; is++) {
                map_derivs(nn, X, Y, c->Xs[is], c->Xs[js], &xpsi, &ypsi, &xeta, &yeta);
                c->Jac[IDX_JAC(c, is, js, ic)] = xpsi * yeta - xeta * ypsi;
            }
        for (int jfp = 0; jfp < N; jfp++)
            for (int ifp = 0; ifp <= N; ifp++) {
                map_derivs(nn, X, Y, c->Xf[ifp], c->Xs[jfp], &xpsi, &ypsi, &xeta, &yeta);
                c->S1[IDX_S1(c, 0, 0, ifp, jfp, ic)] = yeta;
                c->S1[IDX_S1(c, 0, 1, ifp, jfp, ic)] = -xeta;
                c->S1[IDX_S1(c, 1, 0, ifp, jfp, ic)] = -ypsi;
                c->S1[IDX_S1(c, 1, 1, ifp, jfp, ic)] = xpsi;
            }
        for (int jfp = 0; jfp <= N; jfp++)
            for (int ifp = 0; ifp < N; ifp++) {
                map_derivs(nn, X, Y, c->Xs[ifp], c->Xf[jfp], &xpsi, &ypsi, &xeta, &yeta);
                c->S2[IDX_S2(c, 0, 0, ifp, jfp, ic)] = yeta;
                c->S2[IDX_S2(c, 0, 1, ifp, jfp, ic)] = -xeta;
                c->S2[IDX_S2(c, 1, 0, ifp, jfp, ic)] = -ypsi;
                c->S2[IDX_S2(c, 1, 1, ifp, jfp, ic)] = xpsi;
            }
        /* solution-point coordinates, MapBaseFunc / MapBaseFunc8 (mapder.f90:79-104) */
        for (int js = 0; js < N; js++)
            for (int is = 0; is < N; is++) {
                double xi = c->Xs[is], eta = c->Xs[js];
                double fn[8] = {(1.0 - xi) * (1.0 - eta), (1.0 - eta) * xi, xi * eta, (1.0 - xi) * eta};
                if (nn == 8) base8(xi, eta, fn);
                double xx = 0.0, yy = 0.0;
                for (int nb = 0; nb < nn; nb++) xx = xx + X[nb] * fn[nb];
                for (int nb = 0; nb < nn; nb++) yy = yy + Y[nb] * fn[nb];
                c->XXsolu[(IDX_JAC(c, is, js, ic)) * 2 + 0] = xx;
                c->XXsolu[(IDX_JAC(c, is, js, ic)) * 2 + 1] = yy;
            }
    }
}

/* face2fpconnec (initsetup.f90:343-379), 0-based flux-point indices */
static void face2fp(orc_ctx *c)
{
    int N = c->N;
    for (int nfp = 0; nfp < N; nfp++) {
        c->iface2fp[0 * 16 + nfp] = nfp;           c->jface2fp[0 * 16 + nfp] = 0;
        c->iface2fp[1 * 16 + nfp] = N;             c->jface2fp[1 * 16 + nfp] = nfp;
        c->iface2fp[2 * 16 + nfp] = N - 1 - nfp;   c->jface2fp[2 * 16 + nfp] = N;
        c->iface2fp[3 * 16 + nfp] = 0;             c->jface2fp[3 * 16 + nfp] = N - 1 - nfp;
    }
}

static double *dalloc(size_t n) { return calloc(n ? n : 1, sizeof(double)); }

orc_ctx *orc_create(const orc_params *p, int ncell, int nvert, const int *ivcell,
                    const double *xv, const double *yv, const orc_topo *t)
{
    if (p->N < 2 || p->N > 10) return NULL;
    orc_ctx *c = calloc(1, sizeof(orc_ctx));
    c->p = *p;
    c->N = p->N; c->ncell = ncell; c->nvert = nvert;
    c->nface = t->nface; c->ninter = t->ninter; c->npair = t->npair; c->nfree = t->nfree;
    c->gam = p->gam;
    c->lambda = -2.0 / 3.0;     /* setup2d.f90:70 */
    c->prandt = 0.72;           /* setup2d.f90:71 */
    c->pi = atan(1.0) * 4.0;    /* setup2d.f90:67 */
    c->ivcell = malloc(sizeof(int) * (size_t)ncell * 4);
    memcpy(c->ivcell, ivcell, sizeof(int) * (size_t)ncell * 4);
    c->xv = dalloc((size_t)nvert); memcpy(c->xv, xv, sizeof(double) * (size_t)nvert);
    c->yv = dalloc((size_t)nvert); memcpy(c->yv, yv, sizeof(double) * (size_t)nvert);
    c->if2c = malloc(sizeof(int) * ((size_t)t->nface * 4 + 1));
    memcpy(c->if2c, t->if2c, sizeof(int) * (size_t)t->nface * 4);
    c->ibfinter = malloc(sizeof(int) * ((size_t)t->ninter + 1));
    memcpy(c->ibfinter, t->ibfinter, sizeof(int) * (size_t)t->ninter);
    c->cyc2fpair = malloc(sizeof(int) * ((size_t)t->npair * 2 + 1));
    memcpy(c->cyc2fpair, t->cyc2fpair, sizeof(int) * (size_t)t->npair * 2);
    c->ibfree = malloc(sizeof(int) * ((size_t)t->nfree + 1));
    memcpy(c->ibfree, t->ibfree, sizeof(int) * (size_t)t->nfree);
    {
        const int *src[5] = {t->ibfree, t->ibfout, t->ibfsymp, t->ibfwal, t->ibfinl};
        const int cnt[5] = {t->nfree, t->nout, t->nsymp, t->nwal, t->ninl};
        const int kind[5] = {0, 0, 1, 1, 2};
        int n = 0;
        for (int k = 0; k < 5; k++) n += cnt[k];
        c->ibf = malloc(sizeof(int) * ((size_t)n + 1));
        c->bkind = malloc(sizeof(int) * ((size_t)n + 1));
        c->nbf = 0;
        for (int k = 0; k < 5; k++)
            for (int i = 0; i < cnt[k]; i++) {
                c->ibf[c->nbf] = src[k][i];
                c->bkind[c->nbf++] = kind[k];
            }
    }

    int N = c->N;
    size_t nc = (size_t)ncell;
    c->Xs = dalloc(N); c->Xf = dalloc(N + 1);
    c->Lmat = dalloc((size_t)(N + 1) * N); c->Mmat = dalloc((size_t)(N + 1) * N);
    setsandpoints(N, c->pi, c->Xs, c->Xf);
    compute_LM(c);
    c->Xs_low = dalloc(N - 1); c->Xf_low = dalloc(N);
    setsandpoints(N - 1, c->pi, c->Xs_low, c->Xf_low);
    c->Jac = dalloc(nc * N * N);
    c->S1 = dalloc(nc * 4 * (N + 1) * N);
    c->S2 = dalloc(nc * 4 * (N + 1) * N);
    c->XXsolu = dalloc(nc * 2 * N * N);
    metrics(c);
    face2fp(c);
    c->Q = dalloc(nc * NV * N * N);
    c->resid = dalloc(nc * NV * N * N);
    c->F1 = dalloc(nc * 8 * (N + 1) * N);
    c->G2 = dalloc(nc * 8 * (N + 1) * N);
    c->Fv1 = dalloc(nc * NV * (N + 1) * N);
    c->Gv2 = dalloc(nc * NV * (N + 1) * N);
    c->Qvfi = dalloc(nc * NV * (N + 1) * N);
    c->Qvfj = dalloc(nc * NV * (N + 1) * N);
    c->nablaQs = dalloc(nc * NV * 2 * N * N);
    c->nablaQvfi = dalloc(nc * NV * 2 * (N + 1) * N);
    c->nablaQvfj = dalloc(nc * NV * 2 * (N + 1) * N);
    c->Azfi = dalloc(nc * (N + 1) * N);
    c->Azfj = dalloc(nc * (N + 1) * N);
    c->nablaAs = dalloc(nc * 2 * N * N);
    c->muAV = dalloc(nc * NV);
    c->muAVfi = dalloc(nc * NV * 2);
    c->muAVfj = dalloc(nc * NV * 2);
    c->SM_IND = dalloc(nc * NV);
    return c;
}

int orc_set_nodes8(orc_ctx *c, const double *xm, const double *ym)
{
    if (!c || !xm || !ym) return -1;
    if (!c->xm8) { c->xm8 = dalloc((size_t)c->ncell * 4); c->ym8 = dalloc((size_t)c->ncell * 4); }
    memcpy(c->xm8, xm, sizeof(double) * (size_t)c->ncell * 4);
    memcpy(c->ym8, ym, sizeof(double) * (size_t)c->ncell * 4);
    metrics(c);
    return 0;
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    double *ds[] = {c->xv, c->yv, c->Xs, c->Xf, c->Lmat, c->Mmat, c->Xs_low, c->Xf_low, c->Jac,
                    c->S1, c->S2, c->XXsolu, c->Q, c->resid, c->F1, c->G2, c->Fv1, c->Gv2,
                    c->Qvfi, c->Qvfj, c->nablaQs, c->nablaQvfi, c->nablaQvfj, c->Azfi, c->Azfj,
                    c->nablaAs, c->muAV, c->muAVfi, c->muAVfj, c->SM_IND};
    for (size_t i = 0; i < sizeof(ds) / sizeof(ds[0]); i++) free(ds[i]);
    free(c->ivcell); free(c->if2c); free(c->ibfinter); free(c->cyc2fpair); free(c->ibfree);
    free(c->ibf); free(c->bkind);
    free(c->xm8); free(c->ym8);
    free(c);
}

void orc_set_q(orc_ctx *c, const double *q)
{
    memcpy(c->Q, q, sizeof(double) * (size_t)c->ncell * NV * c->N * c->N);
}

void orc_get_q(const orc_ctx *c, double *q)
{
    memcpy(q, c->Q, sizeof(double) * (size_t)c->ncell * NV * c->N * c->N);
}

/* ======================================================================== */
/* pointwise physics: getrusanov.f90                                          */
/* ======================================================================== */

/* getrusanovflux (getrusanov.f90:1-109) */
static void rusanov(orc_ctx *c, const double *Qfl, const double *Qfr, double *Fnl, double *Fnr,
                    double *Az, const double *normf, int sign_l, int sign_r)
{
    double gam = c->gam;
    double magnorm = sqrt(sq(normf[0]) + sq(normf[1]));
    double unit_x = normf[0] / magnorm * sign_l;
    double unit_y = normf[1] / magnorm * sign_l;
    double rhoi = 1.0 / Qfl[0];
    double ul = rhoi * Qfl[1], vl = rhoi * Qfl[2], wl = rhoi * Qfl[3];
    double Bxl = Qfl[5], Byl = Qfl[6], Bzl = Qfl[7];
    double NORM_Bl = sq(Bxl) + sq(Byl) + sq(Bzl);
    double pl = (gam - 1.0) * (Qfl[4] - 0.5 * Qfl[0] * (sq(ul) + sq(vl) + sq(wl)) - NORM_Bl * 0.5);
    pl = pl + 0.5 * NORM_Bl;
    double Vn_l = ul * unit_x + vl * unit_y;
    double Bn_l = Bxl * unit_x + Byl * unit_y;
    rhoi = 1.0 / Qfr[0];
    double ur = rhoi * Qfr[1], vr = rhoi * Qfr[2], wr = rhoi * Qfr[3];
    double Bxr = Qfr[5], Byr = Qfr[6], Bzr = Qfr[7];
    double NORM_Br = sq(Bxr) + sq(Byr) + sq(Bzr);
    double pr = (gam - 1.0) * (Qfr[4] - 0.5 * Qfr[0] * (sq(ur) + sq(vr) + sq(wr)) - NORM_Br * 0.5);
    pr = pr + NORM_Br * 0.5;
    double Vn_r = ur * unit_x + vr * unit_y;
    double Bn_r = Bxr * unit_x + Byr * unit_y;
    double fl[8], fr[8];
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
    pl = pl - 0.5 * NORM_Bl;
    pr = pr - 0.5 * NORM_Br;
    double Bx = 0.5 * (Bxl + Bxr), By = 0.5 * (Byl + Byr), Bz = 0.5 * (Bzl + Bzr);
    double c_a = gam * (pl + pr) / (Qfl[0] + Qfr[0]);
    double c_d = (sq(Bx) + sq(By) + sq(Bz)) / (0.5 * (Qfl[0] + Qfr[0]));
    double c_n = sq(Bx * unit_x + By * unit_y) / (0.5 * (Qfl[0] + Qfr[0]));
    double c_tmp = sqrt(sq(c_a + c_d) - 4.0 * c_a * c_n);
    double c_f = sqrt(0.5 * (c_a + c_d + c_tmp));
    double eigv = 0.5 * fabs(Vn_l + Vn_r) + c_f;
    double sl = magnorm * sign_l;
    double slr = (double)(sign_l * sign_r);
    for (int k = 0; k < 8; k++) {
        double f = 0.5 * (fl[k] + fr[k] - eigv * (Qfr[k] - Qfl[k]));
        Fnl[k] = sl * f;
        Fnr[k] = slr * Fnl[k];
    }
    *Az = (Vn_l > 0.0) ? Qfl[8] : Qfr[8];
    if (eigv > c->eigvmax) c->eigvmax = eigv;
}

/* getIfluxvectors (getrusanov.f90:111-153) */
static void iflux(double gam, const double *Qi, double *Fi, double *Gi)
{
    double rho = Qi[0], rhou = Qi[1], rhov = Qi[2], rhow = Qi[3], rhoE = Qi[4];
    double Bx = Qi[5], By = Qi[6], Bz = Qi[7];
    double NORM_B = sq(Bx) + sq(By) + sq(Bz);
    double pr = (gam - 1.0) * (rhoE - 0.5 * (sq(rhou) + sq(rhov) + sq(rhow)) / rho - 0.5 * NORM_B);
    Fi[0] = rhou;
    Fi[1] = sq(rhou) / rho + pr + (0.5 * NORM_B - sq(Bx));
    Fi[2] = rhou * rhov / rho - Bx * By;
    Fi[3] = rhou * rhow / rho - Bx * Bz;
    Fi[4] = (rhou / rho) * (rhoE + pr + 0.5 * NORM_B) - (rhou / rho * Bx + rhov / rho * By + rhow / rho * Bz) * Bx;
    Fi[5] = 0.0;
    Fi[6] = rhou / rho * By - rhov / rho * Bx;
    Fi[7] = rhou / rho * Bz - rhow / rho * Bx;
    Gi[0] = rhov;
    Gi[1] = rhou * rhov / rho - Bx * By;
    Gi[2] = sq(rhov) / rho + pr + (0.5 * NORM_B - sq(By));
    Gi[3] = rhow * rhov / rho - By * Bz;
    Gi[4] = (rhov / rho) * (rhoE + pr + 0.5 * NORM_B) - (rhou / rho * Bx + rhov / rho * By + rhow / rho * Bz) * By;
    Gi[5] = rhov / rho * Bx - rhou / rho * By;
    Gi[6] = 0.0;
    Gi[7] = rhov / rho * Bz - rhow / rho * By;
}

/* getVfluxvectors (getrusanov.f90:157-249) with viscosity m and resistivity e;
 * getVfluxvectors_AV (getrusanov.f90:251-361) is the same body with m = mu_art, e = 0 */
static void vflux(const orc_ctx *c, double m, double e, const double *Qv, const double *nQ,
                  double *Fv, double *Gv)
{
    double gam = c->gam, lambda = c->lambda, prandt = c->prandt;
    double rho = Qv[0], rhou = Qv[1], rhov = Qv[2], rhow = Qv[3], rhoE = Qv[4];
    double Bx = Qv[5], By = Qv[6], Bz = Qv[7];
    double u = rhou / rho, v = rhov / rho, w = rhow / rho;
    double NORM_U = sq(u) + sq(v) + sq(w);
    double NORM_B = sq(Bx) + sq(By) + sq(Bz);
    double inte = (rhoE - 0.5 * NORM_U * rho - 0.5 * NORM_B) / rho;
    double drho[2], du[2], dv[2], dw[2], dBx[2], dBy[2], dBz[2], dKE[2], dpre_B[2], de[2];
    for (int d = 0; d < 2; d++) {
        const double *g = nQ + d * NV;     /* nablaQ(1:9,d) */
        drho[d] = g[0];
        du[d] = (g[1] - drho[d] * u) / rho;
        dv[d] = (g[2] - drho[d] * v) / rho;
        dw[d] = (g[3] - drho[d] * w) / rho;
        dBx[d] = g[5]; dBy[d] = g[6]; dBz[d] = g[7];
    }
    for (int d = 0; d < 2; d++) {
        dKE[d] = 0.5 * NORM_U * drho[d] + rho * (u * du[d] + v * dv[d] + w * dw[d]);
        dpre_B[d] = Bx * dBx[d] + By * dBy[d] + Bz * dBz[d];
    }
    for (int d = 0; d < 2; d++)
        de[d] = (nQ[d * NV + 4] - dKE[d] - dpre_B[d] - drho[d] * inte) / rho;
    double ux = du[0], uy = du[1], vx = dv[0], vy = dv[1], wx = dw[0], wy = dw[1];
    double ex = de[0], ey = de[1];
    double Bxy = dBx[1], Byx = dBy[0], Bzx = dBz[0], Bzy = dBz[1];
    Fv[0] = 0.0;
    Fv[1] = m * (2.0 * ux + lambda * (ux + vy));
    Fv[2] = m * (vx + uy);
    Fv[3] = m * wx;
    Fv[4] = u * Fv[1] + v * Fv[2] + w * Fv[3] + m * gam * ex / prandt;
    Gv[0] = 0.0;
    Gv[1] = m * (vx + uy);
    Gv[2] = m * (2.0 * vy + lambda * (ux + vy));
    Gv[3] = m * wy;
    Gv[4] = Gv[1] * u + Gv[2] * v + Gv[3] * w + m * gam * ey / prandt;
    double Jx = Bzy - 0.0, Jy = 0.0 - Bzx, Jz = Byx - Bxy;
    Fv[5] = 0.0;
    Fv[6] = e * Jz;
    Fv[7] = -e * Jy;
    Fv[4] = Fv[4] + Fv[6] * By + Fv[7] * Bz;
    Gv[5] = -e * Jz;
    Gv[6] = 0.0;
    Gv[7] = e * Jx;
    Gv[4] = Gv[4] + Gv[5] * Bx + Gv[7] * Bz;
    Fv[8] = 0.0;
    Gv[8] = 0.0;
}

/* ======================================================================== */
/* flux-point work: compflux.f90, interfaceflux_all.f90, cyclocflux_all.f90,  */
/* bcflux.f90, magnetic_potential.f90                                         */
/* ======================================================================== */

/* inviscid_flux_interior (compflux.f90:51-104) */
static void inviscid_flux_interior(orc_ctx *c)
{
    int N = c->N;
    double Qf[NV], F[8], G[8];
    for (int ic = 0; ic < c->ncell; ic++) {
        for (int jfp = 0; jfp < N; jfp++)
            for (int ifp = 1; ifp < N; ifp++) {
                for (int v = 0; v < NV; v++) Qf[v] = 0.0;
                for (int sp = 0; sp < N; sp++)
                    for (int v = 0; v < NV; v++)
                        Qf[v] = Qf[v] + c->Q[IDX_Q(c, v, sp, jfp, ic)] * LM(c, c->Lmat, ifp, sp);
                if (c->p.vismode == 1) {
                    for (int v = 0; v < NV; v++) c->Qvfi[IDX_FI(c, NV, v, ifp, jfp, ic)] = Qf[v];
                    c->Azfi[IDX_AFI(c, ifp, jfp, ic)] = Qf[8];
                }
                iflux(c->gam, Qf, F, G);
                double s11 = c->S1[IDX_S1(c, 0, 0, ifp, jfp, ic)], s12 = c->S1[IDX_S1(c, 0, 1, ifp, jfp, ic)];
                for (int v = 0; v < 8; v++) c->F1[IDX_FI(c, 8, v, ifp, jfp, ic)] = F[v] * s11 + G[v] * s12;
            }
        for (int jfp = 1; jfp < N; jfp++)
            for (int ifp = 0; ifp < N; ifp++) {
                for (int v = 0; v < NV; v++) Qf[v] = 0.0;
                for (int sp = 0; sp < N; sp++)
                    for (int v = 0; v < NV; v++)
                        Qf[v] = Qf[v] + c->Q[IDX_Q(c, v, ifp, sp, ic)] * LM(c, c->Lmat, jfp, sp);
                if (c->p.vismode == 1) {
                    for (int v = 0; v < NV; v++) c->Qvfj[IDX_FJ(c, NV, v, ifp, jfp, ic)] = Qf[v];
                    c->Azfj[IDX_AFJ(c, ifp, jfp, ic)] = Qf[8];
                }
                iflux(c->gam, Qf, F, G);
                double s21 = c->S2[IDX_S2(c, 1, 0, ifp, jfp, ic)], s22 = c->S2[IDX_S2(c, 1, 1, ifp, jfp, ic)];
                for (int v = 0; v < 8; v++) c->G2[IDX_FJ(c, 8, v, ifp, jfp, ic)] = F[v] * s21 + G[v] * s22;
            }
    }
}

/* trace of Q at flux point (ifp,jfp) of cell ic on face family fam */
static void trace_q(const orc_ctx *c, int fam, int ifp, int jfp, int ic, double *Qf)
{
    int N = c->N;
    for (int v = 0; v < NV; v++) Qf[v] = 0.0;
    if (fam == 1) {
        for (int sp = 0; sp < N; sp++)
            for (int v = 0; v < NV; v++) Qf[v] = Qf[v] + c->Q[IDX_Q(c, v, sp, jfp, ic)] * LM(c, c->Lmat, ifp, sp);
    } else {
        for (int sp = 0; sp < N; sp++)
            for (int v = 0; v < NV; v++) Qf[v] = Qf[v] + c->Q[IDX_Q(c, v, ifp, sp, ic)] * LM(c, c->Lmat, jfp, sp);
    }
}

static void store_face(orc_ctx *c, int fam, int ifp, int jfp, int ic, const double *Fn, double Az)
{
    if (fam == 1) {
        for (int v = 0; v < 8; v++) c->F1[IDX_FI(c, 8, v, ifp, jfp, ic)] = Fn[v];
        c->Azfi[IDX_AFI(c, ifp, jfp, ic)] = Az;
    } else {
        for (int v = 0; v < 8; v++) c->G2[IDX_FJ(c, 8, v, ifp, jfp, ic)] = Fn[v];
        c->Azfj[IDX_AFJ(c, ifp, jfp, ic)] = Az;
    }
}

static void store_qv(orc_ctx *c, int fam, int ifp, int jfp, int ic, const double *Qfl, const double *Qfr)
{
    for (int v = 0; v < NV; v++) {
        double a = 0.5 * Qfl[v] + 0.5 * Qfr[v];
        if (fam == 1) c->Qvfi[IDX_FI(c, NV, v, ifp, jfp, ic)] = a;
        else c->Qvfj[IDX_FJ(c, NV, v, ifp, jfp, ic)] = a;
    }
}

/* one two-sided face: interfaceflux (interfaceflux_all.f90:1-116) and
 * CYCLOCFLUX (cyclocflux_all.f90:1-116) share this body */
static void two_sided_face(orc_ctx *c, int icleft, int icright, int ifacelc, int ifacerc)
{
    int N = c->N;
    int faml = ifacelc % 2 + 1, famr = ifacerc % 2 + 1;
    double Qfl[NV], Qfr[NV], Fnl[8], Fnr[8], normf[2], Az;
    for (int nfp = 0; nfp < N; nfp++) {
        int ifpl = c->iface2fp[(ifacelc - 1) * 16 + nfp], jfpl = c->jface2fp[(ifacelc - 1) * 16 + nfp];
        int sign_l = 1;
        trace_q(c, faml, ifpl, jfpl, icleft, Qfl);
        if (faml == 1) {
            if (ifpl == 0) sign_l = -1;
            normf[0] = c->S1[IDX_S1(c, 0, 0, ifpl, jfpl, icleft)];
            normf[1] = c->S1[IDX_S1(c, 0, 1, ifpl, jfpl, icleft)];
        } else {
            if (jfpl == 0) sign_l = -1;
            normf[0] = c->S2[IDX_S2(c, 1, 0, ifpl, jfpl, icleft)];
            normf[1] = c->S2[IDX_S2(c, 1, 1, ifpl, jfpl, icleft)];
        }
        int ifpr = c->iface2fp[(ifacerc - 1) * 16 + (N - 1 - nfp)];
        int jfpr = c->jface2fp[(ifacerc - 1) * 16 + (N - 1 - nfp)];
        int sign_r = 1;
        trace_q(c, famr, ifpr, jfpr, icright, Qfr);
        if (famr == 1) { if (ifpr == N) sign_r = -1; }
        else { if (jfpr == N) sign_r = -1; }
        if (c->p.vismode == 1) {
            store_qv(c, faml, ifpl, jfpl, icleft, Qfl, Qfr);
            store_qv(c, famr, ifpr, jfpr, icright, Qfl, Qfr);
        }
        rusanov(c, Qfl, Qfr, Fnl, Fnr, &Az, normf, sign_l, sign_r);
        store_face(c, faml, ifpl, jfpl, icleft, Fnl, Az);
        store_face(c, famr, ifpr, jfpr, icright, Fnr, Az);
    }
}

/* Right state of a face without a right cell.  FREE (bcflux.f90:46) and OUTL:
 * the left state.  The reference gives INLE/SYMP/WALL faces no flux
 * (compflux.f90:33-49); this build states: SYMP and WALL mirror the normal
 * momentum and the normal magnetic field about the face's unit normal, INLE
 * takes the prescribed state for Q(1:8) and the cell's own A_z. */
static void boundary_state(const orc_ctx *c, int kind, const double *Qfl, const double *normf, double *Qfr)
{
    for (int v = 0; v < NV; v++) Qfr[v] = Qfl[v];
    if (kind == 1) {
        double mag = sqrt(normf[0] * normf[0] + normf[1] * normf[1]);
        double nx = normf[0] / mag, ny = normf[1] / mag;
        double mn = Qfl[1] * nx + Qfl[2] * ny;
        double bn = Qfl[5] * nx + Qfl[6] * ny;
        Qfr[1] = Qfl[1] - 2.0 * mn * nx;
        Qfr[2] = Qfl[2] - 2.0 * mn * ny;
        Qfr[5] = Qfl[5] - 2.0 * bn * nx;
        Qfr[6] = Qfl[6] - 2.0 * bn * ny;
    } else if (kind == 2) {
        for (int v = 0; v < 8; v++) Qfr[v] = c->p.qin[v];
    }
}

/* BCFLUX (bcflux.f90:1-70) over every face without a right cell: the
 * Rusanov flux against the face kind's right state; the central flux-point
 * values stay the cell's own (Qfr = Qfl, bcflux.f90:46-52) for every kind */
static void bcflux(orc_ctx *c)
{
    int N = c->N;
    double Qfl[NV], Qfr[NV], Fnl[8], Fnr[8], normf[2], Az;
    for (int ifn = 0; ifn < c->nbf; ifn++) {
        int iface = c->ibf[ifn] - 1;
        int icleft = IF2C(c, iface, 0) - 1, ifacelc = IF2C(c, iface, 2);
        int faml = ifacelc % 2 + 1;
        for (int nfp = 0; nfp < N; nfp++) {
            int ifpl = c->iface2fp[(ifacelc - 1) * 16 + nfp], jfpl = c->jface2fp[(ifacelc - 1) * 16 + nfp];
            int sign_l = 1, sign_r = 1;
            trace_q(c, faml, ifpl, jfpl, icleft, Qfl);
            if (faml == 1) {
                if (ifpl == 0) sign_l = -1;
                normf[0] = c->S1[IDX_S1(c, 0, 0, ifpl, jfpl, icleft)];
                normf[1] = c->S1[IDX_S1(c, 0, 1, ifpl, jfpl, icleft)];
            } else {
                if (jfpl == 0) sign_l = -1;
                normf[0] = c->S2[IDX_S2(c, 1, 0, ifpl, jfpl, icleft)];
                normf[1] = c->S2[IDX_S2(c, 1, 1, ifpl, jfpl, icleft)];
            }
            if (c->p.vismode == 1) store_qv(c, faml, ifpl, jfpl, icleft, Qfl, Qfl);
            boundary_state(c, c->bkind[ifn], Qfl, normf, Qfr);
            rusanov(c, Qfl, Qfr, Fnl, Fnr, &Az, normf, sign_l, sign_r);
            store_face(c, faml, ifpl, jfpl, icleft, Fnl, Az);
        }
    }
}

/* comp_inviscid_flux (compflux.f90:33-49); PROCINTFLUX/CYCREMFLUX are empty on one rank */
static void comp_inviscid_flux(orc_ctx *c)
{
    inviscid_flux_interior(c);
    for (int k = 0; k < c->ninter; k++) {
        int f = c->ibfinter[k] - 1;
        two_sided_face(c, IF2C(c, f, 0) - 1, IF2C(c, f, 1) - 1, IF2C(c, f, 2), IF2C(c, f, 3));
    }
    for (int p = 0; p < c->npair; p++) {
        int K1 = c->cyc2fpair[p] - 1, K2 = c->cyc2fpair[c->npair + p] - 1;
        two_sided_face(c, IF2C(c, K1, 0) - 1, IF2C(c, K2, 0) - 1, IF2C(c, K1, 2), IF2C(c, K2, 2));
    }
    bcflux(c);
}

/* nablaAsp (magnetic_potential.f90:1-42) */
static void nablaAsp(orc_ctx *c)
{
    int N = c->N;
    for (int ic = 0; ic < c->ncell; ic++)
        for (int js = 0; js < N; js++)
            for (int is = 0; is < N; is++) {
                double dApsi = 0.0, dAeta = 0.0;
                for (int r = 0; r <= N; r++)
                    dApsi = dApsi + c->Azfi[IDX_AFI(c, r, js, ic)] * LM(c, c->Mmat, r, is) * c->S1[IDX_S1(c, 0, 0, r, js, ic)];
                for (int r = 0; r <= N; r++)
                    dAeta = dAeta + c->Azfj[IDX_AFJ(c, is, r, ic)] * LM(c, c->Mmat, r, js) * c->S2[IDX_S2(c, 1, 0, is, r, ic)];
                double J = c->Jac[IDX_JAC(c, is, js, ic)];
                c->nablaAs[IDX_NAS(c, 0, is, js, ic)] = (dApsi + dAeta) / J;
                dApsi = 0.0; dAeta = 0.0;
                for (int r = 0; r <= N; r++)
                    dApsi = dApsi + c->Azfi[IDX_AFI(c, r, js, ic)] * LM(c, c->Mmat, r, is) * c->S1[IDX_S1(c, 0, 1, r, js, ic)];
                for (int r = 0; r <= N; r++)
                    dAeta = dAeta + c->Azfj[IDX_AFJ(c, is, r, ic)] * LM(c, c->Mmat, r, js) * c->S2[IDX_S2(c, 1, 1, is, r, ic)];
                c->nablaAs[IDX_NAS(c, 1, is, js, ic)] = (dApsi + dAeta) / J;
            }
}

/* nablaQsp(v_sta,v_end) (compflux.f90:107-157); 0-based inclusive range */
static void nablaQsp(orc_ctx *c, int v0, int v1)
{
    int N = c->N;
    double dpsi[NV], deta[NV];
    for (int ic = 0; ic < c->ncell; ic++)
        for (int js = 0; js < N; js++)
            for (int is = 0; is < N; is++) {
                double J = c->Jac[IDX_JAC(c, is, js, ic)];
                for (int d = 0; d < 2; d++) {
                    for (int v = v0; v <= v1; v++) dpsi[v] = 0.0;
                    for (int r = 0; r <= N; r++) {
                        double m = LM(c, c->Mmat, r, is), s = c->S1[IDX_S1(c, 0, d, r, js, ic)];
                        for (int v = v0; v <= v1; v++)
                            dpsi[v] = dpsi[v] + c->Qvfi[IDX_FI(c, NV, v, r, js, ic)] * m * s;
                    }
                    for (int v = v0; v <= v1; v++) deta[v] = 0.0;
                    for (int r = 0; r <= N; r++) {
                        double m = LM(c, c->Mmat, r, js), s = c->S2[IDX_S2(c, 1, d, is, r, ic)];
                        for (int v = v0; v <= v1; v++)
                            deta[v] = deta[v] + c->Qvfj[IDX_FJ(c, NV, v, is, r, ic)] * m * s;
                    }
                    for (int v = v0; v <= v1; v++)
                        c->nablaQs[IDX_NQS(c, v, d, is, js, ic)] = (dpsi[v] + deta[v]) / J;
                }
            }
}

/* nablaQ_interior_FP (compflux.f90:159-196) */
static void nablaQ_interior_FP(orc_ctx *c)
{
    int N = c->N;
    double a[NV], b[NV];
    for (int ic = 0; ic < c->ncell; ic++) {
        for (int jfp = 0; jfp < N; jfp++)
            for (int ifp = 1; ifp < N; ifp++) {
                for (int v = 0; v < NV; v++) { a[v] = 0.0; b[v] = 0.0; }
                for (int sp = 0; sp < N; sp++) {
                    double l = LM(c, c->Lmat, ifp, sp);
                    for (int v = 0; v < NV; v++) {
                        a[v] = a[v] + c->nablaQs[IDX_NQS(c, v, 0, sp, jfp, ic)] * l;
                        b[v] = b[v] + c->nablaQs[IDX_NQS(c, v, 1, sp, jfp, ic)] * l;
                    }
                }
                for (int v = 0; v < NV; v++) {
                    c->nablaQvfi[IDX_NQFI(c, v, 0, ifp, jfp, ic)] = a[v];
                    c->nablaQvfi[IDX_NQFI(c, v, 1, ifp, jfp, ic)] = b[v];
                }
            }
        for (int jfp = 1; jfp < N; jfp++)
            for (int ifp = 0; ifp < N; ifp++) {
                for (int v = 0; v < NV; v++) { a[v] = 0.0; b[v] = 0.0; }
                for (int sp = 0; sp < N; sp++) {
                    double l = LM(c, c->Lmat, jfp, sp);
                    for (int v = 0; v < NV; v++) {
                        a[v] = a[v] + c->nablaQs[IDX_NQS(c, v, 0, ifp, sp, ic)] * l;
                        b[v] = b[v] + c->nablaQs[IDX_NQS(c, v, 1, ifp, sp, ic)] * l;
                    }
                }
                for (int v = 0; v < NV; v++) {
                    c->nablaQvfj[IDX_NQFJ(c, v, 0, ifp, jfp, ic)] = a[v];
                    c->nablaQvfj[IDX_NQFJ(c, v, 1, ifp, jfp, ic)] = b[v];
                }
            }
    }
}

static void trace_grad(const orc_ctx *c, int fam, int ifp, int jfp, int ic, double *g1, double *g2)
{
    int N = c->N;
    for (int v = 0; v < NV; v++) { g1[v] = 0.0; g2[v] = 0.0; }
    for (int sp = 0; sp < N; sp++) {
        if (fam == 1) {
            double l = LM(c, c->Lmat, ifp, sp);
            for (int v = 0; v < NV; v++) {
                g1[v] = g1[v] + c->nablaQs[IDX_NQS(c, v, 0, sp, jfp, ic)] * l;
                g2[v] = g2[v] + c->nablaQs[IDX_NQS(c, v, 1, sp, jfp, ic)] * l;
            }
        } else {
            double l = LM(c, c->Lmat, jfp, sp);
            for (int v = 0; v < NV; v++) {
                g1[v] = g1[v] + c->nablaQs[IDX_NQS(c, v, 0, ifp, sp, ic)] * l;
                g2[v] = g2[v] + c->nablaQs[IDX_NQS(c, v, 1, ifp, sp, ic)] * l;
            }
        }
    }
}

static void store_grad(orc_ctx *c, int fam, int ifp, int jfp, int ic, const double *l1, const double *r1,
                       const double *l2, const double *r2)
{
    for (int v = 0; v < NV; v++) {
        double a = 0.5 * l1[v] + 0.5 * r1[v], b = 0.5 * l2[v] + 0.5 * r2[v];
        if (fam == 1) {
            c->nablaQvfi[IDX_NQFI(c, v, 0, ifp, jfp, ic)] = a;
            c->nablaQvfi[IDX_NQFI(c, v, 1, ifp, jfp, ic)] = b;
        } else {
            c->nablaQvfj[IDX_NQFJ(c, v, 0, ifp, jfp, ic)] = a;
            c->nablaQvfj[IDX_NQFJ(c, v, 1, ifp, jfp, ic)] = b;
        }
    }
}

/* interfaceVISflux (interfaceflux_all.f90:118-199) / CYCLOCVISFLUX (cyclocflux_all.f90:118-208) */
static void two_sided_visface(orc_ctx *c, int icleft, int icright, int ifacelc, int ifacerc)
{
    int N = c->N;
    int faml = ifacelc % 2 + 1, famr = ifacerc % 2 + 1;
    double l1[NV], l2[NV], r1[NV], r2[NV];
    for (int nfp = 0; nfp < N; nfp++) {
        int ifpl = c->iface2fp[(ifacelc - 1) * 16 + nfp], jfpl = c->jface2fp[(ifacelc - 1) * 16 + nfp];
        trace_grad(c, faml, ifpl, jfpl, icleft, l1, l2);
        int ifpr = c->iface2fp[(ifacerc - 1) * 16 + (N - 1 - nfp)];
        int jfpr = c->jface2fp[(ifacerc - 1) * 16 + (N - 1 - nfp)];
        trace_grad(c, famr, ifpr, jfpr, icright, r1, r2);
        store_grad(c, faml, ifpl, jfpl, icleft, l1, r1, l2, r2);
        store_grad(c, famr, ifpr, jfpr, icright, l1, r1, l2, r2);
    }
}

/* BCVISFLUX (bcflux.f90:73-121), over every face without a right cell */
static void bcvisflux(orc_ctx *c)
{
    int N = c->N;
    double l1[NV], l2[NV];
    for (int ifn = 0; ifn < c->nbf; ifn++) {
        int iface = c->ibf[ifn] - 1;
        int icleft = IF2C(c, iface, 0) - 1, ifacelc = IF2C(c, iface, 2);
        int faml = ifacelc % 2 + 1;
        for (int nfp = 0; nfp < N; nfp++) {
            int ifpl = c->iface2fp[(ifacelc - 1) * 16 + nfp], jfpl = c->jface2fp[(ifacelc - 1) * 16 + nfp];
            trace_grad(c, faml, ifpl, jfpl, icleft, l1, l2);
            store_grad(c, faml, ifpl, jfpl, icleft, l1, l1, l2, l2);
        }
    }
}

/* ======================================================================== */
/* artificial viscosity: AV.f90:1-303, 494-564                               */
/* ======================================================================== */
static void comp_artificial_viscosity(orc_ctx *c)
{
    int N = c->N, Nl = N - 1;
    double *qlow = dalloc((size_t)NV * Nl * Nl), *qsm = dalloc((size_t)NV * N * N);
    double Se[NV], se_low[NV], se_high[NV], k_range[NV], s0[NV];
    for (int ic = 0; ic < c->ncell; ic++) {
        for (size_t k = 0; k < (size_t)NV * Nl * Nl; k++) qlow[k] = 0.0;
        for (int jsl = 0; jsl < Nl; jsl++)
            for (int isl = 0; isl < Nl; isl++) {
                double xsi = c->Xs_low[isl], eta = c->Xs_low[jsl];
                for (int js = 0; js < N; js++)
                    for (int is = 0; is < N; is++) {
                        double hx = hhval(N, c->Xs, is, xsi), hy = hhval(N, c->Xs, js, eta);
                        for (int v = 0; v < NV; v++) {
                            double *d = &qlow[((size_t)jsl * Nl + isl) * NV + v];
                            *d = *d + c->Q[IDX_Q(c, v, is, js, ic)] * hx * hy;
                        }
                    }
            }
        for (size_t k = 0; k < (size_t)NV * N * N; k++) qsm[k] = 0.0;
        for (int js = 0; js < N; js++)
            for (int is = 0; is < N; is++) {
                double xsi = c->Xs[is], eta = c->Xs[js];
                for (int jsl = 0; jsl < Nl; jsl++)
                    for (int isl = 0; isl < Nl; isl++) {
                        double hx = hhval(Nl, c->Xs_low, isl, xsi), hy = hhval(Nl, c->Xs_low, jsl, eta);
                        for (int v = 0; v < NV; v++) {
                            double *d = &qsm[((size_t)js * N + is) * NV + v];
                            *d = *d + qlow[((size_t)jsl * Nl + isl) * NV + v] * hx * hy;
                        }
                    }
            }
        for (int v = 0; v < NV; v++) Se[v] = 0.0;
        for (int js = 0; js < N; js++)
            for (int is = 0; is < N; is++)
                for (int v = 0; v < NV; v++)
                    if (v != 3 && v != 7) {
                        double q = c->Q[IDX_Q(c, v, is, js, ic)];
                        Se[v] = Se[v] + sq(qsm[((size_t)js * N + is) * NV + v] - q) / sq(q);
                    }
        Se[3] = 1.0; Se[7] = 1.0;
        for (int v = 0; v < NV; v++) c->SM_IND[(size_t)ic * NV + v] = log10(Se[v]);
    }
    free(qlow); free(qsm);
    for (int v = 0; v < NV; v++) {
        se_low[v] = -6.0; se_high[v] = -1.0;
        if (v == 3 || v == 6 || v == 7 || v == 8) { se_low[v] = 0.0; se_high[v] = 0.0; }
        k_range[v] = (se_high[v] - se_low[v]) * 0.5;
        s0[v] = (se_high[v] + se_low[v]) * 0.5;
    }
    c->eigvmax_g = c->eigvmax;
    double h_msh = sqrt(c->p.dxcycl * c->p.dycycl / (double)c->ncell);
    c->epsilon0 = c->eigvmax_g * h_msh / (double)(N - 1) * 1.0;
    for (int ic = 0; ic < c->ncell; ic++)
        for (int v = 0; v < NV; v++) {
            double s = c->SM_IND[(size_t)ic * NV + v], r;
            if (v == 3 || v == 6 || v == 7 || v == 8) r = 0.0;
            else if (s < se_low[v]) r = 0.0;
            else if (s > se_high[v]) r = c->epsilon0;
            else r = c->epsilon0 / 2.0 * (1.0 + sin(c->pi * (s - s0[v]) / 2.0 / k_range[v]));
            c->muAV[(size_t)ic * NV + v] = r;
        }
}

/* slot of a cell's local face in muAVfi/muAVfj (AV.f90:245-254) */
static double *muslot(orc_ctx *c, int ic, int lf)
{
    switch (lf) {
    case 1: return &MUV(c, c->muAVfj, 0, 0, ic);
    case 2: return &MUV(c, c->muAVfi, 0, 1, ic);
    case 3: return &MUV(c, c->muAVfj, 0, 1, ic);
    default: return &MUV(c, c->muAVfi, 0, 0, ic);
    }
}

static void av_face(orc_ctx *c, int icl, int icr, int lfl, int lfr)
{
    double l[NV], r[NV];
    double *pl = muslot(c, icl, lfl), *pr = muslot(c, icr, lfr);
    for (int v = 0; v < NV; v++) { l[v] = pl[v]; r[v] = pr[v]; }
    for (int v = 0; v < NV; v++) pl[v] = 0.5 * (l[v] + r[v]);
    for (int v = 0; v < NV; v++) pr[v] = 0.5 * (l[v] + r[v]);
}

/* compAVflux + interfaceAVflux + cyclocAVflux (AV.f90:199-303, 494-564) */
static void comp_av_flux(orc_ctx *c)
{
    for (int ic = 0; ic < c->ncell; ic++)
        for (int s = 0; s < 2; s++)
            for (int v = 0; v < NV; v++) {
                MUV(c, c->muAVfi, v, s, ic) = c->muAV[(size_t)ic * NV + v];
                MUV(c, c->muAVfj, v, s, ic) = c->muAV[(size_t)ic * NV + v];
            }
    for (int k = 0; k < c->ninter; k++) {
        int f = c->ibfinter[k] - 1;
        av_face(c, IF2C(c, f, 0) - 1, IF2C(c, f, 1) - 1, IF2C(c, f, 2), IF2C(c, f, 3));
    }
    for (int p = 0; p < c->npair; p++) {
        int K1 = c->cyc2fpair[p] - 1, K2 = c->cyc2fpair[c->npair + p] - 1;
        av_face(c, IF2C(c, K1, 0) - 1, IF2C(c, K2, 0) - 1, IF2C(c, K1, 2), IF2C(c, K2, 2));
    }
}

/* comp_viscous_flux (compflux.f90:198-258) */
static void comp_viscous_flux(orc_ctx *c)
{
    int N = c->N;
    double Fv[NV], Gv[NV], Fa[NV], Ga[NV], nQ[2 * NV];
    for (int ic = 0; ic < c->ncell; ic++) {
        for (int jfp = 0; jfp < N; jfp++)
            for (int ifp = 0; ifp <= N; ifp++) {
                const double *Qs = &c->Qvfi[IDX_FI(c, NV, 0, ifp, jfp, ic)];
                for (int d = 0; d < 2; d++)
                    for (int v = 0; v < NV; v++) nQ[d * NV + v] = c->nablaQvfi[IDX_NQFI(c, v, d, ifp, jfp, ic)];
                vflux(c, c->p.mu, c->p.eta0, Qs, nQ, Fv, Gv);
                double s11 = c->S1[IDX_S1(c, 0, 0, ifp, jfp, ic)], s12 = c->S1[IDX_S1(c, 0, 1, ifp, jfp, ic)];
                double *o = &c->Fv1[IDX_FI(c, NV, 0, ifp, jfp, ic)];
                for (int v = 0; v < NV; v++) o[v] = Fv[v] * s11 + Gv[v] * s12;
                if (c->p.artiv == 1) {
                    double xsi = c->Xf[ifp];
                    double m = MUV(c, c->muAVfi, 0, 0, ic) * (1.0 - xsi) + MUV(c, c->muAVfi, 0, 1, ic) * xsi;
                    vflux(c, m, 0.0, Qs, nQ, Fa, Ga);
                    for (int v = 0; v < NV; v++) o[v] = o[v] + Fa[v] * s11 + Ga[v] * s12;
                }
            }
        for (int jfp = 0; jfp <= N; jfp++)
            for (int ifp = 0; ifp < N; ifp++) {
                const double *Qs = &c->Qvfj[IDX_FJ(c, NV, 0, ifp, jfp, ic)];
                for (int d = 0; d < 2; d++)
                    for (int v = 0; v < NV; v++) nQ[d * NV + v] = c->nablaQvfj[IDX_NQFJ(c, v, d, ifp, jfp, ic)];
                vflux(c, c->p.mu, c->p.eta0, Qs, nQ, Fv, Gv);
                double s21 = c->S2[IDX_S2(c, 1, 0, ifp, jfp, ic)], s22 = c->S2[IDX_S2(c, 1, 1, ifp, jfp, ic)];
                double *o = &c->Gv2[IDX_FJ(c, NV, 0, ifp, jfp, ic)];
                for (int v = 0; v < NV; v++) o[v] = Fv[v] * s21 + Gv[v] * s22;
                if (c->p.artiv == 1) {
                    double eta = c->Xf[jfp];
                    double m = MUV(c, c->muAVfj, 0, 0, ic) * (1.0 - eta) + MUV(c, c->muAVfj, 0, 1, ic) * eta;
                    vflux(c, m, 0.0, Qs, nQ, Fa, Ga);
                    for (int v = 0; v < NV; v++) o[v] = o[v] + Fa[v] * s21 + Ga[v] * s22;
                }
            }
    }
}

/* compflux (compflux.f90:1-31) */
static void compflux(orc_ctx *c)
{
    c->eigvmax = 0.0;
    comp_inviscid_flux(c);
    nablaAsp(c);
    if (c->p.vismode == 1) {
        nablaQsp(c, 0, 8);
        nablaQ_interior_FP(c);
        for (int k = 0; k < c->ninter; k++) {
            int f = c->ibfinter[k] - 1;
            two_sided_visface(c, IF2C(c, f, 0) - 1, IF2C(c, f, 1) - 1, IF2C(c, f, 2), IF2C(c, f, 3));
        }
        bcvisflux(c);
        for (int p = 0; p < c->npair; p++) {
            int K1 = c->cyc2fpair[p] - 1, K2 = c->cyc2fpair[c->npair + p] - 1;
            two_sided_visface(c, IF2C(c, K1, 0) - 1, IF2C(c, K2, 0) - 1, IF2C(c, K1, 2), IF2C(c, K2, 2));
        }
        if (c->p.artiv == 1) {
            comp_artificial_viscosity(c);
            comp_av_flux(c);
        }
        comp_viscous_flux(c);
    }
}

/* compresid (compresid.f90:1-67) */
static void compresid(orc_ctx *c)
{
    int N = c->N, vis = c->p.vismode == 1;
    compflux(c);
    double dF[8], dG[8], dFv[NV], dGv[NV];
    for (int ic = 0; ic < c->ncell; ic++)
        for (int js = 0; js < N; js++)
            for (int is = 0; is < N; is++) {
                const double *Qs = &c->Q[IDX_Q(c, 0, is, js, ic)];
                double au = Qs[1] / Qs[0], av = Qs[2] / Qs[0];
                for (int v = 0; v < 8; v++) dF[v] = 0.0;
                for (int v = 0; v < NV; v++) dFv[v] = 0.0;
                for (int r = 0; r <= N; r++) {
                    double m = LM(c, c->Mmat, r, is);
                    for (int v = 0; v < 8; v++) dF[v] = dF[v] + c->F1[IDX_FI(c, 8, v, r, js, ic)] * m;
                    if (vis)
                        for (int v = 0; v < NV; v++) dFv[v] = dFv[v] + c->Fv1[IDX_FI(c, NV, v, r, js, ic)] * m;
                }
                for (int v = 0; v < 8; v++) dG[v] = 0.0;
                for (int v = 0; v < NV; v++) dGv[v] = 0.0;
                for (int r = 0; r <= N; r++) {
                    double m = LM(c, c->Mmat, r, js);
                    for (int v = 0; v < 8; v++) dG[v] = dG[v] + c->G2[IDX_FJ(c, 8, v, is, r, ic)] * m;
                    if (vis)
                        for (int v = 0; v < NV; v++) dGv[v] = dGv[v] + c->Gv2[IDX_FJ(c, NV, v, is, r, ic)] * m;
                }
                double *R = &c->resid[IDX_Q(c, 0, is, js, ic)];
                for (int v = 0; v < 8; v++) R[v] = -(dF[v] + dG[v]);
                R[8] = -(au * c->nablaAs[IDX_NAS(c, 0, is, js, ic)]) - av * c->nablaAs[IDX_NAS(c, 1, is, js, ic)];
                R[8] = R[8] * c->Jac[IDX_JAC(c, is, js, ic)];
                if (vis)
                    for (int v = 0; v < NV; v++) R[v] = R[v] + (dFv[v] + dGv[v]);
            }
    for (int ic = 0; ic < c->ncell; ic++)
        for (int js = 0; js < N; js++)
            for (int is = 0; is < N; is++) {
                double J = c->Jac[IDX_JAC(c, is, js, ic)];
                double *R = &c->resid[IDX_Q(c, 0, is, js, ic)];
                for (int v = 0; v < NV; v++) R[v] = R[v] / J;
            }
}

void orc_compresid(orc_ctx *c, double *resid_out)
{
    compresid(c);
    if (resid_out)
        memcpy(resid_out, c->resid, sizeof(double) * (size_t)c->ncell * NV * c->N * c->N);
}

/* correct_B_via_curl_Az (setinitialcond.f90:82-92) + update_B (magnetic_potential.f90:44-67) */
void orc_correct_b(orc_ctx *c)
{
    int N = c->N;
    comp_inviscid_flux(c);
    nablaQsp(c, 8, 8);
    for (int ic = 0; ic < c->ncell; ic++)
        for (int js = 0; js < N; js++)
            for (int is = 0; is < N; is++) {
                double Bx = c->nablaQs[IDX_NQS(c, 8, 1, is, js, ic)];
                double By = -c->nablaQs[IDX_NQS(c, 8, 0, is, js, ic)];
                c->Q[IDX_Q(c, 5, is, js, ic)] = Bx;
                c->Q[IDX_Q(c, 6, is, js, ic)] = By;
            }
}

/* RK_S5P3 (iterations.f90:162-217) */
static void rk_s5p3(orc_ctx *c)
{
    const double a30 = 0.355909775063327, a32 = 0.644090224936674;
    const double a40 = 0.367933791638137, a43 = 0.632066208361863;
    const double a52 = 0.237593836598569, a54 = 0.762406163401431;
    const double b10 = 0.377268915331368, b21 = 0.377268915331368, b32 = 0.242995220537396;
    const double b43 = 0.238458932846290, b54 = 0.287632146308408;
    size_t n = (size_t)c->ncell * NV * c->N * c->N;
    double *Qk0 = dalloc(n), *Qka = dalloc(n), *Q = c->Q, *R = c->resid, dt = c->p.dt;
    memcpy(Qk0, Q, n * sizeof(double));
    compresid(c);
    for (size_t i = 0; i < n; i++) Q[i] = Q[i] + b10 * dt * R[i];
    compresid(c);
    for (size_t i = 0; i < n; i++) Q[i] = Q[i] + b21 * dt * R[i];
    memcpy(Qka, Q, n * sizeof(double));
    compresid(c);
    for (size_t i = 0; i < n; i++) Q[i] = a30 * Qk0[i] + a32 * Q[i] + b32 * dt * R[i];
    compresid(c);
    for (size_t i = 0; i < n; i++) Q[i] = a40 * Qk0[i] + a43 * Q[i] + b43 * dt * R[i];
    compresid(c);
    for (size_t i = 0; i < n; i++) Q[i] = a52 * Qka[i] + a54 * Q[i] + b54 * dt * R[i];
    free(Qk0); free(Qka);
}

/* FiveStageRk (iterations.f90:92-160) */
static void five_stage_rk(orc_ctx *c)
{
    const double a11 = 0.39175222700392;
    const double a21 = 0.44437049406734, a22 = 0.55562950593266, a23 = 0.36841059262959;
    const double a31 = 0.6201018513854, a32 = 0.3798981486146, a33 = 0.25189177424738;
    const double a41 = 0.17807995410773, a42 = 0.82192004589227, a43 = 0.54497475021237;
    const double a51 = 0.00683325884039, a52 = 0.51723167208978, a53 = 0.12759831133288;
    const double a54 = 0.34833675773694, a55 = 0.08460416338212, a56 = 0.22600748319395;
    size_t n = (size_t)c->ncell * NV * c->N * c->N;
    double *q0 = dalloc(n), *q2 = dalloc(n), *q3 = dalloc(n), *r0 = dalloc(n);
    double *Q = c->Q, *R = c->resid, dt = c->p.dt;
    memcpy(q0, Q, n * sizeof(double));
    compresid(c);
    double rdt = dt * a11;
    for (size_t i = 0; i < n; i++) Q[i] = q0[i] + rdt * R[i];
    compresid(c);
    for (size_t i = 0; i < n; i++) Q[i] = a21 * q0[i] + a22 * Q[i] + a23 * R[i] * dt;
    memcpy(q2, Q, n * sizeof(double));
    compresid(c);
    for (size_t i = 0; i < n; i++) Q[i] = a31 * q0[i] + a32 * Q[i] + a33 * dt * R[i];
    memcpy(q3, Q, n * sizeof(double));
    compresid(c);
    for (size_t i = 0; i < n; i++) Q[i] = a41 * q0[i] + a42 * Q[i] + a43 * dt * R[i];
    memcpy(r0, R, n * sizeof(double));
    compresid(c);
    for (size_t i = 0; i < n; i++)
        Q[i] = a51 * q0[i] + a52 * q2[i] + a53 * q3[i] + a54 * Q[i] + a55 * dt * r0[i] + a56 * dt * R[i];
    free(q0); free(q2); free(q3); free(r0);
}

void orc_steps(orc_ctx *c, int nsteps, double *residglob, double *divbglob, double *divbmax)
{
    int N = c->N;
    double rn = 0.0, db = 0.0, dmax = 0.0;
    for (int it = 0; it < nsteps; it++) {
        if (c->p.k_stage == 5) five_stage_rk(c);
        else rk_s5p3(c);
        orc_correct_b(c);
        /* diagnostic flux + gradient of iterations.f90:35-36 */
        comp_inviscid_flux(c);
        nablaQsp(c, 0, 8);
    }
    for (int ic = 0; ic < c->ncell; ic++)
        for (int js = 0; js < N; js++)
            for (int is = 0; is < N; is++) {
                for (int k = 0; k < NV; k++) rn = rn + sq(c->resid[IDX_Q(c, k, is, js, ic)]);
                double d = c->nablaQs[IDX_NQS(c, 5, 0, is, js, ic)] + c->nablaQs[IDX_NQS(c, 6, 1, is, js, ic)];
                db = db + sq(d);
                if (fabs(d) > dmax) dmax = fabs(d);
            }
    if (residglob) *residglob = sqrt(rn / (c->ncell * N * N * 5));
    if (divbglob) *divbglob = sqrt(db / (c->ncell * N * N));
    if (divbmax) *divbmax = dmax;
}

/* SETINITIALCOND, Orszag-Tang (setinitialcond.f90:1-80) */
void orc_setinitialcond_ot(orc_ctx *c)
{
    int N = c->N;
    double gam = c->gam;
    for (int pass = 0; pass < 2; pass++) {
        for (int ic = 0; ic < c->ncell; ic++)
            for (int js = 0; js < N; js++)
                for (int is = 0; is < N; is++) {
                    double x = c->XXsolu[IDX_JAC(c, is, js, ic) * 2 + 0];
                    double y = c->XXsolu[IDX_JAC(c, is, js, ic) * 2 + 1];
                    /* the reference build (flang -O3) fuses sin(y), cos(y) into one sincos call,
                     * whose results differ from separate sin/cos calls in the last bit for some y */
                    double sy, cy;
                    sincos(y, &sy, &cy);
                    double rho = sq(gam), u = -sy, v = sin(x), w = 0.0, pr = gam;
                    double *q = &c->Q[IDX_Q(c, 0, is, js, ic)];
                    double Bx, By, Bz;
                    /* first pass: sin(2x), cos(2x) fused into sincos too; second pass: cos(2x) */
                    double s2x, c2x;
                    if (pass == 0) { sincos(x * 2, &s2x, &c2x); Bx = -sy; By = s2x; Bz = 0.0; }
                    else { c2x = cos(x * 2); Bx = q[5]; By = q[6]; Bz = q[7]; }
                    double Az = 0.5 * c2x + cy;
                    double NORM_U = sq(u) + sq(v) + sq(w);
                    double NORM_B = sq(Bx) + sq(By) + sq(Bz);
                    q[0] = rho; q[1] = rho * u; q[2] = rho * v; q[3] = rho * w;
                    q[4] = pr / (gam - 1) + 0.5 * rho * NORM_U + 0.5 * NORM_B;
                    q[5] = Bx; q[6] = By; q[7] = Bz; q[8] = Az;
                }
        if (pass == 0) orc_correct_b(c);
    }
}

const double *orc_array(const orc_ctx *c, const char *name, long *count)
{
    size_t nc = (size_t)c->ncell, N = (size_t)c->N;
    struct { const char *n; const double *p; size_t k; } t[] = {
        {"Q", c->Q, nc * NV * N * N}, {"resid", c->resid, nc * NV * N * N},
        {"F1", c->F1, nc * 8 * (N + 1) * N}, {"G2", c->G2, nc * 8 * (N + 1) * N},
        {"Qvfi", c->Qvfi, nc * NV * (N + 1) * N}, {"Qvfj", c->Qvfj, nc * NV * (N + 1) * N},
        {"Azfi", c->Azfi, nc * (N + 1) * N}, {"Azfj", c->Azfj, nc * (N + 1) * N},
        {"nablaAs", c->nablaAs, nc * 2 * N * N}, {"nablaQs", c->nablaQs, nc * NV * 2 * N * N},
        {"nablaQvfi", c->nablaQvfi, nc * NV * 2 * (N + 1) * N},
        {"nablaQvfj", c->nablaQvfj, nc * NV * 2 * (N + 1) * N},
        {"Fv1", c->Fv1, nc * NV * (N + 1) * N}, {"Gv2", c->Gv2, nc * NV * (N + 1) * N},
        {"Jac", c->Jac, nc * N * N}, {"S1", c->S1, nc * 4 * (N + 1) * N},
        {"S2", c->S2, nc * 4 * (N + 1) * N}, {"XXsolu", c->XXsolu, nc * 2 * N * N},
        {"Xs", c->Xs, N}, {"Xf", c->Xf, N + 1}, {"Lmat", c->Lmat, (N + 1) * N},
        {"Mmat", c->Mmat, (N + 1) * N}, {"Xs_low", c->Xs_low, N - 1},
        {"muAV", c->muAV, nc * NV}, {"muAVfi", c->muAVfi, nc * NV * 2},
        {"muAVfj", c->muAVfj, nc * NV * 2}, {"SM_IND", c->SM_IND, nc * NV},
    };
    for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); i++)
        if (strcmp(t[i].n, name) == 0) { if (count) *count = (long)t[i].k; return t[i].p; }
    if (count) *count = 0;
    return NULL;
}

double orc_scalar(const orc_ctx *c, const char *name)
{
    if (strcmp(name, "eigvmax") == 0) return c->eigvmax;
    if (strcmp(name, "eigvmax_g") == 0) return c->eigvmax_g;
    if (strcmp(name, "epsilon0") == 0) return c->epsilon0;
    return 0.0;
}
