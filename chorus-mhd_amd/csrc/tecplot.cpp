// Tecplot output of the solution: the plot-node topology of tecplotter2dsetup
// and the binary (TDV111) writer of tecpost_parallel, single partition.  The
// node values themselves are formed on the device (k_plot_nodes).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "chorus_internal.hpp"
#include "chorus_mhd.h"

namespace chorus {

namespace {
// xyCoor_atGps with MapBaseFunc (mapder.f90:79-87, 255-273): bilinear map of
// the 4 vertices, summed in vertex order
void xy_at(const double xx[2][4], double xi, double eta, double rx[2]) {
    const double f[4] = {(1.0 - xi) * (1.0 - eta), (1.0 - eta) * xi, xi * eta, (1.0 - xi) * eta};
    for (int i = 0; i < 2; i++) {
        double a = 0.0;
        for (int k = 0; k < 4; k++) a = a + xx[i][k] * f[k];
        rx[i] = a;
    }
}
}  // namespace

// tecplotter2dsetup (tecplotter2dsetup.f90:1-164), CURVE_WALL = 0: vertex nodes
// keep the vertex numbers; face nodes are created by the first cell that visits
// the face (local faces 1..4 in order) and found by distance (<= 1e-5) by the
// other; interior nodes per cell; NRE^2 plot cells per cell.  All inputs are
// the reference's 1-based arrays.
int build_plot_topology(int ncell, int nvert, int nface, const double *xv, const double *yv, const int *ivcell,
                        const int *if2c, int nre, PlotTopology &t, std::string &err) {
    const int Np = nre;
    if (Np < 1) { err = "NRE must be >= 1"; return CHORUS_ERR_ARG; }
    auto IVC = [&](int c, int k) { return ivcell[(size_t)k * ncell + c]; };   // 1-based vertex
    // IC2F from IF2C (left cell, right cell, left local face, right local face)
    std::vector<int> ic2f((size_t)ncell * 4, 0);
    for (int f = 0; f < nface; f++) {
        const int cl = if2c[f], cr = if2c[(size_t)nface + f], lfl = if2c[2 * (size_t)nface + f],
                  lfr = if2c[3 * (size_t)nface + f];
        if (cl < 1 || cl > ncell || lfl < 1 || lfl > 4) { err = "IF2C: bad left cell/face"; return CHORUS_ERR_TOPOLOGY; }
        ic2f[(size_t)(cl - 1) * 4 + lfl - 1] = f + 1;
        if (cr > 0) {
            if (cr > ncell || lfr < 1 || lfr > 4) { err = "IF2C: bad right cell/face"; return CHORUS_ERR_TOPOLOGY; }
            ic2f[(size_t)(cr - 1) * 4 + lfr - 1] = f + 1;
        }
    }
    for (size_t i = 0; i < ic2f.size(); i++)
        if (!ic2f[i]) { err = "a local face of a cell has no face"; return CHORUS_ERR_TOPOLOGY; }

    const size_t maxn = (size_t)(Np + 1) * (Np + 1) * ncell;
    t.nre = nre;
    t.px.assign(maxn + 1, 0.0);   // 1-based node ids
    t.py.assign(maxn + 1, 0.0);
    t.factor.assign(maxn + 1, 0);
    t.gnum.assign((size_t)(Np + 1) * (Np + 1) * ncell, 0);   // [c][jj][ii], 0-based ii, jj
    std::vector<int> isfacenod((size_t)nface + 1, 0), gfacenode((size_t)(Np > 1 ? Np - 1 : 1) * (nface + 1), 0);
    auto G = [&](int ii, int jj, int c) -> int & { return t.gnum[((size_t)c * (Np + 1) + (jj - 1)) * (Np + 1) + (ii - 1)]; };
    int plotnodes = 0;
    for (int c = 0; c < ncell; c++) {
        const int v[4] = {IVC(c, 0), IVC(c, 1), IVC(c, 2), IVC(c, 3)};
        G(1, 1, c) = v[0];
        G(Np + 1, 1, c) = v[1];
        G(Np + 1, Np + 1, c) = v[2];
        G(1, Np + 1, c) = v[3];
        for (int k = 0; k < 4; k++) {
            const int g = v[k];
            if (g < 1 || g > nvert) { err = "IVCELL: vertex out of range"; return CHORUS_ERR_TOPOLOGY; }
            if (t.factor[g] == 0) {
                plotnodes++;
                t.px[g] = xv[g - 1];
                t.py[g] = yv[g - 1];
            }
            t.factor[g]++;
        }
    }
    if (plotnodes != nvert) { err = "plot set-up: not every vertex belongs to a cell"; return CHORUS_ERR_TOPOLOGY; }
    const double tolplot = 1e-5;
    for (int c = 0; c < ncell; c++) {
        double xx[2][4];
        for (int j = 0; j < 4; j++) { xx[0][j] = xv[IVC(c, j) - 1]; xx[1][j] = yv[IVC(c, j) - 1]; }
        for (int k = 1; k <= 4; k++) {
            const int iface = ic2f[(size_t)c * 4 + k - 1];
            for (int enode = 1; enode <= Np - 1; enode++) {
                int ii, jj;
                switch (k) {
                case 1: ii = enode + 1; jj = 1; break;
                case 2: ii = Np + 1; jj = enode + 1; break;
                case 3: ii = Np + 1 - enode; jj = Np + 1; break;
                default: ii = 1; jj = Np + 1 - enode; break;
                }
                const double xsi = double(ii - 1) / double(Np), eta = double(jj - 1) / double(Np);
                double rx[2];
                xy_at(xx, xsi, eta, rx);
                if (isfacenod[iface] == 1) {
                    int match = 0;
                    for (int e = 1; e <= Np - 1; e++) {
                        const int g = gfacenode[(size_t)iface * (Np - 1) + e - 1];
                        const double d = std::fabs(std::sqrt((rx[0] - t.px[g]) * (rx[0] - t.px[g]) +
                                                             (rx[1] - t.py[g]) * (rx[1] - t.py[g])));
                        if (d <= tolplot) {
                            G(ii, jj, c) = g;
                            t.factor[g]++;
                            match = 1;
                        }
                    }
                    if (!match) { err = "plot set-up: face node without a partner"; return CHORUS_ERR_TOPOLOGY; }
                } else {
                    plotnodes++;
                    t.px[plotnodes] = rx[0];
                    t.py[plotnodes] = rx[1];
                    G(ii, jj, c) = plotnodes;
                    gfacenode[(size_t)iface * (Np - 1) + enode - 1] = plotnodes;
                    t.factor[plotnodes]++;
                }
            }
            if (isfacenod[iface] == 0) isfacenod[iface] = 1;
        }
        for (int ii = 2; ii <= Np; ii++)
            for (int jj = 2; jj <= Np; jj++) {
                const double xsi = double(ii - 1) / double(Np), eta = double(jj - 1) / double(Np);
                double rx[2];
                xy_at(xx, xsi, eta, rx);
                plotnodes++;
                t.px[plotnodes] = rx[0];
                t.py[plotnodes] = rx[1];
                G(ii, jj, c) = plotnodes;
                t.factor[plotnodes]++;
            }
    }
    t.nnode = plotnodes;
    t.ncellp = ncell * Np * Np;
    t.c2n.assign((size_t)4 * t.ncellp, 0);
    size_t pc = 0;
    for (int c = 0; c < ncell; c++)
        for (int ii = 1; ii <= Np; ii++)
            for (int jj = 1; jj <= Np; jj++) {
                t.c2n[4 * pc + 0] = G(ii, jj, c);
                t.c2n[4 * pc + 1] = G(ii + 1, jj, c);
                t.c2n[4 * pc + 2] = G(ii + 1, jj + 1, c);
                t.c2n[4 * pc + 3] = G(ii, jj + 1, c);
                pc++;
            }
    // contributions to each node in the order tecplotter2d adds them
    // (tecplotter.f90:23-40: cells ascending, then jj, then ii)
    t.cstart.assign((size_t)plotnodes + 1, 0);
    for (int c = 0; c < ncell; c++)
        for (int jj = 1; jj <= Np + 1; jj++)
            for (int ii = 1; ii <= Np + 1; ii++) t.cstart[G(ii, jj, c)]++;
    for (int n = 1; n <= plotnodes; n++) t.cstart[n] += t.cstart[n - 1];
    t.contrib.assign((size_t)2 * t.cstart[plotnodes], 0);
    std::vector<int> fill(t.cstart.begin(), t.cstart.end() - 1);
    for (int c = 0; c < ncell; c++)
        for (int jj = 1; jj <= Np + 1; jj++)
            for (int ii = 1; ii <= Np + 1; ii++) {
                const int k = fill[G(ii, jj, c) - 1]++;     // 0-based node
                t.contrib[2 * (size_t)k] = c;
                t.contrib[2 * (size_t)k + 1] = (jj - 1) * (Np + 1) + (ii - 1);
            }
    t.px.resize((size_t)plotnodes + 1);
    t.py.resize((size_t)plotnodes + 1);
    t.factor.resize((size_t)plotnodes + 1);
    return 0;
}

namespace {
void put_i32(std::vector<unsigned char> &b, int v) {
    unsigned char c[4];
    std::memcpy(c, &v, 4);
    b.insert(b.end(), c, c + 4);
}
void put_f32(std::vector<unsigned char> &b, float v) {
    unsigned char c[4];
    std::memcpy(c, &v, 4);
    b.insert(b.end(), c, c + 4);
}
void put_f64(std::vector<unsigned char> &b, double v) {
    unsigned char c[8];
    std::memcpy(c, &v, 8);
    b.insert(b.end(), c, c + 8);
}
// write_tecplot_string (tecplotter.f90:401-416): each character as an int, then 0
void put_str(std::vector<unsigned char> &b, const char *s) {
    for (const char *p = s; *p; p++) put_i32(b, (int)(unsigned char)*p);
    put_i32(b, 0);
}
}  // namespace

// tecpost_parallel (tecplotter.f90:92-399) for one partition: header, the 13
// variables as FP32 blocks, the 0-based connectivity.  q[v*nnode + n] holds the
// 9 plot variables (rho u v w p Bx By Bz Az) and av[k*nnode + n] muAV(1), muAV(6).
int write_tecplot_file(const char *path, double ctime, const PlotTopology &t, const double *q, const double *av,
                       std::string &err) {
    const int nvar = 13, nn = t.nnode;
    std::vector<const double *> col(nvar);
    std::vector<double> X(t.px.begin() + 1, t.px.end()), Y(t.py.begin() + 1, t.py.end());
    col[0] = X.data();
    col[1] = Y.data();
    for (int v = 0; v < 9; v++) col[2 + v] = q + (size_t)v * nn;
    col[11] = av;
    col[12] = av + nn;
    std::vector<unsigned char> b;
    b.reserve(1024 + (size_t)nn * nvar * 4 + (size_t)t.ncellp * 16);
    const char magic[] = "#!TDV111";
    b.insert(b.end(), magic, magic + 8);
    put_i32(b, 1);
    put_i32(b, 0);
    put_str(b, "step data");
    put_i32(b, nvar);
    const char *names[nvar] = {"X", "Y", "RHO", "U", "V", "W", "P", "Bx", "By", "Bz", "Az", "muAV", "etaAV"};
    for (int v = 0; v < nvar; v++) put_str(b, names[v]);
    put_f32(b, 299.0f);
    put_str(b, "ZONE 001");
    put_i32(b, -1);
    put_i32(b, 0);
    put_f64(b, ctime);
    const int zone[12] = {-1, 3, 0, 0, 0, 0, nn, t.ncellp, 0, 0, 0, 0};
    for (int z : zone) put_i32(b, z);
    put_f32(b, 357.0f);
    put_f32(b, 299.0f);
    for (int v = 0; v < nvar; v++) put_i32(b, 1);
    put_i32(b, 0);
    put_i32(b, 0);
    put_i32(b, -1);
    // minval / maxval of each variable (FP64), minmax(2,nvar) column-major
    for (int v = 0; v < nvar; v++) {
        double lo = col[v][0], hi = col[v][0];
        for (int n = 1; n < nn; n++) {
            lo = col[v][n] < lo ? col[v][n] : lo;
            hi = col[v][n] > hi ? col[v][n] : hi;
        }
        put_f64(b, lo);
        put_f64(b, hi);
    }
    for (int v = 0; v < nvar; v++)
        for (int n = 0; n < nn; n++) put_f32(b, (float)col[v][n]);
    for (int k : t.c2n) put_i32(b, k - 1);
    FILE *fp = std::fopen(path, "wb");
    if (!fp) { err = std::string("cannot open ") + path; return CHORUS_ERR_ARG; }
    const size_t w = std::fwrite(b.data(), 1, b.size(), fp);
    const int rc = std::fclose(fp);
    if (w != b.size() || rc != 0) { err = std::string("write failed: ") + path; return CHORUS_ERR_ARG; }
    return 0;
}

}  // namespace chorus
