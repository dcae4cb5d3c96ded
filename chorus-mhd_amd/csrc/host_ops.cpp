// Host-side set-up arithmetic of the reference, built with -ffp-contract=off:
// solution/flux points and the extrapolation/derivative operators.
#include <cmath>

#include "chorus_internal.hpp"

namespace chorus {

namespace {
const double kPi = std::atan(1.0) * 4.0;   // setup2d.f90:67

// Gauss-Legendre roots of the flux points for np = 7..10, as listed to 12
// digits in setsandpoints (initsetup.f90:211-252), outermost first
const double kRoots[4][4] = {
    {0.932469514203, 0.661209386466, 0.238619186083, 0.0},
    {0.949107912343, 0.741531185599, 0.405845151377, 0.0},
    {0.960289856498, 0.796666477414, 0.525532409916, 0.183434642496},
    {0.968160239508, 0.836031107327, 0.613371432701, 0.324253423404},
};

// setsandpoints (initsetup.f90:167-255)
void points(int np, double *xs, double *xf) {
    for (int i = 1; i <= np; i++)
        xs[i - 1] = 0.5 * (1.0 - std::cos((double)(2 * i - 1) * kPi / (2 * (double)np)));
    if (!xf) return;
    double *f = xf;
    switch (np) {
    case 1: f[0] = 0.0; f[1] = 1.0; return;
    case 2: f[0] = 0.0; f[1] = 0.5; f[2] = 1.0; return;
    case 3:
        f[0] = 0.0; f[1] = (1.0 - std::sqrt(3.0) / 3.0) / 2.0;
        f[2] = (1.0 + std::sqrt(3.0) / 3.0) / 2.0; f[3] = 1.0;
        return;
    case 4:
        f[0] = 0.0; f[1] = (1.0 - std::sqrt(15.0) / 5.0) / 2.0; f[2] = 0.5;
        f[3] = (1.0 + std::sqrt(15.0) / 5.0) / 2.0; f[4] = 1.0;
        return;
    case 5:
        f[0] = 0.0;
        f[1] = (1.0 - std::sqrt(525.0 + 70.0 * std::sqrt(30.0)) / 35.0) / 2.0;
        f[2] = (1.0 - std::sqrt(525.0 - 70.0 * std::sqrt(30.0)) / 35.0) / 2.0;
        f[3] = (1.0 + std::sqrt(525.0 - 70.0 * std::sqrt(30.0)) / 35.0) / 2.0;
        f[4] = (1.0 + std::sqrt(525.0 + 70.0 * std::sqrt(30.0)) / 35.0) / 2.0;
        f[5] = 1.0;
        return;
    case 6:
        f[0] = 0.0;
        f[1] = (1.0 - std::sqrt(245.0 + 14.0 * std::sqrt(70.0)) / 21.0) / 2.0;
        f[2] = (1.0 - std::sqrt(245.0 - 14.0 * std::sqrt(70.0)) / 21.0) / 2.0;
        f[3] = 0.5;
        f[4] = (1.0 + std::sqrt(245.0 - 14.0 * std::sqrt(70.0)) / 21.0) / 2.0;
        f[5] = (1.0 + std::sqrt(245.0 + 14.0 * std::sqrt(70.0)) / 21.0) / 2.0;
        f[6] = 1.0;
        return;
    default: {
        const double *r = kRoots[np - 7];
        int h = (np - 1) / 2;
        f[0] = 0.0;
        for (int k = 0; k < h; k++) f[1 + k] = (1.0 - r[k]) / 2.0;
        if ((np - 1) % 2 == 1) f[1 + h] = 0.5;
        for (int k = 0; k < h; k++) f[np - 1 - k] = (1.0 + r[k]) / 2.0;
        f[np] = 1.0;
        return;
    }
    }
}

// hhval (initsetup.f90:318-340)
double lagrange(int np, const double *xs, int i, double xval) {
    double hvaln = 1.0, hvald = 1.0;
    for (int s = 0; s < np; s++)
        if (s != i) { hvaln = hvaln * (xval - xs[s]); hvald = hvald * (xs[i] - xs[s]); }
    return hvaln / hvald;
}
}  // namespace

void host_solution_points(int N, double *Xs) { points(N, Xs, nullptr); }

void host_av_tables(int N, double *H1, double *H2) {
    double xs[kMaxN], xl[kMaxN];
    points(N, xs, nullptr);
    points(N - 1, xl, nullptr);
    for (int isl = 0; isl < N - 1; isl++)
        for (int is = 0; is < N; is++) H1[isl * N + is] = lagrange(N, xs, is, xl[isl]);
    for (int is = 0; is < N; is++)
        for (int isl = 0; isl < N - 1; isl++) H2[is * (N - 1) + isl] = lagrange(N - 1, xl, isl, xs[is]);
}

void host_plot_basis(int N, int nre, double *H) {
    double xs[kMaxN];
    points(N, xs, nullptr);
    for (int ii = 0; ii <= nre; ii++)
        for (int is = 0; is < N; is++) H[ii * N + is] = lagrange(N, xs, is, double(ii) / double(nre));
}

// computeLandM (initsetup.f90:258-315)
void host_operators(int N, double *Xs, double *Xf, double *L, double *M) {
    points(N, Xs, Xf);
    for (int ifp = 0; ifp <= N; ifp++)
        for (int is = 0; is < N; is++) L[ifp * N + is] = lagrange(N, Xs, is, Xf[ifp]);
    for (int ifp = 0; ifp <= N; ifp++)
        for (int is = 0; is < N; is++) {
            double xval = Xs[is], sum = 0.0;
            for (int k = 0; k <= N; k++) {
                if (k == ifp) continue;
                double num = 1.0, den = 1.0;
                for (int s = 0; s <= N; s++) {
                    if (s != ifp && s != k) num = num * (xval - Xf[s]);
                    if (s != ifp) den = den * (Xf[ifp] - Xf[s]);
                }
                sum = sum + num / den;
            }
            M[ifp * N + is] = sum;
        }
}

}  // namespace chorus
