// Sanitizer driver (CPU only): the library's host code (chorus_api.cpp,
// host_ops.cpp, tecplot.cpp) and the C oracle built with AddressSanitizer and
// UndefinedBehaviorSanitizer, run on one case read from a binary file.
//
//   san_main case.bin out.bin
//
// case.bin (little-endian): int32 ncell, nvert, N, artiv, k_stage, steps, nbnd, nre;
// float64 mu, eta0, gam, dt, dxcycl, dycycl; float64 xv[nvert], yv[nvert];
// int32 ivcell[4*ncell] (column-major, 1-based); int32 bnd[2*nbnd] (column-major);
// int32 btype[nbnd]; float64 q0[9*N*N*ncell].
// out.bin: float64 resid[9*N*N*ncell], qend[9*N*N*ncell] of the oracle;
// int32 nface, nnode, ncellp, init_rc of the library.
//
// The oracle runs compresid and `steps` time steps; the library runs its host
// topology (checked against the oracle's), its plot-node topology, and
// chorus_init, whose host set-up (neighbour table, face list, vertex
// coordinates) runs before the device is looked for: on a machine without a
// GPU it returns CHORUS_ERR_HIP after that work.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "chorus_mhd.h"
#include "chorus_oracle.h"

template <class T> static bool rd(FILE *f, T *p, size_t n) { return fread(p, sizeof(T), n, f) == n; }
template <class T> static bool wr(FILE *f, const T *p, size_t n) { return fwrite(p, sizeof(T), n, f) == n; }

static int die(const char *m) {
    fprintf(stderr, "san_main: %s\n", m);
    return 2;
}

int main(int argc, char **argv) {
    if (argc != 3) return die("usage: san_main case.bin out.bin");
    FILE *f = fopen(argv[1], "rb");
    if (!f) return die("cannot open case");
    int h[8];
    double p[6];
    if (!rd(f, h, 8) || !rd(f, p, 6)) return die("short header");
    const int ncell = h[0], nvert = h[1], N = h[2], artiv = h[3], kst = h[4], steps = h[5], nbnd = h[6], nre = h[7];
    std::vector<double> xv(nvert), yv(nvert), q0((size_t)9 * N * N * ncell);
    std::vector<int> iv((size_t)4 * ncell), bnd((size_t)2 * nbnd), bt(nbnd);
    if (!rd(f, xv.data(), xv.size()) || !rd(f, yv.data(), yv.size()) || !rd(f, iv.data(), iv.size()) ||
        !rd(f, bnd.data(), bnd.size()) || !rd(f, bt.data(), bt.size()) || !rd(f, q0.data(), q0.size()))
        return die("short body");
    fclose(f);

    // oracle
    orc_topo t;
    if (orc_topology(ncell, nvert, iv.data(), xv.data(), yv.data(), nbnd, bnd.data(), bt.data(), p[4], p[5], &t))
        return die("orc_topology failed");
    orc_params op{N, 1, artiv, kst, p[0], p[1], p[2], p[3], p[4], p[5]};
    orc_ctx *o = orc_create(&op, ncell, nvert, iv.data(), xv.data(), yv.data(), &t);
    if (!o) return die("orc_create failed");
    std::vector<double> resid(q0.size()), qend(q0.size());
    orc_set_q(o, q0.data());
    orc_compresid(o, resid.data());
    orc_set_q(o, q0.data());
    double r1, d1, m1;
    orc_steps(o, steps, &r1, &d1, &m1);
    orc_get_q(o, qend.data());
    orc_destroy(o);

    // library host code: topology against the oracle's
    std::vector<int> if2c((size_t)16 * ncell), ibfinter((size_t)4 * ncell), cyc((size_t)8 * ncell),
        ibfree((size_t)4 * ncell);
    int nface = 0, ninter = 0, npair = 0, nfree = 0;
    int rc = chorus_topology(ncell, nvert, iv.data(), xv.data(), yv.data(), nbnd, bnd.data(), bt.data(), p[4], p[5],
                             &nface, if2c.data(), &ninter, ibfinter.data(), &npair, cyc.data(), &nfree,
                             ibfree.data());
    if (rc) return die("chorus_topology failed");
    if (nface != t.nface || ninter != t.ninter || npair != t.npair || nfree != t.nfree) return die("topology counts differ");
    if (memcmp(if2c.data(), t.if2c, sizeof(int) * 4 * nface) ||
        (ninter && memcmp(ibfinter.data(), t.ibfinter, sizeof(int) * ninter)) ||
        (npair && memcmp(cyc.data(), t.cyc2fpair, sizeof(int) * 2 * npair)) ||
        (nfree && memcmp(ibfree.data(), t.ibfree, sizeof(int) * nfree)))
        return die("topology arrays differ from the oracle's");
    chorus_mesh m{ncell, nvert, N, xv.data(), yv.data(), iv.data(), nface, if2c.data(), ninter, ibfinter.data(),
                  npair, cyc.data(), nfree, ibfree.data()};
    int nnode = 0, ncellp = 0;
    std::vector<double> pxy((size_t)2 * (nre + 1) * (nre + 1) * ncell);
    std::vector<int> c2n((size_t)4 * nre * nre * ncell);
    if (chorus_plot_topology(&m, nre, &nnode, &ncellp, pxy.data(), c2n.data())) return die("chorus_plot_topology failed");
    chorus_params cp{1, artiv, kst, p[0], p[1], p[2], p[3], p[4], p[5], ncell, CHORUS_FP64};
    chorus_ctx *c = nullptr;
    int init_rc = chorus_init(&m, &cp, &c);
    chorus_free(c);
    orc_topo_free(&t);

    FILE *g = fopen(argv[2], "wb");
    if (!g) return die("cannot open output");
    int tail[4] = {nface, nnode, ncellp, init_rc};
    if (!wr(g, resid.data(), resid.size()) || !wr(g, qend.data(), qend.size()) || !wr(g, tail, 4)) return die("write failed");
    fclose(g);
    return 0;
}
