// C-ABI layer (include/chorus_mhd.h): context, device memory, on-device face
// connectivity, the SSP steppers and the host-side topology helper.
#include "chorus_mhd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "chorus_internal.hpp"

using namespace chorus;

namespace {
// Tecplot output of one mesh (chorus_plot_setup / chorus_batch_plot_setup):
// the host topology and its device copies, the basis and the work arrays
struct PlotState {
    PlotTopology *t = nullptr;
    int *cstart = nullptr, *contrib = nullptr, *factor = nullptr;
    double *H = nullptr, *q = nullptr, *av = nullptr, *V = nullptr;
    void release() {
        for (void *p : {(void *)cstart, (void *)contrib, (void *)factor, (void *)H, (void *)q, (void *)av, (void *)V})
            (void)hipFree(p);
        cstart = contrib = factor = nullptr;
        H = q = av = V = nullptr;
        delete t;
        t = nullptr;
    }
};
}  // namespace

struct chorus_ctx {
    std::string err;
    chorus_params p{};
    int N = 0, ncell = 0, ncellglob = 0;
    bool ideal = true;             // mu = eta0 = 0 and no AV: viscous terms vanish exactly
    hipStream_t stream = nullptr;
    double *cellxy = nullptr;      // per cell: x of the ng geometry nodes, then their y
    int ng = 4;                    // geometry nodes per cell: 4 vertices, or 8 (chorus_set_nodes8)
    int2 *nbr = nullptr;
    double *X[4] = {nullptr, nullptr, nullptr, nullptr};
    double *R = nullptr;           // residual output / FiveStageRk resid0
    double *partial = nullptr;     // per-block reduction partials
    int nface = 0;                 // faces carrying a flux (interior, CYCL pairs, FREE)
    int4 *faces = nullptr;         // {cellL, lfiL, cellR (-1: FREE), lfiR} ordered by left cell
    double *TQ = nullptr;          // traces of the current state at the face points (mixed form:
                                   // written only for the div B diagnostic)
    MxTraces mt{};                 // mixed form: the traces the face passes read (compact form)
    void *mt_buf = nullptr;        // mixed form: the allocation holding mt
    void *FF = nullptr;            // per cell face point: common inviscid flux + A_z
    double *FA = nullptr;          // mixed form: per cell face point: FP64 upwind A_z
    void *FFv = nullptr;           // viscous path: per cell face point: common viscous flux
    void *gtr = nullptr;           // viscous path: own nablaQs face traces
    void *fvi = nullptr;           // viscous path: the interior viscous/AV fluxes of the stage
    double *sm = nullptr;          // viscous path: AV indicator per cell
    double *eig = nullptr;         // two slots (of ninst values) of the per-stage maximum wave speed
    int ninst = 1;                 // instances sharing this context (batched groups)
    int *inst = nullptr;           // batched groups: instance of each cell (device)
    double *hinst = nullptr;       // batched groups: AV length scale per instance (device)
    int tile = 2;                  // mixed form: tile shape of the cell kernels (the autotuner's usual pick)
    int stage_parity = 0;          // slot used by the next stage
    int last_parity = -1;          // slot written by the last stage
    double h_msh = 0.0;            // AV length scale (AV.f90:147)
    double qin[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // INLE faces: the prescribed state (chorus_bc)
    int nblk = 0;
    int device = -1;               // the HIP device current at chorus_init
    double *scratch = nullptr;     // that device's diagnostics scratch array (chorus_diag_noalloc)
    int cur = 0;                   // buffer holding the current state
    int last_in = -1;              // buffer holding the last stage's input state
    std::vector<hipEvent_t> ev;    // timing events (chorus_time_steps)
    bool timing = false;
    size_t ev_used = 0;
    // Tecplot output (chorus_plot_setup / chorus_write_tecplot)
    PlotState plot;
    double *mucell = nullptr;      // viscous path with AV: muAV(1) per cell of the last stage's input
    double *sm6 = nullptr;         // after chorus_plot_setup (AV): By indicator of the last stage's input
    // restart dump image (chorus_write_restart), allocated on first use
    unsigned *rs_img = nullptr;
    // one time step captured as a hipGraph, per buffer rotation, tile shape and
    // whether the stages write the Tecplot AV columns; post = the bookkeeping
    // (cur, last_in, stage_parity, last_parity) after the step
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        std::array<int, 4> post{};
    };
    std::map<std::array<int, 4>, StepGraph> step_graphs;
};

namespace {

int fail(chorus_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(c, expr)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(c, CHORUS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

size_t state_len(const chorus_ctx *c) { return (size_t)c->ncell * 9 * c->N * c->N; }

// reference Q(9,N,N,NCELL) <-> device Q[cell][var][js][is]
void to_device_layout(const chorus_ctx *c, const double *q, std::vector<double> &out) {
    const int N = c->N, NP = N * N;
    out.resize(state_len(c));
    for (int ic = 0; ic < c->ncell; ic++)
        for (int p = 0; p < NP; p++)
            for (int v = 0; v < 9; v++) out[((size_t)ic * 9 + v) * NP + p] = q[((size_t)ic * NP + p) * 9 + v];
}
void from_device_layout(const chorus_ctx *c, const std::vector<double> &in, double *q) {
    const int N = c->N, NP = N * N;
    for (int ic = 0; ic < c->ncell; ic++)
        for (int p = 0; p < NP; p++)
            for (int v = 0; v < 9; v++) q[((size_t)ic * NP + p) * 9 + v] = in[((size_t)ic * 9 + v) * NP + p];
}

bool g_ops_uploaded = false;

FaceArgs face_args(chorus_ctx *c) {
    FaceArgs f{};
    f.nface = c->nface;
    f.faces = c->faces;
    f.cellxy = c->cellxy;
    f.TQ = c->TQ;
    f.mt = c->mt;
    f.FF = c->FF;
    f.FA = c->FA;
    f.FFv = c->FFv;
    f.gtr = c->gtr;
    f.sm = c->sm;
    f.mucell = c->mucell;
    f.inst = c->inst;
    f.hinst = c->hinst;
    f.artiv = c->p.artiv;
    f.gam = c->p.gam;
    f.mu = c->p.mu;
    f.eta0 = c->p.eta0;
    f.h_msh = c->h_msh;
    f.pi = std::atan(1.0) * 4.0;
    for (int v = 0; v < 8; v++) f.qin[v] = c->qin[v];
    f.ng = c->ng;
    return f;
}

int launch_stage(chorus_ctx *c, const StageArgs &a) {
    hipError_t e;
    bool t = c->timing && c->ev_used + 2 <= c->ev.size();
    if (t) (void)hipEventRecord(c->ev[c->ev_used], c->stream);
    FaceArgs f = face_args(c);
    if (c->ideal) {
        if (c->p.precision == CHORUS_MIXED) e = launch_stage_ideal_mixed(c->N, a, f, c->stream, c->tile);
        else e = launch_stage_ideal_fp64(c->N, a, f, c->stream);
    } else {
        ViscArgs v{};
        v.ncell = c->ncell;
        v.cellxy = c->cellxy;
        v.nbr = c->nbr;
        v.Qin = a.Qin;
        v.gtr = c->gtr;
        v.fvi = c->fvi;
        v.inst = c->inst;
        v.hinst = c->hinst;
        v.sm = c->sm;
        v.eig = c->eig + (size_t)c->stage_parity * c->ninst;
        v.eig_next = c->eig + (size_t)(1 - c->stage_parity) * c->ninst;
        v.ninst = c->ninst;
        v.TQ = c->TQ;
        v.mt = c->mt;
        v.mucell = c->mucell;
        v.sm6 = c->sm6;
        v.artiv = c->p.artiv;
        v.mu = c->p.mu;
        v.eta0 = c->p.eta0;
        v.gam = c->p.gam;
        v.h_msh = c->h_msh;
        v.pi = std::atan(1.0) * 4.0;
        for (int q = 0; q < 8; q++) v.qin[q] = c->qin[q];
        f.eig = v.eig;
        c->last_parity = c->stage_parity;
        c->stage_parity = 1 - c->stage_parity;
        if (c->p.precision == CHORUS_MIXED) e = launch_stage_visc_mixed(c->N, a, v, f, c->stream, c->tile);
        else e = launch_stage_visc_fp64(c->N, a, v, f, c->stream);
    }
    if (t) { (void)hipEventRecord(c->ev[c->ev_used + 1], c->stream); c->ev_used += 2; }
    if (e != hipSuccess) return fail(c, CHORUS_ERR_HIP, std::string("stage launch: ") + hipGetErrorString(e));
    return 0;
}

StageArgs base_args(chorus_ctx *c) {
    StageArgs a{};
    a.ncell = c->ncell;
    a.cellxy = c->cellxy;
    a.nbr = c->nbr;
    a.gam = c->p.gam;
    a.FF = c->FF;
    a.FA = c->FA;
    a.FFv = c->FFv;
    a.TQ = c->TQ;
    a.mt = c->mt;
    a.sm = c->sm;
    a.ng = c->ng;
    return a;
}

bool mixed_traces(const chorus_ctx *c) { return c->mt.dev != nullptr; }

// the traces of Q that the face passes read (and, with AV, its indicator)
int run_trace(chorus_ctx *c, const double *Q) {
    double *sm = c->p.artiv ? c->sm : nullptr;
    hipError_t e = mixed_traces(c) ? launch_trace_mixed(c->N, c->ncell, Q, c->mt, sm, c->stream)
                                   : launch_trace(c->N, c->ncell, Q, c->TQ, sm, c->stream);
    if (e != hipSuccess) return fail(c, CHORUS_ERR_HIP, std::string("trace launch: ") + hipGetErrorString(e));
    return 0;
}

int run_ct(chorus_ctx *c, double *Q, int which, double *partial = nullptr) {
    CTArgs a{};
    a.ncell = c->ncell;
    a.cellxy = c->cellxy;
    a.nbr = c->nbr;
    a.Q = Q;
    a.which = which;
    a.partial = partial ? partial : c->partial;
    a.TQ = c->TQ;
    a.TQout = c->TQ;
    a.ng = c->ng;
    hipError_t e = hipSuccess;
    if (mixed_traces(c)) {
        // the correction reads the A_z traces and writes the Bx, By traces of
        // the compact form; the diagnostic reads FP64 Bx, By traces, made here
        if (which == 0) a.mt = c->mt;
        else e = launch_trace(c->N, c->ncell, Q, c->TQ, nullptr, c->stream);
    }
    if (e == hipSuccess) e = launch_ct(c->N, a, c->stream);
    if (e != hipSuccess) return fail(c, CHORUS_ERR_HIP, std::string("ct launch: ") + hipGetErrorString(e));
    return 0;
}

// one RK_S5P3 step (iterations.f90:162-217) + correct_B_via_curl_Az
int step_s5p3(chorus_ctx *c) {
    const double a30 = 0.355909775063327, a32 = 0.644090224936674;
    const double a40 = 0.367933791638137, a43 = 0.632066208361863;
    const double a52 = 0.237593836598569, a54 = 0.762406163401431;
    const double b10 = 0.377268915331368, b21 = 0.377268915331368, b32 = 0.242995220537396;
    const double b43 = 0.238458932846290, b54 = 0.287632146308408;
    const double dt = c->p.dt;
    const int ia = c->cur, ib = (c->cur + 1) % 3, ic = (c->cur + 2) % 3;
    double *A = c->X[ia], *B = c->X[ib], *C = c->X[ic];
    StageArgs s = base_args(c);
    int rc;
    s.Qin = A; s.Qout = B; s.kind = KIND_AXPY; s.c0 = b10 * dt;
    if ((rc = launch_stage(c, s))) return rc;
    s.Qin = B; s.Qout = C; s.kind = KIND_AXPY; s.c0 = b21 * dt;
    if ((rc = launch_stage(c, s))) return rc;
    s.Qin = C; s.Qout = B; s.Qa = A; s.kind = KIND_3A; s.c1 = a30; s.c2 = a32; s.c0 = b32 * dt;
    if ((rc = launch_stage(c, s))) return rc;
    s.Qin = B; s.Qout = A; s.Qa = A; s.kind = KIND_3A; s.c1 = a40; s.c2 = a43; s.c0 = b43 * dt;
    if ((rc = launch_stage(c, s))) return rc;
    s.Qin = A; s.Qout = C; s.Qa = C; s.kind = KIND_3A; s.c1 = a52; s.c2 = a54; s.c0 = b54 * dt;
    if ((rc = launch_stage(c, s))) return rc;
    c->last_in = ia;
    c->cur = ic;
    return 0;
}

// one FiveStageRk step (iterations.f90:92-160)
int step_s54(chorus_ctx *c) {
    const double a11 = 0.39175222700392;
    const double a21 = 0.44437049406734, a22 = 0.55562950593266, a23 = 0.36841059262959;
    const double a31 = 0.6201018513854, a32 = 0.3798981486146, a33 = 0.25189177424738;
    const double a41 = 0.17807995410773, a42 = 0.82192004589227, a43 = 0.54497475021237;
    const double a51 = 0.00683325884039, a52 = 0.51723167208978, a53 = 0.12759831133288;
    const double a54 = 0.34833675773694, a55 = 0.08460416338212, a56 = 0.22600748319395;
    const double dt = c->p.dt;
    const int ia = c->cur, ib = (c->cur + 1) % 4, ic = (c->cur + 2) % 4, id = (c->cur + 3) % 4;
    double *A = c->X[ia], *B = c->X[ib], *C = c->X[ic], *D = c->X[id];
    StageArgs s = base_args(c);
    int rc;
    s.Qin = A; s.Qout = B; s.kind = KIND_AXPY; s.c0 = dt * a11;
    if ((rc = launch_stage(c, s))) return rc;
    s.Qin = B; s.Qout = C; s.Qa = A; s.kind = KIND_3B; s.c1 = a21; s.c2 = a22; s.c3 = a23; s.c4 = dt;
    if ((rc = launch_stage(c, s))) return rc;
    s.Qin = C; s.Qout = D; s.Qa = A; s.kind = KIND_3A; s.c1 = a31; s.c2 = a32; s.c0 = a33 * dt;
    if ((rc = launch_stage(c, s))) return rc;
    s.Qin = D; s.Qout = B; s.Qa = A; s.Rout = c->R; s.kind = KIND_S54_4; s.c1 = a41; s.c2 = a42; s.c0 = a43 * dt;
    if ((rc = launch_stage(c, s))) return rc;
    s.Qin = B; s.Qout = A; s.Qa = A; s.Qb = C; s.Qc = D; s.Rprev = c->R; s.Rout = nullptr;
    s.kind = KIND_S54_5; s.c1 = a51; s.c2 = a52; s.c3 = a53; s.c4 = a54; s.c5 = a55 * dt; s.c0 = a56 * dt;
    if ((rc = launch_stage(c, s))) return rc;
    c->last_in = ib;
    c->cur = ia;
    return 0;
}

int one_step(chorus_ctx *c) {
    int rc = c->p.k_stage == 5 ? step_s54(c) : step_s5p3(c);
    if (rc) return rc;
    bool t = c->timing && c->ev_used + 2 <= c->ev.size();
    if (t) (void)hipEventRecord(c->ev[c->ev_used], c->stream);
    rc = run_ct(c, c->X[c->cur], 0);
    if (t) { (void)hipEventRecord(c->ev[c->ev_used + 1], c->stream); c->ev_used += 2; }
    return rc;
}

// one_step as one hipGraph launch: the step's 11-26 kernel launches cost one
// launch, which matters for small meshes; captured on first use per key
int graph_step(chorus_ctx *c) {
    const std::array<int, 4> key{c->cur, c->stage_parity, c->tile, c->sm6 ? 1 : 0};
    auto it = c->step_graphs.find(key);
    if (it == c->step_graphs.end()) {
        chorus_ctx::StepGraph g;
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = one_step(c);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(c, CHORUS_ERR_HIP, std::string("step capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(c, CHORUS_ERR_HIP, std::string("step graph: ") + hipGetErrorString(e));
        g.post = {c->cur, c->last_in, c->stage_parity, c->last_parity};
        // the capture advanced the host-side bookkeeping; rewind it to replay below
        c->cur = key[0];
        c->stage_parity = key[1];
        it = c->step_graphs.emplace(key, g).first;
    }
    HIP_TRY(c, hipGraphLaunch(it->second.exec, c->stream));
    const auto &p = it->second.post;
    c->cur = p[0]; c->last_in = p[1]; c->stage_parity = p[2]; c->last_parity = p[3];
    return 0;
}

// ---------------------------------------------------------------------------
// host topology: CONNECTIVITY (connectivity.f90:1-233), CONNECT_BDRY
// (connect_bdry.f90:1-354), MAPCYCLOC (map_cyc_loc.f90:1-42), one rank
// ---------------------------------------------------------------------------
struct Topo {
    std::vector<int> if2c, if2v, ibfinter, cyc, ibfree;   // column-major as the reference
    int nface = 0, npair = 0;
};

int build_topology(int ncell, int nvert, const int *ivcell, const double *xv, const double *yv, int nbnd,
                   const int *bnd, const int *btype, double dxcycl, double dycycl, Topo &t, std::string &err) {
    static const int IVF[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    const double tol = 1e-5;   // tolCYC, setup2d.f90:64
    auto IVC = [&](int ic, int k) { return ivcell[(size_t)k * ncell + ic]; };
    std::vector<int> fill(nvert, 0), sta(nvert + 1, 0);
    for (int ic = 0; ic < ncell; ic++)
        for (int k = 0; k < 4; k++) {
            int v = IVC(ic, k);
            if (v < 1 || v > nvert) { err = "IVCELL vertex id out of range"; return CHORUS_ERR_TOPOLOGY; }
            fill[v - 1]++;
        }
    for (int v = 0; v < nvert; v++)
        if (!fill[v]) { err = "vertex " + std::to_string(v + 1) + " belongs to no cell"; return CHORUS_ERR_TOPOLOGY; }
    for (int v = 0; v < nvert; v++) sta[v + 1] = sta[v] + fill[v];
    std::vector<int> pos(sta.begin(), sta.end() - 1), icvert(sta[nvert]);
    for (int ic = 0; ic < ncell; ic++)
        for (int k = 0; k < 4; k++) icvert[pos[IVC(ic, k) - 1]++] = ic;
    std::vector<int> ic2f((size_t)ncell * 4, 0), f2c, f2v;
    f2c.reserve((size_t)ncell * 4 * 2);
    f2v.reserve((size_t)ncell * 4 * 2);
    int nface = 0;
    for (int ic = 0; ic < ncell; ic++)
        for (int fa = 0; fa < 4; fa++) {
            if (ic2f[ic * 4 + fa]) continue;
            int vb = IVC(ic, IVF[fa][0]), v1 = IVC(ic, IVF[fa][1]);
            bool found = false;
            for (int k = sta[vb - 1]; k < sta[vb] && !found; k++) {
                int icc = icvert[k];
                if (icc == ic) continue;
                for (int ffa = 0; ffa < 4; ffa++) {
                    int w1 = IVC(icc, IVF[ffa][0]), w2 = IVC(icc, IVF[ffa][1]);
                    if ((vb == w1 && v1 == w2) || (vb == w2 && v1 == w1)) {
                        nface++;
                        ic2f[icc * 4 + ffa] = nface;
                        f2c.push_back(ic + 1); f2c.push_back(icc + 1);
                        f2v.push_back(vb); f2v.push_back(v1);
                        found = true;
                        break;
                    }
                }
            }
            if (!found) {
                nface++;
                f2c.push_back(ic + 1); f2c.push_back(0);
                f2v.push_back(vb); f2v.push_back(v1);
            }
            ic2f[ic * 4 + fa] = nface;
        }
    t.nface = nface;
    t.if2c.assign((size_t)nface * 4, 0);
    t.if2v.assign((size_t)nface * 2, 0);
    for (int f = 0; f < nface; f++) {
        t.if2c[f] = f2c[2 * f]; t.if2c[nface + f] = f2c[2 * f + 1];
        t.if2v[f] = f2v[2 * f]; t.if2v[nface + f] = f2v[2 * f + 1];
        int c1 = f2c[2 * f] - 1, c2 = f2c[2 * f + 1] - 1, v1 = f2v[2 * f], v2 = f2v[2 * f + 1];
        int a = IVC(c1, 0), b = IVC(c1, 1), cc = IVC(c1, 2), d = IVC(c1, 3), lf = 0;
        if (a == v1 && b == v2) lf = 1;
        else if (b == v1 && cc == v2) lf = 2;
        else if (cc == v1 && d == v2) lf = 3;
        else if (d == v1 && a == v2) lf = 4;
        if (!lf) { err = "connectivity: face " + std::to_string(f + 1) + " not on its left cell"; return CHORUS_ERR_TOPOLOGY; }
        t.if2c[2 * nface + f] = lf;
        if (c2 >= 0) {
            a = IVC(c2, 0); b = IVC(c2, 1); cc = IVC(c2, 2); d = IVC(c2, 3);
            int rf = 0;
            if (b == v1 && a == v2) rf = 1;
            else if (cc == v1 && b == v2) rf = 2;
            else if (d == v1 && cc == v2) rf = 3;
            else if (a == v1 && d == v2) rf = 4;
            t.if2c[3 * nface + f] = rf;
        }
    }
    std::vector<int> cycl;
    for (int ib = 0; ib < nbnd; ib++) {
        int b1 = bnd[ib], b3 = bnd[nbnd + ib], icc = -1;
        if (b1 < 1 || b1 > nvert || b3 < 1 || b3 > nvert) continue;
        for (int k1 = sta[b1 - 1]; k1 < sta[b1]; k1++)
            for (int k3 = sta[b3 - 1]; k3 < sta[b3]; k3++)
                if (icvert[k1] == icvert[k3]) icc = icvert[k1];
        if (icc < 0) continue;
        int fa;
        for (fa = 0; fa < 4; fa++) {
            int v1 = IVC(icc, IVF[fa][0]), v2 = IVC(icc, IVF[fa][1]);
            if ((v1 == b1 || v2 == b1) && (v1 == b3 || v2 == b3)) break;
        }
        if (fa == 4) continue;
        int k = ic2f[icc * 4 + fa];
        if (btype[ib] == 1) cycl.push_back(k);
        else if (btype[ib] == 2) t.ibfree.push_back(k);
    }
    for (int f = 0; f < nface; f++)
        if (t.if2c[nface + f]) t.ibfinter.push_back(f + 1);
    auto mid = [&](int K, double &x, double &y) {
        x = (xv[t.if2v[K] - 1] + xv[t.if2v[nface + K] - 1]) / 2;
        y = (yv[t.if2v[K] - 1] + yv[t.if2v[nface + K] - 1]) / 2;
    };
    std::vector<int> cycloc;
    for (size_t ib = 0; ib < cycl.size(); ib++) {
        double xi, yi;
        mid(cycl[ib] - 1, xi, yi);
        bool hit = false;
        for (size_t ib2 = 0; ib2 < cycl.size() && !hit; ib2++) {
            double xo, yo;
            mid(cycl[ib2] - 1, xo, yo);
            double ex = std::fabs(std::fabs(xi - xo) - dxcycl), ey = std::fabs(std::fabs(yi - yo) - dycycl);
            if ((std::fabs(xi - xo) < tol && ey < tol) || (std::fabs(yi - yo) < tol && ex < tol)) hit = true;
        }
        if (!hit) { err = "periodic face without a partner (remote periodic faces need more than one rank)"; return CHORUS_ERR_TOPOLOGY; }
        cycloc.push_back(cycl[ib]);
    }
    std::vector<int> p1, p2;
    for (size_t nf = 0; nf < cycloc.size(); nf++) {
        double xi, yi;
        mid(cycloc[nf] - 1, xi, yi);
        for (size_t nf2 = nf + 1; nf2 < cycloc.size(); nf2++) {
            double xo, yo;
            mid(cycloc[nf2] - 1, xo, yo);
            double ex = std::fabs(std::fabs(xi - xo) - dxcycl), ey = std::fabs(std::fabs(yi - yo) - dycycl);
            if ((ex < tol && std::fabs(yi - yo) < tol) || (ey < tol && std::fabs(xi - xo) < tol)) {
                p1.push_back(cycloc[nf]);
                p2.push_back(cycloc[nf2]);
            }
        }
    }
    if (p1.size() * 2 != cycloc.size()) { err = "error in matching CYCLOC faces (map_cyc_loc.f90:37-40)"; return CHORUS_ERR_TOPOLOGY; }
    t.npair = (int)p1.size();
    t.cyc = p1;
    t.cyc.insert(t.cyc.end(), p2.begin(), p2.end());
    return 0;
}

}  // namespace

int init_ctx(chorus_ctx *c, const chorus_mesh *m, const chorus_params *p, const chorus_bc *bc = nullptr);

// ===========================================================================
extern "C" {

const char *chorus_last_error(const chorus_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int chorus_topology(int ncell, int nvert, const int *ivcell, const double *xv, const double *yv, int nbnd,
                    const int *bnd, const int *btype, double dxcycl, double dycycl, int *nface, int *if2c,
                    int *ninter, int *ibfinter, int *npair, int *cyc2fpair, int *nfree, int *ibfree) {
    if (ncell <= 0 || nvert <= 0 || !ivcell || !xv || !yv || (nbnd > 0 && (!bnd || !btype)))
        return CHORUS_ERR_ARG;
    Topo t;
    std::string err;
    int rc = build_topology(ncell, nvert, ivcell, xv, yv, nbnd, bnd, btype, dxcycl, dycycl, t, err);
    if (rc) {
        std::fprintf(stderr, "chorus_topology: %s\n", err.c_str());
        return rc;
    }
    *nface = t.nface;
    *ninter = (int)t.ibfinter.size();
    *npair = t.npair;
    *nfree = (int)t.ibfree.size();
    std::copy(t.if2c.begin(), t.if2c.end(), if2c);
    std::copy(t.ibfinter.begin(), t.ibfinter.end(), ibfinter);
    std::copy(t.cyc.begin(), t.cyc.end(), cyc2fpair);
    std::copy(t.ibfree.begin(), t.ibfree.end(), ibfree);
    return 0;
}

int chorus_init(const chorus_mesh *m, const chorus_params *p, chorus_ctx **out) {
    return chorus_init_bc(m, p, nullptr, out);
}

int chorus_init_bc(const chorus_mesh *m, const chorus_params *p, const chorus_bc *bc, chorus_ctx **out) {
    if (!out) return CHORUS_ERR_ARG;
    *out = nullptr;
    chorus_ctx *c = new chorus_ctx();
    *out = c;
    return init_ctx(c, m, p, bc);
}

}  // extern "C"

// set-up shared by chorus_init and the batched groups (c->ninst set before)
int init_ctx(chorus_ctx *c, const chorus_mesh *m, const chorus_params *p, const chorus_bc *bc) {
    if (!m || !p) return fail(c, CHORUS_ERR_ARG, "null mesh or params");
    if (m->N < 2 || m->N > kMaxN) return fail(c, CHORUS_ERR_ARG, "N must be in 2..8");
    if (m->ncell <= 0) return fail(c, CHORUS_ERR_ARG, "empty mesh");
    // the kernels number face-point records (cell, local face, point) in 32 bits
    if ((long long)m->ncell * 4 * m->N >= (1LL << 31))
        return fail(c, CHORUS_ERR_ARG, "mesh too large: NCELL*4*N face-point records must stay below 2^31");
    if (p->vismode != 1)
        return fail(c, CHORUS_ERR_ARG, "vismode must be 1: with vismode=0 the reference reads uninitialised "
                                      "flux-point A_z (compflux.f90:72-75, magnetic_potential.f90:17)");
    if (p->k_stage != 3 && p->k_stage != 5)
        return fail(c, CHORUS_ERR_ARG, "k_stage can be only 3 or 5 (iterations.f90:22-31)");
    if (p->precision != CHORUS_FP64 && p->precision != CHORUS_MIXED)
        return fail(c, CHORUS_ERR_ARG, "precision must be CHORUS_FP64 or CHORUS_MIXED");
    if (p->artiv != 0 && p->artiv != 1) return fail(c, CHORUS_ERR_ARG, "ARTIV must be 0 or 1");
    if (p->artiv == 1 && m->N < 2) return fail(c, CHORUS_ERR_ARG, "AV needs N >= 2");
    c->p = *p;
    // with mu = eta0 = 0 and no AV every viscous term is an exact zero
    // (getrusanov.f90:218-247), so the stage needs no gradient pass
    c->ideal = (p->artiv == 0 && p->mu == 0.0 && p->eta0 == 0.0);
    c->N = m->N;
    c->ncell = m->ncell;
    c->ncellglob = p->ncellglob > 0 ? p->ncellglob : m->ncell;
    c->h_msh = std::sqrt(p->dxcycl * p->dycycl / (double)c->ncellglob);

    // on-device face connectivity: per cell and local face, the partner cell,
    // its local face and whether this cell is the reference's left or right cell
    // ({partner, lf | right << 3}); a face without a right cell holds
    // {-1, 1 << 4 | kind << 5} with its BoundaryKind
    std::vector<int2> nbr((size_t)m->ncell * 4, make_int2(-2, 0));
    auto IF = [&](int f, int k) { return m->if2c[(size_t)k * m->nface + f]; };
    auto link = [&](int l, int lfl, int r, int lfr) -> bool {
        if (l < 0 || l >= m->ncell || r < 0 || r >= m->ncell || lfl < 1 || lfl > 4 || lfr < 1 || lfr > 4) return false;
        nbr[(size_t)l * 4 + lfl - 1] = make_int2(r, lfr);
        nbr[(size_t)r * 4 + lfr - 1] = make_int2(l, lfl | (1 << 3));
        return true;
    };
    for (int k = 0; k < m->ninter; k++) {
        int f = m->ibfinter[k] - 1;
        if (f < 0 || f >= m->nface || !link(IF(f, 0) - 1, IF(f, 2), IF(f, 1) - 1, IF(f, 3)))
            return fail(c, CHORUS_ERR_TOPOLOGY, "bad interior face " + std::to_string(k + 1));
    }
    for (int k = 0; k < m->npair; k++) {
        int f1 = m->cyc2fpair[k] - 1, f2 = m->cyc2fpair[m->npair + k] - 1;
        if (f1 < 0 || f1 >= m->nface || f2 < 0 || f2 >= m->nface || !link(IF(f1, 0) - 1, IF(f1, 2), IF(f2, 0) - 1, IF(f2, 2)))
            return fail(c, CHORUS_ERR_TOPOLOGY, "bad periodic pair " + std::to_string(k + 1));
    }
    // faces without a right cell: FREE, and the kinds of chorus_bc; every pass
    // but the inviscid face flux treats them alike (own trace)
    std::vector<int> bkind((size_t)m->ncell * 4, BK_ZEROGRAD);
    auto boundary = [&](int nlist, const int *list, int kind, const char *what) -> int {
        if (nlist > 0 && !list) return fail(c, CHORUS_ERR_ARG, std::string("null ") + what + " list");
        for (int k = 0; k < nlist; k++) {
            int f = list[k] - 1;
            if (f < 0 || f >= m->nface) return fail(c, CHORUS_ERR_TOPOLOGY, std::string("bad ") + what + " face");
            int l = IF(f, 0) - 1, lf = IF(f, 2);
            if (l < 0 || l >= m->ncell || lf < 1 || lf > 4)
                return fail(c, CHORUS_ERR_TOPOLOGY, std::string("bad ") + what + " face");
            nbr[(size_t)l * 4 + lf - 1] = make_int2(-1, (1 << 4) | (kind << 5));
            bkind[(size_t)l * 4 + lf - 1] = kind;
        }
        return 0;
    };
    int rcb = boundary(m->nfree, m->ibfree, BK_ZEROGRAD, "FREE");
    if (!rcb && bc) {
        rcb = boundary(bc->noutlet, bc->ibfout, BK_ZEROGRAD, "OUTL");
        if (!rcb) rcb = boundary(bc->nsymp, bc->ibfsymp, BK_MIRROR, "SYMP");
        if (!rcb) rcb = boundary(bc->nwall, bc->ibfwal, BK_MIRROR, "WALL");
        if (!rcb) rcb = boundary(bc->ninlet, bc->ibfinl, BK_INLET, "INLE");
        if (!rcb && bc->ninlet > 0 && !(bc->qin[0] > 0.0))
            rcb = fail(c, CHORUS_ERR_ARG, "INLE faces need an inlet state with positive density (chorus_bc.qin)");
        for (int v = 0; v < 8; v++) c->qin[v] = bc->qin[v];
    }
    if (rcb) return rcb;
    for (size_t i = 0; i < nbr.size(); i++)
        if (nbr[i].x == -2)
            return fail(c, CHORUS_ERR_TOPOLOGY,
                        "cell " + std::to_string(i / 4 + 1) + " face " + std::to_string(i % 4 + 1) +
                            " has no flux: it is in none of IBFINTER, CYC2FPAIR, IBFREE and the chorus_bc lists");
    // face list for the face-ordered passes: each face once, from its left cell;
    // a face without a right cell carries its boundary kind in .w
    std::vector<int4> faces;
    faces.reserve((size_t)m->ncell * 2 + 16);
    for (int ic = 0; ic < m->ncell; ic++)
        for (int lfi = 0; lfi < 4; lfi++) {
            const int2 in = nbr[(size_t)ic * 4 + lfi];
            if ((in.y >> 4) & 1) faces.push_back(make_int4(ic, lfi, -1, bkind[(size_t)ic * 4 + lfi]));
            else if (!((in.y >> 3) & 1)) faces.push_back(make_int4(ic, lfi, in.x, (in.y & 7) - 1));
        }
    std::vector<double> xy((size_t)m->ncell * 8);
    for (int ic = 0; ic < m->ncell; ic++)
        for (int k = 0; k < 4; k++) {
            int v = m->ivcell[(size_t)k * m->ncell + ic];
            if (v < 1 || v > m->nvert) return fail(c, CHORUS_ERR_ARG, "IVCELL vertex out of range");
            xy[(size_t)ic * 8 + k] = m->xv[v - 1];
            xy[(size_t)ic * 8 + 4 + k] = m->yv[v - 1];
        }

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(c, CHORUS_ERR_HIP, "no HIP device: this library has no CPU path");
    if (!g_ops_uploaded) {
        HIP_TRY(c, upload_operators_fp64());
        HIP_TRY(c, upload_operators_mixed());
        g_ops_uploaded = true;
    }
    // the scratch array of chorus_diag_noalloc on this device, resolved now
    // with the device current (once per device)
    HIP_TRY(c, hipGetDevice(&c->device));
    HIP_TRY(c, diag_scratch(c->device, &c->scratch));
    HIP_TRY(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const size_t ns = state_len(c);
    HIP_TRY(c, hipMalloc(&c->cellxy, xy.size() * sizeof(double)));
    HIP_TRY(c, hipMalloc(&c->nbr, nbr.size() * sizeof(int2)));
    for (int i = 0; i < 4; i++) HIP_TRY(c, hipMalloc(&c->X[i], ns * sizeof(double)));
    HIP_TRY(c, hipMalloc(&c->R, ns * sizeof(double)));
    c->nblk = stage_blocks(c->N, c->ncell);
    HIP_TRY(c, hipMalloc(&c->partial, (size_t)c->nblk * 2 * sizeof(double)));
    HIP_TRY(c, hipMalloc(&c->eig, 2 * (size_t)c->ninst * sizeof(double)));
    HIP_TRY(c, hipMemsetAsync(c->eig, 0, 2 * (size_t)c->ninst * sizeof(double), c->stream));
    c->nface = (int)faces.size();
    const size_t nfp = (size_t)c->ncell * 4 * c->N;
    HIP_TRY(c, hipMalloc(&c->faces, faces.size() * sizeof(int4)));
    HIP_TRY(c, hipMalloc(&c->TQ, nfp * 9 * sizeof(double)));
    if (c->p.precision == CHORUS_MIXED) {
        // dev [nfp][8] float, az [nfp] double, ref [ncell][8] double
        HIP_TRY(c, hipMalloc(&c->mt_buf, nfp * 8 * sizeof(float) + nfp * sizeof(double) +
                                             (size_t)c->ncell * 8 * sizeof(double)));
        c->mt.dev = (float *)c->mt_buf;
        c->mt.az = (double *)(c->mt.dev + nfp * 8);
        c->mt.ref = c->mt.az + nfp;
    }
    HIP_TRY(c, hipMalloc(&c->FF, nfp * 9 * sizeof(double)));
    if (c->p.precision == CHORUS_MIXED) HIP_TRY(c, hipMalloc(&c->FA, nfp * sizeof(double)));
    // (the mixed form folds the viscous face flux into its FF records)
    if (!c->ideal && c->p.precision != CHORUS_MIXED) HIP_TRY(c, hipMalloc(&c->FFv, nfp * 9 * sizeof(double)));
    HIP_TRY(c, hipMemcpyAsync(c->faces, faces.data(), faces.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    if (!c->ideal) {
        const size_t nf = nfp;
        HIP_TRY(c, hipMalloc(&c->gtr, nf * 16 * sizeof(double)));
        HIP_TRY(c, hipMalloc(&c->fvi, (size_t)c->ncell * 2 * 9 * c->N * (c->N - 1) * sizeof(double)));
        HIP_TRY(c, hipMalloc(&c->sm, (size_t)c->ncell * sizeof(double)));
        if (c->p.artiv) HIP_TRY(c, hipMalloc(&c->mucell, (size_t)c->ncell * sizeof(double)));
    }
    HIP_TRY(c, hipMemcpyAsync(c->cellxy, xy.data(), xy.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->nbr, nbr.data(), nbr.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" {

void chorus_free(chorus_ctx *c) {
    if (!c) return;
    for (auto e : c->ev) (void)hipEventDestroy(e);
    (void)hipFree(c->cellxy);
    (void)hipFree(c->nbr);
    for (int i = 0; i < 4; i++) (void)hipFree(c->X[i]);
    (void)hipFree(c->R);
    (void)hipFree(c->partial);
    (void)hipFree(c->eig);
    (void)hipFree(c->faces);
    (void)hipFree(c->TQ);
    (void)hipFree(c->mt_buf);
    (void)hipFree(c->FF);
    (void)hipFree(c->FA);
    (void)hipFree(c->FFv);
    (void)hipFree(c->gtr);
    (void)hipFree(c->fvi);
    (void)hipFree(c->sm);
    (void)hipFree(c->inst);
    (void)hipFree(c->hinst);
    c->plot.release();
    (void)hipFree(c->mucell);
    (void)hipFree(c->sm6);
    (void)hipFree(c->rs_img);
    for (auto &g : c->step_graphs) (void)hipGraphExecDestroy(g.second.exec);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int chorus_set_state(chorus_ctx *c, const double *q) {
    if (!c || !q) return fail(c, CHORUS_ERR_ARG, "null argument");
    std::vector<double> h;
    to_device_layout(c, q, h);
    HIP_TRY(c, hipMemcpyAsync(c->X[c->cur], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = run_trace(c, c->X[c->cur]);
    if (rc) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // no stage has evaluated the new state yet: the residual norm, the wave
    // speed and the AV output columns are those of no stage (NaN / zeros)
    c->last_in = -1;
    c->last_parity = -1;
    return 0;
}

int chorus_get_state(chorus_ctx *c, double *q) {
    if (!c || !q) return fail(c, CHORUS_ERR_ARG, "null argument");
    std::vector<double> h(state_len(c));
    HIP_TRY(c, hipMemcpyAsync(h.data(), c->X[c->cur], h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    from_device_layout(c, h, q);
    return 0;
}

int chorus_compresid(chorus_ctx *c, double *resid) {
    if (!c) return CHORUS_ERR_ARG;
    StageArgs s = base_args(c);
    s.Qin = c->X[c->cur];
    s.Rout = c->R;
    s.kind = KIND_RESID;
    int rc = launch_stage(c, s);
    if (rc) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (resid) {
        std::vector<double> h(state_len(c));
        HIP_TRY(c, hipMemcpy(h.data(), c->R, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        from_device_layout(c, h, resid);
    }
    return 0;
}

int chorus_set_nodes8(chorus_ctx *c, const double *xm, const double *ym) {
    if (!c || !xm || !ym) return fail(c, CHORUS_ERR_ARG, "null argument");
    if (c->plot.t) return fail(c, CHORUS_ERR_ARG, "Tecplot output of 8-node cells is not supported");
    const size_t nc = (size_t)c->ncell;
    for (size_t i = 0; i < 4 * nc; i++)
        if (!std::isfinite(xm[i]) || !std::isfinite(ym[i])) return fail(c, CHORUS_ERR_ARG, "edge node not finite");
    // the corners stay those of the mesh (the first 4 x and the first 4 y of
    // each cell's current geometry)
    std::vector<double> old(nc * 2 * c->ng), xy(nc * 16);
    HIP_TRY(c, hipMemcpyAsync(old.data(), c->cellxy, old.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (size_t ic = 0; ic < nc; ic++)
        for (int k = 0; k < 4; k++) {
            xy[ic * 16 + k] = old[ic * 2 * c->ng + k];
            xy[ic * 16 + 8 + k] = old[ic * 2 * c->ng + c->ng + k];
            xy[ic * 16 + 4 + k] = xm[(size_t)k * nc + ic];
            xy[ic * 16 + 12 + k] = ym[(size_t)k * nc + ic];
        }
    double *d = nullptr;
    HIP_TRY(c, hipMalloc(&d, xy.size() * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d, xy.data(), xy.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(c, CHORUS_ERR_HIP, std::string("chorus_set_nodes8: ") + hipGetErrorString(e));
    }
    (void)hipFree(c->cellxy);
    c->cellxy = d;
    c->ng = 8;
    // the captured steps hold the old geometry and the 4-node kernels
    for (auto &g : c->step_graphs) (void)hipGraphExecDestroy(g.second.exec);
    c->step_graphs.clear();
    return 0;
}

int chorus_correct_b(chorus_ctx *c) {
    if (!c) return CHORUS_ERR_ARG;
    int rc = run_ct(c, c->X[c->cur], 0);
    if (rc) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

int chorus_step(chorus_ctx *c, int nsteps) {
    if (!c || nsteps < 0) return fail(c, CHORUS_ERR_ARG, "bad arguments");
    for (int i = 0; i < nsteps; i++) {
        int rc = graph_step(c);
        if (rc) return rc;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

int chorus_diag(chorus_ctx *c, double *residglob, double *divbglob, double *divbmax, double *eigvmax,
                double *epsilon0) {
    if (!c) return CHORUS_ERR_ARG;
    const int N = c->N;
    const double npts = (double)c->ncellglob * N * N;
    if (residglob) {
        // the reference's norm is over the residual of the last stage evaluated
        // (iterations.f90:42-45): re-evaluate it from that stage's input state
        *residglob = NAN;
        if (c->last_in >= 0) {
            StageArgs s = base_args(c);
            s.Qin = c->X[c->last_in];
            s.Rout = c->R;
            s.kind = KIND_RESID;
            // the face passes read the traces of that state; those of the
            // current state are restored afterwards
            int rc = run_trace(c, s.Qin);
            if (!rc) rc = launch_stage(c, s);
            if (!rc) rc = run_trace(c, c->X[c->cur]);
            if (rc) return rc;
            std::vector<double> h(state_len(c));
            HIP_TRY(c, hipMemcpyAsync(h.data(), c->R, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            // cell, point, variable order of the reference's loop
            const int NP = N * N;
            double s2 = 0.0;
            for (int ic = 0; ic < c->ncell; ic++)
                for (int p = 0; p < NP; p++)
                    for (int v = 0; v < 9; v++) {
                        double r = h[((size_t)ic * 9 + v) * NP + p];
                        s2 = s2 + r * r;
                    }
            *residglob = std::sqrt(s2 / (npts * 5));
        }
    }
    if (divbglob || divbmax) {
        int rc = run_ct(c, c->X[c->cur], 1);
        if (rc) return rc;
        std::vector<double> part((size_t)c->nblk * 2);
        HIP_TRY(c, hipMemcpyAsync(part.data(), c->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        double s2 = 0.0, mx = 0.0;
        for (int b = 0; b < c->nblk; b++) { s2 += part[2 * b]; mx = std::max(mx, part[2 * b + 1]); }
        if (divbglob) *divbglob = std::sqrt(s2 / npts);
        if (divbmax) *divbmax = mx;
    }
    if (eigvmax || epsilon0) {
        double e = NAN;
        if (!c->ideal && c->last_parity >= 0) {
            HIP_TRY(c, hipMemcpyAsync(&e, c->eig + c->last_parity, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
        }
        if (eigvmax) *eigvmax = e;
        if (epsilon0) *epsilon0 = (c->p.artiv == 1) ? e * c->h_msh / (double)(N - 1) * 1.0 : NAN;
    }
    return 0;
}

int chorus_diag_noalloc(chorus_ctx *c, double *residglob, double *divbglob, double *divbmax, double *eigvmax,
                        double *epsilon0) {
    if (!c) return CHORUS_ERR_ARG;
    if (!c->scratch) return fail(c, CHORUS_ERR_HIP, "chorus_diag_noalloc: no scratch array on this device");
    // scratch: the div B partials of every cell block, the residual partials, 3 results
    const size_t need = 2 * (size_t)c->nblk + (size_t)kDiagResBlocks + 3;
    if (need > kDiagScratchDoubles)
        return fail(c, CHORUS_ERR_ARG, "chorus_diag_noalloc: the mesh needs more scratch than the fixed array "
                                       "holds (2 x cell blocks + 1027 doubles > 1048576); use chorus_diag");
    const int N = c->N;
    const double npts = (double)c->ncellglob * N * N;
    const bool resid = residglob && c->last_in >= 0;
    double out[3] = {0.0, 0.0, 0.0};
    if (resid || divbglob || divbmax) {
        if (resid) {
            // the residual of the last stage's input (as chorus_diag); the face
            // passes read that state's traces, the current state's are restored
            StageArgs s = base_args(c);
            s.Qin = c->X[c->last_in];
            s.Rout = c->R;
            s.kind = KIND_RESID;
            int rc = run_trace(c, s.Qin);
            if (!rc) rc = launch_stage(c, s);
            if (!rc) rc = run_trace(c, c->X[c->cur]);
            if (rc) return rc;
        }
        int rc = run_ct(c, c->X[c->cur], 1, c->scratch);
        if (rc) return rc;
        HIP_TRY(c, launch_diag_reduce(resid ? c->R : nullptr, state_len(c), c->nblk, c->scratch, c->stream));
        HIP_TRY(c, hipMemcpyAsync(out, c->scratch + 2 * (size_t)c->nblk + kDiagResBlocks, sizeof(out),
                                  hipMemcpyDeviceToHost, c->stream));
    }
    double e = NAN;
    if ((eigvmax || epsilon0) && !c->ideal && c->last_parity >= 0)
        HIP_TRY(c, hipMemcpyAsync(&e, c->eig + c->last_parity, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (residglob) *residglob = resid ? std::sqrt(out[2] / (npts * 5)) : NAN;
    if (divbglob) *divbglob = std::sqrt(out[0] / npts);
    if (divbmax) *divbmax = out[1];
    if (eigvmax) *eigvmax = e;
    if (epsilon0) *epsilon0 = (c->p.artiv == 1) ? e * c->h_msh / (double)(N - 1) * 1.0 : NAN;
    return 0;
}

int chorus_plot_topology(const chorus_mesh *m, int nre, int *nnode, int *ncellp, double *xy, int *c2n) {
    if (!m || !nnode || !ncellp || !m->xv || !m->yv || !m->ivcell || !m->if2c) return CHORUS_ERR_ARG;
    PlotTopology t;
    std::string err;
    const int rc = build_plot_topology(m->ncell, m->nvert, m->nface, m->xv, m->yv, m->ivcell, m->if2c, nre, t, err);
    if (rc) return rc;
    *nnode = t.nnode;
    *ncellp = t.ncellp;
    if (xy)
        for (int n = 0; n < t.nnode; n++) { xy[n] = t.px[n + 1]; xy[t.nnode + n] = t.py[n + 1]; }
    if (c2n) std::copy(t.c2n.begin(), t.c2n.end(), c2n);
    return 0;
}

namespace {
// build the plot topology of mesh m (N points per direction, nre plot cells per
// edge) into ps, replacing what it held
int plot_state_setup(PlotState &ps, const chorus_mesh *m, int N, int nre, std::string &err) {
    auto *t = new PlotTopology;
    int rc = build_plot_topology(m->ncell, m->nvert, m->nface, m->xv, m->yv, m->ivcell, m->if2c, nre, *t, err);
    if (rc) { delete t; return rc; }
    ps.release();
    ps.t = t;
    const size_t nn = (size_t)t->nnode;
    std::vector<double> H((size_t)(nre + 1) * N);
    host_plot_basis(N, nre, H.data());
    hipError_t e = hipSuccess;
    auto ok = [&](hipError_t r) { if (e == hipSuccess) e = r; return e == hipSuccess; };
    ok(hipMalloc(&ps.cstart, (nn + 1) * sizeof(int))) && ok(hipMalloc(&ps.contrib, t->contrib.size() * sizeof(int))) &&
        ok(hipMalloc(&ps.factor, t->factor.size() * sizeof(int))) && ok(hipMalloc(&ps.H, H.size() * sizeof(double))) &&
        ok(hipMalloc(&ps.q, nn * 9 * sizeof(double))) && ok(hipMalloc(&ps.av, nn * 2 * sizeof(double))) &&
        ok(hipMalloc(&ps.V, (size_t)m->ncell * (nre + 1) * (nre + 1) * 9 * sizeof(double))) &&
        ok(hipMemcpy(ps.cstart, t->cstart.data(), (nn + 1) * sizeof(int), hipMemcpyHostToDevice)) &&
        ok(hipMemcpy(ps.contrib, t->contrib.data(), t->contrib.size() * sizeof(int), hipMemcpyHostToDevice)) &&
        ok(hipMemcpy(ps.factor, t->factor.data(), t->factor.size() * sizeof(int), hipMemcpyHostToDevice)) &&
        ok(hipMemcpy(ps.H, H.data(), H.size() * sizeof(double), hipMemcpyHostToDevice));
    if (e != hipSuccess) {
        ps.release();
        err = std::string("plot set-up: ") + hipGetErrorString(e);
        return CHORUS_ERR_HIP;
    }
    return 0;
}

// from here on every viscous stage of c also keeps the By indicator of its
// input, which the output of the last stage's values needs (muAV(1) per cell
// is kept by every AV stage)
hipError_t plot_av_columns(chorus_ctx *c) {
    if (c->ideal || !c->p.artiv || c->sm6) return hipSuccess;
    hipError_t e = hipMalloc(&c->sm6, (size_t)c->ncell * sizeof(double));
    if (e == hipSuccess) e = hipMemset(c->sm6, 0, (size_t)c->ncell * sizeof(double));
    return e;
}

// node values of a.Q (ncell cells of c's order) on the device, then the TDV111
// file; a carries the per-mesh pointers, the rest comes from ps and c
int plot_state_write(const PlotState &ps, chorus_ctx *c, PlotArgs a, const char *path, double ctime, std::string &err) {
    const PlotTopology &t = *ps.t;
    a.nnode = t.nnode;
    a.N = c->N;
    a.nre = t.nre;
    a.cstart = ps.cstart;
    a.contrib = ps.contrib;
    a.factor = ps.factor;
    a.H = ps.H;
    a.gam = c->p.gam;
    a.pi = std::atan(1.0) * 4.0;
    a.q = ps.q;
    a.av = ps.av;
    a.V = ps.V;
    hipError_t e = launch_plot_nodes(a, c->stream);
    std::vector<double> q((size_t)t.nnode * 9), avh((size_t)t.nnode * 2);
    if (e == hipSuccess) e = hipMemcpyAsync(q.data(), ps.q, q.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(avh.data(), ps.av, avh.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        err = std::string("plot: ") + hipGetErrorString(e);
        return CHORUS_ERR_HIP;
    }
    return write_tecplot_file(path, ctime, t, q.data(), avh.data(), err);
}
}  // namespace

int chorus_plot_setup(chorus_ctx *c, const chorus_mesh *m, int nre) {
    if (!c || !m) return fail(c, CHORUS_ERR_ARG, "null argument");
    if (nre < 1) return fail(c, CHORUS_ERR_ARG, "NRE must be >= 1");
    if (c->ng == 8) return fail(c, CHORUS_ERR_ARG, "Tecplot output of 8-node cells is not supported");
    if (m->ncell != c->ncell || m->N != c->N || !m->xv || !m->yv || !m->ivcell || !m->if2c)
        return fail(c, CHORUS_ERR_ARG, "mesh does not match the context");
    std::string err;
    const int rc = plot_state_setup(c->plot, m, c->N, nre, err);
    if (rc) return fail(c, rc, err);
    HIP_TRY(c, plot_av_columns(c));
    return 0;
}

int chorus_write_tecplot(chorus_ctx *c, const char *path, double ctime) {
    if (!c || !path) return fail(c, CHORUS_ERR_ARG, "null argument");
    if (!c->plot.t) return fail(c, CHORUS_ERR_ARG, "chorus_plot_setup was not called");
    const bool av = !c->ideal && c->p.artiv && c->last_parity >= 0;
    PlotArgs a{};
    a.ncell = c->ncell;
    a.artiv = av ? 1 : 0;
    a.Q = c->X[c->cur];
    a.mucell = c->mucell;
    a.sm6 = c->sm6;
    a.eig = av ? c->eig + (size_t)c->last_parity * c->ninst : nullptr;
    a.h_msh = c->h_msh;
    std::string err;
    const int rc = plot_state_write(c->plot, c, a, path, ctime, err);
    return rc ? fail(c, rc, err) : 0;
}

namespace {
// restart dump of ncell cells of c's N starting at device state Q; the image
// buffer c->rs_img holds all of c's cells (allocated on first use)
int restart_dump(chorus_ctx *c, int ncell, const double *Q, const char *path, int iter, double ctime,
                 const int *icg, std::string &err) {
    auto hip_fail = [&](const char *what, hipError_t e) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return CHORUS_ERR_HIP;
    };
    const size_t rec = 1 + 2 * (size_t)9 * c->N * c->N, words = rec * ncell;
    hipError_t e;
    if (!c->rs_img && (e = hipMalloc(&c->rs_img, rec * c->ncell * sizeof(unsigned))) != hipSuccess)
        return hip_fail("restart image", e);
    if ((e = launch_restart_image(ncell, c->N, Q, c->rs_img, c->stream)) != hipSuccess)
        return hip_fail("restart launch", e);
    std::vector<unsigned> img(words);
    if ((e = hipMemcpyAsync(img.data(), c->rs_img, words * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(c->stream)) != hipSuccess)
        return hip_fail("restart copy", e);
    if (icg)
        for (int i = 0; i < ncell; i++) img[(size_t)i * rec] = (unsigned)icg[i];
    // header: int32 iter, FP64 ctime (write_data.f90:47-50), then the records
    unsigned char head[12];
    const int32_t it = iter;
    std::memcpy(head, &it, 4);
    std::memcpy(head + 4, &ctime, 8);
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        err = std::string("cannot open ") + path;
        return CHORUS_ERR_ARG;
    }
    const bool ok = std::fwrite(head, 1, 12, f) == 12 && std::fwrite(img.data(), sizeof(unsigned), words, f) == words;
    if (std::fclose(f) != 0 || !ok) {
        err = std::string("write failed: ") + path;
        return CHORUS_ERR_ARG;
    }
    return 0;
}
}  // namespace

int chorus_write_restart(chorus_ctx *c, const char *path, int iter, double ctime, const int *icg) {
    if (!c || !path) return fail(c, CHORUS_ERR_ARG, "null argument");
    std::string err;
    const int rc = restart_dump(c, c->ncell, c->X[c->cur], path, iter, ctime, icg, err);
    return rc ? fail(c, rc, err) : 0;
}

int chorus_time_steps(chorus_ctx *c, int nsteps, double *ms_total, double *ms_stage_avg, double *ms_ct_avg) {
    if (!c || nsteps < 1) return fail(c, CHORUS_ERR_ARG, "bad arguments");
    const size_t need = (size_t)nsteps * 12 + 2;
    while (c->ev.size() < need) {
        hipEvent_t e;
        HIP_TRY(c, hipEventCreate(&e));
        c->ev.push_back(e);
    }
    hipEvent_t t0 = c->ev[need - 2], t1 = c->ev[need - 1];
    c->ev_used = 0;
    c->timing = true;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipEventRecord(t0, c->stream));
    int rc = 0;
    for (int i = 0; i < nsteps && !rc; i++) rc = one_step(c);
    c->timing = false;
    if (rc) return rc;
    HIP_TRY(c, hipEventRecord(t1, c->stream));
    HIP_TRY(c, hipEventSynchronize(t1));
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, t0, t1));
    if (ms_total) *ms_total = ms;
    // per step: 5 stage launches then 1 CT launch, each bracketed by an event pair
    double st = 0.0, ct = 0.0;
    int nst = 0, nct = 0;
    for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
        float d = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&d, c->ev[i], c->ev[i + 1]));
        if ((i / 2) % 6 == 5) { ct += d; nct++; } else { st += d; nst++; }
    }
    if (ms_stage_avg) *ms_stage_avg = nst ? st / nst : 0.0;
    if (ms_ct_avg) *ms_ct_avg = nct ? ct / nct : 0.0;
    return 0;
}

}  // extern "C"

// ===========================================================================
// Batched instances: groups of instances with equal numerics share one
// context whose mesh is the disjoint union of their meshes; a per-cell
// instance index selects the per-instance AV quantities.  One step of the
// batch is one hipGraph launch whose branches are the groups.
// ===========================================================================

struct chorus_batch {
    std::string err;
    int ninst = 0;
    struct Member {
        int group, off, ncell, N;
        int li = 0;          // index among the group's instances
        double h = 0.0;      // AV length scale (AV.f90:147)
        bool av_valid = false;   // a stage has evaluated the member's current state
        PlotState plot;      // Tecplot output (chorus_batch_plot_setup)
    };
    std::vector<Member> mem;
    std::vector<chorus_ctx *> groups;
    std::vector<hipStream_t> gstream;   // capture branch of each group
    hipStream_t stream = nullptr;       // capture and launch stream
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    struct Graph {
        hipGraphExec_t exec = nullptr;
        std::vector<std::array<int, 4>> post;   // per group: cur, last_in, stage_parity, last_parity
    };
    std::map<std::vector<int>, Graph> graphs;
    bool stale = true;                  // traces / AV indicators not yet evaluated
};

namespace {

int bfail(chorus_batch *b, int code, const std::string &msg) {
    if (b) b->err = msg;
    return code;
}

#define HIP_TRYB(b, expr)                                                                    \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return bfail(b, CHORUS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

bool same_numerics(const chorus_params &a, const chorus_params &b, int Na, int Nb) {
    return Na == Nb && a.precision == b.precision && a.k_stage == b.k_stage && a.vismode == b.vismode &&
           a.artiv == b.artiv && a.mu == b.mu && a.eta0 == b.eta0 && a.gam == b.gam && a.dt == b.dt;
}

void drop_graphs(chorus_batch *b) {
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second.exec);
    b->graphs.clear();
}

int refresh(chorus_batch *b) {
    if (!b->stale) return 0;
    for (chorus_ctx *c : b->groups) {
        int rc = run_trace(c, c->X[c->cur]);
        if (rc) return bfail(b, rc, c->err);
    }
    for (chorus_ctx *c : b->groups) HIP_TRYB(b, hipStreamSynchronize(c->stream));
    b->stale = false;
    return 0;
}

int batch_one_step(chorus_batch *b) {
    std::vector<int> key;
    for (chorus_ctx *c : b->groups) { key.push_back(c->cur); key.push_back(c->stage_parity); }
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        chorus_batch::Graph g;
        HIP_TRYB(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        (void)hipEventRecord(b->fork, b->stream);
        int rc = 0;
        for (size_t k = 0; k < b->groups.size() && !rc; k++) {
            chorus_ctx *c = b->groups[k];
            (void)hipStreamWaitEvent(b->gstream[k], b->fork, 0);
            hipStream_t own = c->stream;
            c->stream = b->gstream[k];
            rc = one_step(c);
            c->stream = own;
            if (rc) b->err = c->err;
            (void)hipEventRecord(b->join[k], b->gstream[k]);
            (void)hipStreamWaitEvent(b->stream, b->join[k], 0);
        }
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &graph);
        if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
        if (e != hipSuccess) return bfail(b, CHORUS_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return bfail(b, CHORUS_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        for (chorus_ctx *c : b->groups) g.post.push_back({c->cur, c->last_in, c->stage_parity, c->last_parity});
        // the capture advanced the host-side bookkeeping; rewind it to replay below
        for (size_t k = 0; k < b->groups.size(); k++) {
            b->groups[k]->cur = key[2 * k];
            b->groups[k]->stage_parity = key[2 * k + 1];
        }
        it = b->graphs.emplace(key, g).first;
    }
    HIP_TRYB(b, hipGraphLaunch(it->second.exec, b->stream));
    for (size_t k = 0; k < b->groups.size(); k++) {
        chorus_ctx *c = b->groups[k];
        const auto &p = it->second.post[k];
        c->cur = p[0]; c->last_in = p[1]; c->stage_parity = p[2]; c->last_parity = p[3];
    }
    return 0;
}

}  // namespace

extern "C" {

const char *chorus_batch_last_error(const chorus_batch *b) { return b ? b->err.c_str() : "null batch"; }

void chorus_batch_free(chorus_batch *b) {
    if (!b) return;
    drop_graphs(b);
    for (auto &m : b->mem) m.plot.release();
    for (chorus_ctx *c : b->groups) chorus_free(c);
    for (auto s : b->gstream) (void)hipStreamDestroy(s);
    for (auto e : b->join) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int chorus_batch_init(int ninst, const chorus_mesh *meshes, const chorus_params *params, chorus_batch **out) {
    if (!out) return CHORUS_ERR_ARG;
    *out = nullptr;
    chorus_batch *b = new chorus_batch();
    *out = b;
    if (ninst < 1 || !meshes || !params) return bfail(b, CHORUS_ERR_ARG, "empty batch");
    b->ninst = ninst;
    b->mem.resize(ninst);
    // group assignment in order of first appearance
    std::vector<std::vector<int>> members;
    for (int i = 0; i < ninst; i++) {
        if (params[i].k_stage != params[0].k_stage)
            return bfail(b, CHORUS_ERR_ARG, "all instances of a batch must use the same k_stage");
        int g = -1;
        for (size_t k = 0; k < members.size() && g < 0; k++) {
            const int j = members[k][0];
            if (same_numerics(params[i], params[j], meshes[i].N, meshes[j].N)) g = (int)k;
        }
        if (g < 0) { g = (int)members.size(); members.emplace_back(); }
        members[g].push_back(i);
    }
    for (size_t g = 0; g < members.size(); g++) {
        // disjoint union of the members' meshes (reference layouts, offsets applied)
        int NC = 0, NV_ = 0, NF = 0, NI = 0, NPr = 0, NFr = 0;
        for (int i : members[g]) {
            const chorus_mesh &m = meshes[i];
            if (!m.xv || !m.yv || !m.ivcell || !m.if2c || m.ncell <= 0)
                return bfail(b, CHORUS_ERR_ARG, "instance " + std::to_string(i) + ": incomplete mesh");
            NC += m.ncell; NV_ += m.nvert; NF += m.nface; NI += m.ninter; NPr += m.npair; NFr += m.nfree;
        }
        std::vector<double> xv(NV_), yv(NV_);
        std::vector<int> ivc((size_t)NC * 4), if2c((size_t)NF * 4), ibi(NI), cyc((size_t)NPr * 2), ibf(NFr), inst(NC);
        std::vector<double> hinst;
        int coff = 0, voff = 0, foff = 0, ioff = 0, poff = 0, froff = 0, li = 0;
        for (int i : members[g]) {
            const chorus_mesh &m = meshes[i];
            const chorus_params &p = params[i];
            b->mem[i] = {(int)g, coff, m.ncell, m.N};
            for (int v = 0; v < m.nvert; v++) { xv[voff + v] = m.xv[v]; yv[voff + v] = m.yv[v]; }
            for (int k = 0; k < 4; k++)
                for (int c = 0; c < m.ncell; c++) ivc[(size_t)k * NC + coff + c] = m.ivcell[(size_t)k * m.ncell + c] + voff;
            for (int k = 0; k < 4; k++)
                for (int f = 0; f < m.nface; f++) {
                    int v = m.if2c[(size_t)k * m.nface + f];
                    if (k < 2 && v != 0) v += coff;
                    if2c[(size_t)k * NF + foff + f] = v;
                }
            for (int k = 0; k < m.ninter; k++) ibi[ioff + k] = m.ibfinter[k] + foff;
            for (int k = 0; k < 2; k++)
                for (int q = 0; q < m.npair; q++) cyc[(size_t)k * NPr + poff + q] = m.cyc2fpair[(size_t)k * m.npair + q] + foff;
            for (int k = 0; k < m.nfree; k++) ibf[froff + k] = m.ibfree[k] + foff;
            for (int c = 0; c < m.ncell; c++) inst[coff + c] = li;
            const int ncg = p.ncellglob > 0 ? p.ncellglob : m.ncell;
            hinst.push_back(std::sqrt(p.dxcycl * p.dycycl / (double)ncg));   // AV.f90:147
            b->mem[i].li = li;
            b->mem[i].h = hinst.back();
            coff += m.ncell; voff += m.nvert; foff += m.nface; ioff += m.ninter; poff += m.npair; froff += m.nfree;
            li++;
        }
        chorus_mesh mm{};
        mm.ncell = NC; mm.nvert = NV_; mm.N = meshes[members[g][0]].N;
        mm.xv = xv.data(); mm.yv = yv.data(); mm.ivcell = ivc.data();
        mm.nface = NF; mm.if2c = if2c.data();
        mm.ninter = NI; mm.ibfinter = ibi.data();
        mm.npair = NPr; mm.cyc2fpair = cyc.data();
        mm.nfree = NFr; mm.ibfree = ibf.data();
        chorus_params pp = params[members[g][0]];
        pp.ncellglob = NC;
        chorus_ctx *c = new chorus_ctx();
        c->ninst = (int)members[g].size();
        b->groups.push_back(c);
        int rc = init_ctx(c, &mm, &pp);
        if (rc) return bfail(b, rc, "group " + std::to_string(g) + ": " + c->err);
        if (c->ninst > 1) {
            HIP_TRYB(b, hipMalloc(&c->inst, inst.size() * sizeof(int)));
            HIP_TRYB(b, hipMalloc(&c->hinst, hinst.size() * sizeof(double)));
            HIP_TRYB(b, hipMemcpy(c->inst, inst.data(), inst.size() * sizeof(int), hipMemcpyHostToDevice));
            HIP_TRYB(b, hipMemcpy(c->hinst, hinst.data(), hinst.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {
            c->h_msh = hinst[0];
        }
        hipStream_t s;
        HIP_TRYB(b, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        b->gstream.push_back(s);
        hipEvent_t e;
        HIP_TRYB(b, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->join.push_back(e);
    }
    HIP_TRYB(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIP_TRYB(b, hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    return 0;
}

int chorus_batch_set_state(chorus_batch *b, int i, const double *q) {
    if (!b || !q || i < 0 || i >= b->ninst) return bfail(b, CHORUS_ERR_ARG, "bad instance");
    const chorus_batch::Member &m = b->mem[i];
    chorus_ctx *c = b->groups[m.group];
    const int NP = m.N * m.N;
    std::vector<double> h((size_t)m.ncell * 9 * NP);
    for (int ic = 0; ic < m.ncell; ic++)
        for (int p = 0; p < NP; p++)
            for (int v = 0; v < 9; v++) h[((size_t)ic * 9 + v) * NP + p] = q[((size_t)ic * NP + p) * 9 + v];
    HIP_TRYB(b, hipMemcpy(c->X[c->cur] + (size_t)m.off * 9 * NP, h.data(), h.size() * sizeof(double),
                          hipMemcpyHostToDevice));
    c->last_in = -1;
    b->mem[i].av_valid = false;   // its AV output columns belong to the previous state
    b->stale = true;
    return 0;
}

int chorus_batch_get_state(chorus_batch *b, int i, double *q) {
    if (!b || !q || i < 0 || i >= b->ninst) return bfail(b, CHORUS_ERR_ARG, "bad instance");
    const chorus_batch::Member &m = b->mem[i];
    chorus_ctx *c = b->groups[m.group];
    const int NP = m.N * m.N;
    HIP_TRYB(b, hipStreamSynchronize(b->stream));
    std::vector<double> h((size_t)m.ncell * 9 * NP);
    HIP_TRYB(b, hipMemcpy(h.data(), c->X[c->cur] + (size_t)m.off * 9 * NP, h.size() * sizeof(double),
                          hipMemcpyDeviceToHost));
    for (int ic = 0; ic < m.ncell; ic++)
        for (int p = 0; p < NP; p++)
            for (int v = 0; v < 9; v++) q[((size_t)ic * NP + p) * 9 + v] = h[((size_t)ic * 9 + v) * NP + p];
    return 0;
}

int chorus_batch_write_restart(chorus_batch *b, int i, const char *path, int iter, double ctime, const int *icg) {
    if (!b || !path || i < 0 || i >= b->ninst) return bfail(b, CHORUS_ERR_ARG, "bad instance");
    const chorus_batch::Member &m = b->mem[i];
    chorus_ctx *c = b->groups[m.group];
    HIP_TRYB(b, hipStreamSynchronize(b->stream));
    std::string err;
    const int rc = restart_dump(c, m.ncell, c->X[c->cur] + (size_t)m.off * 9 * m.N * m.N, path, iter, ctime, icg, err);
    return rc ? bfail(b, rc, err) : 0;
}

int chorus_batch_plot_setup(chorus_batch *b, int i, const chorus_mesh *mesh, int nre) {
    if (!b || !mesh || i < 0 || i >= b->ninst) return bfail(b, CHORUS_ERR_ARG, "bad instance");
    if (nre < 1) return bfail(b, CHORUS_ERR_ARG, "NRE must be >= 1");
    chorus_batch::Member &m = b->mem[i];
    if (mesh->ncell != m.ncell || mesh->N != m.N || !mesh->xv || !mesh->yv || !mesh->ivcell || !mesh->if2c)
        return bfail(b, CHORUS_ERR_ARG, "mesh does not match instance " + std::to_string(i));
    chorus_ctx *c = b->groups[m.group];
    HIP_TRYB(b, hipStreamSynchronize(b->stream));
    std::string err;
    const int rc = plot_state_setup(m.plot, mesh, m.N, nre, err);
    if (rc) return bfail(b, rc, err);
    if (!c->ideal && c->p.artiv && !c->sm6) {
        HIP_TRYB(b, plot_av_columns(c));
        drop_graphs(b);   // the captured stages do not yet write the new AV columns
    }
    return 0;
}

int chorus_batch_write_tecplot(chorus_batch *b, int i, const char *path, double ctime) {
    if (!b || !path || i < 0 || i >= b->ninst) return bfail(b, CHORUS_ERR_ARG, "bad instance");
    const chorus_batch::Member &m = b->mem[i];
    if (!m.plot.t) return bfail(b, CHORUS_ERR_ARG, "chorus_batch_plot_setup was not called for this instance");
    chorus_ctx *c = b->groups[m.group];
    HIP_TRYB(b, hipStreamSynchronize(b->stream));
    const bool av = !c->ideal && c->p.artiv && c->last_parity >= 0 && c->sm6 && m.av_valid;
    PlotArgs a{};
    a.ncell = m.ncell;
    a.artiv = av ? 1 : 0;
    a.Q = c->X[c->cur] + (size_t)m.off * 9 * m.N * m.N;
    a.mucell = av ? c->mucell + m.off : nullptr;
    a.sm6 = av ? c->sm6 + m.off : nullptr;
    a.eig = av ? c->eig + (size_t)c->last_parity * c->ninst + m.li : nullptr;
    a.h_msh = m.h;
    std::string err;
    const int rc = plot_state_write(m.plot, c, a, path, ctime, err);
    return rc ? bfail(b, rc, err) : 0;
}

int chorus_batch_step(chorus_batch *b, int nsteps) {
    if (!b || nsteps < 0) return bfail(b, CHORUS_ERR_ARG, "bad arguments");
    int rc = refresh(b);
    for (int k = 0; k < nsteps && !rc; k++) rc = batch_one_step(b);
    if (rc) return rc;
    HIP_TRYB(b, hipStreamSynchronize(b->stream));
    if (nsteps > 0)
        for (auto &m : b->mem) m.av_valid = true;
    return 0;
}

int chorus_batch_time_steps(chorus_batch *b, int nsteps, double *ms_total) {
    if (!b || nsteps < 1) return bfail(b, CHORUS_ERR_ARG, "bad arguments");
    int rc = refresh(b);
    if (rc) return rc;
    hipEvent_t t0, t1;
    HIP_TRYB(b, hipEventCreate(&t0));
    HIP_TRYB(b, hipEventCreate(&t1));
    HIP_TRYB(b, hipEventRecord(t0, b->stream));
    for (int k = 0; k < nsteps && !rc; k++) rc = batch_one_step(b);
    (void)hipEventRecord(t1, b->stream);
    (void)hipEventSynchronize(t1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc) return rc;
    for (auto &m : b->mem) m.av_valid = true;
    if (ms_total) *ms_total = ms;
    return 0;
}

int chorus_batch_compresid(chorus_batch *b, int i, double *resid) {
    if (!b || i < 0 || i >= b->ninst) return bfail(b, CHORUS_ERR_ARG, "bad instance");
    int rc = refresh(b);
    if (rc) return rc;
    const chorus_batch::Member &m = b->mem[i];
    chorus_ctx *c = b->groups[m.group];
    HIP_TRYB(b, hipStreamSynchronize(b->stream));
    StageArgs s = base_args(c);
    s.Qin = c->X[c->cur];
    s.Rout = c->R;
    s.kind = KIND_RESID;
    rc = launch_stage(c, s);
    if (rc) return bfail(b, rc, c->err);
    HIP_TRYB(b, hipStreamSynchronize(c->stream));
    for (auto &o : b->mem)
        if (o.group == m.group) o.av_valid = true;
    if (resid) {
        const int NP = m.N * m.N;
        std::vector<double> h((size_t)m.ncell * 9 * NP);
        HIP_TRYB(b, hipMemcpy(h.data(), c->R + (size_t)m.off * 9 * NP, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int ic = 0; ic < m.ncell; ic++)
            for (int p = 0; p < NP; p++)
                for (int v = 0; v < 9; v++) resid[((size_t)ic * NP + p) * 9 + v] = h[((size_t)ic * 9 + v) * NP + p];
    }
    return 0;
}

int chorus_batch_divb(chorus_batch *b, double *divbmax) {
    if (!b || !divbmax) return bfail(b, CHORUS_ERR_ARG, "bad arguments");
    int rc = refresh(b);
    if (rc) return rc;
    HIP_TRYB(b, hipStreamSynchronize(b->stream));
    // the diagnostic reads FP64 Bx, By traces: the mixed form's groups make them
    for (chorus_ctx *c : b->groups)
        if (mixed_traces(c)) HIP_TRYB(b, launch_trace(c->N, c->ncell, c->X[c->cur], c->TQ, nullptr, c->stream));
    for (int i = 0; i < b->ninst; i++) {
        const chorus_batch::Member &m = b->mem[i];
        chorus_ctx *c = b->groups[m.group];
        CTArgs a{};
        a.c_first = m.off;
        a.ncell = m.off + m.ncell;
        a.cellxy = c->cellxy;
        a.nbr = c->nbr;
        a.Q = c->X[c->cur];
        a.which = 1;
        a.partial = c->partial;
        a.TQ = c->TQ;
        a.TQout = c->TQ;
        HIP_TRYB(b, launch_ct(m.N, a, c->stream));
        const int nb = ct_blocks(m.N, m.ncell);
        std::vector<double> part((size_t)nb * 2);
        HIP_TRYB(b, hipMemcpyAsync(part.data(), c->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRYB(b, hipStreamSynchronize(c->stream));
        double mx = 0.0;
        for (int k = 0; k < nb; k++) mx = std::max(mx, part[2 * k + 1]);
        divbmax[i] = mx;
    }
    return 0;
}

int chorus_batch_correct_b(chorus_batch *b) {
    if (!b) return CHORUS_ERR_ARG;
    int rc = refresh(b);
    if (rc) return rc;
    HIP_TRYB(b, hipStreamSynchronize(b->stream));
    for (chorus_ctx *c : b->groups) {
        rc = run_ct(c, c->X[c->cur], 0);
        if (rc) return bfail(b, rc, c->err);
    }
    for (chorus_ctx *c : b->groups) HIP_TRYB(b, hipStreamSynchronize(c->stream));
    return 0;
}

int chorus_batch_ngroups(const chorus_batch *b) { return b ? (int)b->groups.size() : 0; }

int chorus_batch_group(const chorus_batch *b, int g, int *N, int *ninst, int *ncell, int *tile) {
    if (!b || g < 0 || g >= (int)b->groups.size()) return CHORUS_ERR_ARG;
    const chorus_ctx *c = b->groups[g];
    if (N) *N = c->N;
    if (ninst) *ninst = c->ninst;
    if (ncell) *ncell = c->ncell;
    if (tile) *tile = c->tile;
    return 0;
}

}  // extern "C"

namespace {
// times KIND_RESID stage launches (the cell kernels of a stage without the
// state update) of context c for every compiled tile shape of the mixed form
// and keeps the fastest; the FP64 form has one shape
int autotune_ctx(chorus_ctx *c, hipEvent_t t0, hipEvent_t t1, int *tile) {
    int best = 0, rc = 0;
    if (c->p.precision == CHORUS_MIXED) {
        float best_ms = 1e30f;
        for (int t = 0; t < tile_count() && !rc; t++) {
            c->tile = t;
            StageArgs s = base_args(c);
            s.Qin = c->X[c->cur];
            s.Rout = c->R;
            s.kind = KIND_RESID;
            if ((rc = launch_stage(c, s))) break;   // warm-up
            if (hipEventRecord(t0, c->stream) != hipSuccess) return CHORUS_ERR_HIP;
            for (int r = 0; r < 3 && !rc; r++) rc = launch_stage(c, s);
            if (hipEventRecord(t1, c->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess)
                return CHORUS_ERR_HIP;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, t0, t1) != hipSuccess) return CHORUS_ERR_HIP;
            if (!rc && ms < best_ms) { best_ms = ms; best = t; }
        }
    }
    c->tile = best;
    if (tile) *tile = best;
    return rc;
}
}  // namespace

extern "C" {

int chorus_autotune(chorus_ctx *c, int *tile) {
    if (!c) return CHORUS_ERR_ARG;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    hipEvent_t t0, t1;
    HIP_TRY(c, hipEventCreate(&t0));
    HIP_TRY(c, hipEventCreate(&t1));
    const int rc = autotune_ctx(c, t0, t1, tile);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return rc ? fail(c, rc, "autotune failed") : 0;
}

int chorus_set_tile(chorus_ctx *c, int tile) {
    if (!c) return CHORUS_ERR_ARG;
    if (tile < 0 || tile >= tile_count()) return fail(c, CHORUS_ERR_ARG, "tile index out of range");
    if (c->p.precision != CHORUS_MIXED && tile != 0) return fail(c, CHORUS_ERR_ARG, "the FP64 form has one tile shape");
    c->tile = tile;
    return 0;
}

int chorus_batch_autotune(chorus_batch *b, int *tiles) {
    if (!b) return CHORUS_ERR_ARG;
    int rc = refresh(b);
    if (rc) return rc;
    HIP_TRYB(b, hipStreamSynchronize(b->stream));
    drop_graphs(b);
    hipEvent_t t0, t1;
    HIP_TRYB(b, hipEventCreate(&t0));
    HIP_TRYB(b, hipEventCreate(&t1));
    for (size_t g = 0; g < b->groups.size() && !rc; g++)
        rc = autotune_ctx(b->groups[g], t0, t1, tiles ? tiles + g : nullptr);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return rc ? bfail(b, rc, "autotune failed") : 0;
}

}  // extern "C"
