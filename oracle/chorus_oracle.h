/* Oracle: plain-C FP64 restatement of the CHORUS-MHD residual + time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker; the
 * product (libchorus_mhd.so) never links or calls it.
 *
 * Every array keeps the reference's Fortran column-major layout and 1-based
 * integer contents (setup2d.f90:39-60), so dumps from the compiled reference
 * compare element by element.  Expressions follow the reference's evaluation
 * order and the library is built with -ffp-contract=off, like the reference
 * build (no FMA), so results are bitwise comparable.
 */
#ifndef CHORUS_ORACLE_H
#define CHORUS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* Mesh topology, CONNECTIVITY + CONNECT_BDRY + MAPCYCLOC restated
 * (connectivity.f90:1-233, connect_bdry.f90:1-354, map_cyc_loc.f90:1-42).
 * Inputs (1-based ids): ivcell(NCELL,4) column-major, bnd(NBND,2) column-major
 * vertex pairs, btype[NBND] in {0 other, 1 CYCL, 2 FREE, 3 INLE, 4 OUTL,
 * 5 SYMP, 6 WALL}.  Outputs are allocated by the oracle and freed with
 * orc_topo_free.  Returns 0 or a negative code. */
typedef struct {
    int nface, ninter, npair, nfree;
    int *if2c;       /* IF2C(NFACE,4) */
    int *if2v;       /* IF2V(NFACE,2) */
    int *ibfinter;   /* IBFINTER(NINTER) */
    int *cyc2fpair;  /* CYC2FPAIR(NPAIR,2) */
    int *ibfree;     /* IBFREE(NFREE) */
    /* IBFINL, IBFOUT, IBFSYMP, IBFWAL in boundary-file order (connect_bdry.f90:186-205) */
    int ninl, nout, nsymp, nwal;
    int *ibfinl, *ibfout, *ibfsymp, *ibfwal;
} orc_topo;

int  orc_topology(int ncell, int nvert, const int *ivcell, const double *xv, const double *yv,
                  int nbnd, const int *bnd, const int *btype, double dxcycl, double dycycl,
                  orc_topo *out);
void orc_topo_free(orc_topo *t);

typedef struct {
    int N;            /* solution points per direction */
    int vismode;      /* 0|1 (readinput.f90:21) */
    int artiv;        /* 0|1 */
    int k_stage;      /* 3 -> RK_S5P3, 5 -> FiveStageRk */
    double mu, eta0, gam, dt, dxcycl, dycycl;
    double qin[9];    /* INLE faces: the prescribed state (the reference has none, see bcflux) */
} orc_params;

/* INIT_SETUP restated (initsetup.f90:1-165 minus I/O): operators, metrics,
 * coordinates, face maps.  Copies every input array. */
orc_ctx *orc_create(const orc_params *p, int ncell, int nvert, const int *ivcell,
                    const double *xv, const double *yv, const orc_topo *topo);
void orc_destroy(orc_ctx *c);

/* 8-node cells: the mid-edge nodes 5..8 of every cell (node 5 on edge 1-2, 6 on
 * 2-3, 7 on 3-4, 8 on 4-1: the orientation of mapder.f90:106-115), XM(NCELL,4)
 * and YM(NCELL,4) column-major; the metrics and solution-point coordinates are
 * recomputed with the 8-node map (MAPDER_8_NODE, mapder.f90:117-253, summed as
 * calcjacob.f90 sums the 4-node map).  Returns 0. */
int orc_set_nodes8(orc_ctx *c, const double *xm, const double *ym);
/* MAPDER_8_NODE's tables for N points: dmdxs(8,2,N,N), dmdxf1(8,2,N+1,N),
 * dmdxf2(8,2,N,N+1), column-major */
void orc_mapder8(int N, double *dmdxs, double *dmdxf1, double *dmdxf2);

/* state Q(9,N,N,NCELL) */
void orc_set_q(orc_ctx *c, const double *q);
void orc_get_q(const orc_ctx *c, double *q);
/* Orszag-Tang IC + B := curl A_z + energy recomputed (setinitialcond.f90:1-80) */
void orc_setinitialcond_ot(orc_ctx *c);
/* compresid (compresid.f90:1-67); resid(9,N,N,NCELL) copied out if non-NULL */
void orc_compresid(orc_ctx *c, double *resid_out);
/* correct_B_via_curl_Az (setinitialcond.f90:82-92) */
void orc_correct_b(orc_ctx *c);
/* nsteps iterations of the iterations.f90 loop body: stepper + CT + the
 * diagnostic flux/gradient; returns RMS resid and RMS divB of the last step
 * (iterations.f90:66-67) and the max |divB| */
void orc_steps(orc_ctx *c, int nsteps, double *residglob, double *divbglob, double *divbmax);
/* read-only views of intermediates (reference layouts) */
const double *orc_array(const orc_ctx *c, const char *name, long *count);
double orc_scalar(const orc_ctx *c, const char *name);

#ifdef __cplusplus
}
#endif
#endif
