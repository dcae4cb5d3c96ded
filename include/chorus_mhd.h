/* chorus_mhd.h -- C-ABI of the MI355X-native CHORUS-MHD residual + time step.
 *
 * The reference exposes no plugin or FFI API: its hot path is a set of
 * argument-less Fortran subroutines operating on `module setup2d`
 * (setup2d.f90:1-99).  Each entry point below replaces one of them and is what
 * an `iso_c_binding` interface in the reference's Fortran driver binds to
 * (chorus-mhd_amd/fortran/chorus_hip.f90).  All arrays cross the boundary in
 * the reference's own layouts: Fortran column-major, 1-based connectivity,
 * Q(9,N,N,NCELL) FP64.  Host arrays are borrowed for the duration of a call
 * (copied); the context owns all device memory and its HIP stream.
 *
 * Every function returns 0 on success and a negative code on error; the
 * message is available from chorus_last_error().  A context is not
 * thread-safe; calls are synchronous with respect to the host.
 */
#ifndef CHORUS_MHD_H
#define CHORUS_MHD_H

#ifdef __cplusplus
extern "C" {
#endif

#define CHORUS_OK            0
#define CHORUS_ERR_ARG      -1   /* invalid argument / unsupported configuration */
#define CHORUS_ERR_HIP      -2   /* HIP runtime error (no device, out of memory, launch) */
#define CHORUS_ERR_TOPOLOGY -3   /* connectivity / periodic matching failure */

/* precision forms */
#define CHORUS_FP64   0   /* FP64 drop-in: every operation in FP64, reference order */
#define CHORUS_MIXED  1   /* FP32 flux/contraction arithmetic; FP64 RK sums, CT, metrics */

typedef struct chorus_ctx chorus_ctx;

/* Mesh after the reference's set-up chain (main.f90:16-42): vertex
 * coordinates XV/YV (readGRIDinfo.f90:283-290), IVCELL(NCELL,4)
 * (counter-clockwise, connectivity.f90:236-283), IF2C(NFACE,4) left cell,
 * right cell, left local face, right local face (connectivity.f90:156-215),
 * IBFINTER(NINTER) (connect_bdry.f90:228-244), CYC2FPAIR(NPAIR,2)
 * (map_cyc_loc.f90:12-35) and IBFREE(NFREE) (connect_bdry.f90:217-221). */
typedef struct {
    int ncell, nvert, N;          /* N = solution points per direction (2..8) */
    const double *xv, *yv;        /* XV(NVERT), YV(NVERT) */
    const int *ivcell;            /* IVCELL(NCELL,4) */
    int nface;
    const int *if2c;              /* IF2C(NFACE,4) */
    int ninter;
    const int *ibfinter;          /* IBFINTER(NINTER) */
    int npair;
    const int *cyc2fpair;         /* CYC2FPAIR(NPAIR,2) */
    int nfree;
    const int *ibfree;            /* IBFREE(NFREE) */
} chorus_mesh;

/* Run parameters of QUAD.INP (readinput.f90:16-33) plus the compile-time
 * constants the reference fixes in setup2d.f90:64-71 (gam). */
typedef struct {
    int vismode;                  /* must be 1 (the reference reads Qv/Az flux-point data only then) */
    int artiv;                    /* 0|1 artificial viscosity (AV.f90) */
    int k_stage;                  /* 3 -> RK_S5P3, 5 -> FiveStageRk (iterations.f90:22-31) */
    double mu, eta0;              /* viscosity, resistivity */
    double gam;                   /* ratio of specific heats (setup2d.f90:68 uses 5/3) */
    double dt;
    double dxcycl, dycycl;        /* periodic shifts; also the AV length scale (AV.f90:147) */
    int ncellglob;                /* NCELLGLOB (AV.f90:147, iterations.f90:66-67); 0 -> ncell */
    int precision;                /* CHORUS_FP64 | CHORUS_MIXED */
} chorus_params;

/* INIT_SETUP replacement (initsetup.f90:1-153): uploads the mesh, builds the
 * on-device face connectivity, operators and metrics.  Replaces the set-up half
 * of initsetup.f90; the initial state is supplied with chorus_set_state. */
int chorus_init(const chorus_mesh *mesh, const chorus_params *params, chorus_ctx **out);
void chorus_free(chorus_ctx *ctx);
const char *chorus_last_error(const chorus_ctx *ctx);

/* Boundary faces of the four kinds the reference lists (IBFINL, IBFOUT,
 * IBFSYMP, IBFWAL: connect_bdry.f90:88-110, 186-205) but never gives a flux
 * (compflux.f90:33-49 treats only interior, CYCL and FREE faces).  This build
 * states their physics as the right state of the inviscid Rusanov flux
 * (getrusanov.f90:1-109) at the face point, formed from the cell's own trace
 * Q_L and the face's unit normal n:
 *   OUTL        zero gradient: Q_R = Q_L (as FREE, bcflux.f90:46)
 *   SYMP, WALL  mirror (inviscid slip wall / symmetry plane): Q_R = Q_L with
 *               the normal momentum and the normal magnetic field reversed,
 *               m_R = m_L - 2 (m_L.n) n, B_R = B_L - 2 (B_L.n) n, so the
 *               normal mass and energy fluxes vanish
 *   INLE        prescribed state: Q_R(1:8) = qin(1:8), A_z from the cell
 * Everything else at these faces is the FREE treatment (bcflux.f90:73-121,
 * AV.f90:199-303): the central flux-point values and gradients of the viscous
 * path, the AV viscosity and the A_z of the CT correction take the cell's own
 * trace.  Face numbers are 1-based, as in IF2C. */
typedef struct {
    int ninlet;  const int *ibfinl;   /* IBFINL(NINLET) */
    int noutlet; const int *ibfout;   /* IBFOUT(NOUTLET) */
    int nsymp;   const int *ibfsymp;  /* IBFSYMP(NSYMP) */
    int nwall;   const int *ibfwal;   /* IBFWAL(NWALL) */
    double qin[9];                    /* inlet state (conserved variables, Q(1:9) order; qin[8] unused) */
} chorus_bc;

/* chorus_init for meshes with INLE/OUTL/SYMP/WALL faces (bc may be NULL:
 * then it is chorus_init).  Single contexts only: the batched API takes
 * interior, CYCL and FREE faces. */
int chorus_init_bc(const chorus_mesh *mesh, const chorus_params *params, const chorus_bc *bc, chorus_ctx **out);

/* Curved cells (CURVE_WALL = 1): replaces the bilinear map of each cell by
 * the 8-node serendipity map of MapBaseFunc8 / MAPDER_8_NODE
 * (mapder.f90:91-253), nodes 1-4 the cell's vertices and 5-8 the edge
 * nodes xm, ym given here: XM(NCELL,4) column-major, column k the node on
 * edge k of the cell (1: vertices 1-2, 2: 2-3, 3: 3-4, 4: 4-1).  The metric
 * terms and Jacobians of the residual and of the CT correction then come
 * from the 8-node derivatives (the sums of calcjacob.f90:20-28 over 8
 * nodes).  The reference prepares these tables and stops
 * (initsetup.f90:71-78); this completes that branch.  Both precision forms,
 * the ideal, viscous and AV paths; no Tecplot output (chorus_plot_setup
 * returns CHORUS_ERR_ARG); call before the first step. */
int chorus_set_nodes8(chorus_ctx *ctx, const double *xm, const double *ym);

/* Q(9,N,N,NCELL) in the reference layout (setup2d.f90:54).  Upload/download. */
int chorus_set_state(chorus_ctx *ctx, const double *q);
int chorus_get_state(chorus_ctx *ctx, double *q);

/* compresid (compresid.f90:1-67): one residual evaluation of the current
 * state; resid(9,N,N,NCELL) copied out when resid != NULL. */
int chorus_compresid(chorus_ctx *ctx, double *resid);

/* correct_B_via_curl_Az (setinitialcond.f90:82-92 -> magnetic_potential.f90:44-67). */
int chorus_correct_b(chorus_ctx *ctx);

/* nsteps iterations of the time loop body (iterations.f90:22-33): the SSP
 * stepper of k_stage (RK_S5P3, iterations.f90:162-217, or FiveStageRk,
 * iterations.f90:92-160) followed by correct_B_via_curl_Az.  Each step is one
 * hipGraph launch on the context's stream (captured on first use per buffer
 * rotation). */
int chorus_step(chorus_ctx *ctx, int nsteps);

/* Diagnostics of iterations.f90:35-67 for the current state: RMS residual of
 * the last stage evaluated, RMS div B (reference definitions), max |div B|,
 * the last stage's maximum Rusanov wave speed and AV amplitude epsilon0
 * (getrusanov.f90:105-107, AV.f90:145-148).  Any pointer may be NULL. */
int chorus_diag(chorus_ctx *ctx, double *residglob, double *divbglob, double *divbmax,
                double *eigvmax, double *epsilon0);

/* Allocation-free variant of chorus_diag (no counterpart in the reference; for
 * callers that cannot allocate device memory or use heap-allocating library
 * routines during the time loop: serving loops and harnesses that forbid
 * hipMalloc).  Same outputs as chorus_diag: div B RMS and max bit for bit,
 * the residual norm summed in another order (agreeing to about 1e-15
 * relative).  The kernels' partials live in a fixed 8 MiB device array per
 * GPU (held until the process ends) instead of context memory, and only the
 * final scalars come back to the host.  Limit: 2 x (cell blocks) + 1027
 * doubles must fit in 1048576 (cell blocks = ceil(NCELL / floor(256/N^2)),
 * e.g. NCELL <= 5.2M at N=5); larger meshes get CHORUS_ERR_ARG.  Every call
 * on one GPU shares the array: calls that overlap on one GPU (several host
 * threads, or contexts whose streams run concurrently) race.  Synchronous, so
 * not for stream capture. */
int chorus_diag_noalloc(chorus_ctx *ctx, double *residglob, double *divbglob, double *divbmax,
                        double *eigvmax, double *epsilon0);

/* Host-side topology helper for drivers that do not run the reference's own
 * connectivity chain: CONNECTIVITY + CONNECT_BDRY + MAPCYCLOC
 * (connectivity.f90:1-233, connect_bdry.f90:1-354, map_cyc_loc.f90:1-42) for
 * one rank.  btype[NBND]: 1 CYCL, 2 FREE, 0 other.  Output arrays are sized by
 * the caller: if2c >= 4*4*NCELL ints, ibfinter/ibfree >= 4*NCELL,
 * cyc2fpair >= 2*4*NCELL.  Counts are returned through nface..nfree. */
int chorus_topology(int ncell, int nvert, const int *ivcell, const double *xv, const double *yv,
                    int nbnd, const int *bnd, const int *btype, double dxcycl, double dycycl,
                    int *nface, int *if2c, int *ninter, int *ibfinter, int *npair,
                    int *cyc2fpair, int *nfree, int *ibfree);

/* Measurement hook (no reference counterpart): runs nsteps steps on the
 * context's stream bracketed by HIP events and returns the elapsed time of the
 * whole loop and the average duration of one stage-kernel launch, timed by
 * events recorded around every stage launch on that same stream. */
int chorus_time_steps(chorus_ctx *ctx, int nsteps, double *ms_total, double *ms_stage_avg,
                      double *ms_ct_avg);

/* Tecplot output, replacing tecplotter2dsetup (tecplotter2dsetup.f90:1-164),
 * tecplotter2d (tecplotter.f90:1-86) and tecpost_parallel (tecplotter.f90:
 * 92-418) for one partition.  chorus_plot_setup builds the plot-node topology
 * of NRE x NRE plot cells per cell from the same mesh as chorus_init; call it
 * before stepping: with ARTIV=1 the viscous stages then also keep muAV(1) and
 * the By indicator of their input (the muAV and etaAV columns hold the last
 * residual evaluation's values, as the reference's module array muAV does;
 * with ARTIV=0 they are zeros, where the reference reads an unallocated
 * array).  chorus_write_tecplot forms the node values of the current state on
 * the device and writes the TDV111 binary file at path (the reference writes
 * ./TEC/TEC<int(ctime):7><frac(ctime):4>.plt; the caller chooses the path)
 * with solution time ctime.  The plot set-up holds about NCELL (NRE+1)^2 x 9 + 13 x nnode
 * doubles of device memory (each cell's contributions, node values). */
int chorus_plot_setup(chorus_ctx *ctx, const chorus_mesh *mesh, int nre);
/* Host only (no device): the plot nodes of chorus_plot_setup.  nnode and
 * ncellp receive the node and plot-cell counts; xy (2 x (NRE+1)^2 NCELL
 * doubles, may be null) the node coordinates X then Y; c2n (4 NRE^2 NCELL
 * ints, may be null) the 1-based nodes of each plot cell. */
int chorus_plot_topology(const chorus_mesh *mesh, int nre, int *nnode, int *ncellp, double *xy, int *c2n);
int chorus_write_tecplot(chorus_ctx *ctx, const char *path, double ctime);

/* Restart dump, replacing WRITE_ALL_DATA_PARALLEL (write_data.f90:5-83) for one
 * partition: the current state is packed into the dump's record layout on the
 * device and written to path (the reference's name is restart.%08d.dat of
 * iter): int32 iter, FP64 ctime, then per cell an int32 global cell number and
 * Q(1:9,1:N,1:N) in Fortran order.  icg[NCELL] holds the global cell numbers
 * (IC2ICG); null writes 1..NCELL.  The file is truncated first (the reference's
 * MPI_MODE_CREATE leaves the tail of a longer existing file in place). */
int chorus_write_restart(chorus_ctx *ctx, const char *path, int iter, double ctime, const int *icg);

/* Tile-shape autotuner (no reference counterpart): times the mixed form's
 * cell kernels in every compiled tile on this context's own mesh and state
 * (residual evaluations, no state change), keeps the fastest for later steps
 * and returns its index in *tile.  A tile gives the shape of the stage kernel
 * and that of the gradient kernel, each one of three threads-per-block
 * targets (ideal/viscous kernels: 256/128, 128/64, 512/256): tiles 0-2 use
 * shape 0, 1 or 2 for both, tiles 3-8 the pairs (stage, gradient) = (0,1),
 * (0,2), (1,0), (1,2), (2,0), (2,1).  The FP64 form has one shape (*tile = 0). */
int chorus_autotune(chorus_ctx *ctx, int *tile);

/* Fixes the tile shape of the mixed form's cell kernels (an index as returned
 * by chorus_autotune), e.g. to reproduce a measurement without re-tuning.
 * CHORUS_ERR_ARG for an index out of range or a nonzero index in the FP64 form. */
int chorus_set_tile(chorus_ctx *ctx, int tile);

/* ---- Batched ensembles (no reference counterpart: the reference runs one
 * mesh per MPI job; the loop of main.f90:44-50 / iterations.f90:18-70 is what
 * each instance executes).  ninst independent instances, each with its own
 * mesh, state and parameters.  Instances whose N, precision, k_stage, vismode,
 * artiv, mu, eta0, gam and dt agree share one group: their cells are
 * concatenated, so every launch of a stage covers the whole group, with
 * per-instance AV quantities (wave-speed maximum, epsilon0, length scale).
 * All instances must use the same k_stage.  One time step of the whole batch
 * is one hipGraph launch (captured on first use, one graph per buffer
 * rotation). */
typedef struct chorus_batch chorus_batch;

int chorus_batch_init(int ninst, const chorus_mesh *meshes, const chorus_params *params, chorus_batch **out);
void chorus_batch_free(chorus_batch *b);
const char *chorus_batch_last_error(const chorus_batch *b);
/* Q(9,N,N,NCELL) of instance i (reference layout), as chorus_set_state / chorus_get_state */
int chorus_batch_set_state(chorus_batch *b, int i, const double *q);
int chorus_batch_get_state(chorus_batch *b, int i, double *q);
int chorus_batch_step(chorus_batch *b, int nsteps);
/* residual of instance i's current state (compresid.f90:1-67), as chorus_compresid */
int chorus_batch_compresid(chorus_batch *b, int i, double *resid);
/* restart dump of instance i (write_data.f90:5-83), as chorus_write_restart */
int chorus_batch_write_restart(chorus_batch *b, int i, const char *path, int iter, double ctime, const int *icg);
/* Tecplot output of instance i (tecplotter2dsetup / tecplotter2d / tecpost_parallel),
 * as chorus_plot_setup / chorus_write_tecplot: mesh is the instance's own (the
 * one passed to chorus_batch_init); set up before stepping for the muAV columns */
int chorus_batch_plot_setup(chorus_batch *b, int i, const chorus_mesh *mesh, int nre);
int chorus_batch_write_tecplot(chorus_batch *b, int i, const char *path, double ctime);
/* correct_B_via_curl_Az (setinitialcond.f90:82-92) of every instance */
int chorus_batch_correct_b(chorus_batch *b);
/* max |div B| per instance (iterations.f90:35-48), divbmax[ninst] */
int chorus_batch_divb(chorus_batch *b, double *divbmax);
/* groups: ngroups, and for group g its N, member count, cells and tile shape */
int chorus_batch_ngroups(const chorus_batch *b);
int chorus_batch_group(const chorus_batch *b, int g, int *N, int *ninst, int *ncell, int *tile);
/* Tile-shape autotuner of the mixed-form cell kernels: per group, times every
 * compiled tile shape on the group's own cells (residual evaluations, no state
 * change) and keeps the fastest; tiles[ngroups] receives the choice. */
int chorus_batch_autotune(chorus_batch *b, int *tiles);
/* Measurement hook: nsteps steps bracketed by HIP events on the batch stream. */
int chorus_batch_time_steps(chorus_batch *b, int nsteps, double *ms_total);

#ifdef __cplusplus
}
#endif
#endif
