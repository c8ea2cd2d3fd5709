/*
 * cles.h -- C ABI of the MI355X (gfx950) right-hand-side and time-advance path
 * for Chimera_Comp_LES.
 *
 * The entry points replace the reference's CPU call chain
 *   Flow_Solver -> RungeKutta -> formRHS / FormBCS / EqOfState
 *   and Compute_Smagorinsky_Constant_1 / Compute_eddy_viscosity_LES
 * with one context per grid block whose fields stay resident in HBM.
 *
 * Data exchanged with the host uses the reference's PETSc DMDA layout:
 *   - node order i fastest, then j, then k (DMDAVecGetArray [k][j][i]);
 *   - conserved variables interleaved {rho, rhoU, rhoV, rhoW, rhoE}
 *     (CompVars, variables.h:22-24); coordinates and metric triples {x,y,z}
 *     (Cmpnts, variables.h:18-20);
 *   - layout CLES_LAYOUT_GLOBAL: the mx*my*mz owned nodes (a global Vec);
 *     layout CLES_LAYOUT_LOCAL: the ghosted local Vec of a single-rank DMDA
 *     with stencil width 2 (ghosts only in periodic directions,
 *     init.c:156-228), so a PETSc host may pass DMDAVecGetArray base pointers.
 *
 * Conventions: every function returns 0 on success and a negative CLES_E*
 * code on failure (the reference returns PetscErrorCode 0); no function
 * calls exit().  All calls are synchronous on return.  A context is not
 * thread-safe; use one context per host thread.
 */
#ifndef CLES_H
#define CLES_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLES_OK 0
#define CLES_EINVAL (-1)     /* bad argument / unsupported option combination */
#define CLES_ENOMEM (-2)     /* device allocation failed */
#define CLES_EDEVICE (-3)    /* HIP runtime error */
#define CLES_ESTATE (-4)     /* call out of sequence (e.g. no state uploaded) */

#define CLES_LAYOUT_GLOBAL 0
#define CLES_LAYOUT_LOCAL 1

/* Arithmetic of the face reconstruction and flux kernels. */
#define CLES_PREC_FP64 0       /* plain FP64 drop-in */
#define CLES_PREC_FP32BULK 1   /* FP32 reconstruction/flux, FP64 metrics, RK accumulation and EOS */

/* Run options.  Names and defaults follow the reference's PETSc options
 * (init.c:85-96, main.c:40-72, rhs.c:2581-2585, les.c:151). */
typedef struct cles_params {
  int mx, my, mz;        /* IM JM KM of grid.dat (nodes incl. boundary)      */
  int bctype[6];         /* bcs.dat: i-, i+, j-, j+, k-, k+ (rhs.c:456-475)  */
  double ren;            /* -ren   Reynolds number                            */
  double Ma;             /* -Ma    Mach number                                */
  double Pr;             /* -Pr    Prandtl number                             */
  double gamma;          /* -gamma ratio of specific heats                    */
  double dt;             /* -dt    time step                                  */
  double st;             /* user->st (1)                                      */
  double Tref;           /* -Tref  Sutherland reference temperature           */
  double max_cs;         /* -max_cs clip of the dynamic Smagorinsky constant  */
  double Twall;          /* -Twall isothermal wall temperature                */
  double s_x, s_y, s_z;  /* -s_x -s_y -s_z hybrid-flux weights                */
  double s_x_min, s_y_min, s_z_min;
  double sigma_thrsh;    /* -sigma_thrsh                                      */
  double ratio;          /* channel forcing ratio (0 without -channel)        */
  int les;               /* -les: 0 off, 1 constant Cs, >=2 dynamic           */
  int invicid;           /* -inv                                              */
  int fluxorder;         /* -fluxorder: 1 Godunov, 2 MUSCL(+RK2), 3 WENO(+RK3)*/
  int eno_order;         /* -ENOorder                                         */
  int central;           /* -central: 2 skew-symmetric 4-point (default)      */
  int sutherland;        /* -visT                                             */
  int isothermal;        /* -isothermal                                       */
  int calc_sigma;        /* -calc_sigma: shock sensor of the hybrid LES flux  */
  int abc;               /* -ABC                                              */
  int twoD;              /* -twoD: 0, or 1/2/3 = drop x/y/z (1: no xi fluxes,
                          * u = 0, no rhoU residual, rhs.c:1349, 1487, 2135,
                          * 2646, 3168, 3465, 4215; all: IB normal and
                          * surface velocity, ibm_iso.c:2662, 2744)          */
  int testfilter_ik;     /* -testfilter_ik                                    */
  int tistart;           /* -rstart value (0 = fresh start); the time loop
                            sees it incremented (main.c:890-891)            */
  int rstart;            /* restart flag                                      */
  int precision;         /* CLES_PREC_*                                       */
  int averaging;         /* -averaging (main.c:779): Do_averaging after steps */
  int i_homo_filter;     /* -i_homo_filter -j_homo_filter -k_homo_filter      */
  int j_homo_filter;     /* (main.c:791-793): Cs from LM/MM averaged over the  */
  int k_homo_filter;     /* homogeneous directions (les.c:934-1116)           */
  /* Build-only schemes with no reference counterpart (FP64 form only):   */
  int weno5;             /* WENO5 (Jiang-Shu) where the reference uses WENO3  */
  int roe;               /* Roe flux (Harten entropy fix) for the Rusanov flux */
} cles_params;

typedef struct cles_ctx cles_ctx;

/* Fill p with the reference defaults for an mx*my*mz block. */
void cles_default_params(cles_params *p, int mx, int my, int mz);

/* Create a context on the current HIP device, upload the node coordinates
 * (global layout, 3 per node) and form the metrics on the device.
 * Replaces MG_Initial's allocation (init.c:956-1138) followed by
 * FormMetrics (metrics.c:3-782).  *err receives a CLES_E* code. */
cles_ctx *cles_create(const cles_params *p, const double *coor, int *err);
void cles_destroy(cles_ctx *c);

/* Immersed-boundary node flags (Nvert, one per node, global layout):
 * 0 fluid, 1 IB node, 3 solid (ibm_iso.c:762-769).  Default: all fluid. */
int cles_set_nvert(cles_ctx *c, const double *nvert);
/* Download Nvert (one per node, global layout). */
int cles_get_nvert(cles_ctx *c, double *nvert);

/* One interpolation node of the immersed body (IBMInfo, variables.h:46-89),
 * in the order the reference's list visits them. */
typedef struct cles_ibm_info {
  int ni, nj, nk;        /* the node                                       */
  int cell;              /* nearest surface triangle                       */
  int imode;             /* donor triangle of the 3x3 face (0..7), or -100 */
  int i1, j1, k1, i2, j2, k2, i3, j3, k3; /* donor nodes                    */
  double pmin[3];        /* nearest surface point                          */
  double d_s, d_i;       /* node-surface, node-donor-point distances       */
  double cs1, cs2, cs3;  /* barycentric weights of pmin in the triangle    */
  double cr1, cr2, cr3;  /* weights of the donor point in the donor nodes  */
} cles_ibm_info;

/* Classify the nodes against one closed triangulated surface and build the
 * interpolation stencils: ibm_search_advanced (ibm_iso.c:87-795) for body 0
 * with the surface as ibm_read_ucd leaves it (ibm_io.c:1530-1760).
 * x_bp,y_bp,z_bp: n_v vertices; nv1,nv2,nv3: n_elmt triangles (0-based,
 * outward normal by the right-hand rule); nvert_o: the previous Nvert for the
 * phase-change test, or NULL.  Nvert is reset (values below 5) and replaced.
 * With -isothermal 1 or 2 the boundary nodes get interpolation stencils; from
 * then on every RungeKutta ends with ibm_interpolation_advanced and
 * FormBCS(0) (solvers.c:279-286).  *n_ib (may be NULL) receives the number
 * of interpolation nodes.  Fails with CLES_EINVAL when a ray needs more than
 * 1024 directions or crosses more than 48 triangles, or a boundary node has
 * no stencil. */
int cles_ibm_search(cles_ctx *c, int n_v, const double *x_bp, const double *y_bp, const double *z_bp,
                    int n_elmt, const int *nv1, const int *nv2, const int *nv3, const double *nvert_o,
                    int *n_ib);
/* Surface velocity now and one step ago (ibm->u, ibm->uold), 3 per vertex.
 * Zero after cles_ibm_search (a body at rest). */
int cles_ibm_set_velocity(cles_ctx *c, const double *u, const double *uold);
/* Copy the n interpolation nodes (n must equal the count) to out. */
int cles_ibm_get_list(cles_ctx *c, cles_ibm_info *out, int n);
/* Surface forces of the body, Calc_forces_SI2 (fsi.c:1312-2470) as
 * Struc_Solver calls it at the start of a step (solvers.c:124): pressure at
 * the interpolation nodes and the viscous stress of the last CalcUCat's
 * velocity on their fluid-side faces, moments about center (fsi->x_c..z_c).
 * out: F x,y,z; Cp x,y,z (pressure part); A x,y,z (projected areas); A_tot;
 * M x,y,z; Mdpdn x,y,z; Ia x,y,z. */
#define CLES_FORCE_N 19
int cles_ibm_forces(cles_ctx *c, const double center[3], double out[CLES_FORCE_N]);
/* The same outputs, computed without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc), whereas
 * cles_ibm_forces allocates its work buffers at its first call on a body.
 * The per-node records live in a static device array that each GPU holds
 * (about 7 MB) until the process ends, shared by every call on that GPU:
 * calls that overlap on one GPU race.  Bodies with more than
 * CLES_FORCE_NOALLOC_MAX_NODES interpolation nodes are rejected with
 * CLES_EINVAL. */
#define CLES_FORCE_NOALLOC_MAX_NODES 16384
int cles_ibm_forces_noalloc(cles_ctx *c, const double center[3], double out[CLES_FORCE_N]);

/* Upload the conserved state; sets Q, Qold and the pressure by EqOfState.
 * Replaces the tail of SetInitialCondition (rhs.c:424-433). */
int cles_set_state(cles_ctx *c, const double *Q, int layout);
/* Upload Q and P as given (no EqOfState); Q becomes Qold too.  The state
 * RestartApplication (main.c:627-650) sets from the field files. */
int cles_set_state_qp(cles_ctx *c, const double *Q, const double *P, int layout);
/* Download Q (5 per node) and/or P (1 per node); either may be NULL. */
int cles_get_state(cles_ctx *c, double *Q, double *P, int layout);

/* Binary field files of the reference (PETSc binary Vec format, single-rank
 * natural ordering, 32-bit PetscInt; see cles_io.hip).
 * cles_write_fields: Q_P_Binary_Output (main.c:246-300) for the current time
 * step: Qfield%05d_%1d.dat (+ .info), with LES sigma (+ .info), pfield,
 * nvfield, with -averaging (ti != 0) Qfield_ave, su2_Sq, su_cross, psum,
 * with LES nu_t; dir may be NULL (current directory).
 * cles_read_fields: Q_P_Binary_Input (main.c:215-243) + RestartApplication
 * (main.c:627-650): reads pfield, nvfield, Qfield of step ti, sets Nvert, Q,
 * Qold = Q, P as read and the time step to ti.  An immersed body is
 * classified again afterwards (cles_ibm_search), as main.c:870-877 does. */
int cles_write_fields(cles_ctx *c, const char *dir, int block);
/* VTKOut of the post-processor (data.c:385-790): dir/Result%2.2d_%6.6d.vts
 * (block, time step), an XML StructuredGrid with raw appended little-endian
 * Float64 arrays Position, nvert, p, rho, E, u and with LES Sigma, nu_T, each
 * after a 4-byte byte count (so each array stays under 2 GiB); -twoD takes
 * one x/y/z node as the reference's writer does. */
int cles_write_vtk(cles_ctx *c, const char *dir, int block);
/* VTKOut's options (data.c:423-429): -skip (keep every skip-th x node, at
 * least 1), -theta (rotation of block 2's positions and velocities about z,
 * radians), -Pres (no file: surface-pressure lines) and -airfoil (with
 * -Pres: the airfoil's lines too; without it no effect). */
typedef struct cles_vtk_opts {
  int skip;
  double theta;
  int pres, airfoil;
} cles_vtk_opts;
/* VTKOut with its options (data.c:385-855).  Without pres: the file of
 * cles_write_vtk with mx / skip x nodes (node i from x node i skip), and for
 * block 2 positions and velocities rotated by theta; *tke (if not NULL) the
 * "TKE" sum the reference prints, sum of (u^2 + v^3 + w^2) / nodes over the
 * written nodes.  With pres: no file; the text the reference prints
 * (" PL \n ", then for block 2 along x node 20 and then with airfoil along
 * x node 8: the fluid nodes whose next node along z (block 2) or y (airfoil)
 * is solid, " PU \n ", those whose previous one is) goes to text
 * (NUL-terminated, at most text_len bytes with the NUL; *text_used its full
 * length); CLES_EINVAL when that x node is not on the grid. */
int cles_write_vtk_opts(cles_ctx *c, const char *dir, int block, const cles_vtk_opts *o, double *tke,
                        char *text, size_t text_len, size_t *text_used);
/* QCriteria (data.c:4815-5053) of the post-processor's -qcr branch on the
 * current state: q (the Q criterion, (|W|^2 - |S|^2) / 2 of the velocity
 * gradient) and lambda2 (the middle of the reference's Jacobi eigenvalue
 * estimates of its S^2 + W^2 products), each mx*my*mz in the global layout.
 * Nodes 1..m-2 of each direction are computed; on the boundary nodes q is
 * the pressure (the reference writes Q into P there only) and lambda2 0. */
int cles_q_criteria(cles_ctx *c, double *q, double *lambda2);
/* VTKOutQ (data.c:1562-1711): dir/QCriteria%2.2d_%5.5d.vtk (block, step), a
 * legacy ASCII structured grid of the nodes 0..m-2 with q, l2 and nvert:
 * 8-node cell means as POINT_DATA (vc 0), or the values of nodes 1..m-2 as
 * CELL_DATA (vc 1, the reference's -vc). */
int cles_write_q_vtk(cles_ctx *c, const char *dir, int block, int vc);
int cles_read_fields(cles_ctx *c, const char *dir, int ti, int block);
/* One PETSc binary Vec file: class id 1211214 and length as big-endian
 * int32, then n big-endian doubles; block_size > 1 also writes <path>.info
 * ("-vecload_block_size"). Host only; no context or device needed. */
int cles_petsc_vec_write(const char *path, const double *x, long long n, int block_size);
/* Read one PETSc binary Vec file of exactly n values (CLES_EINVAL otherwise). */
int cles_petsc_vec_read(const char *path, double *x, long long n);
/* Metrics as formed by FormMetrics: csi, eta, zet (3 per node), aj. */
int cles_get_metrics(cles_ctx *c, double *csi, double *eta, double *zet, double *aj);
/* LES fields: Cs (CS), nu_t (Nu_t), sigma (Sigma, 3 per node); any may be NULL. */
int cles_get_les(cles_ctx *c, double *cs, double *nu_t, double *sigma);

/* Running sums of Do_averaging (les.c:1664-1747), kept with -averaging and
 * advanced after every cles_flow_step / cles_advance step: Q_sum (5 per
 * node), P_sum, the squares (u^2, v^2, w^2) and cross products (uv, vw, wu)
 * of the velocity of the last CalcUCat; any pointer may be NULL. */
int cles_get_averages(cles_ctx *c, double *q_sum, double *p_sum, double *u2_sum, double *ucross_sum);

/* EqOfState, rhs.c:4820: P from Q on every node. */
int cles_eq_of_state(cles_ctx *c);
/* formRHS, rhs.c:4703: Rhs = aj (Visc - Conv) on interior fluid nodes, 0
 * elsewhere; istage selects the sigma refresh of formConvectionLES
 * (rhs.c:3327).  rhs (5 per node, global layout) may be NULL. */
int cles_form_rhs(cles_ctx *c, int istage, double *rhs);
/* RungeKutta, solvers.c:223: fluxorder stages of formRHS, stage update,
 * FormBCS, EqOfState; with an immersed body then the interpolation and
 * FormBCS(0). */
int cles_runge_kutta(cles_ctx *c);
/* Compute_Smagorinsky_Constant_1 (les.c:135) followed by
 * Compute_eddy_viscosity_LES (les.c:1477). */
int cles_les_update(cles_ctx *c);
/* One time step of the reference loop (main.c:900-974 with Flow_Solver,
 * solvers.c:13-64): ti += 1, LES model, RungeKutta, Qold = Q. */
int cles_flow_step(cles_ctx *c);
/* nsteps calls of cles_flow_step without host synchronisation in between. */
int cles_advance(cles_ctx *c, int nsteps);

/* Inf-norms of the last stage: out[0] = |Rhs|inf, out[1] = |Q|inf
 * (VecNorm NORM_INFINITY, solvers.c:251-252). */
int cles_norms(cles_ctx *c, double out[2]);
int cles_get_time(cles_ctx *c);
int cles_set_time(cles_ctx *c, int ti);

/* Kernel timing for measurement: runs the fused stage kernel nrep times
 * back to back on the context's stream (a stage-1 update into a scratch
 * buffer, so the state is unchanged; its input is the last stage's input
 * once a step has run, else the current state, with the current state as
 * Qold) between two HIP events and returns the mean duration per launch in
 * milliseconds. */
int cles_time_stage_kernel(cles_ctx *c, int nrep, double *ms);
/* Same for one LES model update (all of its kernels); the LES fields are
 * recomputed from the current state, nothing else changes. */
int cles_time_les(cles_ctx *c, int nrep, double *ms);

/* Domain integrals of the current state over the unique interior nodes,
 * volume weighted (sum of Q/aj): out = {mass, x-, y-, z-momentum, total
 * energy, kinetic energy}.  No counterpart in the reference: it serves the
 * conservation tests and the kinetic-energy dissipation rate of the
 * Taylor-Green runs.  Deterministic summation order. */
int cles_integrals(cles_ctx *c, double out[6]);

/* ---------------------------------------------------------------------------
 * Batched serving.  No counterpart in the reference (its time loop,
 * solvers.c:223-294, runs one grid): several independent periodic solves
 * (any size, Mach, Reynolds number, precision, with or without the LES model
 * and the shock sensor; no immersed body, no statistics; RK2 or RK3, one
 * fluxorder per batch) are advanced together, one step captured as a
 * hipGraph: the LES model and sensor of each member, then either batched
 * stage and seam launches (default) or one persistent kernel per
 * (precision, model) group that runs every stage of every member, its
 * workgroups pulling tile and seam items from per-XCD queues and waiting on
 * per-member stage counters (cles_batch_set_mode).  Members are contexts
 * with their state uploaded; each member's result equals what cles_advance
 * gives for it alone.  From cles_batch_launch to cles_batch_wait every call
 * on a member returns CLES_ESTATE; a call that changes a member makes the
 * next launch rebuild
 * the batched step; cles_destroy of a member removes it from its batch.
 * ------------------------------------------------------------------------ */
typedef struct cles_batch cles_batch;
/* Device bytes cles_create allocates for p (the admission footprint). */
size_t cles_footprint(const cles_params *p);
/* A batch that admits members while their footprints fit hbm_budget bytes. */
cles_batch *cles_batch_create(size_t hbm_budget, int *err);
void cles_batch_destroy(cles_batch *b);
/* Admit / release a member; CLES_ENOMEM when c does not fit the budget. */
int cles_batch_add(cles_batch *b, cles_ctx *c);
int cles_batch_remove(cles_batch *b, cles_ctx *c);
int cles_batch_size(const cles_batch *b);
size_t cles_batch_used(const cles_batch *b);
/* Stage-kernel tile shape of the batched launches: 0 32x4x2, 1 16x4x4,
 * 2 16x8x2, 3 64x2x2, 4 6x6x6, 5 8x5x5 (256-thread workgroups), 6 4x4x3
 * (64 threads), 7 8x4x3 (128 threads); chosen by the autotuner of the
 * serving loop. */
int cles_batch_set_tile(cles_batch *b, int tile);
/* Stages of the batched step: one batched launch per group and stage plus
 * one seam launch per stage (default, the faster form on gfx950), or one
 * persistent kernel per group. */
#define CLES_BATCH_PERSISTENT 0
#define CLES_BATCH_LAUNCHES 1
int cles_batch_set_mode(cles_batch *b, int mode);
/* Polls of a stage dependency wait of the persistent kernel before it gives
 * up (default 4194304, about a second): the kernel then stops and the next
 * cles_batch_wait / cles_batch_advance returns CLES_EDEVICE, the member
 * states undefined. */
int cles_batch_set_spin_limit(cles_batch *b, unsigned polls);
/* Replay the batched step as a captured hipGraph (on, the default) or launch
 * every step's kernels on the batch stream (off: no capture and instantiation
 * after a membership change, which pays where members come and go often). */
int cles_batch_set_graph(cles_batch *b, int on);
/* Advance every member by nsteps time steps (one graph replay per step) and
 * wait for them. */
int cles_batch_advance(cles_batch *b, int nsteps);
/* The same, returning once the steps are queued; the host may then set up or
 * download other contexts.  Members must not be added, removed or read
 * before cles_batch_wait. */
int cles_batch_launch(cles_batch *b, int nsteps);
int cles_batch_wait(cles_batch *b);
/* cles_batch_advance(b, nrep) timed with HIP events: mean ms per step. */
int cles_batch_time_step(cles_batch *b, int nrep, double *ms);
const char *cles_batch_last_error(const cles_batch *b);
/* Stage-kernel tile shape of a single context (same codes as above). */
int cles_set_tile(cles_ctx *c, int tile);

/* Device bytes held by the context. */
size_t cles_device_bytes(const cles_ctx *c);
/* Last error message of the context (static storage when c is NULL). */
const char *cles_last_error(const cles_ctx *c);

#ifdef __cplusplus
}
#endif
#endif
