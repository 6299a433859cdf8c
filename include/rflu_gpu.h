/*
 * rflu_gpu.h -- C ABI of the MI355X (gfx950) RocfluMP explicit residual path.
 *
 * One handle = one Rocflu region resident in HBM.  Every entry point returns
 * 0 on success and a non-zero RFLU_GPU_E* code on failure; the message is in
 * rflu_gpu_last_error(h) (h may be NULL for errors raised before a handle
 * exists).  No entry point falls back to the CPU: without a usable GPU,
 * rflu_gpu_create fails with RFLU_GPU_ENODEV.
 *
 * Host arrays use plain pointers and sizes.  Cell arrays are SoA with the
 * variable index slowest: cv[k*nCells + i] for k = rho, rho u, rho v, rho w,
 * rho E; dv[k*nCells + i] for k = p, T, c.  Gradients are returned in the
 * reference's gradCell(3,5,N) order (grad[(i*5+k)*3+d]).
 *
 * Each call below names the reference routine it replaces (file:line in the
 * Rocflu sources).  The handle's stream serialises all work; calls that
 * return data synchronise on it.
 */
#ifndef RFLU_GPU_H
#define RFLU_GPU_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define RFLU_GPU_OK 0
#define RFLU_GPU_EINVAL 1     /* bad argument / unsupported option */
#define RFLU_GPU_ENODEV 2     /* no HIP device */
#define RFLU_GPU_EHIP 3       /* HIP runtime error */
#define RFLU_GPU_EMESH 4      /* mesh pipeline rejected the grid */
#define RFLU_GPU_ENAN 5       /* RFLU_CheckValidity: NaN in the state */
#define RFLU_GPU_ENEGATIVE 6  /* RFLU_CheckPositivity: rho or p <= 0 */

/* spaceDiscr values (modfloflu/ModParameters.F90:663-666) */
#define RFLU_DISCR_ROE 1
#define RFLU_DISCR_HLLC 3
#define RFLU_DISCR_AUSM 4
#define RFLU_DISCR_AUSMUP 5 /* AUSM+up (modflu/RFLU_ModAUSMPlusUpFlux.F90) */
/* reconst values (modfloflu/ModParameters.F90:682) */
#define RFLU_RECONST_NONE 0
#define RFLU_RECONST_WENO_XYZ 2 /* WENO-XYZ average + positivity (the deck default) */
#define RFLU_RECONST_VENKAT 11
/* boundary kinds on the path (rocflu/RFLU_CentralFirstPatch.F90:575,790) */
/* Patch kinds: BC_SLIPWALL, BC_OUTFLOW, BC_NOSLIPWALL_HFLUX/_TEMP (the same
 * inviscid flux), BC_FARFIELD, BC_INJECTION, BC_INFLOW_VELTEMP and
 * BC_INFLOW_TOTANG of modfloflu/ModParameters.F90:397-434 (first- and
 * second-order patch fluxes: rocflu/RFLU_CentralFirstPatch.F90:230-1134,
 * RFLU_CentralSecondPatch.F90:241-1351). */
#define RFLU_BC_SLIPWALL 1
#define RFLU_BC_OUTFLOW 2
#define RFLU_BC_NOSLIPWALL 3
#define RFLU_BC_FARFIELD 4
#define RFLU_BC_INJECTION 5
#define RFLU_BC_INFLOW_VELTEMP 6
#define RFLU_BC_INFLOW_TOTANG 7
/* Values per patch in rflu_mesh_desc.bcVals (RFLU_BCVALS each, uniform over
 * the patch; the reference's pPatch%mixt%vals with distrib = 0):
 *   FARFIELD        mach, angle of attack, slip angle (radians), p, T
 *   INJECTION       mass flux per area (> 0: burning surface), T
 *   INFLOW_VELTEMP  u, v, w, T, p (p unless bcOpt is subsonic)
 *   INFLOW_TOTANG   p_tot, T_tot, beta_h, beta_v (radians), mach, fixed flag */
#define RFLU_BCVALS 6
#define RFLU_BCOPT_SUPERSONIC 0
#define RFLU_BCOPT_SUBSONIC 1
#define RFLU_BCOPT_MIXED 2

typedef struct rflu_handle rflu_handle;

/* Grid as the reference reads it (RFLU_ReadGridWrapper): node coordinates,
 * cell-to-vertex lists and per-patch boundary-face vertex lists. */
typedef struct rflu_mesh_desc {
  int nVert;
  const double *xyz;      /* 3*nVert, (x,y,z) per node */
  /* 4: all tets.  8: hexes, or mixed cells: each row holds a tet (4), pyramid
   * (5), prism (6) or hex (8) in Rocflu's local vertex order, padded with -1
   * (tet2v/pyr2v/pri2v/hex2v); rows in Rocflu's global order (tets, hexes,
   * prisms, pyramids) reproduce its face numbering. */
  int nvc;
  int nCells;
  const int *c2v;         /* nvc*nCells, 0-based */
  int nPatches;
  const int *bcType;      /* RFLU_BC_* per patch */
  const int *bcOpt;       /* RFLU_BCOPT_* per patch (outflow) */
  const double *pout;     /* outflow static pressure per patch */
  const int *nBFaces;     /* faces per patch */
  int nvf;                /* vertices per boundary face (3 or 4) */
  const int *bf2v;        /* nvf per face, patches concatenated; with nvf 4 a
                             triangle is padded with -1 (bTri2v before bQuad2v
                             within a patch, as Rocflu orders them) */
  const double *bcVals;   /* RFLU_BCVALS per patch (NULL: all zero) */
} rflu_mesh_desc;

/* The input-deck values the path reads (NUMERICS, TIMESTEP, REFERENCE,
 * ACCELERATION sections; libfloflu/Read*Section.F90). */
typedef struct rflu_params {
  int spaceDiscr;         /* RFLU_DISCR_* */
  int spaceOrder;         /* 1 or 2 */
  int reconst;            /* RFLU_RECONST_* (order 2) */
  int patchOrder;         /* ORDERBF: 1 (default) or 2 */
  double cfl;             /* capped at 2 for RK3 (RFLU_DerivedInputValues.F90:395) */
  double dtImposed;       /* TIMESTEP; used while time <= 1e-8 */
  double dtMinLimit;      /* DTMINLIMIT */
  double dTimeSystem;     /* run length for the dt clip (<= 0: none) */
  double refCp;
  double refGamma;
  double epsentr;         /* ENTROPY (Roe) */
  double dissFact;        /* DISSFACT (Roe) */
  int accelOn;
  double accel[3];
  /* 0 (reference): Venkatakrishnan term = numer/(denom + 1e-12)
   * (modflu/RFLU_ModLimiters.F90:459-505), which flips and amplifies the
   * gradient when denom lies in (-1e-12, 0).  1: add 1e-12 with the sign of
   * denom (a deliberate deviation, off by default). */
  int limiterGuard;
  /* 0 (reference): FP64 throughout.  1: gradients (LSQ + limiter) and
   * interior second-order face fluxes evaluated in FP32 (each face flux is
   * the FP64 physical flux of the left cell's state plus an FP32 difference
   * of the scheme's flux to it), with FP64 kept for the residual
   * accumulation, the conserved-state RK update, p/T recovery, boundary
   * patches and the dt reduction.  Needs spaceOrder 2, no particles.  No
   * reference counterpart: an added precision mode. */
  int fluxPrecision;
  /* AUSM+up constants (NUMERICS SIGMA, KP, KU, MINF; the reference's
   * defaults are 1, 0.25, 0.75, 0.1, libfloflu/ReadNumericsSection.F90:236-275).
   * The TCP flux function sets f_a = 1, so mInf has no effect there. */
  double sigma, kp, ku, mInf;
} rflu_params;

/* RFLU_InitFlowSolver grid/geometry/stencil set-up
 * (rocflu/RFLU_InitFlowSolver.F90:606-607,769-770,1008-1068). */
int rflu_gpu_create(const rflu_mesh_desc *mesh, const rflu_params *prm, rflu_handle **out);
int rflu_gpu_destroy(rflu_handle *h);
const char *rflu_gpu_last_error(const rflu_handle *h);

int rflu_gpu_sizes(const rflu_handle *h, int *nCells, int *nFaces, int *nBFaces, int *nStencil);
/* Mesh products of the set-up (for parity checks); any pointer may be NULL.
 * f2c[2F], fn[4(F+B)], fc[3(F+B)] (interior then boundary faces), bf2c[B],
 * vol[N], cofg[3N], stencil CSR off[N+1], memb[nStencil], moms[10N]. */
int rflu_gpu_get_mesh(const rflu_handle *h, int *f2c, double *fn, double *fc, int *bf2c,
                      double *vol, double *cofg);
int rflu_gpu_get_stencils(const rflu_handle *h, int *off, int *memb, double *moms);

/* Upload cv and set gv/dv (RFLU_SetVarsWrapper at init). */
int rflu_gpu_set_state(rflu_handle *h, const double *cv);
int rflu_gpu_get_state(rflu_handle *h, double *cv, double *dv);
int rflu_gpu_get_rhs(rflu_handle *h, double *rhs);
int rflu_gpu_get_grad(rflu_handle *h, double *grad);

/* rocflu/RFLU_TimeStepInviscid.F90:75-335: per-cell dt and its minimum.
 * dtCell may be NULL. */
int rflu_gpu_time_step_inviscid(rflu_handle *h, double *dtCell, double *dtMinLocal);
/* rocflu/RFLU_MinimumTimeStep.F90:78-365 (+ the run-length clip of
 * rocflu/RFLU_TimeStepping.F90:489-493) on the device-held time. */
int rflu_gpu_minimum_time_step(rflu_handle *h, double *dtMin);

/* libfloflu/RkInitMP.F90:74-185.  The device keeps cv untouched during a
 * stage (round trips happen in registers), so cvOld == cv: nothing to copy. */
int rflu_gpu_rk_init(rflu_handle *h, int istage);
/* libfloflu/CellGradientsMP.F90:75-527: LSQ gradients + limiter. */
int rflu_gpu_cell_gradients(rflu_handle *h);
/* libfloflu/ConvectiveFluxesMP.F90:73-215: rhs = -diss + face fluxes. */
int rflu_gpu_convective_fluxes(rflu_handle *h);
/* libfloflu/SourceTermsMP.F90:73-218: gravity source (its cons/prim round
 * trip only alters cv bits that RKUpdateMP overwrites). */
int rflu_gpu_source_terms(rflu_handle *h);
/* libfloflu/IntegrateSourceTermsMP.F90:75-154: a no-op in the reference. */
int rflu_gpu_integrate_source_terms(rflu_handle *h);
/* libfloflu/RkUpdateMP.F90:75-210 + RFLU_SetVarsContWrapper + AfterUpdateMP. */
int rflu_gpu_rk_update(rflu_handle *h, int istage);
/* One fused RK stage: gradients, fluxes, sources, update (same bits as the
 * five calls above in sequence). */
int rflu_gpu_rk_stage(rflu_handle *h, int istage);
/* nSteps iterations of the RFLU_TimeStepping loop body (dt, 3 RK stages,
 * time advance); useGraph = 1 replays one captured hipGraph per step. */
int rflu_gpu_step(rflu_handle *h, int nSteps, int useGraph);
int rflu_gpu_get_time(rflu_handle *h, double *currentTime, double *dtMin);
/* global%currentTime as a restart sets it from the flow file
 * (modflu/RFLU_ModReadWriteFlow.F90:1219-1224); time-since-restart is reset. */
int rflu_gpu_set_time(rflu_handle *h, double currentTime);
int rflu_gpu_sync(rflu_handle *h);

/* Asynchronous stepping for hosts that drive several regions at once (no
 * reference counterpart; the serving loop uses them).  rflu_gpu_step_async
 * enqueues nSteps graph replays on the handle's stream and returns at once;
 * rflu_gpu_query returns 1 when the stream is idle and 0 while work is
 * pending; rflu_gpu_wait blocks, then reports AfterUpdateMP's checks like
 * rflu_gpu_step.  rflu_gpu_device_bytes: device memory the handle holds;
 * rflu_gpu_mem_info: free and total device memory. */
int rflu_gpu_step_async(rflu_handle *h, int nSteps);
int rflu_gpu_query(rflu_handle *h);
int rflu_gpu_wait(rflu_handle *h);
int rflu_gpu_device_bytes(const rflu_handle *h, size_t *bytes);
int rflu_gpu_mem_info(size_t *freeBytes, size_t *totalBytes);

/* Batched ensemble (no reference counterpart: Rocflu runs one region per
 * rank).  nInst independent regions of one cell type under one rflu_params
 * (gas, scheme, CFL) become one handle: the union of their meshes, each
 * instance's cells contiguous on the device, so every kernel launch covers
 * all instances; each instance keeps its own dt (RFLU_MinimumTimeStep over
 * its own cells) and time.  Results equal those of each region run alone,
 * bit for bit.  Step it with rflu_gpu_step / rflu_gpu_step_async; the
 * routine-level dt calls reject batches.  States are per instance, in that
 * instance's own cell numbering (SoA, 5 x nCells). */
int rflu_batch_create(int nInst, const rflu_mesh_desc *meshes, const rflu_params *prm, rflu_handle **out);
int rflu_batch_sizes(const rflu_handle *h, int inst, int *nCells, int *nInst);
int rflu_batch_set_state(rflu_handle *h, int inst, const double *cv);
int rflu_batch_get_state(rflu_handle *h, int inst, double *cv, double *dv);
int rflu_batch_get_time(rflu_handle *h, int inst, double *currentTime, double *dtMin);

/* Autotuner (no reference counterpart): times every gradient kernel that
 * gives the same bits on this handle's own mesh (reps launches each) and
 * keeps the fastest for later stages; then, on a region without particles,
 * times the stage after the gradients as two kernels (face fluxes, then the
 * gather and RK update) against the fused kernel (below) and keeps the
 * faster (the state is saved and put back).  The captured step graph is
 * rebuilt.  Reports the gradient choice and its time per launch. */
int rflu_gpu_autotune(rflu_handle *h, int reps, int *gradVariant, double *msPerLaunch);

/* Fused stages (no reference counterpart): on = 1 runs stages 2 and 3 of
 * every rflu_gpu_step as the gradients plus one kernel in which each cell
 * evaluates its own faces' fluxes (every interior face twice) and updates,
 * so no per-face flux array passes through HBM.  Bit-identical to on = 0.
 * Regions without particles only (RFLU_GPU_EINVAL otherwise); the
 * routine-by-routine calls are unaffected. */
int rflu_gpu_set_fusion(rflu_handle *h, int on);
int rflu_gpu_get_fusion(const rflu_handle *h, int *on);

/* Mean kernel duration (ms) per launch of the named kernel over the last
 * rflu_gpu_step call with profiling enabled; for measurement only. */
int rflu_gpu_set_profiling(rflu_handle *h, int on);
int rflu_gpu_kernel_time(rflu_handle *h, const char *name, double *msPerLaunch, int *launches);
/* Streaming device-to-device copy of `bytes` (read + write counted), repeated
 * `reps` times; reports GB/s.  A calibration aid for HBM traffic counters. */
int rflu_gpu_bandwidth_probe(size_t bytes, int reps, double *gbps);

#ifdef __cplusplus
}
#endif
#endif
