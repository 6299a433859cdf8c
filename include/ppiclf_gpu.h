/*
 * ppiclf_gpu.h -- ppiclF's Fortran 77 entry points on the MI355X particle
 * engine, for a Rocflu host whose rocpicl/ calls ppiclF unchanged.
 *
 * The symbols carry gfortran's external names (lowercase + '_'; every
 * argument by reference), so rocpicl/PICL_TEMP_InitSolver.F90 and
 * PICL_TEMP_Runge.F90 link against this library in place of libppiclF.
 * As in ppiclF, there is one instance per process (ppiclF keeps its state in
 * COMMON blocks); here it lives on the GPU beside one Rocflu region, which
 * the host names first with ppiclf_gpu_bind together with the user-physics
 * switches that rocpicl sets in common /RFLU_ppiclF/
 * (PICL_TEMP_InitSolver.F90:139-219).
 *
 * Call order, as Rocflu's (PICL_TEMP_InitSolver.F90:275, 682-694;
 * PICL_TEMP_Runge.F90:863-896, 926-953):
 *   ppiclf_gpu_bind(region, params)
 *   ppiclf_solve_Initialize -> InitParticle -> InitBoxFilter (or
 *   InitGaussianFilter) -> InitNeighborBin -> ppiclf_comm_InitOverlapMesh
 *   per RK stage: InterpFieldUser (each field) -> IntegrateParticle ->
 *   GetProFldIJKEF (each node, element and field)
 *
 * InitOverlapMesh attaches the particles to the region exactly as
 * picl_gpu_init does (picl_gpu.h), including InitSolve's projection.
 * IntegrateParticle runs ppiclf_solve_IntegrateRK3s_Rocflu's stage (its stage
 * counter cycles 1, 2, 3 over calls, as the reference's SAVE counter does)
 * from the staged fields; the read-back into Rocflu's rhs and piclVF stays
 * with the host, which reads GetProFldIJKEF.
 *
 * Errors: ppiclF stops the run with "PPICLF: ERROR ..." (ppiclf_exittr,
 * ppiclf_op.f:288-318); so does this library, unless
 * ppiclf_gpu_set_abort(0) asks for errors to be recorded instead
 * (ppiclf_gpu_last_error) and the failing call to return without effect.
 */
#ifndef PPICLF_GPU_H
#define PPICLF_GPU_H
#include "picl_gpu.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Set-up of the default instance (no ppiclF counterpart): the Rocflu region
 * the particles live in (all-hex, spaceOrder 2, no particles yet) and the
 * user-physics switches.  projMode, filterWidth, filterAlpha and
 * periodicX/Y/Z are taken from the filter and Initialize calls instead. */
int ppiclf_gpu_bind(rflu_handle *h, const picl_params *pp);
int ppiclf_gpu_reset(void);
void ppiclf_gpu_set_abort(int on);
const char *ppiclf_gpu_last_error(void);
rflu_handle *ppiclf_gpu_handle(void);

/* ppiclf_solve_Initialize (libpicl/ppiclF/source/ppiclf_solve.f:1): linear
 * periodicity flags and bounds per axis (the bounds must be the grid's
 * vertex extremes, as PICL_TEMP_InitSolver.F90:234-262 computes them);
 * angular periodicity (ai1 = 1) is not built. */
void ppiclf_solve_initialize_(const int *xi1, const double *xpmin, const double *xpmax, const int *yi1,
                              const double *ypmin, const double *ypmax, const int *zi1, const double *zpmin,
                              const double *zpmax, const int *ai1, const double *apa, const double *apxa,
                              const double *aprin, const double *aprout);
/* ppiclf_solve_InitParticle (ppiclf_solve.f:184-188): imethod 2, ndim 3;
 * y(PPICLF_LRS=12, npart), rprop(PPICLF_LRP=48, npart) (JRHOP, JDP, JVOLP
 * and JSPL are read). */
void ppiclf_solve_initparticle_(const int *imethod, const int *ndim, const int *iendian, const int *npart,
                                const double *y, const double *rprop, const double *filt2, const double *filt3);
/* ppiclf_solve_InitBoxFilter (ppiclf_solve.f:1579): iwallm 0, sngl_elem 1. */
void ppiclf_solve_initboxfilter_(const double *filt, const int *iwallm, const int *sngl_elem);
/* ppiclf_solve_InitGaussianFilter (ppiclf_solve.f:1520): width, cut-off
 * alpha, iwallm 0. */
void ppiclf_solve_initgaussianfilter_(const double *filt, const double *alpha, const int *iwallm);
/* ppiclf_solve_InitNeighborBin (ppiclf_solve.f:339). */
void ppiclf_solve_initneighborbin_(const double *rwidth);
/* ppiclf_comm_InitOverlapMesh (ppiclf_comm.f:926): ncell hexes with
 * lx1 = ly1 = lz1 = 2 nodes, coordinates node-fastest; they must be the
 * bound region's grid. */
void ppiclf_comm_initoverlapmesh_(const int *ncell, const int *lx1, const int *ly1, const int *lz1,
                                  const double *xgrid, const double *ygrid, const double *zgrid);
/* ppiclf_solve_InterpFieldUser (ppiclf_solve.f:2020): infld(2,2,2,ncell)
 * for rprop index jp. */
void ppiclf_solve_interpfielduser_(const int *jp, const double *infld);
/* ppiclf_solve_IntegrateParticle (ppiclf_solve.f:1694): one RK stage. */
void ppiclf_solve_integrateparticle_(const int *istep, const int *iostep, const double *dt, const double *time);
/* ppiclf_solve_GetProFldIJKEF (ppiclf_solve.f:3630): ppiclf_pro_fld(i,j,k,e,m),
 * m = PPICLF_P_JPHIP, JFX, JFY, JFZ, JE (1..5). */
void ppiclf_solve_getprofldijkef_(const int *i, const int *j, const int *k, const int *e, const int *m, double *fld);

#ifdef __cplusplus
}
#endif
#endif
