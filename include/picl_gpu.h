/*
 * picl_gpu.h -- C ABI of the MI355X ppiclF coupling (particles in a Rocflu
 * region).
 *
 * The particle cloud lives in HBM, attached to a Rocflu handle
 * (rflu_gpu.h).  Once picl_gpu_init succeeds, every RK stage of that handle
 * (rflu_gpu_rk_stage, rflu_gpu_step) runs the coupled stage of
 * libfloflu/RungeKuttaMP.F90:219-280 with PICL_TEMP_Runge between
 * SourceTermsMP and RKUpdateMP; picl_gpu_runge is that call on its own, for
 * hosts that drive the stage routine by routine.
 *
 * Physics restated (rocpicl/, libpicl/): the quasi-steady drag laws of
 * ppiclf_user_QuasiSteady.f (Parmar, Osnes, modified Parmar, Gidaspow) plus
 * Schiller-Naumann, AM Parmar added mass, the pressure-gradient and
 * viscous-unsteady forces, normal collisions and heat transfer on an Euler
 * flow, Sutherland or fixed viscosity, IDW-27 or trilinear interpolation,
 * single-element box or Gaussian-filter projection, linear periodicity,
 * two-way feedback.  Requirements: all-hex grid, spaceOrder 2, patchOrder 1.
 *
 * Particle arrays use ppiclF's layout: y(PPICLF_LRS=12, n) -> y[12*p + j]
 * (x, y, z, vx, vy, vz, T, ox, oy, oz, metal, oxide).  Element ids are Rocflu
 * cell indices (0-based, the host's numbering); -1 marks a particle that lies
 * in no element (ppiclF removes such particles; here they stay in place and
 * are skipped).
 */
#ifndef PICL_GPU_H
#define PICL_GPU_H
#include "rflu_gpu.h"
#ifdef __cplusplus
extern "C" {
#endif

/* The ppiclF user-physics switches the path reads
 * (rocpicl/PICL_TEMP_InitSolver.F90:139-219, common /RFLU_ppiclF/). */
typedef struct picl_params {
  int rmuFlag;       /* 0: fixed viscosity rmuRef; 1: Sutherland */
  double rmuRef;     /* reference viscosity */
  double tref;       /* Sutherland reference temperature */
  double suth;       /* Sutherland constant */
  double rcpPart;    /* particle heat capacity */
  int qsFlag;        /* ppiclf_user.f:473-485: 1 Parmar, 2 Osnes, 3 modified
                        Parmar, 4 Gidaspow; 5 Schiller-Naumann (rocinteract's
                        INRT_DRAG_MODEL_SN law in the same slot, not a
                        ppiclF choice) */
  int amFlag;        /* 0 off, 1 AM Parmar, 2 AM Briney (neighborWidth below) */
  int pgFlag;        /* 0 off, 1 pressure-gradient force */
  int feedbackFlag;  /* 0 one-way, 1 two-way */
  int interpMode;    /* 0: IDW over the 27 nearest element centroids
                        (ppiclf_solve_LocalInterp, ppiclf_solve.f:2234-2409,
                        the reference's choice); 1: trilinear evaluation of
                        the element's nodal field at the particle's (r,s,t)
                        (pfgslib_findpts_eval_local, ppiclf_solve.f:2198-2208) */
  int projMode;      /* 0: single-element box (InitBoxFilter(filt,0,1), the
                        reference's choice); 1: Gaussian filter
                        (ppiclf_solve_InitGaussianFilter, ppiclf_solve.f:1520-1576,
                        ProjectParticleGrid :3103-3246, no wall mirror) */
  double filterWidth;  /* projMode 1: filter width, the Gaussian's full width at half maximum */
  double filterAlpha;  /* projMode 1: cut-off; node weights below alpha are dropped (0 < alpha < 1) */
  /* HEATTRANSFER (ppiclf_user_HeatTransfer.f): 0 off; Nusselt law 1 Stokes,
   * 2 Ranz-Marshall, 3 Gunn, 4 Fox.  The particle temperature y(JT) then
   * evolves and the heat flux joins the energy feedback. */
  int heatTransfer;
  /* PERIODICX/Y/Z (global%piclPeriodicXFlag..., PICL_TEMP_InitSolver.F90:
   * 234-262, 715-734): 1 makes the axis periodic for the particles over the
   * mesh's vertex minimum to maximum.  After every RK stage a particle beyond
   * one bound re-enters as far inside the other
   * (ppiclf_solve_InvokeLinearPeriodic, ppiclf_solve.f:1319-1353), and the
   * Gaussian projection also deposits the particles' images shifted by one
   * period (the ghost particles of ppiclf_comm_CreateGhost,
   * ppiclf_comm.f:1115-1530). */
  int periodicX, periodicY, periodicZ;
  /* ViscousUnsteady_flag (global%piclViscousUnsteady,
   * PICL_TEMP_InitSolver.F90:205): 1 adds the viscous-unsteady force of
   * ppiclf_user_ViscousUnsteady.f, a history kernel over the last 5 steps'
   * d(rho u)/dt of gas and particle (PPICLF_VU = 5). */
  int viscousUnsteady;
  /* COLLISIONS (collisional_flag, PICL_TEMP_InitSolver.F90:204): 0 off; 1 the
   * normal soft-sphere force between overlapping particles
   * (ppiclf_user_EvalNearestNeighbor.f:118-233, Cundall and Strack), with
   * periodic images on periodic axes; 2 adds the tangential force (Coulomb
   * friction or dashpot, the smaller), the contact and rolling-resistance
   * torques and the angular velocity y(JOX..JOZ) they drive; 3 adds the
   * hydrodynamic torque and the undisturbed-flow torque
   * (ppiclf_user_Torque.f:72-197) from the gas vorticity and its material
   * derivative, interpolated like the other fields (PICL_TEMP_Runge.F90:
   * 618-663, 747-809, with the feedback force's gradient); 4 adds the Saffman
   * and Magnus lift (ppiclf_user_Lift.f:61-181).  KSP and EREST (global%piclKsp, piclERest, :222-223):
   * the spring stiffness is capped at rmass pi^2 / (ksp dt)^2, erest in (0, 1]
   * sets the damping. */
  int collisionalFlag;
  double ksp, erest;
  /* amFlag 2 (AMFORCE 2, Briney's added mass,
   * libpicl/user_files/ppiclf_user_AddedMass.f:164-400 with the pair terms of
   * ppiclf_user_EvalNearestNeighbor.f:115-159): the neighbour width
   * ppiclf_d2chk(3) within which other particles (and their periodic images)
   * add pair terms.  rocpicl passes max(4 d_max, NEIGHBORWIDTH)
   * (PICL_TEMP_InitSolver.F90:655-661); 0 here means 4 d_max. */
  double neighborWidth;
  /* QSFLUCT (qs_fluct_flag, ppiclf_user_Fluctuations.f:32-418): 0 off; 1 the
   * Lattanzi and 2 the Osnes fluctuating quasi-steady force, an SDE step per
   * RK stage from the granular temperature of the neighbours within a box of
   * half-width fluctHalfWidth (ppiclf_d2chk(2); 0 means what rocpicl's
   * filter set-up leaves there: with the box projection the last element's
   * largest extent, rocpicl's filter 4 MAXVAL(Max_EleLen)/2
   * (PICL_TEMP_InitSolver.F90:598-646) halved by InitBoxFilter
   * (ppiclf_solve.f:1611); with the Gaussian one its cut-off radius
   * (ppiclf_solve.f:1558-1559)) inside the neighbour width.  qsFluctAdapt 1 (QSFLUCTFILTADAPT) widens the
   * box to the particle's adaptive filter.  The box filter only: the
   * Gaussian one (QSFLUCTFILT 1) reads a variable the reference never sets
   * (ppiclf_user_EvalNearestNeighbor.f:50,364).  The normal draws come from
   * a counter-based generator seeded by fluctSeed, not from gfortran's
   * RANDOM_NUMBER stream. */
  int qsFluct;
  int qsFluctAdapt;
  double fluctHalfWidth;
  long long fluctSeed;
  /* Angular periodicity about the z axis (ANGULARPERIODIC 1; ppiclF's
   * ang_per_* arguments of ppiclf_solve_Initialize, ppiclf_solve.f:1-104,
   * and ppiclf_solve_InitAngularPeriodic, :1172-1230): a wedge from the polar
   * angle angXAngle to angXAngle + angAngle (radians, as ppiclF receives
   * them; the deck's degrees times acos(-1)/180, PICL_TEMP_InitSolver.F90:
   * 271-273), an opening of less than 90, of 90 or of 180 degrees, radii
   * angRin < angRout.  After each stage's update a particle whose polar angle
   * left the wedge has its position and velocity rotated back
   * (ppiclf_solve_InvokeAngularPeriodic with the reference's rotation
   * matrix, :2852-2881); with periodicZ the z axis wraps and a particle whose
   * x lies outside [angRin, angRout] ([-angRout, angRout] for the half
   * cylinder) is removed at the next stage (InvokeLinearPeriodic's second
   * case, :1362-1394).  Excludes periodicX and periodicY; needs the box
   * projection and no pair models (ppiclF's rotated ghost particles,
   * AngularCreateGhost, are not built). */
  int angularPeriodic;
  double angAngle, angXAngle, angRin, angRout;
} picl_params;

/* PICL_TEMP_InitSolver (rocpicl/PICL_TEMP_InitSolver.F90:538-835):
 * ppiclf_solve_InitParticle (libpicl/ppiclF/source/ppiclf_solve.f:184-188)
 * + ppiclf_comm_InitOverlapMesh (ppiclf_comm.f:926-998, its InitSolve
 * projects the particles), then piclVF from the projection, cv scaled by
 * (1 - piclVF) and dv recomputed.  The handle's state must hold the pure-gas
 * cv (rflu_gpu_set_state).  y: 12 doubles per particle; pconst: (rho_p, d_p,
 * V_p, SPL) per particle. */
int picl_gpu_init(rflu_handle *h, const picl_params *pp, int nPart, const double *y, const double *pconst);

/* PICL_TEMP_Runge (rocpicl/PICL_TEMP_Runge.F90:623-1234) for RK stage
 * istage: field assembly, ppiclf_solve_IntegrateParticle
 * (ppiclf_solve.f:1694-1936: locate, LocalInterp, CreateGhost map,
 * ProjectParticleGrid, SetYdot, RK3), feedback into rhs and the new piclVF.
 * Call after rflu_gpu_source_terms and before rflu_gpu_rk_update. */
int picl_gpu_runge(rflu_handle *h, int istage);

int picl_gpu_sizes(const rflu_handle *h, int *nPart, int *nLost);
/* Particle state (any pointer may be NULL): y, y1, ydot, ydotc 12 per
 * particle; elem one per particle (-1: lost). */
int picl_gpu_get_particles(rflu_handle *h, double *y, double *y1, double *ydot, double *ydotc, int *elem);
/* Per cell (any pointer may be NULL): piclVF; limited grad(rho_g) (3,
 * piclgradRhog); projected node values of PHIP, FX, FY, FZ, E (5,
 * ppiclf_pro_fld via GetProFldIJKEF, ppiclf_solve.f:3630-3647); the 20
 * interpolation inputs of the last stage, in rprop order 2,6,7,8,10,11,12,9,
 * 24,5,13,14,15,16,17,18,19,37,38,39 (PPICLF_USER.h). */
int picl_gpu_get_cells(rflu_handle *h, double *phi, double *gradRhog, double *proFld, double *fields);
/* Projected values at the 8 nodes of every element (ppiclf_pro_fld via
 * GetProFldIJKEF): 40 per cell, (PHIP, FX, FY, FZ, E) x ppiclF node
 * n = i + 2j + 4k, field-major. */
int picl_gpu_get_nodes(rflu_handle *h, double *proNode);
/* With collisionalFlag >= 3: the vorticity and its material derivative per
 * cell at the last stage, 6 per cell (rprop JXVOR..JZVOR, JSDOX..JSDOZ: the
 * fields PICL_TEMP_Runge.F90:747-809 assembles and :880-882,890-892 hand to
 * ppiclf_solve_InterpFieldUser). */
int picl_gpu_get_rot_fields(rflu_handle *h, double *rot6);
/* Values LocalInterp produced at the last stage, 20 per particle in the
 * order above (kept only while picl_gpu_set_debug(h, 1) is on). */
int picl_gpu_get_interp(rflu_handle *h, double *rprop20);
int picl_gpu_set_debug(rflu_handle *h, int on);
/* Element holding each of n points (-1: none), by the library's location
 * rule (lowest-numbered element whose trilinear map contains the point). */
int picl_gpu_locate(rflu_handle *h, int n, const double *xyz, int *elem);

/* picl_gpu_locate without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The points
 * and answers are staged in a static device array that each GPU holds until
 * the process ends, set up by picl_gpu_init; n may not exceed
 * PICL_LOCATE_STATIC_MAX (RFLU_GPU_EINVAL otherwise).  Calls that overlap on
 * one GPU share that array and race. */
#define PICL_LOCATE_STATIC_MAX 262144
int picl_gpu_locate_static(rflu_handle *h, int n, const double *xyz, int *elem);

#ifdef __cplusplus
}
#endif
#endif
