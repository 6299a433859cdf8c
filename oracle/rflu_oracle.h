/*
 * CPU oracle for the RocfluMP explicit residual path -- TEST INFRASTRUCTURE.
 *
 * This is a plain-C restatement of the reference's Fortran loops, kept in the
 * reference's loop and summation order (face-loop scatter, in-place
 * cons<->prim round trips, two R/gamma paths).  It exists only so tests,
 * smoke() and bench.py's cpu_baseline leg can check and time the HIP library;
 * nothing in the product links or calls it.
 *
 * Layouts follow the reference's Fortran arrays (column-major, i.e. AoS per
 * cell): cv(5,N) -> cv[5*i+k], dv(3,N) -> dv[3*i+k] (p,T,c),
 * gradCell(3,5,N) -> grad[15*i+3*k+d], fn(4,F) -> fn[4*f+d], fc(3,F),
 * cofg(3,N), f2c(2,F) -> f2c[2*f+s] (0-based cells).
 */
#ifndef RFLU_ORACLE_H
#define RFLU_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_DISCR_ROE = 1, ORC_DISCR_HLLC = 3, ORC_DISCR_AUSM = 4, ORC_DISCR_AUSMUP = 5 };
enum { ORC_RECONST_NONE = 0, ORC_RECONST_WENO_XYZ = 2, ORC_RECONST_VENKAT = 11 };
/* patch kinds (the reference's BC_SLIPWALL, BC_OUTFLOW, BC_NOSLIPWALL_HFLUX/
 * _TEMP, BC_FARFIELD, BC_INJECTION, BC_INFLOW_VELTEMP, BC_INFLOW_TOTANG) */
enum { ORC_BC_SLIPWALL = 1, ORC_BC_OUTFLOW = 2, ORC_BC_NOSLIPWALL = 3, ORC_BC_FARFIELD = 4, ORC_BC_INJECTION = 5,
       ORC_BC_INFLOW_VELTEMP = 6, ORC_BC_INFLOW_TOTANG = 7 };
/* values per patch (uniform over the patch):
 *   FARFIELD        mach, angle of attack, slip angle (radians), p, T
 *   INJECTION       mass flux (> 0: burning), T
 *   INFLOW_VELTEMP  u, v, w, T, p (p used unless bcOpt is subsonic)
 *   INFLOW_TOTANG   p_tot, T_tot, beta_h, beta_v (radians), mach, fixed (0/1) */
enum { ORC_BCVALS = 6 };
enum { ORC_BCOPT_SUPERSONIC = 0, ORC_BCOPT_SUBSONIC = 1, ORC_BCOPT_MIXED = 2 };

typedef struct orc_params {
  int spaceDiscr;     /* 1 Roe, 3 HLLC, 4 AUSM+, 5 AUSM+up */
  int spaceOrder;     /* 1 or 2 */
  int reconst;        /* 0 none, 2 WENO-XYZ, 11 Venkatakrishnan */
  int patchOrder;     /* 1 or 2: order of all boundary patches (ORDERBF) */
  double cfl;
  double dtImposed;
  double dtMinLimit;
  double dTimeSystem; /* run length; dt is clipped to it (<=0: no clip) */
  double refCp;
  double refGamma;
  double epsentr;
  double dissFact;
  int accelOn;
  double accel[3];
  int limiterGuard;   /* 0: reference denom+TINY; 1: TINY with the sign of denom */
  /* AUSM+up constants (NUMERICS SIGMA, KP, KU, MINF; defaults 1, 0.25, 0.75,
   * 0.1 in libfloflu/ReadNumericsSection.F90:236-275) */
  double sigma, kp, ku, mInf;
} orc_params;

typedef struct orc_region orc_region;

/* Build the region: face list, geometry, stencils (order 2), moments.
 * c2v is nCells x nvc, 0-based: nvc 4 = tets; nvc 8 = hexes, or mixed cells
 * (tet 4, pyramid 5, prism 6, hex 8 vertices) padded with -1; boundary
 * triangles among quads (nvf 4) are padded with -1 too.
 * Patches: patch p has pNBF[p] faces with nvf vertices each, concatenated in
 * bf2v.  Returns NULL on error (message via orc_last_error). */
orc_region *orc_create(int nVert, const double *xyz, int nvc, int nCells,
                       const int *c2v, int nPatches, const int *pBcType,
                       const int *pBcOpt, const double *pPout, const int *pNBF,
                       int nvf, const int *bf2v, const orc_params *prm);
void orc_destroy(orc_region *r);
const char *orc_last_error(void);
void orc_set_patch_vals(orc_region *r, int ip, const double *vals);

int orc_nfaces(const orc_region *r);
int orc_nbfaces(const orc_region *r);
/* Mesh getters (arrays sized as documented above). */
void orc_get_faces(const orc_region *r, int *f2c, int *f2v, double *fn, double *fc);
void orc_get_bfaces(const orc_region *r, int *bf2c, double *bfn, double *bfc);
void orc_get_cells(const orc_region *r, double *vol, double *cofg);
int orc_stencil_total(const orc_region *r);
void orc_get_stencils(const orc_region *r, int *off, int *memb, double *moms);

void orc_set_cv(orc_region *r, const double *cv);  /* also sets dv, gv */
void orc_get_cv(const orc_region *r, double *cv);
void orc_get_dv(const orc_region *r, double *dv);
void orc_get_rhs(const orc_region *r, double *rhs);
void orc_get_grad(const orc_region *r, double *grad);

/* RFLU_TimeStepInviscid on the current state: per-cell dt, returns MINVAL. */
double orc_timestep_inviscid(orc_region *r, double *dtCell);
/* CellGradientsMP + ConvectiveFluxesMP (+ SourceTermsMP) on the current
 * state, leaving rhs; cv keeps the round-tripped bits as in the reference. */
void orc_residual(orc_region *r);
/* Full steps: dt, RFLU_MinimumTimeStep, RungeKuttaMP (3 Wray stages). */
int orc_step(orc_region *r, int nSteps);
double orc_time(const orc_region *r);
/* One RK stage with the current dtMin, and a setter for dtMin (diagnostics). */
int orc_rk_stage(orc_region *r, int istage);
void orc_set_dtmin(orc_region *r, double dt);
double orc_last_dt(const orc_region *r);

/* Point flux functions (for closed-form tests). */
void orc_flux_ausm(const double *nrm, double fs, const double *L, const double *R, double *flx);
void orc_flux_point(const orc_params *p, int discr, const double *nrm, const double *L, const double *R,
                    double *flx);

#ifdef __cplusplus
}
#endif
#endif
