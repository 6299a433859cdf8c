/*
 * CPU oracle for the ppiclF coupling of RocfluMP -- TEST INFRASTRUCTURE.
 *
 * Restates PICL_TEMP_InitSolver / PICL_TEMP_Runge (rocpicl/) and the ppiclF
 * routines they drive for one rank: element location, LocalInterp (IDW over
 * the 27 nearest element centroids) or the trilinear findpts_eval_local,
 * ppiclf_user_SetYdot with the QS laws of ppiclf_user_QuasiSteady.f, AM
 * Parmar and the pressure-gradient force, IntegrateRK3s_Rocflu, and the
 * single-element box or Gaussian-filter projection with its read-back into
 * rhs and piclVF.
 * Only tests/ and bench.py's CPU-baseline leg use it, as the checker.
 *
 * Particle arrays use ppiclF's layout: y(PPICLF_LRS=12, n) -> y[12*p + j],
 * rprop(PPICLF_LRP=48, n) -> rprop[48*p + j] (0-based j = index - 1).
 */
#ifndef PICL_ORACLE_H
#define PICL_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_region orc_region;

typedef struct orc_picl_params {
  int rmuFlag;       /* 0: fixed viscosity rmuRef; 1: Sutherland law */
  double rmuRef;     /* reference viscosity */
  double tref;       /* Sutherland reference temperature */
  double suth;       /* Sutherland constant */
  double rcpPart;    /* particle heat capacity (ppiclf_rcp_part) */
  int qsFlag;        /* 1 Parmar, 2 Osnes, 3 modified Parmar, 4 Gidaspow,
                        5 Schiller-Naumann (INRT_DRAG_MODEL_SN in ppiclF's slot) */
  int amFlag;        /* 0: off, 1: AM Parmar, 2: AM Briney (unary + binary terms) */
  int pgFlag;        /* 0: off, 1: pressure-gradient force */
  int feedbackFlag;  /* 0: one-way, 1: two-way coupling */
  int interpMode;    /* 0: IDW-27 (LocalInterp); 1: trilinear (findpts_eval_local) */
  int projMode;      /* 0: single-element box; 1: Gaussian filter */
  double filterWidth;  /* Gaussian filter width (full width at half maximum) */
  double filterAlpha;  /* Gaussian cut-off: weights below alpha are dropped */
  int heatTransfer;  /* 0: off; Nusselt law 1 Stokes, 2 Ranz-Marshall, 3 Gunn, 4 Fox */
  int periodicX, periodicY, periodicZ; /* PERIODICX/Y/Z: particles wrap across the mesh's
                                          vertex bounds on that axis */
  int viscousUnsteady; /* 0: off; 1: viscous-unsteady force with its history kernel */
  int collisionalFlag; /* 0: off; 1: normal soft-sphere collision force (Cundall-Strack);
                          2: + tangential force, contact and rolling torques */
  double ksp;          /* KSP: collision time = ksp * dt sets the spring stiffness cap */
  double erest;        /* ERESTITUTION: restitution coefficient of the damping */
  double neighborWidth; /* amFlag 2: ppiclf_d2chk(3), the pair terms' reach; 0 means 4 x the
                           largest diameter (PICL_TEMP_InitSolver.F90:655-661 with NEIGHBORWIDTH 0) */
  int qsFluct;          /* QSFLUCT: 0 off, 1 Lattanzi, 2 Osnes quasi-steady force fluctuations */
  int qsFluctAdapt;     /* QSFLUCTFILTADAPT: 1 widens the box to the particle's adaptive filter */
  double fluctHalfWidth; /* ppiclf_d2chk(2), the box filter's half-width; 0 means what the filter's
                            set-up leaves: the last element's largest extent (box projection) or
                            the Gaussian filter's cut-off radius */
  long long fluctSeed;  /* seed of the uniform draws that stand in for RANDOM_NUMBER */
  int angularPeriodic;  /* ANGULARPERIODIC: particles wrap across a wedge about the z axis */
  double angAngle, angXAngle; /* the wedge's opening and lower face angle (radians) */
  double angRin, angRout;     /* its radii (the removal range with periodicZ) */
} orc_picl_params;

/* Interpolated fields, in slot order, with their ppiclF rprop index
 * (1-based, libpicl/user_files/PPICLF_USER.h). */
#define ORC_PICL_NFLD 20
extern const int orc_picl_field_rprop[ORC_PICL_NFLD];

/* PICL_TEMP_InitSolver (rocpicl/PICL_TEMP_InitSolver.F90:538-835) on a
 * region whose cv holds the pure-gas state: locate the particles, project
 * them (ydotc = 0), set piclVF, scale cv by (1 - piclVF) and recompute dv.
 * y: 12 per particle; pconst: (rho_p, d_p, V_p, SPL) per particle.
 * Requires an all-hex, second-order region.  Returns 0 or -1. */
int orc_picl_init(orc_region *r, const orc_picl_params *pp, int nPart, const double *y,
                  const double *pconst);
int orc_picl_npart(const orc_region *r);
int orc_picl_nlost(const orc_region *r);
void orc_picl_get_particles(const orc_region *r, double *y, double *y1, double *ydot, double *ydotc,
                            int *elem);
void orc_picl_get_rprop(const orc_region *r, double *rprop);
/* Per cell: piclVF, limited grad(rho_g) (3), projected node values of
 * (PHIP, FX, FY, FZ, E) (5), interpolation inputs of the last stage (20). */
void orc_picl_get_cells(const orc_region *r, double *phi, double *gradRhog, double *proFld,
                        double *fields);
/* Projected values at the 8 nodes of every element: 40 per cell, field-major
 * ((PHIP, FX, FY, FZ, E) x ppiclF node n = i + 2j + 4k). */
void orc_picl_get_nodes(const orc_region *r, double *proNode);
/* The quasi-steady coefficient beta (force = beta * v_rel) of model qsFlag
 * for one particle, from SetYdot's capped inputs. */
double orc_picl_qs_beta(int qsFlag, double rmachp, double rphip, double rep, double rmu, double dp,
                        double rhof, double vmag);
/* The viscous-unsteady force of one particle from its history: entries
 * 1..nTimeBH of the ages timeBH and of d(rho u)/dt of the gas (mixt) and the
 * particle (plag), 5 per component ([c * 5 + k]); SetYdot's capped inputs. */
void orc_picl_vu_force(int nTimeBH, const double *timeBH, const double *mixt, const double *plag, double dt,
                       double rmu, double rhof, double dp, double vmag, double rphip, double *fvu);
/* Viscous-unsteady bookkeeping: returns nTimeBH, fills the 5 ages and the
 * particles' histories (30 per particle: gas then particle, [c * 5 + k]). */
int orc_picl_vu_state(const orc_region *r, double *timeBH, double *hist);
/* The collision force of particle j on particle i (COLLISIONS 1) for the
 * given positions, velocities, diameters, densities and volumes. */
void orc_picl_collision_force(const double *xi, const double *vi, double dpi, double rhopi, double volpi,
                              const double *xj, const double *vj, double dpj, double rhopj, double volpj, double dt,
                              double ksp, double erest, double *fc);
/* COLLISIONS 2: force and torque of particle j on particle i from their
 * 12-component states (x, v, T, omega, ...). */
void orc_picl_collision_force2(const double *yi12, const double *yj12, double dpi, double rhopi, double volpi,
                               double dpj, double rhopj, double volpj, double dt, double ksp, double erest,
                               double *fc, double *tq);
/* Element containing each point (-1: none), by the oracle's location rule. */
void orc_picl_locate(const orc_region *r, int n, const double *xyz, int *elem);
/* Elements LocalInterp sums for point xp, nearest first; returns how many. */
int orc_picl_neighbors(const orc_region *r, const double *xp, int *ids, double *d2);
void orc_picl_fluct_means(const orc_region *r, int i, double out[7]);

/* Briney's added-mass model (libpicl/user_files/Added_Mass/): the four
 * fitted curves, the distant-neighbour corrections (analytical, and the
 * Simpson sums over the radial distribution function), and the pair's
 * resistance matrix R[k][l] (l < 3: B11, l >= 3: B12, rotated to the pair). */
double orc_am_b11_11(double d, double alpha);
double orc_am_b11_22(double d, double alpha);
double orc_am_b12_11(double d, double alpha);
double orc_am_b12_22(double d, double alpha);
double orc_am_ia_analytical(double rmax, double rad, double alpha);
double orc_am_ii_analytical(double rmax, double rad, double alpha);
double orc_am_rdf(double r, double alpha);
double orc_am_ia_numerical(double rmax, double rad, double alpha);
double orc_am_ii_numerical(double rmax, double rad, double alpha);
void orc_am_resistance_pair(double x, double y, double z, double alpha, double rad, double R[3][6]);

/* COLLISIONS 3-4 laws on given inputs: Torque_Hydro (returns -1 where ppiclF
 * stops, Re_r >= 100) and Saffman + Magnus lift (rep already capped). */
int orc_picl_torque_hydro(const double *vort, const double *spin, double rhof, double dp, double rmu, double rmass,
                          double rhop, double *t);
void orc_picl_lift(const double *vort, const double *spin, const double *vrel, double dp, double rep, double rmu,
                   double rhof, double *lift);

/* QS fluctuations: the uniform draw k of a particle's call, and one SDE step
 * (flag 1 Lattanzi, 2 Osnes) from the box means m (sums of u, v, w, u^2, v^2,
 * w^2 and the count), the previous force F and the uniforms U. */
double orc_fluct_uniform(long long seed, int call, int pid, int k);
void orc_fluct_force(int flag, const double m[7], const double F[3], const double U[6], double rphip,
                     double rmachp, double rep, double vmag, const double vrel[3], double rmu, double dp,
                     double fac, double out[3]);

#ifdef __cplusplus
}
#endif
#endif
