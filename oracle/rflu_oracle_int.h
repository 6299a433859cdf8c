/*
 * Internal layout of the oracle region, shared by rflu_oracle.c and
 * picl_oracle.c -- TEST INFRASTRUCTURE (see rflu_oracle.h).
 */
#ifndef RFLU_ORACLE_INT_H
#define RFLU_ORACLE_INT_H
#include "rflu_oracle.h"
#include "picl_oracle.h"

struct orc_region {
  orc_params p;
  int nVert, nvc, nCells;
  double *xyz;
  int *c2v;
  /* interior faces (reference order) */
  int nFaces;
  int *f2c, *f2v;
  double *fn, *fc;
  /* boundary faces, patch by patch */
  int nPatches, nvf, nBFaces;
  int *pBcType, *pBcOpt, *pNBF, *pOff;
  double *pPout;
  double *pVals;  /* ORC_BCVALS per patch */
  int *bf2v, *bf2c;
  double *bfn, *bfc;
  /* cells */
  double *vol, *cofg;
  /* c2c stencils */
  int *csOff, *csMem;
  double *moms; /* 10 per cell: 11,12,22,13,23,33,14,24,34,44 */
  /* state */
  double *cv, *cvOld, *dv, *rhs, *rhsSum, *diss, *grad, *lim, *dt;
  double cpGv, molGv; /* gv(CP), gv(MOL): constant for TCPERF */
  /* PICL coupling (NULL when particles are off): piclVF per cell, the
   * limited gradient of rho_g = rho/(1-phi) (piclgradRhog), particle state */
  double *phi, *gradRhog;
  struct orc_picl *picl;
  double currentTime, timeSinceRestart, dtMin;
};

/* Hooks between the two translation units. */
void orc__set_dependent_vars(orc_region *r);
int orc__picl_runge(orc_region *r, int istage); /* PICL_TEMP_Runge (picl_oracle.c) */
void orc__picl_free(struct orc_picl *p);
/* CellGradientsMP's building blocks, on any cell field (rflu_oracle.c) */
void orc__gradient_field(orc_region *r, const double *v, int vs, double *out, int os);
void orc__weno_xyz(orc_region *r, double *grad, int stride);
void orc__limit_simple_field(orc_region *r, double *grad, int stride, const double *vals, int vstride);

#endif
