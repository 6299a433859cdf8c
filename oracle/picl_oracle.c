/*
 * CPU oracle of the ppiclF coupling -- TEST INFRASTRUCTURE (picl_oracle.h).
 *
 * Only tests/ and bench.py's cpu_baseline leg load this, as the checker.
 * One rank, one region, all-hex mesh.  Every expression keeps the Fortran
 * evaluation order, and single-precision literals of the ppiclF user code are
 * written as (double)<literal>f: the user files are compiled without
 * -fdefault-real-8, so 0.15*re is (real*4 0.15 widened) * re.
 * Integer powers follow gfortran's expansion: x**2 = x*x, x**3 = (x*x)*x,
 * x**4 = (x*x)*(x*x).
 *
 * Not restated (the enabled models never read them): the feedback-force
 * gradient and the vorticity / DOmega/Dt / p / div(tau) fields of
 * PICL_TEMP_Runge (they feed only lift, torque, burn rate and the viscous
 * models), and the 17 projected fields Rocflu does not read back.
 */
#include "rflu_oracle_int.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* PPICLF_USER.h indices, 0-based */
enum { JX = 0, JY, JZ, JVX, JVY, JVZ, JT };
enum {
  R_JRHOP = 0, R_JRHOF = 1, R_JDP = 2, R_JVOLP = 3, R_JPHIP = 4, R_JUX = 5, R_JUY = 6, R_JUZ = 7,
  R_JCS = 8, R_JDPDX = 9, R_JDPDY = 10, R_JDPDZ = 11, R_JSDRX = 12, R_JSDRY = 13, R_JSDRZ = 14,
  R_JRHSR = 15, R_JPGCX = 16, R_JPGCY = 17, R_JPGCZ = 18, R_JSPL = 21, R_JT = 23,
  R_FLUCTFX = 24, R_WDOTX = 27, R_JXVOR = 30, R_JSDOX = 42, R_JRHOGX = 36, R_JRHOGY = 37, R_JRHOGZ = 38
};
enum { P_JPHIP = 0, P_JFX, P_JFY, P_JFZ, P_JE, NPRO };
#define LRS 12
#define LRP 48

const int orc_picl_field_rprop[ORC_PICL_NFLD] = {2, 6, 7, 8, 10, 11, 12, 9, 24, 5,
                                                 13, 14, 15, 16, 17, 18, 19, 37, 38, 39};

/* ppiclF node (i,j,k) of a Rocflu hex, n = i + 2j + 4k:
 * hex2v 1,4,5,8,2,3,6,7 (rocpicl/PICL_TEMP_InitSolver.F90:562-597) */
static const int NODE2V[8] = {0, 3, 4, 7, 1, 2, 5, 6};

#define F32(x) ((double)(x##f))

struct orc_picl {
  orc_picl_params pp;
  int n, ne;
  double *y, *y1, *ydot, *ydotc, *rprop;
  int *elem;
  char *lost;
  double *xm1b;    /* 24 per element: node n, coordinate d at [3n + d] */
  double *centeri; /* LocalInterp element centroid */
  double *bbox;    /* min xyz, max xyz */
  double d2Max[3]; /* (1.5 * extent of the last element)^2 per axis */
  double *rexp;    /* 1 / (12-tet hex volume) */
  double *proFld;  /* 8 * NPRO per element: field jj at node n in [8 * jj + n] */
  double gRdum, gMult, gD2chk; /* Gaussian filter: 1/(-2 sigma^2), normalisation, cut-off^2 */
  /* viscous-unsteady history (ppiclF's rprop3, drudtMixt/drudtPlag): per
   * particle VU_N values of d(rho u)/dt of the gas at [c * VU_N + k] and of
   * the particle at [3 * VU_N + c * VU_N + k], k = 0 the newest */
  double *hist;
  int nTimeBH;               /* ppiclf_nTimeBH */
  double timeBH[5];          /* ppiclf_timeBH: age of each history entry */
  int per[3];                  /* linear periodicity per axis */
  double perLo[3], perHi[3];   /* ppiclf_xdrange: the mesh's vertex bounds */
  double *fld;     /* ORC_PICL_NFLD interpolation inputs per element */
  double nbW;      /* ppiclf_d2chk(3): the added-mass pair terms' reach */
  double *fldRot;  /* COLLISIONS >= 3: vorticity and D(omega)/Dt per element */
  int rotErr;      /* a rotational Reynolds number reached 100 this stage */
  double fluctW;   /* ppiclf_d2chk(2): the fluctuation box filter's half-width */
  int nCalls;      /* SetYdot calls so far: the counter of the uniform draws */
  int curStage;    /* the RK stage of the current SetYdot call */
  /* angular periodicity (ang_per_flag 1): ppiclF's case, the wedge's polar
   * angles, the sign switch 0.5 (xangle + angle), ppiclf_xdrange(:,1), and
   * the rotation matrices of the local angles +angle, -angle, 0 */
  int ang, angCase;
  double angX, angA, angXd[2], angR[3][3][3];
  char *mark;      /* removed by the x range test: gone from the next stage */
};

/* The rotational fields (rprop JXVOR..JZVOR, JSDOX..JSDOZ, 0-based) */
static const int ROT_RPROP[6] = {R_JXVOR, R_JXVOR + 1, R_JXVOR + 2, R_JSDOX, R_JSDOX + 1, R_JSDOX + 2};

void orc__picl_free(struct orc_picl *p) {
  if (!p) return;
  void *ptrs[] = {p->y, p->y1, p->ydot, p->ydotc, p->rprop, p->elem, p->lost, p->xm1b,
                  p->centeri, p->bbox, p->rexp, p->proFld, p->fld, p->hist, p->fldRot, p->mark};
  for (size_t q = 0; q < sizeof ptrs / sizeof ptrs[0]; q++) free(ptrs[q]);
  free(p);
}

/* ------------------------------------------------------------------------ */
/* Overlap mesh                                                              */
/* ------------------------------------------------------------------------ */
/* Single-element box volume: ppiclf_solve.f:3285-3394 (12 tetrahedra from
 * each face's two triangles to the element centroid). */
static double hex_volume(const double *X) {
  double centroid[3];
  for (int ix = 0; ix < 3; ix++) centroid[ix] = 0.0;
  for (int ix = 0; ix < 3; ix++)
    for (int n = 0; n < 8; n++) centroid[ix] = centroid[ix] + X[3 * n + ix];
  for (int ix = 0; ix < 3; ix++) centroid[ix] = centroid[ix] / 8.0;
  double evol = 0.0;
  for (int ix = 0; ix < 3; ix++) {
    /* face_map: dimension ix takes the face side, ia = max(3-ix,1) the
     * first in-face index, ib = min(5-ix,3) the second (1-based) */
    int ia = (3 - (ix + 1) > 1 ? 3 - (ix + 1) : 1) - 1;
    int ib = (5 - (ix + 1) < 3 ? 5 - (ix + 1) : 3) - 1;
    for (int iface = 0; iface < 2; iface++) {
      double face[2][2][3];
      for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
          int a[3];
          a[ix] = iface; a[ia] = i; a[ib] = j;
          int n = a[0] + 2 * a[1] + 4 * a[2];
          for (int d = 0; d < 3; d++) face[i][j][d] = X[3 * n + d];
        }
      double v1[3], v2[3], v3[3], cross[3];
      for (int d = 0; d < 3; d++) {
        v1[d] = face[0][1][d] - face[1][0][d];
        v2[d] = centroid[d] - face[1][0][d];
      }
      cross[0] = v1[1] * v2[2] - v1[2] * v2[1];
      cross[1] = v1[2] * v2[0] - v1[0] * v2[2];
      cross[2] = v1[0] * v2[1] - v1[1] * v2[0];
      for (int inode = 0; inode < 2; inode++) {
        for (int d = 0; d < 3; d++) v3[d] = face[inode][inode][d] - face[1][0][d];
        double voltet = 0.0;
        for (int d = 0; d < 3; d++) voltet = voltet + v3[d] * cross[d];
        evol = evol + fabs(voltet);
      }
    }
  }
  return evol / 6.0;
}

static int setup_mesh(orc_region *r, struct orc_picl *p) {
  int ne = r->nCells;
  p->ne = ne;
  p->xm1b = malloc(sizeof(double) * 24 * (size_t)ne);
  p->centeri = malloc(sizeof(double) * 3 * (size_t)ne);
  p->bbox = malloc(sizeof(double) * 6 * (size_t)ne);
  p->rexp = malloc(sizeof(double) * (size_t)ne);
  p->proFld = calloc(8 * NPRO * (size_t)ne, sizeof(double));
  p->fld = calloc(ORC_PICL_NFLD * (size_t)ne, sizeof(double));
  for (int e = 0; e < ne; e++) {
    double *X = p->xm1b + 24 * (size_t)e;
    for (int n = 0; n < 8; n++)
      for (int d = 0; d < 3; d++) X[3 * n + d] = r->xyz[3 * (size_t)r->c2v[8 * (size_t)e + NODE2V[n]] + d];
    for (int d = 0; d < 3; d++) {
      double lo = X[d], hi = X[d];
      for (int n = 1; n < 8; n++) { lo = fmin(lo, X[3 * n + d]); hi = fmax(hi, X[3 * n + d]); }
      p->bbox[6 * (size_t)e + d] = lo;
      p->bbox[6 * (size_t)e + 3 + d] = hi;
    }
    p->rexp[e] = 1.0 / hex_volume(X);
  }
  /* LocalInterp set-up, ppiclf_solve.f:2253-2289: centroid = node mean; the
   * extent threshold is reset for every element, so the last one's stays */
  for (int e = 0; e < ne; e++) {
    const double *X = p->xm1b + 24 * (size_t)e;
    for (int l = 0; l < 3; l++) {
      double c = 0.0, mx = -1.0e10, mn = 1.0e10;
      for (int n = 0; n < 8; n++) {
        c = c + X[3 * n + l];
        if (X[3 * n + l] > mx) mx = X[3 * n + l];
        if (X[3 * n + l] < mn) mn = X[3 * n + l];
      }
      double d2i = (mx - mn) * (mx - mn), d2m = 0.0;
      if (d2i > d2m) d2m = d2i;
      p->d2Max[l] = d2m;
      p->centeri[3 * (size_t)e + l] = c / 8;
    }
  }
  for (int l = 0; l < 3; l++) p->d2Max[l] = p->d2Max[l] * (1.5 * 1.5);
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Element location.  gslib findpts is not restated: the rule here is the    */
/* lowest-numbered element whose trilinear map holds the point within        */
/* |r|,|s|,|t| <= 1 + 1e-10 (Newton from the element centre).                */
/* ------------------------------------------------------------------------ */
static int trilinear_inside(const double *X, const double *xp, double *rst) {
  double r = 0.0, s = 0.0, t = 0.0, prev = 0.0;
  for (int it = 0; it < 25; it++) {
    double a[2] = {0.5 * (1.0 - r), 0.5 * (1.0 + r)};
    double b[2] = {0.5 * (1.0 - s), 0.5 * (1.0 + s)};
    double c[2] = {0.5 * (1.0 - t), 0.5 * (1.0 + t)};
    const double da[2] = {-0.5, 0.5};
    double x[3] = {0.0, 0.0, 0.0}, J[3][3] = {{0.0}};
    for (int n = 0; n < 8; n++) {
      int i = n & 1, j = (n >> 1) & 1, k = n >> 2;
      double w = a[i] * b[j] * c[k];
      double wr = da[i] * b[j] * c[k], ws = a[i] * da[j] * c[k], wt = a[i] * b[j] * da[k];
      for (int d = 0; d < 3; d++) {
        double v = X[3 * n + d];
        x[d] = x[d] + w * v;
        J[d][0] = J[d][0] + wr * v;
        J[d][1] = J[d][1] + ws * v;
        J[d][2] = J[d][2] + wt * v;
      }
    }
    double f0 = xp[0] - x[0], f1 = xp[1] - x[1], f2 = xp[2] - x[2];
    double m00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double m01 = J[1][0] * J[2][2] - J[1][2] * J[2][0];
    double m02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * m00 - J[0][1] * m01 + J[0][2] * m02;
    if (!(fabs(det) > 0.0)) return 0;
    double dr = (f0 * m00 - J[0][1] * (f1 * J[2][2] - J[1][2] * f2) + J[0][2] * (f1 * J[2][1] - J[1][1] * f2)) / det;
    double ds = (J[0][0] * (f1 * J[2][2] - J[1][2] * f2) - f0 * m01 + J[0][2] * (J[1][0] * f2 - f1 * J[2][0])) / det;
    double dt = (J[0][0] * (J[1][1] * f2 - f1 * J[2][1]) - J[0][1] * (J[1][0] * f2 - f1 * J[2][0]) + f0 * m02) / det;
    r = r + dr; s = s + ds; t = t + dt;
    /* converged, or the steps no longer shrink: they are rounding noise,   */
    /* about ulp(x) / (element size) in reference coordinates              */
    double step = fabs(dr) + fabs(ds) + fabs(dt);
    if (step < 1.0e-12 || (it >= 2 && step >= 0.5 * prev)) break;
    prev = step;
  }
  if (rst) { rst[0] = r; rst[1] = s; rst[2] = t; }
  const double lim = 1.0 + 1.0e-10;
  return fabs(r) <= lim && fabs(s) <= lim && fabs(t) <= lim;
}

static int locate(const struct orc_picl *p, const double *xp) {
  for (int e = 0; e < p->ne; e++) {
    const double *bb = p->bbox + 6 * (size_t)e;
    int in = 1;
    for (int d = 0; d < 3; d++) {
      double tol = 1.0e-9 * (bb[3 + d] - bb[d]);
      if (xp[d] < bb[d] - tol || xp[d] > bb[3 + d] + tol) in = 0;
    }
    if (in && trilinear_inside(p->xm1b + 24 * (size_t)e, xp, NULL)) return e;
  }
  return -1;
}

/* ------------------------------------------------------------------------ */
/* LocalInterp: ppiclf_solve.f:2234-2409                                     */
/* ------------------------------------------------------------------------ */
static int nearest27(const struct orc_picl *p, const double *xp, int *inearest, double *d2) {
  int nnearest = 0;
  double d2s[28];
  int ins[28];
  for (int q = 0; q < 28; q++) { ins[q] = -1; d2s[q] = F32(1E20); }
  for (int ie = 0; ie < p->ne; ie++) {
    double d2i = 0.0;
    int farAway = 0;
    for (int l = 0; l < 3; l++) {
      double t = p->centeri[3 * (size_t)ie + l] - xp[l];
      double d2l = t * t;
      d2i = d2i + d2l;
      if (d2l > p->d2Max[l]) farAway = 1;
    }
    if (farAway) continue;
    int added = 0;
    for (int i = 1; i <= 27; i++) {
      int j = 27 - i; /* 0-based slot of the reference's j = 27 - i + 1 */
      if (d2i < d2s[j]) {
        d2s[j + 1] = d2s[j];
        ins[j + 1] = ins[j];
        d2s[j] = d2i;
        ins[j] = ie;
        added = 1;
      } else {
        break;
      }
    }
    if (added) nnearest = nnearest + 1;
  }
  if (nnearest > 27) nnearest = 27;
  for (int q = 0; q < nnearest; q++) { inearest[q] = ins[q]; d2[q] = d2s[q]; }
  return nnearest;
}

static void local_interp(struct orc_picl *p, int ip) {
  const double eps = F32(1.0e-12);
  int inear[27];
  double d2[27];
  const double *xp = p->y + LRS * (size_t)ip;
  int nn = nearest27(p, xp, inear, d2);
  if (nn < 1) return; /* unreachable for located particles */
  for (int f = 0; f < ORC_PICL_NFLD; f++) {
    int j = orc_picl_field_rprop[f] - 1;
    double v = 0.0, wsum = 0.0;
    for (int k = 0; k < nn; k++) {
      double w = 1.0 / (sqrt(d2[k]) + eps);
      v = v + w * p->fld[ORC_PICL_NFLD * (size_t)inear[k] + f];
      wsum = wsum + w;
    }
    p->rprop[LRP * (size_t)ip + j] = v / wsum;
  }
  if (p->fldRot)
    for (int f = 0; f < 6; f++) {
      double v = 0.0, wsum = 0.0;
      for (int k = 0; k < nn; k++) {
        double w = 1.0 / (sqrt(d2[k]) + eps);
        v = v + w * p->fldRot[6 * (size_t)inear[k] + f];
        wsum = wsum + w;
      }
      p->rprop[LRP * (size_t)ip + ROT_RPROP[f]] = v / wsum;
    }
}

/* Trilinear mode: gslib findpts_eval_local, the evaluation ppiclF made before
 * LocalInterp (ppiclf_solve.f:2198-2208), on the element's 2x2x2 nodal field,
 * which PICL_TEMP_Runge fills with the cell value on every node.  Lagrange
 * weights on the nodes -1, 1: h0 = (1-r)/2, h1 = (1+r)/2; contracted along r,
 * then s, then t.  (r,s,t) is the location rule's Newton solution, not
 * gslib's. */
static void trilinear_interp(struct orc_picl *p, int ip) {
  const double *xp = p->y + LRS * (size_t)ip;
  int e = p->elem[ip];
  double rst[3];
  trilinear_inside(p->xm1b + 24 * (size_t)e, xp, rst);
  double h[3][2];
  for (int d = 0; d < 3; d++) {
    h[d][0] = 0.5 * (1.0 - rst[d]);
    h[d][1] = 0.5 * (1.0 + rst[d]);
  }
  for (int f = 0; f < ORC_PICL_NFLD; f++) {
    double v = p->fld[ORC_PICL_NFLD * (size_t)e + f];
    double u = h[0][0] * v + h[0][1] * v; /* every (j,k) line holds the same values */
    double w = h[1][0] * u + h[1][1] * u;
    p->rprop[LRP * (size_t)ip + orc_picl_field_rprop[f] - 1] = h[2][0] * w + h[2][1] * w;
  }
  if (p->fldRot)
    for (int f = 0; f < 6; f++) {
      double v = p->fldRot[6 * (size_t)e + f];
      double u = h[0][0] * v + h[0][1] * v;
      double w = h[1][0] * u + h[1][1] * u;
      p->rprop[LRP * (size_t)ip + ROT_RPROP[f]] = h[2][0] * w + h[2][1] * w;
    }
}

/* ------------------------------------------------------------------------ */
/* InitSolve: locate, interpolate, project the map built from the current    */
/* ydotc, zero ydotc.  ppiclf_solve.f:1954-1992; map:                        */
/* ppiclf_user_MapProjPart.f:37-66 via ppiclf_comm.f:1190-1217; deposit:     */
/* ppiclf_solve.f:3285-3394 (sngl_elem); gather :3431-3447.                  */
/* ------------------------------------------------------------------------ */
static void init_solve(struct orc_picl *p, int interp) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int ip = 0; ip < p->n; ip++) {
    if (p->lost[ip]) continue;
    if (p->mark && p->mark[ip]) { /* RemoveParticle: marked in the last stage (iprop 3), gone now */
      p->lost[ip] = 1;
      p->mark[ip] = 0;
      continue;
    }
    int e = locate(p, p->y + LRS * (size_t)ip);
    p->elem[ip] = e;
    if (e < 0) p->lost[ip] = 1;
  }
  if (interp)
#pragma omp parallel for schedule(dynamic, 64)
    for (int ip = 0; ip < p->n; ip++)
      if (!p->lost[ip]) {
        if (p->pp.interpMode == 1)
          trilinear_interp(p, ip);
        else
          local_interp(p, ip);
      }
  double *fldb = calloc(8 * NPRO * (size_t)p->ne, sizeof(double));
  double *map = malloc(sizeof(double) * NPRO * (size_t)(p->n > 0 ? p->n : 1));
  for (int ip = 0; ip < p->n; ip++) {
    const double *rp = p->rprop + LRP * (size_t)ip, *yc = p->ydotc + LRS * (size_t)ip;
    double *m = map + NPRO * (size_t)ip;
    m[P_JPHIP] = rp[R_JVOLP] * rp[R_JSPL];
    m[P_JFX] = yc[JVX] * rp[R_JSPL];
    m[P_JFY] = yc[JVY] * rp[R_JSPL];
    m[P_JFZ] = yc[JVZ] * rp[R_JSPL];
    m[P_JE] = yc[JT] * rp[R_JSPL];
  }
  if (p->pp.projMode == 1) {
    /* Gaussian filter, ppiclf_solve.f:3103-3109, 3118-3130, 3196-3246 with
     * iwallm = 0: every element node sums the particles within the cut-off.
     * On a periodic axis the particles' images shifted by +-(hi - lo) count
     * too: ppiclf_comm_CreateGhost (ppiclf_comm.f:1115-1530) makes them as
     * ghost particles, shifted by ppiclf_comm_CheckPeriodicBC (:2136-2206).
     * The sums are compensated (Neumaier), so that their value does not
     * depend on the order of the terms (ppiclF's order over its bins and
     * ghost particles is not restated); here: ascending particle index. */
    /* the images that exist: shift -1, 0, +1 periods on periodic axes,
     * 0 elsewhere, in ascending image order (img = sx+1 + 3(sy+1) + 9(sz+1)) */
    int nimg = 0, imgSd[27][3];
    for (int img = 0; img < 27; img++) {
      int skip = 0;
      for (int d = 0, q = img; d < 3; d++, q /= 3) {
        imgSd[nimg][d] = q % 3 - 1;
        if (imgSd[nimg][d] != 0 && !p->per[d]) skip = 1;
      }
      if (!skip) nimg++;
    }
#pragma omp parallel for schedule(dynamic, 8)
    for (int e = 0; e < p->ne; e++)
      for (int n = 0; n < 8; n++) {
        const double *xn = p->xm1b + 24 * (size_t)e + 3 * n;
        double *nb = fldb + 8 * NPRO * (size_t)e;
        double cmp[NPRO] = {0.0};
        for (int ip = 0; ip < p->n; ip++) {
          if (p->lost[ip]) continue;
          const double *y = p->y + LRS * (size_t)ip;
          for (int im = 0; im < nimg; im++) {
            double xp[3];
            for (int d = 0; d < 3; d++) {
              int sd = imgSd[im][d];
              xp[d] = sd == 0 ? y[d] : y[d] + (double)sd * (p->perHi[d] - p->perLo[d]);
            }
            double rdist2 = (xn[0] - xp[0]) * (xn[0] - xp[0]) + (xn[1] - xp[1]) * (xn[1] - xp[1]);
            rdist2 = rdist2 + (xn[2] - xp[2]) * (xn[2] - xp[2]);
            if (rdist2 > p->gD2chk) continue;
            double rexp = exp(rdist2 * p->gRdum);
            for (int jj = 0; jj < NPRO; jj++) {
              double x = (map[NPRO * (size_t)ip + jj] * p->gMult) * rexp, s = nb[8 * jj + n], t = s + x;
              if (fabs(s) >= fabs(x))
                cmp[jj] = cmp[jj] + ((s - t) + x);
              else
                cmp[jj] = cmp[jj] + ((x - t) + s);
              nb[8 * jj + n] = t;
            }
          }
        }
        for (int jj = 0; jj < NPRO; jj++) nb[8 * jj + n] = nb[8 * jj + n] + cmp[jj];
      }
  } else {
    for (int ip = 0; ip < p->n; ip++) {
      if (p->lost[ip]) continue;
      int e = p->elem[ip];
      double rexp = p->rexp[e];
      for (int jj = 0; jj < NPRO; jj++)
        for (int n = 0; n < 8; n++)
          fldb[8 * (NPRO * (size_t)e + jj) + n] =
              fldb[8 * (NPRO * (size_t)e + jj) + n] + (map[NPRO * (size_t)ip + jj] * 1.0) * rexp;
    }
  }
  for (size_t q = 0; q < 8 * NPRO * (size_t)p->ne; q++) p->proFld[q] = 0.0 + fldb[q];
  free(map);
  free(fldb);
  memset(p->ydotc, 0, sizeof(double) * LRS * (size_t)p->n);
}

/* ------------------------------------------------------------------------ */
/* Forces                                                                    */
/* ------------------------------------------------------------------------ */
/* ppiclf_user_QS_Parmar: libpicl/user_files/ppiclf_user_QuasiSteady.f:22-122 */
static double parmar_rcd1(double mp, double re) {
  double rcd1;
  if (re < F32(1E-14)) {
    rcd1 = 1.0;
  } else {
    double rmacr = F32(0.6);
    double rcd_mcr = (1.0 + F32(0.15) * pow(re, F32(0.684))) +
                     (re / 24.0) * (F32(0.513) / (1.0 + 483.0 / pow(re, F32(0.669))));
    if (mp <= rmacr) {
      double rcd_std = (1.0 + F32(0.15) * pow(re, F32(0.687))) +
                       (re / 24.0) * (F32(0.42) / (1.0 + 42500.0 / pow(re, F32(1.16))));
      double rmach_rat = mp / rmacr;
      rcd1 = rcd_std + (rcd_mcr - rcd_std) * rmach_rat;
    } else if (mp <= 1.0) {
      double rcd_M1 = (1.0 + F32(0.118) * pow(re, F32(0.813))) +
                      (re / 24.0) * F32(0.69) / (1.0 + 3550.0 / pow(re, F32(.793)));
      double C1 = F32(6.48), C2 = F32(9.28), C3 = F32(12.21);
      double m2 = mp * mp, m3 = mp * mp * mp;
      double f1M = -F32(1.884) + F32(8.422) * mp - F32(13.70) * m2 + F32(8.162) * m3;
      double f2M = -F32(2.228) + F32(10.35) * mp - F32(16.96) * m2 + F32(9.840) * m3;
      double f3M = F32(4.362) - F32(16.91) * mp + F32(19.84) * m2 - F32(6.296) * m3;
      double lrep = log(re);
      double factor = f1M * (lrep - C2) * (lrep - C3) / ((C1 - C2) * (C1 - C3)) +
                      f2M * (lrep - C1) * (lrep - C3) / ((C2 - C1) * (C2 - C3)) +
                      f3M * (lrep - C1) * (lrep - C2) / ((C3 - C1) * (C3 - C2));
      rcd1 = rcd_mcr + (rcd_M1 - rcd_mcr) * factor;
    } else if (mp < 1.75) {
      double rcd_M1 = (1.0 + F32(0.118) * pow(re, F32(0.813))) +
                      (re / 24.0) * F32(0.69) / (1.0 + 3550.0 / pow(re, F32(.793)));
      double rcd_M2 = (1.0 + F32(0.107) * pow(re, F32(0.867))) +
                      (re / 24.0) * F32(0.646) / (1.0 + 861.0 / pow(re, F32(.634)));
      double C1 = F32(6.48), C2 = F32(8.93), C3 = F32(12.21);
      double m2 = mp * mp, m3 = mp * mp * mp;
      double f1M = -F32(2.963) + F32(4.392) * mp - F32(1.169) * m2 - F32(0.027) * m3 -
                   F32(0.233) * exp((1.0 - mp) / F32(0.011));
      double f2M = -F32(6.617) + F32(12.11) * mp - F32(6.501) * m2 + F32(1.182) * m3 -
                   F32(0.174) * exp((1.0 - mp) / F32(0.010));
      double f3M = -F32(5.866) + F32(11.57) * mp - F32(6.665) * m2 + F32(1.312) * m3 -
                   F32(0.350) * exp((1.0 - mp) / F32(0.012));
      double lrep = log(re);
      double factor = f1M * (lrep - C2) * (lrep - C3) / ((C1 - C2) * (C1 - C3)) +
                      f2M * (lrep - C1) * (lrep - C3) / ((C2 - C1) * (C2 - C3)) +
                      f3M * (lrep - C1) * (lrep - C2) / ((C3 - C1) * (C3 - C2));
      rcd1 = rcd_M1 + (rcd_M2 - rcd_M1) * factor;
    } else {
      rcd1 = (1.0 + F32(0.107) * pow(re, F32(0.867))) +
             (re / 24.0) * F32(0.646) / (1.0 + 861.0 / pow(re, F32(.634)));
    }
  }
  return rcd1;
}

static double qs_parmar(double rmachp, double rphip, double rep, double rmu, double dp) {
  const double rpi = acos(-1.0);
  double mp = fmax(rmachp, 0.01), re = fmax(rep, 0.1);
  double rcd1 = parmar_rcd1(mp, re);
  double phi_corr = 1.0 + F32(5.94) * fmin(rphip, 0.5);
  double beta = rcd1 * 3.0 * rpi * rmu * dp;
  return beta * phi_corr;
}

/* Integer powers as gfortran expands them (x**2 = x*x, x**3 = x2*x,
 * x**4 = x2*x2, x**5 = x2*x3, x**(-3) = 1/x**3). */
static double ip2(double x) { return x * x; }
static double ip3(double x) { return (x * x) * x; }
static double ip4(double x) { double x2 = x * x; return x2 * x2; }
static double ip5(double x) { double x2 = x * x; return x2 * (x2 * x); }

/* Osnes et al. volume-fraction terms, ppiclf_user_QuasiSteady.f:208-218, 346-356 */
static void osnes_b(double phi, double re, double mp, double *b1, double *b2, double *b3) {
  double om = 1.0 - phi;
  *b1 = F32(5.81) * phi / ip2(om) + F32(0.48) * pow(phi, 1.0 / 3.0) / ip3(om);
  *b2 = ip2(om) * ip3(phi) * re * (F32(0.95) + F32(0.61) * ip3(phi) / (om * 2.0));
  *b3 = fmin(sqrt(20.0 * mp), 1.0) * (F32(5.65) * phi - 22.0 * ip2(phi) + F32(23.4) * ip3(phi)) *
        (1.0 + tanh((mp - (F32(0.65) - F32(0.24) * phi)) / F32(0.35)));
}

/* Parmar's rcd1 (Re and Mach corrections), shared by QS_Parmar and
 * QS_ModifiedParmar: ppiclf_user_QuasiSteady.f:43-100, 151-205 */
/* ppiclf_user_QS_ModifiedParmar: ppiclf_user_QuasiSteady.f:125-230 */
static double qs_modified_parmar(double rmachp, double rphip, double rep, double rmu, double dp) {
  const double rpi = acos(-1.0);
  double mp = fmax(rmachp, 0.01), phi = fmax(rphip, 0.0001), re = fmax(rep, 0.1);
  double rcd1 = parmar_rcd1(mp, re);
  double b1, b2, b3;
  osnes_b(phi, re, mp, &b1, &b2, &b3);
  double cd = (24.0 / re) * rcd1;
  cd = cd / (1.0 - phi) + b3 + (24.0 / re) * (1.0 - phi) * (b1 + b2);
  return 3.0 * rpi * rmu * dp * (re / 24.0) * cd;
}

/* ppiclf_user_QS_Osnes (Loth et al. 2021 + Osnes et al. 2023):
 * ppiclf_user_QuasiSteady.f:233-362 */
static double qs_osnes(double rmachp, double rphip, double rep, double rmu, double dp) {
  const double rpi = acos(-1.0), gamma = 1.4;
  double mp = fmax(rmachp, 0.01), phi = fmax(rphip, 0.0001), re = fmax(rep, 0.1);
  double cd_loth;
  if (re <= 45.0) {
    double Knp = sqrt(0.5 * rpi * gamma) * mp / re, fKn;
    if (Knp > F32(0.01))
      fKn = 1.0 / (1.0 + Knp * (2.514 + 0.8 * exp(-0.55 / Knp)));
    else
      fKn = 1.0 / (1.0 + Knp * (2.514 + 0.8 * exp(-0.55 / F32(0.01))));
    double CD1 = (24.0 / re) * (1.0 + 0.15 * pow(re, 0.687)) * fKn;
    double s = mp * sqrt(0.5 * gamma);
    double sgby2 = sqrt(0.5 * gamma);
    double JMt;
    if (mp <= 1.0)
      JMt = 2.26 * ip4(mp) + 0.14 * mp;
    else
      JMt = 1.6 * ip4(mp) + 0.25 * ip3(mp) + 0.11 * ip2(mp) + 0.44 * mp;
    double CD2 = (1.0 + 2.0 * ip2(s)) * exp(-ip2(s)) * mp / (ip3(sgby2) * sqrt(rpi)) +
                 (4.0 * ip4(s) + 4.0 * ip2(s) - 1.0) * erf(s) / (2.0 * ip4(sgby2)) +
                 (2.0 * ip3(mp) / (3.0 * sgby2)) * sqrt(rpi);
    CD2 = CD2 / (1.0 + (((CD2 / JMt) - 1.0) * sqrt(re / 45.0)));
    cd_loth = CD1 / (1.0 + ip4(mp)) + CD2 / (1.0 + ip4(mp));
  } else {
    double CM, GM, HM;
    if (mp < 1.5)
      CM = 1.65 + 0.65 * tanh(4.0 * mp - 3.4);
    else
      CM = 2.18 - 0.12913149918318745 * tanh(0.9 * mp - 2.7);
    if (mp < F32(0.8))
      GM = 166.0 * ip3(mp) + 3.29 * ip2(mp) - 10.9 * mp + 20.0;
    else
      GM = 5.0 + 47.809331200000017 * (1.0 / ip3(mp));
    if (mp < 1.0)
      HM = 0.0239 * ip3(mp) + 0.212 * ip2(mp) - 0.074 * mp + 1.0;
    else
      HM = 0.93967777777777772 + 1.0 / (3.5 + ip5(mp));
    cd_loth = (24.0 / re) * (1.0 + F32(0.15) * pow(re, F32(0.687))) * HM +
              F32(0.42) * CM / (1.0 + 42500.0 / pow(re, F32(1.16) * CM) + GM / sqrt(re));
  }
  double b1, b2, b3;
  osnes_b(phi, re, mp, &b1, &b2, &b3);
  double cd = cd_loth / (1.0 - phi) + b3 + (24.0 / re) * (1.0 - phi) * (b1 + b2);
  return 3.0 * rpi * rmu * dp * (re / 24.0) * cd;
}

/* ppiclf_user_QS_Gidaspow (Ergun / Wen-Yu, per particle):
 * ppiclf_user_QuasiSteady.f:364-397 */
static double qs_gidaspow(double rphip, double rphif, double rep, double rmu, double dp, double rhof,
                          double vmag) {
  const double rpi = acos(-1.0);
  double phi = fmax(rphip, 0.0001), phif = fmax(rphif, 0.0001), re = fmax(rep, 0.1);
  double phifRep = phif * re, cd, beta;
  if (phifRep < 1000.0)
    cd = 24.0 / phifRep * (1.0 + F32(0.15) * pow(phifRep, F32(0.687)));
  else
    cd = F32(0.44);
  if (phif < F32(0.8))
    beta = 150.0 * (ip2(phi) * rmu) / (phif * ip2(dp)) + 1.75 * (rhof * phi * vmag / dp);
  else
    beta = F32(0.75) * cd * phi * rhof * vmag / (dp * pow(phif, F32(1.65)));
  return beta * (rpi * ip3(dp)) / (6.0 * phi);
}

/* Schiller-Naumann: rocinteract's INRT_DRAG_MODEL_SN correction
 * psi = 1 + 0.15 Re^0.687 (rocinteract/INRT_CalcDrag.F90:189-191) in
 * ppiclF's quasi-steady slot, beta = psi * 3 pi mu d_p (the form of
 * ppiclf_user_QuasiSteady.f:113); no volume-fraction correction.  Not a
 * ppiclF model: SetYdot rejects qs_flag 5. */
static double qs_sn(double rep, double rmu, double dp) {
  const double rpi = acos(-1.0);
  double re = fmax(rep, 0.1);
  double psi = 1.0 + 0.15 * pow(re, 0.687);
  return psi * 3.0 * rpi * rmu * dp;
}

double orc_picl_qs_beta(int qsFlag, double rmachp, double rphip, double rep, double rmu, double dp,
                        double rhof, double vmag) {
  switch (qsFlag) {
    case 1: return qs_parmar(rmachp, rphip, rep, rmu, dp);
    case 2: return qs_osnes(rmachp, rphip, rep, rmu, dp);
    case 3: return qs_modified_parmar(rmachp, rphip, rep, rmu, dp);
    case 4: return qs_gidaspow(rphip, 1.0 - rphip, rep, rmu, dp, rhof, vmag);
    case 5: return qs_sn(rep, rmu, dp);
  }
  return 0.0;
}

/* ppiclf_user_HT_driver (libpicl/user_files/ppiclf_user_HeatTransfer.f:17-222):
 * Q = pi kappa d (T_gas - T_p) Nu with the Stokes, Ranz-Marshall, Gunn or Fox
 * Nusselt law; ppiclF fixes Pr = 0.70 and cp_gas = 1004.64
 * (ppiclf_user.f:177-178, 310). */
static double heat_transfer(int flag, double rkappa, double dp, double tgas, double tpart, double rep,
                            double rphif, double rmachp) {
  const double rpi = 3.141592653589793116, rpr = 0.70, OneThird = 1.0 / 3.0;
  double Q_conv = rpi * rkappa * dp * (tgas - tpart);
  double Nuss = 0.0;
  switch (flag) {
    case 1: Nuss = 2.0; break;
    case 2: Nuss = 2.0 + 0.6 * pow(rep, 0.5) * pow(rpr, OneThird); break;
    case 3: {
      double vg = rphif;
      Nuss = (7.0 - 10.0 * vg + 5.0 * vg * vg) * (1.0 + 0.7 * pow(rep, 0.2) * pow(rpr, OneThird)) +
             (1.33 - 2.4 * vg + 1.2 * vg * vg) * pow(rep, 0.7) * pow(rpr, OneThird);
      break;
    }
    case 4:
      Nuss = 2.0 * exp(-rmachp) / (1.0 + 17.0 * rmachp / rep) +
             0.495 * pow(rpr, OneThird) * pow(rep, 0.55) * ((1.0 + 0.5 * exp(-17.0 * rmachp / rep)) / 1.5);
      break;
  }
  return 0.0 + Q_conv * Nuss;
}

/* ------------------------------------------------------------------------ */
/* Viscous-unsteady force: ppiclf_user_ViscousUnsteady.f (PPICLF_VU = 5      */
/* history entries, PPICLF_USER.h:12)                                        */
/* ------------------------------------------------------------------------ */
#define VU_N 5

/* ppiclf_user_VU_Rocflu (:19-90): the history kernel summed over entries
 * 2..nTimeBH-1, half weight on the oldest one; entry 1 (this stage's) is
 * not used.  (A+B)**(-2) as gfortran expands it, 1/((A+B)*(A+B)). */
static void vu_sum(int nTimeBH, const double *timeBH, const double *mixt, const double *plag, double dt, double rmu,
                   double rhof, double dp, double vmag, double rphip, double *fvu) {
  const double rpi = acos(-1.0);
  double reyL = dp * vmag * rhof / rmu; /* ppiclf_user.f:365-366 */
  double rnu = rmu / rhof;
  double fH = 0.75 + 0.105 * reyL;
  double factor = 3.0 * rpi * rnu * dp * dt * (1.0 + F32(2.28) * fmin(rphip, 0.5));
  fvu[0] = fvu[1] = fvu[2] = 0.0;
  if (nTimeBH > 1) {
    for (int iT = 2; iT <= nTimeBH; iT++) {
      double time = timeBH[iT - 1];
      double A = pow(4.0 * rpi * time * rnu / ip2(dp), 0.25);
      double B = pow(0.5 * rpi * ip3(vmag) * ip2(time) / (0.5 * dp * rnu * ip3(fH)), 0.5);
      double kernelVU = (iT < nTimeBH ? factor : 0.5 * factor) * (1.0 / ((A + B) * (A + B)));
      for (int c = 0; c < 3; c++) fvu[c] = fvu[c] + kernelVU * (mixt[c * VU_N + iT - 1] - plag[c * VU_N + iT - 1]);
    }
  }
}

static void vu_force(const struct orc_picl *p, int i, double dt, double rmu, double rhof, double dp, double vmag,
                     double rphip, double *fvu) {
  const double *mixt = p->hist + 6 * VU_N * (size_t)i, *plag = mixt + 3 * VU_N;
  vu_sum(p->nTimeBH, p->timeBH, mixt, plag, dt, rmu, rhof, dp, vmag, rphip, fvu);
}

void orc_picl_vu_force(int nTimeBH, const double *timeBH, const double *mixt, const double *plag, double dt,
                       double rmu, double rhof, double dp, double vmag, double rphip, double *fvu) {
  if (nTimeBH < 1 || nTimeBH > VU_N) { fvu[0] = fvu[1] = fvu[2] = 0.0; return; }
  vu_sum(nTimeBH, timeBH, mixt, plag, dt, rmu, rhof, dp, vmag, rphip, fvu);
}

int orc_picl_vu_state(const orc_region *r, double *timeBH, double *hist) {
  const struct orc_picl *p = r->picl;
  if (!p) return -1;
  if (timeBH) memcpy(timeBH, p->timeBH, sizeof p->timeBH);
  if (hist) memcpy(hist, p->hist, sizeof(double) * 6 * VU_N * (size_t)p->n);
  return p->nTimeBH;
}

/* ppiclf_user_UpdatePlag (:146-206): this stage's d(rho u)/dt of the gas
 * (material derivative of rho_g u_g along the particle) and of the particle,
 * into entry 1 */
static void vu_update(struct orc_picl *p, int i, const double *rp, const double *y, const double *yd, double vx,
                      double vy, double vz, double rhof, double rphif) {
  double *mixt = p->hist + 6 * VU_N * (size_t)i, *plag = mixt + 3 * VU_N;
  double SDrho = rp[R_JRHSR] + y[JVX] * rp[R_JPGCX] + y[JVY] * rp[R_JPGCY] + y[JVZ] * rp[R_JPGCZ];
  SDrho = SDrho / rphif;
  double vgradrho = vx * rp[R_JRHOGX] + vy * rp[R_JRHOGY] + vz * rp[R_JRHOGZ];
  const double u[3] = {rp[R_JUX], rp[R_JUY], rp[R_JUZ]};
  const double sdr[3] = {rp[R_JSDRX], rp[R_JSDRY], rp[R_JSDRZ]};
  for (int c = 0; c < 3; c++) {
    mixt[c * VU_N] = u[c] * (SDrho + vgradrho) + rhof * sdr[c];
    plag[c * VU_N] = y[JVX + c] * SDrho + rhof * yd[JVX + c];
  }
}

/* ppiclf_user_ShiftUnsteadyData (:93-143), after the third stage's SetYdot
 * (ppiclf_user.f:1030-1036): entries move one older, the count grows to
 * VU_N, and each entry's age is its younger neighbour's plus dt */
static void vu_shift(struct orc_picl *p, double dt) {
  for (int i = 0; i < p->n; i++) {
    double *h = p->hist + 6 * VU_N * (size_t)i;
    for (int c = 0; c < 6; c++)
      for (int k = VU_N - 1; k >= 1; k--) h[c * VU_N + k] = h[c * VU_N + k - 1];
  }
  if (p->nTimeBH < VU_N) p->nTimeBH = p->nTimeBH + 1;
  for (int iT = p->nTimeBH; iT >= 2; iT--) p->timeBH[iT - 1] = p->timeBH[iT - 2] + dt;
}

/* ------------------------------------------------------------------------ */
/* Collisions (COLLISIONS 1): ppiclf_user_EvalNearestNeighbor.f:118-233      */
/* ------------------------------------------------------------------------ */
/* The normal soft-sphere force of neighbour j on particle i when they
 * overlap (Cundall and Strack): spring min(rmass_i pi^2 / (ksp dt)^2, Hertz
 * stiffness of the overlap), damping from the restitution coefficient;
 * tangential part zero (Ftmin = 0).  Added into fc as ppiclF adds it into
 * ydotc(JVX..JVZ, i). */
static void collision_pair(int flag, const double *yi, const double *rpi_, const double *yj, const double *rpj,
                           double rmass, double dt, double ksp, double erest, double *fc, double *tq) {
  const double rpi = acos(-1.0), pi2 = rpi * rpi;
  double rthresh = 0.5 * (rpi_[R_JDP] + rpj[R_JDP]);
  double rxdiff = yj[JX] - yi[JX], rydiff = yj[JY] - yi[JY], rzdiff = yj[JZ] - yi[JZ];
  double rdiff = ip2(rxdiff) + ip2(rydiff) + ip2(rzdiff);
  rdiff = sqrt(rdiff);
  const double eps = 0.0;
  if (!(rdiff < rthresh + eps)) return;
  double ksp1 = rmass * rpi * rpi / ip2(ksp * dt);
  const double E1 = 1.0e9, E2 = 1.0e9, nu1 = 0.35, nu2 = 0.35;
  double Estar = (1.0 - nu1 * nu1) / E1 + (1.0 - nu2 * nu2) / E2;
  Estar = 1.0 / Estar;
  double r1 = 0.5 * rpi_[R_JDP], r2 = 0.5 * rpj[R_JDP];
  double Rstar = r1 * r2 / (r1 + r2);
  double ksp2 = (4.0 / 3.0) * Estar * sqrt(Rstar);
  ksp2 = ksp2 * sqrt(fabs(rdiff - rthresh));
  double ksp_min = fmin(ksp1, ksp2);
  double rm1 = rpi_[R_JRHOP] * rpi_[R_JVOLP], rm2 = rpj[R_JRHOP] * rpj[R_JVOLP];
  double rmult = (rm1 * rm2) / (rm1 + rm2);
  double le = log(erest);
  double eta_n = -2.0 * sqrt(ksp_min) * le / sqrt(ip2(le) + pi2) * sqrt(rmult);
  double rbot = 1.0 / rdiff;
  double rn[3] = {rxdiff * rbot, rydiff * rbot, rzdiff * rbot};
  double u12[3] = {yi[JVX] - yj[JVX], yi[JVY] - yj[JVY], yi[JVZ] - yj[JVZ]};
  const double *oi = yi + 7, *oj = yj + 7; /* JOX..JOZ */
  if (flag >= 2) { /* contact-point velocity of the rotations */
    double rad1 = 0.5 * rpi_[R_JDP], rad2 = 0.5 * rpj[R_JDP];
    double A12[3];
    for (int c = 0; c < 3; c++) A12[c] = rad1 * oi[c] + rad2 * oj[c];
    u12[0] = u12[0] + (A12[1] * rn[2] - A12[2] * rn[1]);
    u12[1] = u12[1] + (A12[2] * rn[0] - A12[0] * rn[2]);
    u12[2] = u12[2] + (A12[0] * rn[1] - A12[1] * rn[0]);
  }
  double rv12_mag = u12[0] * rn[0] + u12[1] * rn[1] + u12[2] * rn[2];
  double rdelta12 = rthresh - rdiff;
  double rksp_max = ksp_min * rdelta12;
  double rv12_mage = rv12_mag * eta_n;
  double rnmag = -rksp_max - rv12_mage;
  double Fn_mag = fabs(rnmag);
  double ut[3];
  for (int c = 0; c < 3; c++) ut[c] = u12[c] - rv12_mag * rn[c];
  double ut_mag = sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
  ut_mag = fmax(ut_mag, 1.0e-8);
  double rt[3] = {ut[0] / ut_mag, ut[1] / ut_mag, ut[2] / ut_mag};
  double Ftmin = 0.0;
  if (flag >= 2 && ut_mag > 0.0) { /* Coulomb friction or tangential dashpot, the smaller */
    const double mu_c = 0.4;
    double eta_t = eta_n;
    Ftmin = -fmin(mu_c * Fn_mag, eta_t * ut_mag);
  }
  double tc[3] = {0.0, 0.0, 0.0}, tr[3] = {0.0, 0.0, 0.0};
  if (flag >= 2) {
    double Ft[3] = {Ftmin * rt[0], Ftmin * rt[1], Ftmin * rt[2]};
    double rad1 = 0.5 * rpi_[R_JDP];
    tc[0] = rad1 * (rn[1] * Ft[2] - rn[2] * Ft[1]);
    tc[1] = rad1 * (rn[2] * Ft[0] - rn[0] * Ft[2]);
    tc[2] = rad1 * (rn[0] * Ft[1] - rn[1] * Ft[0]);
    const double thetar = F32(0.06); /* rolling resistance, a real*4 literal */
    double dp1 = rpi_[R_JDP], dp2 = rpj[R_JDP];
    double r12 = 0.5 * (dp1 * dp2) / (dp1 + dp2);
    double omgr[3] = {oi[0] - oj[0], oi[1] - oj[1], oi[2] - oj[2]};
    double omgr_mag = sqrt(omgr[0] * omgr[0] + omgr[1] * omgr[1] + omgr[2] * omgr[2]);
    omgr_mag = fmax(omgr_mag, 1.0e-8);
    for (int c = 0; c < 3; c++) tr[c] = -thetar * Fn_mag * r12 * omgr[c] / omgr_mag;
  }
  for (int c = 0; c < 3; c++) fc[c] = fc[c] + rnmag * rn[c] + Ftmin * rt[c];
  for (int c = 0; c < 3; c++) tq[c] = tq[c] + tc[c] + tr[c];
}

void orc_picl_collision_force(const double *xi, const double *vi, double dpi, double rhopi, double volpi,
                              const double *xj, const double *vj, double dpj, double rhopj, double volpj, double dt,
                              double ksp, double erest, double *fc) {
  double yi[LRS] = {0}, yj[LRS] = {0}, ri[LRP] = {0}, rj[LRP] = {0};
  for (int c = 0; c < 3; c++) {
    yi[JX + c] = xi[c];
    yi[JVX + c] = vi[c];
    yj[JX + c] = xj[c];
    yj[JVX + c] = vj[c];
  }
  ri[R_JDP] = dpi;
  ri[R_JRHOP] = rhopi;
  ri[R_JVOLP] = volpi;
  rj[R_JDP] = dpj;
  rj[R_JRHOP] = rhopj;
  rj[R_JVOLP] = volpj;
  double tq[3] = {0.0, 0.0, 0.0};
  fc[0] = fc[1] = fc[2] = 0.0;
  collision_pair(1, yi, ri, yj, rj, volpi * rhopi, dt, ksp, erest, fc, tq);
}

void orc_picl_collision_force2(const double *yi12, const double *yj12, double dpi, double rhopi, double volpi,
                               double dpj, double rhopj, double volpj, double dt, double ksp, double erest,
                               double *fc, double *tq) {
  double ri[LRP] = {0}, rj[LRP] = {0};
  ri[R_JDP] = dpi;
  ri[R_JRHOP] = rhopi;
  ri[R_JVOLP] = volpi;
  rj[R_JDP] = dpj;
  rj[R_JRHOP] = rhopj;
  rj[R_JVOLP] = volpj;
  fc[0] = fc[1] = fc[2] = 0.0;
  tq[0] = tq[1] = tq[2] = 0.0;
  collision_pair(2, yi12, ri, yj12, rj, volpi * rhopi, dt, ksp, erest, fc, tq);
}

/* ppiclf_solve_NearestNeighbor (ppiclf_solve.f:470-560) for the collision
 * force: the other particles in ascending index, then the ghost particles,
 * here the periodic images, by particle and then image. */
static void collision_sum(const struct orc_picl *p, int i, double rmass, double dt, double *fc, double *tq) {
  const double *yi = p->y + LRS * (size_t)i, *rpi_ = p->rprop + LRP * (size_t)i;
  const orc_picl_params *P = &p->pp;
  fc[0] = fc[1] = fc[2] = 0.0;
  tq[0] = tq[1] = tq[2] = 0.0;
  const int flag = P->collisionalFlag;
  for (int j = 0; j < p->n; j++) {
    if (j == i || p->lost[j]) continue;
    collision_pair(flag, yi, rpi_, p->y + LRS * (size_t)j, p->rprop + LRP * (size_t)j, rmass, dt, P->ksp, P->erest,
                   fc, tq);
  }
  if (!(p->per[0] || p->per[1] || p->per[2])) return;
  for (int j = 0; j < p->n; j++) {
    if (p->lost[j]) continue;
    for (int img = 0; img < 27; img++) {
      if (img == 13) continue; /* the particle itself */
      double yj[LRS];
      memcpy(yj, p->y + LRS * (size_t)j, sizeof yj);
      int skip = 0;
      for (int d = 0, q = img; d < 3; d++, q /= 3) {
        int sd = q % 3 - 1;
        if (sd != 0 && !p->per[d]) skip = 1;
        if (sd != 0) yj[d] = yj[d] + (double)sd * (p->perHi[d] - p->perLo[d]);
      }
      if (skip) continue;
      collision_pair(flag, yi, rpi_, yj, p->rprop + LRP * (size_t)j, rmass, dt, P->ksp, P->erest, fc, tq);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Briney added mass (AMFORCE 2): libpicl/user_files/Added_Mass/            */
/* ppiclf_user_AddedMass_functions.f, _subroutines.f; the unary and binary  */
/* terms of ppiclf_user_AddedMass.f:164-400; the pair terms of              */
/* ppiclf_user_EvalNearestNeighbor.f:115-159.                               */
/* Integer powers as gcc expands them: x**5 = x^3 x^2, x**6 = x^3 x^3,      */
/* x**8 = x^4 x^4, x**9 = x^6 x^3 (x^3 = x^2 x, x^4 = x^2 x^2).             */
/* ------------------------------------------------------------------------ */
static double ip6(double x) { double x3 = (x * x) * x; return x3 * x3; }
static double ip8(double x) { double x2 = x * x, x4 = x2 * x2; return x4 * x4; }
static double ip9(double x) { double x3 = (x * x) * x, x6 = x3 * x3; return x6 * x3; }

/* the four fitted curves of the binary model (_functions.f:48-148) */
double orc_am_b11_11(double d, double alpha) {
  double t = 2.0 / d;
  double asym = 0.5 * ((double)(3.0f / 64.0f) * ip6(t) + (double)(9.0f / 256.0f) * ip8(t));
  double hot = F32(0.059) / (((d - F32(1.9098)) * ip2(d - F32(0.4782))) * ip3(d));
  double alpha_corr = ((alpha * alpha - F32(0.0902) * alpha) * F32(70.7731)) / ip6(d + F32(2.0936));
  return (asym + hot) + alpha_corr;
}

double orc_am_b11_22(double d, double alpha) {
  double t = 2.0 / d;
  double asym = 0.5 * ((double)(3.0f / 256.0f) * ip6(t) + (double)(3.0f / 256.0f) * ip8(t));
  double A = (F32(0.0003) + (F32(0.0262) * alpha) * alpha) - F32(0.0012) * alpha;
  double B = (F32(1.3127) + (F32(1.0401) * alpha) * alpha) - F32(1.2519) * alpha;
  double hot = A / ip6(d - B);
  return asym + hot;
}

double orc_am_b12_11(double d, double alpha) {
  double t = 2.0 / d;
  double asym = 0.5 * (-((double)(3.0f / 8.0f) * ip3(t)) - (double)(3.0f / 512.0f) * ip9(t));
  double hot = -(F32(0.0006) / ip5(d - F32(1.5428)));
  double alpha_corr = -((F32(0.7913) * alpha) * exp(-((F32(0.9801) - F32(0.1075) * alpha) * d)));
  return (asym + hot) + alpha_corr;
}

double orc_am_b12_22(double d, double alpha) {
  double t = 2.0 / d;
  double asym = 0.5 * ((double)(3.0f / 16.0f) * ip3(t) + (double)(3.0f / 4096.0f) * ip9(t));
  double hot = (-F32(0.1985) + F32(16.7372) * alpha) / ip6(d) + (F32(1.3907) - F32(48.2604) * alpha) / ip8(d);
  return asym + hot;
}

/* distant-neighbour corrections with g(r) = 1 (_functions.f:160-245) */
double orc_am_ia_analytical(double rmax, double rad, double alpha) {
  double m = rmax / rad;
  double term1 = (5.0 * ip2(m) + 9.0) / (10.0 * ip5(m));
  double l1 = log(m - F32(1.9098)), l2 = log(m - F32(0.4782));
  double num = (((F32(0.00720826) - F32(0.0150737) * m) * l1 - (F32(0.120023) * m) * l2) + F32(0.057395) * l2 +
                (F32(0.135097) * m - F32(0.0646033)) * log(m)) - F32(0.0861828);
  double term2 = num / (m - F32(0.4782));
  double term3 = ((ip2(alpha) - F32(0.0902) * alpha) * ((F32(0.333333) * ip2(m) + F32(0.348933) * m) + F32(0.146105))) /
                 ip5(m + F32(2.0936));
  double B11 = (term1 + term2) + term3;
  term1 = (5.0 * ip2(m) + 12.0) / (40.0 * ip5(m));
  double numerator = (F32(0.0262) * ip2(alpha) - F32(0.0012) * alpha) + F32(0.0003);
  double c = (-(F32(1.0401) * ip2(alpha)) + F32(1.2519) * alpha) - F32(1.3127);
  term2 = (numerator * ((10.0 * ip2(m) + (5.0 * m) * c) + ip2(c))) / (30.0 * ip5(m + c));
  double B22 = term1 + term2;
  return alpha * (B11 + 2.0 * B22);
}

double orc_am_ii_analytical(double rmax, double rad, double alpha) {
  double m = rmax / rad;
  double term1 = -(1.0 / (4.0 * ip6(m)));
  double term2 = -((F32(6.0e-4) * ((0.5 * ip2(m) - F32(0.516067) * m) + F32(0.199744))) / ip4(F32(1.5482) - m));
  double prefactor = -(F32(0.7913) * alpha);
  double c = F32(0.9801) - F32(0.1075) * alpha;
  double term3 = ((prefactor * ((m * c) * ((m * c) + 2.0) + 2.0)) * exp(-(m * c))) / ip3(c);
  double B11 = (term1 + term2) + term3;
  term1 = 1.0 / (32.0 * ip6(m));
  term2 = (-F32(0.1985) + F32(16.7372) * alpha) / (3.0 * ip3(m));
  term3 = (F32(1.3907) - F32(48.2604) * alpha) / (5.0 * ip5(m));
  double B22 = (term1 + term2) + term3;
  return alpha * (B11 + 2.0 * B22);
}

/* radial distribution function of hard spheres (Trokhymchuk et al. 2005,
 * with its erratum), r = distance / diameter (_functions.f:262-347) */
double orc_am_rdf(double r, double alpha) {
  double eta = alpha;
  double d = pow((2.0 * eta) * (((eta * eta - 3.0 * eta) - 3.0) +
                                sqrt(3.0 * ((((ip4(eta) - 2.0 * ip3(eta)) + ip2(eta)) + 6.0 * eta) + 3.0))),
                 (double)(1.0f / 3.0f));
  double f = (2.0 * eta) / (1.0 - eta);
  double mu = f * ((-1.0 - d / (2.0 * eta)) + eta / d);
  double alpha0 = f * ((-1.0 + d / (4.0 * eta)) - eta / (2.0 * d));
  double beta0 = (f * (double)sqrtf(3.0f)) * (-(d / (4.0 * eta)) - eta / (2.0 * d));
  double a2b2 = ip2(alpha0) + ip2(beta0);
  double denom_gamma = ((a2b2 - (2.0 * mu) * alpha0) * (1.0 + 0.5 * eta)) - mu * (1.0 + 2.0 * eta);
  double X = (alpha0 * a2b2 - mu * (ip2(alpha0) - ip2(beta0))) * (1.0 + 0.5 * eta) +
             (a2b2 - mu * alpha0) * (1.0 + 2.0 * eta);
  double gamma = atan(-(((1.0 / beta0) * X) / denom_gamma));
  double omega = (-(F32(0.682) * exp(-(F32(24.697) * eta))) + F32(4.72)) + F32(4.45) * eta;
  double kappa = F32(4.674) * exp(-(F32(3.935) * eta)) + F32(3.536) * exp(-(F32(56.27) * eta));
  double alpha1 = ((F32(44.554) + F32(79.868) * eta) + (F32(116.432) * eta) * eta) - F32(44.652) * exp(2.0 * eta);
  double beta = (-F32(5.022) + F32(5.857) * eta) + F32(5.089) * exp(-(4.0 * eta));
  double rstar = (F32(2.0116) - F32(1.0647) * eta) + (F32(0.0538) * eta) * eta;
  double gm = ((((F32(1.0286) - F32(0.6095) * eta) + F32(3.5781) * ip2(eta)) - F32(21.3651) * ip3(eta)) +
               F32(42.6344) * ip4(eta)) - F32(33.8485) * ip5(eta);
  const double f23 = (double)(2.0f / 3.0f);
  double gsig = (1.0 / (4.0 * eta)) *
                (((((1.0 + eta) + ip2(eta)) - f23 * ip3(eta)) - f23 * ip4(eta)) / ip3(1.0 - eta) - 1.0);
  double Brad = ((gm - (gsig / rstar) * exp(mu * (rstar - 1.0))) * rstar) /
                (cos(beta * (rstar - 1.0) + gamma) * exp(alpha1 * (rstar - 1.0)) - cos(gamma) * exp(mu * (rstar - 1.0)));
  double Arad = gsig - Brad * cos(gamma);
  double delta = -(omega * rstar) - atan((kappa * rstar + 1.0) / (omega * rstar));
  double Crad = ((rstar * (gm - 1.0)) * exp(kappa * rstar)) / cos(omega * rstar + delta);
  double g2 = (Arad / r) * exp(mu * (r - 1.0)) + ((Brad / r) * cos(beta * (r - 1.0) + gamma)) * exp(alpha1 * (r - 1.0));
  if (r > rstar)
    g2 = 1.0 + ((Crad / r) * cos(omega * r + delta)) * exp(-(kappa * r));
  else if (r < 1.0)
    g2 = 0.0;
  return g2;
}

/* the integrands and Simpson sums used when rmax / rad < 3.49
 * (_functions.f:356-513); pi is 4*ATAN(1.0) in single precision */
static double am_integrand(int induced, double d, double rad, double alpha) {
  const double pi = (double)(4.0f * atanf(1.0f));
  double g = orc_am_rdf(d / (2.0 * rad), alpha);
  double f11 = induced ? orc_am_b12_11(d / rad, alpha) : orc_am_b11_11(d / rad, alpha);
  double f22 = induced ? orc_am_b12_22(d / rad, alpha) : orc_am_b11_22(d / rad, alpha);
  double rho = alpha / ((double)(4.0f / 3.0f) * pi);
  double w = ((((double)(1.0f / 3.0f) * rho) * 4.0) * pi) * ip2(d / rad);
  return (w * g) * (f11 + 2.0 * f22);
}

static double am_numerical(int induced, double rmax, double rad, double alpha) {
  const int npts = 1001;
  const double coef[2] = {4.0, 2.0};
  double maxr = rad * 20.0;
  if (!(maxr > rmax)) return 0.0;
  double dr = (maxr - rmax) / (double)(npts + 1);
  double integral = 0.0, r = rmax;
  integral = integral + am_integrand(induced, r, rad, alpha);
  integral = integral + am_integrand(induced, maxr, rad, alpha);
  for (int i = 2; i <= npts - 1; i++) {
    r = r + dr;
    double f = am_integrand(induced, r, rad, alpha);
    integral = integral + coef[i % 2] * f; /* 4, 2, 4, 2, ... */
  }
  return (dr * integral) / 3.0;
}

double orc_am_ia_numerical(double rmax, double rad, double alpha) { return am_numerical(0, rmax, rad, alpha); }
double orc_am_ii_numerical(double rmax, double rad, double alpha) { return am_numerical(1, rmax, rad, alpha); }

/* resistance_pair (_subroutines.f:37-143): B11 and B12 of a pair at
 * (x, y, z) = x_j - x_i rotated from the pair axis, Qt (B Q) with matmul's
 * k-ascending sums.  R[k][l] = the (k, l) entry of B11 (l < 3) or B12 (l >= 3). */
void orc_am_resistance_pair(double x, double y, double z, double alpha, double rad, double R[3][6]) {
  double gam = atan2(y, x), bet = -atan2(z, sqrt(x * x + y * y));
  double Q[3][3] = {{cos(bet) * cos(gam), cos(bet) * sin(gam), -sin(bet)},
                    {-sin(gam), cos(gam), 0.0},
                    {sin(bet) * cos(gam), sin(bet) * sin(gam), cos(bet)}};
  double dist = sqrt(x * x + y * y + z * z) / rad;
  double b1[3] = {orc_am_b11_11(dist, alpha), orc_am_b11_22(dist, alpha), 0.0};
  double b2[3] = {orc_am_b12_11(dist, alpha), orc_am_b12_22(dist, alpha), 0.0};
  b1[2] = b1[1];
  b2[2] = b2[1];
  for (int blk = 0; blk < 2; blk++) {
    const double *b = blk ? b2 : b1;
    double M[3][3]; /* matmul(B, Q): B diagonal */
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s = s + (k == i ? b[i] : 0.0) * Q[k][j];
        M[i][j] = s;
      }
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s = s + Q[k][i] * M[k][j];
        R[i][3 * blk + j] = s;
      }
  }
}

/* EvalNearestNeighbor's added-mass part for neighbour j of i: the pair
 * terms into Fam and the neighbour's acceleration into Wsum. */
static void am_pair(const double *yi, const double *rpi_, const double *yj, const double *rpj, double rphip,
                    double Fam[3], double Wsum[3], int *nn) {
  double rthresh = 0.5 * (rpi_[R_JDP] + rpj[R_JDP]);
  double rx = yj[JX] - yi[JX], ry = yj[JY] - yi[JY], rz = yj[JZ] - yi[JZ];
  double rdiff = sqrt((ip2(rx) + ip2(ry)) + ip2(rz));
  if (rdiff < rthresh) rx = ry = rz = rthresh; /* overlapping: the reference's substitute */
  double alpha_local = fmin(F32(0.4), rphip);
  double rad = 0.5 * rpi_[R_JDP];
  double R[3][6];
  orc_am_resistance_pair(rx, ry, rz, alpha_local, rad, R);
  *nn += 1;
  for (int k = 0; k < 3; k++) {
    for (int l = 0; l < 3; l++) {
      Fam[k] = Fam[k] + R[k][l] * rpi_[R_WDOTX + l];
      Fam[k] = Fam[k] + R[k][l + 3] * rpj[R_WDOTX + l];
    }
    Wsum[k] = Wsum[k] + rpj[R_WDOTX + k];
  }
}

/* ppiclf_solve_NearestNeighbor (ppiclf_solve.f:470-560) for the added mass:
 * the other particles by ascending index, then the periodic images, those
 * within the neighbour width (the squared-distance filter, then
 * EvalNearestNeighbor's rdiff <= d2chk(3)). */
static int am_neighbours(const struct orc_picl *p, int i, double rphip, double Fam[3], double Wsum[3]) {
  const double *yi = p->y + LRS * (size_t)i, *rpi_ = p->rprop + LRP * (size_t)i;
  const double w = p->nbW, dist2 = w * w;
  int nn = 0;
  for (int pass = 0; pass < 2; pass++) {
    for (int j = 0; j < p->n; j++) {
      if (p->lost[j] || (pass == 0 && j == i)) continue;
      for (int img = 0; img < (pass ? 27 : 1); img++) {
        if (pass && img == 13) continue;
        double yj[LRS];
        memcpy(yj, p->y + LRS * (size_t)j, sizeof yj);
        int skip = 0;
        if (pass)
          for (int d = 0, q = img; d < 3; d++, q /= 3) {
            int sd = q % 3 - 1;
            if (sd != 0 && !p->per[d]) skip = 1;
            if (sd != 0) yj[d] = yj[d] + (double)sd * (p->perHi[d] - p->perLo[d]);
          }
        if (skip) continue;
        double tot = (ip2(yi[JX] - yj[JX]) + ip2(yi[JY] - yj[JY])) + ip2(yi[JZ] - yj[JZ]);
        if (tot > dist2) continue;
        double rdiff = sqrt((ip2(yj[JX] - yi[JX]) + ip2(yj[JY] - yi[JY])) + ip2(yj[JZ] - yi[JZ]));
        if (!(rdiff <= w)) continue;
        am_pair(yi, rpi_, yj, p->rprop + LRP * (size_t)j, rphip, Fam, Wsum, &nn);
      }
    }
    if (!(p->per[0] || p->per[1] || p->per[2])) break;
  }
  return nn;
}

/* ------------------------------------------------------------------------ */
/* COLLISIONS 3-4: hydrodynamic and undisturbed torque                       */
/* (ppiclf_user_Torque.f:72-197), Saffman and Magnus lift                    */
/* (ppiclf_user_Lift.f:61-181).  rp holds the interpolated vorticity         */
/* (JXVOR..) and its material derivative (JSDOX..), y the angular velocity.  */
/* ------------------------------------------------------------------------ */
/* Torque_Hydro: -1 when the rotational Reynolds number reaches 100, where
 * ppiclF stops ("Re rotational too large") */
static int torque_hydro(const double *rp, const double *y, double rhof, double dp, double rmu, double rmass,
                        double t[3]) {
  const double rpi = acos(-1.0);
  double omgrx = 0.5 * rp[R_JXVOR] - y[7];
  double omgry = 0.5 * rp[R_JXVOR + 1] - y[8];
  double omgrz = 0.5 * rp[R_JXVOR + 2] - y[9];
  double omgr_mag = sqrt(omgrx * omgrx + omgry * omgry + omgrz * omgrz);
  double reyr = rhof * dp * dp * omgr_mag / (4.0 * rmu);
  reyr = fmax(reyr, 0.001);
  double Ct1, Ct2, Ct3;
  if (reyr < 1.0) {
    Ct1 = 0.0; Ct2 = 16.0 * rpi; Ct3 = 0.0;
  } else if (reyr < 10.0) {
    Ct1 = 0.0; Ct2 = 16.0 * rpi; Ct3 = 0.0418;
  } else if (reyr < 20.0) {
    Ct1 = 5.32; Ct2 = 37.2; Ct3 = 0.0;
  } else if (reyr < 50.0) {
    Ct1 = 6.44; Ct2 = 32.2; Ct3 = 0.0;
  } else if (reyr < 100.0) {
    Ct1 = 6.45; Ct2 = 32.1; Ct3 = 0.0;
  } else {
    return -1;
  }
  double Ct = Ct1 / sqrt(reyr) + Ct2 / reyr + Ct3 * reyr;
  double beta = rp[R_JRHOP] / rhof;
  double rIp = rmass * dp * dp / 10.0;
  double factor = rIp * 60.0 * Ct * omgr_mag / (64.0 * rpi * beta);
  t[0] = factor * omgrx;
  t[1] = factor * omgry;
  t[2] = factor * omgrz;
  return 0;
}

/* Lift_Saffman then Lift_Magnus, added into lift (vrel: the relative
 * velocity vx, vy, vz; rep already capped at 0.1, rnu = rmu / rhof) */
static void lift_forces(const double *rp, const double *y, const double vrel[3], double vmag, double dp, double rep,
                        double rmu, double rhof, double rnu, double lift[3]) {
  if (vmag < 1.0e-8) return;
  double vx = vrel[0], vy = vrel[1], vz = vrel[2];
  { /* Saffman, with Mei's correction */
    double omgx = rp[R_JXVOR], omgy = rp[R_JXVOR + 1], omgz = rp[R_JXVOR + 2];
    double omg_mag = sqrt(omgx * omgx + omgy * omgy + omgz * omgz);
    double omg_star = omg_mag * dp / vmag;
    double epi = sqrt(omg_star / rep);
    double d1 = 1.0 + tanh(2.5 * (log10(epi) + 0.191));
    double d2 = 0.667 + tanh(6.0 * epi - 1.92);
    double Jepi = 0.3 * d1 * d2;
    double factor = 1.615 * rmu * (dp * dp) * vmag * sqrt(omg_mag / rnu);
    double elx = vy * omgz - vz * omgy, ely = vz * omgx - vx * omgz, elz = vx * omgy - vy * omgx;
    double elm = sqrt(elx * elx + ely * ely + elz * elz);
    elm = fmax(1.0e-20, elm);
    double ielm = 1.0 / elm;
    lift[0] = lift[0] + factor * Jepi * elx * ielm;
    lift[1] = lift[1] + factor * Jepi * ely * ielm;
    lift[2] = lift[2] + factor * Jepi * elz * ielm;
  }
  { /* Magnus, from the particle's spin */
    double omgx = y[7], omgy = y[8], omgz = y[9];
    double omg_mag = sqrt(omgx * omgx + omgy * omgy + omgz * omgz);
    double epi = omg_mag * dp / vmag;
    double d1 = 0.675 + 0.15 * (1.0 + tanh(0.28 * (epi - 2.0)));
    double CL = 1.0 - d1 * tanh(F32(0.18) * sqrt(rep));
    double factor = 0.125 * dp * dp * dp * rhof;
    double elx = vy * omgz - vz * omgy, ely = vz * omgx - vx * omgz, elz = vx * omgy - vy * omgx;
    lift[0] = lift[0] + factor * CL * elx;
    lift[1] = lift[1] + factor * CL * ely;
    lift[2] = lift[2] + factor * CL * elz;
  }
}

/* test hooks: the torque and lift laws on given inputs (vort: the gas
 * vorticity, spin: the particle's angular velocity) */
int orc_picl_torque_hydro(const double *vort, const double *spin, double rhof, double dp, double rmu, double rmass,
                          double rhop, double *t) {
  double rp[LRP] = {0}, y[LRS] = {0};
  for (int c = 0; c < 3; c++) {
    rp[R_JXVOR + c] = vort[c];
    y[7 + c] = spin[c];
  }
  rp[R_JRHOP] = rhop;
  return torque_hydro(rp, y, rhof, dp, rmu, rmass, t);
}
void orc_picl_lift(const double *vort, const double *spin, const double *vrel, double dp, double rep, double rmu,
                   double rhof, double *lift) {
  double rp[LRP] = {0}, y[LRS] = {0};
  for (int c = 0; c < 3; c++) {
    rp[R_JXVOR + c] = vort[c];
    y[7 + c] = spin[c];
  }
  double vmag = sqrt(vrel[0] * vrel[0] + vrel[1] * vrel[1] + vrel[2] * vrel[2]);
  lift[0] = lift[1] = lift[2] = 0.0;
  lift_forces(rp, y, vrel, vmag, dp, rep, rmu, rhof, rmu / rhof, lift);
}

/* ------------------------------------------------------------------------ */
/* Quasi-steady force fluctuations (QSFLUCT 1 Lattanzi, 2 Osnes):           */
/* ppiclf_user_Fluctuations.f:32-418, the neighbour means of                 */
/* ppiclf_user.f:411-468 and ppiclf_user_EvalNearestNeighbor.f:318-360.      */
/* ------------------------------------------------------------------------ */
/* The six uniform draws of a particle's call stand in for RANDOM_NUMBER
 * (libgfortran's generator and its call order, which no restatement can
 * reproduce): a splitmix64 hash of (seed, call, original index, draw),
 * uniform on (0, 1) with 53 bits. */
double orc_fluct_uniform(long long seed, int call, int pid, int k) {
  uint64_t n = ((((uint64_t)(uint32_t)call << 32) | (uint32_t)pid) << 3) + (uint64_t)(k + 1);
  uint64_t z = (uint64_t)seed + 0x9E3779B97F4A7C15ull * n;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return ((double)(z >> 11) + 0.5) * 0x1.0p-53;
}

/* TwoSum accumulation: the rounded s + c does not depend on the order of
 * the terms (in all but contrived cases), which the means need: their
 * variance theta = <u^2> - <u>^2 cancels to rounding noise when the
 * neighbours move together. */
static void fl_acc(double *s, double *c, double x) {
  double t = *s + x, z = t - *s;
  *c = *c + ((*s - (t - z)) + (x - z));
  *s = t;
}

/* The box-filter means' sums: the particle itself, then every neighbour
 * (and periodic image) that passes ppiclf_solve_NearestNeighbor's filter
 * (d2chk(3)) and lies within the box of half-width d2chk(2) (adaptive:
 * max(d2chk(2), (10 d^3 / max(1e-4, phi))^(1/3) / 2)).  out = the sums of
 * u, v, w, u^2, v^2, w^2, then the count. */
static void fluct_means(const struct orc_picl *p, int i, double out[7]) {
  const double *yi = p->y + LRS * (size_t)i, *rpi_ = p->rprop + LRP * (size_t)i;
  const double w = p->nbW, dist2 = w * w;
  double box = p->fluctW;
  if (p->pp.qsFluctAdapt >= 1) {
    double dpl = rpi_[R_JDP], phip = rpi_[R_JPHIP];
    double adptfilter = pow(10.0 * ip3(dpl) / fmax(F32(1.e-4), phip), (double)(1.0f / 3.0f));
    adptfilter = adptfilter / 2.0;
    box = fmax(p->fluctW, adptfilter);
  }
  double S[6], Cc[6] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 3; k++) {
    S[k] = yi[JVX + k];
    S[3 + k] = S[k] * S[k];
  }
  int cnt = 1;
  for (int pass = 0; pass < 2; pass++) {
    for (int j = 0; j < p->n; j++) {
      if (p->lost[j] || (pass == 0 && j == i)) continue;
      for (int img = 0; img < (pass ? 27 : 1); img++) {
        if (pass && img == 13) continue;
        double yj[LRS];
        memcpy(yj, p->y + LRS * (size_t)j, sizeof yj);
        int skip = 0;
        if (pass)
          for (int d = 0, q = img; d < 3; d++, q /= 3) {
            int sd = q % 3 - 1;
            if (sd != 0 && !p->per[d]) skip = 1;
            if (sd != 0) yj[d] = yj[d] + (double)sd * (p->perHi[d] - p->perLo[d]);
          }
        if (skip) continue;
        double tot = (ip2(yi[JX] - yj[JX]) + ip2(yi[JY] - yj[JY])) + ip2(yi[JZ] - yj[JZ]);
        if (tot > dist2) continue;
        if (fabs(yi[JX] - yj[JX]) > box || fabs(yi[JY] - yj[JY]) > box || fabs(yi[JZ] - yj[JZ]) > box) continue;
        for (int k = 0; k < 3; k++) {
          fl_acc(&S[k], &Cc[k], yj[JVX + k]);
          fl_acc(&S[3 + k], &Cc[3 + k], ip2(yj[JVX + k]));
        }
        cnt++;
      }
    }
    if (!(p->per[0] || p->per[1] || p->per[2])) break;
  }
  for (int k = 0; k < 6; k++) out[k] = S[k] + Cc[k];
  out[6] = (double)cnt;
}

/* ppiclf_user_QS_fluct_Lattanzi (flag 1) and _Osnes (flag 2): the SDE step
 * of the fluctuating force from its previous value F (rprop FLUCTFX..Z),
 * the means m, the six uniforms U and fac = rk3ark(stage) dt. */
void orc_fluct_force(int flag, const double m[7], const double F[3], const double U[6], double rphip,
                     double rmachp, double rep, double vmag, const double vrel[3], double rmu, double dp,
                     double fac, double out[3]) {
  const double rpi = acos(-1.0), TwoPi = 2.0 * acos(-1.0);
  double denum = fmax(m[6], 1.0);
  double upmean = m[0] / denum, vpmean = m[1] / denum, wpmean = m[2] / denum;
  double u2pmean = m[3] / denum, v2pmean = m[4] / denum, w2pmean = m[5] / denum;
  double fq = F32(6.52) * rphip - F32(22.56) * ip2(rphip) + F32(49.90) * ip3(rphip);
  double Fs = 3.0 * rpi * rmu * dp * (1.0 + F32(0.15) * pow(rep * (1.0 - rphip), F32(0.687))) * (1.0 - rphip) * vmag;
  double bq = fmin(sqrt(20.0 * rmachp), 1.0) * F32(0.55) * pow(rphip, F32(0.7)) *
              (1.0 + tanh((rmachp - 0.5) / F32(0.2)));
  double theta = ((u2pmean + v2pmean + w2pmean) - (ip2(upmean) + ip2(vpmean) + ip2(wpmean))) / 3.0;
  if (theta <= 1.0e-12) theta = 0.0;
  double chi = (1.0 + 2.50 * rphip + F32(4.51) * ip2(rphip) + F32(4.52) * ip3(rphip)) /
               pow(1.0 - ip3(rphip / F32(0.64)), F32(0.68));
  double tF_inv = (24.0 * rphip * chi / dp) * sqrt(theta / rpi);
  double aSDE = tF_inv;
  if (flag == 1) {
    double sigF = (fq + bq) * Fs;
    double bSDE = sigF * sqrt(2.0 * tF_inv);
    for (int k = 0; k < 3; k++) {
      double Z = sqrt(-2.0 * log(U[2 * k])) * cos(TwoPi * U[2 * k + 1]);
      double dW = sqrt(fac) * Z;
      out[k] = (1.0 - aSDE * fac) * F[k] + bSDE * dW;
    }
    return;
  }
  double sigD = (fq + bq) * Fs;
  double bSDE_CD = sigD * sqrt(2.0 * tF_inv);
  /* below 1e-8 m/s the axis is x and the transverse amplitude vanishes
   * (bSDE_CD above keeps the unguarded sigD) (ppiclf_user_Fluctuations.f:317-328) */
  double avec[3] = {1.0, 0.0, 0.0}, CD_frac = 0.0;
  if (vmag > 1.0e-8) {
    for (int k = 0; k < 3; k++) avec[k] = vrel[k] / vmag;
    double CD_prime = F[0] * avec[0] + F[1] * avec[1] + F[2] * avec[2];
    CD_frac = CD_prime / sigD;
  } else {
    sigD = 0.0;
  }
  double sigmoid_cf = 1.0 / (1.0 + exp(-CD_frac));
  double f_CF = F32(0.39356905) * sigmoid_cf + F32(0.43758848);
  double sigT = f_CF * sigD;
  double bSDE_CL = sigT * sqrt(2.0 * tF_inv);
  double eunit[3] = {1.0, 0.0, 0.0};
  if (fabs(avec[1]) + fabs(avec[2]) <= 1.0e-8) {
    eunit[0] = 0.0; eunit[1] = 1.0;
  } else if (fabs(avec[0]) + fabs(avec[2]) <= 1.0e-8) {
    eunit[0] = 0.0; eunit[2] = 1.0;
  }
  double bvec[3] = {avec[1] * eunit[2] - avec[2] * eunit[1], avec[2] * eunit[0] - avec[0] * eunit[2],
                    avec[0] * eunit[1] - avec[1] * eunit[0]};
  double dn = fmax(1.0e-8, sqrt(ip2(bvec[0]) + ip2(bvec[1]) + ip2(bvec[2])));
  for (int k = 0; k < 3; k++) bvec[k] = bvec[k] / dn;
  double cvec[3] = {avec[1] * bvec[2] - avec[2] * bvec[1], avec[2] * bvec[0] - avec[0] * bvec[2],
                    avec[0] * bvec[1] - avec[1] * bvec[0]};
  dn = fmax(1.0e-8, sqrt(ip2(cvec[0]) + ip2(cvec[1]) + ip2(cvec[2])));
  for (int k = 0; k < 3; k++) cvec[k] = cvec[k] / dn;
  double Z1 = sqrt(-2.0 * log(U[0])) * cos(TwoPi * U[1]);
  double Z2 = sqrt(-2.0 * log(U[2])) * cos(TwoPi * U[3]);
  double dW1 = sqrt(fac) * Z1, dW2 = sqrt(fac) * Z2;
  double cosrand = cos(TwoPi * U[4]), sinrand = sin(TwoPi * U[4]);
  double dvec[3];
  for (int k = 0; k < 3; k++) dvec[k] = bvec[k] * cosrand + cvec[k] * sinrand;
  dn = fmax(1.0e-8, sqrt(ip2(dvec[0]) + ip2(dvec[1]) + ip2(dvec[2])));
  for (int k = 0; k < 3; k++) dvec[k] = dvec[k] / dn;
  for (int k = 0; k < 3; k++) out[k] = (1.0 - aSDE * fac) * F[k] + bSDE_CD * dW1 * avec[k] + bSDE_CL * dW2 * dvec[k];
}

/* ppiclf_user_SetYdot (libpicl/user_files/ppiclf_user.f:39-1042) for one
 * particle with the quasi-steady, added-mass (ppiclf_user_AddedMass.f:16-124),
 * pressure-gradient, viscous-unsteady and normal collision forces and heat
 * transfer, Euler flow (no viscous stress term), no tangential collision
 * force, lift, torque or burning. */
static void set_ydot(struct orc_picl *p, int i, double dt) {
  const orc_picl_params *P = &p->pp;
  double *y = p->y + LRS * (size_t)i, *yd = p->ydot + LRS * (size_t)i, *yc = p->ydotc + LRS * (size_t)i;
  const double *rp = p->rprop + LRP * (size_t)i;
  const double rcp_part = P->rcpPart;
  double rmu;
  if (P->rmuFlag == 0) {
    rmu = P->rmuRef;
  } else {
    double temp = rp[R_JT];
    rmu = P->rmuRef * sqrt(temp / P->tref) * (1.0 + P->suth / P->tref) / (1.0 + P->suth / temp);
  }
  double rmass = rp[R_JVOLP] * rp[R_JRHOP];
  double vx = rp[R_JUX] - y[JVX];
  double vy = rp[R_JUY] - y[JVY];
  double vz = rp[R_JUZ] - y[JVZ];
  double vmag = sqrt(vx * vx + vy * vy + vz * vz);
  double rhof = rp[R_JRHOF];
  double dp = rp[R_JDP];
  double rep = vmag * dp * rhof / rmu;
  double rphip = rp[R_JPHIP];
  double rphif = 1.0 - rp[R_JPHIP];
  double asndf = rp[R_JCS];
  double rmachp = vmag / asndf;
  if (vmag <= 1.0e-3) return; /* keeps the previous ydot; ydotc stays 0 */
  rep = fmax(0.1, rep);
  rphip = fmin(rphip, 0.62);
  rphif = 1.0 - rphip;
  double famx = 0.0, famy = 0.0, famz = 0.0, rmass_add = 0.0;
  double fdpdx = 0.0, fdpdy = 0.0, fdpdz = 0.0;
  double fcx = 0.0, fcy = 0.0, fcz = 0.0;
  double fvux = 0.0, fvuy = 0.0, fvuz = 0.0;
  double Fam[3] = {0.0, 0.0, 0.0}, Wsum[3] = {0.0, 0.0, 0.0};
  int nneighbors = 0;
  if (P->amFlag == 2) {
    /* Step 1a: the unary term of Briney's model (ppiclf_user_AM_Briney_Unary,
     * ppiclf_user_AddedMass.f:164-240): Mach correction only, no volume
     * factor yet */
    double rcd_am;
    if (rmachp > F32(0.6)) {
      rcd_am = 0x1.5104d6p+1; /* 1.0 + 1.8*((0.6)**2) + 7.6*((0.6)**4) folded in single precision */
    } else {
      double m2 = rmachp * rmachp;
      rcd_am = 1.0 + F32(1.8) * m2 + F32(7.6) * (m2 * m2);
    }
    rcd_am = rcd_am * 0.5;
    rmass_add = rhof * rp[R_JVOLP] * rcd_am;
    double SDrho = rp[R_JRHSR] + y[JVX] * rp[R_JPGCX] + y[JVY] * rp[R_JPGCY] + y[JVZ] * rp[R_JPGCZ];
    SDrho = SDrho / rphif;
    double vgradrho = vx * rp[R_JRHOGX] + vy * rp[R_JRHOGY] + vz * rp[R_JRHOGZ];
    famx = rcd_am * (vx * SDrho + rhof * rp[R_JSDRX] + rp[R_JUX] * vgradrho);
    famy = rcd_am * (vy * SDrho + rhof * rp[R_JSDRY] + rp[R_JUY] * vgradrho);
    famz = rcd_am * (vz * SDrho + rhof * rp[R_JSDRZ] + rp[R_JUZ] * vgradrho);
    Fam[0] = famx;
    Fam[1] = famy;
    Fam[2] = famz;
    /* Step 1b: the pair terms (ppiclf_user.f:411-468) */
    nneighbors = am_neighbours(p, i, rphip, Fam, Wsum);
  }
  double fmeans[7];
  if (P->qsFluct >= 1) fluct_means(p, i, fmeans); /* Step 1b, the fluctuation means */
  double tqc[3] = {0.0, 0.0, 0.0}; /* collision torques (ydotc(JOX..JOZ)) */
  if (P->collisionalFlag >= 1) { /* Step 1b and Step 6, ppiclf_user.f:411-470, 587-600 */
    double fc[3];
    collision_sum(p, i, rmass, dt, fc, tqc);
    fcx = fc[0];
    fcy = fc[1];
    fcz = fc[2];
  }
  double lift[3] = {0.0, 0.0, 0.0};
  if (P->collisionalFlag == 4) { /* Step 9b (:640-646) */
    double vrel[3] = {vx, vy, vz};
    lift_forces(rp, y, vrel, vmag, dp, rep, rmu, rhof, rmu / rhof, lift);
  }
  const double liftx = lift[0], lifty = lift[1], liftz = lift[2];
  double beta; /* ppiclf_user.f:473-485 */
  switch (P->qsFlag) {
    case 1: beta = qs_parmar(rmachp, rphip, rep, rmu, dp); break;
    case 2: beta = qs_osnes(rmachp, rphip, rep, rmu, dp); break;
    case 3: beta = qs_modified_parmar(rmachp, rphip, rep, rmu, dp); break;
    case 4: beta = qs_gidaspow(rphip, rphif, rep, rmu, dp, rhof, vmag); break;
    default: beta = qs_sn(rep, rmu, dp); break;
  }
  double fqsx = beta * vx, fqsy = beta * vy, fqsz = beta * vz;
  double fl[3] = {0.0, 0.0, 0.0};
  if (P->qsFluct >= 1) { /* Step 3 (ppiclf_user.f:487-510) */
    static const double rk3ark[3] = {8.0 / 15.0, 5.0 / 12.0, 0.75};
    double fac = rk3ark[p->curStage - 1] * dt, U[6], vrel[3] = {vx, vy, vz};
    for (int k = 0; k < 6; k++) U[k] = orc_fluct_uniform(P->fluctSeed, p->nCalls, i, k);
    orc_fluct_force(P->qsFluct, fmeans, rp + R_FLUCTFX, U, rphip, rmachp, rep, vmag, vrel, rmu, dp, fac, fl);
  }
  fqsx = fqsx + fl[0];
  fqsy = fqsy + fl[1];
  fqsz = fqsz + fl[2];
  if (P->qsFluct >= 1)
    for (int k = 0; k < 3; k++) p->rprop[LRP * (size_t)i + R_FLUCTFX + k] = fl[k];
  if (P->amFlag == 1) {
    double rcd_am;
    if (rmachp > F32(0.6)) {
      /* 1.0 + 1.8*((0.6)**2) + 7.6*((0.6)**4): a real*4 constant expression
       * that the compiler folds in single precision */
      rcd_am = 0x1.5104d6p+1;
    } else {
      double m2 = rmachp * rmachp;
      rcd_am = 1.0 + F32(1.8) * m2 + F32(7.6) * (m2 * m2);
    }
    rcd_am = rcd_am * 0.5;
    rcd_am = rcd_am * (1.0 + F32(0.68) * (rphip * rphip));
    rmass_add = rhof * rp[R_JVOLP] * rcd_am;
    double SDrho = rp[R_JRHSR] + y[JVX] * rp[R_JPGCX] + y[JVY] * rp[R_JPGCY] + y[JVZ] * rp[R_JPGCZ];
    SDrho = SDrho / rphif;
    double vgradrho = vx * rp[R_JRHOGX] + vy * rp[R_JRHOGY] + vz * rp[R_JRHOGZ];
    double ug = rp[R_JUX], vg = rp[R_JUY], wg = rp[R_JUZ];
    famx = rcd_am * rp[R_JVOLP] * (vx * SDrho + rhof * rp[R_JSDRX] + ug * vgradrho);
    famy = rcd_am * rp[R_JVOLP] * (vy * SDrho + rhof * rp[R_JSDRY] + vg * vgradrho);
    famz = rcd_am * rp[R_JVOLP] * (vz * SDrho + rhof * rp[R_JSDRZ] + wg * vgradrho);
  } else if (P->amFlag == 2) { /* Step 4 (ppiclf_user.f:527-552) */
    if (nneighbors > 0) {      /* ppiclf_user_AM_Briney_Binary (ppiclf_user_AddedMass.f:290-360) */
      double rad = rp[R_JDP] * 0.5;
      double dr_max = p->nbW;
      double alpha = fmin(F32(0.4), rphip);
      double IA, II;
      if (dr_max / rad >= F32(3.49)) {
        IA = orc_am_ia_analytical(dr_max, rad, alpha);
        II = orc_am_ii_analytical(dr_max, rad, alpha);
      } else {
        IA = orc_am_ia_numerical(dr_max, rad, alpha);
        II = orc_am_ii_numerical(dr_max, rad, alpha);
      }
      for (int c = 0; c < 3; c++) {
        Fam[c] = Fam[c] + IA * rp[R_WDOTX + c];
        Fam[c] = Fam[c] + II * Wsum[c] / (double)nneighbors;
      }
      for (int c = 0; c < 3; c++) Fam[c] = rp[R_JVOLP] * Fam[c];
      famx = Fam[0];
      famy = Fam[1];
      famz = Fam[2];
    } else {
      famx = famx * rp[R_JVOLP];
      famy = famy * rp[R_JVOLP];
      famz = famz * rp[R_JVOLP];
    }
  }
  if (P->pgFlag == 1) {
    fdpdx = -rp[R_JVOLP] * rp[R_JDPDX];
    fdpdy = -rp[R_JVOLP] * rp[R_JDPDY];
    fdpdz = -rp[R_JVOLP] * rp[R_JDPDZ];
    fdpdx = fdpdx + 0.0; /* + fdpvdx (Euler flow) */
    fdpdy = fdpdy + 0.0;
    fdpdz = fdpdz + 0.0;
  }
  if (P->viscousUnsteady == 1) { /* Step 7, ppiclf_user.f:605-607 */
    double fvu[3];
    vu_force(p, i, dt, rmu, rhof, dp, vmag, rphip, fvu);
    fvux = fvu[0];
    fvuy = fvu[1];
    fvuz = fvu[2];
  }
  double rmass_therm = rmass * rcp_part;
  double qq = 0.0;
  if (P->heatTransfer >= 1) {
    double rkappa = 1004.64 * rmu / 0.70; /* rcp_fluid*rmu/rpr (ppiclf_user.f:310) */
    qq = heat_transfer(P->heatTransfer, rkappa, dp, rp[R_JT], y[JT], rep, rphif, rmachp);
  }
  double rmass_omega = rmass * dp * dp / 10.0;
  double taux = 0.0, tauy = 0.0, tauz = 0.0;
  double tauh[3] = {0.0, 0.0, 0.0}, tauu[3] = {0.0, 0.0, 0.0};
  if (P->collisionalFlag >= 2) { /* Step 9a, :638-646: Torque_driver */
    if (P->collisionalFlag >= 3) {
      if (torque_hydro(rp, y, rhof, dp, rmu, rmass, tauh)) p->rotErr = 1;
      double rIf = rhof * dp * dp * rp[R_JVOLP] / 10.0; /* Torque_Undisturbed */
      for (int c = 0; c < 3; c++) tauu[c] = 0.5 * rIf * rp[R_JSDOX + c];
    }
    taux = tqc[0] + tauh[0] + tauu[0];
    tauy = tqc[1] + tauh[1] + tauu[1];
    tauz = tqc[2] + tauh[2] + tauu[2];
  }
  yd[JX] = y[JVX];
  yd[JY] = y[JVY];
  yd[JZ] = y[JVZ];
  yd[JVX] = (fqsx + famx + fdpdx + fvux + liftx + fcx) / (rmass + rmass_add);
  yd[JVY] = (fqsy + famy + fdpdy + fvuy + lifty + fcy) / (rmass + rmass_add);
  yd[JVZ] = (fqsz + famz + fdpdz + fvuz + liftz + fcz) / (rmass + rmass_add);
  yd[JT] = qq / rmass_therm;
  yd[7] = taux / rmass_omega;
  yd[8] = tauy / rmass_omega;
  yd[9] = tauz / rmass_omega;
  yd[10] = 0.0;
  yd[11] = 0.0;
  if (P->viscousUnsteady >= 1) vu_update(p, i, rp, y, yd, vx, vy, vz, rhof, rphif); /* :669-672 */
  if (P->feedbackFlag == 1) {
    const double tauxh = tauh[0], tauyh = tauh[1], tauzh = tauh[2];
    yc[JVX] = rp[R_JSPL] * (yd[JVX] * rmass - fcx);
    yc[JVY] = rp[R_JSPL] * (yd[JVY] * rmass - fcy);
    yc[JVZ] = rp[R_JSPL] * (yd[JVZ] * rmass - fcz);
    yc[JT] = rp[R_JSPL] * ((fqsx + fvux + liftx) * y[JVX] + (fqsy + fvuy + lifty) * y[JVY] +
                           (fqsz + fvuz + liftz) * y[JVZ] + famx * rp[R_JUX] + famy * rp[R_JUY] +
                           famz * rp[R_JUZ] + tauxh * y[7] + tauyh * y[8] + tauzh * y[9] + qq);
  } else { /* ydotc keeps what the collisions added into it */
    yc[JVX] = fcx;
    yc[JVY] = fcy;
    yc[JVZ] = fcz;
    yc[JT] = 0.0;
  }
}

/* ppiclf_user.f:934-990 (am_flag 2, after every particle's forces): each
 * particle's density-weighted acceleration W = D(rho_g u_g)/Dt - rho_g dv/dt,
 * read by the next stage's pair terms (rprop(WDOTX..WDOTZ)).  The loop reads
 * two variables the particle loop left behind: rphif of the last particle
 * (every particle's SDrho is divided by it), and vx, vy, vz, which it sets
 * only after using them in vgradrhog, so particle i's gradient term takes the
 * previous particle's relative velocity (the first particle, the last one's). */
static void am_store_wdot(struct orc_picl *p) {
  int last = -1;
  for (int i = p->n - 1; i >= 0 && last < 0; i--)
    if (!p->lost[i]) last = i;
  if (last < 0) return;
  const double *rl = p->rprop + LRP * (size_t)last, *yl = p->y + LRS * (size_t)last;
  double vx = rl[R_JUX] - yl[JVX], vy = rl[R_JUY] - yl[JVY], vz = rl[R_JUZ] - yl[JVZ];
  double rphif = 1.0 - rl[R_JPHIP]; /* ppiclf_user.f:324, then :360-361 unless the particle was skipped */
  if (!(sqrt(vx * vx + vy * vy + vz * vz) <= 1.0e-3)) rphif = 1.0 - fmin(rl[R_JPHIP], 0.62);
  for (int i = 0; i < p->n; i++) {
    if (p->lost[i]) continue;
    double *rp = p->rprop + LRP * (size_t)i;
    const double *y = p->y + LRS * (size_t)i, *yd = p->ydot + LRS * (size_t)i;
    double SDrho = rp[R_JRHSR] + y[JVX] * rp[R_JPGCX] + y[JVY] * rp[R_JPGCY] + y[JVZ] * rp[R_JPGCZ];
    SDrho = SDrho / rphif;
    double vgradrhog = vx * rp[R_JRHOGX] + vy * rp[R_JRHOGY] + vz * rp[R_JRHOGZ];
    double rhof = rp[R_JRHOF];
    vx = rp[R_JUX] - y[JVX];
    vy = rp[R_JUY] - y[JVY];
    vz = rp[R_JUZ] - y[JVZ];
    double ug = rp[R_JUX], vg = rp[R_JUY], wg = rp[R_JUZ];
    rp[R_WDOTX] = vx * SDrho + rhof * rp[R_JSDRX] + ug * vgradrhog - rhof * yd[JVX];
    rp[R_WDOTX + 1] = vy * SDrho + rhof * rp[R_JSDRY] + vg * vgradrhog - rhof * yd[JVY];
    rp[R_WDOTX + 2] = vz * SDrho + rhof * rp[R_JSDRZ] + wg * vgradrhog - rhof * yd[JVZ];
  }
}

/* ppiclf_solve_InvokeLinearPeriodic (ppiclf_solve.f:1319-1353), reached
 * from RemoveParticle (ppiclf.f:13140-13144) after every stage's update
 * (ppiclf_solve.f:1933): each axis on its own, a particle beyond one bound
 * re-enters the same distance inside the other.  A non-periodic axis keeps
 * ppiclF's +-1e20 range, so nothing happens there. */
static void linear_periodic(const struct orc_picl *p, double *y) {
  for (int d = 0; d < 3; d++) {
    if (!p->per[d]) continue;
    if (y[d] < p->perLo[d])
      y[d] = p->perHi[d] - fabs(p->perLo[d] - y[d]);
    else if (y[d] > p->perHi[d])
      y[d] = p->perLo[d] + fabs(y[d] - p->perHi[d]);
  }
}

/* ppiclf_solve_Initialize's angular case (ppiclf_solve.f:60-104, with the
 * REAL*4 pi = ACOS(-1.0): angled = angle * 180 / pi; below 90 case 1, NINT 90
 * case 2, NINT 180 case 3, anything else stops in InitAngularPeriodic) and
 * ppiclf_solve_InitAngularPeriodic (:1172-1230: rin < rout; xdrange(:,1) =
 * (rin, rout), for the half cylinder (-rout, rout)).  The rotation matrices of
 * ppiclf_solve_InvokeAngularPeriodic (:1396-1500) for the local angles
 * +angle, -angle and 0, axis e_z, written as the reference writes them. */
static int angular_init(struct orc_picl *p, const orc_picl_params *pp) {
  double pi = (double)acosf(-1.0f);
  double angled = pp->angAngle * 180.0 / pi;
  int cs = 0;
  if (angled < 90.0) cs = 1;
  if (lround(angled) == 90) cs = 2;
  if (lround(angled) == 180) cs = 3;
  if (cs == 0 || !(pp->angAngle > 0.0) || !(pp->angRin < pp->angRout)) return -1;
  p->ang = 1;
  p->angCase = cs;
  p->angX = pp->angXAngle;
  p->angA = pp->angAngle;
  p->angXd[0] = cs == 3 ? -1.0 * pp->angRout : pp->angRin;
  p->angXd[1] = pp->angRout;
  const double ex = 0.0, ey = 0.0, ez = 1.0;
  for (int k = 0; k < 3; k++) {
    double local_angle = k == 0 ? pp->angAngle : (k == 1 ? -1.0 * pp->angAngle : 0.0);
    double ct = cos(local_angle), st = sin(local_angle);
    double (*R)[3] = p->angR[k];
    R[0][0] = ct + (1.0 - ct) * ex * ex;
    R[0][1] = (1.0 - ct) * ex * ey - st * ez;
    R[0][2] = (1.0 - ct) * ex * ez + st * ey;
    R[1][0] = (1.0 - ct) * ey * ex + st * ez;
    R[1][1] = ct + (1.0 - ct) * ey * ey;
    R[1][2] = (1.0 - ct) * ey * ez - st * ex;
    R[2][0] = (1.0 - ct) * ez * ex - st * ey;
    R[2][1] = (1.0 - ct) * ez * ey + st * ex;
    R[2][2] = ct + (1.0 - ct) * ez * ez;
  }
  return 0;
}

/* MATMUL(rotmat, v) as gfortran forms it: each row summed over the columns
 * in order from zero */
static void rotate3(double R[3][3], double *v) {
  double o[3];
  for (int i = 0; i < 3; i++) {
    double sum = 0.0;
    for (int j = 0; j < 3; j++) sum = sum + R[i][j] * v[j];
    o[i] = sum;
  }
  for (int i = 0; i < 3; i++) v[i] = o[i];
}

/* ppiclf_solve_RemoveParticle for one particle after the update (:2826-2962):
 * with angular periodicity, a polar angle per_alpha = atan2(y, x) outside
 * [xangle, xangle + angle] calls InvokeAngularPeriodic (register 1): local
 * angle +angle, -angle when per_alpha > 0.5 (xangle + angle), 0 for the half
 * cylinder when per_alpha < xangle; position and velocity rotated.  Then
 * InvokeLinearPeriodic (:1319-1394): without angular periodicity each
 * periodic axis wraps; with it and periodicZ, an x outside xdrange(:,1)
 * marks the particle for removal (MarkForRemoval, iprop 3), otherwise z
 * wraps. */
static void remove_particle(struct orc_picl *p, int i, double *y) {
  if (!p->ang) {
    linear_periodic(p, y);
    return;
  }
  double per_alpha = atan2(y[1], y[0]);
  if (per_alpha < p->angX || per_alpha > p->angX + p->angA) {
    int k = 0;
    if (per_alpha > 0.5 * (p->angX + p->angA)) k = 1;
    if (p->angCase == 3 && per_alpha < p->angX) k = 2;
    rotate3(p->angR[k], y);
    rotate3(p->angR[k], y + 3);
  }
  if (p->per[2]) {
    if (y[0] < p->angXd[0] || y[0] > p->angXd[1]) {
      p->mark[i] = 1;
      return;
    }
    if (y[2] < p->perLo[2])
      y[2] = p->perHi[2] - fabs(p->perLo[2] - y[2]);
    else if (y[2] > p->perHi[2])
      y[2] = p->perLo[2] + fabs(y[2] - p->perHi[2]);
  }
}

/* ppiclf_solve_IntegrateRK3s_Rocflu + SetRK3Coeff (imethod 2):
 * ppiclf_solve.f:1845-1936, 2774-2793 */
static void integrate_particle(struct orc_picl *p, int istage, double dt) {
  double c[3][3] = {{0.0, 1.0, dt * 8.0 / 15.0},
                    {dt * 17.0 / 60.0, 1.0, dt * 5.0 / 12.0},
                    {dt * 5.0 / 12.0, 1.0, dt * 3.0 / 4.0}};
  init_solve(p, 1);
  p->curStage = istage;
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < p->n; i++)
    if (!p->lost[i]) set_ydot(p, i, dt);
  p->nCalls++;
  if (p->pp.amFlag == 2) am_store_wdot(p);
  if (p->pp.viscousUnsteady == 1 && istage == 3) vu_shift(p, dt);
  if (istage == 1) memset(p->y1, 0, sizeof(double) * LRS * (size_t)p->n);
  const double *k = c[istage - 1];
  for (int i = 0; i < p->n; i++) {
    if (p->lost[i]) continue;
    double *y = p->y + LRS * (size_t)i, *y1 = p->y1 + LRS * (size_t)i, *yd = p->ydot + LRS * (size_t)i;
    for (int j = 0; j < LRS; j++) y[j] = -(k[0] * y1[j]) + k[1] * y[j] + k[2] * yd[j];
    for (int j = 0; j < LRS; j++) y1[j] = yd[j];
    remove_particle(p, i, y);
  }
}

/* ------------------------------------------------------------------------ */
/* PICL_TEMP_Runge: rocpicl/PICL_TEMP_Runge.F90:623-1234                     */
/* ------------------------------------------------------------------------ */
static double sum8(double v) {
  double s = 0.0;
  for (int q = 0; q < 8; q++) s = s + v;
  return s;
}
/* Sum over the element's nodes in the read-back's lz, ly, lx loop order */
static double sum_nodes(const double *v8) {
  double s = 0.0;
  for (int q = 0; q < 8; q++) s = s + v8[q];
  return s;
}
/* Rocflu's read-back of a projected volume fraction, :826-842 and :1217-1234 */
static double phip_of(const double *vfP8, double vol) {
  double s = 0.0;
  for (int q = 0; q < 8; q++) s = s + (0.125 * vfP8[q]) * vol;
  s = s / vol;
  if (s > F32(0.62)) s = F32(0.62);
  return s;
}

int orc__picl_runge(orc_region *r, int istage) {
  struct orc_picl *p = r->picl;
  int N = r->nCells;
  double *fb = NULL, *fbg = NULL;
  if (p->fldRot) {
    /* :618-663: the feedback force per cell from the previous projection,
     * its LSQ gradient, WENO-XYZ and RFLU_LimitGradCellsSimple with the
     * variable flags (1, 2, 3): only the first (x) is VAR_INFO_POS */
    fb = malloc(sizeof(double) * 3 * (size_t)N);
    fbg = malloc(sizeof(double) * 9 * (size_t)N);
    for (int i = 0; i < N; i++) {
      const double *pf = p->proFld + 8 * NPRO * (size_t)i;
      for (int k = 0; k < 3; k++) fb[3 * (size_t)i + k] = sum_nodes(pf + 8 * (P_JFX + k)) * 0.125 * r->vol[i];
    }
    for (int k = 0; k < 3; k++) orc__gradient_field(r, fb + k, 3, fbg + 3 * k, 9);
    orc__weno_xyz(r, fbg, 9);
    orc__limit_simple_field(r, fbg, 9, fb, 3);
  }
  /* field assembly :712-842 (one value per element; LocalInterp's node
   * average of the 8 identical copies, :2363-2373, is sum8(v)/8) */
  for (int i = 0; i < N; i++) {
    const double *cv = r->cv + 5 * (size_t)i, *dv = r->dv + 3 * (size_t)i;
    const double *rhs = r->rhs + 5 * (size_t)i, *g = r->grad + 15 * (size_t)i;
    double vol = r->vol[i];
    double PhiP = phip_of(p->proFld + 8 * (NPRO * (size_t)i + P_JPHIP), vol);
    double ug[3] = {cv[1] / cv[0], cv[2] / cv[0], cv[3] / cv[0]};
    double sdr[3];
    for (int m = 0; m < 3; m++) {
      double dot = 0.0; /* DOT_PRODUCT(ug, pGc(:,2+m,i)) */
      for (int d = 0; d < 3; d++) dot = dot + ug[d] * g[3 * (1 + m) + d];
      sdr[m] = (-(rhs[1 + m] / vol) + ug[m] * rhs[0] / vol) / cv[0] + dot;
    }
    double vFrac = 1.0 - r->phi[i];
    double f[ORC_PICL_NFLD];
    f[0] = cv[0] / vFrac;
    f[1] = cv[1] / cv[0];
    f[2] = cv[2] / cv[0];
    f[3] = cv[3] / cv[0];
    f[4] = g[12]; f[5] = g[13]; f[6] = g[14];
    f[7] = dv[2];
    f[8] = dv[1];
    f[9] = PhiP;
    f[10] = sdr[0]; f[11] = sdr[1]; f[12] = sdr[2];
    f[13] = -(rhs[0] / vol);
    f[14] = g[0]; f[15] = g[1]; f[16] = g[2];
    f[17] = r->gradRhog[3 * (size_t)i];
    f[18] = r->gradRhog[3 * (size_t)i + 1];
    f[19] = r->gradRhog[3 * (size_t)i + 2];
    /* IDW reads LocalInterp's node average; the trilinear evaluation reads
     * the nodes themselves */
    for (int q = 0; q < ORC_PICL_NFLD; q++)
      p->fld[ORC_PICL_NFLD * (size_t)i + q] = p->pp.interpMode == 1 ? f[q] : sum8(f[q]) / 8;
    if (p->fldRot) {
      /* :747-809: vorticity and its material derivative (stretching,
       * dilatation, baroclinic, the feedback force's curl and its
       * misalignment with grad(rho phi)) */
      double dudx = g[3], dudy = g[4], dudz = g[5], dvdx = g[6], dvdy = g[7], dvdz = g[8];
      double dwdx = g[9], dwdy = g[10], dwdz = g[11];
      double omgx = dwdy - dvdz, omgy = dudz - dwdx, omgz = dvdx - dudy;
      double divu = dudx + dvdy + dwdz;
      double dprdx = g[0], dprdy = g[1], dprdz = g[2], dpdx = g[12], dpdy = g[13], dpdz = g[14];
      const double *fg = fbg + 9 * (size_t)i, *F = fb + 3 * (size_t)i;
      double ir = 1.0 / cv[0], ir2 = ir * ir;
      double dodx = omgx * dudx + omgy * dudy + omgz * dudz;
      double dody = omgx * dvdx + omgy * dvdy + omgz * dvdz;
      double dodz = omgx * dwdx + omgy * dwdy + omgz * dwdz;
      dodx = dodx - omgx * divu;
      dody = dody - omgy * divu;
      dodz = dodz - omgz * divu;
      dodx = dodx + (dprdy * dpdz - dprdz * dpdy) * ir2;
      dody = dody + (dprdz * dpdx - dprdx * dpdz) * ir2;
      dodz = dodz + (dprdx * dpdy - dprdy * dpdx) * ir2;
      dodx = dodx + (fg[6 + 1] - fg[3 + 2]) * ir; /* dFz/dy - dFy/dz */
      dody = dody + (fg[0 + 2] - fg[6 + 0]) * ir; /* dFx/dz - dFz/dx */
      dodz = dodz + (fg[3 + 0] - fg[0 + 1]) * ir; /* dFy/dx - dFx/dy */
      dodx = dodx + (dprdy * F[2] - dprdz * F[1]) * ir2;
      dody = dody + (dprdz * F[0] - dprdx * F[2]) * ir2;
      dodz = dodz + (dprdx * F[1] - dprdy * F[0]) * ir2;
      double fr[6] = {omgx, omgy, omgz, dodx, dody, dodz};
      for (int q = 0; q < 6; q++) p->fldRot[6 * (size_t)i + q] = p->pp.interpMode == 1 ? fr[q] : sum8(fr[q]) / 8;
    }
  }
  free(fb);
  free(fbg);
  p->rotErr = 0;
  integrate_particle(p, istage, r->dtMin);
  if (p->rotErr) return -1; /* ppiclf_exittr: "Re rotational too large" */
  if (p->pp.feedbackFlag == 1) { /* :926-1038 */
    for (int i = 0; i < N; i++) {
      const double *pf = p->proFld + 8 * NPRO * (size_t)i;
      double vol = r->vol[i];
      double jfx = sum_nodes(pf + 8 * P_JFX) * 0.125 * vol;
      double jfy = sum_nodes(pf + 8 * P_JFY) * 0.125 * vol;
      double jfz = sum_nodes(pf + 8 * P_JFZ) * 0.125 * vol;
      double jfe = sum_nodes(pf + 8 * P_JE) * 0.125 * vol;
      double *rhs = r->rhs + 5 * (size_t)i;
      rhs[1] = rhs[1] + jfx;
      rhs[2] = rhs[2] + jfy;
      rhs[3] = rhs[3] + jfz;
      rhs[4] = rhs[4] + jfe;
    }
  }
  for (int i = 0; i < N; i++) r->phi[i] = phip_of(p->proFld + 8 * (NPRO * (size_t)i + P_JPHIP), r->vol[i]);
  return 0;
}

/* ------------------------------------------------------------------------ */
int orc_picl_init(orc_region *r, const orc_picl_params *pp, int nPart, const double *y,
                  const double *pconst) {
  if (r->nvc != 8 || r->p.spaceOrder != 2) return -1;
  for (size_t q = 0; q < (size_t)8 * r->nCells; q++)
    if (r->c2v[q] < 0) return -1; /* all hexes */
  if (pp->qsFlag < 1 || pp->qsFlag > 5 || pp->interpMode < 0 || pp->interpMode > 1 || pp->projMode < 0 ||
      pp->projMode > 1 || pp->heatTransfer < 0 || pp->heatTransfer > 4 || pp->viscousUnsteady < 0 ||
      pp->viscousUnsteady > 1 || pp->collisionalFlag < 0 || pp->collisionalFlag > 4 || pp->amFlag < 0 ||
      pp->amFlag > 2 || pp->neighborWidth < 0.0 || pp->qsFluct < 0 || pp->qsFluct > 2 || pp->qsFluctAdapt < 0 ||
      pp->qsFluctAdapt > 1 || pp->fluctHalfWidth < 0.0)
    return -1;
  if (pp->collisionalFlag && !(pp->ksp > 0.0 && pp->erest > 0.0 && pp->erest <= 1.0)) return -1;
  const int perFlag[3] = {pp->periodicX, pp->periodicY, pp->periodicZ};
  double vlo[3] = {1e300, 1e300, 1e300}, vhi[3] = {-1e300, -1e300, -1e300};
  for (int v = 0; v < r->nVert; v++)
    for (int d = 0; d < 3; d++) {
      double x = r->xyz[3 * (size_t)v + d];
      if (x < vlo[d]) vlo[d] = x;
      if (x > vhi[d]) vhi[d] = x;
    }
  for (int d = 0; d < 3; d++) /* ppiclf_solve_InitPeriodicX/Y/Z: xl < xr */
    if (perFlag[d] < 0 || perFlag[d] > 1 || (perFlag[d] && !(vlo[d] < vhi[d]))) return -1;
  if (pp->projMode == 1 && !(pp->filterWidth > 0.0 && pp->filterAlpha > 0.0 && pp->filterAlpha < 1.0)) return -1;
  /* angular periodicity: not with periodicX/Y (PICL_TEMP_InitSolver.F90:
   * 700-705); the box projection and no pair models (no rotated ghosts) */
  if (pp->angularPeriodic < 0 || pp->angularPeriodic > 1) return -1;
  if (pp->angularPeriodic && (perFlag[0] || perFlag[1] || pp->projMode != 0 || pp->collisionalFlag ||
                              pp->amFlag == 2 || pp->qsFluct))
    return -1;
  struct orc_picl *p = calloc(1, sizeof *p);
  p->pp = *pp;
  if (pp->angularPeriodic) {
    if (angular_init(p, pp)) {
      free(p);
      return -1;
    }
    p->mark = calloc((size_t)nPart, 1);
  }
  /* PICL_TEMP_InitSolver.F90:234-262: the range of a periodic axis is the
   * mesh's vertex minimum and maximum on it */
  for (int d = 0; d < 3; d++) {
    p->per[d] = perFlag[d];
    p->perLo[d] = vlo[d];
    p->perHi[d] = vhi[d];
  }
  if (pp->projMode == 1) {
    /* ppiclf_solve_InitGaussianFilter (:1520-1576) and the constants of
     * ProjectParticleGrid (:3103-3109) */
    const double pi = 4.0 * atan(1.0);
    double rsig = pp->filterWidth / (2.0 * sqrt(2.0 * log(2.0)));
    double d2chk = rsig * sqrt(-2.0 * log(pp->filterAlpha));
    double s2 = sqrt(2.0 * pi);
    double multfci = 1.0 / ((s2 * s2) * (rsig * rsig));
    p->gMult = pow(multfci, 1.5);
    p->gRdum = 1.0 / (-2.0 * (rsig * rsig));
    p->gD2chk = d2chk * d2chk;
  }
  p->n = nPart;
  size_t n = nPart > 0 ? (size_t)nPart : 1;
  p->y = malloc(sizeof(double) * LRS * n);
  memcpy(p->y, y, sizeof(double) * LRS * (size_t)nPart);
  p->y1 = calloc(LRS * n, sizeof(double));
  p->ydot = calloc(LRS * n, sizeof(double));
  p->ydotc = calloc(LRS * n, sizeof(double));
  p->rprop = calloc(LRP * n, sizeof(double));
  p->elem = malloc(sizeof(int) * n);
  p->lost = calloc(n, 1);
  p->hist = calloc(6 * VU_N * n, sizeof(double)); /* ppiclf_solve.f:432-434 */
  p->nTimeBH = 1;                                  /* PICL_TEMP_InitSolver.F90:290 */
  if (pp->amFlag == 2 || pp->qsFluct >= 1) { /* PICL_TEMP_InitSolver.F90:655-661: max(4 d_max, NEIGHBORWIDTH) */
    double dmax = 0.0;
    for (int i = 0; i < nPart; i++) dmax = fmax(dmax, pconst[4 * (size_t)i + 1]);
    p->nbW = pp->neighborWidth > 0.0 ? pp->neighborWidth : 4.0 * dmax;
  }
  for (int i = 0; i < nPart; i++) {
    double *rp = p->rprop + LRP * (size_t)i;
    rp[R_JRHOP] = pconst[4 * (size_t)i];
    rp[R_JDP] = pconst[4 * (size_t)i + 1];
    rp[R_JVOLP] = pconst[4 * (size_t)i + 2];
    rp[R_JSPL] = pconst[4 * (size_t)i + 3];
  }
  setup_mesh(r, p);
  if (pp->collisionalFlag >= 3) p->fldRot = calloc(6 * (size_t)p->ne, sizeof(double));
  if (pp->qsFluct >= 1) {
    /* PICL_TEMP_InitSolver.F90:598-646: filter = 4 MAXVAL(Max_EleLen) / 2,
     * where Max_EleLen is reset for every element, so the last element's
     * extents decide (those outside [1e-7, 1e-2] count as 1.23456789e-10) */
    double mx = 0.0;
    const double *X = p->xm1b + 24 * (size_t)(p->ne - 1);
    for (int d = 0; d < 3; d++) {
      double lo = 1.0e10, hi = -1.0e10;
      for (int n = 0; n < 8; n++) {
        if (X[3 * n + d] > hi) hi = X[3 * n + d];
        if (X[3 * n + d] < lo) lo = X[3 * n + d];
      }
      double len = fabs(hi - lo), mel = 1.23456789e-10;
      if (!(len > 1e-2 || len < 1e-7) && len > mel) mel = len;
      if (d == 0 || mel > mx) mx = mel;
    }
    double filter = 4.0 * mx;
    filter = filter / 2.0;
    /* ppiclf_solve_InitBoxFilter(filter) halves it (ppiclf_solve.f:1611);
     * the Gaussian filter's set-up puts its cut-off radius there instead
     * (:1558-1559) */
    double w2 = filter / 2.0;
    if (pp->projMode == 1) {
      double rsig = pp->filterWidth / (2.0 * sqrt(2.0 * log(2.0)));
      w2 = rsig * sqrt(-2.0 * log(pp->filterAlpha));
    }
    p->fluctW = pp->fluctHalfWidth > 0.0 ? pp->fluctHalfWidth : w2;
  }
  r->picl = p;
  int N = r->nCells;
  r->phi = calloc((size_t)N, sizeof(double));
  r->gradRhog = calloc(3 * (size_t)N, sizeof(double));
  init_solve(p, 0); /* ppiclf_comm_InitOverlapMesh -> InitSolve (no fields yet) */
  /* PICL_TEMP_InitSolver.F90:789-835 */
  for (int i = 0; i < N; i++) r->phi[i] = phip_of(p->proFld + 8 * (NPRO * (size_t)i + P_JPHIP), r->vol[i]);
  for (int i = 0; i < N; i++) {
    double vFrac = 1.0 - r->phi[i];
    for (int k = 0; k < 5; k++) r->cv[5 * (size_t)i + k] = vFrac * r->cv[5 * (size_t)i + k];
  }
  orc__set_dependent_vars(r); /* RFLU_SetVarsWrapper after the PICL set-up */
  return 0;
}

int orc_picl_npart(const orc_region *r) { return r->picl ? r->picl->n : 0; }
int orc_picl_nlost(const orc_region *r) {
  int c = 0;
  if (r->picl)
    for (int i = 0; i < r->picl->n; i++) c += r->picl->lost[i];
  return c;
}
void orc_picl_get_particles(const orc_region *r, double *y, double *y1, double *ydot, double *ydotc,
                            int *elem) {
  const struct orc_picl *p = r->picl;
  size_t b = sizeof(double) * LRS * (size_t)p->n;
  if (y) memcpy(y, p->y, b);
  if (y1) memcpy(y1, p->y1, b);
  if (ydot) memcpy(ydot, p->ydot, b);
  if (ydotc) memcpy(ydotc, p->ydotc, b);
  if (elem)
    for (int i = 0; i < p->n; i++) elem[i] = (p->lost[i] || (p->mark && p->mark[i])) ? -1 : p->elem[i];
}
void orc_picl_get_rprop(const orc_region *r, double *rprop) {
  memcpy(rprop, r->picl->rprop, sizeof(double) * LRP * (size_t)r->picl->n);
}
void orc_picl_get_cells(const orc_region *r, double *phi, double *gradRhog, double *proFld,
                        double *fields) {
  int N = r->nCells;
  if (phi) memcpy(phi, r->phi, sizeof(double) * (size_t)N);
  if (gradRhog) memcpy(gradRhog, r->gradRhog, sizeof(double) * 3 * (size_t)N);
  if (proFld)
    for (int i = 0; i < N; i++)
      for (int jj = 0; jj < NPRO; jj++) proFld[NPRO * (size_t)i + jj] = r->picl->proFld[8 * (NPRO * (size_t)i + jj)];
  if (fields) memcpy(fields, r->picl->fld, sizeof(double) * ORC_PICL_NFLD * (size_t)N);
}
void orc_picl_get_nodes(const orc_region *r, double *proNode) {
  memcpy(proNode, r->picl->proFld, sizeof(double) * 8 * NPRO * (size_t)r->nCells);
}
void orc_picl_locate(const orc_region *r, int n, const double *xyz, int *elem) {
  for (int i = 0; i < n; i++) elem[i] = locate(r->picl, xyz + 3 * (size_t)i);
}
int orc_picl_neighbors(const orc_region *r, const double *xp, int *ids, double *d2) {
  return nearest27(r->picl, xp, ids, d2);
}
/* Test hook: particle i's fluctuation box sums (u, v, w, u^2, v^2, w^2,
 * count) over the current positions and velocities. */
void orc_picl_fluct_means(const orc_region *r, int i, double out[7]) { fluct_means(r->picl, i, out); }
