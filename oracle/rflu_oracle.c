/*
 * CPU oracle for the RocfluMP explicit residual path -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library, and only as the checker.  It is a plain-C restatement
 * of the reference's loops; every function cites the Fortran it follows.
 * Expression order follows the Fortran parse (left-to-right for equal
 * precedence) and the library is compiled with -ffp-contract=off because the
 * reference is built without FMA contraction (gfortran -O on x86-64).
 *
 * Parity pins for each part are listed in oracle/README.md.
 */
#include "rflu_oracle_int.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static char g_err[512];
const char *orc_last_error(void) { return g_err; }
#define FAIL(...) do { snprintf(g_err, sizeof g_err, __VA_ARGS__); goto fail; } while (0)

/* modflu/RFLU_ModGrid.F90:92-96 (0-based) */
static const int F2V_TET[4][4] = {{0, 1, 2, -1}, {1, 3, 2, -1}, {0, 2, 3, -1}, {0, 3, 1, -1}};
/* prisms and pyramids: f2vPri, f2vPyr of modflu/RFLU_ModGrid.F90:97-102 */
static const int F2V_PRI[5][4] = {{0, 2, 1, -1}, {0, 1, 4, 3}, {1, 2, 5, 4}, {0, 3, 5, 2}, {3, 4, 5, -1}};
static const int F2V_PYR[5][4] = {{0, 3, 2, 1}, {0, 1, 4, -1}, {1, 2, 4, -1}, {2, 3, 4, -1}, {0, 4, 3, -1}};
static const int F2V_HEX[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                  {2, 3, 7, 6}, {0, 4, 7, 3}, {4, 5, 6, 7}};

/* Cells: rows of nvc entries; with nvc = 8 a row may hold a tetrahedron (4),
 * pyramid (5), prism (6) or hexahedron (8) with -1 padding.  Returns the
 * vertex count and the local face table (-1 marks a triangle's 4th vertex). */
static int cell_nv(const orc_region *r, int i) {
  if (r->nvc == 4) return 4;
  int n = 0;
  for (int q = 0; q < 8; q++) n += r->c2v[(size_t)i * 8 + q] >= 0;
  return n;
}
static const int (*cell_faces(int nv, int *nlf))[4] {
  switch (nv) {
    case 4: *nlf = 4; return F2V_TET;
    case 5: *nlf = 5; return F2V_PYR;
    case 6: *nlf = 5; return F2V_PRI;
    default: *nlf = 6; return F2V_HEX;
  }
}

/* struct orc_region: rflu_oracle_int.h */

/* ------------------------------------------------------------------------ */
/* Sorting helpers                                                           */
/* ------------------------------------------------------------------------ */
static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* modfloflu/ModSortSearch.F90:898-990 (QuickSortRFREALInteger).  The pair
 * order among equal keys is whatever this partition scheme produces; the
 * reference's c2c stencil truncation depends on it, so the steps are kept,
 * written with 1-based indices as in the Fortran. */
static void quicksort_real_int(double *a0, int *b0, int n) {
  double *a = a0 - 1;
  int *b = b0 - 1;
  int stackl[51], stackr[51];
  int s = 1, l, r, i, j, k;
  stackl[1] = 1;
  stackr[1] = n;
  if (n < 1) return;
L10:
  l = stackl[s];
  r = stackr[s];
  s = s - 1;
L20:
  i = l;
  j = r;
  k = (l + r) / 2;
  {
    double x = a[k];
    for (;;) {
      while (a[i] < x) i = i + 1;
      while (x < a[j]) j = j - 1;
      if (i <= j) {
        double w = a[i]; a[i] = a[j]; a[j] = w;
        int v = b[i]; b[i] = b[j]; b[j] = v;
        i = i + 1;
        j = j - 1;
        if (i > j) break;
      } else {
        break;
      }
    }
  }
  if (j - l >= r - i) {
    if (l < j) { s = s + 1; stackl[s] = l; stackr[s] = j; }
    l = i;
  } else {
    if (i < r) { s = s + 1; stackl[s] = i; stackr[s] = r; }
    r = j;
  }
  if (l < r) goto L20;
  if (s != 0) goto L10;
}

/* ------------------------------------------------------------------------ */
/* Face list: modflu/RFLU_ModFaceList.F90:784-1585 + RFLU_ModHashTable.F90   */
/* Faces are numbered in first-insertion order over cells and local faces;   */
/* dedup on the three smallest sorted vertex ids; c1 = first cell, c2 = the  */
/* second; interior faces keep that order, boundary faces go to patches.     */
/* ------------------------------------------------------------------------ */
static uint64_t key3(const int *v) {
  return ((uint64_t)(uint32_t)v[0] * 0x9E3779B97F4A7C15ull) ^
         ((uint64_t)(uint32_t)v[1] * 0xC2B2AE3D27D4EB4Full) ^
         ((uint64_t)(uint32_t)v[2] * 0x165667B19E3779F9ull);
}

static void sort_small(int *v, int n) {
  for (int i = 1; i < n; i++) {
    int x = v[i], j = i - 1;
    while (j >= 0 && v[j] > x) { v[j + 1] = v[j]; j--; }
    v[j + 1] = x;
  }
}

static int build_faces(orc_region *r) {
  int maxF = r->nCells * 6;
  int *c1 = malloc(sizeof(int) * maxF), *c2 = malloc(sizeof(int) * maxF);
  int *lf = malloc(sizeof(int) * maxF), *k3 = malloc(sizeof(int) * 3 * maxF);
  size_t tsz = 1;
  while (tsz < (size_t)maxF * 2) tsz <<= 1;
  int *tab = malloc(sizeof(int) * tsz);
  int nF = 0, ok = 1;
  for (size_t t = 0; t < tsz; t++) tab[t] = -1;
  for (int icl = 0; icl < r->nCells; icl++) {
    int nlf;
    const int(*tbl)[4] = cell_faces(cell_nv(r, icl), &nlf);
    for (int ifl = 0; ifl < nlf; ifl++) {
      int fv[4], nfv = tbl[ifl][3] < 0 ? 3 : 4;
      for (int q = 0; q < nfv; q++) fv[q] = r->c2v[(size_t)icl * r->nvc + tbl[ifl][q]];
      sort_small(fv, nfv);
      size_t h = key3(fv) & (tsz - 1);
      for (;;) {
        int e = tab[h];
        if (e < 0) {
          tab[h] = nF;
          c1[nF] = icl; c2[nF] = -1; lf[nF] = ifl;
          k3[3 * nF] = fv[0]; k3[3 * nF + 1] = fv[1]; k3[3 * nF + 2] = fv[2];
          nF++;
          break;
        }
        if (k3[3 * e] == fv[0] && k3[3 * e + 1] == fv[1] && k3[3 * e + 2] == fv[2]) {
          c2[e] = icl;
          break;
        }
        h = (h + 1) & (tsz - 1);
      }
    }
  }
  /* interior faces in insertion order */
  int nInt = 0;
  for (int f = 0; f < nF; f++) if (c2[f] >= 0) nInt++;
  r->nFaces = nInt;
  r->f2c = malloc(sizeof(int) * 2 * (nInt > 0 ? nInt : 1));
  r->f2v = malloc(sizeof(int) * 4 * (nInt > 0 ? nInt : 1));
  int *old2int = malloc(sizeof(int) * nF);
  int fi = 0;
  for (int f = 0; f < nF; f++) {
    old2int[f] = -1;
    if (c2[f] < 0) continue;
    old2int[f] = fi;
    r->f2c[2 * fi] = c1[f];
    r->f2c[2 * fi + 1] = c2[f];
    int nlf;
    const int(*tbl)[4] = cell_faces(cell_nv(r, c1[f]), &nlf);
    for (int q = 0; q < 4; q++)
      r->f2v[4 * fi + q] = tbl[lf[f]][q] >= 0 ? r->c2v[(size_t)c1[f] * r->nvc + tbl[lf[f]][q]] : -1;
    fi++;
  }
  /* boundary faces: RFLU_UnHashBFace -> bf2c = c1 of the matching face */
  r->bf2c = malloc(sizeof(int) * (r->nBFaces > 0 ? r->nBFaces : 1));
  for (int b = 0; b < r->nBFaces; b++) {
    int fv[4], nfv = (r->nvf == 4 && r->bf2v[(size_t)b * 4 + 3] < 0) ? 3 : r->nvf;
    for (int q = 0; q < nfv; q++) fv[q] = r->bf2v[(size_t)b * r->nvf + q];
    sort_small(fv, nfv);
    size_t h = key3(fv) & (tsz - 1);
    int found = -1;
    for (;;) {
      int e = tab[h];
      if (e < 0) break;
      if (k3[3 * e] == fv[0] && k3[3 * e + 1] == fv[1] && k3[3 * e + 2] == fv[2]) { found = e; break; }
      h = (h + 1) & (tsz - 1);
    }
    if (found < 0 || c2[found] >= 0) { ok = 0; snprintf(g_err, sizeof g_err, "boundary face %d not matched", b); break; }
    r->bf2c[b] = c1[found];
  }
  /* every face seen once must be on a patch */
  if (ok) {
    int nB = 0;
    for (int f = 0; f < nF; f++) if (c2[f] < 0) nB++;
    if (nB != r->nBFaces) { ok = 0; snprintf(g_err, sizeof g_err, "%d open faces but %d patch faces", nB, r->nBFaces); }
  }
  free(c1); free(c2); free(lf); free(k3); free(tab); free(old2int);
  return ok;
}

/* ------------------------------------------------------------------------ */
/* Geometry: modflu/RFLU_ModGeometry.F90:307-640, libfloflu/FaceVector.F90, */
/* libfloflu/FaceCentroid.F90, RFLU_ComputeApproxCentroids (:1277-1497).     */
/* ------------------------------------------------------------------------ */
static void face_vector_tria(const double *x1, const double *x2, const double *x3, double *fv) {
  double dxa = x1[0] - x2[0], dya = x1[1] - x2[1], dza = x1[2] - x2[2];
  double dxb = x1[0] - x3[0], dyb = x1[1] - x3[1], dzb = x1[2] - x3[2];
  fv[0] = 0.5 * (dya * dzb - dyb * dza);
  fv[1] = -0.5 * (dxa * dzb - dxb * dza);
  fv[2] = 0.5 * (dxa * dyb - dxb * dya);
}
static void face_vector_quad(const double *x1, const double *x2, const double *x3, const double *x4, double *fv) {
  double dxa = x3[0] - x1[0], dya = x3[1] - x1[1], dza = x3[2] - x1[2];
  double dxb = x2[0] - x4[0], dyb = x2[1] - x4[1], dzb = x2[2] - x4[2];
  fv[0] = 0.5 * (dza * dyb - dya * dzb);
  fv[1] = 0.5 * (dxa * dzb - dza * dxb);
  fv[2] = 0.5 * (dya * dxb - dxa * dyb);
}
static void face_centroid(const double *const *x, int n, double *c) {
  const double thrd = 1.0 / 3.0;
  for (int d = 0; d < 3; d++) {
    if (n == 3) c[d] = thrd * (x[0][d] + x[1][d] + x[2][d]);
    else c[d] = 0.25 * (x[0][d] + x[1][d] + x[2][d] + x[3][d]);
  }
}

static void face_geom(const orc_region *r, const int *v, int nv, double *fvec, double *fcen) {
  const double *x[4];
  for (int q = 0; q < nv; q++) x[q] = r->xyz + 3 * (size_t)v[q];
  if (nv == 3) face_vector_tria(x[0], x[1], x[2], fvec);
  else face_vector_quad(x[0], x[1], x[2], x[3], fvec);
  face_centroid(x, nv, fcen);
}

static void build_geometry(orc_region *r) {
  int N = r->nCells;
  double *app = malloc(sizeof(double) * 3 * N);
  for (int i = 0; i < N; i++) {
    r->vol[i] = 0.0;
    r->cofg[3 * i] = r->cofg[3 * i + 1] = r->cofg[3 * i + 2] = 0.0;
    const int *v = r->c2v + (size_t)i * r->nvc;
    int nv = cell_nv(r, i);
    /* RFLU_ComputeApproxCentroids: term * (sum in vertex order), term = 1/nv */
    double term = nv == 4 ? 1.0 / 4.0 : nv == 5 ? 1.0 / 5.0 : nv == 6 ? 1.0 / 6.0 : 1.0 / 8.0;
    for (int d = 0; d < 3; d++) {
      double s = r->xyz[3 * v[0] + d];
      for (int q = 1; q < nv; q++) s = s + r->xyz[3 * v[q] + d];
      app[3 * i + d] = term * s;
    }
  }
  for (int f = 0; f < r->nFaces; f++) {
    double fv[3], fc[3];
    face_geom(r, r->f2v + 4 * f, r->f2v[4 * f + 3] < 0 ? 3 : 4, fv, fc);
    double fVecM = sqrt(fv[0] * fv[0] + fv[1] * fv[1] + fv[2] * fv[2]);
    double term = 1.0 / fVecM;
    r->fn[4 * f] = fv[0] * term;
    r->fn[4 * f + 1] = fv[1] * term;
    r->fn[4 * f + 2] = fv[2] * term;
    r->fn[4 * f + 3] = fVecM;
    r->fc[3 * f] = fc[0]; r->fc[3 * f + 1] = fc[1]; r->fc[3 * f + 2] = fc[2];
    int c1 = r->f2c[2 * f], c2 = r->f2c[2 * f + 1];
    term = (fc[0] - app[3 * c1]) * fv[0] + (fc[1] - app[3 * c1 + 1]) * fv[1] + (fc[2] - app[3 * c1 + 2]) * fv[2];
    r->vol[c1] = r->vol[c1] + term;
    term = fc[0] * fv[0] + fc[1] * fv[1] + fc[2] * fv[2];
    r->cofg[3 * c1] = r->cofg[3 * c1] + term * fc[0];
    r->cofg[3 * c1 + 1] = r->cofg[3 * c1 + 1] + term * fc[1];
    r->cofg[3 * c1 + 2] = r->cofg[3 * c1 + 2] + term * fc[2];
    term = (fc[0] - app[3 * c2]) * fv[0] + (fc[1] - app[3 * c2 + 1]) * fv[1] + (fc[2] - app[3 * c2 + 2]) * fv[2];
    r->vol[c2] = r->vol[c2] - term;
    term = fc[0] * fv[0] + fc[1] * fv[1] + fc[2] * fv[2];
    r->cofg[3 * c2] = r->cofg[3 * c2] - term * fc[0];
    r->cofg[3 * c2 + 1] = r->cofg[3 * c2 + 1] - term * fc[1];
    r->cofg[3 * c2 + 2] = r->cofg[3 * c2 + 2] - term * fc[2];
  }
  for (int b = 0; b < r->nBFaces; b++) {
    double fv[3], fc[3];
    const int *bv = r->bf2v + (size_t)b * r->nvf;
    face_geom(r, bv, (r->nvf == 4 && bv[3] < 0) ? 3 : r->nvf, fv, fc);
    double fVecM = sqrt(fv[0] * fv[0] + fv[1] * fv[1] + fv[2] * fv[2]);
    double term = 1.0 / fVecM;
    r->bfn[4 * b] = fv[0] * term;
    r->bfn[4 * b + 1] = fv[1] * term;
    r->bfn[4 * b + 2] = fv[2] * term;
    r->bfn[4 * b + 3] = fVecM;
    r->bfc[3 * b] = fc[0]; r->bfc[3 * b + 1] = fc[1]; r->bfc[3 * b + 2] = fc[2];
    int c1 = r->bf2c[b];
    term = (fc[0] - app[3 * c1]) * fv[0] + (fc[1] - app[3 * c1 + 1]) * fv[1] + (fc[2] - app[3 * c1 + 2]) * fv[2];
    r->vol[c1] = r->vol[c1] + term;
    term = fc[0] * fv[0] + fc[1] * fv[1] + fc[2] * fv[2];
    r->cofg[3 * c1] = r->cofg[3 * c1] + term * fc[0];
    r->cofg[3 * c1 + 1] = r->cofg[3 * c1 + 1] + term * fc[1];
    r->cofg[3 * c1 + 2] = r->cofg[3 * c1 + 2] + term * fc[2];
  }
  const double THRD = 1.0 / 3.0;
  for (int i = 0; i < N; i++) {
    r->vol[i] = THRD * r->vol[i];
    double term = 1.0 / (4.0 * r->vol[i]);
    r->cofg[3 * i] = term * r->cofg[3 * i];
    r->cofg[3 * i + 1] = term * r->cofg[3 * i + 1];
    r->cofg[3 * i + 2] = term * r->cofg[3 * i + 2];
  }
  free(app);
}

/* ------------------------------------------------------------------------ */
/* Singular-value test of the scaled stencil matrix (libflu/                 */
/* RFLU_InvertMatrixSVD.F90: s1 > 100 eps and s(i) > 0.01 s(i-1)), with the  */
/* singular values taken as square roots of the eigenvalues of A^T A         */
/* (cyclic Jacobi).  Returns 0 when the reference would flag singularity.    */
/* ------------------------------------------------------------------------ */
static int stencil_regular(double (*a)[4], int nRows, double *minRatio) {
  double colMax[4];
  for (int c = 0; c < 4; c++) {
    colMax[c] = -1.0e308;
    for (int q = 0; q < nRows; q++) colMax[c] = fmax(colMax[c], fabs(a[q][c]));
    for (int q = 0; q < nRows; q++) a[q][c] = a[q][c] / colMax[c];
  }
  double m[4][4] = {{0}}, v[4][4] = {{0}};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double s = 0.0;
      for (int q = 0; q < nRows; q++) s += a[q][i] * a[q][j];
      m[i][j] = s;
    }
  for (int i = 0; i < 4; i++) v[i][i] = 1.0;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.0;
    for (int i = 0; i < 4; i++) for (int j = i + 1; j < 4; j++) off += m[i][j] * m[i][j];
    if (off < 1e-30) break;
    for (int p = 0; p < 4; p++)
      for (int q = p + 1; q < 4; q++) {
        if (fabs(m[p][q]) < 1e-300) continue;
        double th = 0.5 * atan2(2.0 * m[p][q], m[q][q] - m[p][p]);
        double c = cos(th), s = sin(th);
        for (int k = 0; k < 4; k++) {
          double mkp = m[k][p], mkq = m[k][q];
          m[k][p] = c * mkp - s * mkq;
          m[k][q] = s * mkp + c * mkq;
        }
        for (int k = 0; k < 4; k++) {
          double mpk = m[p][k], mqk = m[q][k];
          m[p][k] = c * mpk - s * mqk;
          m[q][k] = s * mpk + c * mqk;
        }
      }
  }
  double sv[4];
  for (int i = 0; i < 4; i++) sv[i] = sqrt(fmax(m[i][i], 0.0));
  for (int i = 0; i < 4; i++)
    for (int j = i + 1; j < 4; j++)
      if (sv[j] > sv[i]) { double t = sv[i]; sv[i] = sv[j]; sv[j] = t; }
  double eps = 2.220446049250313e-16;
  if (!(sv[0] > 100.0 * eps)) return 0;
  double mr = 1e300;
  for (int i = 1; i < 4; i++) {
    double ratio = sv[i] / sv[i - 1];
    if (ratio < mr) mr = ratio;
    if (!(sv[i] > 0.01 * sv[i - 1])) { *minRatio = mr; return 0; }
  }
  if (mr < *minRatio) *minRatio = mr;
  return 1;
}

/* ------------------------------------------------------------------------ */
/* c2c stencils: modflu/RFLU_ModStencilsCells.F90:750-1182 (BuildC2CStencil) */
/* and :1207-1416 (BuildC2CStencilBasic); sizes from RFLU_SetInfoC2CStencil  */
/* (:2830-2873: min 4, max 10*4=40, 6 layers).                               */
/* Moments: modflu/RFLU_ModWeights.F90:1025-1160, 1770-1894.                 */
/* ------------------------------------------------------------------------ */
static int build_stencils(orc_region *r) {
  const int SMAX = 40, SMIN = 4;
  int N = r->nCells, nvc = r->nvc;
  int *v2cCnt = calloc(r->nVert + 1, sizeof(int));
  for (size_t q = 0; q < (size_t)N * nvc; q++) if (r->c2v[q] >= 0) v2cCnt[r->c2v[q] + 1]++;
  for (int v = 0; v < r->nVert; v++) v2cCnt[v + 1] += v2cCnt[v];
  int *v2c = malloc(sizeof(int) * (size_t)N * nvc), *fill = calloc(r->nVert, sizeof(int));
  for (int i = 0; i < N; i++)
    for (int q = 0; q < nvc; q++) {
      int v = r->c2v[(size_t)i * nvc + q];
      if (v >= 0) v2c[v2cCnt[v] + fill[v]++] = i;
    }
  free(fill);
  r->csOff = malloc(sizeof(int) * (N + 1));
  r->csMem = malloc(sizeof(int) * (size_t)N * SMAX);
  r->moms = malloc(sizeof(double) * 10 * (size_t)N);
  int cap = 1024, *tmp = malloc(sizeof(int) * cap);
  double *frac = malloc(sizeof(double) * cap);
  double (*a)[4] = malloc(sizeof(double) * 4 * SMAX);
  double minRatio = 1e300;
  int ok = 1, tot = 0;
  for (int icg = 0; icg < N; icg++) {
    int n = 0, nv = cell_nv(r, icg);
    double iNVert = 1.0 / (double)nv;  /* RFLU_BuildC2CStencilBasic: 1/nVert of the cell itself */
    for (int q = 0; q < nv; q++) {
      int v = r->c2v[(size_t)icg * nvc + q];
      for (int t = v2cCnt[v]; t < v2cCnt[v + 1]; t++) {
        if (n >= cap) { cap *= 2; tmp = realloc(tmp, sizeof(int) * cap); frac = realloc(frac, sizeof(double) * cap); }
        tmp[n++] = v2c[t];
      }
    }
    qsort(tmp, n, sizeof(int), cmp_int);
    int n2 = 0;
    for (int t = 0; t < n; t++) if (n2 == 0 || tmp[t] != tmp[n2 - 1]) tmp[n2++] = tmp[t];
    /* remove icg itself (BinarySearchInteger + RemoveInteger) */
    int m = 0;
    for (int t = 0; t < n2; t++) if (tmp[t] != icg) tmp[m++] = tmp[t];
    n2 = m;
    int vt[8];
    for (int q = 0; q < nv; q++) vt[q] = r->c2v[(size_t)icg * nvc + q];
    sort_small(vt, nv);
    for (int t = 0; t < n2; t++) {
      int v2[8], nc = 0, nv2 = cell_nv(r, tmp[t]);
      for (int q = 0; q < nv2; q++) v2[q] = r->c2v[(size_t)tmp[t] * nvc + q];
      sort_small(v2, nv2);
      int ia = 0, ib = 0;
      while (ia < nv && ib < nv2) {
        if (vt[ia] == v2[ib]) { nc++; ia++; ib++; }
        else if (vt[ia] < v2[ib]) ia++;
        else ib++;
      }
      frac[t] = (double)nc * iNVert;
    }
    quicksort_real_int(frac, tmp, n2);
    int degr = 0, lo = n2 - SMAX + 1;
    if (lo < 1) lo = 1;
    int *mem = r->csMem + tot;
    for (int q = n2; q >= lo; q--) mem[degr++] = tmp[q - 1];
    qsort(mem, degr, sizeof(int), cmp_int);
    if (degr < SMIN) { ok = 0; snprintf(g_err, sizeof g_err, "cell %d: stencil of %d < %d needs layer addition (unsupported)", icg, degr, SMIN); break; }
    for (int q = 0; q < degr; q++) {
      double dx = r->cofg[3 * mem[q]] - r->cofg[3 * icg];
      double dy = r->cofg[3 * mem[q] + 1] - r->cofg[3 * icg + 1];
      double dz = r->cofg[3 * mem[q] + 2] - r->cofg[3 * icg + 2];
      double term = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
      a[q][0] = term; a[q][1] = term * dx; a[q][2] = term * dy; a[q][3] = term * dz;
    }
    if (!stencil_regular(a, degr, &minRatio)) {
      ok = 0;
      snprintf(g_err, sizeof g_err, "cell %d: singular stencil needs layer addition (unsupported)", icg);
      break;
    }
    r->csOff[icg] = tot;
    tot += degr;
    /* RFLU_ComputeStencilMoments3D1 */
    double *xm = r->moms + 10 * (size_t)icg;
    for (int q = 0; q < 10; q++) xm[q] = 0.0;
    for (int q = 0; q < degr; q++) {
      double dx = r->cofg[3 * mem[q]] - r->cofg[3 * icg];
      double dy = r->cofg[3 * mem[q] + 1] - r->cofg[3 * icg + 1];
      double dz = r->cofg[3 * mem[q] + 2] - r->cofg[3 * icg + 2];
      double wt = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
      dx = wt * dx; dy = wt * dy; dz = wt * dz;
      xm[0] = xm[0] + dx * dx;
      xm[1] = xm[1] + dx * dy;
      xm[2] = xm[2] + dy * dy;
      xm[3] = xm[3] + dx * dz;
      xm[4] = xm[4] + dy * dz;
      xm[5] = xm[5] + dz * dz;
      xm[6] = xm[6] + wt * dx;
      xm[7] = xm[7] + wt * dy;
      xm[8] = xm[8] + wt * dz;
      xm[9] = xm[9] + wt * wt;
    }
    double r11 = sqrt(xm[0]);
    double ir11 = 1.0 / r11;
    double r12 = ir11 * xm[1];
    double r22 = sqrt(xm[2] - r12 * r12);
    double ir22 = 1.0 / r22;
    double r13 = ir11 * xm[3];
    double r23 = ir22 * (xm[4] - r12 * r13);
    double r33 = sqrt(xm[5] - (r13 * r13 + r23 * r23));
    double ir33 = 1.0 / r33;
    double r14 = ir11 * xm[6];
    double r24 = ir22 * (xm[7] - r12 * r14);
    double r34 = ir33 * (xm[8] - (r13 * r14 + r23 * r24));
    double r44 = sqrt(xm[9] - (r14 * r14 + r24 * r24 + r34 * r34));
    xm[0] = r11; xm[1] = r12; xm[2] = r22; xm[3] = r13; xm[4] = r23;
    xm[5] = r33; xm[6] = r14; xm[7] = r24; xm[8] = r34; xm[9] = r44;
  }
  r->csOff[N] = tot;
  free(v2cCnt); free(v2c); free(tmp); free(frac); free(a);
  return ok;
}

/* ------------------------------------------------------------------------ */
/* Thermodynamics: libfloflu/MixtPerf_*.F90                                  */
/* ------------------------------------------------------------------------ */
static inline double R_CpG(double cp, double g) { return cp - cp / g; }
static inline double R_M(double m) { return 8314.3 / m; }
static inline double M_R(double R) { return 8314.3 / R; }
static inline double G_CpR(double cp, double R) { return cp / (cp - R); }
static inline double T_DPR(double D, double P, double R) { return P / (D * R); }
static inline double C_GRT(double G, double R, double T) { return sqrt(G * R * T); }
static inline double C_DGP(double D, double G, double P) { return sqrt(G * P / D); }
static inline double C_GHoVm2(double G, double Ho, double Vm2) { return sqrt((G - 1.0) * (Ho - 0.5 * Vm2)); }
static inline double Ho_CpTUVW(double Cp, double T, double U, double V, double W, double K) {
  return Cp * T + 0.5 * (U * U + V * V + W * W) + K;
}
static inline double Eo_DGPUVW(double D, double G, double P, double U, double V, double W, double K) {
  return P / (D * (G - 1.0)) + 0.5 * (U * U + V * V + W * W) + K;
}
static inline double P_DEoGVm2(double D, double Eo, double G, double Vm2, double K) {
  return (G - 1.0) * D * (Eo - 0.5 * Vm2 - K);
}

/* libflu/RFLU_SetGasVars.F90:173-179 (TCPERF): gv(MOL)=8314.3/(cp-cp/g) */
static void set_gas_vars(orc_region *r) {
  double gc = R_CpG(r->p.refCp, r->p.refGamma);
  r->molGv = M_R(gc);
  r->cpGv = r->p.refCp;
}

/* libflu/RFLU_SetDependentVars.F90:256-287 */
static void set_dependent_vars(orc_region *r) {
  const double ksg = 0.0;
  for (int i = 0; i < r->nCells; i++) {
    const double *cv = r->cv + 5 * (size_t)i;
    double gc = R_M(r->molGv);
    double g = G_CpR(r->cpGv, gc);
    double rr = cv[0];
    double ir = 1.0 / rr;
    double Eo = ir * cv[4];
    double Vm2 = ir * ir * (cv[1] * cv[1] + cv[2] * cv[2] + cv[3] * cv[3]);
    /* PICL: p, T, c of the gas phase use rho/(1-phi) (:276-281) */
    if (r->phi) rr = rr / (1.0 - r->phi[i]);
    double *dv = r->dv + 3 * (size_t)i;
    dv[0] = P_DEoGVm2(rr, Eo, g, Vm2, ksg);
    dv[1] = T_DPR(rr, dv[0], gc);
    dv[2] = C_GRT(g, gc, dv[1]);
  }
}

/* ------------------------------------------------------------------------ */
/* dt: rocflu/RFLU_TimeStepInviscid.F90:152-327 (static grid: fs = 0)        */
/* ------------------------------------------------------------------------ */
double orc_timestep_inviscid(orc_region *r, double *dtCell) {
  double *dt = r->dt;
  const double fsu = 0.0;
  for (int i = 0; i < r->nCells; i++) dt[i] = 0.0;
  for (int f = 0; f < r->nFaces; f++) {
    int c1 = r->f2c[2 * f], c2 = r->f2c[2 * f + 1];
    double nx = r->fn[4 * f], ny = r->fn[4 * f + 1], nz = r->fn[4 * f + 2], nm = r->fn[4 * f + 3];
    const double *q1 = r->cv + 5 * (size_t)c1, *q2 = r->cv + 5 * (size_t)c2;
    double ir1 = 1.0 / q1[0], ir2 = 1.0 / q2[0];
    double u1 = q1[1] * ir1, v1 = q1[2] * ir1, w1 = q1[3] * ir1, a1 = r->dv[3 * (size_t)c1 + 2];
    double u2 = q2[1] * ir2, v2 = q2[2] * ir2, w2 = q2[3] * ir2, a2 = r->dv[3 * (size_t)c2 + 2];
    double vc = 0.5 * ((u1 + u2) * nx + (v1 + v2) * ny + (w1 + w2) * nz);
    double as = 0.5 * (a1 + a2);
    double srad = (fabs(vc - fsu) + as) * nm;
    dt[c1] = dt[c1] + srad;
    dt[c2] = dt[c2] + srad;
  }
  for (int b = 0; b < r->nBFaces; b++) {
    int c1 = r->bf2c[b];
    double nx = r->bfn[4 * b], ny = r->bfn[4 * b + 1], nz = r->bfn[4 * b + 2], nm = r->bfn[4 * b + 3];
    const double *q1 = r->cv + 5 * (size_t)c1;
    double ir1 = 1.0 / q1[0];
    double u1 = q1[1] * ir1, v1 = q1[2] * ir1, w1 = q1[3] * ir1, a1 = r->dv[3 * (size_t)c1 + 2];
    double vc = u1 * nx + v1 * ny + w1 * nz;
    dt[c1] = dt[c1] + (fabs(vc - fsu) + a1) * nm;
  }
  double dmin = 1.0e308;
  for (int i = 0; i < r->nCells; i++) {
    dt[i] = r->vol[i] / dt[i];
    if (dt[i] < dmin) dmin = dt[i];
  }
  if (dtCell) memcpy(dtCell, dt, sizeof(double) * r->nCells);
  return dmin;
}

/* ------------------------------------------------------------------------ */
/* cons <-> prim: modflu/RFLU_ModConvertCv.F90:221-231, 554-592 (no PICL)    */
/* ------------------------------------------------------------------------ */
static void cons2prim(orc_region *r) {
  for (int i = 0; i < r->nCells; i++) {
    double *cv = r->cv + 5 * (size_t)i;
    double ir = 1.0 / cv[0];
    cv[1] = ir * cv[1];
    cv[2] = ir * cv[2];
    cv[3] = ir * cv[3];
    cv[4] = r->dv[3 * (size_t)i];
  }
}
static void prim2cons(orc_region *r) {
  const double ksg = 0.0;
  for (int i = 0; i < r->nCells; i++) {
    double *cv = r->cv + 5 * (size_t)i;
    double rr = cv[0], u = cv[1], v = cv[2], w = cv[3], p = r->dv[3 * (size_t)i];
    cv[1] = rr * u;
    cv[2] = rr * v;
    cv[3] = rr * w;
    double gc = R_M(r->molGv);
    double g = G_CpR(r->cpGv, gc);
    /* PICL: energy from the gas-phase density r/(1-phi) (:584-586) */
    double rPure = r->phi ? rr / (1.0 - r->phi[i]) : rr;
    cv[4] = rr * Eo_DGPUVW(rPure, g, p, u, v, w, ksg);
  }
}

/* ------------------------------------------------------------------------ */
/* LSQ gradients: modflu/RFLU_ModDifferentiationCells.F90:955-1174           */
/* ------------------------------------------------------------------------ */
static void gradients(orc_region *r) {
#pragma omp parallel for schedule(static)
  for (int icg = 0; icg < r->nCells; icg++) {
    const double *m = r->moms + 10 * (size_t)icg;
    double r11 = m[0], r12 = m[1], r22 = m[2], r13 = m[3], r23 = m[4], r33 = m[5];
    double c11 = 1.0 / r11, c22 = 1.0 / r22, c33 = 1.0 / r33;
    double c12 = -(c11 * r12);
    double c13 = -(c11 * r13 + c12 * c22 * r23);
    double c23 = -(c22 * r23);
    double c1111 = c11 * c11, c2222 = c22 * c22, c3333 = c33 * c33;
    double *g = r->grad + 15 * (size_t)icg;
    for (int q = 0; q < 15; q++) g[q] = 0.0;
    const double *x0 = r->cofg + 3 * (size_t)icg;
    const double *v0 = r->cv + 5 * (size_t)icg;
    for (int s = r->csOff[icg]; s < r->csOff[icg + 1]; s++) {
      int j = r->csMem[s];
      double dx = r->cofg[3 * (size_t)j] - x0[0];
      double dy = r->cofg[3 * (size_t)j + 1] - x0[1];
      double dz = r->cofg[3 * (size_t)j + 2] - x0[2];
      double term = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
      dx = term * dx; dy = term * dy; dz = term * dz;
      double term1 = c1111 * (dx);
      double term2 = c2222 * (dy + c12 * dx);
      double term3 = c3333 * (dz + c23 * dy + c13 * dx);
      double wtx = term * (term1 + c12 * term2 + c13 * term3);
      double wty = term * (term2 + c23 * term3);
      double wtz = term * (term3);
      const double *vj = r->cv + 5 * (size_t)j;
      for (int k = 0; k < 5; k++) {
        double dVar = vj[k] - v0[k];
        g[3 * k] = g[3 * k] + wtx * dVar;
        g[3 * k + 1] = g[3 * k + 1] + wty * dVar;
        g[3 * k + 2] = g[3 * k + 2] + wtz * dVar;
      }
    }
  }
}

/* One face's Venkatakrishnan term, modflu/RFLU_ModLimiters.F90:459-505:
 * dm the neighbour extreme minus the cell value, d2 the gradient's
 * extrapolation to the face centroid, epsq = (K vol^(1/3))^3; TINY carries
 * tinyp for denom >= 0 and tinyn otherwise. */
static inline double venkat_term(double dm, double d2, double epsq, double tinyp, double tinyn) {
  double numer = (dm * dm + epsq) * d2 + 2.0 * d2 * d2 * dm;
  double denom = d2 * (dm * dm + 2.0 * d2 * d2 + dm * d2 + epsq);
  return numer / (denom + (denom >= 0.0 ? tinyp : tinyn));
}

/* Venkatakrishnan: modflu/RFLU_ModLimiters.F90:339-532, 707-760 */
static void limiter_venkat(orc_region *r) {
  int N = r->nCells;
  const double THRD = 1.0 / 3.0, K = 1.0;
  const double TINYP = 1.0e-12, TINYN = r->p.limiterGuard ? -1.0e-12 : 1.0e-12;
  double *vmin = malloc(sizeof(double) * 5 * N), *vmax = malloc(sizeof(double) * 5 * N);
  for (int i = 0; i < N; i++)
    for (int k = 0; k < 5; k++) {
      vmax[5 * i + k] = vmin[5 * i + k] = r->cv[5 * (size_t)i + k];
      r->lim[5 * (size_t)i + k] = 1.0;
    }
  for (int f = 0; f < r->nFaces; f++) {
    int c1 = r->f2c[2 * f], c2 = r->f2c[2 * f + 1];
    for (int k = 0; k < 5; k++) {
      double a = r->cv[5 * (size_t)c1 + k], b = r->cv[5 * (size_t)c2 + k];
      vmin[5 * c1 + k] = fmin(vmin[5 * c1 + k], b);
      vmin[5 * c2 + k] = fmin(vmin[5 * c2 + k], a);
      vmax[5 * c1 + k] = fmax(vmax[5 * c1 + k], b);
      vmax[5 * c2 + k] = fmax(vmax[5 * c2 + k], a);
    }
  }
  double *tbuf = malloc(sizeof(double) * 10 * (size_t)(r->nFaces > 0 ? r->nFaces : 1));
#pragma omp parallel for schedule(static)
  for (int f = 0; f < r->nFaces; f++) {
    int c1 = r->f2c[2 * f], c2 = r->f2c[2 * f + 1];
    double xc = r->fc[3 * f], yc = r->fc[3 * f + 1], zc = r->fc[3 * f + 2];
    double dx1 = r->cofg[3 * c1] - xc, dy1 = r->cofg[3 * c1 + 1] - yc, dz1 = r->cofg[3 * c1 + 2] - zc;
    double dx2 = r->cofg[3 * c2] - xc, dy2 = r->cofg[3 * c2 + 1] - yc, dz2 = r->cofg[3 * c2 + 2] - zc;
    double ds1 = pow(r->vol[c1], THRD), ds2 = pow(r->vol[c2], THRD);
    double epsq1 = (K * ds1) * (K * ds1) * (K * ds1);
    double epsq2 = (K * ds2) * (K * ds2) * (K * ds2);
    for (int k = 0; k < 5; k++) {
      double var1 = r->cv[5 * (size_t)c1 + k], var2 = r->cv[5 * (size_t)c2 + k];
      double d1min1 = vmin[5 * c1 + k] - var1, d1max1 = vmax[5 * c1 + k] - var1;
      double d1min2 = vmin[5 * c2 + k] - var2, d1max2 = vmax[5 * c2 + k] - var2;
      const double *g1 = r->grad + 15 * (size_t)c1 + 3 * k, *g2 = r->grad + 15 * (size_t)c2 + 3 * k;
      double d21 = g1[0] * dx1 + g1[1] * dy1 + g1[2] * dz1;
      double d22 = g2[0] * dx2 + g2[1] * dy2 + g2[2] * dz2;
      /* +inf stands for "no term" (d = 0): fmin(+inf, l) = l */
      double *t = tbuf + 10 * (size_t)f;
      t[k] = d21 > 0.0 ? venkat_term(d1max1, d21, epsq1, TINYP, TINYN)
           : d21 < 0.0 ? venkat_term(d1min1, d21, epsq1, TINYP, TINYN) : INFINITY;
      t[5 + k] = d22 > 0.0 ? venkat_term(d1max2, d22, epsq2, TINYP, TINYN)
               : d22 < 0.0 ? venkat_term(d1min2, d22, epsq2, TINYP, TINYN) : INFINITY;
    }
  }
  /* the per-face terms were formed in parallel; the minima in face order */
  for (int f = 0; f < r->nFaces; f++) {
    int c1 = r->f2c[2 * f], c2 = r->f2c[2 * f + 1];
    const double *t = tbuf + 10 * (size_t)f;
    for (int k = 0; k < 5; k++) {
      double *l1 = &r->lim[5 * (size_t)c1 + k], *l2 = &r->lim[5 * (size_t)c2 + k];
      *l1 = fmin(t[k], *l1);
      *l2 = fmin(t[5 + k], *l2);
    }
  }
  free(tbuf);
  for (int i = 0; i < N; i++)
    for (int k = 0; k < 5; k++) {
      double l = r->lim[5 * (size_t)i + k];
      double *g = r->grad + 15 * (size_t)i + 3 * k;
      g[0] = l * g[0]; g[1] = l * g[1]; g[2] = l * g[2];
    }
  free(vmin); free(vmax);
}

/* LSQ gradient of cv(1) alone: RFLU_ComputeGradCellsFast_3D with
 * iBegVar = iEndVar = 1 (libfloflu/CellGradientsMP.F90:375-378) */
/* The LSQ gradient (RFLU_ComputeGradCellsWrapper) of one cell field,
 * v[vs * cell], into out[os * cell + 0..2]. */
void orc__gradient_field(orc_region *r, const double *v, int vs, double *out, int os) {
  for (int icg = 0; icg < r->nCells; icg++) {
    const double *m = r->moms + 10 * (size_t)icg;
    double r11 = m[0], r12 = m[1], r22 = m[2], r13 = m[3], r23 = m[4], r33 = m[5];
    double c11 = 1.0 / r11, c22 = 1.0 / r22, c33 = 1.0 / r33;
    double c12 = -(c11 * r12);
    double c13 = -(c11 * r13 + c12 * c22 * r23);
    double c23 = -(c22 * r23);
    double c1111 = c11 * c11, c2222 = c22 * c22, c3333 = c33 * c33;
    double *g = out + (size_t)os * icg;
    g[0] = g[1] = g[2] = 0.0;
    const double *x0 = r->cofg + 3 * (size_t)icg;
    double v0 = v[(size_t)vs * icg];
    for (int s = r->csOff[icg]; s < r->csOff[icg + 1]; s++) {
      int j = r->csMem[s];
      double dx = r->cofg[3 * (size_t)j] - x0[0];
      double dy = r->cofg[3 * (size_t)j + 1] - x0[1];
      double dz = r->cofg[3 * (size_t)j + 2] - x0[2];
      double term = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
      dx = term * dx; dy = term * dy; dz = term * dz;
      double term1 = c1111 * (dx);
      double term2 = c2222 * (dy + c12 * dx);
      double term3 = c3333 * (dz + c23 * dy + c13 * dx);
      double wtx = term * (term1 + c12 * term2 + c13 * term3);
      double wty = term * (term2 + c23 * term3);
      double wtz = term * (term3);
      double dVar = v[(size_t)vs * j] - v0;
      g[0] = g[0] + wtx * dVar;
      g[1] = g[1] + wty * dVar;
      g[2] = g[2] + wtz * dVar;
    }
  }
}

/* the gradient of cv(1) (rho_g while CellGradientsMP holds it there) */
static void gradient_var1(orc_region *r, double *out) { orc__gradient_field(r, r->cv, 5, out, 3); }

/* RFLU_WENOGradCellsXYZFast_3D, one gradient (modflu/RFLU_ModWENO.F90:1402-1588):
 * per component, the average of the cell's own and its stencil members'
 * gradients weighted by 1/(g^2 + 1e-15), accumulated self first, then the
 * members in stencil order; all cells read the unmodified input. */
static void weno_xyz_n(orc_region *r, double *grad, int stride) {
  int N = r->nCells;
  double *eno = malloc(sizeof(double) * (size_t)stride * (size_t)N);
  for (int icg = 0; icg < N; icg++) {
    for (int d = 0; d < stride; d++) {
      double term = grad[(size_t)stride * icg + d];
      double sum = 1.0 / (term * term + 1.0E-15);
      double loc = sum * term;
      for (int s = r->csOff[icg]; s < r->csOff[icg + 1]; s++) {
        term = grad[(size_t)stride * r->csMem[s] + d];
        double w = 1.0 / (term * term + 1.0E-15);
        sum = sum + w;
        loc = loc + w * term;
      }
      eno[(size_t)stride * icg + d] = loc / sum;
    }
  }
  memcpy(grad, eno, sizeof(double) * (size_t)stride * (size_t)N);
  free(eno);
}
static void weno_xyz(orc_region *r, double *grad) { weno_xyz_n(r, grad, 3); }
void orc__weno_xyz(orc_region *r, double *grad, int stride) { weno_xyz_n(r, grad, stride); }

/* RFLU_LimitGradCellsSimple, one VAR_INFO_POS variable
 * (modflu/RFLU_ModLimiters.F90:793-921): the gradient is zeroed when the
 * extrapolation to any face centroid of the cell (interior or boundary) is
 * <= 0.  The face loop exits at the first such face and only then changes
 * the gradient, so the outcome does not depend on the face order. */
/* grad: the variable's 3 components at grad[stride * cell]; var: its index in cv */
static void limit_simple_var(orc_region *r, double *grad, int stride, int var) {
  int N = r->nCells;
  char *zero = calloc((size_t)N, 1);
  for (int f = 0; f < r->nFaces + r->nBFaces; f++) {
    int nc = f < r->nFaces ? 2 : 1;
    const double *fc = f < r->nFaces ? r->fc + 3 * (size_t)f : r->bfc + 3 * (size_t)(f - r->nFaces);
    for (int s = 0; s < nc; s++) {
      int c = f < r->nFaces ? r->f2c[2 * f + s] : r->bf2c[f - r->nFaces];
      const double *x = r->cofg + 3 * (size_t)c, *g = grad + (size_t)stride * c;
      double dx = fc[0] - x[0], dy = fc[1] - x[1], dz = fc[2] - x[2];
      double varf = r->cv[5 * (size_t)c + var] + g[0] * dx + g[1] * dy + g[2] * dz;
      if (varf <= 0.0) zero[c] = 1;
    }
  }
  for (int i = 0; i < N; i++)
    if (zero[i]) grad[(size_t)stride * i] = grad[(size_t)stride * i + 1] = grad[(size_t)stride * i + 2] = 0.0;
  free(zero);
}
static void limit_simple(orc_region *r, double *grad) { limit_simple_var(r, grad, 3, 0); }

/* The same for a cell field that is not in cv: vals[vstride * cell]. */
void orc__limit_simple_field(orc_region *r, double *grad, int stride, const double *vals, int vstride) {
  int N = r->nCells;
  char *zero = calloc((size_t)N, 1);
  for (int f = 0; f < r->nFaces + r->nBFaces; f++) {
    int nc = f < r->nFaces ? 2 : 1;
    const double *fc = f < r->nFaces ? r->fc + 3 * (size_t)f : r->bfc + 3 * (size_t)(f - r->nFaces);
    for (int s = 0; s < nc; s++) {
      int c = f < r->nFaces ? r->f2c[2 * f + s] : r->bf2c[f - r->nFaces];
      const double *x = r->cofg + 3 * (size_t)c, *g = grad + (size_t)stride * c;
      double dx = fc[0] - x[0], dy = fc[1] - x[1], dz = fc[2] - x[2];
      double varf = vals[(size_t)vstride * c] + g[0] * dx + g[1] * dy + g[2] * dz;
      if (varf <= 0.0) zero[c] = 1;
    }
  }
  for (int i = 0; i < N; i++)
    if (zero[i]) grad[(size_t)stride * i] = grad[(size_t)stride * i + 1] = grad[(size_t)stride * i + 2] = 0.0;
  free(zero);
}

/* CellGradientsMP: libfloflu/CellGradientsMP.F90:158-253, and the PICL
 * block :355-390 (gradient of rho_g = rho/(1-phi), WENO-XYZ, positivity
 * limiter; cv(1) is divided and multiplied back in place, so it keeps the
 * round-tripped bits) */
static void cell_gradients(orc_region *r) {
  if (r->p.spaceOrder > 1) {
    cons2prim(r);
    gradients(r);
    if (r->p.reconst == ORC_RECONST_VENKAT) limiter_venkat(r);
    if (r->p.reconst == ORC_RECONST_WENO_XYZ) {
      /* :210-215: RFLU_WENOGradCellsXYZWrapper, then RFLU_LimitGradCellsSimple
       * on the VAR_INFO_POS variables of the DUVWP state, rho and p */
      weno_xyz_n(r, r->grad, 15);
      limit_simple_var(r, r->grad, 15, 0);
      limit_simple_var(r, r->grad + 12, 15, 4);
    }
    prim2cons(r);
  }
  if (r->phi) {
    int N = r->nCells;
    double *vfg = malloc(sizeof(double) * (size_t)N);
    for (int i = 0; i < N; i++) {
      vfg[i] = 1.0 - r->phi[i];
      r->cv[5 * (size_t)i] = r->cv[5 * (size_t)i] / vfg[i];
    }
    gradient_var1(r, r->gradRhog);
    weno_xyz(r, r->gradRhog);
    limit_simple(r, r->gradRhog);
    for (int i = 0; i < N; i++) r->cv[5 * (size_t)i] = r->cv[5 * (size_t)i] * vfg[i];
    free(vfg);
  }
}

/* ------------------------------------------------------------------------ */
/* Flux functions                                                            */
/* ------------------------------------------------------------------------ */
/* modflu/RFLU_ModAUSMFlux.F90:282-367 */
static void ausm_flux_function(double nx, double ny, double nz, double nm, double fs,
                               double rl, double ul, double vl, double wl, double pl, double Hl, double al,
                               double rr, double ur, double vr, double wr, double pr, double Hr, double ar,
                               double *flx) {
  double ql = ul * nx + vl * ny + wl * nz - fs;
  double qr = ur * nx + vr * ny + wr * nz - fs;
  double af = 0.5 * (al + ar);
  double ml = ql / af, mla = fabs(ml);
  double mr = qr / af, mra = fabs(mr);
  double mlp, wtl, mrm, wtr;
  if (mla <= 1.0) {
    mlp = 0.25 * (ml + 1.0) * (ml + 1.0) + 0.125 * (ml * ml - 1.0) * (ml * ml - 1.0);
    wtl = 0.25 * (ml + 1.0) * (ml + 1.0) * (2.0 - ml) + 0.1875 * ml * (ml * ml - 1.0) * (ml * ml - 1.0);
  } else {
    mlp = 0.5 * (ml + mla);
    wtl = 0.5 * (1.0 + ml / mla);
  }
  if (mra <= 1.0) {
    mrm = -0.25 * (mr - 1.0) * (mr - 1.0) - 0.125 * (mr * mr - 1.0) * (mr * mr - 1.0);
    wtr = 0.25 * (mr - 1.0) * (mr - 1.0) * (2.0 + mr) - 0.1875 * mr * (mr * mr - 1.0) * (mr * mr - 1.0);
  } else {
    mrm = 0.5 * (mr - mra);
    wtr = 0.5 * (1.0 - mr / mra);
  }
  double mf = mlp + mrm, mfa = fabs(mf);
  double mfp = 0.5 * (mf + mfa), mfm = 0.5 * (mf - mfa);
  double pf = wtl * pl + wtr * pr;
  flx[0] = (af * (mfp * rl + mfm * rr)) * nm;
  flx[1] = (af * (mfp * rl * ul + mfm * rr * ur) + pf * nx) * nm;
  flx[2] = (af * (mfp * rl * vl + mfm * rr * vr) + pf * ny) * nm;
  flx[3] = (af * (mfp * rl * wl + mfm * rr * wr) + pf * nz) * nm;
  flx[4] = (af * (mfp * rl * Hl + mfm * rr * Hr) + pf * fs) * nm;
}

void orc_flux_ausm(const double *n, double fs, const double *L, const double *R, double *flx) {
  ausm_flux_function(n[0], n[1], n[2], n[3], fs, L[0], L[1], L[2], L[3], L[4], L[5], L[6],
                     R[0], R[1], R[2], R[3], R[4], R[5], R[6], flx);
}

/* modflu/RFLU_ModEntropyFixes.F90:95-106 */
static inline double entropy_fix_hh(double l, double d) {
  return (l > d) ? l : (l * l + d * d) / (2.0 * d);
}

/* Roe: modflu/RFLU_ModRoeFlux.F90:2783-2855 (shared by :1723 first order) */
static void roe_flux(const orc_params *p, double nx, double ny, double nz, double nm, double fs,
                     double rl, double ul, double vl, double wl, double pl, double Hl, double ql,
                     double rr, double ur, double vr, double wr, double pr, double Hr, double qr,
                     double *f) {
  double g = p->refGamma;
  double rh = sqrt(rl * rr);
  double wt = rl / (rl + rh);
  double uh = wt * ul + (1.0 - wt) * ur;
  double vh = wt * vl + (1.0 - wt) * vr;
  double wh = wt * wl + (1.0 - wt) * wr;
  double Hh = wt * Hl + (1.0 - wt) * Hr;
  double qh = uh * nx + vh * ny + wh * nz - fs;
  double sh = 0.5 * (uh * uh + vh * vh + wh * wh);
  double ah = C_GHoVm2(g, Hh, 2.0 * sh);
  double l1 = fabs(qh - ah), l2 = fabs(qh), l5 = fabs(qh + ah);
  double efc = p->epsentr * l5;
  l1 = entropy_fix_hh(l1, efc);
  l2 = entropy_fix_hh(l2, efc);
  l5 = entropy_fix_hh(l5, efc);
  double dr = rr - rl, du = ur - ul, de = vr - vl, dw = wr - wl, dp = pr - pl;
  double dq = du * nx + de * ny + dw * nz;
  double dv1 = (dp - rh * ah * dq) / (2.0 * ah * ah);
  double dv5 = (dp + rh * ah * dq) / (2.0 * ah * ah);
  double dv2 = dr - dp / (ah * ah);
  double t1 = l1 * dv1, t2 = l2 * dv2, t3 = l2 * rh, t5 = l5 * dv5;
  double fc[5], fd[5];
  fc[0] = 0.5 * (ql * rl + qr * rr) * nm;
  fc[1] = 0.5 * (ql * rl * ul + pl * nx + qr * rr * ur + pr * nx) * nm;
  fc[2] = 0.5 * (ql * rl * vl + pl * ny + qr * rr * vr + pr * ny) * nm;
  fc[3] = 0.5 * (ql * rl * wl + pl * nz + qr * rr * wr + pr * nz) * nm;
  fc[4] = 0.5 * (ql * rl * Hl + pl * fs + qr * rr * Hr + pr * fs) * nm;
  double df = p->dissFact;
  fd[0] = 0.5 * df * (t1 + t2 + t5) * nm;
  fd[1] = 0.5 * df * (t1 * (uh - nx * ah) + t2 * uh + t3 * (du - nx * dq) + t5 * (uh + nx * ah)) * nm;
  fd[2] = 0.5 * df * (t1 * (vh - ny * ah) + t2 * vh + t3 * (de - ny * dq) + t5 * (vh + ny * ah)) * nm;
  fd[3] = 0.5 * df * (t1 * (wh - nz * ah) + t2 * wh + t3 * (dw - nz * dq) + t5 * (wh + nz * ah)) * nm;
  fd[4] = 0.5 * df * (t1 * (Hh - qh * ah) + t2 * sh + t3 * (uh * du + vh * de + wh * dw - qh * dq) + t5 * (Hh + qh * ah)) * nm;
  for (int k = 0; k < 5; k++) f[k] = fc[k] - fd[k];
}

/* HLLC: modflu/RFLU_ModHLLCFlux.F90:2338-2409 (shared with :937 first order) */
static void hllc_flux(const orc_params *p, double nx, double ny, double nz, double nm, double fs,
                      double rl, double ul, double vl, double wl, double pl, double el, double al, double ql, double Hl,
                      double rr, double ur, double vr, double wr, double pr, double er, double ar, double qr, double Hr,
                      double *flx) {
  double g = p->refGamma;
  double wt1 = sqrt(rr / rl);
  double wt2 = 1.0 / (1.0 + wt1);
  double us = wt2 * (ul + wt1 * ur);
  double vs = wt2 * (vl + wt1 * vr);
  double ws = wt2 * (wl + wt1 * wr);
  double qs = wt2 * (ql + wt1 * qr);
  double Hs = wt2 * (Hl + wt1 * Hr);
  double Vms = us * us + vs * vs + ws * ws;
  double as = C_GHoVm2(g, Hs, Vms);
  double sl = fmin(ql - al, qs - as);
  double sr = fmax(qr + ar, qs + as);
  if (sl > 0.0) {
    flx[0] = ql * rl * nm;
    flx[1] = (ql * rl * ul + pl * nx) * nm;
    flx[2] = (ql * rl * vl + pl * ny) * nm;
    flx[3] = (ql * rl * wl + pl * nz) * nm;
    flx[4] = (ql * (rl * el + pl) + pl * fs) * nm;
  } else if (sr < 0.0) {
    flx[0] = qr * rr * nm;
    flx[1] = (qr * rr * ur + pr * nx) * nm;
    flx[2] = (qr * rr * vr + pr * ny) * nm;
    flx[3] = (qr * rr * wr + pr * nz) * nm;
    flx[4] = (qr * (rr * er + pr) + pr * fs) * nm;
  } else {
    double sm = (rr * qr * (sr - qr) - rl * ql * (sl - ql) - pr + pl) / (rr * (sr - qr) - rl * (sl - ql));
    double ps = pl + rl * (ql - sl) * (ql - sm);
    double rs, rus, rvs, rws, res, term;
    if (sm > 0.0) {
      term = 1.0 / (sl - sm);
      rs = term * (sl - ql) * rl;
      rus = term * ((sl - ql) * rl * ul + (ps - pl) * nx);
      rvs = term * ((sl - ql) * rl * vl + (ps - pl) * ny);
      rws = term * ((sl - ql) * rl * wl + (ps - pl) * nz);
      res = term * ((sl - ql) * rl * el + (ps * sm - pl * ql));
    } else {
      term = 1.0 / (sr - sm);
      rs = term * (sr - qr) * rr;
      rus = term * ((sr - qr) * rr * ur + (ps - pr) * nx);
      rvs = term * ((sr - qr) * rr * vr + (ps - pr) * ny);
      rws = term * ((sr - qr) * rr * wr + (ps - pr) * nz);
      res = term * ((sr - qr) * rr * er + (ps * sm - pr * qr));
    }
    flx[0] = sm * rs * nm;
    flx[1] = (sm * rus + ps * nx) * nm;
    flx[2] = (sm * rvs + ps * ny) * nm;
    flx[3] = (sm * rws + ps * nz) * nm;
    flx[4] = (sm * (res + ps) - ps * fs) * nm;
  }
}

/* RFLU_AUSMPlusUp_FluxFunction2 (modflu/RFLU_ModAUSMPlusUpFlux.F90:469-588),
 * the function both TCP drivers call.  It overwrites the Mach-number
 * function fa with 1 (:519), so mInf and the M_o terms have no effect and
 * alpha = 3/16; vFracF scales the convected part (the face's gas volume
 * fraction with particles, else 1). */
static void ausmup_flux_function(const orc_params *p, double nx, double ny, double nz, double nm, double fs,
                                 double rl, double ul, double vl, double wl, double pl, double Hl, double al,
                                 double rr, double ur, double vr, double wr, double pr, double Hr, double ar,
                                 double vFracF, double *flx) {
  double ql = ul * nx + vl * ny + wl * nz - fs;
  double qr = ur * nx + vr * ny + wr * nz - fs;
  double af = 0.5 * (al + ar);
  double ml = ql / af, mla = fabs(ml);
  double mr = qr / af, mra = fabs(mr);
  double mbarsq = 0.5 * (ml * ml + mr * mr);
  double fa = 1.0;
  double alpha = 0.1875 * (5.0 * fa * fa - 4.0);
  double mlp, wtl, mrm, wtr;
  if (mla <= 1.0) {
    mlp = 0.25 * (ml + 1.0) * (ml + 1.0) + 0.125 * (ml * ml - 1.0) * (ml * ml - 1.0);
    wtl = 0.25 * (ml + 1.0) * (ml + 1.0) * (2.0 - ml) + alpha * ml * (ml * ml - 1.0) * (ml * ml - 1.0);
  } else {
    mlp = 0.5 * (ml + mla);
    wtl = 0.5 * (1.0 + ml / mla);
  }
  if (mra <= 1.0) {
    mrm = -0.25 * (mr - 1.0) * (mr - 1.0) - 0.125 * (mr * mr - 1.0) * (mr * mr - 1.0);
    wtr = 0.25 * (mr - 1.0) * (mr - 1.0) * (2.0 + mr) - alpha * mr * (mr * mr - 1.0) * (mr * mr - 1.0);
  } else {
    mrm = 0.5 * (mr - mra);
    wtr = 0.5 * (1.0 - mr / mra);
  }
  double rf = 0.5 * (rl + rr);
  double sigmabar = 1.0 - p->sigma * mbarsq;
  /* the pressure diffusion term M_p and the velocity diffusion term p_u */
  double mp = -(p->kp * fmax(sigmabar, 0.0) * (pr - pl)) / (fa * rf * af * af);
  double pu = -(p->ku * wtl * wtr * fa * af * (rl + rr) * (qr - ql));
  double mf = mlp + mrm + mp;
  double mfa = fabs(mf);
  double mfp = 0.5 * (mf + mfa), mfm = 0.5 * (mf - mfa);
  double pf = wtl * pl + wtr * pr + pu;
  double av = af * vFracF;
  flx[0] = (av * (mfp * rl + mfm * rr)) * nm;
  flx[1] = (av * (mfp * rl * ul + mfm * rr * ur) + pf * nx) * nm;
  flx[2] = (av * (mfp * rl * vl + mfm * rr * vr) + pf * ny) * nm;
  flx[3] = (av * (mfp * rl * wl + mfm * rr * wr) + pf * nz) * nm;
  flx[4] = (av * (mfp * rl * Hl + mfm * rr * Hr) + pf * fs) * nm;
}

/* Point fluxes for closed-form tests.  L/R = (rho,u,v,w,p); H, e and a are
 * derived with the flux path's perfect-gas relations (refCp, refGamma). */
void orc_flux_point(const orc_params *p, int discr, const double *n, const double *L, const double *R,
                    double *flx) {
  double cp = p->refCp, g = p->refGamma, gc = R_CpG(cp, g), fs = 0.0;
  double tl = T_DPR(L[0], L[4], gc), tr = T_DPR(R[0], R[4], gc);
  double Hl = Ho_CpTUVW(cp, tl, L[1], L[2], L[3], 0.0), Hr = Ho_CpTUVW(cp, tr, R[1], R[2], R[3], 0.0);
  double ql = L[1] * n[0] + L[2] * n[1] + L[3] * n[2] - fs, qr = R[1] * n[0] + R[2] * n[1] + R[3] * n[2] - fs;
  if (discr == ORC_DISCR_AUSM) {
    ausm_flux_function(n[0], n[1], n[2], n[3], fs, L[0], L[1], L[2], L[3], L[4], Hl, C_GRT(g, gc, tl),
                       R[0], R[1], R[2], R[3], R[4], Hr, C_GRT(g, gc, tr), flx);
  } else if (discr == ORC_DISCR_AUSMUP) {
    ausmup_flux_function(p, n[0], n[1], n[2], n[3], fs, L[0], L[1], L[2], L[3], L[4], Hl, C_GRT(g, gc, tl),
                         R[0], R[1], R[2], R[3], R[4], Hr, C_GRT(g, gc, tr), 1.0, flx);
  } else if (discr == ORC_DISCR_ROE) {
    roe_flux(p, n[0], n[1], n[2], n[3], fs, L[0], L[1], L[2], L[3], L[4], Hl, ql,
             R[0], R[1], R[2], R[3], R[4], Hr, qr, flx);
  } else {
    double el = Eo_DGPUVW(L[0], g, L[4], L[1], L[2], L[3], 0.0), er = Eo_DGPUVW(R[0], g, R[4], R[1], R[2], R[3], 0.0);
    hllc_flux(p, n[0], n[1], n[2], n[3], fs, L[0], L[1], L[2], L[3], L[4], el, C_DGP(L[0], g, L[4]), ql,
              el + L[4] / L[0], R[0], R[1], R[2], R[3], R[4], er, C_DGP(R[0], g, R[4]), qr, er + R[4] / R[0], flx);
  }
}

/* reconstruction at the face centroid, modflu/RFLU_ModAUSMFlux.F90:2735-2783 */
static void reconstruct(const orc_region *r, int c, double xc, double yc, double zc,
                        double *rr, double *uu, double *vv, double *ww, double *pp) {
  const double *cv = r->cv + 5 * (size_t)c;
  double ir = 1.0 / cv[0];
  double rho = cv[0], u = cv[1] * ir, v = cv[2] * ir, w = cv[3] * ir, p = r->dv[3 * (size_t)c];
  double dx = xc - r->cofg[3 * (size_t)c], dy = yc - r->cofg[3 * (size_t)c + 1], dz = zc - r->cofg[3 * (size_t)c + 2];
  const double *g = r->grad + 15 * (size_t)c;
  *rr = rho + g[0] * dx + g[1] * dy + g[2] * dz;
  *uu = u + g[3] * dx + g[4] * dy + g[5] * dz;
  *vv = v + g[6] * dx + g[7] * dy + g[8] * dz;
  *ww = w + g[9] * dx + g[10] * dy + g[11] * dz;
  *pp = p + g[12] * dx + g[13] * dy + g[14] * dz;
}

static void add_rhs(orc_region *r, int c1, int c2, const double *flx) {
  double *a = r->rhs + 5 * (size_t)c1, *b = r->rhs + 5 * (size_t)c2;
  for (int k = 0; k < 5; k++) a[k] = a[k] + flx[k];
  for (int k = 0; k < 5; k++) b[k] = b[k] - flx[k];
}

/* Interior faces: RFLU_{AUSM,ROE,HLLC}_ComputeFlux{1,2}_TCP face loops */
static void interior_fluxes(orc_region *r) {
  const orc_params *p = &r->p;
  const double fs = 0.0, ksg = 0.0;
  double cp = p->refCp, g = p->refGamma, gc = R_CpG(cp, g);
  /* the face fluxes in parallel, then added into the cells in face order */
  double *fbuf = malloc(sizeof(double) * 5 * (size_t)(r->nFaces > 0 ? r->nFaces : 1));
#pragma omp parallel for schedule(static)
  for (int f = 0; f < r->nFaces; f++) {
    int c1 = r->f2c[2 * f], c2 = r->f2c[2 * f + 1];
    double nx = r->fn[4 * f], ny = r->fn[4 * f + 1], nz = r->fn[4 * f + 2], nm = r->fn[4 * f + 3];
    const double *q1 = r->cv + 5 * (size_t)c1, *q2 = r->cv + 5 * (size_t)c2;
    const double *d1 = r->dv + 3 * (size_t)c1, *d2 = r->dv + 3 * (size_t)c2;
    double flx[5];
    if (p->spaceOrder == 1) {
      double rl = q1[0], irl = 1.0 / rl, ul = q1[1] * irl, vl = q1[2] * irl, wl = q1[3] * irl;
      double pl = d1[0], tl = d1[1], al = d1[2];
      double rr = q2[0], irr = 1.0 / rr, ur = q2[1] * irr, vr = q2[2] * irr, wr = q2[3] * irr;
      double pr = d2[0], tr = d2[1], ar = d2[2];
      if (p->spaceDiscr == ORC_DISCR_AUSM) {
        /* RFLU_ModAUSMFlux.F90:673-744 */
        double Hl = q1[4] * irl + pl / rl;
        double Hr = q2[4] * irr + pr / rr;
        ausm_flux_function(nx, ny, nz, nm, fs, rl, ul, vl, wl, pl, Hl, al, rr, ur, vr, wr, pr, Hr, ar, flx);
      } else if (p->spaceDiscr == ORC_DISCR_AUSMUP) {
        /* RFLU_AUSMPlusUp_ComputeFlux1 (RFLU_ModAUSMPlusUpFlux.F90:894-1366) */
        double Hl = q1[4] * irl + pl / rl;
        double Hr = q2[4] * irr + pr / rr;
        ausmup_flux_function(p, nx, ny, nz, nm, fs, rl, ul, vl, wl, pl, Hl, al, rr, ur, vr, wr, pr, Hr, ar, 1.0,
                             flx);
      } else if (p->spaceDiscr == ORC_DISCR_ROE) {
        /* RFLU_ModRoeFlux.F90:1723 (first order TCP) */
        double Hl = Ho_CpTUVW(cp, tl, ul, vl, wl, ksg);
        double ql = ul * nx + vl * ny + wl * nz - fs;
        double Hr = Ho_CpTUVW(cp, tr, ur, vr, wr, ksg);
        double qr = ur * nx + vr * ny + wr * nz - fs;
        roe_flux(p, nx, ny, nz, nm, fs, rl, ul, vl, wl, pl, Hl, ql, rr, ur, vr, wr, pr, Hr, qr, flx);
      } else {
        /* RFLU_ModHLLCFlux.F90:937 (first order TCP) */
        double el = q1[4] * irl, er = q2[4] * irr;
        double ql = ul * nx + vl * ny + wl * nz - fs, qr = ur * nx + vr * ny + wr * nz - fs;
        double Hl = el + irl * pl, Hr = er + irr * pr;
        hllc_flux(p, nx, ny, nz, nm, fs, rl, ul, vl, wl, pl, el, al, ql, Hl, rr, ur, vr, wr, pr, er, ar, qr, Hr, flx);
      }
    } else {
      double xc = r->fc[3 * f], yc = r->fc[3 * f + 1], zc = r->fc[3 * f + 2];
      double rl, ul, vl, wl, pl, rr, ur, vr, wr, pr;
      reconstruct(r, c1, xc, yc, zc, &rl, &ul, &vl, &wl, &pl);
      reconstruct(r, c2, xc, yc, zc, &rr, &ur, &vr, &wr, &pr);
      if (p->spaceDiscr == ORC_DISCR_AUSMUP) {
        /* RFLU_AUSMPlusUp_ComputeFlux2_TCP (RFLU_ModAUSMPlusUpFlux.F90:3255-3632):
         * with particles the reconstructed density is divided by the
         * cell's gas fraction before T, a and H, and the face's gas fraction
         * (the mean of the two cells') scales the convected flux */
        double vFracF = 1.0;
        if (r->phi) {
          rl = rl / (1.0 - r->phi[c1]);
          rr = rr / (1.0 - r->phi[c2]);
          vFracF = 0.5 * ((1.0 - r->phi[c1]) + (1.0 - r->phi[c2]));
        }
        double tl = T_DPR(rl, pl, gc), al = C_GRT(g, gc, tl), Hl = Ho_CpTUVW(cp, tl, ul, vl, wl, ksg);
        double tr = T_DPR(rr, pr, gc), ar = C_GRT(g, gc, tr), Hr = Ho_CpTUVW(cp, tr, ur, vr, wr, ksg);
        ausmup_flux_function(p, nx, ny, nz, nm, fs, rl, ul, vl, wl, pl, Hl, al, rr, ur, vr, wr, pr, Hr, ar, vFracF,
                             flx);
      } else if (p->spaceDiscr == ORC_DISCR_AUSM) {
        /* RFLU_ModAUSMFlux.F90:2784-2843 */
        double tl = T_DPR(rl, pl, gc), al = C_GRT(g, gc, tl), Hl = Ho_CpTUVW(cp, tl, ul, vl, wl, ksg);
        double tr = T_DPR(rr, pr, gc), ar = C_GRT(g, gc, tr), Hr = Ho_CpTUVW(cp, tr, ur, vr, wr, ksg);
        ausm_flux_function(nx, ny, nz, nm, fs, rl, ul, vl, wl, pl, Hl, al, rr, ur, vr, wr, pr, Hr, ar, flx);
      } else if (p->spaceDiscr == ORC_DISCR_ROE) {
        /* RFLU_ModRoeFlux.F90:2700-2781 */
        double tl = T_DPR(rl, pl, gc), Hl = Ho_CpTUVW(cp, tl, ul, vl, wl, ksg);
        double ql = ul * nx + vl * ny + wl * nz - fs;
        double tr = T_DPR(rr, pr, gc), Hr = Ho_CpTUVW(cp, tr, ur, vr, wr, ksg);
        double qr = ur * nx + vr * ny + wr * nz - fs;
        roe_flux(p, nx, ny, nz, nm, fs, rl, ul, vl, wl, pl, Hl, ql, rr, ur, vr, wr, pr, Hr, qr, flx);
      } else {
        /* RFLU_ModHLLCFlux.F90:2255-2337: H uses 1/cv(1), not the
         * reconstructed density */
        double irl = 1.0 / q1[0], irr = 1.0 / q2[0];
        double el = Eo_DGPUVW(rl, g, pl, ul, vl, wl, ksg), al = C_DGP(rl, g, pl);
        double ql = ul * nx + vl * ny + wl * nz - fs, Hl = el + irl * pl;
        double er = Eo_DGPUVW(rr, g, pr, ur, vr, wr, ksg), ar = C_DGP(rr, g, pr);
        double qr = ur * nx + vr * ny + wr * nz - fs, Hr = er + irr * pr;
        hllc_flux(p, nx, ny, nz, nm, fs, rl, ul, vl, wl, pl, el, al, ql, Hl, rr, ur, vr, wr, pr, er, ar, qr, Hr, flx);
      }
    }
    memcpy(fbuf + 5 * (size_t)f, flx, sizeof flx);
  }
  for (int f = 0; f < r->nFaces; f++) add_rhs(r, r->f2c[2 * f], r->f2c[2 * f + 1], fbuf + 5 * (size_t)f);
  free(fbuf);
}

/* libfloflu/BcondOutflowPerf.F90:88-160 */
static void bcond_outflow_perf(int bcOpt, double pout, double sxn, double syn, double szn,
                               double cpgas, double mol, double rho, double rhou, double rhov,
                               double rhow, double rhoe, double press, double *b) {
  const double ksg = 0.0;
  double rgas = R_M(mol);
  double gamma = G_CpR(cpgas, rgas);
  double u = rhou / rho, v = rhov / rho, w = rhow / rho;
  double csound = C_DGP(rho, gamma, press);
  double mach = sqrt(u * u + v * v + w * w) / csound;
  if (mach < 1.0 && (bcOpt == ORC_BCOPT_SUBSONIC || bcOpt == ORC_BCOPT_MIXED)) {
    double rrhoc = 1.0 / (rho * csound);
    double deltp = press - pout;
    double rhob = rho - deltp / (csound * csound);
    double ub = u + sxn * deltp * rrhoc;
    double vb = v + syn * deltp * rrhoc;
    double wb = w + szn * deltp * rrhoc;
    double vnd = ub * sxn + vb * syn + wb * szn;
    if (vnd < 0.0) {
      ub = copysign(1.0, u) * fmax(fabs(ub), fabs(u));
      vb = copysign(1.0, v) * fmax(fabs(vb), fabs(v));
      wb = copysign(1.0, w) * fmax(fabs(wb), fabs(w));
    }
    b[0] = rhob;
    b[1] = rhob * ub;
    b[2] = rhob * vb;
    b[3] = rhob * wb;
    b[4] = rhob * Eo_DGPUVW(rhob, gamma, pout, ub, vb, wb, ksg);
  } else {
    b[0] = rho; b[1] = rhou; b[2] = rhov; b[3] = rhow; b[4] = rhoe;
  }
}

/* libfloflu/MixtPerf_*.F90 relations the patch rind states use */
static inline double D_PRT(double P, double R, double T) { return P / (R * T); }
static inline double P_DRT(double D, double R, double T) { return D * R * T; }
static inline double Eo_DGPVm(double D, double G, double P, double Vm, double K) {
  return P / (D * (G - 1.0)) + 0.5 * Vm * Vm + K;
}
static inline double C2_GRT(double G, double R, double T) { return G * R * T; }
static inline double C_Co2GUVW(double Co2, double G, double U, double V, double W) {
  return sqrt(Co2 - 0.5 * (G - 1.0) * (U * U + V * V + W * W));
}
static inline double T_CGR(double C, double G, double R) { return C * C / (G * R); }
static inline double P_GPoTTo(double G, double Po, double T, double To) { return Po * pow(T / To, G / (G - 1.0)); }
static inline double Vm_C2Co2G(double C2, double Co2, double G) {
  return Co2 > C2 ? sqrt(2.0 / (G - 1.0) * (Co2 - C2)) : 0.0;
}
static inline double P_GMaPo(double G, double Ma, double Po) {
  return Po / pow(1.0 + 0.5 * (G - 1.0) * Ma * Ma, G / (G - 1.0));
}
static inline double T_GMaTo(double G, double Ma, double To) { return To / (1.0 + 0.5 * (G - 1.0) * Ma * Ma); }

/* RFLU_SetRindStateFarfieldPerf without the lift correction
 * (modflu/RFLU_ModRindStates.F90:249-421); pl is recomputed from rel. */
static void rind_farfield(double cp, double mm, double nx, double ny, double nz, double machInf, double pInf,
                          double tInf, double alphaInf, double betaInf, double rl, double rul, double rvl, double rwl,
                          double rel, double *o) {
  const double ksg = 0.0;
  double rGas = R_M(mm), gGas = G_CpR(cp, rGas);
  double ul = rul / rl, vl = rvl / rl, wl = rwl / rl;
  double ql = ul * nx + vl * ny + wl * nz;
  double el = rel / rl, sl2 = ul * ul + vl * vl + wl * wl;
  double pl = P_DEoGVm2(rl, el, gGas, sl2, ksg);
  double ri = D_PRT(pInf, rGas, tInf), pi = pInf;
  double qi = machInf * C_GRT(gGas, rGas, tInf);
  double ui = qi * cos(alphaInf) * cos(betaInf);
  double vi = qi * sin(alphaInf) * cos(betaInf);
  double wi = qi * sin(betaInf);
  if (machInf < 1.0) {
    double al = C_DGP(rl, gGas, pl), pb, rb, ub, vb, wb;
    if (ql < 0.0) {
      pb = 0.5 * (pi + pl - rl * al * ((ui - ul) * nx + (vi - vl) * ny + (wi - wl) * nz));
      rb = ri - (pi - pb) / (al * al);
      ub = ui - nx * (pi - pb) / (rl * al);
      vb = vi - ny * (pi - pb) / (rl * al);
      wb = wi - nz * (pi - pb) / (rl * al);
    } else {
      pb = pi;
      rb = rl - (pl - pi) / (al * al);
      ub = ul + nx * (pl - pi) / (rl * al);
      vb = vl + ny * (pl - pi) / (rl * al);
      wb = wl + nz * (pl - pi) / (rl * al);
    }
    o[0] = rb; o[1] = rb * ub; o[2] = rb * vb; o[3] = rb * wb;
    o[4] = rb * Eo_DGPUVW(rb, gGas, pb, ub, vb, wb, ksg);
    o[5] = pb;
  } else if (ql < 0.0) {
    o[0] = ri; o[1] = ri * ui; o[2] = ri * vi; o[3] = ri * wi;
    o[4] = ri * Eo_DGPVm(ri, gGas, pi, qi, ksg);
    o[5] = pi;
  } else {
    o[0] = rl; o[1] = rul; o[2] = rvl; o[3] = rwl; o[4] = rel; o[5] = pl;
  }
}

/* RFLU_SetRindStateSlipWallPerf (modflu/RFLU_ModRindStates.F90:539-600):
 * the wall pressure from the one-dimensional wave towards the wall */
static double rind_slipwall_p(double cp, double mm, double nx, double ny, double nz, double rl, double rul,
                              double rvl, double rwl, double fs, double pl) {
  double rGas = R_M(mm), gGas = G_CpR(cp, rGas);
  double irl = 1.0 / rl;
  double ul = irl * rul, vl = irl * rvl, wl = irl * rwl;
  double ql = ul * nx + vl * ny + wl * nz - fs;
  double al = C_DGP(rl, gGas, pl);
  if (ql < 0.0) {
    double term = 1.0 + 0.5 * (gGas - 1.0) * ql / al;
    return pl * pow(term, 2.0 * gGas / (gGas - 1.0));
  }
  double term = (gGas + 1.0) / 4.0;
  return pl + term * rl * ql * (ql + sqrt(al * al + term * term * ql * ql) / term);
}

/* libfloflu/BcondInflowPerf.F90: total pressure, total temperature and flow
 * angles.  Where the reference leaves the state undefined (a subsonic-only
 * patch seeing |q| >= c) the stagnation state at rest is used. */
static void bcond_inflow_perf(int opt, int fixed, double ptot, double ttot, double betah, double betav,
                              double mach, double sxn, double syn, double szn, double cpgas, double mm, double rl,
                              double rul, double rvl, double rwl, double *o) {
  const double ksg = 0.0;
  double rgas = R_M(mm), g = G_CpR(cpgas, rgas);
  double rr, ur, vr, wr, pr, rer;
  int supersonicState = 0;
  if (opt == ORC_BCOPT_SUBSONIC || opt == ORC_BCOPT_MIXED) {
    double gm1 = g - 1.0;
    double ul = rul / rl, vl = rvl / rl, wl = rwl / rl;
    double a02 = C2_GRT(g, rgas, ttot);
    double al = C_Co2GUVW(a02, g, ul, vl, wl);
    double ql = ul * sxn + vl * syn + wl * szn;
    if (fabs(ql) < al) {
      double sl = sqrt(ul * ul + vl * vl + wl * wl), eta;
      if (!fixed) eta = sl > 1.0E-6 ? ql / sl : -1.0;
      else eta = -1.0;
      double Rm = al - 0.5 * gm1 * ql;
      double disc = 0.5 * gm1 * (eta * eta) * (a02 / (Rm * Rm) * (1.0 + 0.5 * gm1 * (eta * eta)) - 1.0);
      double ar, tr, sr;
      if (disc < 0.0) {
        ar = sqrt(a02);
        tr = ttot;
        pr = ptot;
        sr = 0.0;
      } else {
        ar = Rm / (1.0 + 0.5 * gm1 * eta * eta) * (1.0 + sqrt(disc));
        tr = T_CGR(ar, g, rgas);
        pr = P_GPoTTo(g, ptot, tr, ttot);
        sr = Vm_C2Co2G(ar * ar, a02, g);
      }
      (void)ar;
      rr = D_PRT(pr, rgas, tr);
      ur = sr * cos(betah) * cos(betav);
      vr = sr * sin(betah);
      wr = sr * cos(betah) * sin(betav);
      rer = rr * Eo_DGPVm(rr, g, pr, sr, ksg);
    } else if (opt == ORC_BCOPT_MIXED) {
      supersonicState = 1;
    } else {  /* undefined in the reference: stagnation state */
      pr = ptot;
      rr = D_PRT(pr, rgas, ttot);
      ur = vr = wr = 0.0;
      rer = rr * Eo_DGPUVW(rr, g, pr, ur, vr, wr, ksg);
    }
  } else {
    supersonicState = 1;
  }
  if (supersonicState) {
    pr = P_GMaPo(g, mach, ptot);
    double tr = T_GMaTo(g, mach, ttot);
    rr = D_PRT(pr, rgas, tr);
    double ar = C_GRT(g, rgas, tr);
    ur = mach * ar * cos(betah) * cos(betav);
    vr = mach * ar * sin(betah);
    wr = mach * ar * cos(betah) * sin(betav);
    rer = rr * Eo_DGPUVW(rr, g, pr, ur, vr, wr, ksg);
  }
  o[0] = rr; o[1] = rr * ur; o[2] = rr * vr; o[3] = rr * wr; o[4] = rer; o[5] = pr;
}

/* the central average of the interior and rind fluxes the inflow, outflow
 * and farfield patches share (RFLU_CentralFirstPatch.F90:700-712) */
static void central_patch_flux(double nx, double ny, double nz, double nm, double fs, double rl, double rul,
                               double rvl, double rwl, double rel, double pl, const double *b, double *flx) {
  double rr = b[0], rur = b[1], rvr = b[2], rwr = b[3], rer = b[4], pr = b[5];
  double ql = (rul * nx + rvl * ny + rwl * nz) / rl - fs;
  double qr = (rur * nx + rvr * ny + rwr * nz) / rr - fs;
  flx[0] = 0.5 * (ql * rl + qr * rr) * nm;
  flx[1] = 0.5 * (ql * rul + pl * nx + qr * rur + pr * nx) * nm;
  flx[2] = 0.5 * (ql * rvl + pl * ny + qr * rvr + pr * ny) * nm;
  flx[3] = 0.5 * (ql * rwl + pl * nz + qr * rwr + pr * nz) * nm;
  flx[4] = 0.5 * (ql * (rel + pl) + pl * fs + qr * (rer + pr) + pr * fs) * nm;
}

/* Patches: rocflu/RFLU_CentralFirstPatch.F90:230-1134 and
 * rocflu/RFLU_CentralSecondPatch.F90:241-1351 (inflow, outflow, slip and
 * no-slip walls, farfield, injection) */
static void patch_fluxes(orc_region *r) {
  const double fs = 0.0, ksg = 0.0;
  double cp = r->cpGv, mm = r->molGv, rgas = R_M(mm), g = G_CpR(cp, rgas);
  for (int ip = 0; ip < r->nPatches; ip++) {
    int type = r->pBcType[ip], opt = r->pBcOpt[ip];
    const double *pv = r->pVals + ORC_BCVALS * (size_t)ip;
    for (int b = r->pOff[ip]; b < r->pOff[ip + 1]; b++) {
      int c1 = r->bf2c[b];
      double nx = r->bfn[4 * b], ny = r->bfn[4 * b + 1], nz = r->bfn[4 * b + 2], nm = r->bfn[4 * b + 3];
      const double *cv = r->cv + 5 * (size_t)c1;
      double *rh = r->rhs + 5 * (size_t)c1;
      double rl, rul, rvl, rwl, rel, pl, ul, vl, wl;
      if (r->p.patchOrder == 1) {
        rl = cv[0]; rul = cv[1]; rvl = cv[2]; rwl = cv[3]; rel = cv[4];
        pl = r->dv[3 * (size_t)c1];
        ul = rul / rl; vl = rvl / rl; wl = rwl / rl;
      } else {
        reconstruct(r, c1, r->bfc[3 * b], r->bfc[3 * b + 1], r->bfc[3 * b + 2], &rl, &ul, &vl, &wl, &pl);
        rul = rl * ul; rvl = rl * vl; rwl = rl * wl;
        if (type == ORC_BC_OUTFLOW || type == ORC_BC_FARFIELD || type == ORC_BC_INFLOW_VELTEMP ||
            type == ORC_BC_INFLOW_TOTANG)
          rel = rl * Eo_DGPUVW(rl, g, pl, ul, vl, wl, ksg);
        else
          rel = 0.0;
      }
      double flx[5], bb[6];
      switch (type) {
        case ORC_BC_SLIPWALL:
        case ORC_BC_NOSLIPWALL:  /* Euler: the pressure force only, no mass term */
          rh[1] = rh[1] + pl * nx * nm;
          rh[2] = rh[2] + pl * ny * nm;
          rh[3] = rh[3] + pl * nz * nm;
          rh[4] = rh[4] + pl * fs * nm;
          break;
        case ORC_BC_OUTFLOW: {
          double pr = (opt != ORC_BCOPT_SUPERSONIC) ? r->pPout[ip] : pl;
          bcond_outflow_perf(opt, pr, nx, ny, nz, cp, mm, rl, rul, rvl, rwl, rel, pl, bb);
          bb[5] = pr;
          central_patch_flux(nx, ny, nz, nm, fs, rl, rul, rvl, rwl, rel, pl, bb, flx);
          for (int k = 0; k < 5; k++) rh[k] = rh[k] + flx[k];
          break;
        }
        case ORC_BC_FARFIELD:
          rind_farfield(cp, mm, nx, ny, nz, pv[0], pv[3], pv[4], pv[1], pv[2], rl, rul, rvl, rwl, rel, bb);
          central_patch_flux(nx, ny, nz, nm, fs, rl, rul, rvl, rwl, rel, pl, bb, flx);
          for (int k = 0; k < 5; k++) rh[k] = rh[k] + flx[k];
          break;
        case ORC_BC_INFLOW_VELTEMP: {
          double ur = pv[0], vr = pv[1], wr = pv[2], tr = pv[3], rr, pr;
          if (opt != ORC_BCOPT_SUBSONIC) {
            pr = pv[4];
            rr = D_PRT(pr, rgas, tr);
          } else {
            rr = rl;
            pr = P_DRT(rr, rgas, tr);
          }
          bb[0] = rr; bb[4] = rr * Eo_DGPUVW(rr, g, pr, ur, vr, wr, ksg);
          bb[1] = rr * ur; bb[2] = rr * vr; bb[3] = rr * wr; bb[5] = pr;
          central_patch_flux(nx, ny, nz, nm, fs, rl, rul, rvl, rwl, rel, pl, bb, flx);
          for (int k = 0; k < 5; k++) rh[k] = rh[k] + flx[k];
          break;
        }
        case ORC_BC_INFLOW_TOTANG:
          bcond_inflow_perf(opt, pv[5] != 0.0, pv[0], pv[1], pv[2], pv[3], opt != ORC_BCOPT_SUBSONIC ? pv[4] : 0.0,
                            nx, ny, nz, cp, mm, rl, rul, rvl, rwl, bb);
          central_patch_flux(nx, ny, nz, nm, fs, rl, rul, rvl, rwl, rel, pl, bb, flx);
          for (int k = 0; k < 5; k++) rh[k] = rh[k] + flx[k];
          break;
        case ORC_BC_INJECTION: {
          double minj = pv[0];
          if (minj > 0.0) {  /* RFLU_SetRindStateInjectPerf (RFLU_ModRindStates.F90:460-505) */
            double tinj = pv[1];
            double ri = D_PRT(pl, rgas, tinj);
            double qi = -minj / ri + fs;
            double ui = qi * nx, vi = qi * ny, wi = qi * nz;
            double Hi = Ho_CpTUVW(cp, tinj, ui, vi, wi, 0.0);
            flx[0] = -minj * nm;
            flx[1] = (-minj * ui + pl * nx) * nm;
            flx[2] = (-minj * vi + pl * ny) * nm;
            flx[3] = (-minj * wi + pl * nz) * nm;
            flx[4] = (-minj * Hi + pl * fs) * nm;
          } else {
            double pw = rind_slipwall_p(cp, mm, nx, ny, nz, rl, rul, rvl, rwl, fs, pl);
            flx[0] = 0.0;
            flx[1] = pw * nx * nm;
            flx[2] = pw * ny * nm;
            flx[3] = pw * nz * nm;
            flx[4] = pw * fs * nm;
          }
          for (int k = 0; k < 5; k++) rh[k] = rh[k] + flx[k];
          break;
        }
      }
      (void)ul; (void)vl; (void)wl;
    }
  }
}

/* ConvectiveFluxes: libfloflu/ConvectiveFluxes.F90:133-141 + RFLU_ComputeFluxInv */
static void convective_fluxes(orc_region *r) {
  for (size_t q = 0; q < 5 * (size_t)r->nCells; q++) r->rhs[q] = -r->diss[q];
  interior_fluxes(r);
  patch_fluxes(r);
}

/* SourceTermsMP: libfloflu/SourceTermsMP.F90:73-218 (gravity, round trip) */
static void source_terms(orc_region *r) {
  if (r->p.accelOn) {
    /* libfloflu/SourceTerms.F90:127-136 */
    const double *a = r->p.accel;
    for (int i = 0; i < r->nCells; i++) {
      const double *cv = r->cv + 5 * (size_t)i;
      double *rh = r->rhs + 5 * (size_t)i;
      double rhoVol = r->vol[i] * cv[0];
      rh[1] = rh[1] - a[0] * rhoVol;
      rh[2] = rh[2] - a[1] * rhoVol;
      rh[3] = rh[3] - a[2] * rhoVol;
      rh[4] = rh[4] - r->vol[i] * (a[0] * cv[1] + a[1] * cv[2] + a[2] * cv[3]);
    }
  }
  cons2prim(r);
  prim2cons(r);
}

void orc_residual(orc_region *r) {
  cell_gradients(r);
  convective_fluxes(r);
  source_terms(r);
}

/* RFLU_MinimumTimeStep: rocflu/RFLU_MinimumTimeStep.F90:125-241 plus the
 * clip to the run length of RFLU_TimeStepping.F90:489-493 */
static double minimum_timestep(orc_region *r, double dtLocal) {
  double cfl = r->p.cfl < 2.0 ? r->p.cfl : 2.0; /* RFLU_DerivedInputValues.F90:395 */
  double wI = 0.0, wC = 0.0;
  double d = cfl * dtLocal;
  if (r->currentTime <= 1.0e-8) wI = 1.0; else wC = 1.0;
  d = wI / r->p.dtImposed + wC / d;
  d = 1.0 / d;
  if (d <= 0.0) d = 6.0e-12;
  d = fmax(d, r->p.dtMinLimit);
  if (r->p.dTimeSystem > 0.0 && r->timeSinceRestart + d > r->p.dTimeSystem)
    d = r->p.dTimeSystem - r->timeSinceRestart;
  return d;
}

/* RkUpdateGeneric Wray branch: libfloflu/RkUpdateGeneric.F90:285-314;
 * coefficients libflu/RFLU_DerivedInputValues.F90:374-386 */
static void rk_update(orc_region *r, int istage) {
  const double ark[3] = {8.0 / 15.0, 5.0 / 12.0, 0.75};
  const double grk[3] = {0.0, 17.0 / 25.0, 5.0 / 9.0};
  double fac = ark[istage - 1] * r->dtMin;
  for (int i = 0; i < r->nCells; i++) {
    double adtv = fac / r->vol[i];
    double *cv = r->cv + 5 * (size_t)i, *co = r->cvOld + 5 * (size_t)i;
    double *rh = r->rhs + 5 * (size_t)i, *rs = r->rhsSum + 5 * (size_t)i;
    for (int k = 0; k < 5; k++) {
      if (istage == 1) {
        cv[k] = co[k] - adtv * rh[k];
        rs[k] = rh[k];
      } else if (istage == 2) {
        cv[k] = co[k] - adtv * (rh[k] - grk[1] * rs[k]);
        rs[k] = rh[k];
      } else {
        cv[k] = co[k] - adtv * (rh[k] - grk[2] * rs[k]);
      }
    }
  }
}

/* AfterUpdateMP checks: libflu/RFLU_CheckValidity.F90, RFLU_EnforceBounds.F90,
 * RFLU_CheckPositivity.F90 (TCPERF branches) */
static int after_update(orc_region *r) {
  const double ksg = 0.0;
  double rgas = R_M(r->molGv), gamma = G_CpR(r->cpGv, rgas);
  for (int i = 0; i < r->nCells; i++) {
    double *cv = r->cv + 5 * (size_t)i, *dv = r->dv + 3 * (size_t)i;
    double rho = cv[0], rrho = 1.0 / rho;
    double u = rrho * cv[1], v = rrho * cv[2], w = rrho * cv[3];
    if (isnan(rho) || isnan(u) || isnan(v) || isnan(w) || isnan(dv[0]) || isnan(dv[1])) {
      snprintf(g_err, sizeof g_err, "invalid variables in cell %d", i);
      return -1;
    }
  }
  for (int i = 0; i < r->nCells; i++) {
    double *cv = r->cv + 5 * (size_t)i;
    double rr = cv[0], ir = 1.0 / rr;
    double u = ir * cv[1], v = ir * cv[2], w = ir * cv[3], Eo = ir * cv[4];
    double Vm2 = u * u + v * v + w * w;
    double p = P_DEoGVm2(rr, Eo, gamma, Vm2, ksg);
    if (p <= 0.0 && fabs(p) <= 1.0e-12) {
      p = 1.0e-12;
      cv[4] = rr * Eo_DGPUVW(rr, gamma, p, u, v, w, ksg);
    }
  }
  for (int i = 0; i < r->nCells; i++) {
    double *cv = r->cv + 5 * (size_t)i, *dv = r->dv + 3 * (size_t)i;
    double rho = cv[0], rrho = 1.0 / rho;
    double u = rrho * cv[1], v = rrho * cv[2], w = rrho * cv[3];
    double Vm2 = u * u + v * v + w * w;
    double p = dv[0];
    if (p <= 0.0) {
      double Eo = 1.0e4 + 0.5 * Vm2 + ksg;
      p = P_DEoGVm2(rho, Eo, gamma, Vm2, ksg);
      dv[0] = p;
      dv[1] = T_DPR(rho, p, rgas);
      cv[4] = rho * Eo;
    }
    if (rho <= 0.0 || p <= 0.0) {
      snprintf(g_err, sizeof g_err, "negative positive-definite variables in cell %d", i);
      return -2;
    }
  }
  return 0;
}

/* One RungeKuttaMP stage with the current dtMin (diagnostics). */
int orc_rk_stage(orc_region *r, int istage) {
  memcpy(r->cvOld, r->cv, sizeof(double) * 5 * r->nCells);
  memset(r->diss, 0, sizeof(double) * 5 * r->nCells);
  orc_residual(r);
  if (r->picl) {
    int e = orc__picl_runge(r, istage);
    if (e) { snprintf(g_err, sizeof g_err, "ppiclF: Re rotational too large"); return e; }
  }
  rk_update(r, istage);
  set_dependent_vars(r);
  return after_update(r);
}
void orc_set_dtmin(orc_region *r, double dt) { r->dtMin = dt; }
void orc__set_dependent_vars(orc_region *r) { set_dependent_vars(r); }

int orc_step(orc_region *r, int nSteps) {
  for (int it = 0; it < nSteps; it++) {
    double dtl = orc_timestep_inviscid(r, NULL);
    r->dtMin = minimum_timestep(r, dtl);
    for (int istage = 1; istage <= 3; istage++) {
      memcpy(r->cvOld, r->cv, sizeof(double) * 5 * r->nCells);           /* RkInitGeneric */
      memset(r->diss, 0, sizeof(double) * 5 * r->nCells);
      orc_residual(r);
      if (r->picl) {                                                       /* RungeKuttaMP.F90:243-247 */
        int e = orc__picl_runge(r, istage);
        if (e) { snprintf(g_err, sizeof g_err, "ppiclF: Re rotational too large"); return e; }
      }
      rk_update(r, istage);
      set_dependent_vars(r);
      int e = after_update(r);
      if (e) return e;
    }
    r->currentTime = r->currentTime + r->dtMin;
    r->timeSinceRestart = r->timeSinceRestart + r->dtMin;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
orc_region *orc_create(int nVert, const double *xyz, int nvc, int nCells, const int *c2v,
                       int nPatches, const int *pBcType, const int *pBcOpt, const double *pPout,
                       const int *pNBF, int nvf, const int *bf2v, const orc_params *prm) {
  orc_region *r = calloc(1, sizeof *r);
  g_err[0] = 0;
  if (nvc != 4 && nvc != 8) FAIL("nvc must be 4 (tets) or 8 (hexes, or mixed cells padded with -1)");
  r->p = *prm;
  r->nVert = nVert; r->nvc = nvc; r->nCells = nCells; r->nPatches = nPatches; r->nvf = nvf;
  r->xyz = malloc(sizeof(double) * 3 * nVert);
  memcpy(r->xyz, xyz, sizeof(double) * 3 * nVert);
  r->c2v = malloc(sizeof(int) * (size_t)nvc * nCells);
  memcpy(r->c2v, c2v, sizeof(int) * (size_t)nvc * nCells);
  r->pBcType = malloc(sizeof(int) * nPatches); r->pBcOpt = malloc(sizeof(int) * nPatches);
  r->pPout = malloc(sizeof(double) * nPatches); r->pNBF = malloc(sizeof(int) * nPatches);
  r->pVals = calloc((size_t)ORC_BCVALS * (nPatches > 0 ? nPatches : 1), sizeof(double));
  r->pOff = malloc(sizeof(int) * (nPatches + 1));
  r->pOff[0] = 0;
  for (int p = 0; p < nPatches; p++) {
    r->pBcType[p] = pBcType[p]; r->pBcOpt[p] = pBcOpt[p]; r->pPout[p] = pPout[p]; r->pNBF[p] = pNBF[p];
    r->pOff[p + 1] = r->pOff[p] + pNBF[p];
  }
  r->nBFaces = r->pOff[nPatches];
  r->bf2v = malloc(sizeof(int) * (size_t)nvf * (r->nBFaces + 1));
  memcpy(r->bf2v, bf2v, sizeof(int) * (size_t)nvf * r->nBFaces);
  if (!build_faces(r)) goto fail;
  size_t F = r->nFaces > 0 ? r->nFaces : 1, B = r->nBFaces > 0 ? r->nBFaces : 1, N = nCells;
  r->fn = malloc(sizeof(double) * 4 * F); r->fc = malloc(sizeof(double) * 3 * F);
  r->bfn = malloc(sizeof(double) * 4 * B); r->bfc = malloc(sizeof(double) * 3 * B);
  r->vol = malloc(sizeof(double) * N); r->cofg = malloc(sizeof(double) * 3 * N);
  build_geometry(r);
  if (prm->spaceOrder > 1 && !build_stencils(r)) goto fail;
  r->cv = calloc(5 * N, sizeof(double)); r->cvOld = calloc(5 * N, sizeof(double));
  r->dv = calloc(3 * N, sizeof(double)); r->rhs = calloc(5 * N, sizeof(double));
  r->rhsSum = calloc(5 * N, sizeof(double)); r->diss = calloc(5 * N, sizeof(double));
  r->grad = calloc(15 * N, sizeof(double)); r->lim = calloc(5 * N, sizeof(double));
  r->dt = calloc(N, sizeof(double));
  set_gas_vars(r);
  return r;
fail:
  orc_destroy(r);
  return NULL;
}

void orc_destroy(orc_region *r) {
  if (!r) return;
  void *ptrs[] = {r->xyz, r->c2v, r->f2c, r->f2v, r->fn, r->fc, r->pBcType, r->pBcOpt, r->pNBF, r->pOff,
                  r->pPout, r->bf2v, r->bf2c, r->bfn, r->bfc, r->vol, r->cofg, r->csOff, r->csMem, r->moms,
                  r->cv, r->cvOld, r->dv, r->rhs, r->rhsSum, r->diss, r->grad, r->lim, r->dt,
                  r->phi, r->gradRhog, r->pVals};
  for (size_t q = 0; q < sizeof ptrs / sizeof ptrs[0]; q++) free(ptrs[q]);
  orc__picl_free(r->picl);
  free(r);
}

void orc_set_patch_vals(orc_region *r, int ip, const double *vals) {
  if (ip >= 0 && ip < r->nPatches) memcpy(r->pVals + ORC_BCVALS * (size_t)ip, vals, sizeof(double) * ORC_BCVALS);
}

int orc_nfaces(const orc_region *r) { return r->nFaces; }
int orc_nbfaces(const orc_region *r) { return r->nBFaces; }
void orc_get_faces(const orc_region *r, int *f2c, int *f2v, double *fn, double *fc) {
  if (f2c) memcpy(f2c, r->f2c, sizeof(int) * 2 * r->nFaces);
  if (f2v) memcpy(f2v, r->f2v, sizeof(int) * 4 * r->nFaces);
  if (fn) memcpy(fn, r->fn, sizeof(double) * 4 * r->nFaces);
  if (fc) memcpy(fc, r->fc, sizeof(double) * 3 * r->nFaces);
}
void orc_get_bfaces(const orc_region *r, int *bf2c, double *bfn, double *bfc) {
  if (bf2c) memcpy(bf2c, r->bf2c, sizeof(int) * r->nBFaces);
  if (bfn) memcpy(bfn, r->bfn, sizeof(double) * 4 * r->nBFaces);
  if (bfc) memcpy(bfc, r->bfc, sizeof(double) * 3 * r->nBFaces);
}
void orc_get_cells(const orc_region *r, double *vol, double *cofg) {
  if (vol) memcpy(vol, r->vol, sizeof(double) * r->nCells);
  if (cofg) memcpy(cofg, r->cofg, sizeof(double) * 3 * r->nCells);
}
int orc_stencil_total(const orc_region *r) { return r->csOff ? r->csOff[r->nCells] : 0; }
void orc_get_stencils(const orc_region *r, int *off, int *memb, double *moms) {
  if (!r->csOff) return;
  if (off) memcpy(off, r->csOff, sizeof(int) * (r->nCells + 1));
  if (memb) memcpy(memb, r->csMem, sizeof(int) * r->csOff[r->nCells]);
  if (moms) memcpy(moms, r->moms, sizeof(double) * 10 * r->nCells);
}
void orc_set_cv(orc_region *r, const double *cv) {
  memcpy(r->cv, cv, sizeof(double) * 5 * r->nCells);
  set_gas_vars(r);
  set_dependent_vars(r);
}
void orc_get_cv(const orc_region *r, double *cv) { memcpy(cv, r->cv, sizeof(double) * 5 * r->nCells); }
void orc_get_dv(const orc_region *r, double *dv) { memcpy(dv, r->dv, sizeof(double) * 3 * r->nCells); }
void orc_get_rhs(const orc_region *r, double *rhs) { memcpy(rhs, r->rhs, sizeof(double) * 5 * r->nCells); }
void orc_get_grad(const orc_region *r, double *g) { memcpy(g, r->grad, sizeof(double) * 15 * r->nCells); }
double orc_time(const orc_region *r) { return r->currentTime; }
double orc_last_dt(const orc_region *r) { return r->dtMin; }

/* ------------------------------------------------------------------------ */
/* Test hooks: the routines above applied to closed-form inputs.            */
/* ------------------------------------------------------------------------ */
/* rk_update on y' = lam y for one cell (RKInitMP copies cv to cvOld before
 * every stage; the residual of dU/dt = lam U is rhs = -lam vol U), dt = 1;
 * y[k] receives the value after stage k+1. */
void orc_test_rk3_linear(double lam, double y0, double *y) {
  orc_region r;
  memset(&r, 0, sizeof r);
  double vol = 0.75, cv[5], cvOld[5], rhs[5], rhsSum[5] = {0};
  for (int k = 0; k < 5; k++) cv[k] = y0;
  r.nCells = 1; r.vol = &vol; r.cv = cv; r.cvOld = cvOld; r.rhs = rhs; r.rhsSum = rhsSum; r.dtMin = 1.0;
  for (int st = 1; st <= 3; st++) {
    for (int k = 0; k < 5; k++) { cvOld[k] = cv[k]; rhs[k] = -lam * vol * cv[k]; }
    rk_update(&r, st);
    y[st - 1] = cv[0];
  }
}

double orc_test_entropy_fix(double l, double d) { return entropy_fix_hh(l, d); }

double orc_test_venkat_term(double dm, double d2, double epsq, int guard) {
  return venkat_term(dm, d2, epsq, 1.0e-12, guard ? -1.0e-12 : 1.0e-12);
}

/* Threads of the oracle's independent per-cell, per-element and per-particle
 * loops (every such loop writes only its own entries, so the results do not
 * depend on the thread count). */
void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
