// Host-side mesh pipeline (see rflu_mesh.hpp).  Compiled with
// -ffp-contract=off so the geometry and moment arithmetic rounds like the
// reference's non-FMA build.
#include "rflu_mesh.hpp"

#include <algorithm>
#include <thread>
#include <array>
#include <cmath>
#include <cstdio>
#include <numeric>

namespace rflu {

namespace {

// Local face tables: modflu/RFLU_ModGrid.F90:92-102, 0-based (-1: the 4th
// vertex of a triangle).
const int kTetFace[4][4] = {{0, 1, 2, -1}, {1, 3, 2, -1}, {0, 2, 3, -1}, {0, 3, 1, -1}};
const int kHexFace[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                            {2, 3, 7, 6}, {0, 4, 7, 3}, {4, 5, 6, 7}};
const int kPriFace[5][4] = {{0, 2, 1, -1}, {0, 1, 4, 3}, {1, 2, 5, 4}, {0, 3, 5, 2}, {3, 4, 5, -1}};
const int kPyrFace[5][4] = {{0, 3, 2, 1}, {0, 1, 4, -1}, {1, 2, 4, -1}, {2, 3, 4, -1}, {0, 4, 3, -1}};

// Cells are rows of nvc vertex ids; with nvc = 8 a row holds a tetrahedron
// (4), pyramid (5), prism (6) or hexahedron (8), padded with -1.
int row_nv(const MeshInput &in, int i) {
  if (in.nvc == 4) return 4;
  int n = 0;
  for (int q = 0; q < 8; q++) n += in.c2v[(size_t)i * 8 + q] >= 0;
  return n;
}
const int (*faces_of(int nv, int &nlf))[4] {
  switch (nv) {
    case 4: nlf = 4; return kTetFace;
    case 5: nlf = 5; return kPyrFace;
    case 6: nlf = 5; return kPriFace;
    default: nlf = 6; return kHexFace;
  }
}

struct FaceKey {
  int k0, k1, k2;  // three smallest vertex ids, ascending
  int rank;        // insertion rank = cell*6 + localFace
  bool operator<(const FaceKey &o) const {
    if (k0 != o.k0) return k0 < o.k0;
    if (k1 != o.k1) return k1 < o.k1;
    if (k2 != o.k2) return k2 < o.k2;
    return rank < o.rank;
  }
};

template <int N>
std::array<int, N> sorted_verts(const int *v) {
  std::array<int, N> a;
  for (int q = 0; q < N; q++) a[q] = v[q];
  std::sort(a.begin(), a.end());
  return a;
}

// Face vector and centroid: libfloflu/FaceVector.F90, FaceCentroid.F90.
void face_vec_cen(const double *xyz, const int *v, int nv, double *fv, double *fcen) {
  const double *x1 = xyz + 3 * (size_t)v[0], *x2 = xyz + 3 * (size_t)v[1], *x3 = xyz + 3 * (size_t)v[2];
  if (nv == 3) {
    double dxa = x1[0] - x2[0], dya = x1[1] - x2[1], dza = x1[2] - x2[2];
    double dxb = x1[0] - x3[0], dyb = x1[1] - x3[1], dzb = x1[2] - x3[2];
    fv[0] = 0.5 * (dya * dzb - dyb * dza);
    fv[1] = -0.5 * (dxa * dzb - dxb * dza);
    fv[2] = 0.5 * (dxa * dyb - dxb * dya);
    const double thrd = 1.0 / 3.0;
    for (int d = 0; d < 3; d++) fcen[d] = thrd * (x1[d] + x2[d] + x3[d]);
  } else {
    const double *x4 = xyz + 3 * (size_t)v[3];
    double dxa = x3[0] - x1[0], dya = x3[1] - x1[1], dza = x3[2] - x1[2];
    double dxb = x2[0] - x4[0], dyb = x2[1] - x4[1], dzb = x2[2] - x4[2];
    fv[0] = 0.5 * (dza * dyb - dya * dzb);
    fv[1] = 0.5 * (dxa * dzb - dza * dxb);
    fv[2] = 0.5 * (dya * dxb - dxa * dyb);
    for (int d = 0; d < 3; d++) fcen[d] = 0.25 * (x1[d] + x2[d] + x3[d] + x4[d]);
  }
}

// Pair sort of (key, payload) in the partition order of the reference's
// QuickSortRFREALInteger (modfloflu/ModSortSearch.F90:898-990).  The c2c
// stencil keeps the 40 members with the largest shared-vertex fraction and
// the members that survive among ties are decided by this exact scheme, so
// the same steps are used; indices are 1-based as in the Fortran.
void pair_quicksort(double *key, int *val, int n) {
  if (n < 1) return;
  std::array<int, 64> sl{}, sr{};
  int top = 1;
  sl[1] = 1;
  sr[1] = n;
  auto A = [&](int i) -> double & { return key[i - 1]; };
  auto B = [&](int i) -> int & { return val[i - 1]; };
  while (top != 0) {
    int lo = sl[top], hi = sr[top];
    top--;
    while (true) {
      int i = lo, j = hi;
      double piv = A((lo + hi) / 2);
      while (true) {
        while (A(i) < piv) ++i;
        while (piv < A(j)) --j;
        if (i > j) break;
        std::swap(A(i), A(j));
        std::swap(B(i), B(j));
        ++i;
        --j;
        if (i > j) break;
      }
      if (j - lo >= hi - i) {
        if (lo < j) { ++top; sl[top] = lo; sr[top] = j; }
        lo = i;
      } else {
        if (i < hi) { ++top; sl[top] = i; sr[top] = hi; }
        hi = j;
      }
      if (!(lo < hi)) break;
    }
  }
}

// Singular values of the column-scaled n x 4 stencil matrix by one-sided
// Jacobi; returns the smallest ratio s(i)/s(i-1) of the descending values.
double min_sv_ratio(std::vector<std::array<double, 4>> &a, double &s0) {
  int n = (int)a.size();
  for (int c = 0; c < 4; c++) {
    double m = -1e308;
    for (int q = 0; q < n; q++) m = std::max(m, std::fabs(a[q][c]));
    for (int q = 0; q < n; q++) a[q][c] = a[q][c] / m;
  }
  for (int sweep = 0; sweep < 50; sweep++) {
    double off = 0.0;
    for (int p = 0; p < 4; p++)
      for (int q = p + 1; q < 4; q++) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int k = 0; k < n; k++) {
          alpha += a[k][p] * a[k][p];
          beta += a[k][q] * a[k][q];
          gamma += a[k][p] * a[k][q];
        }
        if (std::fabs(gamma) <= 1e-300) continue;
        off = std::max(off, std::fabs(gamma) / std::sqrt(alpha * beta));
        double zeta = (beta - alpha) / (2.0 * gamma);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        for (int k = 0; k < n; k++) {
          double ap = a[k][p], aq = a[k][q];
          a[k][p] = c * ap - s * aq;
          a[k][q] = s * ap + c * aq;
        }
      }
    if (off < 1e-15) break;
  }
  std::array<double, 4> sv;
  for (int c = 0; c < 4; c++) {
    double s = 0;
    for (int k = 0; k < n; k++) s += a[k][c] * a[k][c];
    sv[c] = std::sqrt(s);
  }
  std::sort(sv.begin(), sv.end(), std::greater<double>());
  s0 = sv[0];
  double r = 1e300;
  for (int i = 1; i < 4; i++) r = std::min(r, sv[i] / sv[i - 1]);
  return r;
}

}  // namespace

std::string build_mesh(const MeshInput &in, bool needStencils, MeshHost &m) {
  if (in.nvc != 4 && in.nvc != 8) return "nvc must be 4 (tets) or 8 (hexes, or mixed cells padded with -1)";
  if (in.nvf != 3 && in.nvf != 4) return "nvf must be 3 or 4";
  if (in.nvc == 4 && in.nvf != 3) return "tet meshes have triangular boundary faces (nvf 3)";
  const int N = in.nCells;
  m.nCells = N;
  m.nvc = in.nvc;
  std::vector<unsigned char> cnv(N);
  for (int i = 0; i < N; i++) {
    int nv = row_nv(in, i);
    if (nv != 4 && nv != 5 && nv != 6 && nv != 8) return "a cell row must hold 4, 5, 6 or 8 vertices";
    for (int q = 0; q < nv; q++)
      if (in.c2v[(size_t)i * in.nvc + q] < 0) return "a cell row's vertices must precede its -1 padding";
    cnv[i] = (unsigned char)nv;
  }
  auto lface = [&](int icl, int ifl, int q) {  // -1: the 4th vertex of a triangle
    int nlf;
    int lv = faces_of(cnv[icl], nlf)[ifl][q];
    return lv < 0 ? -1 : in.c2v[(size_t)icl * in.nvc + lv];
  };
  auto bface_nv = [&](int b) { return (in.nvf == 4 && in.bf2v[(size_t)b * 4 + 3] < 0) ? 3 : in.nvf; };

  // ---- face list: sort all (cell, local face) keys; a face's number is the
  // rank of its first occurrence (RFLU_ModFaceList.F90:888-1015 numbers faces
  // by first insertion, cells in their global order, which Rocflu assigns by
  // type: tets, hexes, prisms, pyramids); duplicates pair into interior faces.
  std::vector<FaceKey> keys;
  keys.reserve((size_t)N * 6);
  for (int icl = 0; icl < N; icl++) {
    int nlf;
    faces_of(cnv[icl], nlf);
    for (int ifl = 0; ifl < nlf; ifl++) {
      int v[4], nfv = lface(icl, ifl, 3) < 0 ? 3 : 4;
      for (int q = 0; q < nfv; q++) v[q] = lface(icl, ifl, q);
      std::sort(v, v + nfv);
      keys.push_back(FaceKey{v[0], v[1], v[2], icl * 6 + ifl});
    }
  }
  std::sort(keys.begin(), keys.end());
  struct Grp { int first, second; };
  std::vector<Grp> interior;
  std::vector<FaceKey> open;  // boundary candidates, still sorted by key
  for (size_t a = 0; a < keys.size();) {
    size_t b = a + 1;
    while (b < keys.size() && keys[b].k0 == keys[a].k0 && keys[b].k1 == keys[a].k1 && keys[b].k2 == keys[a].k2) b++;
    if (b - a == 1) open.push_back(keys[a]);
    else if (b - a == 2) interior.push_back({keys[a].rank, keys[a + 1].rank});
    else return "non-manifold face shared by more than two cells";
    a = b;
  }
  std::sort(interior.begin(), interior.end(), [](const Grp &x, const Grp &y) { return x.first < y.first; });
  const int F = (int)interior.size();
  m.nFaces = F;
  m.f2c.resize(2 * (size_t)F);
  m.f2v.assign(4 * (size_t)F, -1);
  for (int f = 0; f < F; f++) {
    int c1 = interior[f].first / 6, l1 = interior[f].first % 6;
    m.f2c[2 * f] = c1;
    m.f2c[2 * f + 1] = interior[f].second / 6;
    for (int q = 0; q < 4; q++) m.f2v[4 * f + q] = lface(c1, l1, q);
  }
  // boundary faces -> owning cell
  int nB = 0;
  for (int p = 0; p < in.nPatches; p++) nB += in.nBFaces[p];
  if (nB != (int)open.size()) {
    char buf[128];
    snprintf(buf, sizeof buf, "%zu open faces but %d patch faces", open.size(), nB);
    return buf;
  }
  m.nBFaces = nB;
  m.bf2c.resize(nB);
  m.bPatch.resize(nB);
  {
    int b = 0;
    for (int p = 0; p < in.nPatches; p++)
      for (int k = 0; k < in.nBFaces[p]; k++, b++) {
        int v[4], nfv = bface_nv(b);
        for (int q = 0; q < nfv; q++) v[q] = in.bf2v[(size_t)b * in.nvf + q];
        std::sort(v, v + nfv);
        FaceKey probe{v[0], v[1], v[2], -1};
        auto it = std::lower_bound(open.begin(), open.end(), probe);
        if (it == open.end() || it->k0 != v[0] || it->k1 != v[1] || it->k2 != v[2])
          return "boundary face not found among open cell faces";
        m.bf2c[b] = it->rank / 6;
        m.bPatch[b] = p;
      }
  }

  // ---- cell -> face-slot gather lists in ascending slot order (the order in
  // which the reference's face loop, then patch loop, touches each cell).
  m.cfOff.assign(N + 1, 0);
  for (int f = 0; f < F; f++) { m.cfOff[m.f2c[2 * f] + 1]++; m.cfOff[m.f2c[2 * f + 1] + 1]++; }
  for (int b = 0; b < nB; b++) m.cfOff[m.bf2c[b] + 1]++;
  for (int i = 0; i < N; i++) m.cfOff[i + 1] += m.cfOff[i];
  m.cfIdx.resize(m.cfOff[N]);
  {
    std::vector<int> fill(m.cfOff.begin(), m.cfOff.end() - 1);
    for (int f = 0; f < F; f++) {
      m.cfIdx[fill[m.f2c[2 * f]]++] = f << 1;
      m.cfIdx[fill[m.f2c[2 * f + 1]]++] = (f << 1) | 1;
    }
    for (int b = 0; b < nB; b++) m.cfIdx[fill[m.bf2c[b]]++] = (F + b) << 1;
  }

  // ---- geometry (RFLU_ModGeometry.F90:307-640), gathered per cell in the
  // reference's face order so every cell sums its terms in the same order.
  const int S = F + nB;
  m.fn.resize(4 * (size_t)S);
  m.fc.resize(3 * (size_t)S);
  std::vector<double> fvec(3 * (size_t)S);
  for (int s = 0; s < S; s++) {
    const int *v = s < F ? &m.f2v[4 * (size_t)s] : in.bf2v + (size_t)(s - F) * in.nvf;
    int nfv = s < F ? (v[3] < 0 ? 3 : 4) : bface_nv(s - F);
    double fv[3], fcen[3];
    face_vec_cen(in.xyz, v, nfv, fv, fcen);
    double mag = std::sqrt(fv[0] * fv[0] + fv[1] * fv[1] + fv[2] * fv[2]);
    double term = 1.0 / mag;
    m.fn[4 * (size_t)s] = fv[0] * term;
    m.fn[4 * (size_t)s + 1] = fv[1] * term;
    m.fn[4 * (size_t)s + 2] = fv[2] * term;
    m.fn[4 * (size_t)s + 3] = mag;
    for (int d = 0; d < 3; d++) { m.fc[3 * (size_t)s + d] = fcen[d]; fvec[3 * (size_t)s + d] = fv[d]; }
  }
  m.vol.resize(N);
  m.cofg.resize(3 * (size_t)N);
  const double THRD = 1.0 / 3.0;
  for (int i = 0; i < N; i++) {
    const int *cv = in.c2v + (size_t)i * in.nvc;
    const int nv = cnv[i];
    // RFLU_ComputeApproxCentroids (RFLU_ModGeometry.F90): term * (sum in vertex order)
    const double term0 = nv == 4 ? 1.0 / 4.0 : nv == 5 ? 1.0 / 5.0 : nv == 6 ? 1.0 / 6.0 : 1.0 / 8.0;
    double app[3];
    for (int d = 0; d < 3; d++) {
      double s = in.xyz[3 * (size_t)cv[0] + d];
      for (int q = 1; q < nv; q++) s = s + in.xyz[3 * (size_t)cv[q] + d];
      app[d] = term0 * s;
    }
    double vol = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int e = m.cfOff[i]; e < m.cfOff[i + 1]; e++) {
      int s = m.cfIdx[e] >> 1;
      bool second = m.cfIdx[e] & 1;
      const double *fv = &fvec[3 * (size_t)s], *fc = &m.fc[3 * (size_t)s];
      double t1 = (fc[0] - app[0]) * fv[0] + (fc[1] - app[1]) * fv[1] + (fc[2] - app[2]) * fv[2];
      double t2 = fc[0] * fv[0] + fc[1] * fv[1] + fc[2] * fv[2];
      if (!second) {
        vol = vol + t1;
        cx = cx + t2 * fc[0]; cy = cy + t2 * fc[1]; cz = cz + t2 * fc[2];
      } else {
        vol = vol - t1;
        cx = cx - t2 * fc[0]; cy = cy - t2 * fc[1]; cz = cz - t2 * fc[2];
      }
    }
    vol = THRD * vol;
    double term = 1.0 / (4.0 * vol);
    m.vol[i] = vol;
    m.cofg[3 * (size_t)i] = term * cx;
    m.cofg[3 * (size_t)i + 1] = term * cy;
    m.cofg[3 * (size_t)i + 2] = term * cz;
    if (!(vol > 0.0)) return "non-positive cell volume";
  }
  m.epsq.resize(N);
  for (int i = 0; i < N; i++) {
    // modflu/RFLU_ModLimiters.F90:459-462 with K = 1 (static per cell)
    double ds = std::pow(m.vol[i], THRD);
    m.epsq[i] = (1.0 * ds) * (1.0 * ds) * (1.0 * ds);
  }
  if (!needStencils) return "";

  // ---- c2c stencils (RFLU_ModStencilsCells.F90:750-1416): vertex
  // neighbours ranked by shared-vertex fraction, at most 40, then ascending.
  const int SMAX = 40, SMIN = 4;
  std::vector<int> v2cOff(in.nVert + 1, 0), v2c((size_t)N * in.nvc);
  for (int i = 0; i < N; i++)
    for (int q = 0; q < cnv[i]; q++) v2cOff[in.c2v[(size_t)i * in.nvc + q] + 1]++;
  for (int v = 0; v < in.nVert; v++) v2cOff[v + 1] += v2cOff[v];
  {
    std::vector<int> fill(v2cOff.begin(), v2cOff.end() - 1);
    for (int i = 0; i < N; i++)
      for (int q = 0; q < cnv[i]; q++) v2c[fill[in.c2v[(size_t)i * in.nvc + q]]++] = i;
  }
  m.csOff.assign(N + 1, 0);
  m.csMem.clear();
  m.moms.assign(10 * (size_t)N, 0.0);
  m.stencilMinSvRatio = 1e300;
  // Cells are independent: chunks of cells run on host threads, each with
  // its own member list, merged in cell order (the result does not depend
  // on the number of threads).
  struct Part {
    std::vector<int> mem, cnt;
    double minRatio = 1e300;
    int maxKeep = 0;
    std::string err;
  };
  int nThr = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (N < 20000) nThr = 1;
  std::vector<Part> parts(nThr);
  auto work = [&](int tid) {
    Part &P = parts[tid];
    const int c0 = (int)((long long)N * tid / nThr), c1 = (int)((long long)N * (tid + 1) / nThr);
    P.cnt.reserve(c1 - c0);
    P.mem.reserve((size_t)(c1 - c0) * 30);
    std::vector<int> cand, mem;
    std::vector<double> frac;
    std::vector<std::array<double, 4>> amat;
    for (int i = c0; i < c1; i++) {
      const int *vi = in.c2v + (size_t)i * in.nvc;
      const int nvi = cnv[i];
      const double iNVert = 1.0 / (double)nvi;  // 1/nVert of the cell itself (RFLU_BuildC2CStencilBasic)
      cand.clear();
      for (int q = 0; q < nvi; q++)
        for (int t = v2cOff[vi[q]]; t < v2cOff[vi[q] + 1]; t++)
          if (v2c[t] != i) cand.push_back(v2c[t]);
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      int own[8];
      std::copy(vi, vi + nvi, own);
      std::sort(own, own + nvi);
      frac.resize(cand.size());
      for (size_t t = 0; t < cand.size(); t++) {
        int oth[8], nvo = cnv[cand[t]];
        std::copy(in.c2v + (size_t)cand[t] * in.nvc, in.c2v + (size_t)cand[t] * in.nvc + nvo, oth);
        std::sort(oth, oth + nvo);
        int common = 0;
        for (int a = 0, b = 0; a < nvi && b < nvo;) {
          if (own[a] == oth[b]) { common++; a++; b++; }
          else if (own[a] < oth[b]) a++;
          else b++;
        }
        frac[t] = (double)common * iNVert;
      }
      int n2 = (int)cand.size();
      pair_quicksort(frac.data(), cand.data(), n2);
      int keep = std::min(n2, SMAX);
      mem.assign(cand.end() - keep, cand.end());
      std::sort(mem.begin(), mem.end());
      if (keep < SMIN) { P.err = "stencil smaller than 4 cells needs layer addition (not supported)"; return; }
      const double *x0 = &m.cofg[3 * (size_t)i];
      amat.resize(keep);
      for (int q = 0; q < keep; q++) {
        const double *xj = &m.cofg[3 * (size_t)mem[q]];
        double dx = xj[0] - x0[0], dy = xj[1] - x0[1], dz = xj[2] - x0[2];
        double term = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
        amat[q] = {term, term * dx, term * dy, term * dz};
      }
      double s0;
      double ratio = min_sv_ratio(amat, s0);
      P.minRatio = std::min(P.minRatio, ratio);
      if (!(s0 > 100.0 * 2.220446049250313e-16) || !(ratio > 0.01)) {
        P.err = "singular c2c stencil needs layer addition (not supported)";
        return;
      }
      P.maxKeep = std::max(P.maxKeep, keep);
      // RFLU_ComputeStencilMoments3D1 (RFLU_ModWeights.F90:1770-1894)
      double xm[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int q = 0; q < keep; q++) {
        const double *xj = &m.cofg[3 * (size_t)mem[q]];
        double dx = xj[0] - x0[0], dy = xj[1] - x0[1], dz = xj[2] - x0[2];
        double wt = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
        dx = wt * dx; dy = wt * dy; dz = wt * dz;
        xm[0] = xm[0] + dx * dx; xm[1] = xm[1] + dx * dy; xm[2] = xm[2] + dy * dy;
        xm[3] = xm[3] + dx * dz; xm[4] = xm[4] + dy * dz; xm[5] = xm[5] + dz * dz;
        xm[6] = xm[6] + wt * dx; xm[7] = xm[7] + wt * dy; xm[8] = xm[8] + wt * dz;
        xm[9] = xm[9] + wt * wt;
      }
      double r11 = std::sqrt(xm[0]);
      double ir11 = 1.0 / r11;
      double r12 = ir11 * xm[1];
      double r22 = std::sqrt(xm[2] - r12 * r12);
      double ir22 = 1.0 / r22;
      double r13 = ir11 * xm[3];
      double r23 = ir22 * (xm[4] - r12 * r13);
      double r33 = std::sqrt(xm[5] - (r13 * r13 + r23 * r23));
      double ir33 = 1.0 / r33;
      double r14 = ir11 * xm[6];
      double r24 = ir22 * (xm[7] - r12 * r14);
      double r34 = ir33 * (xm[8] - (r13 * r14 + r23 * r24));
      double r44 = std::sqrt(xm[9] - (r14 * r14 + r24 * r24 + r34 * r34));
      double *mo = &m.moms[10 * (size_t)i];
      mo[0] = r11; mo[1] = r12; mo[2] = r22; mo[3] = r13; mo[4] = r23;
      mo[5] = r33; mo[6] = r14; mo[7] = r24; mo[8] = r34; mo[9] = r44;
      P.mem.insert(P.mem.end(), mem.begin(), mem.end());
      P.cnt.push_back(keep);
    }
  };
  if (nThr == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nThr; t++) th.emplace_back(work, t);
    for (auto &t : th) t.join();
  }
  size_t total = 0;
  for (auto &P : parts) {
    if (!P.err.empty()) return P.err;
    total += P.mem.size();
  }
  m.csMem.reserve(total);
  int i = 0;
  for (auto &P : parts) {
    m.stencilMinSvRatio = std::min(m.stencilMinSvRatio, P.minRatio);
    m.stencilMax = std::max(m.stencilMax, P.maxKeep);
    m.csMem.insert(m.csMem.end(), P.mem.begin(), P.mem.end());
    for (int c : P.cnt) { m.csOff[i + 1] = m.csOff[i] + c; i++; }
  }
  return "";
}

}  // namespace rflu
