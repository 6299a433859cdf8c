// MI355X (gfx950) kernels and C ABI of the ppiclF coupling (include/picl_gpu.h).
// Included at the end of rflu_gpu.hip: one translation unit, so the particle
// stage shares the region's handle, stream, graph and device helpers.
//
// Layout in HBM:
//   particles, SoA in "slot" order: grouped by element, ascending original
//     (ppiclF) index within an element, lost particles last.  Rows: position,
//     velocity, the RK register y1 and ydot of those six components, ydotc
//     (vx, vy, vz, T) of the previous stage (the projection map); constants
//     (rho_p, d_p, V_p, SPL); original index; element.  Two copies: each
//     stage's counting sort by element writes the particles in their new
//     order into the other copy (a device-side parity flag selects the
//     current one, so a captured graph stays valid).  Lanes of a wave thus
//     read contiguous particle data and share their elements' neighbourhood,
//     and the deposit sums each element's particles in ppiclF's order;
//   cells (device numbering): piclVF, rho_g, grad rho_g (raw and final, 4
//     doubles per cell), the 20 interpolation inputs (160 B per cell), the 5
//     projected fields (SoA), 1/V_hex, ppiclF centroid (4 per cell), the 8
//     nodes (24 per cell) and bounding box for point location;
//   two uniform bin grids: element bounding boxes (location) and element
//     centroids (interpolation neighbours).
//
// Integer results (element ids, neighbour sets and order, summation order of
// the deposit) follow the oracle's rules exactly.  The drag law calls pow,
// exp and log, whose device versions may differ from glibc by an ulp, so the
// particle state is compared to the oracle with a tolerance.

#define PICL_NFLD 20
#define PICL_LRS 12  // PPICLF_LRS: doubles per particle in y
#define PICL_LRP 48  // PPICLF_LRP: doubles per particle in rprop
#define F32(x) ((double)(x##f))

namespace {

struct PiclPhys {
  int rmuFlag, amFlag, pgFlag, feedbackFlag, qsFlag, interpMode;
  double rmuRef, tref, suth, rcpPart;
  int heatTransfer;
  int per[3];                 // linear periodicity per axis
  double perLo[3], perHi[3];  // ppiclf_xdrange: the mesh's vertex bounds
  int vu;                     // viscous-unsteady force
  int col;                    // normal collision force
  double ksp, erest;          // collision time / dt, restitution coefficient
  double nbW;                 // ppiclf_d2chk(3): the neighbour width (Briney's pair terms, fluctuation means)
  int qsFluct, qsFluctAdapt;  // QS fluctuations: 1 Lattanzi, 2 Osnes; adaptive box
  double fluctW;              // ppiclf_d2chk(2): the fluctuation box's half-width
  long long fluctSeed;        // seed of the uniform draws
  // angular periodicity about z: the wedge [angX, angUp] of polar angles,
  // the sign switch angMid = 0.5 (xangle + angle), ppiclF's case (1 wedge
  // below 90 degrees, 2 quarter, 3 half cylinder), the rotation matrices of
  // +angle, -angle and 0 (row-major), and the x range of the removal test
  int ang = 0, angCase = 0;
  double angX = 0, angUp = 0, angMid = 0, angR[3][9] = {}, angXd[2] = {0, 0};
};

// the element a particle removed by the angular-periodic range test carries
// into the next stage
#define E_REMOVED (-3)

// Rows of PBuf::st
// particle state rows: x, y, z, u, v, w (and their y1, ydot), the feedback
// map, then the temperature's y, y1 and ydot, touched only with heat
// transfer.  With tangential collisions the angular velocity's y, y1 and
// ydot follow (ROW_O..), with the viscous-unsteady force VU_N entries of
// d(rho u)/dt of the gas (ROW_VU + c * VU_N + k) and of the particle
// (ROW_VU + 3 * VU_N + c * VU_N + k), k = 0 the newest (ppiclF's rprop3).
#define VU_N 5
enum { ROW_X = 0, ROW_Y1 = 6, ROW_YD = 12, ROW_YCP = 18, ROW_T = 22, ROW_TY1 = 23, ROW_TYD = 24, NROW = 25,
       ROW_O = 25, ROW_OY1 = 28, ROW_OYD = 31, ROW_VU = 34, NROW_VU = ROW_VU + 6 * VU_N };

struct PBuf {
  double *st = nullptr;  // [nRow][nP]
  double *pc = nullptr;  // [4][nP] rho_p, d_p, V_p, SPL
  int *pid = nullptr;    // original (ppiclF) index
  int *elem = nullptr;   // element (device cell) of the current stage, -1 lost
  double *w = nullptr;   // [3][nP] W = D(rho_g u_g)/Dt - rho_g dv/dt (rprop WDOTX..Z; amFlag 2)
  double *fq = nullptr;  // [3][nP] the fluctuating quasi-steady force (rprop FLUCTFX..Z; qsFluct)
};

struct PiclDev {
  int nP = 0, nC = 0;
  int ht = 0;  // heat transfer on: the temperature rows move with the particles
  int vu = 0;  // viscous-unsteady force on: the history rows move with the particles
  int rot = 0;  // tangential collisions on: the angular-velocity rows move with the particles
  int nRow = NROW;
  double *vuT = nullptr;  // [VU_N] ppiclf_timeBH, then ppiclf_nTimeBH
  PBuf buf[2];
  int *par = nullptr;     // buf[par[0]] is current
  double *ycN = nullptr;  // [4][nP] ydotc of this stage, by current slot
  int *src = nullptr;     // new position -> current slot (active particles)
  int *dest = nullptr;    // current slot -> new position
  int *srt = nullptr;     // new position -> current slot, each element's particles by original index
  int *cnt = nullptr, *off = nullptr, *bsum = nullptr, *nAct = nullptr, *lostCur = nullptr;
  double *fld = nullptr;   // [nC][20]
  double *gRho = nullptr;  // [nC][4]
  double *proj = nullptr;  // [5][nC]
  double *rexp = nullptr, *cen = nullptr, *X = nullptr, *bbox = nullptr;
  int *refId = nullptr, *inv = nullptr;
  int lbN[3] = {1, 1, 1};
  double lbLo[3] = {0, 0, 0}, lbH[3] = {1, 1, 1};
  int *lbOff = nullptr, *lbMem = nullptr;
  int cbN[3] = {1, 1, 1};
  double cbLo[3] = {0, 0, 0}, cbH[3] = {1, 1, 1}, R[3] = {0, 0, 0};
  int *cbOff = nullptr;
  double *cbRec = nullptr;  // per bin member: cx, cy, cz, (reference id << 32 | device id)
  double d2Max[3] = {0, 0, 0};
  double *dbg = nullptr;  // [20][nP] interpolated values (debug)
  // Gaussian-filter projection (gauss = 1): node values per mesh vertex
  int gauss = 0, nV = 0, nCB = 0;
  double *vX = nullptr;     // [3][nV] vertex coordinates
  int *ev = nullptr;        // [8][nC] vertex of ppiclF node n of each (device) element
  int *ecb = nullptr;       // centroid bin of each element
  int *binCnt = nullptr;    // particles per centroid bin (this stage)
  int *gBox = nullptr;      // [6] lowest x, y, z and highest x, y, z centroid bin holding particles (this stage)
  int *gvList = nullptr;    // vertices within reach of a bin holding particles (this stage), gvList[nV] = count
  double *gq = nullptr;     // by sorted position: [nP][4] x, y, z, map0 * multfci, then [nP][4] map1..4 * multfci
  double *xo = nullptr;     // [3][nP] position at the start of the stage, by current slot
  double *vproj = nullptr;  // [5][nV]
  int *vperm = nullptr;     // vertices in Morton order (the order of the deposit's vertex list)
  double gRdum = 0, gMult = 0, gD2chk = 0, gReach = 0, gReach2 = 0;
  // element test of the deposit: an element's particles lie within gHalf of
  // its centroid per axis; it is a candidate when that box comes within
  // the cut-off (gCut2 = the squared cut-off with a relative margin)
  double gHalf[3] = {0, 0, 0}, gCut2 = 0;
  // periodic images of the Gaussian deposit: shift hi - lo per periodic
  // axis (0 elsewhere); cbTop = the largest centroid bin coordinate
  double gShift[3] = {0, 0, 0}, cbTop[3] = {0, 0, 0};
  int gImg = 0;  // any periodic axis
  // collisions: a grid of cubes of side colH >= the largest diameter over
  // the mesh's bounding box, x fastest, coordinates clamped into it (one
  // stage's located particles); cell lists by colOff / colMem
  int col = 0, colNb = 0, colN[3] = {1, 1, 1};
  double colLo[3] = {0, 0, 0}, colHi[3] = {0, 0, 0}, colH = 1, colR = 0, colDmax = 0;
  int *colCnt = nullptr, *colOff = nullptr, *colFill = nullptr, *colMem = nullptr, *colB = nullptr,
      *colBsum = nullptr;
  double *fc = nullptr;  // [3][nP] collision force by current slot
  double *tq = nullptr;  // [3][nP] collision torque by current slot (tangential collisions)
  int grid = 0;          // the grid of cubes is built each stage (collisions or the added-mass pairs)
  // Briney's added mass (amFlag 2), per stage
  int am2 = 0;
  double *vint = nullptr;    // [20][nP] interpolated values by current slot
  double *famB = nullptr;    // [4][nP] famx, famy, famz, rmass_add by current slot
  double *snum = nullptr;    // [nP] SDrho's numerator by current slot
  double *vrelP = nullptr;   // [3][nP] relative velocity by original index
  double *rphifP = nullptr;  // [nP] the particle loop's rphif by original index
  int *actP = nullptr;       // [nP] located this stage, by original index
  int *lastP = nullptr;      // [1] 1 + the largest original index located this stage
  // COLLISIONS >= 3: vorticity and its material derivative, interpolated
  int rot3 = 0;
  double *fb = nullptr;      // [3][nC] feedback force per cell (previous projection)
  double *fbRaw = nullptr;   // [nC][9] its LSQ gradient
  double *fbg = nullptr;     // [nC][9] after WENO-XYZ and the x-force positivity limiter
  double *fldRot = nullptr;  // [nC][6] vorticity, D(omega)/Dt per element
  int nVint = PICL_NFLD;     // values per particle in vint (20, 26 with the rotational fields)
  // angular periodicity with z periodic: a particle whose x left the wedge's
  // range after the update is marked (rmMark by current slot) and enters the
  // next stage with element E_REMOVED, which the location turns into lost
  int *rmMark = nullptr;
  // QS fluctuations: the box means by current slot, the SetYdot call counter
  int fl = 0;
  int pairs = 0;             // the two-phase stage with k_pairs (Briney or fluctuations)
  double *fmean = nullptr;   // [7][nP] sums of u, v, w, u^2, v^2, w^2 and the count
  int *nCall = nullptr;      // [1]
};

// ---------------------------------------------------------------------------
// Point location: the lowest-numbered element whose trilinear map holds the
// point within |r|,|s|,|t| <= 1 + 1e-10 (Newton from the element centre);
// the stand-in for gslib findpts.  Same operations as the oracle.
// ---------------------------------------------------------------------------
__device__ bool trilinear_inside(const double *__restrict__ X, const double xp[3], double *maxrst = nullptr,
                                 double *rst = nullptr) {
  double r = 0.0, s = 0.0, t = 0.0, prev = 0.0;
  for (int it = 0; it < 25; it++) {
    double a[2] = {0.5 * (1.0 - r), 0.5 * (1.0 + r)};
    double b[2] = {0.5 * (1.0 - s), 0.5 * (1.0 + s)};
    double c[2] = {0.5 * (1.0 - t), 0.5 * (1.0 + t)};
    const double da[2] = {-0.5, 0.5};
    double x[3] = {0.0, 0.0, 0.0};
    double J[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const int i = n & 1, j = (n >> 1) & 1, k = n >> 2;
      double w = a[i] * b[j] * c[k];
      double wr = da[i] * b[j] * c[k], ws = a[i] * da[j] * c[k], wt = a[i] * b[j] * da[k];
#pragma unroll
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
    if (!(fabs(det) > 0.0)) return false;
    double dr = (f0 * m00 - J[0][1] * (f1 * J[2][2] - J[1][2] * f2) + J[0][2] * (f1 * J[2][1] - J[1][1] * f2)) / det;
    double ds = (J[0][0] * (f1 * J[2][2] - J[1][2] * f2) - f0 * m01 + J[0][2] * (J[1][0] * f2 - f1 * J[2][0])) / det;
    double dt = (J[0][0] * (J[1][1] * f2 - f1 * J[2][1]) - J[0][1] * (J[1][0] * f2 - f1 * J[2][0]) + f0 * m02) / det;
    r = r + dr; s = s + ds; t = t + dt;
    // converged, or the steps no longer shrink: they are rounding noise,
    // about ulp(x) / (element size) in reference coordinates
    double step = fabs(dr) + fabs(ds) + fabs(dt);
    if (step < 1.0e-12 || (it >= 2 && step >= 0.5 * prev)) break;
    prev = step;
  }
  const double lim = 1.0 + 1.0e-10;
  if (maxrst) *maxrst = fmax(fabs(r), fmax(fabs(s), fabs(t)));
  if (rst) { rst[0] = r; rst[1] = s; rst[2] = t; }
  return fabs(r) <= lim && fabs(s) <= lim && fabs(t) <= lim;
}

// hint: an element that probably holds the point (the particle's element of
// the previous stage).  A point well inside it (|r|,|s|,|t| <= 1 - 1e-8) lies
// in no other element, not even within the 1e-10 tolerance, so the rule's
// answer is the hint itself; otherwise the full search decides.
// rst (optional): the point's reference coordinates in the element found
// (the same Newton solution a separate call on that element gives).
__device__ int picl_locate(const PiclDev &P, const double xp[3], int hint = -1, double *rst = nullptr) {
  if (!(isfinite(xp[0]) && isfinite(xp[1]) && isfinite(xp[2]))) return -1;
  if (hint >= 0) {
    double m = 2.0;
    if (trilinear_inside(P.X + 24 * (size_t)hint, xp, &m, rst) && m <= 1.0 - 1.0e-8) return hint;
  }
  int b[3];
#pragma unroll
  for (int l = 0; l < 3; l++) {
    double t = (xp[l] - P.lbLo[l]) / P.lbH[l];
    t = fmin(fmax(t, 0.0), (double)(P.lbN[l] - 1));
    b[l] = (int)t;
  }
  int bin = b[0] + P.lbN[0] * (b[1] + P.lbN[1] * b[2]);
  for (int q = P.lbOff[bin]; q < P.lbOff[bin + 1]; q++) {
    int e = P.lbMem[q];
    const double *bb = P.bbox + 8 * (size_t)e;
    bool in = true;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double tol = 1.0e-9 * (bb[4 + d] - bb[d]);
      if (xp[d] < bb[d] - tol || xp[d] > bb[4 + d] + tol) in = false;
    }
    if (in && trilinear_inside(P.X + 24 * (size_t)e, xp, nullptr, rst)) return e;
  }
  return -1;
}

// ---------------------------------------------------------------------------
// Gradient of rho_g: WENO-XYZ (modflu/RFLU_ModWENO.F90:1402-1588) over the
// cell and its stencil in stencil order, then the positivity limiter
// (modflu/RFLU_ModLimiters.F90:793-921) over the cell's faces.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLK) k_rhog(Dev d, PiclDev P) {
  int i = xcd_block() * BLK + threadIdx.x;
  if (i >= d.nC) return;
  const unsigned *mem = reinterpret_cast<const unsigned *>(d.csMem);
  const double2 *gr = reinterpret_cast<const double2 *>(d.gRaw);
  int sb = d.csOff[i], se = d.csOff[i + 1];
  // one pass over the stencil, the next member's gradient in flight while
  // the current one is weighted; each component keeps its own sequence
  // (the cell first, then the members in stencil order)
  double2 a = gr[2 * (size_t)i], b = gr[2 * (size_t)i + 1];
  double g0 = a.x, g1 = a.y, g2 = b.x;
  double s0 = 1.0 / (g0 * g0 + 1.0E-15), s1 = 1.0 / (g1 * g1 + 1.0E-15), s2 = 1.0 / (g2 * g2 + 1.0E-15);
  double l0 = s0 * g0, l1 = s1 * g1, l2 = s2 * g2;
  double2 na = make_double2(0.0, 0.0), nb = make_double2(0.0, 0.0);
  if (sb < se) {
    unsigned j = mem[sb] & CS_IDX;
    na = gr[2 * (size_t)j];
    nb = gr[2 * (size_t)j + 1];
  }
  unsigned jn = (sb + 1 < se) ? (mem[sb + 1] & CS_IDX) : 0u;
  for (int s = sb; s < se; s++) {
    double t0 = na.x, t1 = na.y, t2 = nb.x;
    if (s + 1 < se) {
      na = gr[2 * (size_t)jn];
      nb = gr[2 * (size_t)jn + 1];
      if (s + 2 < se) jn = mem[s + 2] & CS_IDX;
    }
    double w0 = 1.0 / (t0 * t0 + 1.0E-15), w1 = 1.0 / (t1 * t1 + 1.0E-15), w2 = 1.0 / (t2 * t2 + 1.0E-15);
    s0 = s0 + w0; s1 = s1 + w1; s2 = s2 + w2;
    l0 = l0 + w0 * t0; l1 = l1 + w1 * t1; l2 = l2 + w2 * t2;
  }
  double g[3] = {l0 / s0, l1 / s1, l2 / s2};
  int N = d.nC;
  double varc = d.rhog[i];
  double xc = d.cofg[i], yc = d.cofg[N + i], zc = d.cofg[2 * N + i];
  bool zero = false;
  for (int e = d.cfOff[i]; e < d.cfOff[i + 1]; e++) {
    int slot = d.cfIdx[e] >> 1;
    double dx = d.fcx[slot] - xc, dy = d.fcy[slot] - yc, dz = d.fcz[slot] - zc;
    double varf = varc + g[0] * dx + g[1] * dy + g[2] * dz;
    if (varf <= 0.0) zero = true;
  }
  if (zero) g[0] = g[1] = g[2] = 0.0;
  double2 *o = reinterpret_cast<double2 *>(P.gRho + 4 * (size_t)i);
  o[0] = make_double2(g[0], g[1]);
  o[1] = make_double2(g[2], 0.0);
}

// The same WENO-XYZ + positivity limiter over the gradient tiles: the tile's
// slots (its cells and their stencil members) carry w = 1/(g^2 + 1e-15) and
// w*g per component, formed once per slot in LDS instead of once per
// stencil appearance; each cell then adds them in stencil order.  Same
// operations on the same values, same bits as k_rhog.
template <int TB>
__global__ void __launch_bounds__(TB) k_rhog_tile(Dev d, PiclDev P) {
  extern __shared__ double2 sW[];  // [nU][3]: (w0, w0*g0), (w1, w1*g1), (w2, w2*g2)
  int tile = xcd_block();
  int u0 = d.tOff[tile], nU = d.tOff[tile + 1] - u0;
  const double2 *gr = reinterpret_cast<const double2 *>(d.gRaw);
  for (int u = threadIdx.x; u < nU; u += TB) {
    size_t c = d.tMem[u0 + u];
    double2 a = gr[2 * c], b = gr[2 * c + 1];
    double t0 = a.x, t1 = a.y, t2 = b.x;
    double w0 = 1.0 / (t0 * t0 + 1.0E-15), w1 = 1.0 / (t1 * t1 + 1.0E-15), w2 = 1.0 / (t2 * t2 + 1.0E-15);
    sW[3 * u] = make_double2(w0, w0 * t0);
    sW[3 * u + 1] = make_double2(w1, w1 * t1);
    sW[3 * u + 2] = make_double2(w2, w2 * t2);
  }
  __syncthreads();
  int i = tile * TB + threadIdx.x;
  if (i >= d.nC) return;
  double2 o0 = sW[3 * threadIdx.x], o1 = sW[3 * threadIdx.x + 1], o2 = sW[3 * threadIdx.x + 2];
  double s0 = o0.x, s1 = o1.x, s2 = o2.x, l0 = o0.y, l1 = o1.y, l2 = o2.y;
  int sb = d.csOff[i], se = d.csOff[i + 1];
  for (int s = sb; s < se; s++) {
    int u = d.sSlot[s] & 0x7fffu;
    double2 a = sW[3 * u], b = sW[3 * u + 1], c = sW[3 * u + 2];
    s0 = s0 + a.x; s1 = s1 + b.x; s2 = s2 + c.x;
    l0 = l0 + a.y; l1 = l1 + b.y; l2 = l2 + c.y;
  }
  double g[3] = {l0 / s0, l1 / s1, l2 / s2};
  int N = d.nC;
  double varc = d.rhog[i];
  double xc = d.cofg[i], yc = d.cofg[N + i], zc = d.cofg[2 * N + i];
  bool zero = false;
  for (int e = d.cfOff[i]; e < d.cfOff[i + 1]; e++) {
    int slot = d.cfIdx[e] >> 1;
    double dx = d.fcx[slot] - xc, dy = d.fcy[slot] - yc, dz = d.fcz[slot] - zc;
    double varf = varc + g[0] * dx + g[1] * dy + g[2] * dz;
    if (varf <= 0.0) zero = true;
  }
  if (zero) g[0] = g[1] = g[2] = 0.0;
  double2 *o = reinterpret_cast<double2 *>(P.gRho + 4 * (size_t)i);
  o[0] = make_double2(g[0], g[1]);
  o[1] = make_double2(g[2], 0.0);
}

// ---------------------------------------------------------------------------
// PICL_TEMP_Runge field assembly (rocpicl/PICL_TEMP_Runge.F90:712-842): one
// value per element, already averaged over its 8 identical node copies the way
// LocalInterp does (sum of eight, divided by 8, ppiclf_solve.f:2363-2373).
// The cv the reference reads here has been through CellGradientsMP's and
// SourceTermsMP's round trips; they are replayed in registers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sum8(double v) {
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 8; q++) s = s + v;
  return s;
}
// Projected field k at ppiclF node n of (device) element i: the box deposit
// gives all 8 nodes the element's value; the Gaussian filter one per vertex.
__device__ __forceinline__ double node_val(const PiclDev &P, int k, int i, int n) {
  return P.gauss ? P.vproj[(size_t)k * P.nV + P.ev[(size_t)n * P.nC + i]] : P.proj[(size_t)k * P.nC + i];
}
// Rocflu's read-backs loop over the nodes lz, ly, lx (PICL_TEMP_Runge.F90:923-1000, 1217-1234)
__device__ __forceinline__ double phip_of(const PiclDev &P, int i, double vol) {
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 8; q++) s = s + (0.125 * node_val(P, 0, i, q)) * vol;
  s = s / vol;
  if (s > F32(0.62)) s = F32(0.62);
  return s;
}
__device__ __forceinline__ double sum_nodes(const PiclDev &P, int k, int i) {
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 8; q++) s = s + node_val(P, k, i, q);
  return s;
}

// tri = 1 (trilinear interpolation) keeps the node value itself: the
// evaluation reads the nodes, not LocalInterp's average.
__global__ void __launch_bounds__(BLK) k_picl_fields(Dev d, Gas G, PiclDev P, int tri) {
  int i = xcd_block() * BLK + threadIdx.x;
  if (i >= d.nC) return;
  int N = d.nC;
  double r[8];
  load_rec(d.rec, i, r);
  double vol = d.vol[i], phi = d.phi[i];
  double vfg = 1.0 - phi;
  double rho = r[0];
  double rP = rho / vfg * vfg;                    // CellGradientsMP rho_g round trip
  double ir = 1.0 / rP;                           // SourceTermsMP cons->prim ...
  double uB = ir * (rho * r[1]), vB = ir * (rho * r[2]), wB = ir * (rho * r[3]);
  double cvB[4] = {rP, rP * uB, rP * vB, rP * wB};  // ... and prim->cons
  const double *g = d.grad + 15 * (size_t)i;
  double gg[15];
#pragma unroll
  for (int q = 0; q < 15; q++) gg[q] = g[q];
  double rh[4];
#pragma unroll
  for (int k = 0; k < 4; k++) rh[k] = d.rhs[k * N + i];
  double ug[3] = {cvB[1] / cvB[0], cvB[2] / cvB[0], cvB[3] / cvB[0]};
  double sdr[3];
#pragma unroll
  for (int m = 0; m < 3; m++) {
    double dot = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) dot = dot + ug[k] * gg[3 * (1 + m) + k];
    sdr[m] = (-(rh[1 + m] / vol) + ug[m] * rh[0] / vol) / cvB[0] + dot;
  }
  double T = r[4] / (rho / (1.0 - phi) * G.gcD);  // dv from RFLU_SetDependentVars
  double f[PICL_NFLD];
  f[0] = cvB[0] / (1.0 - phi);
  f[1] = cvB[1] / cvB[0];
  f[2] = cvB[2] / cvB[0];
  f[3] = cvB[3] / cvB[0];
  f[4] = gg[12]; f[5] = gg[13]; f[6] = gg[14];
  f[7] = sqrt(G.gD * G.gcD * T);
  f[8] = T;
  f[9] = phip_of(P, i, vol);
  f[10] = sdr[0]; f[11] = sdr[1]; f[12] = sdr[2];
  f[13] = -(rh[0] / vol);
  f[14] = gg[0]; f[15] = gg[1]; f[16] = gg[2];
  const double *q = P.gRho + 4 * (size_t)i;
  f[17] = q[0]; f[18] = q[1]; f[19] = q[2];
  double2 *o = reinterpret_cast<double2 *>(P.fld + PICL_NFLD * (size_t)i);
#pragma unroll
  for (int k = 0; k < PICL_NFLD / 2; k++)
    o[k] = tri ? make_double2(f[2 * k], f[2 * k + 1]) : make_double2(sum8(f[2 * k]) / 8, sum8(f[2 * k + 1]) / 8);
  if (P.rot3) {
    // PICL_TEMP_Runge.F90:747-809: vorticity and its material derivative
    // (stretching, dilatation, baroclinic, the feedback force's curl and its
    // misalignment with grad(rho phi))
    double dudx = gg[3], dudy = gg[4], dudz = gg[5], dvdx = gg[6], dvdy = gg[7], dvdz = gg[8];
    double dwdx = gg[9], dwdy = gg[10], dwdz = gg[11];
    double omgx = dwdy - dvdz, omgy = dudz - dwdx, omgz = dvdx - dudy;
    double divu = dudx + dvdy + dwdz;
    double dprdx = gg[0], dprdy = gg[1], dprdz = gg[2], dpdx = gg[12], dpdy = gg[13], dpdz = gg[14];
    const double *fg = P.fbg + 9 * (size_t)i;
    double F[3] = {P.fb[i], P.fb[(size_t)N + i], P.fb[2 * (size_t)N + i]};
    double ir = 1.0 / cvB[0], ir2 = ir * ir;
    double dodx = omgx * dudx + omgy * dudy + omgz * dudz;
    double dody = omgx * dvdx + omgy * dvdy + omgz * dvdz;
    double dodz = omgx * dwdx + omgy * dwdy + omgz * dwdz;
    dodx = dodx - omgx * divu;
    dody = dody - omgy * divu;
    dodz = dodz - omgz * divu;
    dodx = dodx + (dprdy * dpdz - dprdz * dpdy) * ir2;
    dody = dody + (dprdz * dpdx - dprdx * dpdz) * ir2;
    dodz = dodz + (dprdx * dpdy - dprdy * dpdx) * ir2;
    dodx = dodx + (fg[7] - fg[5]) * ir;  // dFz/dy - dFy/dz
    dody = dody + (fg[2] - fg[6]) * ir;  // dFx/dz - dFz/dx
    dodz = dodz + (fg[3] - fg[1]) * ir;  // dFy/dx - dFx/dy
    dodx = dodx + (dprdy * F[2] - dprdz * F[1]) * ir2;
    dody = dody + (dprdz * F[0] - dprdx * F[2]) * ir2;
    dodz = dodz + (dprdx * F[1] - dprdy * F[0]) * ir2;
    double fr[6] = {omgx, omgy, omgz, dodx, dody, dodz};
    double *ro = P.fldRot + 6 * (size_t)i;
#pragma unroll
    for (int q = 0; q < 6; q++) ro[q] = tri ? fr[q] : sum8(fr[q]) / 8;
  }
}

// ---------------------------------------------------------------------------
// Forces: ppiclf_user_QS_Parmar (libpicl/user_files/ppiclf_user_QuasiSteady.f:22-122);
// real*4 literals widened as the reference's compiler does.
// ---------------------------------------------------------------------------
__device__ double parmar_rcd1(double mp, double re) {
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
    } else {
      double rcd_M1 = (1.0 + F32(0.118) * pow(re, F32(0.813))) +
                      (re / 24.0) * F32(0.69) / (1.0 + 3550.0 / pow(re, F32(.793)));
      double m2 = mp * mp, m3 = mp * mp * mp;
      double lrep = log(re);
      if (mp <= 1.0) {
        double C1 = F32(6.48), C2 = F32(9.28), C3 = F32(12.21);
        double f1M = -F32(1.884) + F32(8.422) * mp - F32(13.70) * m2 + F32(8.162) * m3;
        double f2M = -F32(2.228) + F32(10.35) * mp - F32(16.96) * m2 + F32(9.840) * m3;
        double f3M = F32(4.362) - F32(16.91) * mp + F32(19.84) * m2 - F32(6.296) * m3;
        double factor = f1M * (lrep - C2) * (lrep - C3) / ((C1 - C2) * (C1 - C3)) +
                        f2M * (lrep - C1) * (lrep - C3) / ((C2 - C1) * (C2 - C3)) +
                        f3M * (lrep - C1) * (lrep - C2) / ((C3 - C1) * (C3 - C2));
        rcd1 = rcd_mcr + (rcd_M1 - rcd_mcr) * factor;
      } else {
        double rcd_M2 = (1.0 + F32(0.107) * pow(re, F32(0.867))) +
                        (re / 24.0) * F32(0.646) / (1.0 + 861.0 / pow(re, F32(.634)));
        if (mp < 1.75) {
          double C1 = F32(6.48), C2 = F32(8.93), C3 = F32(12.21);
          double f1M = -F32(2.963) + F32(4.392) * mp - F32(1.169) * m2 - F32(0.027) * m3 -
                       F32(0.233) * exp((1.0 - mp) / F32(0.011));
          double f2M = -F32(6.617) + F32(12.11) * mp - F32(6.501) * m2 + F32(1.182) * m3 -
                       F32(0.174) * exp((1.0 - mp) / F32(0.010));
          double f3M = -F32(5.866) + F32(11.57) * mp - F32(6.665) * m2 + F32(1.312) * m3 -
                       F32(0.350) * exp((1.0 - mp) / F32(0.012));
          double factor = f1M * (lrep - C2) * (lrep - C3) / ((C1 - C2) * (C1 - C3)) +
                          f2M * (lrep - C1) * (lrep - C3) / ((C2 - C1) * (C2 - C3)) +
                          f3M * (lrep - C1) * (lrep - C2) / ((C3 - C1) * (C3 - C2));
          rcd1 = rcd_M1 + (rcd_M2 - rcd_M1) * factor;
        } else {
          rcd1 = rcd_M2;
        }
      }
    }
  }
  return rcd1;
}

__device__ double qs_parmar(double rmachp, double rphip, double rep, double rmu, double dp) {
  const double rpi = 3.141592653589793116;  // acos(-1.0d0)
  double mp = fmax(rmachp, 0.01), re = fmax(rep, 0.1);
  double rcd1 = parmar_rcd1(mp, re);
  double phi_corr = 1.0 + F32(5.94) * fmin(rphip, 0.5);
  double beta = rcd1 * 3.0 * rpi * rmu * dp;
  return beta * phi_corr;
}

// Integer powers as gfortran expands them (x**3 = x2*x, x**4 = x2*x2,
// x**5 = x2*x3, x**(-3) = 1/x**3).
__device__ __forceinline__ double ip2(double x) { return x * x; }
__device__ __forceinline__ double ip3(double x) { return (x * x) * x; }
__device__ __forceinline__ double ip4(double x) { double x2 = x * x; return x2 * x2; }
__device__ __forceinline__ double ip5(double x) { double x2 = x * x; return x2 * (x2 * x); }

// Osnes et al. volume-fraction terms (ppiclf_user_QuasiSteady.f:208-218, 346-356)
__device__ void osnes_b(double phi, double re, double mp, double &b1, double &b2, double &b3) {
  double om = 1.0 - phi;
  b1 = F32(5.81) * phi / ip2(om) + F32(0.48) * pow(phi, 1.0 / 3.0) / ip3(om);
  b2 = ip2(om) * ip3(phi) * re * (F32(0.95) + F32(0.61) * ip3(phi) / (om * 2.0));
  b3 = fmin(sqrt(20.0 * mp), 1.0) * (F32(5.65) * phi - 22.0 * ip2(phi) + F32(23.4) * ip3(phi)) *
       (1.0 + tanh((mp - (F32(0.65) - F32(0.24) * phi)) / F32(0.35)));
}

// ppiclf_user_QS_ModifiedParmar (ppiclf_user_QuasiSteady.f:125-230)
__device__ double qs_modified_parmar(double rmachp, double rphip, double rep, double rmu, double dp) {
  const double rpi = 3.141592653589793116;  // acos(-1.0d0)
  double mp = fmax(rmachp, 0.01), phi = fmax(rphip, 0.0001), re = fmax(rep, 0.1);
  double rcd1 = parmar_rcd1(mp, re);
  double b1, b2, b3;
  osnes_b(phi, re, mp, b1, b2, b3);
  double cd = (24.0 / re) * rcd1;
  cd = cd / (1.0 - phi) + b3 + (24.0 / re) * (1.0 - phi) * (b1 + b2);
  return 3.0 * rpi * rmu * dp * (re / 24.0) * cd;
}

// ppiclf_user_QS_Osnes, Loth et al. drag with Osnes et al. volume fraction
// terms (ppiclf_user_QuasiSteady.f:233-362)
__device__ double qs_osnes(double rmachp, double rphip, double rep, double rmu, double dp) {
  const double rpi = 3.141592653589793116;  // acos(-1.0d0)
  const double gamma = 1.4;
  double mp = fmax(rmachp, 0.01), phi = fmax(rphip, 0.0001), re = fmax(rep, 0.1);
  double cd_loth;
  if (re <= 45.0) {
    double Knp = sqrt(0.5 * rpi * gamma) * mp / re, fKn;
    if (Knp > F32(0.01))
      fKn = 1.0 / (1.0 + Knp * (2.514 + 0.8 * exp(-0.55 / Knp)));
    else
      fKn = 1.0 / (1.0 + Knp * (2.514 + 0.8 * exp(-0.55 / F32(0.01))));
    double CD1 = (24.0 / re) * (1.0 + 0.15 * pow(re, 0.687)) * fKn;
    double sq = mp * sqrt(0.5 * gamma);
    double sgby2 = sqrt(0.5 * gamma);
    double JMt;
    if (mp <= 1.0)
      JMt = 2.26 * ip4(mp) + 0.14 * mp;
    else
      JMt = 1.6 * ip4(mp) + 0.25 * ip3(mp) + 0.11 * ip2(mp) + 0.44 * mp;
    double CD2 = (1.0 + 2.0 * ip2(sq)) * exp(-ip2(sq)) * mp / (ip3(sgby2) * sqrt(rpi)) +
                 (4.0 * ip4(sq) + 4.0 * ip2(sq) - 1.0) * erf(sq) / (2.0 * ip4(sgby2)) +
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
  osnes_b(phi, re, mp, b1, b2, b3);
  double cd = cd_loth / (1.0 - phi) + b3 + (24.0 / re) * (1.0 - phi) * (b1 + b2);
  return 3.0 * rpi * rmu * dp * (re / 24.0) * cd;
}

// ppiclf_user_QS_Gidaspow, Ergun / Wen-Yu per particle (ppiclf_user_QuasiSteady.f:364-397)
__device__ double qs_gidaspow(double rphip, double rphif, double rep, double rmu, double dp, double rhof,
                              double vmag) {
  const double rpi = 3.141592653589793116;  // acos(-1.0d0)
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

// Schiller-Naumann: rocinteract's INRT_DRAG_MODEL_SN correction
// psi = 1 + 0.15 Re^0.687 (rocinteract/INRT_CalcDrag.F90:189-191) in ppiclF's
// quasi-steady slot, beta = psi * 3 pi mu d_p (ppiclf_user_QuasiSteady.f:113),
// no volume-fraction correction.
__device__ double qs_sn(double rep, double rmu, double dp) {
  const double rpi = 3.141592653589793116;  // acos(-1.0d0)
  double re = fmax(rep, 0.1);
  double psi = 1.0 + 0.15 * pow(re, 0.687);
  return psi * 3.0 * rpi * rmu * dp;
}

// ---------------------------------------------------------------------------
// One particle per lane, in element-grouped order: locate, IDW-27
// interpolation (ppiclf_solve.f:2234-2409), ppiclf_user_SetYdot
// (libpicl/user_files/ppiclf_user.f:39-1042 with QS Parmar, AM Parmar
// ppiclf_user_AddedMass.f:16-124, pressure gradient; Euler flow) and the
// IntegrateRK3s_Rocflu update (ppiclf_solve.f:1845-1936, 2774-2793).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool key_less(double a, int ra, double b, int rb) {
  return a < b || (a == b && ra < rb);
}

__device__ __forceinline__ PBuf cur_buf(const PiclDev &P) { return P.par[0] ? P.buf[1] : P.buf[0]; }

// TRI: trilinear interpolation instead of IDW-27; QS: the quasi-steady law
// (ppiclf_user.f:473-485 numbering, 5 = Schiller-Naumann)
// ppiclf_user_HT_driver (libpicl/user_files/ppiclf_user_HeatTransfer.f:17-222):
// Q = pi kappa d (T_gas - T_p) Nu, Nusselt law 1 Stokes, 2 Ranz-Marshall,
// 3 Gunn, 4 Fox; ppiclF fixes Pr = 0.70 (ppiclf_user.f:177).
__device__ double heat_transfer(int flag, double rkappa, double dp, double tgas, double tpart, double rep,
                                double rphif, double rmachp) {
  const double rpi = 3.141592653589793116, rpr = 0.70, OneThird = 1.0 / 3.0;
  double Q_conv = rpi * rkappa * dp * (tgas - tpart);
  double Nuss = 0.0;
  if (flag == 1) {
    Nuss = 2.0;
  } else if (flag == 2) {
    Nuss = 2.0 + 0.6 * pow(rep, 0.5) * pow(rpr, OneThird);
  } else if (flag == 3) {
    double vg = rphif;
    Nuss = (7.0 - 10.0 * vg + 5.0 * vg * vg) * (1.0 + 0.7 * pow(rep, 0.2) * pow(rpr, OneThird)) +
           (1.33 - 2.4 * vg + 1.2 * vg * vg) * pow(rep, 0.7) * pow(rpr, OneThird);
  } else {
    Nuss = 2.0 * exp(-rmachp) / (1.0 + 17.0 * rmachp / rep) +
           0.495 * pow(rpr, OneThird) * pow(rep, 0.55) * ((1.0 + 0.5 * exp(-17.0 * rmachp / rep)) / 1.5);
  }
  return 0.0 + Q_conv * Nuss;
}

// ---------------------------------------------------------------------------
// Collisions (COLLISIONS 1, 2).  The search: every located particle goes into
// the grid cell of its cube (side colH, a third of the largest diameter or
// more; coordinates clamped into the grid, so a particle a rounding outside
// the box lands in an edge cell); a particle of diameter d scans the cells
// within (d + dmax) / 2 of itself (no neighbour it touches lies further), and
// of its images one period away on periodic axes, with the same clamping,
// and keeps the neighbours that overlap it.
// Their forces are summed in ppiclf_solve_NearestNeighbor's order on one
// rank (ppiclf_solve.f:470-560): the other particles by ascending index,
// then the ghost (periodic image) particles by index and image.
// ---------------------------------------------------------------------------
#define COL_CAP 32
__device__ __forceinline__ int col_coord(const PiclDev &P, double x, int l) {
  double c = floor((x - P.colLo[l]) / P.colH);
  return (int)fmin(fmax(c, 0.0), (double)(P.colN[l] - 1));
}

// The stage's location (k_particles repeats it and gets the same element)
// and the bucket counts
// cnt[e] += 1 for every active lane with e >= 0, one atomic per run of
// equal elements in consecutive lanes (the particles are mostly in element
// order from the last stage's sort).  Every active lane must call it;
// lanes that already left the kernel end any run.
__device__ __forceinline__ void count_by_runs(int *cnt, int e) {
  const int lane = threadIdx.x & 63;
  const unsigned long long act = __ballot(true);
  const int below = __shfl_up(e, 1, 64);
  const bool starts = e >= 0 && (lane == 0 || below != e);
  const unsigned long long brk = __ballot(e < 0 || lane == 0 || below != e) | ~act;
  if (starts) {
    const unsigned long long after = brk & ~((2ull << lane) - 1ull);
    const int end = after ? __builtin_ctzll(after) : 64;
    atomicAdd(&cnt[e], end - lane);
  }
}

// Slots from a per-key counter, one atomic per run of equal keys in
// consecutive lanes: returns ctr[key]'s old value plus the lane's place in
// its run (-1 for key < 0).  Every active lane must call it.
__device__ __forceinline__ int reserve_by_runs(int *ctr, int key) {
  const int lane = threadIdx.x & 63;
  const unsigned long long act = __ballot(true);
  const int below = __shfl_up(key, 1, 64);
  const bool starts = key >= 0 && (lane == 0 || below != key);
  const unsigned long long heads = __ballot(starts);
  const unsigned long long brk = __ballot(key < 0 || lane == 0 || below != key) | ~act;
  const int h = 63 - __builtin_clzll((heads & ((2ull << lane) - 1ull)) | 1ull);
  int base = 0;
  if (starts) {
    const unsigned long long after = brk & ~((2ull << lane) - 1ull);
    const int end = after ? __builtin_ctzll(after) : 64;
    base = atomicAdd(&ctr[key], end - lane);
  }
  base = __shfl(base, h, 64);
  return key >= 0 ? base + (lane - h) : -1;
}

__global__ void __launch_bounds__(BLK) k_col_count(PiclDev P) {
  int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= P.nAct[0]) return;
  const size_t nP = P.nP;
  PBuf C = cur_buf(P);
  double xp[3] = {C.st[t], C.st[nP + t], C.st[2 * nP + t]};
  int e = picl_locate(P, xp, C.elem[t]);
  C.elem[t] = e;
  int b = -1;
  if (e >= 0) {
    int c[3];
#pragma unroll
    for (int l = 0; l < 3; l++) c[l] = col_coord(P, xp[l], l);
    b = c[0] + P.colN[0] * (c[1] + P.colN[1] * c[2]);
  }
  P.colB[t] = b;
  count_by_runs(P.colCnt, b);
}

__global__ void __launch_bounds__(BLK) k_col_fill(PiclDev P) {
  int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= P.nAct[0]) return;
  int b = P.colB[t];
  int q = reserve_by_runs(P.colFill, b);  // the cube's order is fixed by k_col_sort
  if (b >= 0) P.colMem[P.colOff[b] + q] = t;
}

// ppiclf_user_EvalNearestNeighbor, collisional part (:118-233), for
// collisional_flag 1: the normal soft-sphere force of j on i (Cundall and
// Strack) with spring min(rmass_i pi^2 / (ksp dt)^2, the overlap's Hertz
// stiffness) and damping from the restitution coefficient; Ftmin = 0.
// Same operations as the oracle.
__device__ void col_pair(int flag, const double xi[3], const double vi[3], const double oi[3], double dpi, double rmi,
                         const double xj[3], const double vj[3], const double oj[3], double dpj, double rmj,
                         double rmass, double dt, double ksp, double erest, double fc[3], double tq[3]) {
  const double rpi = 3.141592653589793116, pi2 = rpi * rpi;
  double rthresh = 0.5 * (dpi + dpj);
  double rxdiff = xj[0] - xi[0], rydiff = xj[1] - xi[1], rzdiff = xj[2] - xi[2];
  double rdiff = rxdiff * rxdiff + rydiff * rydiff + rzdiff * rzdiff;
  rdiff = sqrt(rdiff);
  double ksp1 = rmass * rpi * rpi / ((ksp * dt) * (ksp * dt));
  const double E1 = 1.0e9, E2 = 1.0e9, nu1 = 0.35, nu2 = 0.35;
  double Estar = (1.0 - nu1 * nu1) / E1 + (1.0 - nu2 * nu2) / E2;
  Estar = 1.0 / Estar;
  double r1 = 0.5 * dpi, r2 = 0.5 * dpj;
  double Rstar = r1 * r2 / (r1 + r2);
  double ksp2 = (4.0 / 3.0) * Estar * sqrt(Rstar);
  ksp2 = ksp2 * sqrt(fabs(rdiff - rthresh));
  double ksp_min = fmin(ksp1, ksp2);
  double rmult = (rmi * rmj) / (rmi + rmj);
  double le = log(erest);
  double eta_n = -2.0 * sqrt(ksp_min) * le / sqrt(le * le + pi2) * sqrt(rmult);
  double rbot = 1.0 / rdiff;
  double rn[3] = {rxdiff * rbot, rydiff * rbot, rzdiff * rbot};
  double u12[3] = {vi[0] - vj[0], vi[1] - vj[1], vi[2] - vj[2]};
  if (flag >= 2) {  // contact-point velocity of the rotations
    double rad1 = 0.5 * dpi, rad2 = 0.5 * dpj;
    double A12[3];
#pragma unroll
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
#pragma unroll
  for (int c = 0; c < 3; c++) ut[c] = u12[c] - rv12_mag * rn[c];
  double ut_mag = sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
  ut_mag = fmax(ut_mag, 1.0e-8);
  double rt[3] = {ut[0] / ut_mag, ut[1] / ut_mag, ut[2] / ut_mag};
  double Ftmin = 0.0;
  if (flag >= 2 && ut_mag > 0.0) {  // Coulomb friction or tangential dashpot, the smaller
    const double mu_c = 0.4;
    double eta_t = eta_n;
    Ftmin = -fmin(mu_c * Fn_mag, eta_t * ut_mag);
  }
#pragma unroll
  for (int c = 0; c < 3; c++) fc[c] = fc[c] + rnmag * rn[c] + Ftmin * rt[c];
  if (flag >= 2) {  // contact torque and rolling resistance into ydotc(JOX..JOZ)
    double Ft[3] = {Ftmin * rt[0], Ftmin * rt[1], Ftmin * rt[2]};
    double rad1 = 0.5 * dpi;
    double tc[3] = {rad1 * (rn[1] * Ft[2] - rn[2] * Ft[1]), rad1 * (rn[2] * Ft[0] - rn[0] * Ft[2]),
                    rad1 * (rn[0] * Ft[1] - rn[1] * Ft[0])};
    const double thetar = F32(0.06);  // a real*4 literal
    double r12 = 0.5 * (dpi * dpj) / (dpi + dpj);
    double omgr[3] = {oi[0] - oj[0], oi[1] - oj[1], oi[2] - oj[2]};
    double omgr_mag = sqrt(omgr[0] * omgr[0] + omgr[1] * omgr[1] + omgr[2] * omgr[2]);
    omgr_mag = fmax(omgr_mag, 1.0e-8);
#pragma unroll
    for (int c = 0; c < 3; c++) tq[c] = tq[c] + tc[c] + -thetar * Fn_mag * r12 * omgr[c] / omgr_mag;
  }
}

__global__ void __launch_bounds__(BLK) k_contacts(Dev d, PiclDev P, PiclPhys Ph) {
  int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= P.nAct[0]) return;
  const size_t nP = P.nP;
  PBuf C = cur_buf(P);
  const double *st = C.st;
  double fc[3] = {0.0, 0.0, 0.0}, tq[3] = {0.0, 0.0, 0.0};
  if (C.elem[t] >= 0) {
    double xi[3] = {st[t], st[nP + t], st[2 * nP + t]};
    double dpi = C.pc[nP + t];
    const double Ri = 0.5 * (dpi + P.colDmax) * (1.0 + 1.0e-9);
    // overlapping neighbours: key = (ghost << 62) | (index << 5) | image
    long long key[COL_CAP];
    int slot[COL_CAP];
    int nk = 0;
    bool over = false;
    // per axis, the shifts whose query box meets the grid (sd = 0 always;
    // -1, +1 on a periodic axis near a face) with their cell ranges
    int nsd[3], sdv[3][3], clo[3][3], chi[3][3];
    double shv[3][3];
#pragma unroll
    for (int l = 0; l < 3; l++) {
      nsd[l] = 0;
      for (int sd = -1; sd <= 1; sd++) {
        if (sd != 0 && !Ph.per[l]) continue;
        double shl = sd == 0 ? 0.0 : (double)sd * (Ph.perHi[l] - Ph.perLo[l]);
        double q = xi[l] - shl;
        if (q + Ri < P.colLo[l] - Ri || q - Ri > P.colHi[l] + Ri) continue;
        int a = col_coord(P, q - Ri, l), b = col_coord(P, q + Ri, l);
        sdv[l][nsd[l]] = sd;
        shv[l][nsd[l]] = shl;
        clo[l][nsd[l]] = a;
        chi[l][nsd[l]] = b;
        nsd[l]++;
      }
    }
    for (int iz = 0; iz < nsd[2]; iz++)
    for (int iy = 0; iy < nsd[1]; iy++)
    for (int ix = 0; ix < nsd[0]; ix++) {
      // image index (sx + 1) + 3 (sy + 1) + 9 (sz + 1); 13 is the particle itself
      int im = (sdv[0][ix] + 1) + 3 * (sdv[1][iy] + 1) + 9 * (sdv[2][iz] + 1);
      double sh[3] = {shv[0][ix], shv[1][iy], shv[2][iz]};
      int lo[3] = {clo[0][ix], clo[1][iy], clo[2][iz]}, hi[3] = {chi[0][ix], chi[1][iy], chi[2][iz]};
      for (int cz = lo[2]; cz <= hi[2]; cz++)
        for (int cy = lo[1]; cy <= hi[1]; cy++) {
          int row = P.colN[0] * (cy + P.colN[1] * cz);
          // the row's cells are contiguous: one span of the member list
          for (int q = P.colOff[row + lo[0]]; q < P.colOff[row + hi[0] + 1]; q++) {
              int sj = P.colMem[q];
              if (im == 13 && sj == t) continue;
              double xj[3];
#pragma unroll
              for (int l = 0; l < 3; l++) xj[l] = sh[l] == 0.0 ? st[l * nP + sj] : st[l * nP + sj] + sh[l];
              double rx = xj[0] - xi[0], ry = xj[1] - xi[1], rz = xj[2] - xi[2];
              double rdiff = sqrt(rx * rx + ry * ry + rz * rz);
              if (!(rdiff < 0.5 * (dpi + C.pc[nP + sj]) + 0.0)) continue;
              if (nk == COL_CAP) { over = true; continue; }
              key[nk] = ((long long)(im != 13) << 62) | ((long long)C.pid[sj] << 5) | im;
              slot[nk] = sj;
              nk++;
            }
          }
    }
    if (over) atomicOr(d.err, E_CONTACT);
    for (int a = 1; a < nk; a++) {  // ascending key
      long long x = key[a];
      int y = slot[a], j = a - 1;
      while (j >= 0 && key[j] > x) { key[j + 1] = key[j]; slot[j + 1] = slot[j]; j--; }
      key[j + 1] = x;
      slot[j + 1] = y;
    }
    double vi[3] = {st[3 * nP + t], st[4 * nP + t], st[5 * nP + t]};
    double oi[3] = {0.0, 0.0, 0.0};
    if (P.rot) {
#pragma unroll
      for (int c = 0; c < 3; c++) oi[c] = st[(ROW_O + c) * nP + t];
    }
    double rmi = C.pc[t] * C.pc[2 * nP + t];
    double rmass = C.pc[2 * nP + t] * C.pc[t];  // volp * rhop (ppiclf_user.f:313)
    double dt = d.scal[S_DTMIN];
    for (int a = 0; a < nk; a++) {
      int sj = slot[a], im = (int)(key[a] & 31);
      double xj[3], vj[3], oj[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int l = 0, r = im; l < 3; l++, r /= 3) {
        int sd = r % 3 - 1;
        double x = st[l * nP + sj];
        xj[l] = sd == 0 ? x : x + (double)sd * (Ph.perHi[l] - Ph.perLo[l]);
        vj[l] = st[(3 + l) * nP + sj];
        if (P.rot) oj[l] = st[(ROW_O + l) * nP + sj];
      }
      col_pair(Ph.col, xi, vi, oi, dpi, rmi, xj, vj, oj, C.pc[nP + sj], C.pc[sj] * C.pc[2 * nP + sj], rmass, dt,
               Ph.ksp, Ph.erest, fc, tq);
    }
  }
#pragma unroll
  for (int c = 0; c < 3; c++) P.fc[c * nP + t] = fc[c];
  if (P.rot) {
#pragma unroll
    for (int c = 0; c < 3; c++) P.tq[c * nP + t] = tq[c];
  }
}

// ---------------------------------------------------------------------------
// Briney added mass (AMFORCE 2): libpicl/user_files/Added_Mass/
// ppiclf_user_AddedMass_functions.f, _subroutines.f, the unary and binary
// terms of ppiclf_user_AddedMass.f:164-400 and the pair terms of
// ppiclf_user_EvalNearestNeighbor.f:115-159.  Same operations as the
// oracle: real*4 literals widened, integer powers as gcc expands them.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ip6(double x) { double x3 = (x * x) * x; return x3 * x3; }
__device__ __forceinline__ double ip8(double x) { double x2 = x * x, x4 = x2 * x2; return x4 * x4; }
__device__ __forceinline__ double ip9(double x) { double x3 = (x * x) * x, x6 = x3 * x3; return x6 * x3; }

__device__ double am_b11_11(double d, double alpha) {
  double t = 2.0 / d;
  double asym = 0.5 * (0.046875 * ip6(t) + 0.03515625 * ip8(t));  // 3.0/64.0, 9.0/256.0
  double hot = F32(0.059) / (((d - F32(1.9098)) * ip2(d - F32(0.4782))) * ip3(d));
  double alpha_corr = ((alpha * alpha - F32(0.0902) * alpha) * F32(70.7731)) / ip6(d + F32(2.0936));
  return (asym + hot) + alpha_corr;
}
__device__ double am_b11_22(double d, double alpha) {
  double t = 2.0 / d;
  double asym = 0.5 * (0.01171875 * ip6(t) + 0.01171875 * ip8(t));  // 3.0/256.0
  double A = (F32(0.0003) + (F32(0.0262) * alpha) * alpha) - F32(0.0012) * alpha;
  double B = (F32(1.3127) + (F32(1.0401) * alpha) * alpha) - F32(1.2519) * alpha;
  return asym + A / ip6(d - B);
}
__device__ double am_b12_11(double d, double alpha) {
  double t = 2.0 / d;
  double asym = 0.5 * (-(0.375 * ip3(t)) - 0.005859375 * ip9(t));  // 3.0/8.0, 3.0/512.0
  double hot = -(F32(0.0006) / ip5(d - F32(1.5428)));
  double alpha_corr = -((F32(0.7913) * alpha) * exp(-((F32(0.9801) - F32(0.1075) * alpha) * d)));
  return (asym + hot) + alpha_corr;
}
__device__ double am_b12_22(double d, double alpha) {
  double t = 2.0 / d;
  double asym = 0.5 * (0.1875 * ip3(t) + 0.000732421875 * ip9(t));  // 3.0/16.0, 3.0/4096.0
  double hot = (-F32(0.1985) + F32(16.7372) * alpha) / ip6(d) + (F32(1.3907) - F32(48.2604) * alpha) / ip8(d);
  return asym + hot;
}

__device__ double am_ia_analytical(double rmax, double rad, double alpha) {
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
  return alpha * (B11 + 2.0 * (term1 + term2));
}
__device__ double am_ii_analytical(double rmax, double rad, double alpha) {
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
  return alpha * (B11 + 2.0 * ((term1 + term2) + term3));
}

// radial distribution function of hard spheres (Trokhymchuk et al. 2005 and
// erratum), r = distance / diameter; sqrt(3.0) and 2/3 in single precision
__device__ double am_rdf(double r, double alpha) {
  double eta = alpha;
  double d = pow((2.0 * eta) * (((eta * eta - 3.0 * eta) - 3.0) +
                                sqrt(3.0 * ((((ip4(eta) - 2.0 * ip3(eta)) + ip2(eta)) + 6.0 * eta) + 3.0))),
                 0x1.555556p-2);
  double f = (2.0 * eta) / (1.0 - eta);
  double mu = f * ((-1.0 - d / (2.0 * eta)) + eta / d);
  double alpha0 = f * ((-1.0 + d / (4.0 * eta)) - eta / (2.0 * d));
  double beta0 = (f * 0x1.bb67aep+0) * (-(d / (4.0 * eta)) - eta / (2.0 * d));
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
  const double f23 = 0x1.555556p-1;
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

// Simpson sums of the distant-neighbour corrections, used when rmax / rad <
// 3.49; pi = 4*ATAN(1.0) in single precision
__device__ double am_numerical(bool induced, double rmax, double rad, double alpha) {
  const int npts = 1001;
  double maxr = rad * 20.0;
  if (!(maxr > rmax)) return 0.0;
  double dr = (maxr - rmax) / (double)(npts + 1);
  auto f = [&](double d) {
    const double pi = 0x1.921fb6p+1;
    double g = am_rdf(d / (2.0 * rad), alpha);
    double f11 = induced ? am_b12_11(d / rad, alpha) : am_b11_11(d / rad, alpha);
    double f22 = induced ? am_b12_22(d / rad, alpha) : am_b11_22(d / rad, alpha);
    double rho = alpha / (0x1.555556p+0 * pi);
    double w = (((0x1.555556p-2 * rho) * 4.0) * pi) * ip2(d / rad);
    return (w * g) * (f11 + 2.0 * f22);
  };
  double integral = 0.0, r = rmax;
  integral = integral + f(r);
  integral = integral + f(maxr);
  for (int i = 2; i <= npts - 1; i++) {
    r = r + dr;
    double fi = f(r);
    integral = integral + (i % 2 == 0 ? 4.0 : 2.0) * fi;
  }
  return (dr * integral) / 3.0;
}

// resistance_pair: B11, B12 rotated onto the pair axis (x, y, z) = x_j - x_i,
// Qt (B Q) with matmul's k-ascending sums; R[k][l]: l < 3 B11, l >= 3 B12
__device__ void am_resistance_pair(double x, double y, double z, double alpha, double rad, double R[3][6]) {
  double gam = atan2(y, x), bet = -atan2(z, sqrt(x * x + y * y));
  double cb = cos(bet), sb = sin(bet), cg = cos(gam), sg = sin(gam);
  double Q[3][3] = {{cb * cg, cb * sg, -sb}, {-sg, cg, 0.0}, {sb * cg, sb * sg, cb}};
  double dist = sqrt(x * x + y * y + z * z) / rad;
  double b[2][3];
  b[0][0] = am_b11_11(dist, alpha);
  b[0][1] = b[0][2] = am_b11_22(dist, alpha);
  b[1][0] = am_b12_11(dist, alpha);
  b[1][1] = b[1][2] = am_b12_22(dist, alpha);
#pragma unroll
  for (int blk = 0; blk < 2; blk++) {
    double M[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) s = s + (k == i ? b[blk][i] : 0.0) * Q[k][j];
        M[i][j] = s;
      }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) s = s + Q[k][i] * M[k][j];
        R[i][3 * blk + j] = s;
      }
  }
}

// One thread per cube: its member list in ascending slot order, so the
// scans that sum in member order (the added-mass pair terms) are
// deterministic whatever order k_col_fill's atomics produced.
__global__ void __launch_bounds__(BLK) k_col_sort(PiclDev P) {
  int b = blockIdx.x * BLK + threadIdx.x;
  if (b >= P.colNb) return;
  int *L = P.colMem + P.colOff[b];
  int n = P.colOff[b + 1] - P.colOff[b];
  for (int a = 1; a < n; a++) {
    int x = L[a], j = a - 1;
    while (j >= 0 && L[j] > x) { L[j + 1] = L[j]; j--; }
    L[j + 1] = x;
  }
}

// ---------------------------------------------------------------------------
// Quasi-steady force fluctuations (QSFLUCT 1 Lattanzi, 2 Osnes):
// ppiclf_user_Fluctuations.f:32-418.  Same operations as the oracle.
// ---------------------------------------------------------------------------
// the uniform draws that stand in for RANDOM_NUMBER: a splitmix64 hash of
// (seed, call, original index, draw), uniform on (0, 1) with 53 bits
__device__ __forceinline__ double fluct_uniform(long long seed, int call, int pid, int k) {
  unsigned long long n = ((((unsigned long long)(unsigned)call << 32) | (unsigned)pid) << 3) + (unsigned long long)(k + 1);
  unsigned long long z = (unsigned long long)seed + 0x9E3779B97F4A7C15ull * n;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return ((double)(z >> 11) + 0.5) * 0x1.0p-53;
}

__device__ __forceinline__ void fl_acc(double &s, double &c, double x) {
  double t = s + x, z = t - s;
  c = c + ((s - (t - z)) + (x - z));
  s = t;
}

// one SDE step from the box means m (sums and count), the previous force F,
// the uniforms U and fac = rk3ark(stage) dt
__device__ void fluct_force(int flag, const double m[7], const double F[3], const double U[6], double rphip,
                            double rmachp, double rep, double vmag, const double vrel[3], double rmu, double dp,
                            double fac, double out[3]) {
  const double rpi = 3.141592653589793116, TwoPi = 6.283185307179586232;
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
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double Z = sqrt(-2.0 * log(U[2 * k])) * cos(TwoPi * U[2 * k + 1]);
      double dW = sqrt(fac) * Z;
      out[k] = (1.0 - aSDE * fac) * F[k] + bSDE * dW;
    }
    return;
  }
  double sigD = (fq + bq) * Fs;
  double bSDE_CD = sigD * sqrt(2.0 * tF_inv);
  // below 1e-8 m/s the axis is x and the transverse amplitude vanishes; bSDE_CD
  // keeps the unguarded sigD (ppiclf_user_Fluctuations.f:317-328)
  double avec[3] = {1.0, 0.0, 0.0}, CD_frac = 0.0;
  if (vmag > 1.0e-8) {
#pragma unroll
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
#pragma unroll
  for (int k = 0; k < 3; k++) bvec[k] = bvec[k] / dn;
  double cvec[3] = {avec[1] * bvec[2] - avec[2] * bvec[1], avec[2] * bvec[0] - avec[0] * bvec[2],
                    avec[0] * bvec[1] - avec[1] * bvec[0]};
  dn = fmax(1.0e-8, sqrt(ip2(cvec[0]) + ip2(cvec[1]) + ip2(cvec[2])));
#pragma unroll
  for (int k = 0; k < 3; k++) cvec[k] = cvec[k] / dn;
  double Z1 = sqrt(-2.0 * log(U[0])) * cos(TwoPi * U[1]);
  double Z2 = sqrt(-2.0 * log(U[2])) * cos(TwoPi * U[3]);
  double dW1 = sqrt(fac) * Z1, dW2 = sqrt(fac) * Z2;
  double cosrand = cos(TwoPi * U[4]), sinrand = sin(TwoPi * U[4]);
  double dvec[3];
#pragma unroll
  for (int k = 0; k < 3; k++) dvec[k] = bvec[k] * cosrand + cvec[k] * sinrand;
  dn = fmax(1.0e-8, sqrt(ip2(dvec[0]) + ip2(dvec[1]) + ip2(dvec[2])));
#pragma unroll
  for (int k = 0; k < 3; k++) dvec[k] = dvec[k] / dn;
#pragma unroll
  for (int k = 0; k < 3; k++) out[k] = (1.0 - aSDE * fac) * F[k] + bSDE_CD * dW1 * avec[k] + bSDE_CL * dW2 * dvec[k];
}

// Step 1b of ppiclf_user_SetYdot (ppiclf_user.f:393-468) over the
// neighbours within the neighbour width d2chk(3) (ppiclf_solve_NearestNeighbor's
// squared-distance filter), periodic images included, then:
// * Briney's added mass (am_flag 2): the unary term (Step 1a), the pair
//   terms of EvalNearestNeighbor (rdiff <= d2chk(3)) and the binary term
//   (Step 4, ppiclf_user.f:527-552), into famB = (famx, famy, famz,
//   rmass_add);
// * the fluctuation means (qs_fluct_flag >= 1; EvalNearestNeighbor.f:
//   318-360): the particle's own velocity, then every neighbour within the
//   box of half-width d2chk(2) (or the adaptive one), TwoSum-accumulated so
//   the result does not depend on the order, into fmean.
// The pair terms are added in the cubes' order (z, y, x, each cube's members
// by slot), not ppiclF's particle-index order: they agree with the oracle's
// to rounding.
__global__ void __launch_bounds__(BLK) k_pairs(Dev d, PiclDev P, PiclPhys Ph) {
  int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= P.nAct[0]) return;
  const size_t nP = P.nP;
  PBuf C = cur_buf(P);
  if (C.elem[t] < 0) return;
  const double *st = C.st;
  double v[PICL_NFLD];
#pragma unroll
  for (int f = 0; f < PICL_NFLD; f++) v[f] = P.vint[f * nP + t];
  double xi[3] = {st[t], st[nP + t], st[2 * nP + t]};
  double yv[3] = {st[3 * nP + t], st[4 * nP + t], st[5 * nP + t]};
  double vxr = v[1] - yv[0], vyr = v[2] - yv[1], vzr = v[3] - yv[2];
  double vmag = sqrt(vxr * vxr + vyr * vyr + vzr * vzr);
  if (vmag <= 1.0e-3) return;  // skipped by SetYdot
  const bool am = P.am2, fl = P.fl;
  double rhof = v[0], rmachp = vmag / v[7];
  double rphip = fmin(v[9], 0.62), rphif = 1.0 - rphip;
  double dpi = C.pc[nP + t], volp = C.pc[2 * nP + t];
  double Fam[3] = {0.0, 0.0, 0.0}, Wi[3] = {0.0, 0.0, 0.0}, Wsum[3] = {0.0, 0.0, 0.0}, rmass_add = 0.0;
  int nn = 0;
  if (am) {  // Step 1a: ppiclf_user_AM_Briney_Unary (Mach correction only, no volume yet)
    double rcd_am;
    if (rmachp > F32(0.6)) {
      rcd_am = 0x1.5104d6p+1;
    } else {
      double m2 = rmachp * rmachp;
      rcd_am = 1.0 + F32(1.8) * m2 + F32(7.6) * (m2 * m2);
    }
    rcd_am = rcd_am * 0.5;
    rmass_add = rhof * volp * rcd_am;
    double SDrho = v[13] + yv[0] * v[14] + yv[1] * v[15] + yv[2] * v[16];
    SDrho = SDrho / rphif;
    double vgradrho = vxr * v[17] + vyr * v[18] + vzr * v[19];
    Fam[0] = rcd_am * (vxr * SDrho + rhof * v[10] + v[1] * vgradrho);
    Fam[1] = rcd_am * (vyr * SDrho + rhof * v[11] + v[2] * vgradrho);
    Fam[2] = rcd_am * (vzr * SDrho + rhof * v[12] + v[3] * vgradrho);
#pragma unroll
    for (int c = 0; c < 3; c++) Wi[c] = C.w[c * nP + t];
  }
  double S[6], Cc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int cnt = 1;
  double box = Ph.fluctW;
  if (fl) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      S[k] = yv[k];
      S[3 + k] = yv[k] * yv[k];
    }
    if (Ph.qsFluctAdapt >= 1) {  // EvalNearestNeighbor.f:325-335
      double adptfilter = pow(10.0 * ip3(dpi) / fmax(F32(1.e-4), v[9]), 0x1.555556p-2);
      adptfilter = adptfilter / 2.0;
      box = fmax(Ph.fluctW, adptfilter);
    }
  }
  const double w = Ph.nbW, dist2 = w * w, Rq = w * (1.0 + 1.0e-9);
  const double alpha_local = fmin(F32(0.4), rphip), rad = 0.5 * dpi;
  int nsd[3], sdv[3][3], clo[3][3], chi[3][3];
  double shv[3][3];
#pragma unroll
  for (int l = 0; l < 3; l++) {
    nsd[l] = 0;
    for (int sd = -1; sd <= 1; sd++) {
      if (sd != 0 && !Ph.per[l]) continue;
      double shl = sd == 0 ? 0.0 : (double)sd * (Ph.perHi[l] - Ph.perLo[l]);
      double q = xi[l] - shl;
      if (q + Rq < P.colLo[l] - Rq || q - Rq > P.colHi[l] + Rq) continue;
      sdv[l][nsd[l]] = sd;
      shv[l][nsd[l]] = shl;
      clo[l][nsd[l]] = col_coord(P, q - Rq, l);
      chi[l][nsd[l]] = col_coord(P, q + Rq, l);
      nsd[l]++;
    }
  }
  for (int iz = 0; iz < nsd[2]; iz++)
  for (int iy = 0; iy < nsd[1]; iy++)
  for (int ix = 0; ix < nsd[0]; ix++) {
    const bool self = sdv[0][ix] == 0 && sdv[1][iy] == 0 && sdv[2][iz] == 0;
    double sh[3] = {shv[0][ix], shv[1][iy], shv[2][iz]};
    for (int cz = clo[2][iz]; cz <= chi[2][iz]; cz++)
      for (int cy = clo[1][iy]; cy <= chi[1][iy]; cy++) {
        int row = P.colN[0] * (cy + P.colN[1] * cz);
        for (int q = P.colOff[row + clo[0][ix]]; q < P.colOff[row + chi[0][ix] + 1]; q++) {
          int sj = P.colMem[q];
          if (self && sj == t) continue;
          double xj[3];
#pragma unroll
          for (int l = 0; l < 3; l++) xj[l] = sh[l] == 0.0 ? st[l * nP + sj] : st[l * nP + sj] + sh[l];
          double tot = ((xi[0] - xj[0]) * (xi[0] - xj[0]) + (xi[1] - xj[1]) * (xi[1] - xj[1])) +
                       (xi[2] - xj[2]) * (xi[2] - xj[2]);
          if (tot > dist2) continue;
          if (am) {
            double rx = xj[0] - xi[0], ry = xj[1] - xi[1], rz = xj[2] - xi[2];
            double rdiff = sqrt((rx * rx + ry * ry) + rz * rz);
            if (rdiff <= w) {
              double rthresh = 0.5 * (dpi + C.pc[nP + sj]);
              if (rdiff < rthresh) rx = ry = rz = rthresh;  // overlapping: the reference's substitute
              double R[3][6];
              am_resistance_pair(rx, ry, rz, alpha_local, rad, R);
              double Wj[3] = {C.w[sj], C.w[nP + sj], C.w[2 * nP + sj]};
              nn++;
#pragma unroll
              for (int k = 0; k < 3; k++) {
#pragma unroll
                for (int l = 0; l < 3; l++) {
                  Fam[k] = Fam[k] + R[k][l] * Wi[l];
                  Fam[k] = Fam[k] + R[k][l + 3] * Wj[l];
                }
                Wsum[k] = Wsum[k] + Wj[k];
              }
            }
          }
          if (fl && fabs(xi[0] - xj[0]) <= box && fabs(xi[1] - xj[1]) <= box && fabs(xi[2] - xj[2]) <= box) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
              double u = st[(3 + k) * nP + sj];
              fl_acc(S[k], Cc[k], u);
              fl_acc(S[3 + k], Cc[3 + k], u * u);
            }
            cnt++;
          }
        }
      }
  }
  if (am) {  // Step 4: ppiclf_user_AM_Briney_Binary, or the unary term times the volume
    if (nn > 0) {
      double IA, II;
      if (w / rad >= F32(3.49)) {
        IA = am_ia_analytical(w, rad, alpha_local);
        II = am_ii_analytical(w, rad, alpha_local);
      } else {
        IA = am_numerical(false, w, rad, alpha_local);
        II = am_numerical(true, w, rad, alpha_local);
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        Fam[c] = Fam[c] + IA * Wi[c];
        Fam[c] = Fam[c] + II * Wsum[c] / (double)nn;
        Fam[c] = volp * Fam[c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) Fam[c] = Fam[c] * volp;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) P.famB[c * nP + t] = Fam[c];
    P.famB[3 * nP + t] = rmass_add;
  }
  if (fl) {
#pragma unroll
    for (int k = 0; k < 6; k++) P.fmean[k * nP + t] = S[k] + Cc[k];
    P.fmean[6 * nP + t] = (double)cnt;
  }
}

// ppiclf_user.f:934-990 after every particle's forces (am_flag 2): each
// particle's W = D(rho_g u_g)/Dt - rho_g dv/dt.  It divides by the rphif the
// particle loop left behind (the last particle's) and, in vgradrhog, uses
// the relative velocity of the previous particle in ppiclF's order (the
// first particle, the last one's): k_particles' first phase recorded both by
// original index, with the pre-update velocities.
__global__ void __launch_bounds__(BLK) k_wdot(PiclDev P) {
  int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= P.nAct[0]) return;
  const size_t nP = P.nP;
  PBuf C = cur_buf(P);
  if (C.elem[t] < 0) return;
  const int last = P.lastP[0] - 1;
  int q = C.pid[t] - 1;
  while (q >= 0 && !P.actP[q]) q--;
  if (q < 0) q = last;
  const double rphif = P.rphifP[last];
  double SDrho = P.snum[t] / rphif;
  double vgradrhog = P.vrelP[q] * P.vint[17 * nP + t] + P.vrelP[nP + q] * P.vint[18 * nP + t] +
                     P.vrelP[2 * nP + q] * P.vint[19 * nP + t];
  double rhof = P.vint[t];
  const int pid = C.pid[t];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double vr = P.vrelP[c * nP + pid], ug = P.vint[(1 + c) * nP + t];
    C.w[c * nP + t] = vr * SDrho + rhof * P.vint[(10 + c) * nP + t] + ug * vgradrhog -
                      rhof * C.st[(ROW_YD + 3 + c) * nP + t];
  }
}

// ---------------------------------------------------------------------------
// COLLISIONS 3-4: hydrodynamic and undisturbed torque
// (ppiclf_user_Torque.f:72-197), Saffman and Magnus lift
// (ppiclf_user_Lift.f:61-181).  vr: the interpolated vorticity, then its
// material derivative; om: the particle's angular velocity.  Same
// operations as the oracle.
// ---------------------------------------------------------------------------
// Torque_Hydro; false when the rotational Reynolds number reaches 100, where
// ppiclF stops ("Re rotational too large")
__device__ bool torque_hydro(const double vr[6], const double om[3], double rhof, double dp, double rmu,
                             double rmass, double rhop, double t[3]) {
  const double rpi = 3.141592653589793116;  // acos(-1.0d0)
  double omgrx = 0.5 * vr[0] - om[0], omgry = 0.5 * vr[1] - om[1], omgrz = 0.5 * vr[2] - om[2];
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
    t[0] = t[1] = t[2] = 0.0;
    return false;
  }
  double Ct = Ct1 / sqrt(reyr) + Ct2 / reyr + Ct3 * reyr;
  double beta = rhop / rhof;
  double rIp = rmass * dp * dp / 10.0;
  double factor = rIp * 60.0 * Ct * omgr_mag / (64.0 * rpi * beta);
  t[0] = factor * omgrx;
  t[1] = factor * omgry;
  t[2] = factor * omgrz;
  return true;
}

// Lift_Saffman then Lift_Magnus into lift (v: relative velocity; rep
// already capped at 0.1; rnu = rmu / rhof)
__device__ void lift_forces(const double vr[6], const double om[3], const double v[3], double vmag, double dp,
                            double rep, double rmu, double rhof, double rnu, double lift[3]) {
  if (vmag < 1.0e-8) return;
  {
    double ox = vr[0], oy = vr[1], oz = vr[2];
    double omg_mag = sqrt(ox * ox + oy * oy + oz * oz);
    double omg_star = omg_mag * dp / vmag;
    double epi = sqrt(omg_star / rep);
    double d1 = 1.0 + tanh(2.5 * (log10(epi) + 0.191));
    double d2 = 0.667 + tanh(6.0 * epi - 1.92);
    double Jepi = 0.3 * d1 * d2;
    double factor = 1.615 * rmu * (dp * dp) * vmag * sqrt(omg_mag / rnu);
    double elx = v[1] * oz - v[2] * oy, ely = v[2] * ox - v[0] * oz, elz = v[0] * oy - v[1] * ox;
    double elm = sqrt(elx * elx + ely * ely + elz * elz);
    elm = fmax(1.0e-20, elm);
    double ielm = 1.0 / elm;
    lift[0] = lift[0] + factor * Jepi * elx * ielm;
    lift[1] = lift[1] + factor * Jepi * ely * ielm;
    lift[2] = lift[2] + factor * Jepi * elz * ielm;
  }
  {
    double ox = om[0], oy = om[1], oz = om[2];
    double omg_mag = sqrt(ox * ox + oy * oy + oz * oz);
    double epi = omg_mag * dp / vmag;
    double d1 = 0.675 + 0.15 * (1.0 + tanh(0.28 * (epi - 2.0)));
    double CL = 1.0 - d1 * tanh(F32(0.18) * sqrt(rep));
    double factor = 0.125 * dp * dp * dp * rhof;
    double elx = v[1] * oz - v[2] * oy, ely = v[2] * ox - v[0] * oz, elz = v[0] * oy - v[1] * ox;
    lift[0] = lift[0] + factor * CL * elx;
    lift[1] = lift[1] + factor * CL * ely;
    lift[2] = lift[2] + factor * CL * elz;
  }
}

// PICL_TEMP_Runge.F90:618-663 for COLLISIONS >= 3: the feedback force per
// cell from the previous stage's projection (the sums k_feedback forms)...
__global__ void __launch_bounds__(BLK) k_fb_cells(Dev d, PiclDev P) {
  int i = xcd_block() * BLK + threadIdx.x;
  if (i >= d.nC) return;
  double vol = d.vol[i];
#pragma unroll
  for (int k = 0; k < 3; k++) P.fb[k * (size_t)d.nC + i] = sum_nodes(P, 1 + k, i) * 0.125 * vol;
}

// ... its LSQ gradient over the cell stencil (RFLU_ComputeGradCellsWrapper,
// k_grad2's weights and order) ...
__global__ void __launch_bounds__(BLK) k_fb_grad(Dev d, PiclDev P) {
  int i = xcd_block() * BLK + threadIdx.x;
  if (i >= d.nC) return;
  const int N = d.nC;
  double r11 = d.rm[i], r12 = d.rm[N + i], r22 = d.rm[2 * N + i];
  double r13 = d.rm[3 * N + i], r23 = d.rm[4 * N + i], r33 = d.rm[5 * N + i];
  double c11 = 1.0 / r11, c22 = 1.0 / r22, c33 = 1.0 / r33;
  double c12 = -(c11 * r12);
  double c13 = -(c11 * r13 + c12 * c22 * r23);
  double c23 = -(c22 * r23);
  double c1111 = c11 * c11, c2222 = c22 * c22, c3333 = c33 * c33;
  const double xi = d.cofg[i], yi = d.cofg[N + i], zi = d.cofg[2 * N + i];
  double fi[3], g[9];
#pragma unroll
  for (int k = 0; k < 3; k++) fi[k] = P.fb[k * (size_t)N + i];
#pragma unroll
  for (int q = 0; q < 9; q++) g[q] = 0.0;
  const unsigned *mem = reinterpret_cast<const unsigned *>(d.csMem);
  for (int s = d.csOff[i]; s < d.csOff[i + 1]; s++) {
    int j = (int)(mem[s] & CS_IDX);
    double dx = d.cofg[j] - xi, dy = d.cofg[N + j] - yi, dz = d.cofg[2 * N + j] - zi;
    double term = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    dx = term * dx; dy = term * dy; dz = term * dz;
    double term1 = c1111 * (dx);
    double term2 = c2222 * (dy + c12 * dx);
    double term3 = c3333 * (dz + c23 * dy + c13 * dx);
    double wtx = term * (term1 + c12 * term2 + c13 * term3);
    double wty = term * (term2 + c23 * term3);
    double wtz = term * (term3);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double dVar = P.fb[k * (size_t)N + j] - fi[k];
      g[3 * k] = g[3 * k] + wtx * dVar;
      g[3 * k + 1] = g[3 * k + 1] + wty * dVar;
      g[3 * k + 2] = g[3 * k + 2] + wtz * dVar;
    }
  }
#pragma unroll
  for (int q = 0; q < 9; q++) P.fbRaw[9 * (size_t)i + q] = g[q];
}

// ... then WENO-XYZ per component (self first, then the stencil members) and
// RFLU_LimitGradCellsSimple with the flags (1, 2, 3): only the x force is
// VAR_INFO_POS, its gradient zeroed when an extrapolation to a face
// centroid is <= 0
__global__ void __launch_bounds__(BLK) k_fb_weno(Dev d, PiclDev P) {
  int i = xcd_block() * BLK + threadIdx.x;
  if (i >= d.nC) return;
  const int N = d.nC;
  const unsigned *mem = reinterpret_cast<const unsigned *>(d.csMem);
  // each member's 9 gradient components are read once (one 72-byte row) and
  // feed the 9 sums, each summed in the reference's member order
  double g[9], sum[9], loc[9];
#pragma unroll
  for (int q = 0; q < 9; q++) {
    double term = P.fbRaw[9 * (size_t)i + q];
    sum[q] = 1.0 / (term * term + 1.0E-15);
    loc[q] = sum[q] * term;
  }
  for (int s = d.csOff[i]; s < d.csOff[i + 1]; s++) {
    const double *row = P.fbRaw + 9 * (size_t)(mem[s] & CS_IDX);
#pragma unroll
    for (int q = 0; q < 9; q++) {
      double term = row[q];
      double w = 1.0 / (term * term + 1.0E-15);
      sum[q] = sum[q] + w;
      loc[q] = loc[q] + w * term;
    }
  }
#pragma unroll
  for (int q = 0; q < 9; q++) g[q] = loc[q] / sum[q];
  double varc = P.fb[i];
  double xc = d.cofg[i], yc = d.cofg[N + i], zc = d.cofg[2 * N + i];
  bool zero = false;
  for (int e = d.cfOff[i]; e < d.cfOff[i + 1]; e++) {
    int slot = d.cfIdx[e] >> 1;
    double dx = d.fcx[slot] - xc, dy = d.fcy[slot] - yc, dz = d.fcz[slot] - zc;
    double varf = varc + g[0] * dx + g[1] * dy + g[2] * dz;
    if (varf <= 0.0) zero = true;
  }
  if (zero) g[0] = g[1] = g[2] = 0.0;
#pragma unroll
  for (int q = 0; q < 9; q++) P.fbg[9 * (size_t)i + q] = g[q];
}

// ppiclf_user_VU_Rocflu (ppiclf_user_ViscousUnsteady.f:19-90): the history
// kernel (Mei and Adrian's, with ppiclF's volume-fraction factor) summed over
// entries 2..nTimeBH-1, half weight on the oldest; entry 1 (this stage's)
// is not used.  (A+B)**(-2) as gfortran expands it, 1/((A+B)*(A+B)).
__device__ void vu_force(const double *st, size_t nP, int t, const double *vuT, double dt, double rmu, double rhof,
                         double dp, double vmag, double rphip, double fvu[3]) {
  const double rpi = 3.141592653589793116;  // acos(-1.0d0)
  const int nT = (int)vuT[VU_N];
  if (nT <= 1) return;
  double reyL = dp * vmag * rhof / rmu;  // ppiclf_user.f:365-366
  double rnu = rmu / rhof;
  double fH = 0.75 + 0.105 * reyL;
  double factor = 3.0 * rpi * rnu * dp * dt * (1.0 + F32(2.28) * fmin(rphip, 0.5));
  double fH3 = (fH * fH) * fH, v3 = (vmag * vmag) * vmag;
  for (int iT = 2; iT <= nT; iT++) {
    double time = vuT[iT - 1];
    double A = pow(4.0 * rpi * time * rnu / (dp * dp), 0.25);
    double B = pow(0.5 * rpi * v3 * (time * time) / (0.5 * dp * rnu * fH3), 0.5);
    double kernelVU = (iT < nT ? factor : 0.5 * factor) * (1.0 / ((A + B) * (A + B)));
#pragma unroll
    for (int c = 0; c < 3; c++)
      fvu[c] = fvu[c] + kernelVU * (st[(ROW_VU + c * VU_N + iT - 1) * nP + t] -
                                    st[(ROW_VU + 3 * VU_N + c * VU_N + iT - 1) * nP + t]);
  }
}

// ppiclf_user_ShiftUnsteadyData's global part (:133-140): the entry count
// grows to VU_N and each entry's age is its younger neighbour's plus dt
__global__ void k_vu_time(Dev d, PiclDev P) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double *T = P.vuT;
  int n = (int)T[VU_N];
  if (n < VU_N) n = n + 1;
  double dt = d.scal[S_DTMIN];
  for (int iT = n; iT >= 2; iT--) T[iT - 1] = T[iT - 2] + dt;
  T[VU_N] = (double)n;
}

// ppiclf_solve_RemoveParticle's angular test (ppiclf_solve.f:2852-2881) and
// ppiclf_solve_InvokeAngularPeriodic (:1396-1500) with register = 1: a
// particle whose polar angle left [xangle, xangle + angle] is rotated about z
// by -angle when it is past 0.5 (xangle + angle), by +angle otherwise, and
// not at all (case 3) below xangle; position and velocity, each as gfortran's
// MATMUL of the 3x3 matrix with the vector (sums over the columns in order).
__device__ __forceinline__ void rot_apply(const double *R, double *v) {
  double o[3];
#pragma unroll
  for (int i = 0; i < 3; i++) o[i] = ((0.0 + R[3 * i] * v[0]) + R[3 * i + 1] * v[1]) + R[3 * i + 2] * v[2];
#pragma unroll
  for (int i = 0; i < 3; i++) v[i] = o[i];
}
__device__ __forceinline__ void angular_wrap(const PiclPhys &Ph, double yn[6]) {
  double alpha = atan2(yn[1], yn[0]);
  if (!(alpha < Ph.angX || alpha > Ph.angUp)) return;
  int c = alpha > Ph.angMid ? 1 : 0;
  if (Ph.angCase == 3 && alpha < Ph.angX) c = 2;
  rot_apply(Ph.angR[c], yn);
  rot_apply(Ph.angR[c], yn + 3);
}

// phase 0: the whole stage.  With Briney's added mass or the QS
// fluctuations the stage runs in two phases around k_pairs (whose sums need
// every particle's interpolated values): phase 1 locates and interpolates
// into vint and records what the W store reads, phase 2 takes vint and the
// pair kernel's sums and does the rest.
// PH1 = 1 instantiates phase 1 alone, without the force code's registers.
template <int TRI, int QS, int PH1 = 0>
__global__ void __launch_bounds__(BLK, 3) k_particles(Dev d, PiclDev P, PiclPhys Ph, int istage, int phaseArg) {
  const int phase = PH1 ? 1 : phaseArg;
  int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= P.nAct[0]) return;
  const size_t nP = P.nP;
  PBuf C = cur_buf(P);
  double *st = C.st;
  double v[PICL_NFLD];
  double vr[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // vorticity, D(omega)/Dt (COLLISIONS >= 3)
  if (phase == 2) {
    if (C.elem[t] < 0) return;
#pragma unroll
    for (int f = 0; f < PICL_NFLD; f++) v[f] = P.vint[f * nP + t];
    if (P.rot3)
#pragma unroll
      for (int q = 0; q < 6; q++) vr[q] = P.vint[(PICL_NFLD + q) * nP + t];
  } else {
  double xp[3] = {st[t], st[nP + t], st[2 * nP + t]};
  double rst[3];
  // a particle marked by the angular-periodic range test in the last stage
  // is gone (ppiclf_solve_RemoveParticle at the start of InitSolve)
  int e = C.elem[t] == E_REMOVED ? -1 : picl_locate(P, xp, C.elem[t], TRI ? rst : nullptr);
  C.elem[t] = e;
  count_by_runs(P.cnt, e);
  if (e < 0) {  // left every element: skipped from now on (ppiclF would remove it)
#pragma unroll
    for (int k = 0; k < 4; k++) P.ycN[k * nP + t] = 0.0;
    if (phase == 1 && P.am2) P.actP[C.pid[t]] = 0;
    return;
  }
  if (P.gauss) {  // the Gaussian projection reads the position at the start of the stage
#pragma unroll
    for (int k = 0; k < 3; k++) P.xo[k * nP + t] = xp[k];
  }
  if (TRI) {
    // gslib findpts_eval_local on the element's 2x2x2 nodal field (every node
    // holds the cell value): Lagrange weights (1-r)/2, (1+r)/2 on the nodes
    // -1, 1, contracted along r, then s, then t
    // rst: the location's Newton solution in element e
    double h[3][2];
#pragma unroll
    for (int l = 0; l < 3; l++) {
      h[l][0] = 0.5 * (1.0 - rst[l]);
      h[l][1] = 0.5 * (1.0 + rst[l]);
    }
    const double2 *fp = reinterpret_cast<const double2 *>(P.fld + PICL_NFLD * (size_t)e);
#pragma unroll
    for (int f = 0; f < PICL_NFLD / 2; f++) {
      double2 a = fp[f];
      double ua = h[0][0] * a.x + h[0][1] * a.x, ub = h[0][0] * a.y + h[0][1] * a.y;
      double wa = h[1][0] * ua + h[1][1] * ua, wb = h[1][0] * ub + h[1][1] * ub;
      v[2 * f] = h[2][0] * wa + h[2][1] * wa;
      v[2 * f + 1] = h[2][0] * wb + h[2][1] * wb;
    }
    if (P.rot3) {
      const double *rp = P.fldRot + 6 * (size_t)e;
#pragma unroll
      for (int q = 0; q < 6; q++) {
        double a = rp[q];
        double ua = h[0][0] * a + h[0][1] * a;
        double wa = h[1][0] * ua + h[1][1] * ua;
        vr[q] = h[2][0] * wa + h[2][1] * wa;
      }
    }
  } else {
  // ---- the 27 nearest element centroids, ordered by (d2, element number)
  double kd[27];
  int kr[27];
#pragma unroll
  for (int k = 0; k < 27; k++) { kd[k] = F32(1E20); kr[k] = -1; }
  int nadd = 0;
  // bins reached within the per-axis radius; the outer bins also hold every
  // centroid beyond the grid (bin indices are clamped when binning)
  int blo[3], bhi[3];
#pragma unroll
  for (int l = 0; l < 3; l++) {
    double top = (double)(P.cbN[l] - 1);
    double a = (xp[l] - P.R[l] - P.cbLo[l]) / P.cbH[l], b = (xp[l] + P.R[l] - P.cbLo[l]) / P.cbH[l];
    a = fmin(fmax(a, 0.0), top);
    b = fmin(fmax(b, 0.0), top);
    blo[l] = (int)a;
    bhi[l] = (int)b;
  }
  {
    for (int bz = blo[2]; bz <= bhi[2]; bz++)
      for (int by = blo[1]; by <= bhi[1]; by++)
        for (int bx = blo[0]; bx <= bhi[0]; bx++) {
          int bin = bx + P.cbN[0] * (by + P.cbN[1] * bz);
          for (int q = P.cbOff[bin]; q < P.cbOff[bin + 1]; q++) {
            const double2 *cp = reinterpret_cast<const double2 *>(P.cbRec + 4 * (size_t)q);
            double2 c01 = cp[0], c23 = cp[1];
            double t0 = c01.x - xp[0], t1 = c01.y - xp[1], t2 = c23.x - xp[2];
            double l0 = t0 * t0, l1 = t1 * t1, l2 = t2 * t2;
            if (l0 > P.d2Max[0] || l1 > P.d2Max[1] || l2 > P.d2Max[2]) continue;
            double d2i = 0.0;
            d2i = d2i + l0;
            d2i = d2i + l1;
            d2i = d2i + l2;
            if (!(d2i < F32(1E20))) continue;
            nadd++;
            int rc = (int)(__double_as_longlong(c23.y) >> 32);
            if (!key_less(d2i, rc, kd[26], kr[26])) continue;
            // insertion into the sorted registers, branch-free (constant indices)
            bool lt[27];
#pragma unroll
            for (int j = 0; j < 27; j++) lt[j] = key_less(d2i, rc, kd[j], kr[j]);
#pragma unroll
            for (int j = 26; j >= 1; j--) {
              double nd = lt[j - 1] ? kd[j - 1] : d2i;
              int nr = lt[j - 1] ? kr[j - 1] : rc;
              kd[j] = lt[j] ? nd : kd[j];
              kr[j] = lt[j] ? nr : kr[j];
            }
            kd[0] = lt[0] ? d2i : kd[0];
            kr[0] = lt[0] ? rc : kr[0];
          }
        }
  }
  int nn = nadd < 27 ? nadd : 27;
  if (nn < 1) {  // ppiclf_solve.f:2331-2341 stops the run here
    atomicOr(d.err, E_INTERP);
    return;
  }
#pragma unroll
  for (int f = 0; f < PICL_NFLD; f++) v[f] = 0.0;
  double wsum = 0.0;
  const double eps = F32(1.0e-12);
#pragma unroll
  for (int k = 0; k < 27; k++) {
    if (k < nn) {
      double w = 1.0 / (sqrt(kd[k]) + eps);
      const double2 *fp = reinterpret_cast<const double2 *>(P.fld + PICL_NFLD * (size_t)P.inv[kr[k]]);
#pragma unroll
      for (int f = 0; f < PICL_NFLD / 2; f++) {
        double2 a = fp[f];
        v[2 * f] = v[2 * f] + w * a.x;
        v[2 * f + 1] = v[2 * f + 1] + w * a.y;
      }
      if (P.rot3) {
        const double *rp = P.fldRot + 6 * (size_t)P.inv[kr[k]];
#pragma unroll
        for (int q = 0; q < 6; q++) vr[q] = vr[q] + w * rp[q];
      }
      wsum = wsum + w;
    }
  }
#pragma unroll
  for (int f = 0; f < PICL_NFLD; f++) v[f] = v[f] / wsum;
  if (P.rot3)
#pragma unroll
    for (int q = 0; q < 6; q++) vr[q] = vr[q] / wsum;
  }
  if (P.dbg) {
    int pid = C.pid[t];
#pragma unroll
    for (int f = 0; f < PICL_NFLD; f++) P.dbg[f * nP + pid] = v[f];
  }
  if (phase == 1) {
#pragma unroll
    for (int f = 0; f < PICL_NFLD; f++) P.vint[f * nP + t] = v[f];
    if (P.rot3)
#pragma unroll
      for (int q = 0; q < 6; q++) P.vint[(PICL_NFLD + q) * nP + t] = vr[q];
    // what ppiclf_user.f:934-990 reads: SDrho's numerator with the
    // start-of-stage velocity, the relative velocity and the rphif that
    // SetYdot's particle loop leaves (ppiclf_user.f:316-361), by index
    if (P.am2) {
      const int pid = C.pid[t];
      double y0 = st[3 * nP + t], y1 = st[4 * nP + t], y2 = st[5 * nP + t];
      P.snum[t] = v[13] + y0 * v[14] + y1 * v[15] + y2 * v[16];
      double vx = v[1] - y0, vy = v[2] - y1, vz = v[3] - y2;
      P.vrelP[pid] = vx;
      P.vrelP[nP + pid] = vy;
      P.vrelP[2 * nP + pid] = vz;
      P.rphifP[pid] = sqrt(vx * vx + vy * vy + vz * vz) <= 1.0e-3 ? 1.0 - v[9] : 1.0 - fmin(v[9], 0.62);
      P.actP[pid] = 1;
      // one address for every particle: read first, so only the few that
      // raise the maximum issue the atomic (a stale read costs one more)
      if (pid + 1 > __hip_atomic_load(P.lastP, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(P.lastP, pid + 1);
    }
    return;
  }
  }
  // slots: 0 rhof 1-3 u 4-6 dp/dx 7 c 8 T 9 phi 10-12 Du/Dt 13 rhsR 14-16 grad(rho phi) 17-19 grad rho_g

  // ---- SetYdot
  double yv[3] = {st[3 * nP + t], st[4 * nP + t], st[5 * nP + t]};
  double rhop = C.pc[t], dpp = C.pc[nP + t], volp = C.pc[2 * nP + t], spl = C.pc[3 * nP + t];
  double rmu;
  if (Ph.rmuFlag == 0) {
    rmu = Ph.rmuRef;
  } else {
    double temp = v[8];
    rmu = Ph.rmuRef * sqrt(temp / Ph.tref) * (1.0 + Ph.suth / Ph.tref) / (1.0 + Ph.suth / temp);
  }
  double rmass = volp * rhop;
  double vxr = v[1] - yv[0], vyr = v[2] - yv[1], vzr = v[3] - yv[2];
  double vmag = sqrt(vxr * vxr + vyr * vyr + vzr * vzr);
  double rhof = v[0];
  double rep = vmag * dpp * rhof / rmu;
  double rphip = v[9];
  double rmachp = vmag / v[7];
  double ydn[6], ydT = 0.0, ydO[3] = {0.0, 0.0, 0.0};
  bool skip = vmag <= 1.0e-3;  // keeps the previous ydot; ydotc stays 0
  double yc[4] = {0.0, 0.0, 0.0, 0.0};
  if (!skip) {
    rep = fmax(0.1, rep);
    rphip = fmin(rphip, 0.62);
    double rphif = 1.0 - rphip;
    double famx = 0.0, famy = 0.0, famz = 0.0, rmass_add = 0.0;
    double fdpdx = 0.0, fdpdy = 0.0, fdpdz = 0.0;
    double beta;  // ppiclf_user.f:473-485
    switch (QS) {
      case 1: beta = qs_parmar(rmachp, rphip, rep, rmu, dpp); break;
      case 2: beta = qs_osnes(rmachp, rphip, rep, rmu, dpp); break;
      case 3: beta = qs_modified_parmar(rmachp, rphip, rep, rmu, dpp); break;
      case 4: beta = qs_gidaspow(rphip, rphif, rep, rmu, dpp, rhof, vmag); break;
      default: beta = qs_sn(rep, rmu, dpp); break;
    }
    double fqsx = beta * vxr, fqsy = beta * vyr, fqsz = beta * vzr;
    if (Ph.qsFluct >= 1) {  // Step 3 (ppiclf_user.f:487-510): one SDE step of the fluctuating force
      const double rk3ark = istage == 1 ? 8.0 / 15.0 : (istage == 2 ? 5.0 / 12.0 : 0.75);
      double fac = rk3ark * d.scal[S_DTMIN];
      double m[7], F[3], U[6], fl[3], vrel[3] = {vxr, vyr, vzr};
#pragma unroll
      for (int k = 0; k < 7; k++) m[k] = P.fmean[k * nP + t];
#pragma unroll
      for (int k = 0; k < 3; k++) F[k] = C.fq[k * nP + t];
      const int call = P.nCall[0], pid = C.pid[t];
#pragma unroll
      for (int k = 0; k < 6; k++) U[k] = fluct_uniform(Ph.fluctSeed, call, pid, k);
      fluct_force(Ph.qsFluct, m, F, U, rphip, rmachp, rep, vmag, vrel, rmu, dpp, fac, fl);
      fqsx = fqsx + fl[0];
      fqsy = fqsy + fl[1];
      fqsz = fqsz + fl[2];
#pragma unroll
      for (int k = 0; k < 3; k++) C.fq[k * nP + t] = fl[k];
    } else {
      fqsx = fqsx + 0.0;
      fqsy = fqsy + 0.0;
      fqsz = fqsz + 0.0;
    }
    if (Ph.amFlag == 1) {
      double rcd_am;
      if (rmachp > F32(0.6)) {
        rcd_am = 0x1.5104d6p+1;  // 1.0 + 1.8*((0.6)**2) + 7.6*((0.6)**4) folded in single precision
      } else {
        double m2 = rmachp * rmachp;
        rcd_am = 1.0 + F32(1.8) * m2 + F32(7.6) * (m2 * m2);
      }
      rcd_am = rcd_am * 0.5;
      rcd_am = rcd_am * (1.0 + F32(0.68) * (rphip * rphip));
      rmass_add = rhof * volp * rcd_am;
      double SDrho = v[13] + yv[0] * v[14] + yv[1] * v[15] + yv[2] * v[16];
      SDrho = SDrho / rphif;
      double vgradrho = vxr * v[17] + vyr * v[18] + vzr * v[19];
      famx = rcd_am * volp * (vxr * SDrho + rhof * v[10] + v[1] * vgradrho);
      famy = rcd_am * volp * (vyr * SDrho + rhof * v[11] + v[2] * vgradrho);
      famz = rcd_am * volp * (vzr * SDrho + rhof * v[12] + v[3] * vgradrho);
    } else if (Ph.amFlag == 2) {  // Briney: unary, pair and binary terms from k_pairs
      famx = P.famB[t];
      famy = P.famB[nP + t];
      famz = P.famB[2 * nP + t];
      rmass_add = P.famB[3 * nP + t];
    }
    if (Ph.pgFlag == 1) {
      fdpdx = -volp * v[4];
      fdpdy = -volp * v[5];
      fdpdz = -volp * v[6];
      fdpdx = fdpdx + 0.0;
      fdpdy = fdpdy + 0.0;
      fdpdz = fdpdz + 0.0;
    }
    double fcv[3] = {0.0, 0.0, 0.0};  // Step 6: the collision force (ppiclf_user.f:587-600)
    if (Ph.col) {
#pragma unroll
      for (int c = 0; c < 3; c++) fcv[c] = P.fc[c * nP + t];
    }
    double fvu[3] = {0.0, 0.0, 0.0};
    if (Ph.vu) vu_force(st, nP, t, P.vuT, d.scal[S_DTMIN], rmu, rhof, dpp, vmag, rphip, fvu);
    double om[3] = {0.0, 0.0, 0.0};  // y(JOX..JOZ) at the start of the stage
    if (P.rot)
#pragma unroll
      for (int c = 0; c < 3; c++) om[c] = st[(ROW_O + c) * nP + t];
    double lift[3] = {0.0, 0.0, 0.0};  // Step 9b (ppiclf_user.f:640-646)
    if (Ph.col == 4) {
      double vrel[3] = {vxr, vyr, vzr};
      lift_forces(vr, om, vrel, vmag, dpp, rep, rmu, rhof, rmu / rhof, lift);
    }
    ydn[0] = yv[0];
    ydn[1] = yv[1];
    ydn[2] = yv[2];
    ydn[3] = (fqsx + famx + fdpdx + fvu[0] + lift[0] + fcv[0]) / (rmass + rmass_add);
    ydn[4] = (fqsy + famy + fdpdy + fvu[1] + lift[1] + fcv[1]) / (rmass + rmass_add);
    ydn[5] = (fqsz + famz + fdpdz + fvu[2] + lift[2] + fcv[2]) / (rmass + rmass_add);
    if (Ph.vu) {  // ppiclf_user_UpdatePlag (ppiclf_user_ViscousUnsteady.f:146-206)
      double SDrho = v[13] + yv[0] * v[14] + yv[1] * v[15] + yv[2] * v[16];
      SDrho = SDrho / rphif;
      double vgradrho = vxr * v[17] + vyr * v[18] + vzr * v[19];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        st[(ROW_VU + c * VU_N) * nP + t] = v[1 + c] * (SDrho + vgradrho) + rhof * v[10 + c];
        st[(ROW_VU + 3 * VU_N + c * VU_N) * nP + t] = yv[c] * SDrho + rhof * ydn[3 + c];
      }
    }
    double qq = 0.0;
    if (Ph.heatTransfer) {
      double rkappa = 1004.64 * rmu / 0.70;  // rcp_fluid*rmu/rpr (ppiclf_user.f:310)
      qq = heat_transfer(Ph.heatTransfer, rkappa, dpp, v[8], st[ROW_T * nP + t], rep, rphif, rmachp);
      ydT = qq / (rmass * Ph.rcpPart);
      st[ROW_TYD * nP + t] = ydT;
    }
    double tauh[3] = {0.0, 0.0, 0.0};  // Torque_Hydro
    if (P.rot) {  // Step 9a (:638-646): ydot(JOX..JOZ) = tau / (rmass dp^2 / 10)
      double rmass_omega = rmass * dpp * dpp / 10.0;
      double tauu[3] = {0.0, 0.0, 0.0};
      if (P.rot3) {  // Torque_driver (ppiclf_user_Torque.f:23-62)
        if (!torque_hydro(vr, om, rhof, dpp, rmu, rmass, rhop, tauh)) atomicOr(d.err, E_REROT);
        double rIf = rhof * dpp * dpp * volp / 10.0;  // Torque_Undisturbed
#pragma unroll
        for (int c = 0; c < 3; c++) tauu[c] = 0.5 * rIf * vr[3 + c];
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        ydO[c] = (P.tq[c * nP + t] + tauh[c] + tauu[c]) / rmass_omega;
        st[(ROW_OYD + c) * nP + t] = ydO[c];
      }
    }
    if (Ph.feedbackFlag == 1) {
      yc[0] = spl * (ydn[3] * rmass - fcv[0]);
      yc[1] = spl * (ydn[4] * rmass - fcv[1]);
      yc[2] = spl * (ydn[5] * rmass - fcv[2]);
      yc[3] = spl * ((fqsx + fvu[0] + lift[0]) * yv[0] + (fqsy + fvu[1] + lift[1]) * yv[1] +
                     (fqsz + fvu[2] + lift[2]) * yv[2] + famx * v[1] + famy * v[2] + famz * v[3] + tauh[0] * om[0] +
                     tauh[1] * om[1] + tauh[2] * om[2] + qq);
    } else {  // ydotc keeps what the collisions added into it
#pragma unroll
      for (int c = 0; c < 3; c++) yc[c] = fcv[c];
    }
#pragma unroll
    for (int k = 0; k < 6; k++) st[(ROW_YD + k) * nP + t] = ydn[k];
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) ydn[k] = st[(ROW_YD + k) * nP + t];
    if (Ph.heatTransfer) ydT = st[ROW_TYD * nP + t];
    if (P.rot) {
#pragma unroll
      for (int c = 0; c < 3; c++) ydO[c] = st[(ROW_OYD + c) * nP + t];
    }
  }
#pragma unroll
  for (int k = 0; k < 4; k++) P.ycN[k * nP + t] = yc[k];
  // ppiclf_user_ShiftUnsteadyData after the third stage's SetYdot
  // (ppiclf_user.f:1030-1036): every entry moves one older
  if (Ph.vu && istage == 3) {
    for (int c = 0; c < 6; c++) {
      double *hrow = st + (size_t)(ROW_VU + c * VU_N) * nP + t;
#pragma unroll
      for (int k = VU_N - 1; k >= 1; k--) hrow[k * nP] = hrow[(k - 1) * nP];
    }
  }

  // ---- RK3 (Rocflu coefficients), y = -c1*y1 + y + c3*ydot, y1 = ydot
  double dt = d.scal[S_DTMIN];
  double k0, k2;
  if (istage == 1) { k0 = 0.0; k2 = dt * 8.0 / 15.0; }
  else if (istage == 2) { k0 = dt * 17.0 / 60.0; k2 = dt * 5.0 / 12.0; }
  else { k0 = dt * 5.0 / 12.0; k2 = dt * 3.0 / 4.0; }
  double yn[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    double y1 = istage == 1 ? 0.0 : st[(ROW_Y1 + k) * nP + t];
    double y = st[(ROW_X + k) * nP + t];
    yn[k] = -(k0 * y1) + 1.0 * y + k2 * ydn[k];
  }
  // ppiclf_solve_RemoveParticle after the update (ppiclf_solve.f:1933): the
  // angular wrap, then InvokeLinearPeriodic (:1319-1394)
  if (Ph.ang) {
    angular_wrap(Ph, yn);
    if (Ph.per[2]) {  // the second case: the x range, then the z wrap
      if (yn[0] < Ph.angXd[0] || yn[0] > Ph.angXd[1])
        P.rmMark[t] = 1;
      else if (yn[2] < Ph.perLo[2])
        yn[2] = Ph.perHi[2] - fabs(Ph.perLo[2] - yn[2]);
      else if (yn[2] > Ph.perHi[2])
        yn[2] = Ph.perLo[2] + fabs(yn[2] - Ph.perHi[2]);
    }
  } else {
    // beyond one bound, re-enter as far inside the other, axis by axis
#pragma unroll
    for (int k = 0; k < 3; k++)
      if (Ph.per[k]) {
        if (yn[k] < Ph.perLo[k])
          yn[k] = Ph.perHi[k] - fabs(Ph.perLo[k] - yn[k]);
        else if (yn[k] > Ph.perHi[k])
          yn[k] = Ph.perLo[k] + fabs(yn[k] - Ph.perHi[k]);
      }
  }
#pragma unroll
  for (int k = 0; k < 6; k++) {
    st[(ROW_X + k) * nP + t] = yn[k];
    st[(ROW_Y1 + k) * nP + t] = ydn[k];
  }
  if (Ph.heatTransfer) {
    double y1 = istage == 1 ? 0.0 : st[ROW_TY1 * nP + t];
    double y = st[ROW_T * nP + t];
    st[ROW_T * nP + t] = -(k0 * y1) + 1.0 * y + k2 * ydT;
    st[ROW_TY1 * nP + t] = ydT;
  }
  if (P.rot) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double y1 = istage == 1 ? 0.0 : st[(ROW_OY1 + c) * nP + t];
      double y = st[(ROW_O + c) * nP + t];
      st[(ROW_O + c) * nP + t] = -(k0 * y1) + 1.0 * y + k2 * ydO[c];
      st[(ROW_OY1 + c) * nP + t] = ydO[c];
    }
  }
}

// Locate only (the set-up's InitSolve: no fields, no forces).
__global__ void __launch_bounds__(BLK) k_locate_init(PiclDev P) {
  int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= P.nAct[0]) return;
  const size_t nP = P.nP;
  PBuf C = cur_buf(P);
  double xp[3] = {C.st[t], C.st[nP + t], C.st[2 * nP + t]};
  int e = picl_locate(P, xp);
  C.elem[t] = e;
  if (e >= 0) atomicAdd(&P.cnt[e], 1);
  else {
#pragma unroll
    for (int k = 0; k < 4; k++) P.ycN[k * nP + t] = 0.0;
  }
}

__global__ void __launch_bounds__(BLK) k_locate_pts(PiclDev P, int n, const double *xyz, int *out) {
  int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= n) return;
  double xp[3] = {xyz[3 * (size_t)t], xyz[3 * (size_t)t + 1], xyz[3 * (size_t)t + 2]};
  int e = picl_locate(P, xp);
  out[t] = e < 0 ? -1 : P.refId[e];
}

__global__ void __launch_bounds__(BLK) k_zero_int(int *a, int n) {
  int t = blockIdx.x * BLK + threadIdx.x;
  if (t < n) a[t] = 0;
}

// ---------------------------------------------------------------------------
// Exclusive scan of the per-element counts (cnt[0..n], cnt[n] = 0) into off;
// the counts are cleared for use as fill cursors.  4096 entries per block.
// ---------------------------------------------------------------------------
#define SCAN_T 1024
#define SCAN_V 4
__global__ void __launch_bounds__(SCAN_T) k_scan1(const int *cnt, int n, int *bsum) {
  __shared__ int sm[SCAN_T / 64];
  size_t base = (size_t)blockIdx.x * SCAN_T * SCAN_V + (size_t)threadIdx.x * SCAN_V;
  int s = 0;
  for (int k = 0; k < SCAN_V; k++)
    if (base + k < (size_t)n) s += cnt[base + k];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int q = 0; q < SCAN_T / 64; q++) tot += sm[q];
    bsum[blockIdx.x] = tot;
  }
}
__global__ void __launch_bounds__(SCAN_T) k_scan2(int *bsum, int nb) {
  __shared__ int sm[SCAN_T];
  int carry = 0;
  for (int c0 = 0; c0 < nb; c0 += SCAN_T) {
    int i = c0 + threadIdx.x;
    int v = i < nb ? bsum[i] : 0;
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < SCAN_T; o <<= 1) {
      int a = threadIdx.x >= o ? sm[threadIdx.x - o] : 0;
      __syncthreads();
      sm[threadIdx.x] += a;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + sm[threadIdx.x] - v;
    int tot = sm[SCAN_T - 1];
    __syncthreads();
    carry += tot;
  }
}
__global__ void __launch_bounds__(SCAN_T) k_scan3(int *cnt, int n, const int *bsum, int *off) {
  __shared__ int sm[SCAN_T];
  size_t base = (size_t)blockIdx.x * SCAN_T * SCAN_V + (size_t)threadIdx.x * SCAN_V;
  int v[SCAN_V], s = 0;
  for (int k = 0; k < SCAN_V; k++) {
    v[k] = base + k < (size_t)n ? cnt[base + k] : 0;
    s += v[k];
  }
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < SCAN_T; o <<= 1) {
    int a = threadIdx.x >= o ? sm[threadIdx.x - o] : 0;
    __syncthreads();
    sm[threadIdx.x] += a;
    __syncthreads();
  }
  int run = bsum[blockIdx.x] + sm[threadIdx.x] - s;
  for (int k = 0; k < SCAN_V; k++)
    if (base + k < (size_t)n) {
      off[base + k] = run;
      run += v[k];
      cnt[base + k] = 0;
    }
}

// Particles into their element's new positions (arbitrary order within an
// element; k_deposit sorts each segment); particles that left every element
// go after the active ones.
// The particles are mostly in element order already (the last stage's
// sort), so a wave's lanes form a few runs of one element each: one atomic
// per run reserves the run's slots.  The order inside a segment is fixed
// afterwards by k_rank, so the reservation order does not matter.
__global__ void __launch_bounds__(BLK) k_fill(PiclDev P) {
  const int t = blockIdx.x * BLK + threadIdx.x;
  const int nA = P.nAct[0];
  PBuf C = cur_buf(P);
  const int e = t < nA ? C.elem[t] : -2;
  if (e == -1) P.dest[t] = P.off[P.nC] + atomicAdd(P.lostCur, 1);
  const int slot = reserve_by_runs(P.cnt, e);
  if (e >= 0) P.src[P.off[e] + slot] = t;
}

// ---------------------------------------------------------------------------
// ProjectParticleGrid, single-element box mode (ppiclf_solve.f:3285-3447):
// each element sums map * (1/V_hex) of its particles in ascending particle
// index; the map (ppiclf_user_MapProjPart.f:37-66) is built from the previous
// stage's ydotc (ppiclf_comm.f:1190-1217 runs before SetYdot).  The segment,
// sorted by original index, fixes the particles' positions in the next order.
// ---------------------------------------------------------------------------
// Gaussian mode (P.gauss): the sorted segments instead receive one record per
// particle, position at the start of the stage and map * multfci, read by
// k_gauss.  posXo = 0 reads the position from the particle rows (set-up).
// Each particle's place in its element's segment: its rank by original index
// among the segment's particles (ppiclF's order within an element, which the
// box sums follow).  One thread per segment slot; the segment's threads read
// its (short) list together.
__global__ void __launch_bounds__(BLK) k_rank(PiclDev P) {
  int q = blockIdx.x * BLK + threadIdx.x;
  if (q >= P.off[P.nC]) return;
  PBuf C = cur_buf(P);
  const int t = P.src[q], e = C.elem[t], b = P.off[e], n = P.off[e + 1] - b, my = C.pid[t];
  int rank = 0;
  for (int j = 0; j < n; j++) rank += C.pid[P.src[b + j]] < my ? 1 : 0;
  P.srt[b + rank] = t;
  P.dest[t] = b + rank;
}

// The Gaussian filter's particle records, in the sorted order: position at
// the start of the stage and the five deposited values times multfci.
__global__ void __launch_bounds__(BLK) k_gq(PiclDev P, int posXo) {
  int q = blockIdx.x * BLK + threadIdx.x;
  if (q >= P.off[P.nC]) return;
  const size_t nP = P.nP;
  PBuf C = cur_buf(P);
  const int t = P.srt[q];
  const double spl = C.pc[3 * nP + t];
  double map[5];
  map[0] = C.pc[2 * nP + t] * spl;
#pragma unroll
  for (int k = 0; k < 4; k++) map[1 + k] = C.st[(ROW_YCP + k) * nP + t] * spl;
  const double *xs = posXo ? P.xo : C.st;
  // the position record is what the deposit's candidate tests read: 32
  // bytes, four to a 128-byte line
  double2 *g = reinterpret_cast<double2 *>(P.gq + 4 * (size_t)q);
  double2 *gm = reinterpret_cast<double2 *>(P.gq + 4 * (nP + (size_t)q));
  g[0] = make_double2(xs[t], xs[nP + t]);
  g[1] = make_double2(xs[2 * nP + t], map[0] * P.gMult);
  gm[0] = make_double2(map[1] * P.gMult, map[2] * P.gMult);
  gm[1] = make_double2(map[3] * P.gMult, map[4] * P.gMult);
}

// The box filter (ppiclf_solve.f:3285-3447 with the single-element box): the
// element's particles summed in the sorted order.
__global__ void __launch_bounds__(BLK) k_deposit_box(PiclDev P) {
  int e = blockIdx.x * BLK + threadIdx.x;
  if (e >= P.nC) return;
  const size_t nP = P.nP;
  PBuf C = cur_buf(P);
  const int b = P.off[e], n = P.off[e + 1] - b;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const double rexp = P.rexp[e];
  for (int q = 0; q < n; q++) {
    const int t = P.srt[b + q];
    const double spl = C.pc[3 * nP + t];
    double map[5];
    map[0] = C.pc[2 * nP + t] * spl;
#pragma unroll
    for (int k = 0; k < 4; k++) map[1 + k] = C.st[(ROW_YCP + k) * nP + t] * spl;
#pragma unroll
    for (int k = 0; k < 5; k++) acc[k] = acc[k] + (map[k] * 1.0) * rexp;
  }
#pragma unroll
  for (int k = 0; k < 5; k++) P.proj[k * (size_t)P.nC + e] = 0.0 + acc[k];
}

// ---------------------------------------------------------------------------
// ProjectParticleGrid with the Gaussian filter (ppiclf_solve.f:3103-3109,
// 3196-3246, iwallm = 0): every mesh vertex sums map * multfci *
// exp(-r^2 / (2 sigma^2)) over the particles within the cut-off.  One thread
// per vertex gathers from the elements whose centroid lies within cut-off +
// element radius (centroid bins; bins without particles are skipped), each
// element's particles in ascending original index.  The sums are compensated
// (Neumaier), which makes them independent of the gather order up to the
// last bit in all but contrived cases; ppiclF's own order over its bins and
// ghost particles is not restated.
__device__ __forceinline__ void neumaier(double &s, double &c, double x) {
  double t = s + x;
  if (fabs(s) >= fabs(x))
    c = c + ((s - t) + x);
  else
    c = c + ((x - t) + s);
  s = t;
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_max_int(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, 64));
  return x;
}
__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__global__ void k_gbox_reset(PiclDev P) {
  int t = threadIdx.x;
  if (t < 6) P.gBox[t] = t < 3 ? 0x7fffffff : -1;
  if (t == 6) P.gvList[P.nV] = 0;
}

// The centroid-bin range a vertex's image reaches, clipped to the bins that
// hold particles; false when it is empty.
__device__ __forceinline__ bool gauss_bins(const PiclDev &P, const double xv[3], int im, double sh[3], int blo[3],
                                           int bhi[3]) {
  bool skip = false;
#pragma unroll
  for (int l = 0, r = im; l < 3; l++, r /= 3) {
    int sd = P.gImg ? r % 3 - 1 : 0;
    if (sd != 0 && P.gShift[l] == 0.0) skip = true;
    sh[l] = sd == 0 ? 0.0 : (double)sd * P.gShift[l];
    double top = (double)(P.cbN[l] - 1);
    double xs = xv[l] - sh[l];
    double a = (xs - P.gReach - P.cbLo[l]) / P.cbH[l], b = (xs + P.gReach - P.cbLo[l]) / P.cbH[l];
    if (b < 0.0 || a > P.cbTop[l]) skip = true;  // no centroid within reach
    blo[l] = max((int)fmin(fmax(a, 0.0), top), P.gBox[l]);
    bhi[l] = min((int)fmin(fmax(b, 0.0), top), P.gBox[3 + l]);
    if (blo[l] > bhi[l]) skip = true;
  }
  return !skip;
}

// The vertices the deposit can reach this stage, appended to gvList (one
// atomic per wave); every other vertex's values are zero.  The list's order
// does not matter: each vertex's sum is computed on its own.
__global__ void __launch_bounds__(BLK) k_gvsel(PiclDev P) {
  const int t = blockIdx.x * BLK + threadIdx.x;
  const int v = t < P.nV ? P.vperm[t] : 0;  // Morton order: the list's neighbours are close in space
  bool sel = false;
  const size_t nV = P.nV;
  if (t < P.nV) {
    const double xv[3] = {P.vX[v], P.vX[nV + v], P.vX[2 * nV + v]};
    const int nImg = P.gImg ? 27 : 1;
    for (int im = 0; im < nImg && !sel; im++) {
      double sh[3];
      int blo[3], bhi[3];
      sel = gauss_bins(P, xv, im, sh, blo, bhi);
    }
    if (!sel) {
#pragma unroll
      for (int k = 0; k < 5; k++) P.vproj[k * nV + v] = 0.0;
    }
  }
  const unsigned long long m = __ballot(sel);
  int base = 0;
  if ((threadIdx.x & 63) == 0 && m) base = atomicAdd(&P.gvList[P.nV], __popcll(m));
  base = __shfl(base, 0, 64);
  if (sel) P.gvList[base + lanes_below(m)] = v;
}

// Particles per centroid bin, and the box of bins that hold any (one atomic
// per wave and bound).
__global__ void __launch_bounds__(BLK) k_gcount(PiclDev P) {
  int e = blockIdx.x * BLK + threadIdx.x;
  int n = e < P.nC ? P.off[e + 1] - P.off[e] : 0;
  int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {-1, -1, -1};
  if (n) {
    int bin = P.ecb[e];
    atomicAdd(&P.binCnt[bin], n);
    int bx = bin % P.cbN[0], by = (bin / P.cbN[0]) % P.cbN[1], bz = bin / (P.cbN[0] * P.cbN[1]);
    lo[0] = hi[0] = bx;
    lo[1] = hi[1] = by;
    lo[2] = hi[2] = bz;
  }
#pragma unroll
  for (int l = 0; l < 3; l++)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lo[l] = min(lo[l], __shfl_xor(lo[l], o, 64));
      hi[l] = max(hi[l], __shfl_xor(hi[l], o, 64));
    }
  if ((threadIdx.x & 63) == 0 && hi[0] >= 0) {
    // a plain read first: once the bounds are nearly final, few waves still
    // tighten them, and the others issue no atomic on these six words
#pragma unroll
    for (int l = 0; l < 3; l++) {
      if (lo[l] < __atomic_load_n(&P.gBox[l], __ATOMIC_RELAXED)) atomicMin(&P.gBox[l], lo[l]);
      if (hi[l] > __atomic_load_n(&P.gBox[3 + l], __ATOMIC_RELAXED)) atomicMax(&P.gBox[3 + l], hi[l]);
    }
  }
}

// One wave per vertex.  The wave scans the vertex's centroid bins 64 at a
// time; elements within reach that hold particles are appended to a list in
// LDS by ballot, in (bin, record) order, so the list's order is fixed.  The
// 64 lanes then stride the list's particles flattened (lane l takes the l-th,
// (l+64)-th, ... particle), so every loop of the wave is uniform.  Each lane
// keeps a compensated sum; the lanes' sums are combined pairwise in lane
// order at the end, so the result does not depend on the schedule.
#define GW_CAP 256
#define GW_WAVES (BLK / 64)
#ifndef GW_BLOCKS
#define GW_BLOCKS 8192  // a multiple of 8: 32 workgroups per CU
#endif
#ifndef GW_OCC
#define GW_OCC 4  // waves per SIMD the register budget allows
#endif

// Compensated accumulation, branch-free: Knuth's TwoSum gives the exact
// rounding error of s + x whatever their magnitudes (the same error the
// Neumaier branch selects), in six additions.
__device__ __forceinline__ void two_sum_acc(double &s, double &c, double x) {
  double t = s + x;
  double z = t - s;
  double e = (s - (t - z)) + (x - z);
  c = c + e;
  s = t;
}

// One batch of queued (particle slot, r^2) pairs within the cut-off: lane l
// takes entry l (l < n).
__device__ __forceinline__ void gauss_batch(const PiclDev &P, const int *Qp, const double *Qr, int n, int lane,
                                            const double *T, double acc[5], double cmp[5]) {
  if (lane < n) {
    const size_t at = Qp[lane];
    const double2 *g = reinterpret_cast<const double2 *>(P.gq + 4 * at);
    const double2 *gm = reinterpret_cast<const double2 *>(P.gq + 4 * ((size_t)P.nP + at));
    double rexp = gx::exp_tab(Qr[lane] * P.gRdum, T);  // the argument lies in [ln alpha, 0]
    double2 g1 = g[1], g2 = gm[0], g3 = gm[1];
    two_sum_acc(acc[0], cmp[0], g1.y * rexp);
    two_sum_acc(acc[1], cmp[1], g2.x * rexp);
    two_sum_acc(acc[2], cmp[2], g2.y * rexp);
    two_sum_acc(acc[3], cmp[3], g3.x * rexp);
    two_sum_acc(acc[4], cmp[4], g3.y * rexp);
  }
}

// The particles of the listed elements (E[j] = the element's first particle
// slot in gq, C[j] = its particle count), two elements at a time as one
// sequence over the 64 lanes (elements hold about 26 particles, so a pair
// mostly fits one pass).  Those within the cut-off go to the queue (Qp, Qr;
// *qn entries, fewer than 64 between calls) in lane order, and every 64 of
// them are summed as one batch.  The chunks (element pair k, k + 1, sequence
// positions pb to pb + 63) are walked in order with the next chunk's
// positions loaded before the current one is tested, so a gather's latency
// overlaps a test.
__device__ __forceinline__ void gauss_flush(const PiclDev &P, const int *E, const int *C, int n, int lane,
                                            const double xv[3], const double sh[3], int *Qp, double *Qr, int &qn,
                                            const double *T, double acc[5], double cmp[5]) {
  int k = 0, pb = 0;
  // the pair's counts and first slots (the same for every lane)
  int c0 = C[0], e0 = E[0], c1 = 1 < n ? C[1] : 0, e1 = 1 < n ? E[1] : 0;
  // the first chunk's positions: lane l holds sequence position l
  bool nin = lane < c0 + c1;
  int npos = lane < c0 ? e0 + lane : e1 + (lane - c0);
  double2 nxy = make_double2(0.0, 0.0);
  double nz = 0.0;
  if (nin) {
    nxy = reinterpret_cast<const double2 *>(P.gq + 4 * (size_t)npos)[0];
    nz = P.gq[4 * (size_t)npos + 2];
  }
  while (true) {
    const bool in = nin;
    const int pos = npos;
    const double2 gxy = nxy;
    const double gz = nz;
    // advance to the next chunk (uniform over the wave) and start its loads
    pb += 64;
    if (pb >= c0 + c1) {
      k += 2;
      pb = 0;
      if (k < n) {
        c0 = C[k];
        e0 = E[k];
        c1 = k + 1 < n ? C[k + 1] : 0;
        e1 = k + 1 < n ? E[k + 1] : 0;
      }
    }
    const bool more = k < n;
    if (more) {
      const int q = pb + lane;
      nin = q < c0 + c1;
      npos = q < c0 ? e0 + q : e1 + (q - c0);
      if (nin) {
        nxy = reinterpret_cast<const double2 *>(P.gq + 4 * (size_t)npos)[0];
        nz = P.gq[4 * (size_t)npos + 2];
      }
    }
    bool take = false;
    double rdist2 = 0.0;
    if (in) {
      // the image's coordinates as ppiclF's ghost holds them, x_p + s
      // (x_p + 0 is x_p, or +0 for -0, which gives the same difference)
      double dx = xv[0] - (gxy.x + sh[0]);
      double dy = xv[1] - (gxy.y + sh[1]);
      double dz = xv[2] - (gz + sh[2]);
      rdist2 = dx * dx + dy * dy;
      rdist2 = rdist2 + dz * dz;
      take = rdist2 <= P.gD2chk;
    }
    const unsigned long long m = __ballot(take);
    if (take) {
      int at = qn + lanes_below(m);
      Qp[at] = pos;
      Qr[at] = rdist2;
    }
    qn += __popcll(m);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (qn >= 64) {
      gauss_batch(P, Qp, Qr, 64, lane, T, acc, cmp);
      // the remainder moves to the front
      int rest = qn - 64, mp = lane < rest ? Qp[64 + lane] : 0;
      double mr = lane < rest ? Qr[64 + lane] : 0.0;
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (lane < rest) {
        Qp[lane] = mp;
        Qr[lane] = mr;
      }
      qn = rest;
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (!more) break;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

__device__ __forceinline__ void gauss_vertex(const PiclDev &P, int v, int lane, int *E, int *C, int *Qp, double *Qr,
                                             const double *T) {
  const size_t nV = P.nV;
  const double xv[3] = {P.vX[v], P.vX[nV + v], P.vX[2 * nV + v]};
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, cmp[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int qn = 0;  // queued pairs of the current image
  // image im: shift (im % 3 - 1, im / 3 % 3 - 1, im / 9 - 1) periods; a
  // particle image x_p + s counts at the vertex when x_p lies within reach
  // of xv - s
  const int nImg = P.gImg ? 27 : 1;
  for (int im = 0; im < nImg; im++) {
    double sh[3];
    int blo[3], bhi[3];
    if (!gauss_bins(P, xv, im, sh, blo, bhi)) continue;
    const int nbx = bhi[0] - blo[0] + 1, nby = bhi[1] - blo[1] + 1;
    // lanes over (x, y) of one z layer: rows of nbx bins, rpr rows per round
    // (chunks of 64 bins of a row when a row is longer)
    const int xstep = nbx <= 64 ? nbx : 64, rpr = nbx <= 64 ? 64 / nbx : 1;
    const int lx = nbx <= 64 ? lane % nbx : lane, lr = nbx <= 64 ? lane / nbx : 0;
    int nList = 0;
    for (int bz = blo[2]; bz <= bhi[2]; bz++)
    for (int yb = 0; yb < nby; yb += rpr)
    for (int xb = 0; xb < nbx; xb += xstep) {
      int q0 = 0, nrec = 0;
      int bx = blo[0] + xb + lx, by = blo[1] + yb + lr;
      if (lr < rpr && bx <= bhi[0] && by <= bhi[1]) {
        int bin = bx + P.cbN[0] * (by + P.cbN[1] * bz);
        if (P.binCnt[bin]) {
          q0 = P.cbOff[bin];
          nrec = P.cbOff[bin + 1] - q0;
        }
      }
      const int maxrec = wave_max_int(nrec);
      for (int r = 0; r < maxrec; r++) {
        bool take = false;
        int first = 0, cnt = 0;
        if (r < nrec) {
          const double2 *cp = reinterpret_cast<const double2 *>(P.cbRec + 4 * (size_t)(q0 + r));
          double2 c01 = cp[0], c23 = cp[1];
          // the element's box of particle positions against the cut-off
          double t0 = fmax(fabs(c01.x - (xv[0] - sh[0])) - P.gHalf[0], 0.0);
          double t1 = fmax(fabs(c01.y - (xv[1] - sh[1])) - P.gHalf[1], 0.0);
          double t2 = fmax(fabs(c23.x - (xv[2] - sh[2])) - P.gHalf[2], 0.0);
          if (t0 * t0 + t1 * t1 + t2 * t2 <= P.gCut2) {
            int e = (int)(unsigned)(__double_as_longlong(c23.y) & 0xffffffffll);
            first = P.off[e];
            cnt = P.off[e + 1] - first;
            take = cnt > 0;
          }
        }
        const unsigned long long m = __ballot(take);
        const int nm = __popcll(m);
        if (nList + nm > GW_CAP) {
          gauss_flush(P, E, C, nList, lane, xv, sh, Qp, Qr, qn, T, acc, cmp);
          nList = 0;
        }
        if (take) {
          int at = nList + lanes_below(m);
          E[at] = first;
          C[at] = cnt;
        }
        nList += nm;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
    }
    if (nList) gauss_flush(P, E, C, nList, lane, xv, sh, Qp, Qr, qn, T, acc, cmp);
    gauss_batch(P, Qp, Qr, qn, lane, T, acc, cmp);  // what is left of this image's queue
    qn = 0;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  // combine the lanes' (sum, compensation) pairs, lower lane first
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    bool upper = (lane & off) != 0;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      double so = __shfl_xor(acc[k], off, 64), co = __shfl_xor(cmp[k], off, 64);
      double a = upper ? so : acc[k], b = upper ? acc[k] : so;
      double ca = upper ? co : cmp[k], cb = upper ? cmp[k] : co;
      double s2 = a + b, bb = s2 - a;
      double e = (a - (s2 - bb)) + (b - bb);
      acc[k] = s2;
      cmp[k] = (ca + cb) + e;
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 5; k++) P.vproj[k * nV + v] = 0.0 + (acc[k] + cmp[k]);
  }
}

__global__ void __launch_bounds__(BLK, GW_OCC) k_gauss_w(PiclDev P) {
  __shared__ double sT[32];
  __shared__ int sE[GW_WAVES][GW_CAP];
  __shared__ int sC[GW_WAVES][GW_CAP];
  __shared__ int sQp[GW_WAVES][128];
  __shared__ double sQr[GW_WAVES][128];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int *E = sE[w], *C = sC[w], *Qp = sQp[w];
  double *Qr = sQr[w];
  if (threadIdx.x < 32) sT[threadIdx.x] = gx::kTab[threadIdx.x];
  __syncthreads();
  const size_t nV = P.nV;
  // the listed vertices are split into 8 contiguous ranges, one per XCD
  // (hardware block b runs on XCD b % 8), walked by that XCD's waves
  const int nList = P.gvList[P.nV];
  const int xcd = blockIdx.x % 8, wavesX = (gridDim.x / 8) * GW_WAVES;
  const int lo = (int)((long long)nList * xcd / 8), hi = (int)((long long)nList * (xcd + 1) / 8);
  if (blockIdx.x >= (gridDim.x / 8) * 8) return;  // blocks beyond a multiple of 8 have no range
  for (int vt = lo + (int)(blockIdx.x / 8) * GW_WAVES + w; vt < hi; vt += wavesX) {
    const int v = P.gvList[vt];
    gauss_vertex(P, v, lane, E, C, Qp, Qr, sT);
  }
}


// Every particle to its new position in the other copy; this stage's ydotc
// becomes the next map.  Particles lost earlier keep their positions.
__global__ void __launch_bounds__(BLK) k_permute(PiclDev P) {
  int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= P.nP) return;
  const size_t nP = P.nP;
  int par = P.par[0];
  PBuf C = par ? P.buf[1] : P.buf[0], X = par ? P.buf[0] : P.buf[1];
  int nA = P.nAct[0];
  int pos = t < nA ? P.dest[t] : t;
#pragma unroll
  for (int r = 0; r < ROW_YCP; r++) X.st[r * nP + pos] = C.st[r * nP + t];
#pragma unroll
  for (int k = 0; k < 4; k++) X.st[(ROW_YCP + k) * nP + pos] = t < nA ? P.ycN[k * nP + t] : C.st[(ROW_YCP + k) * nP + t];
#pragma unroll
  for (int k = 0; k < 4; k++) X.pc[k * nP + pos] = C.pc[k * nP + t];
  if (P.ht)
    for (int r = ROW_T; r < NROW; r++) X.st[r * nP + pos] = C.st[r * nP + t];
  if (P.rot)
    for (int r = ROW_O; r < ROW_VU; r++) X.st[r * nP + pos] = C.st[r * nP + t];
  if (P.vu)
    for (int r = ROW_VU; r < NROW_VU; r++) X.st[r * nP + pos] = C.st[r * nP + t];
  if (P.am2)
#pragma unroll
    for (int c = 0; c < 3; c++) X.w[c * nP + pos] = C.w[c * nP + t];
  if (P.fl)
#pragma unroll
    for (int c = 0; c < 3; c++) X.fq[c * nP + pos] = C.fq[c * nP + t];
  X.pid[pos] = C.pid[t];
  int el = C.elem[t];
  if (P.rmMark && t < nA && P.rmMark[t]) {  // removed by the angular-periodic range test
    el = E_REMOVED;
    P.rmMark[t] = 0;
  }
  X.elem[pos] = el;
}

// the SetYdot call counter that keys the fluctuations' uniform draws
__global__ void k_next_call(PiclDev P) {
  if (threadIdx.x == 0 && blockIdx.x == 0) P.nCall[0] = P.nCall[0] + 1;
}

__global__ void k_flip(PiclDev P) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    P.par[0] = 1 - P.par[0];
    P.nAct[0] = P.off[P.nC];
  }
}

// Rocflu's read-back (PICL_TEMP_Runge.F90:926-1038, 1217-1234): feedback into
// rhs, and piclVF from the projected volume.
__global__ void __launch_bounds__(BLK) k_feedback(Dev d, PiclDev P, int feedback) {
  int i = xcd_block() * BLK + threadIdx.x;
  if (i >= d.nC) return;
  int N = d.nC;
  double vol = d.vol[i];
  if (feedback) {
#pragma unroll
    for (int k = 1; k < 5; k++) {
      double jf = sum_nodes(P, k, i) * 0.125 * vol;
      d.rhs[k * N + i] = d.rhs[k * N + i] + jf;
    }
  }
  d.phi[i] = phip_of(P, i, vol);
}

// PICL_TEMP_InitSolver.F90:789-835: piclVF from the first projection and
// cv scaled to the bulk state.
__global__ void __launch_bounds__(BLK) k_picl_init_cells(Dev d, PiclDev P) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= d.nC) return;
  int N = d.nC;
  double phi = phip_of(P, i, d.vol[i]);
  d.phi[i] = phi;
  double vFrac = 1.0 - phi;
#pragma unroll
  for (int k = 0; k < 5; k++) d.cv[k * N + i] = vFrac * d.cv[k * N + i];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
const int NODE2V[8] = {0, 3, 4, 7, 1, 2, 5, 6};  // ppiclF node n = i+2j+4k <- hex2v 1,4,5,8,2,3,6,7

// Single-element box volume (ppiclf_solve.f:3285-3394), same operations as the oracle.
double hex_volume(const double *X) {
  double centroid[3];
  for (int ix = 0; ix < 3; ix++) centroid[ix] = 0.0;
  for (int ix = 0; ix < 3; ix++)
    for (int n = 0; n < 8; n++) centroid[ix] = centroid[ix] + X[3 * n + ix];
  for (int ix = 0; ix < 3; ix++) centroid[ix] = centroid[ix] / 8.0;
  double evol = 0.0;
  for (int ix = 0; ix < 3; ix++) {
    int ia = (3 - (ix + 1) > 1 ? 3 - (ix + 1) : 1) - 1;
    int ib = (5 - (ix + 1) < 3 ? 5 - (ix + 1) : 3) - 1;
    for (int iface = 0; iface < 2; iface++) {
      double face[2][2][3];
      for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
          int a[3];
          a[ix] = iface; a[ia] = i; a[ib] = j;
          int n = a[0] + 2 * a[1] + 4 * a[2];
          for (int dd = 0; dd < 3; dd++) face[i][j][dd] = X[3 * n + dd];
        }
      double v1[3], v2[3], v3[3], cross[3];
      for (int dd = 0; dd < 3; dd++) {
        v1[dd] = face[0][1][dd] - face[1][0][dd];
        v2[dd] = centroid[dd] - face[1][0][dd];
      }
      cross[0] = v1[1] * v2[2] - v1[2] * v2[1];
      cross[1] = v1[2] * v2[0] - v1[0] * v2[2];
      cross[2] = v1[0] * v2[1] - v1[1] * v2[0];
      for (int inode = 0; inode < 2; inode++) {
        for (int dd = 0; dd < 3; dd++) v3[dd] = face[inode][inode][dd] - face[1][0][dd];
        double voltet = 0.0;
        for (int dd = 0; dd < 3; dd++) voltet = voltet + v3[dd] * cross[dd];
        evol = evol + fabs(voltet);
      }
    }
  }
  return evol / 6.0;
}

}  // namespace

struct PiclState {
  PiclDev P;
  PiclPhys phys;
  std::vector<double> y0;  // the 12 components as given (constant ones are returned as is)
  bool debug = false;
};

namespace {

void picl_free(PiclState *s) {
  API_LOCK;
  if (!s) return;
  PiclDev &P = s->P;
  void *ptrs[] = {P.buf[0].st, P.buf[0].pc, P.buf[0].pid, P.buf[0].elem, P.buf[1].st, P.buf[1].pc, P.buf[1].pid,
                  P.buf[1].elem, P.par, P.ycN, P.src, P.dest, P.cnt, P.off, P.bsum, P.nAct, P.lostCur,
                  P.fld, P.gRho, P.proj, P.rexp, P.cen, P.X, P.bbox, P.refId, P.inv, P.lbOff, P.lbMem, P.cbOff,
                  P.cbRec, P.dbg, P.vX, P.ev, P.ecb, P.binCnt, P.gq, P.xo, P.vproj, P.vperm, P.vuT,
                  P.colCnt, P.colOff, P.colFill, P.colMem, P.colB, P.colBsum, P.fc, P.tq, P.gBox, P.gvList,
                  P.buf[0].w, P.buf[1].w, P.vint, P.famB, P.snum, P.vrelP, P.rphifP, P.actP, P.lastP,
                  P.fb, P.fbRaw, P.fbg, P.fldRot, P.buf[0].fq, P.buf[1].fq, P.fmean, P.nCall, P.srt};
  for (void *q : ptrs) if (q) (void)hipFree(q);
  delete s;
}

int scan_blocks(int n) { return (n + SCAN_T * SCAN_V - 1) / (SCAN_T * SCAN_V); }

// Counting sort of the located particles by element, then the deposit.
// posXo: the Gaussian deposit reads the start-of-stage positions saved by
// k_particles (0 at set-up, where the particle rows hold them).
void picl_sort_and_deposit(rflu_handle *h, int posXo) {
  PiclDev &P = h->picl->P;
  int n = P.nC + 1, nb = scan_blocks(n);
  {
    ProfScope ps(h, "k_scan");
    hipLaunchKernelGGL(k_scan1, dim3(nb), dim3(SCAN_T), 0, h->stream, P.cnt, n, P.bsum);
    hipLaunchKernelGGL(k_scan2, dim3(1), dim3(SCAN_T), 0, h->stream, P.bsum, nb);
    hipLaunchKernelGGL(k_scan3, dim3(nb), dim3(SCAN_T), 0, h->stream, P.cnt, n, P.bsum, P.off);
  }
  {
    ProfScope ps(h, "k_fill");
    hipLaunchKernelGGL(k_fill, dim3(nblk(P.nP)), dim3(BLK), 0, h->stream, P);
  }
  {
    ProfScope ps(h, "k_deposit");
    hipLaunchKernelGGL(k_rank, dim3(nblk(P.nP)), dim3(BLK), 0, h->stream, P);
    if (P.gauss)
      hipLaunchKernelGGL(k_gq, dim3(nblk(P.nP)), dim3(BLK), 0, h->stream, P, posXo);
    else
      hipLaunchKernelGGL(k_deposit_box, dim3(nblk(P.nC)), dim3(BLK), 0, h->stream, P);
  }
  if (P.gauss) {
    {
      ProfScope ps(h, "k_gsel");  // the bins holding particles and the vertices within reach
      hipLaunchKernelGGL(k_zero_int, dim3(nblk(P.nCB)), dim3(BLK), 0, h->stream, P.binCnt, P.nCB);
      hipLaunchKernelGGL(k_gbox_reset, dim3(1), dim3(64), 0, h->stream, P);
      hipLaunchKernelGGL(k_gcount, dim3(nblk(P.nC)), dim3(BLK), 0, h->stream, P);
      hipLaunchKernelGGL(k_gvsel, dim3(nblk(P.nV)), dim3(BLK), 0, h->stream, P);
    }
    ProfScope ps(h, "k_gauss");  // the deposit itself (k_gauss_w)
    hipLaunchKernelGGL(k_gauss_w, dim3(GW_BLOCKS), dim3(BLK), 0, h->stream, P);
  }
  ProfScope ps(h, "k_permute");
  hipLaunchKernelGGL(k_permute, dim3(nblk(P.nP)), dim3(BLK), 0, h->stream, P);
  hipLaunchKernelGGL(k_flip, dim3(1), dim3(64), 0, h->stream, P);
}

int launch_rhog(rflu_handle *h) {
  ProfScope ps(h, "k_rhog");
  size_t lds = h->tileLds / 64 * 48;
  if (h->gradVariant == 4 && h->tileLds && lds <= h->ldsMax)
    TILE_LAUNCH(k_rhog_tile, lds, h->d, h->picl->P);
  else
    hipLaunchKernelGGL(k_rhog, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, h->picl->P);
  return 0;
}

// PICL_TEMP_Runge for one stage (after SourceTermsMP, before RKUpdateMP).
// ppiclf_solve_IntegrateParticle's work for RK stage istage, from the fields
// in P.fld and dt in scal[S_DTMIN]: collisions' neighbour search, locate,
// interpolate, forces, RK3, then the sort by element and the projection.
void picl_particle_stage(rflu_handle *h, int istage) {
  PiclState *s = h->picl;
  PiclDev &P = s->P;
  hipLaunchKernelGGL(k_zero_int, dim3(nblk(P.nC + 1)), dim3(BLK), 0, h->stream, P.cnt, P.nC + 1);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(BLK), 0, h->stream, P.lostCur, 1);
  if (P.grid) {  // the stage's grid of cubes for the neighbour searches
    ProfScope ps(h, "k_contacts");
    int n = P.colNb + 1, nb = scan_blocks(n);
    hipLaunchKernelGGL(k_zero_int, dim3(nblk(n)), dim3(BLK), 0, h->stream, P.colCnt, n);
    hipLaunchKernelGGL(k_zero_int, dim3(nblk(P.colNb)), dim3(BLK), 0, h->stream, P.colFill, P.colNb);
    hipLaunchKernelGGL(k_col_count, dim3(nblk(P.nP)), dim3(BLK), 0, h->stream, P);
    hipLaunchKernelGGL(k_scan1, dim3(nb), dim3(SCAN_T), 0, h->stream, P.colCnt, n, P.colBsum);
    hipLaunchKernelGGL(k_scan2, dim3(1), dim3(SCAN_T), 0, h->stream, P.colBsum, nb);
    hipLaunchKernelGGL(k_scan3, dim3(nb), dim3(SCAN_T), 0, h->stream, P.colCnt, n, P.colBsum, P.colOff);
    hipLaunchKernelGGL(k_col_fill, dim3(nblk(P.nP)), dim3(BLK), 0, h->stream, P);
    if (P.pairs) hipLaunchKernelGGL(k_col_sort, dim3(nblk(P.colNb)), dim3(BLK), 0, h->stream, P);
    if (P.col) hipLaunchKernelGGL(k_contacts, dim3(nblk(P.nP)), dim3(BLK), 0, h->stream, h->d, P, s->phys);
  }
  if (P.pairs) {  // the particle kernel's first phase: locate and interpolate
    ProfScope ps(h, "k_interp");
    using KP = void (*)(Dev, PiclDev, PiclPhys, int, int);
    static const KP kp[2] = {k_particles<0, 1, 1>, k_particles<1, 1, 1>};
    if (P.am2) hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(BLK), 0, h->stream, P.lastP, 1);
    hipLaunchKernelGGL(kp[s->phys.interpMode], dim3(nblk(P.nP)), dim3(BLK), 0, h->stream, h->d, P, s->phys, istage,
                       1);
  }
  if (P.pairs) {
    ProfScope ps(h, "k_pairs");
    hipLaunchKernelGGL(k_pairs, dim3(nblk(P.nP)), dim3(BLK), 0, h->stream, h->d, P, s->phys);
  }
  {
    ProfScope ps(h, "k_particles");
    using KP = void (*)(Dev, PiclDev, PiclPhys, int, int);
    static const KP kp[2][5] = {
        {k_particles<0, 1>, k_particles<0, 2>, k_particles<0, 3>, k_particles<0, 4>, k_particles<0, 5>},
        {k_particles<1, 1>, k_particles<1, 2>, k_particles<1, 3>, k_particles<1, 4>, k_particles<1, 5>}};
    hipLaunchKernelGGL(kp[s->phys.interpMode][s->phys.qsFlag - 1], dim3(nblk(P.nP)), dim3(BLK), 0, h->stream, h->d,
                       P, s->phys, istage, P.pairs ? 2 : 0);
    if (P.am2) hipLaunchKernelGGL(k_wdot, dim3(nblk(P.nP)), dim3(BLK), 0, h->stream, P);
    if (P.fl) hipLaunchKernelGGL(k_next_call, dim3(1), dim3(64), 0, h->stream, P);
  }
  if (s->phys.vu && istage == 3) hipLaunchKernelGGL(k_vu_time, dim3(1), dim3(64), 0, h->stream, h->d, P);
  picl_sort_and_deposit(h, 1);
}

// PICL_TEMP_Runge for RK stage istage: the fields from the gas state, the
// particle stage, Rocflu's read-back into rhs and piclVF.
int picl_stage(rflu_handle *h, int istage) {
  PiclState *s = h->picl;
  PiclDev &P = s->P;
  if (P.rot3) {  // the feedback force's gradient for D(omega)/Dt
    ProfScope ps(h, "k_fb_grad");
    hipLaunchKernelGGL(k_fb_cells, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, P);
    hipLaunchKernelGGL(k_fb_grad, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, P);
    hipLaunchKernelGGL(k_fb_weno, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, P);
  }
  {
    ProfScope ps(h, "k_picl_fields");
    hipLaunchKernelGGL(k_picl_fields, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, h->gas, P,
                       s->phys.interpMode);
  }
  picl_particle_stage(h, istage);
  ProfScope ps(h, "k_feedback");
  hipLaunchKernelGGL(k_feedback, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, P, s->phys.feedbackFlag);
  return 0;
}

// Uniform bin grid over boxes [lo, hi] (element bounding boxes or points):
// members are appended in the order given.
void build_bins(const std::vector<double> &blo, const std::vector<double> &bhi, const std::vector<int> &order,
                const double lo[3], const double H[3], const int nb[3], std::vector<int> &off,
                std::vector<int> &mem) {
  size_t nbin = (size_t)nb[0] * nb[1] * nb[2];
  std::vector<std::vector<int>> lists(nbin);
  for (int e : order) {
    int a[3], b[3];
    for (int l = 0; l < 3; l++) {
      double ta = (blo[3 * (size_t)e + l] - lo[l]) / H[l], tb = (bhi[3 * (size_t)e + l] - lo[l]) / H[l];
      ta = std::min(std::max(ta, 0.0), (double)(nb[l] - 1));
      tb = std::min(std::max(tb, 0.0), (double)(nb[l] - 1));
      a[l] = (int)ta;
      b[l] = (int)tb;
    }
    for (int z = a[2]; z <= b[2]; z++)
      for (int y = a[1]; y <= b[1]; y++)
        for (int x = a[0]; x <= b[0]; x++) lists[x + (size_t)nb[0] * (y + (size_t)nb[1] * z)].push_back(e);
  }
  off.assign(nbin + 1, 0);
  for (size_t q = 0; q < nbin; q++) off[q + 1] = off[q] + (int)lists[q].size();
  mem.clear();
  mem.reserve(off[nbin]);
  for (auto &l : lists) mem.insert(mem.end(), l.begin(), l.end());
  if (mem.empty()) mem.push_back(0);
}

}  // namespace

// Static staging for picl_gpu_locate_static: PICL_LOCATE_STATIC_MAX points
// (3 doubles each) and as many element ids.  Held by every device the library
// runs on until the process ends.
#define SCRATCH_ELEM_OFF ((size_t)PICL_LOCATE_STATIC_MAX * 3 * sizeof(double))
#define SCRATCH_BYTES (SCRATCH_ELEM_OFF + (size_t)PICL_LOCATE_STATIC_MAX * sizeof(int))
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
#define SCRATCH_MAXDEV 64
static unsigned char *g_scratchDev[SCRATCH_MAXDEV];  // each device's address of `scratch`

// Resolve this device's copy of `scratch` once, with the device current.
static int picl_scratch_init(rflu_handle *h) {
  if (h->device < 0 || h->device >= SCRATCH_MAXDEV) return set_err(h, RFLU_GPU_EINVAL, "device ordinal out of range");
  if (g_scratchDev[h->device]) return 0;
  int cur = 0;
  HIPCHK(h, hipGetDevice(&cur));
  if (cur != h->device) HIPCHK(h, hipSetDevice(h->device));
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (cur != h->device) (void)hipSetDevice(cur);
  if (e != hipSuccess || !ptr)
    return set_err(h, RFLU_GPU_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratchDev[h->device] = static_cast<unsigned char *>(ptr);
  return 0;
}

extern "C" {

// ppiclf_solve_Initialize's angular case (ppiclf_solve.f:60-104, its pi the
// single-precision ACOS(-1.0)) and ppiclf_solve_InitAngularPeriodic
// (:1172-1230), then InvokeAngularPeriodic's rotation matrices (about e_z,
// Rodrigues' form with the reference's expressions and cos/sin of the three
// local angles +angle, -angle, 0) computed once here.
static int angular_setup(rflu_handle *h, const picl_params *pp, PiclPhys &A) {
  const double angle = pp->angAngle, xangle = pp->angXAngle;
  const double pif = (double)acosf(-1.0f);
  const double angled = angle * 180.0 / pif;
  int c = 0;
  if (angled < 90.0) c = 1;
  if (std::lround(angled) == 90) c = 2;
  if (std::lround(angled) == 180) c = 3;
  if (!(pp->angRin < pp->angRout))
    return set_err(h, RFLU_GPU_EINVAL, "Angular Per must have rin < rout (ppiclf_solve_InitAngularPeriodic)");
  if (c == 0 || !(angle > 0.0))
    return set_err(h, RFLU_GPU_EINVAL, "Invalid Rotational Case! (the wedge must open by less than 90, by 90 or by "
                                       "180 degrees)");
  A.ang = 1;
  A.angCase = c;
  A.angX = xangle;
  A.angUp = xangle + angle;
  A.angMid = 0.5 * (xangle + angle);
  A.angXd[0] = c == 3 ? -1.0 * pp->angRout : pp->angRin;
  A.angXd[1] = pp->angRout;
  const double ex = 0.0, ey = 0.0, ez = 1.0;
  const double la[3] = {angle, -1.0 * angle, 0.0};
  for (int k = 0; k < 3; k++) {
    double ct = cos(la[k]), st = sin(la[k]);
    double *R = A.angR[k];
    R[0] = ct + (1.0 - ct) * ex * ex;
    R[1] = (1.0 - ct) * ex * ey - st * ez;
    R[2] = (1.0 - ct) * ex * ez + st * ey;
    R[3] = (1.0 - ct) * ey * ex + st * ez;
    R[4] = ct + (1.0 - ct) * ey * ey;
    R[5] = (1.0 - ct) * ey * ez - st * ex;
    R[6] = (1.0 - ct) * ez * ex - st * ey;
    R[7] = (1.0 - ct) * ez * ey + st * ex;
    R[8] = ct + (1.0 - ct) * ez * ez;
  }
  return 0;
}

int picl_gpu_init(rflu_handle *h, const picl_params *pp, int nPart, const double *y, const double *pconst) {
  API_LOCK;
  if (!h || !pp || nPart < 0 || (nPart > 0 && (!y || !pconst))) return RFLU_GPU_EINVAL;
  if (h->picl) return set_err(h, RFLU_GPU_EINVAL, "particles are already attached to this region");
  const rflu::MeshHost &m = h->mesh;
  bool allHex = m.nvc == 8;
  for (size_t q = 0; allHex && q < h->c2v.size(); q++) allHex = h->c2v[q] >= 0;
  if (!allHex) return set_err(h, RFLU_GPU_EINVAL, "PICL needs an all-hex grid (PICL_TEMP_InitSolver.F90:562-597)");
  if (h->prm.spaceOrder != 2) return set_err(h, RFLU_GPU_EINVAL, "PICL needs spaceOrder 2 (its fields use gradients)");
  if (h->prm.patchOrder != 1) return set_err(h, RFLU_GPU_EINVAL, "PICL is built for patchOrder 1");
  if (h->prm.fluxPrecision != 0) return set_err(h, RFLU_GPU_EINVAL, "PICL runs in FP64 (fluxPrecision 0)");
  if (h->d.nInst) return set_err(h, RFLU_GPU_EINVAL, "particles attach to a single region, not a batch");
  if (pp->heatTransfer < 0 || pp->heatTransfer > 4)
    return set_err(h, RFLU_GPU_EINVAL, "heatTransfer must be 0 (off), 1 (Stokes), 2 (Ranz-Marshall), 3 (Gunn) or 4 (Fox)");
  if (pp->qsFlag < 1 || pp->qsFlag > 5)
    return set_err(h, RFLU_GPU_EINVAL, "qsFlag must be 1 (Parmar), 2 (Osnes), 3 (modified Parmar), 4 (Gidaspow) or 5 "
                                       "(Schiller-Naumann)");
  if (pp->interpMode < 0 || pp->interpMode > 1) return set_err(h, RFLU_GPU_EINVAL, "interpMode must be 0 or 1");
  if (pp->projMode < 0 || pp->projMode > 1) return set_err(h, RFLU_GPU_EINVAL, "projMode must be 0 or 1");
  if (pp->projMode == 1 && !(pp->filterWidth > 0.0 && pp->filterAlpha >= 1.0e-300 && pp->filterAlpha < 1.0))
    return set_err(h, RFLU_GPU_EINVAL, "the Gaussian filter needs filterWidth > 0 and 1e-300 <= filterAlpha < 1");
  if (pp->amFlag < 0 || pp->amFlag > 2)
    return set_err(h, RFLU_GPU_EINVAL, "amFlag must be 0 (off), 1 (Parmar) or 2 (Briney)");
  if (!(pp->neighborWidth >= 0.0)) return set_err(h, RFLU_GPU_EINVAL, "neighborWidth must be >= 0");
  if (pp->qsFluct < 0 || pp->qsFluct > 2)
    return set_err(h, RFLU_GPU_EINVAL, "qsFluct must be 0 (off), 1 (Lattanzi) or 2 (Osnes)");
  if (pp->qsFluctAdapt < 0 || pp->qsFluctAdapt > 1) return set_err(h, RFLU_GPU_EINVAL, "qsFluctAdapt must be 0 or 1");
  if (!(pp->fluctHalfWidth >= 0.0)) return set_err(h, RFLU_GPU_EINVAL, "fluctHalfWidth must be >= 0");
  if (pp->viscousUnsteady < 0 || pp->viscousUnsteady > 1)
    return set_err(h, RFLU_GPU_EINVAL, "viscousUnsteady must be 0 or 1");
  if (pp->collisionalFlag < 0 || pp->collisionalFlag > 4)
    return set_err(h, RFLU_GPU_EINVAL,
                   "collisionalFlag must be 0 (off), 1 (normal force), 2 (+ tangential force and contact torques), "
                   "3 (+ hydrodynamic and undisturbed torque) or 4 (+ Saffman and Magnus lift)");
  if (pp->collisionalFlag && !(pp->ksp > 0.0 && pp->erest > 0.0 && pp->erest <= 1.0))
    return set_err(h, RFLU_GPU_EINVAL, "collisions need ksp > 0 and 0 < erest <= 1");
  if (nPart < 1) return set_err(h, RFLU_GPU_EINVAL, "no particles");
  // PICL_TEMP_InitSolver.F90:234-262: a periodic axis spans the mesh's
  // vertex minimum to maximum (ppiclf_solve_InitPeriodicX/Y/Z need xl < xr)
  const int perFlag[3] = {pp->periodicX, pp->periodicY, pp->periodicZ};
  double vlo[3] = {1e300, 1e300, 1e300}, vhi[3] = {-1e300, -1e300, -1e300};
  for (size_t v = 0; v < h->xyz.size() / 3; v++)
    for (int l = 0; l < 3; l++) {
      vlo[l] = std::min(vlo[l], h->xyz[3 * v + l]);
      vhi[l] = std::max(vhi[l], h->xyz[3 * v + l]);
    }
  for (int l = 0; l < 3; l++) {
    if (perFlag[l] < 0 || perFlag[l] > 1)
      return set_err(h, RFLU_GPU_EINVAL, "periodicX/Y/Z must be 0 or 1");
    if (perFlag[l] && !(vlo[l] < vhi[l]))
      return set_err(h, RFLU_GPU_EINVAL, "a periodic axis needs a mesh of positive extent along it");
  }
  PiclPhys angPh;
  if (pp->angularPeriodic < 0 || pp->angularPeriodic > 1)
    return set_err(h, RFLU_GPU_EINVAL, "angularPeriodic must be 0 or 1");
  if (pp->angularPeriodic) {
    // PICL_TEMP_InitSolver.F90:700-705 and ppiclf_solve.f:75-78
    if (perFlag[0] || perFlag[1])
      return set_err(h, RFLU_GPU_EINVAL, "PPICLF: Invalid Periodicity choice (angular with periodicX or periodicY)");
    if (pp->projMode != 0 || pp->collisionalFlag || pp->amFlag == 2 || pp->qsFluct)
      return set_err(h, RFLU_GPU_EINVAL, "angular periodicity needs the box projection and no pair models "
                                         "(ppiclF's rotated ghost particles are not built)");
    if (int e = angular_setup(h, pp, angPh)) return e;
  }
  if (int e = picl_scratch_init(h)) return e;  // the static staging of picl_gpu_locate_static
  const int N = h->d.nC, nP = nPart;
  PiclState *s = new PiclState();
  PiclDev &P = s->P;
  P.nP = nP;
  P.nC = N;
  s->phys = PiclPhys{pp->rmuFlag, pp->amFlag, pp->pgFlag, pp->feedbackFlag, pp->qsFlag, pp->interpMode,
                     pp->rmuRef, pp->tref, pp->suth, pp->rcpPart, pp->heatTransfer};
  for (int l = 0; l < 3; l++) {
    s->phys.per[l] = perFlag[l];
    s->phys.perLo[l] = vlo[l];
    s->phys.perHi[l] = vhi[l];
  }
  if (pp->angularPeriodic) {
    s->phys.ang = 1;
    s->phys.angCase = angPh.angCase;
    s->phys.angX = angPh.angX;
    s->phys.angUp = angPh.angUp;
    s->phys.angMid = angPh.angMid;
    memcpy(s->phys.angR, angPh.angR, sizeof angPh.angR);
    s->phys.angXd[0] = angPh.angXd[0];
    s->phys.angXd[1] = angPh.angXd[1];
  }
  s->P.ht = pp->heatTransfer != 0 ? 1 : 0;
  s->phys.vu = s->P.vu = pp->viscousUnsteady;
  s->phys.col = s->P.col = pp->collisionalFlag;
  s->P.rot = pp->collisionalFlag >= 2 ? 1 : 0;
  s->phys.ksp = pp->ksp;
  s->phys.erest = pp->erest;
  s->P.am2 = pp->amFlag == 2 ? 1 : 0;
  s->P.rot3 = pp->collisionalFlag >= 3 ? 1 : 0;
  s->P.fl = pp->qsFluct >= 1 ? 1 : 0;
  s->P.pairs = s->P.am2 || s->P.fl ? 1 : 0;
  s->phys.qsFluct = pp->qsFluct;
  s->phys.qsFluctAdapt = pp->qsFluctAdapt;
  s->phys.fluctSeed = pp->fluctSeed;
  s->P.nVint = PICL_NFLD + (s->P.rot3 ? 6 : 0);
  s->P.grid = s->P.col || s->P.am2 || s->P.fl ? 1 : 0;
  s->P.nRow = s->P.vu ? NROW_VU : (s->P.rot ? ROW_VU : NROW);
  s->y0.assign(y, y + 12 * (size_t)nP);
  // element geometry in device numbering
  std::vector<double> X(24 * (size_t)N), bbox(8 * (size_t)N, 0.0), cen(4 * (size_t)N, 0.0), rexp(N);
  std::vector<double> blo(3 * (size_t)N), bhi(3 * (size_t)N);
  std::vector<int> refId(N), inv(N);
  double lastLen[3] = {0, 0, 0}, lastExt2[3] = {0, 0, 0}, maxExt[3] = {0, 0, 0}, glo[3] = {1e300, 1e300, 1e300}, ghi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < N; i++) {
    int o = h->perm[i];
    refId[i] = o;
    inv[o] = i;
    double *Xe = X.data() + 24 * (size_t)i;
    for (int n = 0; n < 8; n++)
      for (int dd = 0; dd < 3; dd++) Xe[3 * n + dd] = h->xyz[3 * (size_t)h->c2v[8 * (size_t)o + NODE2V[n]] + dd];
    for (int l = 0; l < 3; l++) {
      double c = 0.0, mx = -1.0e10, mn = 1.0e10;
      for (int n = 0; n < 8; n++) {
        c = c + Xe[3 * n + l];
        if (Xe[3 * n + l] > mx) mx = Xe[3 * n + l];
        if (Xe[3 * n + l] < mn) mn = Xe[3 * n + l];
      }
      cen[4 * (size_t)i + l] = c / 8;
      bbox[8 * (size_t)i + l] = mn;
      bbox[8 * (size_t)i + 4 + l] = mx;
      double tol = 1.0e-9 * (mx - mn);
      blo[3 * (size_t)i + l] = mn - 2.0 * tol;
      bhi[3 * (size_t)i + l] = mx + 2.0 * tol;
      maxExt[l] = std::max(maxExt[l], mx - mn);
      glo[l] = std::min(glo[l], mn);
      ghi[l] = std::max(ghi[l], mx);
      if (o == N - 1) lastLen[l] = std::fabs(mx - mn);
      if (o == N - 1) {  // LocalInterp keeps the last element's extent (ppiclf_solve.f:2263-2289)
        double d2i = (mx - mn) * (mx - mn), d2m = 0.0;
        if (d2i > d2m) d2m = d2i;
        lastExt2[l] = d2m;
      }
    }
    rexp[i] = 1.0 / hex_volume(Xe);
  }
  for (int l = 0; l < 3; l++) P.d2Max[l] = lastExt2[l] * (1.5 * 1.5);
  if (P.fl) {
    // d2chk(2) as the filter's set-up leaves it.  The box filter: rocpicl's
    // filter = 4 MAXVAL(Max_EleLen) / 2 (PICL_TEMP_InitSolver.F90:598-646,
    // Max_EleLen reset for every element, so the last element's extents
    // decide; those outside [1e-7, 1e-2] count as 1.23456789e-10), which
    // ppiclf_solve_InitBoxFilter halves (ppiclf_solve.f:1611).  The Gaussian
    // filter: its cut-off radius (ppiclf_solve.f:1558-1559).
    double w2;
    if (pp->projMode == 1) {
      double rsig = pp->filterWidth / (2.0 * sqrt(2.0 * log(2.0)));
      w2 = rsig * sqrt(-2.0 * log(pp->filterAlpha));
    } else {
      double mxl = 0.0;
      for (int l = 0; l < 3; l++) {
        double mel = 1.23456789e-10;
        if (!(lastLen[l] > 1e-2 || lastLen[l] < 1e-7) && lastLen[l] > mel) mel = lastLen[l];
        if (l == 0 || mel > mxl) mxl = mel;
      }
      double filter = 4.0 * mxl;
      filter = filter / 2.0;
      w2 = filter / 2.0;
    }
    s->phys.fluctW = pp->fluctHalfWidth > 0.0 ? pp->fluctHalfWidth : w2;
  }
  // location bins: one maximal element extent per bin, members in reference order
  std::vector<int> refOrder(N);
  for (int o = 0; o < N; o++) refOrder[o] = inv[o];
  for (int l = 0; l < 3; l++) {
    P.lbLo[l] = glo[l];
    P.lbH[l] = maxExt[l] > 0 ? maxExt[l] : 1.0;
    P.lbN[l] = std::max(1, std::min(4096, (int)std::floor((ghi[l] - glo[l]) / P.lbH[l]) + 1));
  }
  std::vector<int> lbOff, lbMem, cbOff, cbMem;
  build_bins(blo, bhi, refOrder, P.lbLo, P.lbH, P.lbN, lbOff, lbMem);
  // centroid bins: width = the interpolation reach per axis
  std::vector<double> cpts(3 * (size_t)N);
  double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < N; i++)
    for (int l = 0; l < 3; l++) {
      cpts[3 * (size_t)i + l] = cen[4 * (size_t)i + l];
      clo[l] = std::min(clo[l], cen[4 * (size_t)i + l]);
      chi[l] = std::max(chi[l], cen[4 * (size_t)i + l]);
    }
  for (int l = 0; l < 3; l++) {
    double r = std::sqrt(P.d2Max[l]);
    P.R[l] = r * (1.0 + 1.0e-9) + 1.0e-300;
    P.cbLo[l] = clo[l];
    P.cbH[l] = r > 0 ? r : (maxExt[l] > 0 ? maxExt[l] : 1.0);
    P.cbN[l] = std::max(1, std::min(4096, (int)std::floor((chi[l] - clo[l]) / P.cbH[l]) + 1));
    P.cbTop[l] = (chi[l] - clo[l]) / P.cbH[l];
  }
  std::vector<int> idOrder(N);
  for (int i = 0; i < N; i++) idOrder[i] = i;
  build_bins(cpts, cpts, idOrder, P.cbLo, P.cbH, P.cbN, cbOff, cbMem);
  // particles, SoA rows in the given order (the first sort groups them)
  std::vector<double> st(P.nRow * (size_t)nP, 0.0), pcs(4 * (size_t)nP);  // VU history starts at 0 (ppiclf_solve.f:432-434)
  for (int p = 0; p < nP; p++) {
    for (int k = 0; k < 6; k++) st[(ROW_X + k) * (size_t)nP + p] = y[12 * (size_t)p + k];
    st[ROW_T * (size_t)nP + p] = y[12 * (size_t)p + 6];
    if (P.rot)
      for (int k = 0; k < 3; k++) st[(ROW_O + k) * (size_t)nP + p] = y[12 * (size_t)p + 7 + k];
    for (int k = 0; k < 4; k++) pcs[k * (size_t)nP + p] = pconst[4 * (size_t)p + k];
  }
  std::vector<int> ident(nP), minus1(nP, -1);
  for (int p = 0; p < nP; p++) ident[p] = p;
  int zero = 0;
  int rc = 0;
  for (int b = 0; b < 2; b++) {
    rc |= dalloc(h, &P.buf[b].st, P.nRow * (size_t)nP, b == 0 ? st.data() : nullptr);
    rc |= dalloc(h, &P.buf[b].pc, 4 * (size_t)nP, b == 0 ? pcs.data() : nullptr);
    rc |= dalloc(h, &P.buf[b].pid, (size_t)nP, b == 0 ? ident.data() : nullptr);
    rc |= dalloc(h, &P.buf[b].elem, (size_t)nP, minus1.data());
  }
  rc |= dalloc(h, &P.par, 1, &zero);
  if (P.grid) {  // grid of cubes over the bounding box for the neighbour searches
    double dmax = 0.0;
    for (int p = 0; p < nP; p++) dmax = std::max(dmax, pconst[4 * (size_t)p + 1]);
    if (!(dmax > 0.0)) {
      picl_free(s);
      return set_err(h, RFLU_GPU_EINVAL, "collisions and Briney's added mass need positive diameters");
    }
    P.colR = dmax * (1.0 + 1.0e-9);
    P.colDmax = dmax;
    // collisions reach (d + d_max) / 2: cubes of a third of d_max; the
    // added-mass pairs reach d2chk(3): cubes of half of it
    P.colH = dmax / 3.0 * (1.0 + 1.0e-8);
    if (P.pairs) {
      // PICL_TEMP_InitSolver.F90:655-661 passes max(4 d_max, NEIGHBORWIDTH)
      s->phys.nbW = pp->neighborWidth > 0.0 ? pp->neighborWidth : 4.0 * dmax;
      P.colH = std::max(P.colH, 0.5 * s->phys.nbW * (1.0 + 1.0e-8));
    }
    for (int l = 0; l < 3; l++) { P.colLo[l] = vlo[l]; P.colHi[l] = vhi[l]; }
    // at most 2^25 cells and 8 per particle: coarser cubes find the same
    // contacts among more candidates (the span loops take any width)
    const double cap = std::min((double)(1 << 25), 8.0 * nP + 4096.0);
    for (;;) {
      double cells = 1.0;
      for (int l = 0; l < 3; l++) {
        P.colN[l] = std::max(1, (int)std::ceil((vhi[l] - vlo[l]) / P.colH));
        cells *= P.colN[l];
      }
      if (cells <= cap) break;
      P.colH *= 1.25;
    }
    P.colNb = P.colN[0] * P.colN[1] * P.colN[2];
    rc |= dalloc(h, &P.colCnt, (size_t)P.colNb + 1);
    rc |= dalloc(h, &P.colOff, (size_t)P.colNb + 1);
    rc |= dalloc(h, &P.colFill, (size_t)P.colNb);
    rc |= dalloc(h, &P.colMem, (size_t)nP);
    rc |= dalloc(h, &P.colB, (size_t)nP);
    rc |= dalloc(h, &P.colBsum, (size_t)scan_blocks(P.colNb + 1));
    if (P.col) rc |= dalloc(h, &P.fc, 3 * (size_t)nP);
    if (P.rot) rc |= dalloc(h, &P.tq, 3 * (size_t)nP);
  }
  if (P.pairs) rc |= dalloc(h, &P.vint, (size_t)P.nVint * nP);
  if (P.am2) {  // W starts at 0 (the zeroed rprop, ppiclf_solve.f:432-434)
    std::vector<double> z3(3 * (size_t)nP, 0.0);
    std::vector<int> zi((size_t)nP, 0);
    for (int b = 0; b < 2; b++) rc |= dalloc(h, &P.buf[b].w, 3 * (size_t)nP, z3.data());
    rc |= dalloc(h, &P.famB, 4 * (size_t)nP);
    rc |= dalloc(h, &P.snum, (size_t)nP);
    rc |= dalloc(h, &P.vrelP, 3 * (size_t)nP, z3.data());
    rc |= dalloc(h, &P.rphifP, (size_t)nP, z3.data());
    rc |= dalloc(h, &P.actP, (size_t)nP, zi.data());
    rc |= dalloc(h, &P.lastP, 1, zi.data());
  }
  if (P.fl) {  // the fluctuating force starts at 0 (the zeroed rprop)
    std::vector<double> z3(3 * (size_t)nP, 0.0);
    int zero1 = 0;
    for (int b = 0; b < 2; b++) rc |= dalloc(h, &P.buf[b].fq, 3 * (size_t)nP, z3.data());
    rc |= dalloc(h, &P.fmean, 7 * (size_t)nP);
    rc |= dalloc(h, &P.nCall, 1, &zero1);
  }
  if (P.vu) {  // ppiclf_nTimeBH = 1 (PICL_TEMP_InitSolver.F90:290), ages 0
    std::vector<double> t0(VU_N + 1, 0.0);
    t0[VU_N] = 1.0;
    rc |= dalloc(h, &P.vuT, t0.size(), t0.data());
  }
  if (P.rot3) {
    rc |= dalloc(h, &P.fb, 3 * (size_t)N);
    rc |= dalloc(h, &P.fbRaw, 9 * (size_t)N);
    rc |= dalloc(h, &P.fbg, 9 * (size_t)N);
    rc |= dalloc(h, &P.fldRot, 6 * (size_t)N);
  }
  rc |= dalloc(h, &P.ycN, 4 * (size_t)nP);
  if (pp->angularPeriodic && perFlag[2]) {
    std::vector<int> z0(nP, 0);
    rc |= dalloc(h, &P.rmMark, (size_t)nP, z0.data());
  }
  rc |= dalloc(h, &P.src, (size_t)nP);
  rc |= dalloc(h, &P.dest, (size_t)nP);
  rc |= dalloc(h, &P.srt, (size_t)nP);
  rc |= dalloc(h, &P.lostCur, 1, &zero);
  rc |= dalloc(h, &P.cnt, (size_t)N + 1);
  rc |= dalloc(h, &P.off, (size_t)N + 1);
  rc |= dalloc(h, &P.bsum, (size_t)scan_blocks(N + 1));
  rc |= dalloc(h, &P.nAct, 1, &nP);
  rc |= dalloc(h, &P.fld, PICL_NFLD * (size_t)N);
  rc |= dalloc(h, &P.gRho, 4 * (size_t)N);
  rc |= dalloc(h, &P.proj, 5 * (size_t)N);
  rc |= dalloc(h, &P.rexp, (size_t)N, rexp.data());
  rc |= dalloc(h, &P.cen, 4 * (size_t)N, cen.data());
  rc |= dalloc(h, &P.X, 24 * (size_t)N, X.data());
  rc |= dalloc(h, &P.bbox, 8 * (size_t)N, bbox.data());
  rc |= dalloc(h, &P.refId, (size_t)N, refId.data());
  rc |= dalloc(h, &P.inv, (size_t)N, inv.data());
  rc |= dalloc(h, &P.lbOff, lbOff.size(), lbOff.data());
  rc |= dalloc(h, &P.lbMem, lbMem.size(), lbMem.data());
  rc |= dalloc(h, &P.cbOff, cbOff.size(), cbOff.data());
  std::vector<double> cbRec(4 * cbMem.size());
  for (size_t q = 0; q < cbMem.size(); q++) {
    int c = cbMem[q];
    for (int l = 0; l < 3; l++) cbRec[4 * q + l] = cen[4 * (size_t)c + l];
    long long key = ((long long)refId[c] << 32) | (unsigned)c;
    memcpy(&cbRec[4 * q + 3], &key, sizeof key);
  }
  rc |= dalloc(h, &P.cbRec, cbRec.size(), cbRec.data());
  if (pp->projMode == 1) {
    // ppiclf_solve_InitGaussianFilter (:1520-1576) and ProjectParticleGrid's
    // constants (:3103-3109); same operations as the oracle
    const double pi = 4.0 * atan(1.0);
    double rsig = pp->filterWidth / (2.0 * sqrt(2.0 * log(2.0)));
    double d2chk = rsig * sqrt(-2.0 * log(pp->filterAlpha));
    double s2 = sqrt(2.0 * pi);
    double multfci = 1.0 / ((s2 * s2) * (rsig * rsig));
    P.gMult = pow(multfci, 1.5);
    P.gRdum = 1.0 / (-2.0 * (rsig * rsig));
    P.gD2chk = d2chk * d2chk;
    // reach of an element's particles: cut-off + the largest centroid-to-node
    // distance (a located point is a convex combination of the nodes)
    double hmax = 0.0;
    for (int i = 0; i < N; i++)
      for (int n = 0; n < 8; n++) {
        double r2 = 0.0;
        for (int l = 0; l < 3; l++) {
          double t = X[24 * (size_t)i + 3 * n + l] - cen[4 * (size_t)i + l];
          r2 += t * t;
        }
        hmax = std::max(hmax, std::sqrt(r2));
      }
    P.gReach = (d2chk + hmax) * (1.0 + 1.0e-6);
    P.gReach2 = P.gReach * P.gReach;
    for (int l = 0; l < 3; l++) {
      double hl = 0.0;
      for (int i = 0; i < N; i++)
        for (int n = 0; n < 8; n++) hl = std::max(hl, std::fabs(X[24 * (size_t)i + 3 * n + l] - cen[4 * (size_t)i + l]));
      P.gHalf[l] = hl * (1.0 + 1.0e-6);
    }
    P.gCut2 = (d2chk * (1.0 + 1.0e-6)) * (d2chk * (1.0 + 1.0e-6));
    P.gauss = 1;
    // ghost particles across periodic bounds (ppiclf_comm.f:1171-1185,
    // CheckPeriodicBC :2136-2206): images shifted by +-(hi - lo)
    for (int l = 0; l < 3; l++) {
      P.gShift[l] = perFlag[l] ? vhi[l] - vlo[l] : 0.0;
      if (perFlag[l]) P.gImg = 1;
    }
    P.nV = (int)(h->xyz.size() / 3);
    P.nCB = P.cbN[0] * P.cbN[1] * P.cbN[2];
    std::vector<double> vX(3 * (size_t)P.nV);
    for (int v = 0; v < P.nV; v++)
      for (int l = 0; l < 3; l++) vX[l * (size_t)P.nV + v] = h->xyz[3 * (size_t)v + l];
    std::vector<int> ev(8 * (size_t)N), ecb(N);
    for (int i = 0; i < N; i++) {
      for (int n = 0; n < 8; n++) ev[n * (size_t)N + i] = h->c2v[8 * (size_t)refId[i] + NODE2V[n]];
      int b[3];
      for (int l = 0; l < 3; l++) {
        double t = (cen[4 * (size_t)i + l] - P.cbLo[l]) / P.cbH[l];
        t = std::min(std::max(t, 0.0), (double)(P.cbN[l] - 1));
        b[l] = (int)t;
      }
      ecb[i] = b[0] + P.cbN[0] * (b[1] + P.cbN[1] * b[2]);
    }
    rc |= dalloc(h, &P.vX, vX.size(), vX.data());
    rc |= dalloc(h, &P.ev, ev.size(), ev.data());
    rc |= dalloc(h, &P.ecb, ecb.size(), ecb.data());
    rc |= dalloc(h, &P.binCnt, (size_t)P.nCB);
    rc |= dalloc(h, &P.gBox, 6);
    rc |= dalloc(h, &P.gvList, (size_t)P.nV + 1);
    rc |= dalloc(h, &P.gq, 8 * (size_t)nP);
    rc |= dalloc(h, &P.xo, 3 * (size_t)nP);
    rc |= dalloc(h, &P.vproj, 5 * (size_t)P.nV);
    // vertices along a Morton curve: the waves an XCD runs at once then work
    // on one compact region and share its particles in that XCD's L2
    std::vector<int> vperm;
    morton_order(h->xyz.data(), P.nV, vperm);
    rc |= dalloc(h, &P.vperm, vperm.size(), vperm.data());
  }
  rc |= dalloc(h, &h->d.phi, (size_t)N);
  rc |= dalloc(h, &h->d.rhog, (size_t)N);
  rc |= dalloc(h, &h->d.gRaw, 4 * (size_t)N);
  if (rc) { std::string msg = h->err; picl_free(s); return set_err(h, RFLU_GPU_EHIP, msg); }
  HIPCHK(h, hipMemsetAsync(P.buf[1].st, 0, P.nRow * (size_t)nP * sizeof(double), h->stream));
  HIPCHK(h, hipMemsetAsync(P.buf[1].pc, 0, 4 * (size_t)nP * sizeof(double), h->stream));
  HIPCHK(h, hipMemsetAsync(P.buf[1].pid, 0, (size_t)nP * sizeof(int), h->stream));
  HIPCHK(h, hipMemsetAsync(P.ycN, 0, 4 * (size_t)nP * sizeof(double), h->stream));
  HIPCHK(h, hipMemsetAsync(P.src, 0, (size_t)nP * sizeof(int), h->stream));
  HIPCHK(h, hipMemsetAsync(P.dest, 0, (size_t)nP * sizeof(int), h->stream));
  HIPCHK(h, hipMemsetAsync(h->d.gRaw, 0, 4 * (size_t)N * sizeof(double), h->stream));
  HIPCHK(h, hipMemsetAsync(P.gRho, 0, 4 * (size_t)N * sizeof(double), h->stream));
  HIPCHK(h, hipMemsetAsync(P.fld, 0, PICL_NFLD * (size_t)N * sizeof(double), h->stream));
  h->picl = s;
  if (h->graph) { hipGraphExecDestroy(h->graph); h->graph = nullptr; }  // the step changes
  // ppiclf_comm_InitOverlapMesh -> InitSolve: locate and project (ydotc = 0)
  hipLaunchKernelGGL(k_zero_int, dim3(nblk(N + 1)), dim3(BLK), 0, h->stream, P.cnt, N + 1);
  hipLaunchKernelGGL(k_locate_init, dim3(nblk(nP)), dim3(BLK), 0, h->stream, P);
  picl_sort_and_deposit(h, 0);
  hipLaunchKernelGGL(k_picl_init_cells, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, P);
  hipLaunchKernelGGL(k_records, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, h->gas, h->prm.spaceOrder);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int picl_gpu_runge(rflu_handle *h, int istage) {
  if (!h || istage < 1 || istage > 3) return RFLU_GPU_EINVAL;
  if (!h->picl) return set_err(h, RFLU_GPU_EINVAL, "no particles attached (picl_gpu_init)");
  picl_stage(h, istage);
  HIPCHK(h, hipGetLastError());
  return 0;
}

int picl_gpu_sizes(const rflu_handle *h, int *nPart, int *nLost) {
  API_LOCK;
  if (!h || !h->picl) return RFLU_GPU_EINVAL;
  const PiclDev &P = h->picl->P;
  int act = 0;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return RFLU_GPU_EHIP;
  if (hipMemcpy(&act, P.nAct, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return RFLU_GPU_EHIP;
  if (nPart) *nPart = P.nP;
  if (nLost) *nLost = P.nP - act;
  return 0;
}

int picl_gpu_get_particles(rflu_handle *h, double *y, double *y1, double *ydot, double *ydotc, int *elem) {
  API_LOCK;
  if (!h || !h->picl) return RFLU_GPU_EINVAL;
  PiclState *s = h->picl;
  PiclDev &P = s->P;
  size_t nP = P.nP;
  int par = 0;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(&par, P.par, sizeof(int), hipMemcpyDeviceToHost));
  const PBuf &C = P.buf[par];
  std::vector<double> st((size_t)(P.rot ? ROW_VU : NROW) * nP);
  std::vector<int> pid(nP), ev(nP);
  HIPCHK(h, hipMemcpy(st.data(), C.st, st.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(pid.data(), C.pid, nP * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(ev.data(), C.elem, nP * sizeof(int), hipMemcpyDeviceToHost));
  auto fill = [&](double *dst, int row, int ncomp, int first, bool keepConst) {
    if (!dst) return;
    for (size_t t = 0; t < nP; t++) {
      size_t p = pid[t];
      for (int j = 0; j < 12; j++) dst[12 * p + j] = keepConst ? s->y0[12 * p + j] : 0.0;
      for (int k = 0; k < ncomp; k++) dst[12 * p + first + k] = st[(row + k) * nP + t];
    }
  };
  fill(y, ROW_X, 6, 0, true);
  fill(y1, ROW_Y1, 6, 0, false);
  fill(ydot, ROW_YD, 6, 0, false);
  fill(ydotc, ROW_YCP, 4, 3, false);
  if (P.ht) {  // the particle temperature evolves
    for (size_t t = 0; t < nP; t++) {
      size_t p = pid[t];
      if (y) y[12 * p + 6] = st[ROW_T * nP + t];
      if (y1) y1[12 * p + 6] = st[ROW_TY1 * nP + t];
      if (ydot) ydot[12 * p + 6] = st[ROW_TYD * nP + t];
    }
  }
  if (P.rot) {  // the angular velocity evolves
    for (size_t t = 0; t < nP; t++) {
      size_t p = pid[t];
      for (int c = 0; c < 3; c++) {
        if (y) y[12 * p + 7 + c] = st[(ROW_O + c) * nP + t];
        if (y1) y1[12 * p + 7 + c] = st[(ROW_OY1 + c) * nP + t];
        if (ydot) ydot[12 * p + 7 + c] = st[(ROW_OYD + c) * nP + t];
      }
    }
  }
  if (elem)
    for (size_t t = 0; t < nP; t++) elem[pid[t]] = ev[t] < 0 ? -1 : h->perm[ev[t]];
  return 0;
}

int picl_gpu_get_cells(rflu_handle *h, double *phi, double *gradRhog, double *proFld, double *fields) {
  API_LOCK;
  if (!h || !h->picl) return RFLU_GPU_EINVAL;
  PiclDev &P = h->picl->P;
  size_t N = P.nC;
  if (phi)
    if (int e = from_device_cells(h, phi, h->d.phi, 1, false)) return e;
  if (gradRhog) {
    std::vector<double> g4(4 * N);
    if (int e = from_device_cells(h, g4.data(), P.gRho, 4, true)) return e;
    for (size_t i = 0; i < N; i++)
      for (int k = 0; k < 3; k++) gradRhog[3 * i + k] = g4[4 * i + k];
  }
  if (proFld) {
    if (P.gauss) {  // the value at node (1,1,1)
      std::vector<double> nodes(40 * N);
      if (int e = picl_gpu_get_nodes(h, nodes.data())) return e;
      for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 5; k++) proFld[5 * i + k] = nodes[40 * i + 8 * k];
    } else {
      std::vector<double> pf(5 * N);
      if (int e = from_device_cells(h, pf.data(), P.proj, 5, false)) return e;
      for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 5; k++) proFld[5 * i + k] = pf[k * N + i];
    }
  }
  if (fields)
    if (int e = from_device_cells(h, fields, P.fld, PICL_NFLD, true)) return e;
  return 0;
}

int picl_gpu_get_nodes(rflu_handle *h, double *proNode) {
  API_LOCK;
  if (!h || !h->picl || !proNode) return RFLU_GPU_EINVAL;
  PiclDev &P = h->picl->P;
  size_t N = P.nC;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (P.gauss) {
    std::vector<double> vp(5 * (size_t)P.nV);
    HIPCHK(h, hipMemcpy(vp.data(), P.vproj, vp.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t o = 0; o < N; o++)
      for (int k = 0; k < 5; k++)
        for (int n = 0; n < 8; n++) proNode[40 * o + 8 * k + n] = vp[k * (size_t)P.nV + h->c2v[8 * o + NODE2V[n]]];
  } else {
    std::vector<double> pf(5 * N);
    if (int e = from_device_cells(h, pf.data(), P.proj, 5, false)) return e;
    for (size_t o = 0; o < N; o++)
      for (int k = 0; k < 5; k++)
        for (int n = 0; n < 8; n++) proNode[40 * o + 8 * k + n] = pf[k * N + o];
  }
  return 0;
}

int picl_gpu_get_rot_fields(rflu_handle *h, double *rot6) {
  API_LOCK;
  if (!h || !h->picl || !rot6) return RFLU_GPU_EINVAL;
  PiclDev &P = h->picl->P;
  if (!P.rot3) return set_err(h, RFLU_GPU_EINVAL, "the rotational fields exist only with collisionalFlag >= 3");
  return from_device_cells(h, rot6, P.fldRot, 6, true);
}

int picl_gpu_set_debug(rflu_handle *h, int on) {
  API_LOCK;
  if (!h || !h->picl) return RFLU_GPU_EINVAL;
  PiclDev &P = h->picl->P;
  if (on && !P.dbg) {
    if (dalloc(h, &P.dbg, PICL_NFLD * (size_t)P.nP)) return RFLU_GPU_EHIP;
    HIPCHK(h, hipMemsetAsync(P.dbg, 0, PICL_NFLD * (size_t)P.nP * sizeof(double), h->stream));
  } else if (!on && P.dbg) {
    (void)hipFree(P.dbg);
    P.dbg = nullptr;
  }
  if (h->graph) { hipGraphExecDestroy(h->graph); h->graph = nullptr; }
  return 0;
}

int picl_gpu_get_interp(rflu_handle *h, double *rprop20) {
  API_LOCK;
  if (!h || !h->picl || !rprop20) return RFLU_GPU_EINVAL;
  PiclDev &P = h->picl->P;
  if (!P.dbg) return set_err(h, RFLU_GPU_EINVAL, "interpolated values are kept only with picl_gpu_set_debug(h, 1)");
  size_t nP = P.nP;
  std::vector<double> a(PICL_NFLD * nP);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(a.data(), P.dbg, a.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t p = 0; p < nP; p++)
    for (int f = 0; f < PICL_NFLD; f++) rprop20[PICL_NFLD * p + f] = a[f * nP + p];
  return 0;
}

int picl_gpu_locate(rflu_handle *h, int n, const double *xyz, int *elem) {
  API_LOCK;
  if (!h || !h->picl || n < 0 || (n > 0 && (!xyz || !elem))) return RFLU_GPU_EINVAL;
  if (n == 0) return 0;
  double *dx = nullptr;
  int *de = nullptr;
  const size_t before = h->devBytes;
  if (dalloc(h, &dx, 3 * (size_t)n, xyz) || dalloc(h, &de, (size_t)n)) {
    if (dx) (void)hipFree(dx);
    h->devBytes = before;
    return RFLU_GPU_EHIP;
  }
  hipLaunchKernelGGL(k_locate_pts, dim3(nblk(n)), dim3(BLK), 0, h->stream, h->picl->P, n, dx, de);
  hipError_t e1 = hipGetLastError();
  hipError_t e2 = hipMemcpyAsync(elem, de, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, h->stream);
  hipError_t e3 = hipStreamSynchronize(h->stream);
  (void)hipFree(dx);
  (void)hipFree(de);
  h->devBytes = before;  // the staging buffers live for this call only
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return set_err(h, RFLU_GPU_EHIP, "picl_gpu_locate failed");
  return 0;
}

// picl_gpu_locate without allocation: the points and the answers are staged
// in `scratch` (points at byte 0, answers after PICL_LOCATE_STATIC_MAX points'
// coordinates).  Each device's copy of the array is found once, by
// hipGetSymbolAddress with that device current (picl_scratch_init, run by
// picl_gpu_init); host code uses only that address, never the array's name.
// Every call on a device shares its copy, so calls that overlap on one device
// race; the API lock orders calls made through this library.
int picl_gpu_locate_static(rflu_handle *h, int n, const double *xyz, int *elem) {
  API_LOCK;
  if (!h || !h->picl || n < 0 || (n > 0 && (!xyz || !elem))) return RFLU_GPU_EINVAL;
  if (n > PICL_LOCATE_STATIC_MAX)
    return set_err(h, RFLU_GPU_EINVAL, "picl_gpu_locate_static: n exceeds PICL_LOCATE_STATIC_MAX (" +
                                           std::to_string(PICL_LOCATE_STATIC_MAX) + ") points");
  if (n == 0) return 0;
  unsigned char *base = g_scratchDev[h->device];
  if (!base) return set_err(h, RFLU_GPU_EHIP, "picl_gpu_locate_static: the static scratch of this device is not set up");
  double *dx = reinterpret_cast<double *>(base);
  int *de = reinterpret_cast<int *>(base + SCRATCH_ELEM_OFF);
  HIPCHK(h, hipMemcpyAsync(dx, xyz, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_locate_pts, dim3(nblk(n)), dim3(BLK), 0, h->stream, h->picl->P, n, (const double *)dx, de);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(elem, de, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// ppiclF's F77 entry points (gfortran names: lowercase + '_', arguments by
// reference) over one default instance, as ppiclF keeps one instance per
// process in COMMON blocks.  The instance is bound to a Rocflu region
// (ppiclf_gpu_bind); InitOverlapMesh attaches the particles to it
// (picl_gpu_init), InterpFieldUser stages host fields, IntegrateParticle
// runs one RK stage of the particles on the device from those fields, and
// GetProFldIJKEF reads the projection back.  Errors stop the process with
// ppiclF's message, as ppiclf_exittr does (ppiclf_op.f:288-318), unless the
// host asks for them to be recorded instead.
// ---------------------------------------------------------------------------
namespace {
struct F77State {
  rflu_handle *h = nullptr;
  picl_params pp{};
  bool init = false, linit = false, lfilt = false, overlap = false;
  int per[3] = {0, 0, 0};
  int ang = 0;                                             // ang_per_flag
  double angA = 0.0, angX = 0.0, angRin = 0.0, angRout = 0.0;  // radians, radii
  int npart = 0;
  std::vector<double> y, pc;
  int proj = 0;
  double width = 0.0, alpha = 0.0, nbw = 0.0, filt2 = 0.0;
  int stage = 0;  // ppiclf_solve_IntegrateRK3s_Rocflu's SAVE stage counter
  int icnt = 0;   // PPICLF_INT_ICNT
  std::vector<double> fld;  // [PICL_NFLD + 6][nCells], the host's cell numbering
  std::vector<char> have;
  std::vector<double> pro;  // [nCells][5][8] projection read back
  bool proValid = false;
  bool abortOnError = true;
  std::string err;
};
F77State g_f77;

// rprop index (PPICLF_USER.h) of each field the device interpolates, in its order
const int F77_FIELD_RPROP[PICL_NFLD] = {2, 6, 7, 8, 10, 11, 12, 9, 24, 5, 13, 14, 15, 16, 17, 18, 19, 37, 38, 39};
// with COLLISIONS >= 3: the vorticity and its material derivative (JXVOR..JZVOR,
// JSDOX..JSDOZ; PICL_TEMP_Runge.F90:880-882,890-892)
const int F77_ROT_RPROP[6] = {31, 32, 33, 43, 44, 45};

void f77_fail(const std::string &m) {
  g_f77.err = m;
  if (g_f77.abortOnError) {
    fprintf(stderr, "PPICLF: ERROR %s\n", m.c_str());
    fflush(stderr);
    std::exit(1);
  }
}
}  // namespace

extern "C" {

int ppiclf_gpu_bind(rflu_handle *h, const picl_params *pp) {
  API_LOCK;
  if (!h || !pp) return RFLU_GPU_EINVAL;
  if (h->picl) return set_err(h, RFLU_GPU_EINVAL, "ppiclf_gpu_bind: the region already carries particles");
  bool abortOnError = g_f77.abortOnError;
  g_f77 = F77State();
  g_f77.abortOnError = abortOnError;
  g_f77.h = h;
  g_f77.pp = *pp;
  return 0;
}

int ppiclf_gpu_reset(void) {
  API_LOCK;
  bool abortOnError = g_f77.abortOnError;
  g_f77 = F77State();
  g_f77.abortOnError = abortOnError;
  return 0;
}

void ppiclf_gpu_set_abort(int on) {
  API_LOCK;
  g_f77.abortOnError = on != 0;
}

const char *ppiclf_gpu_last_error(void) { return g_f77.err.c_str(); }

rflu_handle *ppiclf_gpu_handle(void) { return g_f77.h; }

// ppiclf_solve.f:1-180
void ppiclf_solve_initialize_(const int *xi1, const double *xpmin, const double *xpmax, const int *yi1,
                              const double *ypmin, const double *ypmax, const int *zi1, const double *zpmin,
                              const double *zpmax, const int *ai1, const double *apa, const double *apxa,
                              const double *aprin, const double *aprout) {
  API_LOCK;
  if (!g_f77.h) return f77_fail("ppiclf_gpu_bind must come before ppiclf_solve_Initialize");
  if (*ai1 != 0 && *ai1 != 1) return f77_fail("ang_per_flag must be 0 or 1");
  if (*ai1 == 1 && (*xi1 == 1 || *yi1 == 1)) return f77_fail("PPICLF: Invalid Periodicity choice");
  g_f77.ang = *ai1;
  g_f77.angA = *apa;
  g_f77.angX = *apxa;
  g_f77.angRin = *aprin;
  g_f77.angRout = *aprout;
  const int fl[3] = {*xi1, *yi1, *zi1};
  const double lo[3] = {*xpmin, *ypmin, *zpmin}, hi[3] = {*xpmax, *ypmax, *zpmax};
  double vlo[3] = {1e300, 1e300, 1e300}, vhi[3] = {-1e300, -1e300, -1e300};
  const std::vector<double> &X = g_f77.h->xyz;
  for (size_t v = 0; v < X.size() / 3; v++)
    for (int l = 0; l < 3; l++) {
      vlo[l] = std::min(vlo[l], X[3 * v + l]);
      vhi[l] = std::max(vhi[l], X[3 * v + l]);
    }
  static const char *ax = "XYZ";
  for (int l = 0; l < 3; l++) {
    g_f77.per[l] = fl[l] == 1 ? 1 : 0;
    if (fl[l] != 1) continue;
    if (lo[l] >= hi[l]) return f77_fail(std::string("Periodic") + ax[l] + " must have min < max");
    if (lo[l] != vlo[l] || hi[l] != vhi[l])
      return f77_fail(std::string("Periodic") + ax[l] +
                      " bounds must be the grid's vertex minimum and maximum (PICL_TEMP_InitSolver.F90:234-262)");
  }
  g_f77.init = true;
}

// ppiclf_solve.f:184-330
void ppiclf_solve_initparticle_(const int *imethod, const int *ndim, const int *iendian, const int *npart,
                                const double *y, const double *rprop, const double *filt2, const double *filt3) {
  API_LOCK;
  (void)iendian;
  if (!g_f77.init) return f77_fail("ppiclf_solve_Initialize must come before InitParticle");
  if (*imethod != 2) return f77_fail("imethod must be 2 (IntegrateRK3s_Rocflu)");
  if (*ndim != 3) return f77_fail("ndim must be 3");
  if (*npart < 1) return f77_fail("InitParticle needs particles");
  const int n = *npart;
  g_f77.npart = n;
  g_f77.y.assign(y, y + (size_t)PICL_LRS * n);
  g_f77.pc.resize(4 * (size_t)n);
  for (int p = 0; p < n; p++) {  // (rho_p, d_p, V_p, SPL): rprop JRHOP, JDP, JVOLP, JSPL
    const double *r = rprop + (size_t)PICL_LRP * p;
    g_f77.pc[4 * (size_t)p] = r[0];
    g_f77.pc[4 * (size_t)p + 1] = r[2];
    g_f77.pc[4 * (size_t)p + 2] = r[3];
    g_f77.pc[4 * (size_t)p + 3] = r[21];
  }
  g_f77.filt2 = *filt2;  // ppiclf_d2chk(2) until the filter's set-up resets it
  g_f77.nbw = *filt3;  // ppiclf_d2chk(3) until InitNeighborBin resets it (ppiclf_solve.f:224-226)
  g_f77.linit = true;
}

// ppiclf_solve.f:1579-1640
void ppiclf_solve_initboxfilter_(const double *filt, const int *iwallm, const int *sngl_elem) {
  API_LOCK;
  if (!g_f77.linit) return f77_fail("InitParticle must be before InitFilter");
  if (*iwallm != 0) return f77_fail("the wall mirror (iwallm 1) is not built");
  if (*sngl_elem != 1) return f77_fail("the multi-element box filter (sngl_elem 0) is not built");
  g_f77.filt2 = *filt / 2.0;  // d2chk(2) (ppiclf_solve.f:1611)
  g_f77.proj = 0;
  g_f77.lfilt = true;
}

// ppiclf_solve.f:1520-1576
void ppiclf_solve_initgaussianfilter_(const double *filt, const double *alpha, const int *iwallm) {
  API_LOCK;
  if (!g_f77.linit) return f77_fail("InitParticle must be before InitFilter");
  if (*iwallm != 0) return f77_fail("the wall mirror (iwallm 1) is not built");
  g_f77.proj = 1;
  g_f77.width = *filt;
  g_f77.alpha = *alpha;
  const double rsig = *filt / (2.0 * sqrt(2.0 * log(2.0)));
  g_f77.filt2 = rsig * sqrt(-2.0 * log(*alpha));  // d2chk(2), the cut-off (ppiclf_solve.f:1558-1559)
  g_f77.lfilt = true;
}

// ppiclf_solve.f:339-380
void ppiclf_solve_initneighborbin_(const double *rwidth) {
  API_LOCK;
  if (!g_f77.linit) return f77_fail("InitParticle must be before InitNeighborBin");
  g_f77.nbw = *rwidth;  // ppiclf_d2chk(3), the added-mass pairs' reach (the collision search sizes its own cubes)
}

// ppiclf_comm.f:926-1000: the element nodes must be the bound region's hex
// grid, node (i,j,k) = c2v entry NODE2V[(i-1) + 2(j-1) + 4(k-1)]
// (PICL_TEMP_InitSolver.F90:538-597)
void ppiclf_comm_initoverlapmesh_(const int *ncell, const int *lx1, const int *ly1, const int *lz1,
                                  const double *xgrid, const double *ygrid, const double *zgrid) {
  API_LOCK;
  rflu_handle *h = g_f77.h;
  if (!g_f77.linit) return f77_fail("InitParticle must be before InitOverlap");
  if (!g_f77.lfilt) return f77_fail("a filter (InitBoxFilter or InitGaussianFilter) must be set before InitOverlap");
  if (*lx1 != 2 || *ly1 != 2 || *lz1 != 2) return f77_fail("LX1, LY1, LZ1 must be 2 (PPICLF_LEX/LEY/LEZ)");
  const int N = h->mesh.nCells;
  if (*ncell != N) return f77_fail("InitOverlap's ncell must be the bound region's cell count");
  if (h->c2v.size() != 8 * (size_t)N) return f77_fail("the bound region must be an all-hex grid");
  for (int e = 0; e < N; e++)
    for (int n = 0; n < 8; n++) {
      const double *xv = &h->xyz[3 * (size_t)h->c2v[8 * (size_t)e + NODE2V[n]]];
      size_t q = 8 * (size_t)e + n;
      if (xgrid[q] != xv[0] || ygrid[q] != xv[1] || zgrid[q] != xv[2])
        return f77_fail("InitOverlap's grid is not the bound region's grid (element " + std::to_string(e + 1) + ")");
    }
  picl_params p = g_f77.pp;
  p.projMode = g_f77.proj;
  p.filterWidth = g_f77.width;
  p.filterAlpha = g_f77.alpha;
  p.periodicX = g_f77.per[0];
  p.periodicY = g_f77.per[1];
  p.periodicZ = g_f77.per[2];
  p.angularPeriodic = g_f77.ang;
  p.angAngle = g_f77.angA;
  p.angXAngle = g_f77.angX;
  p.angRin = g_f77.angRin;
  p.angRout = g_f77.angRout;
  p.neighborWidth = g_f77.nbw;
  p.fluctHalfWidth = g_f77.filt2;
  if (int rc = picl_gpu_init(h, &p, g_f77.npart, g_f77.y.data(), g_f77.pc.data()))
    return f77_fail("InitOverlap: " + h->err + " (code " + std::to_string(rc) + ")");
  g_f77.fld.assign((size_t)(PICL_NFLD + 6) * N, 0.0);
  g_f77.have.assign(PICL_NFLD + 6, 0);
  g_f77.overlap = true;
  g_f77.stage = 0;
  g_f77.icnt = 0;
  g_f77.proValid = false;
}

// ppiclf_solve.f:2020-2058.  A field that no built model reads is accepted
// and dropped.  With IDW interpolation the element value is LocalInterp's
// mean of the 8 node values (ppiclf_solve.f:2363-2373); the trilinear
// evaluation takes element-constant fields (PICL_TEMP_Runge fills all 8
// nodes of a cell with its value).
void ppiclf_solve_interpfielduser_(const int *jp, const double *infld) {
  API_LOCK;
  if (!g_f77.overlap) return f77_fail("InitOverlap must come before InterpFieldUser");
  g_f77.icnt++;
  if (g_f77.icnt > 30) return f77_fail("Interpolating too many fields (PPICLF_LRP_INT 30)");
  if (*jp <= 0 || *jp > PICL_LRP) return f77_fail("Invalid particle array interp. location " + std::to_string(*jp));
  int k = -1;
  for (int q = 0; q < PICL_NFLD; q++)
    if (F77_FIELD_RPROP[q] == *jp) k = q;
  if (g_f77.h->picl->P.rot3)
    for (int q = 0; q < 6; q++)
      if (F77_ROT_RPROP[q] == *jp) k = PICL_NFLD + q;
  if (k < 0) return;
  const int N = g_f77.h->mesh.nCells;
  const bool tri = g_f77.h->picl->phys.interpMode == 1;
  double *dst = g_f77.fld.data() + (size_t)k * N;
  for (int c = 0; c < N; c++) {
    const double *v = infld + 8 * (size_t)c;
    if (tri) {
      for (int n = 1; n < 8; n++)
        if (v[n] != v[0])
          return f77_fail("trilinear interpolation takes element-constant fields (rprop " + std::to_string(*jp) +
                          ", element " + std::to_string(c + 1) + ")");
      dst[c] = v[0];
    } else {
      double s = 0.0;
      for (int n = 0; n < 8; n++) s = s + v[n];
      dst[c] = s / 8;
    }
  }
  g_f77.have[k] = 1;
}

// ppiclf_solve.f:1694-1745 with imethod 2: one RK stage of
// IntegrateRK3s_Rocflu (the stage counter cycles 1, 2, 3 over calls)
void ppiclf_solve_integrateparticle_(const int *istep, const int *iostep, const double *dt, const double *time) {
  API_LOCK;
  (void)istep; (void)iostep; (void)time;
  if (!g_f77.overlap) return f77_fail("InitOverlap must come before IntegrateParticle");
  rflu_handle *h = g_f77.h;
  PiclDev &P = h->picl->P;
  const int nf = PICL_NFLD + (P.rot3 ? 6 : 0);
  for (int k = 0; k < nf; k++)
    if (!g_f77.have[k])
      return f77_fail("IntegrateParticle: field rprop " +
                      std::to_string(k < PICL_NFLD ? F77_FIELD_RPROP[k] : F77_ROT_RPROP[k - PICL_NFLD]) +
                      " was not given by InterpFieldUser this stage");
  const int N = h->mesh.nCells;
  std::vector<double> dev((size_t)PICL_NFLD * N), rot(P.rot3 ? 6 * (size_t)N : 0);
  for (int i = 0; i < N; i++) {
    for (int k = 0; k < PICL_NFLD; k++) dev[(size_t)PICL_NFLD * i + k] = g_f77.fld[(size_t)k * N + h->perm[i]];
    for (int q = 0; q < (int)(rot.size() / N); q++)
      rot[6 * (size_t)i + q] = g_f77.fld[(size_t)(PICL_NFLD + q) * N + h->perm[i]];
  }
  g_f77.stage = g_f77.stage % 3 + 1;
  if (hipMemcpyAsync(P.fld, dev.data(), dev.size() * sizeof(double), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      (P.rot3 && hipMemcpyAsync(P.fldRot, rot.data(), rot.size() * sizeof(double), hipMemcpyHostToDevice,
                                h->stream) != hipSuccess) ||
      hipMemcpyAsync(h->d.scal + S_DTMIN, dt, sizeof(double), hipMemcpyHostToDevice, h->stream) != hipSuccess)
    return f77_fail("IntegrateParticle: upload failed");
  picl_particle_stage(h, g_f77.stage);
  if (hipGetLastError() != hipSuccess) return f77_fail("IntegrateParticle: launch failed");
  if (int rc = check_err(h)) return f77_fail("IntegrateParticle: " + h->err + " (code " + std::to_string(rc) + ")");
  g_f77.icnt = 0;
  std::fill(g_f77.have.begin(), g_f77.have.end(), 0);
  g_f77.proValid = false;
}

// ppiclf_solve.f:3630-3647: ppiclf_pro_fld(i,j,k,e,m); the projected fields
// built are PHIP, FX, FY, FZ, E (m = 1..5)
void ppiclf_solve_getprofldijkef_(const int *i, const int *j, const int *k, const int *e, const int *m, double *fld) {
  API_LOCK;
  if (!g_f77.overlap) return f77_fail("InitOverlap must come before GetProFldIJKEF");
  rflu_handle *h = g_f77.h;
  const int N = h->mesh.nCells;
  if (*i < 1 || *i > 2 || *j < 1 || *j > 2 || *k < 1 || *k > 2 || *e < 1 || *e > N)
    return f77_fail("GetProFldIJKEF: node or element out of range");
  if (*m < 1 || *m > 5)
    return f77_fail("GetProFldIJKEF: projected field " + std::to_string(*m) +
                    " is not built (PHIP, FX, FY, FZ, E are)");
  if (!g_f77.proValid) {
    g_f77.pro.resize(40 * (size_t)N);
    if (int rc = picl_gpu_get_nodes(h, g_f77.pro.data()))
      return f77_fail("GetProFldIJKEF: " + h->err + " (code " + std::to_string(rc) + ")");
    g_f77.proValid = true;
  }
  *fld = g_f77.pro[40 * (size_t)(*e - 1) + 8 * (size_t)(*m - 1) + (*i - 1) + 2 * (*j - 1) + 4 * (*k - 1)];
}

}  // extern "C"
