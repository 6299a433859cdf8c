// MI355X (gfx950) kernels and C ABI of the RocfluMP explicit residual path.
//
// Data layout in HBM (one region per handle):
//   cv, rhsSum, rhs : SoA [5][nCells] doubles (coalesced cell streams)
//   rec             : AoS [nCells][8] doubles, one 64-byte record per cell that
//                     the gather kernels fetch in one segment:
//                     order 1: (rho, rho u, rho v, rho w, rho E, p, T, c)
//                     order 2: (rho, u, v, w, p, x, y, z)  (primitive + cofg)
//   grad            : AoS [nCells][15] limited gradients (order 2)
//   flux            : AoS [nFaces+nBFaces][5] face fluxes of the stage
//   face geometry   : SoA nx, ny, nz, |S|, xc, yc, zc per face slot
//                     (interior faces first, then boundary faces by patch)
//
// Bitwise parity with the reference's FP64 arithmetic rests on three rules:
// the file is compiled with -ffp-contract=off (no FMA, like the reference);
// expressions keep the Fortran evaluation order; and every per-cell sum runs
// over the cell's faces in ascending face-slot order, which is the order the
// reference's face loop and patch loop add into that cell.  Dependent
// variables (p, T, c) and the in-place cons/prim round trips of
// CellGradientsMP are recomputed in registers from the stored state, with the
// same expressions, instead of being stored.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gauss_exp.h"

#include "../../include/rflu_gpu.h"
#include "../../include/picl_gpu.h"
#include "../../include/ppiclf_gpu.h"
#include "rflu_mesh.hpp"

#define BLK 256
#define TBIG 768  // cells per gradient tile when the tile records fit one workgroup's LDS

namespace {

struct Gas {
  double cp, g, gc;   // flux path: refCp, refGamma, R = cp - cp/g (MixtPerf_R_CpG)
  double gD, gcD;     // dependent-variable path: R = 8314.3/(8314.3/R), g = cp/(cp-R)
  double epsentr, dissFact;
  double sigma, kp, ku;  // AUSM+up
  double accel[3];
  int accelOn;
};

struct Dev {
  int nC = 0, nF = 0, nB = 0, S = 0, nBlk = 0;
  int *f2c = nullptr, *bf2c = nullptr, *bkind = nullptr, *cfOff = nullptr, *cfIdx = nullptr;
  int *bkopt = nullptr;  // bcOpt of each boundary face's patch
  int *csOff = nullptr, *csMem = nullptr, *cfNInt = nullptr;
  double *cfdx = nullptr, *cfdy = nullptr, *cfdz = nullptr;
  double *fnx = nullptr, *fny = nullptr, *fnz = nullptr, *fnm = nullptr;
  double *fcx = nullptr, *fcy = nullptr, *fcz = nullptr, *bpout = nullptr;
  int *bpatch = nullptr;   // patch of each boundary face
  double *pcon = nullptr;  // PCON derived constants per patch (inflow, farfield, injection)
  double *vol = nullptr, *epsq = nullptr, *rm = nullptr;
  double *cv = nullptr, *rhsSum = nullptr, *rhs = nullptr, *rec = nullptr, *grad = nullptr;
  double *flux = nullptr, *srad = nullptr, *dtCell = nullptr, *dvOut = nullptr, *partial = nullptr, *scal = nullptr, *cofg = nullptr;
  int *err = nullptr;
  // ppiclF coupling (null without particles): piclVF, rho_g = rho/(1-phi)
  // and its raw LSQ gradient (4 doubles per cell, the 4th unused)
  double *phi = nullptr, *rhog = nullptr, *gRaw = nullptr;
  // gradient tiles: block b owns cells [b*BLK, b*BLK+BLK); tMem lists the
  // records it stages in LDS (its own cells first, then the other stencil
  // members, ascending); sSlot gives each stencil entry's LDS slot, with bit
  // 15 marking face neighbours
  int *tOff = nullptr, *tMem = nullptr;
  unsigned short *sSlot = nullptr;
  unsigned short *cfSlot = nullptr;  // per interior cell-face entry: the LDS slot of the face neighbour
  // FP32 evaluation mode: limited gradients as 16 floats per cell (15 + pad)
  float *gradF = nullptr;
  // WENO-XYZ reconstruction (reconst 2): the unlimited LSQ gradients
  double *gradRaw = nullptr;
  // fused stages (k_fu): the records written by stage 2 and read by stage 3
  // (cells read their neighbours' records while others update theirs)
  double *recAlt = nullptr;
  // batched ensemble (nInst > 0): instance k owns device cells
  // [iOff[k], iOff[k+1]); cInst gives each cell's instance; iScal holds each
  // instance's scal[] slots (4 per instance)
  int nInst = 0;
  int *cInst = nullptr, *iOff = nullptr;
  double *iScal = nullptr;
};

// scal[] slots
enum { S_DTLOCAL = 0, S_DTMIN = 1, S_TIME = 2, S_TSR = 3 };
// err bits
enum { E_NAN = 1, E_NEG = 2, E_POSFIX = 4, E_BOUNDFIX = 8, E_INTERP = 16, E_CONTACT = 32, E_REROT = 64 };
// boundary kinds on device: 0 slip wall, 1..3 outflow (1 + bcOpt)
}  // namespace

struct PiclState;
struct rflu_handle {
  PiclState *picl = nullptr;  // particles attached by picl_gpu_init
  rflu_params prm;
  Gas gas;
  Dev d;
  rflu::MeshHost mesh;
  std::vector<double> xyz;  // the grid as given (the particle set-up reads the hex nodes)
  std::vector<int> c2v;
  hipStream_t stream = nullptr;
  int device = 0;  // the device current at creation; the handle's memory and stream live there
  std::string err;
  hipGraphExec_t graph = nullptr;
  int fuse = 0;         // stages 2 and 3 of a step as one fused kernel after the gradients (k_fu)
  int gradVariant = 4;  // 4: LDS tiles (default when they fit), 2: one thread per cell, 0: plain, 1/3: cooperative
  size_t tileLds = 0;   // LDS bytes per gradient tile (0: tiles not built)
  int tileCells = BLK;  // cells per gradient tile (= threads per tile workgroup)
  int nTiles = 0;
  size_t ldsMax = 65536;
  std::vector<int> perm, inv;  // device cell i holds reference cell perm[i]
  std::vector<double> stage;   // host staging for (un)permuting cell arrays
  bool profiling = false;
  struct Prof { double ms = 0; int n = 0; };
  std::map<std::string, Prof> prof;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> evPool;
  size_t devBytes = 0;  // device memory allocated for this handle
  std::vector<int> instCell, instFace, instBFace;  // batch: per-instance offsets in the union (reference numbering)
};

// Text of the last error of a call that had no handle (creation, memory
// queries), per calling thread: concurrent set-ups never share it.
static thread_local std::string g_err;

// One process-wide lock orders the runtime calls that may not overlap a step
// graph's capture on another thread: allocation and freeing, copies and
// memsets, stream, event and graph creation or destruction, and every
// synchronisation.  capture_step holds it from hipStreamBeginCapture to the
// instantiation.  Kernel launches and graph replays do not take it, so the
// hot path never waits on it.  Recursive: a locked entry point may call
// another (picl_gpu_init allocates through dalloc).
static std::recursive_mutex g_api;
#define API_LOCK std::lock_guard<std::recursive_mutex> api_lock_(g_api)

#define HIPCHK(h, x)                                                                \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      (h)->err = std::string(#x) + ": " + hipGetErrorString(e_);                    \
      return RFLU_GPU_EHIP;                                                         \
    }                                                                               \
  } while (0)

// ---------------------------------------------------------------------------
// Device helpers (thermodynamics: libfloflu/MixtPerf_*.F90)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_rec(const double *__restrict__ rec, int i, double r[8]) {
  const double2 *p = reinterpret_cast<const double2 *>(rec + 8 * (size_t)i);
  double2 a = p[0], b = p[1], c = p[2], e = p[3];
  r[0] = a.x; r[1] = a.y; r[2] = b.x; r[3] = b.y; r[4] = c.x; r[5] = c.y; r[6] = e.x; r[7] = e.y;
}
__device__ __forceinline__ void store_rec(double *__restrict__ rec, int i, const double r[8]) {
  double2 *p = reinterpret_cast<double2 *>(rec + 8 * (size_t)i);
  p[0] = make_double2(r[0], r[1]);
  p[1] = make_double2(r[2], r[3]);
  p[2] = make_double2(r[4], r[5]);
  p[3] = make_double2(r[6], r[7]);
}

// libflu/RFLU_SetDependentVars.F90:256-287 (TCPERF, ksg = 0).  With
// particles the gas-phase density rho/(1-phi) replaces rho in p, T, c
// (:276-281); Eo and |u|^2 still use rho.
struct DV { double p, T, c; };
__device__ __forceinline__ DV dep_vars(const Gas &G, double rho, double m1, double m2, double m3, double E,
                                       const double *phi = nullptr, int i = 0) {
  double ir = 1.0 / rho;
  double Eo = ir * E;
  double Vm2 = ir * ir * (m1 * m1 + m2 * m2 + m3 * m3);
  double r = phi ? rho / (1.0 - phi[i]) : rho;
  DV d;
  d.p = (G.gD - 1.0) * r * (Eo - 0.5 * Vm2 - 0.0);
  d.T = d.p / (r * G.gcD);
  d.c = sqrt(G.gD * G.gcD * d.T);
  return d;
}

// libfloflu/MixtPerf_E.F90: Eo = P/(D(G-1)) + |u|^2/2 + ksg
__device__ __forceinline__ double eo_dgpuvw(double D, double G, double P, double U, double V, double W) {
  return P / (D * (G - 1.0)) + 0.5 * (U * U + V * V + W * W) + 0.0;
}

// Record of a cell from its conserved state (and dv as computed above).
__device__ __forceinline__ void make_rec(int order, double rho, double m1, double m2, double m3,
                                         double E, const DV &dv, double x, double y, double z,
                                         double r[8]) {
  if (order == 1) {
    r[0] = rho; r[1] = m1; r[2] = m2; r[3] = m3; r[4] = E; r[5] = dv.p; r[6] = dv.T; r[7] = dv.c;
  } else {
    double ir = 1.0 / rho;  // modflu/RFLU_ModConvertCv.F90:221-231
    r[0] = rho; r[1] = ir * m1; r[2] = ir * m2; r[3] = ir * m3; r[4] = dv.p; r[5] = x; r[6] = y; r[7] = z;
  }
}

// Speed of sound of a cell from its record (and its phi with particles).
__device__ __forceinline__ double rec_c(const Gas &G, int order, const double r[8], double phi = 0.0) {
  if (order == 1) return r[7];
  double T = r[4] / (r[0] / (1.0 - phi) * G.gcD);
  return sqrt(G.gD * G.gcD * T);
}
// Velocity the reference's dt loop uses: cv(m)*(1/cv(rho)) == cons2prim u.
__device__ __forceinline__ void rec_u(int order, const double r[8], double &u, double &v, double &w) {
  if (order == 1) {
    double ir = 1.0 / r[0];
    u = r[1] * ir; v = r[2] * ir; w = r[3] * ir;
  } else {
    u = r[1]; v = r[2]; w = r[3];
  }
}

// Workgroups are dealt round-robin to the 8 XCDs (hardware block b runs on
// XCD b % 8).  Remap so each XCD walks one contiguous range of cells/faces:
// neighbouring cells then share that XCD's L2 instead of being spread over
// all eight.  Bijective on [0, gridDim.x) for any grid size.
__device__ int g_xcdRemap = 1;
__device__ __forceinline__ int xcd_block() {
  if (!g_xcdRemap) return blockIdx.x;
  const int nx = 8;
  int nb = gridDim.x, b = blockIdx.x;
  int xcd = b % nx, loc = b / nx, q = nb / nx, r = nb % nx;
  return xcd * q + min(xcd, r) + loc;
}

// ---------------------------------------------------------------------------
// Records from cv (RFLU_SetVarsWrapper at start-up)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLK) k_records(Dev d, Gas G, int order, int i0 = 0, int n = -1) {
  int i = i0 + blockIdx.x * BLK + threadIdx.x;
  if (i >= (n < 0 ? d.nC : i0 + n)) return;
  int N = d.nC;
  double rho = d.cv[i], m1 = d.cv[N + i], m2 = d.cv[2 * N + i], m3 = d.cv[3 * N + i], E = d.cv[4 * N + i];
  DV dv = dep_vars(G, rho, m1, m2, m3, E, d.phi, i);
  double r[8];
  make_rec(order, rho, m1, m2, m3, E, dv, d.cofg[i], d.cofg[N + i], d.cofg[2 * N + i], r);
  store_rec(d.rec, i, r);
  if (d.rhog) d.rhog[i] = rho / (1.0 - d.phi[i]);
}

// ---------------------------------------------------------------------------
// dt: RFLU_TimeStepInviscid (per cell, faces in reference order) + MINVAL +
// RFLU_MinimumTimeStep in the last block to finish.
// ---------------------------------------------------------------------------
struct DtPar { double cfl, dtImposed, dtMinLimit, dTimeSystem; int finalize; };

__device__ double minimum_time_step(const DtPar &P, double dtLocal, double t, double tsr) {
  // rocflu/RFLU_MinimumTimeStep.F90:125-146,235-241
  double cfl = P.cfl < 2.0 ? P.cfl : 2.0;
  double dmin = cfl * dtLocal;
  double wI = 0.0, wC = 0.0;
  if (t <= 1.0e-8) wI = 1.0; else wC = 1.0;
  dmin = wI / P.dtImposed + wC / dmin;
  dmin = 1.0 / dmin;
  if (dmin <= 0.0) dmin = 6.0e-12;
  dmin = fmax(dmin, P.dtMinLimit);
  if (P.dTimeSystem > 0.0 && tsr + dmin > P.dTimeSystem) dmin = P.dTimeSystem - tsr;
  return dmin;
}

// Spectral radius of one face (rocflu/RFLU_TimeStepInviscid.F90:165-197 for
// interior faces, :264-285 for boundary faces; static grid, fs = 0).  The
// expression is symmetric in the two cells, as in the reference.
__device__ __forceinline__ double face_srad(const Gas &G, int order, const double L[8], const double *R,
                                            double nx, double ny, double nz, double nm,
                                            double phiL = 0.0, double phiR = 0.0) {
  double u1, v1, w1;
  rec_u(order, L, u1, v1, w1);
  double a1 = rec_c(G, order, L, phiL);
  if (R) {
    double u2, v2, w2;
    rec_u(order, R, u2, v2, w2);
    double a2 = rec_c(G, order, R, phiR);
    double vc = 0.5 * ((u1 + u2) * nx + (v1 + v2) * ny + (w1 + w2) * nz);
    double as = 0.5 * (a1 + a2);
    return (fabs(vc - 0.0) + as) * nm;
  }
  double vc = u1 * nx + v1 * ny + w1 * nz;
  return (fabs(vc - 0.0) + a1) * nm;
}

__global__ void __launch_bounds__(BLK) k_srad(Dev d, Gas G, int order) {
  int s = xcd_block() * BLK + threadIdx.x;
  if (s >= d.S) return;
  double L[8], R[8];
  if (s < d.nF) {
    int c1 = d.f2c[2 * s], c2 = d.f2c[2 * s + 1];
    load_rec(d.rec, c1, L);
    load_rec(d.rec, c2, R);
    double p1 = d.phi ? d.phi[c1] : 0.0, p2 = d.phi ? d.phi[c2] : 0.0;
    d.srad[s] = face_srad(G, order, L, R, d.fnx[s], d.fny[s], d.fnz[s], d.fnm[s], p1, p2);
  } else {
    int c1 = d.bf2c[s - d.nF];
    load_rec(d.rec, c1, L);
    d.srad[s] = face_srad(G, order, L, nullptr, d.fnx[s], d.fny[s], d.fnz[s], d.fnm[s], d.phi ? d.phi[c1] : 0.0);
  }
}

// dt = vol / (sum of face spectral radii in face order) per cell, and each
// workgroup's minimum.  The minimum over workgroups is taken by k_dtfinal in a
// separate launch: the kernel boundary publishes the partial minima across
// XCDs without a device-scope fence (an L2 write-back) in every workgroup.
__global__ void __launch_bounds__(BLK) k_dtmin(Dev d, int storeCell) {
  __shared__ double sm[BLK / 64];
  int i = xcd_block() * BLK + threadIdx.x;
  double dt = __builtin_inf();
  if (i < d.nC) {
    double sum = 0.0;
    for (int e = d.cfOff[i]; e < d.cfOff[i + 1]; e++) sum = sum + d.srad[d.cfIdx[e] >> 1];
    dt = d.vol[i] / sum;
    if (storeCell) d.dtCell[i] = dt;
  }
  for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off, 64));
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sm[wv] = dt;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sm[0];
    for (int q = 1; q < BLK / 64; q++) m = fmin(m, sm[q]);
    d.partial[blockIdx.x] = m;
  }
}

// MINVAL over the workgroup minima, then RFLU_MinimumTimeStep (one workgroup).
__global__ void __launch_bounds__(1024) k_dtfinal(Dev d, DtPar P, int nPart) {
  __shared__ double sm[1024 / 64];
  double m = __builtin_inf();
  for (int q = threadIdx.x; q < nPart; q += 1024) m = fmin(m, d.partial[q]);
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sm[wv] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = sm[0];
    for (int q = 1; q < 1024 / 64; q++) mm = fmin(mm, sm[q]);
    d.scal[S_DTLOCAL] = mm;
    if (P.finalize) d.scal[S_DTMIN] = minimum_time_step(P, mm, d.scal[S_TIME], d.scal[S_TSR]);
  }
}

// Batched ensemble: each instance's MINVAL over its own cells (one workgroup
// per instance over the per-cell dt of k_dtmin), then its
// RFLU_MinimumTimeStep with its own time.
__global__ void __launch_bounds__(1024) k_dt_batch(Dev d, DtPar P) {
  __shared__ double sm[1024 / 64];
  int k = blockIdx.x;
  double m = __builtin_inf();
  for (int i = d.iOff[k] + threadIdx.x; i < d.iOff[k + 1]; i += 1024) m = fmin(m, d.dtCell[i]);
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sm[wv] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = sm[0];
    for (int q = 1; q < 1024 / 64; q++) mm = fmin(mm, sm[q]);
    double *sc = d.iScal + 4 * (size_t)k;
    sc[S_DTLOCAL] = mm;
    sc[S_DTMIN] = minimum_time_step(P, mm, sc[S_TIME], sc[S_TSR]);
  }
}

__global__ void k_min_time_step(Dev d, DtPar P) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    d.scal[S_DTMIN] = minimum_time_step(P, d.scal[S_DTLOCAL], d.scal[S_TIME], d.scal[S_TSR]);
}

__global__ void k_advance_time(Dev d) {
  if (d.nInst) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < d.nInst) {
      double *sc = d.iScal + 4 * (size_t)k;
      sc[S_TIME] = sc[S_TIME] + sc[S_DTMIN];
      sc[S_TSR] = sc[S_TSR] + sc[S_DTMIN];
    }
    return;
  }
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    d.scal[S_TIME] = d.scal[S_TIME] + d.scal[S_DTMIN];
    d.scal[S_TSR] = d.scal[S_TSR] + d.scal[S_DTMIN];
  }
}

// ---------------------------------------------------------------------------
// Gradients + Venkatakrishnan limiter (order 2), one thread per cell.
// modflu/RFLU_ModDifferentiationCells.F90:955-1174 (weights from the R-factor
// each call, members in stencil order) and modflu/RFLU_ModLimiters.F90:339-532
// (min/max over interior-face neighbours, one limiter term per interior face).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLK) k_grad(Dev d, int venkat, int guard) {
  int i = xcd_block() * BLK + threadIdx.x;
  if (i >= d.nC) return;
  int N = d.nC;
  double ri[8];
  load_rec(d.rec, i, ri);
  double r11 = d.rm[i], r12 = d.rm[N + i], r22 = d.rm[2 * N + i];
  double r13 = d.rm[3 * N + i], r23 = d.rm[4 * N + i], r33 = d.rm[5 * N + i];
  double c11 = 1.0 / r11, c22 = 1.0 / r22, c33 = 1.0 / r33;
  double c12 = -(c11 * r12);
  double c13 = -(c11 * r13 + c12 * c22 * r23);
  double c23 = -(c22 * r23);
  double c1111 = c11 * c11, c2222 = c22 * c22, c3333 = c33 * c33;
  double g[15];
#pragma unroll
  for (int q = 0; q < 15; q++) g[q] = 0.0;
  int sb = d.csOff[i], se = d.csOff[i + 1];
  for (int s = sb; s < se; s++) {
    int j = (int)((unsigned)d.csMem[s] & 0x7fffffffu);
    double rj[8];
    load_rec(d.rec, j, rj);
    double dx = rj[5] - ri[5], dy = rj[6] - ri[6], dz = rj[7] - ri[7];
    double term = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    dx = term * dx; dy = term * dy; dz = term * dz;
    double term1 = c1111 * (dx);
    double term2 = c2222 * (dy + c12 * dx);
    double term3 = c3333 * (dz + c23 * dy + c13 * dx);
    double wtx = term * (term1 + c12 * term2 + c13 * term3);
    double wty = term * (term2 + c23 * term3);
    double wtz = term * (term3);
#pragma unroll
    for (int k = 0; k < 5; k++) {
      double dVar = rj[k] - ri[k];
      g[3 * k] = g[3 * k] + wtx * dVar;
      g[3 * k + 1] = g[3 * k + 1] + wty * dVar;
      g[3 * k + 2] = g[3 * k + 2] + wtz * dVar;
    }
  }
  if (venkat) {
    const double TINYP = 1.0e-12, TINYN = guard ? -1.0e-12 : 1.0e-12;
    double vmin[5], vmax[5], lim[5];
#pragma unroll
    for (int k = 0; k < 5; k++) { vmin[k] = ri[k]; vmax[k] = ri[k]; lim[k] = 1.0; }
    int eb = d.cfOff[i], ee = d.cfOff[i + 1];
    for (int e = eb; e < ee; e++) {
      int code = d.cfIdx[e], s = code >> 1;
      if (s >= d.nF) break;  // boundary slots come last
      int o = d.f2c[2 * s + ((code & 1) ? 0 : 1)];
      double ro[8];
      load_rec(d.rec, o, ro);
#pragma unroll
      for (int k = 0; k < 5; k++) { vmin[k] = fmin(vmin[k], ro[k]); vmax[k] = fmax(vmax[k], ro[k]); }
    }
    double epsq = d.epsq[i];
    for (int e = eb; e < ee; e++) {
      int s = d.cfIdx[e] >> 1;
      if (s >= d.nF) break;
      double dx1 = ri[5] - d.fcx[s], dy1 = ri[6] - d.fcy[s], dz1 = ri[7] - d.fcz[s];
#pragma unroll
      for (int k = 0; k < 5; k++) {
        double d21 = g[3 * k] * dx1 + g[3 * k + 1] * dy1 + g[3 * k + 2] * dz1;
        if (d21 > 0.0) {
          double d1 = vmax[k] - ri[k];
          double numer = (d1 * d1 + epsq) * d21 + 2.0 * d21 * d21 * d1;
          double denom = d21 * (d1 * d1 + 2.0 * d21 * d21 + d1 * d21 + epsq);
          lim[k] = fmin(numer / (denom + (denom >= 0.0 ? TINYP : TINYN)), lim[k]);
        } else if (d21 < 0.0) {
          double d1 = vmin[k] - ri[k];
          double numer = (d1 * d1 + epsq) * d21 + 2.0 * d21 * d21 * d1;
          double denom = d21 * (d1 * d1 + 2.0 * d21 * d21 + d1 * d21 + epsq);
          lim[k] = fmin(numer / (denom + (denom >= 0.0 ? TINYP : TINYN)), lim[k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 5; k++) {
      g[3 * k] = lim[k] * g[3 * k];
      g[3 * k + 1] = lim[k] * g[3 * k + 1];
      g[3 * k + 2] = lim[k] * g[3 * k + 2];
    }
  }
  double *o = d.grad + 15 * (size_t)i;
#pragma unroll
  for (int q = 0; q < 15; q++) o[q] = g[q];
}

// Latency-trimmed variant (default).  Static geometry is prepared at set-up:
// stencil entries carry a flag bit (CS_FACE) when the member is also an
// interior-face neighbour, so the limiter's min/max comes out of the member
// loop with no second gather; and each interior cell-face entry stores
// cofg(cell) - fc(face), the same difference the reference forms, so the
// limiter's face loop reads one contiguous stream.  The next member's record is
// fetched while the current one is processed.
#define CS_FACE 0x80000000u
#define CS_IDX 0x7fffffffu
// PICL = 1 also forms the LSQ gradient of rho_g = rho/(1-phi) over the same
// stencil (CellGradientsMP.F90:375-378, before its WENO and positivity
// passes in k_rhog).
template <int MINW, int PICL>
__global__ void __launch_bounds__(BLK, MINW) k_grad2(Dev d, int venkat, int guard) {
  int i = xcd_block() * BLK + threadIdx.x;
  if (i >= d.nC) return;
  int N = d.nC;
  double ri[8];
  load_rec(d.rec, i, ri);
  double r11 = d.rm[i], r12 = d.rm[N + i], r22 = d.rm[2 * N + i];
  double r13 = d.rm[3 * N + i], r23 = d.rm[4 * N + i], r33 = d.rm[5 * N + i];
  double c11 = 1.0 / r11, c22 = 1.0 / r22, c33 = 1.0 / r33;
  double c12 = -(c11 * r12);
  double c13 = -(c11 * r13 + c12 * c22 * r23);
  double c23 = -(c22 * r23);
  double c1111 = c11 * c11, c2222 = c22 * c22, c3333 = c33 * c33;
  double g[15], vmin[5], vmax[5];
#pragma unroll
  for (int q = 0; q < 15; q++) g[q] = 0.0;
#pragma unroll
  for (int k = 0; k < 5; k++) { vmin[k] = ri[k]; vmax[k] = ri[k]; }
  const unsigned *mem = reinterpret_cast<const unsigned *>(d.csMem);
  int sb = d.csOff[i], se = d.csOff[i + 1];
  unsigned jn = mem[sb];
  double rn[8];
  load_rec(d.rec, (int)(jn & CS_IDX), rn);
  double gq0 = 0.0, gq1 = 0.0, gq2 = 0.0, qi = 0.0;
  if (PICL) qi = d.rhog[i];
  for (int s = sb; s < se; s++) {
    double rj[8];
#pragma unroll
    for (int q = 0; q < 8; q++) rj[q] = rn[q];
    unsigned j = jn;
    double qj = 0.0;
    if (PICL) qj = d.rhog[j & CS_IDX];
    if (s + 1 < se) {
      jn = mem[s + 1];
      load_rec(d.rec, (int)(jn & CS_IDX), rn);
    }
    double dx = rj[5] - ri[5], dy = rj[6] - ri[6], dz = rj[7] - ri[7];
    double term = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    dx = term * dx; dy = term * dy; dz = term * dz;
    double term1 = c1111 * (dx);
    double term2 = c2222 * (dy + c12 * dx);
    double term3 = c3333 * (dz + c23 * dy + c13 * dx);
    double wtx = term * (term1 + c12 * term2 + c13 * term3);
    double wty = term * (term2 + c23 * term3);
    double wtz = term * (term3);
#pragma unroll
    for (int k = 0; k < 5; k++) {
      double dVar = rj[k] - ri[k];
      g[3 * k] = g[3 * k] + wtx * dVar;
      g[3 * k + 1] = g[3 * k + 1] + wty * dVar;
      g[3 * k + 2] = g[3 * k + 2] + wtz * dVar;
    }
    if (PICL) {
      double dq = qj - qi;
      gq0 = gq0 + wtx * dq;
      gq1 = gq1 + wty * dq;
      gq2 = gq2 + wtz * dq;
    }
    if (j & CS_FACE) {
#pragma unroll
      for (int k = 0; k < 5; k++) { vmin[k] = fmin(vmin[k], rj[k]); vmax[k] = fmax(vmax[k], rj[k]); }
    }
  }
  if (venkat) {
    const double TINYP = 1.0e-12, TINYN = guard ? -1.0e-12 : 1.0e-12;
    double lim[5];
#pragma unroll
    for (int k = 0; k < 5; k++) lim[k] = 1.0;
    double epsq = d.epsq[i];
    int eb = d.cfOff[i], ee = eb + d.cfNInt[i];
    for (int e = eb; e < ee; e++) {
      double dx1 = d.cfdx[e], dy1 = d.cfdy[e], dz1 = d.cfdz[e];
#pragma unroll
      for (int k = 0; k < 5; k++) {
        double d21 = g[3 * k] * dx1 + g[3 * k + 1] * dy1 + g[3 * k + 2] * dz1;
        if (d21 > 0.0) {
          double d1 = vmax[k] - ri[k];
          double numer = (d1 * d1 + epsq) * d21 + 2.0 * d21 * d21 * d1;
          double denom = d21 * (d1 * d1 + 2.0 * d21 * d21 + d1 * d21 + epsq);
          lim[k] = fmin(numer / (denom + (denom >= 0.0 ? TINYP : TINYN)), lim[k]);
        } else if (d21 < 0.0) {
          double d1 = vmin[k] - ri[k];
          double numer = (d1 * d1 + epsq) * d21 + 2.0 * d21 * d21 * d1;
          double denom = d21 * (d1 * d1 + 2.0 * d21 * d21 + d1 * d21 + epsq);
          lim[k] = fmin(numer / (denom + (denom >= 0.0 ? TINYP : TINYN)), lim[k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 5; k++) {
      g[3 * k] = lim[k] * g[3 * k];
      g[3 * k + 1] = lim[k] * g[3 * k + 1];
      g[3 * k + 2] = lim[k] * g[3 * k + 2];
    }
  }
  double *o = d.grad + 15 * (size_t)i;
#pragma unroll
  for (int q = 0; q < 15; q++) o[q] = g[q];
  if (PICL) {
    double2 *gr = reinterpret_cast<double2 *>(d.gRaw + 4 * (size_t)i);
    gr[0] = make_double2(gq0, gq1);
    gr[1] = make_double2(gq2, 0.0);
  }
}

// Tiled variant (default): a workgroup first stages, once, the records of
// every cell its 256 cells' stencils touch (Morton-contiguous cells share most
// of their members: ~3.3 records per cell instead of ~40 gathers), then each
// lane runs k_grad2's member loop on LDS.  Same operations, same order.
#ifndef GU
#define GU 4
#endif
// LDS records of the gradient tiles, one 64-byte record per slot.  (Rotating
// each slot's 16-byte chunks by (u >> 2) & 3 cut the LDS bank conflicts of
// the member reads 2.7x but added 12 % VALU instructions to a VALU-bound
// loop: measured 2 % slower.)
__device__ __forceinline__ void lds_rec_store(double2 *s, int u, double2 a, double2 b, double2 c, double2 e) {
  s[4 * u] = a;
  s[4 * u + 1] = b;
  s[4 * u + 2] = c;
  s[4 * u + 3] = e;
}
__device__ __forceinline__ void lds_rec_load(const double2 *s, int u, double q[8]) {
  double2 a = s[4 * u], b = s[4 * u + 1], c = s[4 * u + 2], e = s[4 * u + 3];
  q[0] = a.x; q[1] = a.y; q[2] = b.x; q[3] = b.y; q[4] = c.x; q[5] = c.y; q[6] = e.x; q[7] = e.y;
}

// PICL = 1 also stages rho_g = rho/(1-phi) per slot (after the records) and
// forms its LSQ gradient (CellGradientsMP.F90:375-378), as k_grad2<.,1>.
template <int PICL, int TB>
__device__ __forceinline__ void grad_tile_body(Dev d, int venkat, int guard, int nUmax) {
  extern __shared__ double2 sRec[];  // [nU][4]: one 64-byte record per slot; then rho_g[nUmax]
  double *sQ = reinterpret_cast<double *>(sRec + 4 * (size_t)nUmax);
  int tile = xcd_block();
  int u0 = d.tOff[tile], nU = d.tOff[tile + 1] - u0;
  for (int u = threadIdx.x; u < nU; u += TB) {
    int c0 = d.tMem[u0 + u];
    const double2 *g = reinterpret_cast<const double2 *>(d.rec + 8 * (size_t)c0);
    lds_rec_store(sRec, u, g[0], g[1], g[2], g[3]);
    if (PICL) sQ[u] = d.rhog[c0];
  }
  __syncthreads();
  int i = tile * TB + threadIdx.x;
  if (i >= d.nC) return;
  int N = d.nC;
  double ri[8];
  lds_rec_load(sRec, threadIdx.x, ri);
  double r11 = d.rm[i], r12 = d.rm[N + i], r22 = d.rm[2 * N + i];
  double r13 = d.rm[3 * N + i], r23 = d.rm[4 * N + i], r33 = d.rm[5 * N + i];
  double c11 = 1.0 / r11, c22 = 1.0 / r22, c33 = 1.0 / r33;
  double c12 = -(c11 * r12);
  double c13 = -(c11 * r13 + c12 * c22 * r23);
  double c23 = -(c22 * r23);
  double c1111 = c11 * c11, c2222 = c22 * c22, c3333 = c33 * c33;
  double g[15], vmin[5], vmax[5];
#pragma unroll
  for (int q = 0; q < 15; q++) g[q] = 0.0;
#pragma unroll
  for (int k = 0; k < 5; k++) { vmin[k] = ri[k]; vmax[k] = ri[k]; }
  int sb = d.csOff[i], se = d.csOff[i + 1];
  // GU members per iteration: their weights (sqrt, division) are computed
  // independently, then added in stencil order
  double gq0 = 0.0, gq1 = 0.0, gq2 = 0.0, qi = 0.0;
  if (PICL) qi = sQ[threadIdx.x];
  int s = sb;
  for (; s + GU - 1 < se; s += GU) {
    unsigned jj[GU];
    double rr[GU][8], ww[GU][3];
#pragma unroll
    for (int u = 0; u < GU; u++) {
      jj[u] = d.sSlot[s + u];
      lds_rec_load(sRec, (int)(jj[u] & 0x7fffu), rr[u]);
    }
#pragma unroll
    for (int u = 0; u < GU; u++) {
      double dx = rr[u][5] - ri[5], dy = rr[u][6] - ri[6], dz = rr[u][7] - ri[7];
      double term = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
      dx = term * dx; dy = term * dy; dz = term * dz;
      double term1 = c1111 * (dx);
      double term2 = c2222 * (dy + c12 * dx);
      double term3 = c3333 * (dz + c23 * dy + c13 * dx);
      ww[u][0] = term * (term1 + c12 * term2 + c13 * term3);
      ww[u][1] = term * (term2 + c23 * term3);
      ww[u][2] = term * (term3);
    }
#pragma unroll
    for (int u = 0; u < GU; u++) {
#pragma unroll
      for (int k = 0; k < 5; k++) {
        double dVar = rr[u][k] - ri[k];
        g[3 * k] = g[3 * k] + ww[u][0] * dVar;
        g[3 * k + 1] = g[3 * k + 1] + ww[u][1] * dVar;
        g[3 * k + 2] = g[3 * k + 2] + ww[u][2] * dVar;
      }
      if (PICL) {
        double dq = sQ[jj[u] & 0x7fffu] - qi;
        gq0 = gq0 + ww[u][0] * dq;
        gq1 = gq1 + ww[u][1] * dq;
        gq2 = gq2 + ww[u][2] * dq;
      }
    }
  }
  for (; s < se; s++) {
    unsigned j = d.sSlot[s];
    double rj[8];
    lds_rec_load(sRec, (int)(j & 0x7fffu), rj);
    double dx = rj[5] - ri[5], dy = rj[6] - ri[6], dz = rj[7] - ri[7];
    double term = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    dx = term * dx; dy = term * dy; dz = term * dz;
    double term1 = c1111 * (dx);
    double term2 = c2222 * (dy + c12 * dx);
    double term3 = c3333 * (dz + c23 * dy + c13 * dx);
    double wtx = term * (term1 + c12 * term2 + c13 * term3);
    double wty = term * (term2 + c23 * term3);
    double wtz = term * (term3);
#pragma unroll
    for (int k = 0; k < 5; k++) {
      double dVar = rj[k] - ri[k];
      g[3 * k] = g[3 * k] + wtx * dVar;
      g[3 * k + 1] = g[3 * k + 1] + wty * dVar;
      g[3 * k + 2] = g[3 * k + 2] + wtz * dVar;
    }
    if (PICL) {
      double dq = sQ[j & 0x7fffu] - qi;
      gq0 = gq0 + wtx * dq;
      gq1 = gq1 + wty * dq;
      gq2 = gq2 + wtz * dq;
    }
  }
  if (venkat) {
    const double TINYP = 1.0e-12, TINYN = guard ? -1.0e-12 : 1.0e-12;
    double lim[5];
#pragma unroll
    for (int k = 0; k < 5; k++) lim[k] = 1.0;
    double epsq = d.epsq[i];
    int eb = d.cfOff[i], ee = eb + d.cfNInt[i];
    // min/max over the face neighbours (RFLU_ModLimiters.F90:428-438; order-free),
    // their records read from the tile's LDS slots
    for (int e = eb; e < ee; e++) {
      double rn[8];
      lds_rec_load(sRec, (int)d.cfSlot[e], rn);
#pragma unroll
      for (int k = 0; k < 5; k++) { vmin[k] = fmin(vmin[k], rn[k]); vmax[k] = fmax(vmax[k], rn[k]); }
    }
    for (int e = eb; e < ee; e++) {
      double dx1 = d.cfdx[e], dy1 = d.cfdy[e], dz1 = d.cfdz[e];
#pragma unroll
      for (int k = 0; k < 5; k++) {
        double d21 = g[3 * k] * dx1 + g[3 * k + 1] * dy1 + g[3 * k + 2] * dz1;
        if (d21 > 0.0) {
          double d1 = vmax[k] - ri[k];
          double numer = (d1 * d1 + epsq) * d21 + 2.0 * d21 * d21 * d1;
          double denom = d21 * (d1 * d1 + 2.0 * d21 * d21 + d1 * d21 + epsq);
          lim[k] = fmin(numer / (denom + (denom >= 0.0 ? TINYP : TINYN)), lim[k]);
        } else if (d21 < 0.0) {
          double d1 = vmin[k] - ri[k];
          double numer = (d1 * d1 + epsq) * d21 + 2.0 * d21 * d21 * d1;
          double denom = d21 * (d1 * d1 + 2.0 * d21 * d21 + d1 * d21 + epsq);
          lim[k] = fmin(numer / (denom + (denom >= 0.0 ? TINYP : TINYN)), lim[k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 5; k++) {
      g[3 * k] = lim[k] * g[3 * k];
      g[3 * k + 1] = lim[k] * g[3 * k + 1];
      g[3 * k + 2] = lim[k] * g[3 * k + 2];
    }
  }
  double *o = d.grad + 15 * (size_t)i;
#pragma unroll
  for (int q = 0; q < 15; q++) o[q] = g[q];
  if (PICL) {
    double2 *gr = reinterpret_cast<double2 *>(d.gRaw + 4 * (size_t)i);
    gr[0] = make_double2(gq0, gq1);
    gr[1] = make_double2(gq2, 0.0);
  }
}
template <int TB>
__global__ void __launch_bounds__(TB) k_grad_tile_gas(Dev d, int venkat, int guard, int nUmax) {
  grad_tile_body<0, TB>(d, venkat, guard, nUmax);
}
template <int TB>
__global__ void __launch_bounds__(TB) k_grad_tile_picl(Dev d, int venkat, int guard, int nUmax) {
  grad_tile_body<1, TB>(d, venkat, guard, nUmax);
}

// FP32 evaluation of the same gradient (fluxPrecision 1): the tile's records
// are staged in LDS with the state in FP32 and the centroid in FP64 (48 bytes
// a member); centroid differences are taken in FP64 and rounded once, so a
// cell's result does not depend on which tile it falls in.  LSQ weights,
// sums and the limiter run in FP32, GU members per iteration.
struct __align__(16) SlotF {
  float4 s;          // rho, u, v, w
  float p, pad;
  double x;
  double2 yz;
};
template <int TB>
__global__ void __launch_bounds__(TB) k_grad_tile_f32(Dev d, int venkat, int guard) {
  extern __shared__ SlotF sRecF[];
  int tile = xcd_block();
  int u0 = d.tOff[tile], nU = d.tOff[tile + 1] - u0;
  for (int u = threadIdx.x; u < nU; u += TB) {
    double r[8];
    load_rec(d.rec, d.tMem[u0 + u], r);
    SlotF t;
    t.s = make_float4((float)r[0], (float)r[1], (float)r[2], (float)r[3]);
    t.p = (float)r[4];
    t.pad = 0.0f;
    t.x = r[5];
    t.yz = make_double2(r[6], r[7]);
    sRecF[u] = t;
  }
  __syncthreads();
  int i = tile * TB + threadIdx.x;
  if (i >= d.nC) return;
  int N = d.nC;
  const SlotF own = sRecF[threadIdx.x];
  float ri[5] = {own.s.x, own.s.y, own.s.z, own.s.w, own.p};
  const double xi = own.x, yi = own.yz.x, zi = own.yz.y;
  float r11 = (float)d.rm[i], r12 = (float)d.rm[N + i], r22 = (float)d.rm[2 * N + i];
  float r13 = (float)d.rm[3 * N + i], r23 = (float)d.rm[4 * N + i], r33 = (float)d.rm[5 * N + i];
  float c11 = 1.0f / r11, c22 = 1.0f / r22, c33 = 1.0f / r33;
  float c12 = -(c11 * r12);
  float c13 = -(c11 * r13 + c12 * c22 * r23);
  float c23 = -(c22 * r23);
  float c1111 = c11 * c11, c2222 = c22 * c22, c3333 = c33 * c33;
  float g[15], dmin[5], dmax[5];
#pragma unroll
  for (int q = 0; q < 15; q++) g[q] = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) { dmin[k] = 0.0f; dmax[k] = 0.0f; }
  int s = d.csOff[i], se = d.csOff[i + 1];
  auto member = [&](unsigned j) {
    const SlotF &m = sRecF[j & 0x7fffu];
    float4 a = m.s;
    float bp = m.p;
    double2 yz = m.yz;
    float dx = (float)(m.x - xi), dy = (float)(yz.x - yi), dz = (float)(yz.y - zi);
    float term = rsqrtf(dx * dx + dy * dy + dz * dz);
    dx = term * dx; dy = term * dy; dz = term * dz;
    float term1 = c1111 * dx;
    float term2 = c2222 * (dy + c12 * dx);
    float term3 = c3333 * (dz + c23 * dy + c13 * dx);
    float wtx = term * (term1 + c12 * term2 + c13 * term3);
    float wty = term * (term2 + c23 * term3);
    float wtz = term * term3;
    float dv[5] = {a.x - ri[0], a.y - ri[1], a.z - ri[2], a.w - ri[3], bp - ri[4]};
#pragma unroll
    for (int k = 0; k < 5; k++) {
      g[3 * k] += wtx * dv[k];
      g[3 * k + 1] += wty * dv[k];
      g[3 * k + 2] += wtz * dv[k];
    }
  };
  for (; s + GU - 1 < se; s += GU) {
    unsigned jj[GU];
#pragma unroll
    for (int u = 0; u < GU; u++) jj[u] = d.sSlot[s + u];
#pragma unroll
    for (int u = 0; u < GU; u++) member(jj[u]);
  }
  for (; s < se; s++) member(d.sSlot[s]);
  if (venkat) {
    const float TINYP = 1.0e-12f, TINYN = guard ? -1.0e-12f : 1.0e-12f;
    float lim[5];
#pragma unroll
    for (int k = 0; k < 5; k++) lim[k] = 1.0f;
    float epsq = (float)d.epsq[i];
    int eb = d.cfOff[i], ee = eb + d.cfNInt[i];
    for (int e = eb; e < ee; e++) {  // face neighbours' differences, from the tile's LDS slots
      const SlotF &m = sRecF[d.cfSlot[e]];
      float dv[5] = {m.s.x - ri[0], m.s.y - ri[1], m.s.z - ri[2], m.s.w - ri[3], m.p - ri[4]};
#pragma unroll
      for (int k = 0; k < 5; k++) { dmin[k] = fminf(dmin[k], dv[k]); dmax[k] = fmaxf(dmax[k], dv[k]); }
    }
    for (int e = eb; e < ee; e++) {
      float dx1 = (float)d.cfdx[e], dy1 = (float)d.cfdy[e], dz1 = (float)d.cfdz[e];
#pragma unroll
      for (int k = 0; k < 5; k++) {
        float d21 = g[3 * k] * dx1 + g[3 * k + 1] * dy1 + g[3 * k + 2] * dz1;
        if (d21 != 0.0f) {
          float d1 = d21 > 0.0f ? dmax[k] : dmin[k];
          float numer = (d1 * d1 + epsq) * d21 + 2.0f * d21 * d21 * d1;
          float denom = d21 * (d1 * d1 + 2.0f * d21 * d21 + d1 * d21 + epsq);
          lim[k] = fminf(__fdividef(numer, denom + (denom >= 0.0f ? TINYP : TINYN)), lim[k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 5; k++) {
      g[3 * k] *= lim[k];
      g[3 * k + 1] *= lim[k];
      g[3 * k + 2] *= lim[k];
    }
  }
  float4 *o = reinterpret_cast<float4 *>(d.gradF + 16 * (size_t)i);
  o[0] = make_float4(g[0], g[1], g[2], g[3]);
  o[1] = make_float4(g[4], g[5], g[6], g[7]);
  o[2] = make_float4(g[8], g[9], g[10], g[11]);
  o[3] = make_float4(g[12], g[13], g[14], 0.0f);
}

// RFLU_LimitGradCellsSimple on rho and p of the DUVWP record, then the store.
__device__ __forceinline__ void weno_positivity_store(const Dev &d, int i, double g[15]) {
  double r[8];
  load_rec(d.rec, i, r);  // rho u v w p x y z
  double xc = r[5], yc = r[6], zc = r[7];
  bool z0 = false, z4 = false;
  for (int e = d.cfOff[i]; e < d.cfOff[i + 1]; e++) {
    int slot = d.cfIdx[e] >> 1;
    double dx = d.fcx[slot] - xc, dy = d.fcy[slot] - yc, dz = d.fcz[slot] - zc;
    if (r[0] + g[0] * dx + g[1] * dy + g[2] * dz <= 0.0) z0 = true;
    if (r[4] + g[12] * dx + g[13] * dy + g[14] * dz <= 0.0) z4 = true;
  }
  if (z0) g[0] = g[1] = g[2] = 0.0;
  if (z4) g[12] = g[13] = g[14] = 0.0;
  double *o = d.grad + 15 * (size_t)i;
#pragma unroll
  for (int q = 0; q < 15; q++) o[q] = g[q];
}

// RECONST_WENO_XYZ (libfloflu/CellGradientsMP.F90:208-215): the WENO-XYZ
// average of the unlimited LSQ gradients over each cell and its stencil
// (RFLU_WENOGradCellsXYZFast_3D, modflu/RFLU_ModWENO.F90:1402-1588: per
// component, weights 1/(g^2 + 1e-15), the cell first, then the members in
// stencil order), then RFLU_LimitGradCellsSimple on rho and p
// (modflu/RFLU_ModLimiters.F90:793-921: the gradient is zeroed when the
// extrapolation to any face centroid of the cell is <= 0).  Over the
// gradient tiles: for each variable the tile's slots carry w and w*g per
// component in LDS, formed once per slot.
template <int TB>
__global__ void __launch_bounds__(TB) k_weno_tile(Dev d) {
  extern __shared__ double2 sW[];  // [nU][3]
  int tile = xcd_block();
  int u0 = d.tOff[tile], nU = d.tOff[tile + 1] - u0;
  int i = tile * TB + threadIdx.x;
  bool live = i < d.nC;
  int sb = live ? d.csOff[i] : 0, se = live ? d.csOff[i + 1] : 0;
  double g[15];
  for (int k = 0; k < 5; k++) {
    if (k) __syncthreads();  // the previous variable's slots are consumed
    for (int u = threadIdx.x; u < nU; u += TB) {
      const double *gr = d.gradRaw + 15 * (size_t)d.tMem[u0 + u] + 3 * k;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        double t = gr[q];
        double w = 1.0 / (t * t + 1.0E-15);
        sW[3 * u + q] = make_double2(w, w * t);
      }
    }
    __syncthreads();
    if (live) {
      double2 o0 = sW[3 * threadIdx.x], o1 = sW[3 * threadIdx.x + 1], o2 = sW[3 * threadIdx.x + 2];
      double s0 = o0.x, s1 = o1.x, s2 = o2.x, l0 = o0.y, l1 = o1.y, l2 = o2.y;
      for (int s = sb; s < se; s++) {
        int u = d.sSlot[s] & 0x7fffu;
        double2 a = sW[3 * u], b = sW[3 * u + 1], c = sW[3 * u + 2];
        s0 = s0 + a.x; s1 = s1 + b.x; s2 = s2 + c.x;
        l0 = l0 + a.y; l1 = l1 + b.y; l2 = l2 + c.y;
      }
      g[3 * k] = l0 / s0;
      g[3 * k + 1] = l1 / s1;
      g[3 * k + 2] = l2 / s2;
    }
  }
  if (live) weno_positivity_store(d, i, g);
}

// The same average straight from global memory (meshes whose tiles do not
// fit in LDS).
__global__ void __launch_bounds__(BLK) k_weno(Dev d) {
  int i = xcd_block() * BLK + threadIdx.x;
  if (i >= d.nC) return;
  const unsigned *mem = reinterpret_cast<const unsigned *>(d.csMem);
  int sb = d.csOff[i], se = d.csOff[i + 1];
  double g[15];
  for (int c = 0; c < 15; c++) {
    double t = d.gradRaw[15 * (size_t)i + c];
    double sum = 1.0 / (t * t + 1.0E-15), loc = sum * t;
    for (int s = sb; s < se; s++) {
      t = d.gradRaw[15 * (size_t)(mem[s] & CS_IDX) + c];
      double w = 1.0 / (t * t + 1.0E-15);
      sum = sum + w;
      loc = loc + w * t;
    }
    g[c] = loc / sum;
  }
  weno_positivity_store(d, i, g);
}

// Cooperative variant 3: a workgroup of 256 lanes takes 16 cells.
//  A: 16 lanes per cell fetch that cell's stencil members in parallel and
//     stage each member's weights and variable differences in LDS (SoA, so
//     consecutive lanes hit consecutive words); face-neighbour members also
//     feed the limiter's min/max, reduced across the 16 lanes (exact).
//  B: 15 lanes per cell, one per (variable, direction), accumulate over the
//     members in stencil order -- the reference's sequence of roundings.
//  C: 5 lanes per cell apply the limiter from the staged per-face geometry.
#define G3C 16
#ifndef GMAXM
#define GMAXM 40
#endif
__global__ void __launch_bounds__(256) k_grad3(Dev d, int venkat, int guard) {
  constexpr int NS = G3C * GMAXM + 1;  // +1 staggers the 8 arrays across LDS banks
  __shared__ double sA[8][NS];        // wtx, wty, wtz, dVar[5]
  __shared__ double sOwn[G3C][8];
  __shared__ double sC[G3C][6];
  __shared__ double sMin[G3C][5], sMax[G3C][5];
  __shared__ double sG[G3C][15];
  __shared__ int sOff[G3C + 1];
  const int c0 = xcd_block() * G3C;
  const int nCell = min(G3C, d.nC - c0);
  const int tid = threadIdx.x;
  const int N = d.nC;
  if (tid < nCell) {
    int i = c0 + tid;
    double r[8];
    load_rec(d.rec, i, r);
#pragma unroll
    for (int q = 0; q < 8; q++) sOwn[tid][q] = r[q];
    double r11 = d.rm[i], r12 = d.rm[N + i], r22 = d.rm[2 * N + i];
    double r13 = d.rm[3 * N + i], r23 = d.rm[4 * N + i], r33 = d.rm[5 * N + i];
    double c11 = 1.0 / r11, c22 = 1.0 / r22, c33 = 1.0 / r33;
    double c12 = -(c11 * r12);
    double c13 = -(c11 * r13 + c12 * c22 * r23);
    double c23 = -(c22 * r23);
    sC[tid][0] = c11 * c11; sC[tid][1] = c22 * c22; sC[tid][2] = c33 * c33;
    sC[tid][3] = c12; sC[tid][4] = c13; sC[tid][5] = c23;
  }
  if (tid <= nCell) sOff[tid] = d.csOff[c0 + tid] - d.csOff[c0];
  __syncthreads();
  const unsigned *mem = reinterpret_cast<const unsigned *>(d.csMem) + d.csOff[c0];
  {
    const int ci = tid >> 4, l = tid & 15;
    double vmn[5], vmx[5];
    const bool live = ci < nCell;
#pragma unroll
    for (int k = 0; k < 5; k++) { vmn[k] = live ? sOwn[ci][k] : 0.0; vmx[k] = vmn[k]; }
    if (live) {
      const double xi = sOwn[ci][5], yi = sOwn[ci][6], zi = sOwn[ci][7];
      const double c1111 = sC[ci][0], c2222 = sC[ci][1], c3333 = sC[ci][2];
      const double c12 = sC[ci][3], c13 = sC[ci][4], c23 = sC[ci][5];
      const int qb = sOff[ci], qe = sOff[ci + 1];
      for (int q = qb + l; q < qe; q += 16) {
        unsigned j = mem[q];
        double rj[8];
        load_rec(d.rec, (int)(j & CS_IDX), rj);
        double dx = rj[5] - xi, dy = rj[6] - yi, dz = rj[7] - zi;
        double term = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
        dx = term * dx; dy = term * dy; dz = term * dz;
        double term1 = c1111 * (dx);
        double term2 = c2222 * (dy + c12 * dx);
        double term3 = c3333 * (dz + c23 * dy + c13 * dx);
        sA[0][q] = term * (term1 + c12 * term2 + c13 * term3);
        sA[1][q] = term * (term2 + c23 * term3);
        sA[2][q] = term * (term3);
#pragma unroll
        for (int k = 0; k < 5; k++) sA[3 + k][q] = rj[k] - sOwn[ci][k];
        if (j & CS_FACE) {
#pragma unroll
          for (int k = 0; k < 5; k++) { vmn[k] = fmin(vmn[k], rj[k]); vmx[k] = fmax(vmx[k], rj[k]); }
        }
      }
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1)
#pragma unroll
      for (int k = 0; k < 5; k++) {
        vmn[k] = fmin(vmn[k], __shfl_xor(vmn[k], off, 16));
        vmx[k] = fmax(vmx[k], __shfl_xor(vmx[k], off, 16));
      }
    if (live && l == 0)
#pragma unroll
      for (int k = 0; k < 5; k++) { sMin[ci][k] = vmn[k]; sMax[ci][k] = vmx[k]; }
  }
  __syncthreads();
  if (tid < nCell * 15) {
    int ci = tid / 15, comp = tid - 15 * ci, k = comp / 3, dd = comp - 3 * k;
    const double *w = sA[dd], *v = sA[3 + k];
    double g = 0.0;
#pragma unroll 4
    for (int q = sOff[ci]; q < sOff[ci + 1]; q++) g = g + w[q] * v[q];
    sG[ci][comp] = g;
  }
  __syncthreads();
  if (tid < nCell * 5) {
    int ci = tid / 5, k = tid - 5 * ci, i = c0 + ci;
    double gx = sG[ci][3 * k], gy = sG[ci][3 * k + 1], gz = sG[ci][3 * k + 2];
    if (venkat) {
      const double TINYP = 1.0e-12, TINYN = guard ? -1.0e-12 : 1.0e-12;
      double vi = sOwn[ci][k], lim = 1.0;
      double dmax = sMax[ci][k] - vi, dmin = sMin[ci][k] - vi;
      double epsq = d.epsq[i];
      int eb = d.cfOff[i], ee = eb + d.cfNInt[i];
      for (int e = eb; e < ee; e++) {
        double d21 = gx * d.cfdx[e] + gy * d.cfdy[e] + gz * d.cfdz[e];
        if (d21 > 0.0) {
          double numer = (dmax * dmax + epsq) * d21 + 2.0 * d21 * d21 * dmax;
          double denom = d21 * (dmax * dmax + 2.0 * d21 * d21 + dmax * d21 + epsq);
          lim = fmin(numer / (denom + (denom >= 0.0 ? TINYP : TINYN)), lim);
        } else if (d21 < 0.0) {
          double numer = (dmin * dmin + epsq) * d21 + 2.0 * d21 * d21 * dmin;
          double denom = d21 * (dmin * dmin + 2.0 * d21 * d21 + dmin * d21 + epsq);
          lim = fmin(numer / (denom + (denom >= 0.0 ? TINYP : TINYN)), lim);
        }
      }
      gx = lim * gx; gy = lim * gy; gz = lim * gz;
    }
    double *o = d.grad + 15 * (size_t)i + 3 * k;
    o[0] = gx; o[1] = gy; o[2] = gz;
  }
}

// Cooperative variant: one workgroup takes GCPB consecutive cells.  All
// stencil members of those cells are gathered at once (member-parallel), each
// member's LSQ weights and variable differences are staged in LDS, and then
// 15 lanes per cell each run one (variable, direction) accumulation over the
// members in stencil order -- the same sequence of roundings per accumulator
// as the reference's member loop.  The limiter then runs with 5 lanes per
// cell (one per variable; min/max and the per-face minimum are exact, so their
// order is free).
#define GCPB 16
__global__ void __launch_bounds__(BLK) k_grad_lds(Dev d, int venkat, int guard) {
  __shared__ double sW[GCPB * GMAXM][8];  // wtx, wty, wtz, dVar[5]
  __shared__ double sOwn[GCPB][8];        // own record
  __shared__ double sC[GCPB][6];          // c1111, c2222, c3333, c12, c13, c23
  __shared__ double sG[GCPB][15];
  __shared__ int sOff[GCPB + 1];
  const int c0 = xcd_block() * GCPB;
  const int nCell = min(GCPB, d.nC - c0);
  const int tid = threadIdx.x;
  const int N = d.nC;
  if (tid < nCell) {
    int i = c0 + tid;
    double r[8];
    load_rec(d.rec, i, r);
#pragma unroll
    for (int q = 0; q < 8; q++) sOwn[tid][q] = r[q];
    double r11 = d.rm[i], r12 = d.rm[N + i], r22 = d.rm[2 * N + i];
    double r13 = d.rm[3 * N + i], r23 = d.rm[4 * N + i], r33 = d.rm[5 * N + i];
    double c11 = 1.0 / r11, c22 = 1.0 / r22, c33 = 1.0 / r33;
    double c12 = -(c11 * r12);
    double c13 = -(c11 * r13 + c12 * c22 * r23);
    double c23 = -(c22 * r23);
    sC[tid][0] = c11 * c11; sC[tid][1] = c22 * c22; sC[tid][2] = c33 * c33;
    sC[tid][3] = c12; sC[tid][4] = c13; sC[tid][5] = c23;
  }
  if (tid <= nCell) sOff[tid] = d.csOff[c0 + tid] - d.csOff[c0];
  __syncthreads();
  const int base = d.csOff[c0];
  const int nSlot = sOff[nCell];
  for (int q = tid; q < nSlot; q += BLK) {
    int ci = 0;
    while (q >= sOff[ci + 1]) ci++;
    int j = (int)((unsigned)d.csMem[base + q] & 0x7fffffffu);
    double rj[8];
    load_rec(d.rec, j, rj);
    double dx = rj[5] - sOwn[ci][5], dy = rj[6] - sOwn[ci][6], dz = rj[7] - sOwn[ci][7];
    double term = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    dx = term * dx; dy = term * dy; dz = term * dz;
    double c1111 = sC[ci][0], c2222 = sC[ci][1], c3333 = sC[ci][2];
    double c12 = sC[ci][3], c13 = sC[ci][4], c23 = sC[ci][5];
    double term1 = c1111 * (dx);
    double term2 = c2222 * (dy + c12 * dx);
    double term3 = c3333 * (dz + c23 * dy + c13 * dx);
    sW[q][0] = term * (term1 + c12 * term2 + c13 * term3);
    sW[q][1] = term * (term2 + c23 * term3);
    sW[q][2] = term * (term3);
#pragma unroll
    for (int k = 0; k < 5; k++) sW[q][3 + k] = rj[k] - sOwn[ci][k];
  }
  __syncthreads();
  if (tid < nCell * 15) {
    int ci = tid / 15, comp = tid - 15 * ci, k = comp / 3, dd = comp - 3 * k;
    double g = 0.0;
    for (int q = sOff[ci]; q < sOff[ci + 1]; q++) g = g + sW[q][dd] * sW[q][3 + k];
    sG[ci][comp] = g;
  }
  __syncthreads();
  if (tid < nCell * 5) {
    int ci = tid / 5, k = tid - 5 * ci, i = c0 + ci;
    double gx = sG[ci][3 * k], gy = sG[ci][3 * k + 1], gz = sG[ci][3 * k + 2];
    if (venkat) {
      const double TINYP = 1.0e-12, TINYN = guard ? -1.0e-12 : 1.0e-12;
      double vi = sOwn[ci][k], vmin = vi, vmax = vi, lim = 1.0;
      int eb = d.cfOff[i], ee = d.cfOff[i + 1];
      for (int e = eb; e < ee; e++) {
        int code = d.cfIdx[e], s = code >> 1;
        if (s >= d.nF) break;
        int o = d.f2c[2 * s + ((code & 1) ? 0 : 1)];
        double vo = d.rec[8 * (size_t)o + k];
        vmin = fmin(vmin, vo);
        vmax = fmax(vmax, vo);
      }
      double epsq = d.epsq[i];
      for (int e = eb; e < ee; e++) {
        int s = d.cfIdx[e] >> 1;
        if (s >= d.nF) break;
        double dx1 = sOwn[ci][5] - d.fcx[s], dy1 = sOwn[ci][6] - d.fcy[s], dz1 = sOwn[ci][7] - d.fcz[s];
        double d21 = gx * dx1 + gy * dy1 + gz * dz1;
        if (d21 > 0.0) {
          double d1 = vmax - vi;
          double numer = (d1 * d1 + epsq) * d21 + 2.0 * d21 * d21 * d1;
          double denom = d21 * (d1 * d1 + 2.0 * d21 * d21 + d1 * d21 + epsq);
          lim = fmin(numer / (denom + (denom >= 0.0 ? TINYP : TINYN)), lim);
        } else if (d21 < 0.0) {
          double d1 = vmin - vi;
          double numer = (d1 * d1 + epsq) * d21 + 2.0 * d21 * d21 * d1;
          double denom = d21 * (d1 * d1 + 2.0 * d21 * d21 + d1 * d21 + epsq);
          lim = fmin(numer / (denom + (denom >= 0.0 ? TINYP : TINYN)), lim);
        }
      }
      gx = lim * gx; gy = lim * gy; gz = lim * gz;
    }
    double *o = d.grad + 15 * (size_t)i + 3 * k;
    o[0] = gx; o[1] = gy; o[2] = gz;
  }
}

// ---------------------------------------------------------------------------
// Face fluxes
// ---------------------------------------------------------------------------
// modflu/RFLU_ModAUSMFlux.F90:282-367
__device__ __forceinline__ void ausm_flux(double nx, double ny, double nz, double nm,
                                          double rl, double ul, double vl, double wl, double pl, double Hl, double al,
                                          double rr, double ur, double vr, double wr, double pr, double Hr, double ar,
                                          double f[5]) {
  const double fs = 0.0;
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
  f[0] = (af * (mfp * rl + mfm * rr)) * nm;
  f[1] = (af * (mfp * rl * ul + mfm * rr * ur) + pf * nx) * nm;
  f[2] = (af * (mfp * rl * vl + mfm * rr * vr) + pf * ny) * nm;
  f[3] = (af * (mfp * rl * wl + mfm * rr * wr) + pf * nz) * nm;
  f[4] = (af * (mfp * rl * Hl + mfm * rr * Hr) + pf * fs) * nm;
}

// RFLU_AUSMPlusUp_FluxFunction2 (modflu/RFLU_ModAUSMPlusUpFlux.F90:469-588),
// called by both TCP drivers: f_a is overwritten with 1 there, so alpha =
// 3/16 and M_inf has no effect.  vFracF scales the convected part.
__device__ __forceinline__ void ausmup_flux(const Gas &G, double nx, double ny, double nz, double nm,
                                            double rl, double ul, double vl, double wl, double pl, double Hl,
                                            double al, double rr, double ur, double vr, double wr, double pr,
                                            double Hr, double ar, double vFracF, double f[5]) {
  const double fs = 0.0, fa = 1.0;
  double ql = ul * nx + vl * ny + wl * nz - fs;
  double qr = ur * nx + vr * ny + wr * nz - fs;
  double af = 0.5 * (al + ar);
  double ml = ql / af, mla = fabs(ml);
  double mr = qr / af, mra = fabs(mr);
  double mbarsq = 0.5 * (ml * ml + mr * mr);
  const double alpha = 0.1875;  // 0.1875 * (5 fa^2 - 4)
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
  double sigmabar = 1.0 - G.sigma * mbarsq;
  double mp = -(G.kp * fmax(sigmabar, 0.0) * (pr - pl)) / (fa * rf * af * af);  // pressure diffusion
  double pu = -(G.ku * wtl * wtr * fa * af * (rl + rr) * (qr - ql));            // velocity diffusion
  double mf = mlp + mrm + mp, mfa = fabs(mf);
  double mfp = 0.5 * (mf + mfa), mfm = 0.5 * (mf - mfa);
  double pf = wtl * pl + wtr * pr + pu;
  double av = af * vFracF;
  f[0] = (av * (mfp * rl + mfm * rr)) * nm;
  f[1] = (av * (mfp * rl * ul + mfm * rr * ur) + pf * nx) * nm;
  f[2] = (av * (mfp * rl * vl + mfm * rr * vr) + pf * ny) * nm;
  f[3] = (av * (mfp * rl * wl + mfm * rr * wr) + pf * nz) * nm;
  f[4] = (av * (mfp * rl * Hl + mfm * rr * Hr) + pf * fs) * nm;
}

// modflu/RFLU_ModRoeFlux.F90:2783-2855 (+ RFLU_ModEntropyFixes.F90:95-106)
__device__ __forceinline__ double hh_fix(double l, double dd) { return (l > dd) ? l : (l * l + dd * dd) / (2.0 * dd); }
__device__ __forceinline__ void roe_flux(const Gas &G, double nx, double ny, double nz, double nm,
                                         double rl, double ul, double vl, double wl, double pl, double Hl, double ql,
                                         double rr, double ur, double vr, double wr, double pr, double Hr, double qr,
                                         double f[5]) {
  const double fs = 0.0;
  double rh = sqrt(rl * rr);
  double wt = rl / (rl + rh);
  double uh = wt * ul + (1.0 - wt) * ur;
  double vh = wt * vl + (1.0 - wt) * vr;
  double wh = wt * wl + (1.0 - wt) * wr;
  double Hh = wt * Hl + (1.0 - wt) * Hr;
  double qh = uh * nx + vh * ny + wh * nz - fs;
  double sh = 0.5 * (uh * uh + vh * vh + wh * wh);
  double ah = sqrt((G.g - 1.0) * (Hh - 0.5 * (2.0 * sh)));
  double l1 = fabs(qh - ah), l2 = fabs(qh), l5 = fabs(qh + ah);
  double efc = G.epsentr * l5;
  l1 = hh_fix(l1, efc);
  l2 = hh_fix(l2, efc);
  l5 = hh_fix(l5, efc);
  double dr = rr - rl, du = ur - ul, de = vr - vl, dw = wr - wl, dp = pr - pl;
  double dq = du * nx + de * ny + dw * nz;
  double dv1 = (dp - rh * ah * dq) / (2.0 * ah * ah);
  double dv5 = (dp + rh * ah * dq) / (2.0 * ah * ah);
  double dv2 = dr - dp / (ah * ah);
  double t1 = l1 * dv1, t2 = l2 * dv2, t3 = l2 * rh, t5 = l5 * dv5;
  double df = G.dissFact;
  f[0] = 0.5 * (ql * rl + qr * rr) * nm - 0.5 * df * (t1 + t2 + t5) * nm;
  f[1] = 0.5 * (ql * rl * ul + pl * nx + qr * rr * ur + pr * nx) * nm -
         0.5 * df * (t1 * (uh - nx * ah) + t2 * uh + t3 * (du - nx * dq) + t5 * (uh + nx * ah)) * nm;
  f[2] = 0.5 * (ql * rl * vl + pl * ny + qr * rr * vr + pr * ny) * nm -
         0.5 * df * (t1 * (vh - ny * ah) + t2 * vh + t3 * (de - ny * dq) + t5 * (vh + ny * ah)) * nm;
  f[3] = 0.5 * (ql * rl * wl + pl * nz + qr * rr * wr + pr * nz) * nm -
         0.5 * df * (t1 * (wh - nz * ah) + t2 * wh + t3 * (dw - nz * dq) + t5 * (wh + nz * ah)) * nm;
  f[4] = 0.5 * (ql * rl * Hl + pl * fs + qr * rr * Hr + pr * fs) * nm -
         0.5 * df * (t1 * (Hh - qh * ah) + t2 * sh + t3 * (uh * du + vh * de + wh * dw - qh * dq) + t5 * (Hh + qh * ah)) * nm;
}

// modflu/RFLU_ModHLLCFlux.F90:2338-2409
__device__ __forceinline__ void hllc_flux(const Gas &G, double nx, double ny, double nz, double nm,
                                          double rl, double ul, double vl, double wl, double pl, double el, double al, double ql, double Hl,
                                          double rr, double ur, double vr, double wr, double pr, double er, double ar, double qr, double Hr,
                                          double f[5]) {
  const double fs = 0.0;
  double wt1 = sqrt(rr / rl);
  double wt2 = 1.0 / (1.0 + wt1);
  double us = wt2 * (ul + wt1 * ur);
  double vs = wt2 * (vl + wt1 * vr);
  double ws = wt2 * (wl + wt1 * wr);
  double qs = wt2 * (ql + wt1 * qr);
  double Hs = wt2 * (Hl + wt1 * Hr);
  double Vms = us * us + vs * vs + ws * ws;
  double as = sqrt((G.g - 1.0) * (Hs - 0.5 * Vms));
  double sl = fmin(ql - al, qs - as);
  double sr = fmax(qr + ar, qs + as);
  if (sl > 0.0) {
    f[0] = ql * rl * nm;
    f[1] = (ql * rl * ul + pl * nx) * nm;
    f[2] = (ql * rl * vl + pl * ny) * nm;
    f[3] = (ql * rl * wl + pl * nz) * nm;
    f[4] = (ql * (rl * el + pl) + pl * fs) * nm;
  } else if (sr < 0.0) {
    f[0] = qr * rr * nm;
    f[1] = (qr * rr * ur + pr * nx) * nm;
    f[2] = (qr * rr * vr + pr * ny) * nm;
    f[3] = (qr * rr * wr + pr * nz) * nm;
    f[4] = (qr * (rr * er + pr) + pr * fs) * nm;
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
    f[0] = sm * rs * nm;
    f[1] = (sm * rus + ps * nx) * nm;
    f[2] = (sm * rvs + ps * ny) * nm;
    f[3] = (sm * rws + ps * nz) * nm;
    f[4] = (sm * (res + ps) - ps * fs) * nm;
  }
}

// FP32 forms of the three flux functions (fluxPrecision 1), same formulas.
__device__ __forceinline__ void ausm_flux_f(float nx, float ny, float nz, float nm, float rl, float ul, float vl,
                                            float wl, float pl, float Hl, float al, float rr, float ur, float vr,
                                            float wr, float pr, float Hr, float ar, float f[5]) {
  float ql = ul * nx + vl * ny + wl * nz;
  float qr = ur * nx + vr * ny + wr * nz;
  float af = 0.5f * (al + ar);
  float ml = ql / af, mla = fabsf(ml);
  float mr = qr / af, mra = fabsf(mr);
  float mlp, wtl, mrm, wtr;
  if (mla <= 1.0f) {
    mlp = 0.25f * (ml + 1.0f) * (ml + 1.0f) + 0.125f * (ml * ml - 1.0f) * (ml * ml - 1.0f);
    wtl = 0.25f * (ml + 1.0f) * (ml + 1.0f) * (2.0f - ml) + 0.1875f * ml * (ml * ml - 1.0f) * (ml * ml - 1.0f);
  } else {
    mlp = 0.5f * (ml + mla);
    wtl = 0.5f * (1.0f + ml / mla);
  }
  if (mra <= 1.0f) {
    mrm = -0.25f * (mr - 1.0f) * (mr - 1.0f) - 0.125f * (mr * mr - 1.0f) * (mr * mr - 1.0f);
    wtr = 0.25f * (mr - 1.0f) * (mr - 1.0f) * (2.0f + mr) - 0.1875f * mr * (mr * mr - 1.0f) * (mr * mr - 1.0f);
  } else {
    mrm = 0.5f * (mr - mra);
    wtr = 0.5f * (1.0f - mr / mra);
  }
  float mf = mlp + mrm, mfa = fabsf(mf);
  float mfp = 0.5f * (mf + mfa), mfm = 0.5f * (mf - mfa);
  float pf = wtl * pl + wtr * pr;
  f[0] = (af * (mfp * rl + mfm * rr)) * nm;
  f[1] = (af * (mfp * rl * ul + mfm * rr * ur) + pf * nx) * nm;
  f[2] = (af * (mfp * rl * vl + mfm * rr * vr) + pf * ny) * nm;
  f[3] = (af * (mfp * rl * wl + mfm * rr * wr) + pf * nz) * nm;
  f[4] = (af * (mfp * rl * Hl + mfm * rr * Hr)) * nm;
}

__device__ __forceinline__ float hh_fix_f(float l, float dd) { return (l > dd) ? l : (l * l + dd * dd) / (2.0f * dd); }
__device__ __forceinline__ void roe_flux_f(float g1, float epsentr, float dissFact, float nx, float ny, float nz,
                                           float nm, float rl, float ul, float vl, float wl, float pl, float Hl,
                                           float ql, float rr, float ur, float vr, float wr, float pr, float Hr,
                                           float qr, float f[5]) {
  float rh = sqrtf(rl * rr);
  float wt = rl / (rl + rh);
  float uh = wt * ul + (1.0f - wt) * ur;
  float vh = wt * vl + (1.0f - wt) * vr;
  float wh = wt * wl + (1.0f - wt) * wr;
  float Hh = wt * Hl + (1.0f - wt) * Hr;
  float qh = uh * nx + vh * ny + wh * nz;
  float sh = 0.5f * (uh * uh + vh * vh + wh * wh);
  float ah = sqrtf(g1 * (Hh - sh));
  float l1 = fabsf(qh - ah), l2 = fabsf(qh), l5 = fabsf(qh + ah);
  float efc = epsentr * l5;
  l1 = hh_fix_f(l1, efc);
  l2 = hh_fix_f(l2, efc);
  l5 = hh_fix_f(l5, efc);
  float dr = rr - rl, du = ur - ul, de = vr - vl, dw = wr - wl, dp = pr - pl;
  float dq = du * nx + de * ny + dw * nz;
  float dv1 = (dp - rh * ah * dq) / (2.0f * ah * ah);
  float dv5 = (dp + rh * ah * dq) / (2.0f * ah * ah);
  float dv2 = dr - dp / (ah * ah);
  float t1 = l1 * dv1, t2 = l2 * dv2, t3 = l2 * rh, t5 = l5 * dv5;
  float hd = 0.5f * dissFact;
  f[0] = 0.5f * (ql * rl + qr * rr) * nm - hd * (t1 + t2 + t5) * nm;
  f[1] = 0.5f * (ql * rl * ul + pl * nx + qr * rr * ur + pr * nx) * nm -
         hd * (t1 * (uh - nx * ah) + t2 * uh + t3 * (du - nx * dq) + t5 * (uh + nx * ah)) * nm;
  f[2] = 0.5f * (ql * rl * vl + pl * ny + qr * rr * vr + pr * ny) * nm -
         hd * (t1 * (vh - ny * ah) + t2 * vh + t3 * (de - ny * dq) + t5 * (vh + ny * ah)) * nm;
  f[3] = 0.5f * (ql * rl * wl + pl * nz + qr * rr * wr + pr * nz) * nm -
         hd * (t1 * (wh - nz * ah) + t2 * wh + t3 * (dw - nz * dq) + t5 * (wh + nz * ah)) * nm;
  f[4] = 0.5f * (ql * rl * Hl + qr * rr * Hr) * nm -
         hd * (t1 * (Hh - qh * ah) + t2 * sh + t3 * (uh * du + vh * de + wh * dw - qh * dq) + t5 * (Hh + qh * ah)) * nm;
}

__device__ __forceinline__ void hllc_flux_f(float g1, float nx, float ny, float nz, float nm, float rl, float ul,
                                            float vl, float wl, float pl, float el, float al, float ql, float Hl,
                                            float rr, float ur, float vr, float wr, float pr, float er, float ar,
                                            float qr, float Hr, float f[5]) {
  float wt1 = sqrtf(rr / rl);
  float wt2 = 1.0f / (1.0f + wt1);
  float us = wt2 * (ul + wt1 * ur);
  float vs = wt2 * (vl + wt1 * vr);
  float ws = wt2 * (wl + wt1 * wr);
  float qs = wt2 * (ql + wt1 * qr);
  float Hs = wt2 * (Hl + wt1 * Hr);
  float Vms = us * us + vs * vs + ws * ws;
  float as = sqrtf(g1 * (Hs - 0.5f * Vms));
  float sl = fminf(ql - al, qs - as);
  float sr = fmaxf(qr + ar, qs + as);
  if (sl > 0.0f) {
    f[0] = ql * rl * nm;
    f[1] = (ql * rl * ul + pl * nx) * nm;
    f[2] = (ql * rl * vl + pl * ny) * nm;
    f[3] = (ql * rl * wl + pl * nz) * nm;
    f[4] = (ql * (rl * el + pl)) * nm;
  } else if (sr < 0.0f) {
    f[0] = qr * rr * nm;
    f[1] = (qr * rr * ur + pr * nx) * nm;
    f[2] = (qr * rr * vr + pr * ny) * nm;
    f[3] = (qr * rr * wr + pr * nz) * nm;
    f[4] = (qr * (rr * er + pr)) * nm;
  } else {
    float sm = (rr * qr * (sr - qr) - rl * ql * (sl - ql) - pr + pl) / (rr * (sr - qr) - rl * (sl - ql));
    float ps = pl + rl * (ql - sl) * (ql - sm);
    float rs, rus, rvs, rws, res, term;
    if (sm > 0.0f) {
      term = 1.0f / (sl - sm);
      rs = term * (sl - ql) * rl;
      rus = term * ((sl - ql) * rl * ul + (ps - pl) * nx);
      rvs = term * ((sl - ql) * rl * vl + (ps - pl) * ny);
      rws = term * ((sl - ql) * rl * wl + (ps - pl) * nz);
      res = term * ((sl - ql) * rl * el + (ps * sm - pl * ql));
    } else {
      term = 1.0f / (sr - sm);
      rs = term * (sr - qr) * rr;
      rus = term * ((sr - qr) * rr * ur + (ps - pr) * nx);
      rvs = term * ((sr - qr) * rr * vr + (ps - pr) * ny);
      rws = term * ((sr - qr) * rr * wr + (ps - pr) * nz);
      res = term * ((sr - qr) * rr * er + (ps * sm - pr * qr));
    }
    f[0] = sm * rs * nm;
    f[1] = (sm * rus + ps * nx) * nm;
    f[2] = (sm * rvs + ps * ny) * nm;
    f[3] = (sm * rws + ps * nz) * nm;
    f[4] = (sm * (res + ps)) * nm;
  }
}

// FP32 face state: the cell value rounded once plus the FP32 gradient
// correction; the face-minus-centroid vector is formed in FP64.
struct FaceStateF { float r, u, v, w, p; };
__device__ __forceinline__ FaceStateF reconstruct_f(const double rc[8], const float *__restrict__ gp, double xc,
                                                    double yc, double zc) {
  const float4 *g4 = reinterpret_cast<const float4 *>(gp);
  float4 a = g4[0], b = g4[1], c = g4[2], e = g4[3];
  float dx = (float)(xc - rc[5]), dy = (float)(yc - rc[6]), dz = (float)(zc - rc[7]);
  FaceStateF s;
  s.r = (float)rc[0] + (a.x * dx + a.y * dy + a.z * dz);
  s.u = (float)rc[1] + (a.w * dx + b.x * dy + b.y * dz);
  s.v = (float)rc[2] + (b.z * dx + b.w * dy + c.x * dz);
  s.w = (float)rc[3] + (c.y * dx + c.z * dy + c.w * dz);
  s.p = (float)rc[4] + (e.x * dx + e.y * dy + e.z * dz);
  return s;
}

// Face state reconstructed at (xc,yc,zc) from an order-2 record and gradient;
// u is taken through the reference's prim->cons->prim round trip
// (RFLU_ModConvertCv.F90:554-592 then RFLU_ModAUSMFlux.F90:2735-2783).
// rhoP is cv(1) as the flux loop reads it: rho itself, or with particles
// (rho/(1-phi))*(1-phi) after CellGradientsMP's rho_g round trip
// (CellGradientsMP.F90:371-390), while rho*u was formed from rho before it.
struct FaceState { double r, u, v, w, p, ir; };
__device__ __forceinline__ FaceState reconstruct(const double rc[8], const double *__restrict__ gp,
                                                 double xc, double yc, double zc, double rhoP) {
  FaceState s;
  double rho = rc[0];
  double ir = 1.0 / rhoP;
  double u = (rho * rc[1]) * ir, v = (rho * rc[2]) * ir, w = (rho * rc[3]) * ir, p = rc[4];
  double dx = xc - rc[5], dy = yc - rc[6], dz = zc - rc[7];
  double g[15];
#pragma unroll
  for (int q = 0; q < 15; q++) g[q] = gp[q];
  s.r = rhoP + g[0] * dx + g[1] * dy + g[2] * dz;
  s.u = u + g[3] * dx + g[4] * dy + g[5] * dz;
  s.v = v + g[6] * dx + g[7] * dy + g[8] * dz;
  s.w = w + g[9] * dx + g[10] * dy + g[11] * dz;
  s.p = p + g[12] * dx + g[13] * dy + g[14] * dz;
  s.ir = ir;
  return s;
}

// libfloflu/BcondOutflowPerf.F90:88-160 (gas from gv: R = 8314.3/M path)
__device__ __forceinline__ void bcond_outflow(const Gas &G, int opt, double pout, double sxn, double syn, double szn,
                                              double rho, double rhou, double rhov, double rhow, double rhoe,
                                              double press, double b[5]) {
  double gamma = G.gD;
  double u = rhou / rho, v = rhov / rho, w = rhow / rho;
  double csound = sqrt(gamma * press / rho);
  double mach = sqrt(u * u + v * v + w * w) / csound;
  if (mach < 1.0 && (opt == RFLU_BCOPT_SUBSONIC || opt == RFLU_BCOPT_MIXED)) {
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
    b[0] = rhob; b[1] = rhob * ub; b[2] = rhob * vb; b[3] = rhob * wb;
    b[4] = rhob * eo_dgpuvw(rhob, gamma, pout, ub, vb, wb);
  } else {
    b[0] = rho; b[1] = rhou; b[2] = rhov; b[3] = rhow; b[4] = rhoe;
  }
}

struct FluxPar { int order, discr, patchOrder, doSrad; };

// Boundary-face kinds (d.bkind): 0 slip wall, 1 + bcOpt outflow, then
enum { BK_NOSLIP = 4, BK_FARFIELD = 5, BK_INJECTION = 6, BK_VELTEMP = 7, BK_TOTANG = 8 };
#define PCON 24
// Patch constants (host-computed with the reference's expressions, so the
// angle functions and the patch-constant powers match the reference bit for bit):
//   farfield: mach, p_inf, rho_inf, q_inf, u_inf, v_inf, w_inf
//   injection: mass flux, T
//   velocity/temperature inflow: u, v, w, T, p
//   total-pressure inflow: p_tot, T_tot, cos bh, sin bh, cos bv, sin bv, mach,
//     fixed, supersonic state (rho, rho u, rho v, rho w, rho E, p) at 8..13,
//     stagnation state (rho, rho E, p) at 14..16, c_0^2 at 17
enum { PC_TT_SUP = 8, PC_TT_STAG = 14, PC_TT_A02 = 17 };

// RFLU_SetRindStateFarfieldPerf without the lift correction
// (modflu/RFLU_ModRindStates.F90:249-421): b = (rho, rho u, rho v, rho w, rho E, p)
__device__ __forceinline__ void rind_farfield(const Gas &G, const double *pc, double nx, double ny, double nz,
                                              double rl, double rul, double rvl, double rwl, double rel, double b[6]) {
  const double g = G.gD;
  double ul = rul / rl, vl = rvl / rl, wl = rwl / rl;
  double ql = ul * nx + vl * ny + wl * nz;
  double el = rel / rl, sl2 = ul * ul + vl * vl + wl * wl;
  double pl = (g - 1.0) * rl * (el - 0.5 * sl2 - 0.0);
  double machInf = pc[0], pi = pc[1], ri = pc[2], qi = pc[3], ui = pc[4], vi = pc[5], wi = pc[6];
  if (machInf < 1.0) {
    double al = sqrt(g * pl / rl), pb, rb, ub, vb, wb;
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
    b[0] = rb; b[1] = rb * ub; b[2] = rb * vb; b[3] = rb * wb;
    b[4] = rb * eo_dgpuvw(rb, g, pb, ub, vb, wb);
    b[5] = pb;
  } else if (ql < 0.0) {
    b[0] = ri; b[1] = ri * ui; b[2] = ri * vi; b[3] = ri * wi;
    b[4] = ri * (pi / (ri * (g - 1.0)) + 0.5 * qi * qi + 0.0);  // MixtPerf_Eo_DGPVm
    b[5] = pi;
  } else {
    b[0] = rl; b[1] = rul; b[2] = rvl; b[3] = rwl; b[4] = rel; b[5] = pl;
  }
}

// libfloflu/BcondInflowPerf.F90 (total pressure, total temperature, angles);
// a subsonic-only patch seeing |q| >= c, which the reference leaves
// undefined, gets the stagnation state at rest.
__device__ __forceinline__ void rind_totang(const Gas &G, int opt, const double *pc, double nx, double ny, double nz,
                                            double rl, double rul, double rvl, double rwl, double b[6]) {
  const double g = G.gD, rgas = G.gcD;
  if (opt == RFLU_BCOPT_SUBSONIC || opt == RFLU_BCOPT_MIXED) {
    double gm1 = g - 1.0;
    double ul = rul / rl, vl = rvl / rl, wl = rwl / rl;
    double a02 = pc[PC_TT_A02];
    double al = sqrt(a02 - 0.5 * (g - 1.0) * (ul * ul + vl * vl + wl * wl));  // MixtPerf_C_Co2GUVW
    double ql = ul * nx + vl * ny + wl * nz;
    if (fabs(ql) < al) {
      double sl = sqrt(ul * ul + vl * vl + wl * wl), eta;
      if (pc[7] == 0.0) eta = sl > 1.0E-6 ? ql / sl : -1.0;
      else eta = -1.0;
      double Rm = al - 0.5 * gm1 * ql;
      double disc = 0.5 * gm1 * (eta * eta) * (a02 / (Rm * Rm) * (1.0 + 0.5 * gm1 * (eta * eta)) - 1.0);
      double tr, pr, sr;
      if (disc < 0.0) {
        tr = pc[1];
        pr = pc[0];
        sr = 0.0;
      } else {
        double ar = Rm / (1.0 + 0.5 * gm1 * eta * eta) * (1.0 + sqrt(disc));
        tr = ar * ar / (g * rgas);                           // MixtPerf_T_CGR
        pr = pc[0] * pow(tr / pc[1], g / (g - 1.0));         // MixtPerf_P_GPoTTo
        double c2 = ar * ar;                                 // MixtPerf_Vm_C2Co2G
        sr = a02 > c2 ? sqrt(2.0 / (g - 1.0) * (a02 - c2)) : 0.0;
      }
      double rr = pr / (rgas * tr);
      double ur = sr * pc[2] * pc[4], vr = sr * pc[3], wr = sr * pc[2] * pc[5];
      b[0] = rr; b[1] = rr * ur; b[2] = rr * vr; b[3] = rr * wr;
      b[4] = rr * (pr / (rr * (g - 1.0)) + 0.5 * sr * sr + 0.0);  // MixtPerf_Eo_DGPVm
      b[5] = pr;
      return;
    }
    if (opt != RFLU_BCOPT_MIXED) {
      b[0] = pc[PC_TT_STAG]; b[1] = 0.0; b[2] = 0.0; b[3] = 0.0; b[4] = pc[PC_TT_STAG + 1]; b[5] = pc[PC_TT_STAG + 2];
      return;
    }
  }
#pragma unroll
  for (int k = 0; k < 6; k++) b[k] = pc[PC_TT_SUP + k];
}

// The interior face flux of the FP32 mode: AUSM+, Roe or HLLC on FP32 face
// states.  irA, irB: 1/rho of the cells the states belong to (HLLC's
// H = e + p / rho takes the cell's rho, as the FP64 path).
__device__ __forceinline__ void face_flux_f(const Gas &G, int discr, float fx, float fy, float fz, float fm,
                                            const FaceStateF &A, const FaceStateF &B, float irA, float irB,
                                            float ff[5]) {
  float g = (float)G.g, gc = (float)G.gc, cp = (float)G.cp;
  if (discr == RFLU_DISCR_AUSM) {
    float tl = A.p / (A.r * gc), al = sqrtf(g * gc * tl);
    float Hl = cp * tl + 0.5f * (A.u * A.u + A.v * A.v + A.w * A.w);
    float tr = B.p / (B.r * gc), ar = sqrtf(g * gc * tr);
    float Hr = cp * tr + 0.5f * (B.u * B.u + B.v * B.v + B.w * B.w);
    ausm_flux_f(fx, fy, fz, fm, A.r, A.u, A.v, A.w, A.p, Hl, al, B.r, B.u, B.v, B.w, B.p, Hr, ar, ff);
  } else if (discr == RFLU_DISCR_ROE) {
    float Hl = cp * (A.p / (A.r * gc)) + 0.5f * (A.u * A.u + A.v * A.v + A.w * A.w);
    float ql = A.u * fx + A.v * fy + A.w * fz;
    float Hr = cp * (B.p / (B.r * gc)) + 0.5f * (B.u * B.u + B.v * B.v + B.w * B.w);
    float qr = B.u * fx + B.v * fy + B.w * fz;
    roe_flux_f(g - 1.0f, (float)G.epsentr, (float)G.dissFact, fx, fy, fz, fm, A.r, A.u, A.v, A.w, A.p, Hl, ql, B.r,
               B.u, B.v, B.w, B.p, Hr, qr, ff);
  } else {
    float el = A.p / ((g - 1.0f) * A.r) + 0.5f * (A.u * A.u + A.v * A.v + A.w * A.w), al = sqrtf(g * A.p / A.r);
    float ql = A.u * fx + A.v * fy + A.w * fz, Hl = el + A.p * irA;
    float er = B.p / ((g - 1.0f) * B.r) + 0.5f * (B.u * B.u + B.v * B.v + B.w * B.w), ar = sqrtf(g * B.p / B.r);
    float qr = B.u * fx + B.v * fy + B.w * fz, Hr = er + B.p * irB;
    hllc_flux_f(g - 1.0f, fx, fy, fz, fm, A.r, A.u, A.v, A.w, A.p, el, al, ql, Hl, B.r, B.u, B.v, B.w, B.p, er, ar, qr,
                Hr, ff);
  }
}

// The flux of face slot s (interior faces, then boundary faces by patch) as
// the face loop of RFLU_ComputeFluxInv and the patch loops compute it; srad
// (optional) receives the face's spectral radius for RFLU_TimeStepInviscid.
// F32 = 1: interior second-order faces from the FP32 gradients, the flux
// evaluated in FP32 as a difference to the left cell's own state, whose
// physical flux is added in FP64; boundary faces stay FP64 (their gradients
// widened from FP32).
template <int F32>
__device__ __forceinline__ void face_flux(const Dev &d, const Gas &G, const FluxPar &P, int s, double f[5],
                                          double *srad) {
  const double fs = 0.0, ksg = 0.0;
  double nx = d.fnx[s], ny = d.fny[s], nz = d.fnz[s], nm = d.fnm[s];
  if (s < d.nF) {
    int c1 = d.f2c[2 * s], c2 = d.f2c[2 * s + 1];
    double L[8], R[8];
    load_rec(d.rec, c1, L);
    load_rec(d.rec, c2, R);
    double phL = 0.0, phR = 0.0, rpL = L[0], rpR = R[0];
    if (d.phi) {
      phL = d.phi[c1];
      phR = d.phi[c2];
      double vl = 1.0 - phL, vr = 1.0 - phR;
      rpL = L[0] / vl * vl;
      rpR = R[0] / vr * vr;
    }
    if (srad) *srad = face_srad(G, P.order, L, R, nx, ny, nz, nm, phL, phR);
    if (P.order == 1) {
      double rl = L[0], irl = 1.0 / rl, ul = L[1] * irl, vl = L[2] * irl, wl = L[3] * irl;
      double pl = L[5], tl = L[6], al = L[7];
      double rr = R[0], irr = 1.0 / rr, ur = R[1] * irr, vr = R[2] * irr, wr = R[3] * irr;
      double pr = R[5], tr = R[6], ar = R[7];
      if (P.discr == RFLU_DISCR_AUSM) {
        double Hl = L[4] * irl + pl / rl, Hr = R[4] * irr + pr / rr;
        ausm_flux(nx, ny, nz, nm, rl, ul, vl, wl, pl, Hl, al, rr, ur, vr, wr, pr, Hr, ar, f);
      } else if (P.discr == RFLU_DISCR_AUSMUP) {  // RFLU_AUSMPlusUp_ComputeFlux1
        double Hl = L[4] * irl + pl / rl, Hr = R[4] * irr + pr / rr;
        ausmup_flux(G, nx, ny, nz, nm, rl, ul, vl, wl, pl, Hl, al, rr, ur, vr, wr, pr, Hr, ar, 1.0, f);
      } else if (P.discr == RFLU_DISCR_ROE) {
        double Hl = G.cp * tl + 0.5 * (ul * ul + vl * vl + wl * wl) + ksg;
        double Hr = G.cp * tr + 0.5 * (ur * ur + vr * vr + wr * wr) + ksg;
        double ql = ul * nx + vl * ny + wl * nz - fs, qr = ur * nx + vr * ny + wr * nz - fs;
        roe_flux(G, nx, ny, nz, nm, rl, ul, vl, wl, pl, Hl, ql, rr, ur, vr, wr, pr, Hr, qr, f);
      } else {
        double el = L[4] * irl, er = R[4] * irr;
        double ql = ul * nx + vl * ny + wl * nz - fs, qr = ur * nx + vr * ny + wr * nz - fs;
        double Hl = el + irl * pl, Hr = er + irr * pr;
        hllc_flux(G, nx, ny, nz, nm, rl, ul, vl, wl, pl, el, al, ql, Hl, rr, ur, vr, wr, pr, er, ar, qr, Hr, f);
      }
    } else if (F32) {
      double xc = d.fcx[s], yc = d.fcy[s], zc = d.fcz[s];
      FaceStateF A = reconstruct_f(L, d.gradF + 16 * (size_t)c1, xc, yc, zc);
      FaceStateF B = reconstruct_f(R, d.gradF + 16 * (size_t)c2, xc, yc, zc);
      float fx = (float)nx, fy = (float)ny, fz = (float)nz, fm = (float)nm;
      // The face flux is the physical flux of the left cell's own state U in
      // FP64 plus the FP32 difference F(A, B) - F(U, U).  Near a uniform
      // state the two FP32 evaluations round almost alike, so the difference
      // is accurate well below FP32 rounding of the flux itself, and the
      // faces of a closed cell cancel to FP64 rounding whatever their
      // orientation.  Where both face states equal U (zero gradients) the
      // difference is exactly zero and is not evaluated.
      const FaceStateF U = {(float)L[0], (float)L[1], (float)L[2], (float)L[3], (float)L[4]};
      {
        const double r = L[0], u = L[1], v = L[2], w = L[3], p = L[4];
        const double mf = r * (u * nx + v * ny + w * nz) * nm;
        const double H = G.g / (G.g - 1.0) * (p / r) + 0.5 * (u * u + v * v + w * w);
        f[0] = mf;
        f[1] = mf * u + p * nx * nm;
        f[2] = mf * v + p * ny * nm;
        f[3] = mf * w + p * nz * nm;
        f[4] = mf * H;
      }
      const bool uniform = A.r == U.r && A.u == U.u && A.v == U.v && A.w == U.w && A.p == U.p && B.r == U.r &&
                           B.u == U.u && B.v == U.v && B.w == U.w && B.p == U.p;
      if (!uniform) {
        const float irL = (float)(1.0 / L[0]), irR = (float)(1.0 / R[0]);
        float ff[5], fr[5];
        face_flux_f(G, P.discr, fx, fy, fz, fm, A, B, irL, irR, ff);
        face_flux_f(G, P.discr, fx, fy, fz, fm, U, U, irL, irL, fr);
#pragma unroll
        for (int k = 0; k < 5; k++) f[k] = f[k] + (double)(ff[k] - fr[k]);
      }
      return;
    } else {
      double xc = d.fcx[s], yc = d.fcy[s], zc = d.fcz[s];
      FaceState A = reconstruct(L, d.grad + 15 * (size_t)c1, xc, yc, zc, rpL);
      FaceState B = reconstruct(R, d.grad + 15 * (size_t)c2, xc, yc, zc, rpR);
      if (P.discr == RFLU_DISCR_AUSMUP) {
        // RFLU_AUSMPlusUp_ComputeFlux2_TCP: with particles the reconstructed
        // density is divided by the cell's gas fraction before T, a and H, and
        // the face's mean gas fraction scales the convected flux
        double rl = A.r, rr = B.r, vFracF = 1.0;
        if (d.phi) {
          rl = rl / (1.0 - phL);
          rr = rr / (1.0 - phR);
          vFracF = 0.5 * ((1.0 - phL) + (1.0 - phR));
        }
        double tl = A.p / (rl * G.gc), al = sqrt(G.g * G.gc * tl);
        double Hl = G.cp * tl + 0.5 * (A.u * A.u + A.v * A.v + A.w * A.w) + ksg;
        double tr = B.p / (rr * G.gc), ar = sqrt(G.g * G.gc * tr);
        double Hr = G.cp * tr + 0.5 * (B.u * B.u + B.v * B.v + B.w * B.w) + ksg;
        ausmup_flux(G, nx, ny, nz, nm, rl, A.u, A.v, A.w, A.p, Hl, al, rr, B.u, B.v, B.w, B.p, Hr, ar, vFracF, f);
      } else if (P.discr == RFLU_DISCR_AUSM) {
        double tl = A.p / (A.r * G.gc), al = sqrt(G.g * G.gc * tl);
        double Hl = G.cp * tl + 0.5 * (A.u * A.u + A.v * A.v + A.w * A.w) + ksg;
        double tr = B.p / (B.r * G.gc), ar = sqrt(G.g * G.gc * tr);
        double Hr = G.cp * tr + 0.5 * (B.u * B.u + B.v * B.v + B.w * B.w) + ksg;
        ausm_flux(nx, ny, nz, nm, A.r, A.u, A.v, A.w, A.p, Hl, al, B.r, B.u, B.v, B.w, B.p, Hr, ar, f);
      } else if (P.discr == RFLU_DISCR_ROE) {
        double tl = A.p / (A.r * G.gc);
        double Hl = G.cp * tl + 0.5 * (A.u * A.u + A.v * A.v + A.w * A.w) + ksg;
        double ql = A.u * nx + A.v * ny + A.w * nz - fs;
        double tr = B.p / (B.r * G.gc);
        double Hr = G.cp * tr + 0.5 * (B.u * B.u + B.v * B.v + B.w * B.w) + ksg;
        double qr = B.u * nx + B.v * ny + B.w * nz - fs;
        roe_flux(G, nx, ny, nz, nm, A.r, A.u, A.v, A.w, A.p, Hl, ql, B.r, B.u, B.v, B.w, B.p, Hr, qr, f);
      } else {
        double el = eo_dgpuvw(A.r, G.g, A.p, A.u, A.v, A.w), al = sqrt(G.g * A.p / A.r);
        double ql = A.u * nx + A.v * ny + A.w * nz - fs, Hl = el + A.ir * A.p;
        double er = eo_dgpuvw(B.r, G.g, B.p, B.u, B.v, B.w), ar = sqrt(G.g * B.p / B.r);
        double qr = B.u * nx + B.v * ny + B.w * nz - fs, Hr = er + B.ir * B.p;
        hllc_flux(G, nx, ny, nz, nm, A.r, A.u, A.v, A.w, A.p, el, al, ql, Hl, B.r, B.u, B.v, B.w, B.p, er, ar, qr, Hr, f);
      }
    }
  } else {
    // rocflu/RFLU_CentralFirstPatch.F90:575-909 / RFLU_CentralSecondPatch.F90:638-1035
    int b = s - d.nF;
    int c1 = d.bf2c[b], kind = d.bkind[b];
    double L[8];
    load_rec(d.rec, c1, L);
    double phL = d.phi ? d.phi[c1] : 0.0;
    if (srad) *srad = face_srad(G, P.order, L, nullptr, nx, ny, nz, nm, phL);
    double rl, rul, rvl, rwl, rel, pl;
    if (P.patchOrder == 1) {
      if (P.order == 1) {
        rl = L[0]; rul = L[1]; rvl = L[2]; rwl = L[3]; rel = L[4]; pl = L[5];
      } else {  // cv after CellGradientsMP's prim->cons conversion
        rl = L[0]; rul = L[0] * L[1]; rvl = L[0] * L[2]; rwl = L[0] * L[3]; pl = L[4];
        double rE = L[0];
        if (d.phi) {  // energy from rho/(1-phi) (RFLU_ModConvertCv.F90:584-586), then the rho_g round trip
          double vl = 1.0 - phL;
          rE = L[0] / vl;
          rl = L[0] / vl * vl;
        }
        rel = L[0] * eo_dgpuvw(rE, G.gD, L[4], L[1], L[2], L[3]);
      }
    } else {
      double gw[15];
      const double *gp = d.grad + 15 * (size_t)c1;
      if (F32) {
#pragma unroll
        for (int q = 0; q < 15; q++) gw[q] = (double)d.gradF[16 * (size_t)c1 + q];
        gp = gw;
      }
      FaceState A = reconstruct(L, gp, d.fcx[s], d.fcy[s], d.fcz[s], L[0]);
      rl = A.r; pl = A.p;
      rul = rl * A.u; rvl = rl * A.v; rwl = rl * A.w;
      bool needE = kind != 0 && kind != BK_NOSLIP && kind != BK_INJECTION;
      rel = needE ? rl * eo_dgpuvw(rl, G.gD, pl, A.u, A.v, A.w) : 0.0;
    }
    if (kind == 0 || kind == BK_NOSLIP) {  // slip / no-slip wall: the pressure force only
      f[0] = -0.0;  // nothing added to the continuity residual
      f[1] = pl * nx * nm;
      f[2] = pl * ny * nm;
      f[3] = pl * nz * nm;
      f[4] = pl * fs * nm;
    } else if (kind == BK_INJECTION) {  // RFLU_CentralFirstPatch.F90:1049-1133
      const double *pc = d.pcon + PCON * (size_t)d.bpatch[b];
      double minj = pc[0];
      if (minj > 0.0) {  // RFLU_SetRindStateInjectPerf (RFLU_ModRindStates.F90:460-505)
        double tinj = pc[1];
        double ri = pl / (G.gcD * tinj);
        double qi = -minj / ri + fs;
        double ui = qi * nx, vi = qi * ny, wi = qi * nz;
        double Hi = G.cp * tinj + 0.5 * (ui * ui + vi * vi + wi * wi) + 0.0;
        f[0] = -minj * nm;
        f[1] = (-minj * ui + pl * nx) * nm;
        f[2] = (-minj * vi + pl * ny) * nm;
        f[3] = (-minj * wi + pl * nz) * nm;
        f[4] = (-minj * Hi + pl * fs) * nm;
      } else {  // RFLU_SetRindStateSlipWallPerf (RFLU_ModRindStates.F90:539-600)
        const double g = G.gD;
        double irl = 1.0 / rl;
        double ul = irl * rul, vl = irl * rvl, wl = irl * rwl;
        double ql = ul * nx + vl * ny + wl * nz - fs;
        double al = sqrt(g * pl / rl), pw;
        if (ql < 0.0) {
          double term = 1.0 + 0.5 * (g - 1.0) * ql / al;
          pw = pl * pow(term, 2.0 * g / (g - 1.0));
        } else {
          double term = (g + 1.0) / 4.0;
          pw = pl + term * rl * ql * (ql + sqrt(al * al + term * term * ql * ql) / term);
        }
        f[0] = 0.0;
        f[1] = pw * nx * nm;
        f[2] = pw * ny * nm;
        f[3] = pw * nz * nm;
        f[4] = pw * fs * nm;
      }
    } else {  // outflow, farfield and inflow: the central average with a rind state
      double bb[6];
      if (kind <= 3) {
        int opt = kind - 1;
        double pr = (opt != RFLU_BCOPT_SUPERSONIC) ? d.bpout[b] : pl;
        bcond_outflow(G, opt, pr, nx, ny, nz, rl, rul, rvl, rwl, rel, pl, bb);
        bb[5] = pr;
      } else {
        const double *pc = d.pcon + PCON * (size_t)d.bpatch[b];
        if (kind == BK_FARFIELD) {
          rind_farfield(G, pc, nx, ny, nz, rl, rul, rvl, rwl, rel, bb);
        } else if (kind == BK_VELTEMP) {  // RFLU_CentralFirstPatch.F90:377-574
          double ur = pc[0], vr = pc[1], wr = pc[2], tr = pc[3], rr, pr;
          if (d.bkopt[b] != RFLU_BCOPT_SUBSONIC) {
            pr = pc[4];
            rr = pr / (G.gcD * tr);
          } else {
            rr = rl;
            pr = rr * G.gcD * tr;
          }
          bb[0] = rr; bb[4] = rr * eo_dgpuvw(rr, G.gD, pr, ur, vr, wr);
          bb[1] = rr * ur; bb[2] = rr * vr; bb[3] = rr * wr; bb[5] = pr;
        } else {
          rind_totang(G, d.bkopt[b], pc, nx, ny, nz, rl, rul, rvl, rwl, bb);
        }
      }
      double rr = bb[0], rur = bb[1], rvr = bb[2], rwr = bb[3], rer = bb[4], pr = bb[5];
      double ql = (rul * nx + rvl * ny + rwl * nz) / rl - fs;
      double qr = (rur * nx + rvr * ny + rwr * nz) / rr - fs;
      f[0] = 0.5 * (ql * rl + qr * rr) * nm;
      f[1] = 0.5 * (ql * rul + pl * nx + qr * rur + pr * nx) * nm;
      f[2] = 0.5 * (ql * rvl + pl * ny + qr * rvr + pr * ny) * nm;
      f[3] = 0.5 * (ql * rwl + pl * nz + qr * rwr + pr * nz) * nm;
      f[4] = 0.5 * (ql * (rel + pl) + pl * fs + qr * (rer + pr) + pr * fs) * nm;
    }
  }
}

template <int F32>
__global__ void __launch_bounds__(BLK) k_flux(Dev d, Gas G, FluxPar P) {
  int s = xcd_block() * BLK + threadIdx.x;
  if (s >= d.S) return;
  double f[5];
  face_flux<F32>(d, G, P, s, f, P.doSrad ? d.srad + s : nullptr);
  double *o = d.flux + 5 * (size_t)s;
#pragma unroll
  for (int k = 0; k < 5; k++) o[k] = f[k];
}

// ---------------------------------------------------------------------------
// Residual gather, sources, RK update, dependent variables and checks.
// ---------------------------------------------------------------------------
struct UpdPar { int order, istage, fused, storeRhs; double ark, grk; };

// rhs = -diss (diss = 0) + face fluxes in ascending slot order
// (libfloflu/ConvectiveFluxes.F90:133-141 and the face/patch loops).
// Every face's flux is FP64 in flux[], in the FP32 evaluation mode too.
template <int F32>
__device__ __forceinline__ void gather_rhs(const Dev &d, int i, double rhs[5]) {
#pragma unroll
  for (int k = 0; k < 5; k++) rhs[k] = -0.0;
  for (int e = d.cfOff[i]; e < d.cfOff[i + 1]; e++) {
    int code = d.cfIdx[e];
    int slot = code >> 1;
    const double *f = d.flux + 5 * (size_t)slot;
    double f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3], f4 = f[4];
    if (code & 1) {
      rhs[0] = rhs[0] - f0; rhs[1] = rhs[1] - f1; rhs[2] = rhs[2] - f2; rhs[3] = rhs[3] - f3; rhs[4] = rhs[4] - f4;
    } else {
      rhs[0] = rhs[0] + f0; rhs[1] = rhs[1] + f1; rhs[2] = rhs[2] + f2; rhs[3] = rhs[3] + f3; rhs[4] = rhs[4] + f4;
    }
  }
}

// libfloflu/SourceTerms.F90:127-136 on the cv that SourceTermsMP sees (with
// particles, cv(1) carries CellGradientsMP's rho_g round trip)
__device__ __forceinline__ void gravity(const Gas &G, int order, const double r[8], double vol, double rhs[5],
                                        const double *phi = nullptr, int i = 0) {
  double rho = r[0], m1, m2, m3;
  if (order == 1) { m1 = r[1]; m2 = r[2]; m3 = r[3]; }
  else { m1 = rho * r[1]; m2 = rho * r[2]; m3 = rho * r[3]; }
  if (phi) {
    double vf = 1.0 - phi[i];
    rho = rho / vf * vf;
  }
  double rhoVol = vol * rho;
  rhs[1] = rhs[1] - G.accel[0] * rhoVol;
  rhs[2] = rhs[2] - G.accel[1] * rhoVol;
  rhs[3] = rhs[3] - G.accel[2] * rhoVol;
  rhs[4] = rhs[4] - vol * (G.accel[0] * m1 + G.accel[1] * m2 + G.accel[2] * m3);
}

// RKUpdateMP, RFLU_SetDependentVars and AfterUpdateMP for cell i, given its
// residual.
__device__ __forceinline__ void rk_update_cell(const Dev &d, const Gas &G, const UpdPar &P, int i, double vol,
                                               double rhs[5], double *recOut) {
  const int N = d.nC;
  // libfloflu/RkUpdateGeneric.F90:285-314 (Wray RK3, static grid)
  double fac = P.ark * (d.nInst ? d.iScal[4 * (size_t)d.cInst[i] + S_DTMIN] : d.scal[S_DTMIN]);
  double adtv = fac / vol;
  double cv[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    double c0 = d.cv[k * N + i];
    if (P.istage == 1) {
      cv[k] = c0 - adtv * rhs[k];
      d.rhsSum[k * N + i] = rhs[k];
    } else if (P.istage == 2) {
      double rs = d.rhsSum[k * N + i];
      cv[k] = c0 - adtv * (rhs[k] - P.grk * rs);
      d.rhsSum[k * N + i] = rhs[k];
    } else {
      double rs = d.rhsSum[k * N + i];
      cv[k] = c0 - adtv * (rhs[k] - P.grk * rs);
    }
  }
  // RFLU_SetDependentVars, then AfterUpdateMP: RFLU_CheckValidity,
  // RFLU_EnforceBounds, RFLU_CheckPositivity (libflu/*.F90, TCPERF branches)
  DV dv = dep_vars(G, cv[0], cv[1], cv[2], cv[3], cv[4], d.phi, i);
  {
    double rho = cv[0], rrho = 1.0 / rho;
    double u = rrho * cv[1], v = rrho * cv[2], w = rrho * cv[3];
    if (isnan(rho) || isnan(u) || isnan(v) || isnan(w) || isnan(dv.p) || isnan(dv.T)) atomicOr(d.err, E_NAN);
  }
  {
    double rr = cv[0], ir = 1.0 / rr;
    double u = ir * cv[1], v = ir * cv[2], w = ir * cv[3], Eo = ir * cv[4];
    double Vm2 = u * u + v * v + w * w;
    double p = (G.gD - 1.0) * rr * (Eo - 0.5 * Vm2 - 0.0);
    if (p <= 0.0 && fabs(p) <= 1.0e-12) {
      cv[4] = rr * eo_dgpuvw(rr, G.gD, 1.0e-12, u, v, w);
      atomicOr(d.err, E_BOUNDFIX);
    }
  }
  {
    double rho = cv[0], rrho = 1.0 / rho;
    double u = rrho * cv[1], v = rrho * cv[2], w = rrho * cv[3];
    double Vm2 = u * u + v * v + w * w;
    if (dv.p <= 0.0) {
      double Eo = 1.0e4 + 0.5 * Vm2 + 0.0;
      dv.p = (G.gD - 1.0) * rho * (Eo - 0.5 * Vm2 - 0.0);
      dv.T = dv.p / (rho * G.gcD);
      cv[4] = rho * Eo;
      atomicOr(d.err, E_POSFIX);
    }
    if (rho <= 0.0 || dv.p <= 0.0) atomicOr(d.err, E_NEG);
  }
#pragma unroll
  for (int k = 0; k < 5; k++) d.cv[k * N + i] = cv[k];
  double r[8];
  make_rec(P.order, cv[0], cv[1], cv[2], cv[3], cv[4], dv, d.cofg[i], d.cofg[N + i], d.cofg[2 * N + i], r);
  store_rec(recOut, i, r);
  if (d.rhog) d.rhog[i] = cv[0] / (1.0 - d.phi[i]);
}

template <int F32>
__global__ void __launch_bounds__(BLK) k_update(Dev d, Gas G, UpdPar P) {
  int i = xcd_block() * BLK + threadIdx.x;
  if (i >= d.nC) return;
  int N = d.nC;
  double rhs[5];
  double vol = d.vol[i];
  if (P.fused) {
    gather_rhs<F32>(d, i, rhs);
    if (G.accelOn) {
      double r[8];
      load_rec(d.rec, i, r);
      gravity(G, P.order, r, vol, rhs);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 5; k++) rhs[k] = d.rhs[k * N + i];
  }
  rk_update_cell(d, G, P, i, vol, rhs, d.rec);
}

// One fused stage after the gradients (ConvectiveFluxesMP + SourceTermsMP +
// RKUpdateMP for a stage without the dt evaluation, no particles): each cell
// evaluates its own faces' fluxes in its slot order and updates at once, so
// no per-face flux array passes through HBM.  Every interior face is
// evaluated twice (once per cell) by the same function from the same inputs,
// so the residual's bits are those of k_flux followed by k_update.  The
// records are read from rec and the updated ones written to recAlt.
template <int F32>
__global__ void __launch_bounds__(BLK) k_fu(Dev d, Gas G, FluxPar FP, UpdPar P) {
  int i = xcd_block() * BLK + threadIdx.x;
  if (i >= d.nC) return;
  double rhs[5];
#pragma unroll
  for (int k = 0; k < 5; k++) rhs[k] = -0.0;
  for (int e = d.cfOff[i]; e < d.cfOff[i + 1]; e++) {
    int code = d.cfIdx[e];
    double f[5];
    face_flux<F32>(d, G, FP, code >> 1, f, nullptr);
    if (code & 1) {
      rhs[0] = rhs[0] - f[0]; rhs[1] = rhs[1] - f[1]; rhs[2] = rhs[2] - f[2]; rhs[3] = rhs[3] - f[3]; rhs[4] = rhs[4] - f[4];
    } else {
      rhs[0] = rhs[0] + f[0]; rhs[1] = rhs[1] + f[1]; rhs[2] = rhs[2] + f[2]; rhs[3] = rhs[3] + f[3]; rhs[4] = rhs[4] + f[4];
    }
  }
  double vol = d.vol[i];
  if (G.accelOn) {
    double r[8];
    load_rec(d.rec, i, r);
    gravity(G, P.order, r, vol, rhs);
  }
  rk_update_cell(d, G, P, i, vol, rhs, d.recAlt);
}

// ConvectiveFluxesMP / SourceTermsMP as separate steps (rhs kept in HBM).
template <int F32>
__global__ void __launch_bounds__(BLK) k_gather(Dev d) {
  int i = xcd_block() * BLK + threadIdx.x;
  if (i >= d.nC) return;
  double rhs[5];
  gather_rhs<F32>(d, i, rhs);
#pragma unroll
  for (int k = 0; k < 5; k++) d.rhs[k * d.nC + i] = rhs[k];
}

__global__ void __launch_bounds__(BLK) k_source(Dev d, Gas G, int order) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= d.nC) return;
  int N = d.nC;
  double rhs[5], r[8];
#pragma unroll
  for (int k = 0; k < 5; k++) rhs[k] = d.rhs[k * N + i];
  load_rec(d.rec, i, r);
  gravity(G, order, r, d.vol[i], rhs, d.phi, i);
#pragma unroll
  for (int k = 0; k < 5; k++) d.rhs[k * N + i] = rhs[k];
}

// Plain streaming copy; used only to calibrate HBM traffic counters and to
// measure the copy bandwidth the board actually sustains.
__global__ void __launch_bounds__(BLK) k_copy_probe(double2 *__restrict__ dst, const double2 *__restrict__ src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLK) dst[i] = src[i];
}

// dv for output
__global__ void __launch_bounds__(BLK) k_dv_out(Dev d, Gas G, int order, double *dv, int i0 = 0, int n = -1) {
  int i = i0 + blockIdx.x * BLK + threadIdx.x;
  if (i >= (n < 0 ? d.nC : i0 + n)) return;
  int N = d.nC;
  double r[8];
  load_rec(d.rec, i, r);
  if (order == 1) {
    dv[i] = r[5]; dv[N + i] = r[6]; dv[2 * N + i] = r[7];
  } else {
    double T = r[4] / (r[0] / (1.0 - (d.phi ? d.phi[i] : 0.0)) * G.gcD);
    dv[i] = r[4]; dv[N + i] = T; dv[2 * N + i] = sqrt(G.gD * G.gcD * T);
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
  v &= 0x1fffffull;
  v = (v | v << 32) & 0x1f00000000ffffull;
  v = (v | v << 16) & 0x1f0000ff0000ffull;
  v = (v | v << 8) & 0x100f00f00f00f00full;
  v = (v | v << 4) & 0x10c30c30c30c30c3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}

// Order of n points (xyz triples) along a 3-D Morton curve (ties by index).
void morton_order(const double *xyz, int n, std::vector<int> &order) {
  order.resize(n);
  if (n == 0) return;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], xyz[3 * (size_t)i + k]); hi[k] = std::max(hi[k], xyz[3 * (size_t)i + k]); }
  double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
  double sc = ext > 0 ? (double)((1 << 21) - 1) / ext : 0.0;
  std::vector<std::pair<uint64_t, int>> key(n);
  for (int i = 0; i < n; i++) {
    uint64_t c[3];
    for (int k = 0; k < 3; k++) c[k] = (uint64_t)((xyz[3 * (size_t)i + k] - lo[k]) * sc);
    key[i] = {spread3(c[0]) | spread3(c[1]) << 1 | spread3(c[2]) << 2, i};
  }
  std::sort(key.begin(), key.end());
  for (int i = 0; i < n; i++) order[i] = key[i].second;
}

// Gradient tiles of T consecutive (Morton-ordered) cells: per tile the sorted
// union of its cells' stencil members (its own cells first, as slots 0..T-1),
// each stencil entry rewritten as a 16-bit LDS slot (bit 15: the member is
// also a face neighbour), and each interior cell-face entry's neighbour as a
// slot.  Fails when a tile has more than 0x7fff slots or a face neighbour is
// not a stencil member.
bool build_tiles(int T, int N, const std::vector<int> &csOff, const std::vector<int> &csMem,
                 const std::vector<int> &cfOff, const std::vector<int> &cfNInt, const std::vector<int> &cfNb,
                 std::vector<int> &tOff, std::vector<int> &tMem, std::vector<unsigned short> &sSlot,
                 std::vector<unsigned short> &cfSlot, int &nUmax) {
  int nT = (N + T - 1) / T;
  nUmax = 0;
  tOff.assign(nT + 1, 0);
  tMem.clear();
  sSlot.assign(csMem.size(), 0);
  cfSlot.assign(cfNb.size(), 0);
  std::vector<int> halo;
  auto slot_of = [&](int j, int c0, int c1) -> int {
    if (j >= c0 && j < c1) return j - c0;
    auto it = std::lower_bound(halo.begin(), halo.end(), j);
    if (it == halo.end() || *it != j) return -1;
    return (c1 - c0) + (int)(it - halo.begin());
  };
  for (int t = 0; t < nT; t++) {
    int c0 = t * T, c1 = std::min(N, c0 + T);
    halo.clear();
    for (int i = c0; i < c1; i++)
      for (int e = csOff[i]; e < csOff[i + 1]; e++) {
        int j = csMem[e] & 0x7fffffff;
        if (j < c0 || j >= c1) halo.push_back(j);
      }
    std::sort(halo.begin(), halo.end());
    halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
    int nU = (c1 - c0) + (int)halo.size();
    if (nU > 0x7fff) return false;
    nUmax = std::max(nUmax, nU);
    for (int i = c0; i < c1; i++) tMem.push_back(i);
    tMem.insert(tMem.end(), halo.begin(), halo.end());
    tOff[t + 1] = (int)tMem.size();
    for (int i = c0; i < c1; i++) {
      for (int e = csOff[i]; e < csOff[i + 1]; e++) {
        unsigned v = (unsigned)csMem[e];
        int slot = slot_of((int)(v & 0x7fffffffu), c0, c1);
        sSlot[e] = (unsigned short)(slot | ((v & 0x80000000u) ? 0x8000 : 0));
      }
      for (int e = cfOff[i]; e < cfOff[i] + cfNInt[i]; e++) {  // face neighbours, for the limiter's min/max
        int slot = slot_of(cfNb[e], c0, c1);
        if (slot < 0) return false;
        cfSlot[e] = (unsigned short)slot;
      }
    }
  }
  return true;
}

int set_err(rflu_handle *h, int code, const std::string &m) {
  if (h) h->err = m; else g_err = m;
  return code;
}

template <class T>
int dalloc(rflu_handle *h, T **p, size_t n, const T *src = nullptr) {
  API_LOCK;
  if (n == 0) n = 1;
  HIPCHK(h, hipMalloc((void **)p, n * sizeof(T)));
  h->devBytes += n * sizeof(T);
  if (src) {  // on the handle's own stream: nothing here touches the null stream
    HIPCHK(h, hipMemcpyAsync(*p, src, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return 0;
}

inline int nblk(int n) { return (n + BLK - 1) / BLK; }

hipEvent_t ev_get(rflu_handle *h) {
  if (!h->evPool.empty()) { hipEvent_t e = h->evPool.back(); h->evPool.pop_back(); return e; }
  API_LOCK;
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}

struct ProfScope {
  rflu_handle *h; const char *name; hipEvent_t a = nullptr, b = nullptr;
  ProfScope(rflu_handle *hh, const char *n) : h(hh), name(n) {
    if (h->profiling) { a = ev_get(h); b = ev_get(h); hipEventRecord(a, h->stream); }
  }
  ~ProfScope() {
    if (h->profiling) { hipEventRecord(b, h->stream); h->pending.push_back({name, {a, b}}); }
  }
};

void prof_collect(rflu_handle *h) {
  API_LOCK;
  for (auto &p : h->pending) {
    float ms = 0.f;
    hipEventSynchronize(p.second.second);
    hipEventElapsedTime(&ms, p.second.first, p.second.second);
    auto &q = h->prof[p.first];
    q.ms += ms;
    q.n += 1;
    h->evPool.push_back(p.second.first);
    h->evPool.push_back(p.second.second);
  }
  h->pending.clear();
}

DtPar dtpar(const rflu_handle *h, int finalize) {
  return DtPar{h->prm.cfl, h->prm.dtImposed, h->prm.dtMinLimit, h->prm.dTimeSystem, finalize};
}

int launch_dtmin(rflu_handle *h, int finalize, int storeCell) {
  if (h->d.nInst) {
    {
      ProfScope ps(h, "k_dtmin");
      hipLaunchKernelGGL(k_dtmin, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, 1);
    }
    ProfScope ps(h, "k_dtfinal");
    hipLaunchKernelGGL(k_dt_batch, dim3(h->d.nInst), dim3(1024), 0, h->stream, h->d, dtpar(h, 1));
    return 0;
  }
  {
    ProfScope ps(h, "k_dtmin");
    hipLaunchKernelGGL(k_dtmin, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, storeCell);
  }
  ProfScope ps(h, "k_dtfinal");
  hipLaunchKernelGGL(k_dtfinal, dim3(1), dim3(1024), 0, h->stream, h->d, dtpar(h, finalize), h->d.nBlk);
  return 0;
}

int launch_grad_impl(rflu_handle *h, const Dev &dv, int venkat, int guard);

// Launch a tile kernel template K<TB> over the gradient tiles, one workgroup
// of tileCells threads per tile.
#define TILE_LAUNCH(K, lds, ...)                                                                          \
  do {                                                                                                    \
    if (h->tileCells == TBIG)                                                                             \
      hipLaunchKernelGGL((K<TBIG>), dim3(h->nTiles), dim3(TBIG), (lds), h->stream, __VA_ARGS__);          \
    else                                                                                                  \
      hipLaunchKernelGGL((K<BLK>), dim3(h->nTiles), dim3(BLK), (lds), h->stream, __VA_ARGS__);            \
  } while (0)

int launch_grad(rflu_handle *h) {
  if (h->prm.spaceOrder < 2) return 0;
  if (h->prm.reconst == RFLU_RECONST_WENO_XYZ) {
    // unlimited LSQ gradients into gradRaw, then WENO-XYZ + positivity into grad
    Dev dv = h->d;
    dv.grad = h->d.gradRaw;
    {
      ProfScope ps(h, "k_grad");
      launch_grad_impl(h, dv, 0, 0);
    }
    ProfScope ps(h, "k_weno");
    if (h->gradVariant == 4 && h->tileLds && h->tileLds / 64 * 48 <= h->ldsMax)
      TILE_LAUNCH(k_weno_tile, h->tileLds / 64 * 48, h->d);
    else
      hipLaunchKernelGGL(k_weno, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d);
    return 0;
  }
  ProfScope ps(h, "k_grad");
  int venkat = h->prm.reconst == RFLU_RECONST_VENKAT ? 1 : 0, guard = h->prm.limiterGuard ? 1 : 0;
  return launch_grad_impl(h, h->d, venkat, guard);
}

int launch_grad_impl(rflu_handle *h, const Dev &dv, int venkat, int guard) {
  int nUmax = (int)(h->tileLds / 64);
  if (h->prm.fluxPrecision == 1) {
    TILE_LAUNCH(k_grad_tile_f32, h->tileLds / 64 * sizeof(SlotF), dv, venkat, guard);
    return 0;
  }
  if (h->picl && h->gradVariant == 4 && h->tileLds && h->tileLds + 8 * (size_t)nUmax <= h->ldsMax)
    TILE_LAUNCH(k_grad_tile_picl, h->tileLds + 8 * (size_t)nUmax, dv, venkat, guard, nUmax);
  else if (h->picl)
    hipLaunchKernelGGL((k_grad2<1, 1>), dim3(h->d.nBlk), dim3(BLK), 0, h->stream, dv, venkat, guard);
  else if (h->gradVariant == 4 && h->tileLds)
    TILE_LAUNCH(k_grad_tile_gas, h->tileLds, dv, venkat, guard, nUmax);
  else if (h->gradVariant == 2 || h->gradVariant == 4)
    hipLaunchKernelGGL((k_grad2<1, 0>), dim3(h->d.nBlk), dim3(BLK), 0, h->stream, dv, venkat, guard);
  else if (h->gradVariant == 24)
    hipLaunchKernelGGL((k_grad2<4, 0>), dim3(h->d.nBlk), dim3(BLK), 0, h->stream, dv, venkat, guard);
  else if (h->gradVariant == 25)
    hipLaunchKernelGGL((k_grad2<5, 0>), dim3(h->d.nBlk), dim3(BLK), 0, h->stream, dv, venkat, guard);
  else if (h->gradVariant == 26)
    hipLaunchKernelGGL((k_grad2<6, 0>), dim3(h->d.nBlk), dim3(BLK), 0, h->stream, dv, venkat, guard);
  else if (h->gradVariant == 3)
    hipLaunchKernelGGL(k_grad3, dim3((h->d.nC + G3C - 1) / G3C), dim3(256), 0, h->stream, dv, venkat, guard);
  else if (h->gradVariant == 0)
    hipLaunchKernelGGL(k_grad, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, dv, venkat, guard);
  else
    hipLaunchKernelGGL(k_grad_lds, dim3((h->d.nC + GCPB - 1) / GCPB), dim3(BLK), 0, h->stream, dv, venkat, guard);
  return 0;
}

int launch_flux(rflu_handle *h, int doSrad) {
  ProfScope ps(h, "k_flux");
  FluxPar P{h->prm.spaceOrder, h->prm.spaceDiscr, h->prm.patchOrder, doSrad};
  if (h->prm.fluxPrecision == 1)
    hipLaunchKernelGGL((k_flux<1>), dim3(nblk(h->d.S)), dim3(BLK), 0, h->stream, h->d, h->gas, P);
  else
    hipLaunchKernelGGL((k_flux<0>), dim3(nblk(h->d.S)), dim3(BLK), 0, h->stream, h->d, h->gas, P);
  return 0;
}

UpdPar updpar(const rflu_handle *h, int istage, int fused) {
  // libflu/RFLU_DerivedInputValues.F90:374-386
  const double ark[3] = {8.0 / 15.0, 5.0 / 12.0, 0.75};
  const double grk[3] = {0.0, 17.0 / 25.0, 5.0 / 9.0};
  return UpdPar{h->prm.spaceOrder, istage, fused, 0, ark[istage - 1], grk[istage - 1]};
}

int launch_update(rflu_handle *h, int istage, int fused) {
  ProfScope ps(h, "k_update");
  if (h->prm.fluxPrecision == 1)
    hipLaunchKernelGGL((k_update<1>), dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, h->gas, updpar(h, istage, fused));
  else
    hipLaunchKernelGGL((k_update<0>), dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, h->gas, updpar(h, istage, fused));
  return 0;
}

int launch_fu(rflu_handle *h, int istage) {
  ProfScope ps(h, "k_fu");
  FluxPar FP{h->prm.spaceOrder, h->prm.spaceDiscr, h->prm.patchOrder, 0};
  if (h->prm.fluxPrecision == 1)
    hipLaunchKernelGGL((k_fu<1>), dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, h->gas, FP, updpar(h, istage, 1));
  else
    hipLaunchKernelGGL((k_fu<0>), dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, h->gas, FP, updpar(h, istage, 1));
  return 0;
}

int picl_stage(rflu_handle *h, int istage);  // picl_gpu.inl
int launch_rhog(rflu_handle *h);
void picl_free(PiclState *s);

// One RK stage.  The stage-1 flux kernel reads exactly the state
// RFLU_TimeStepInviscid sees, so with withDt it also writes the face spectral
// radii and k_dtmin forms dt and dtMin before the update needs it.  With
// particles the stage follows libfloflu/RungeKuttaMP.F90:219-280 with
// PICL_TEMP_Runge between SourceTermsMP and RKUpdateMP.
void enqueue_stage(rflu_handle *h, int st, bool withDt) {
  if (h->picl) {
    launch_grad(h);
    launch_rhog(h);
    launch_flux(h, withDt);
    if (withDt) launch_dtmin(h, 1, 0);
    {
      ProfScope ps(h, "k_gather");
      hipLaunchKernelGGL((k_gather<0>), dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d);
    }
    if (h->gas.accelOn) hipLaunchKernelGGL(k_source, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, h->gas, h->prm.spaceOrder);
    picl_stage(h, st);
    launch_update(h, st, 0);
    return;
  }
  launch_grad(h);
  launch_flux(h, withDt);
  if (withDt) launch_dtmin(h, 1, 0);
  launch_update(h, st, 1);
}

// Stage 2 or 3 fused (h->fuse): the gradients, then k_fu.  Stage 2 reads the
// records in rec and writes recAlt, stage 3 the other way round, so a step
// ends with the records where it found them.
void enqueue_fused_stage(rflu_handle *h, int st) {
  double *a = h->d.rec, *b = h->d.recAlt;
  if (st == 3) { h->d.rec = b; h->d.recAlt = a; }
  launch_grad(h);
  launch_fu(h, st);
  h->d.rec = a;
  h->d.recAlt = b;
}

// One RFLU_TimeStepping iteration.
int enqueue_step(rflu_handle *h) {
  for (int st = 1; st <= 3; st++) {
    if (h->fuse && !h->picl && st >= 2)
      enqueue_fused_stage(h, st);
    else
      enqueue_stage(h, st, st == 1);
  }
  hipLaunchKernelGGL(k_advance_time, dim3(h->d.nInst ? (h->d.nInst + 255) / 256 : 1), dim3(h->d.nInst ? 256 : 64), 0,
                     h->stream, h->d);
  return 0;
}

int check_err(rflu_handle *h) {
  API_LOCK;
  int e = 0;
  HIPCHK(h, hipMemcpyAsync(&e, h->d.err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (e & E_NAN) return set_err(h, RFLU_GPU_ENAN, "RFLU_CheckValidity: invalid (NaN) variables detected");
  if (e & E_NEG) return set_err(h, RFLU_GPU_ENEGATIVE, "RFLU_CheckPositivity: negative positive-definite variables");
  if (e & E_INTERP) return set_err(h, RFLU_GPU_EINVAL, "ppiclf LocalInterp: a particle has no element centroid within reach");
  if (e & E_CONTACT) return set_err(h, RFLU_GPU_EINVAL, "collisions: a particle overlaps more than 32 neighbours");
  if (e & E_REROT)
    return set_err(h, RFLU_GPU_EINVAL, "ppiclf Torque_Hydro: Re rotational too large (ppiclf_user_Torque.f:125)");
  return 0;
}

}  // namespace (host helpers are file-local; C ABI below)

extern "C" {

const char *rflu_gpu_last_error(const rflu_handle *h) { return h ? h->err.c_str() : g_err.c_str(); }

// groupOff (batched ensemble): the mesh is the disjoint union of instances,
// instance k holding reference cells [groupOff[k], groupOff[k+1]).
// Per-patch constants of the inflow, farfield and injection patches, with the
// reference's expressions (libfloflu/MixtPerf_*.F90, BcondInflowPerf.F90,
// RFLU_ModRindStates.F90); the gas is the patch gas (R = 8314.3/M).
static void patch_constants(const rflu_mesh_desc *md, const Gas &G, double *pcon) {
  const double g = G.gD, rgas = G.gcD;
  for (int p = 0; p < md->nPatches; p++) {
    const double *v = md->bcVals ? md->bcVals + RFLU_BCVALS * (size_t)p : nullptr;
    double *c = pcon + PCON * (size_t)p;
    switch (md->bcType[p]) {
      case RFLU_BC_FARFIELD: {
        double mach = v[0], aoa = v[1], aos = v[2], pInf = v[3], tInf = v[4];
        double qi = mach * sqrt(g * rgas * tInf);
        c[0] = mach; c[1] = pInf; c[2] = pInf / (rgas * tInf); c[3] = qi;
        c[4] = qi * cos(aoa) * cos(aos);
        c[5] = qi * sin(aoa) * cos(aos);
        c[6] = qi * sin(aos);
        break;
      }
      case RFLU_BC_INJECTION: c[0] = v[0]; c[1] = v[1]; break;
      case RFLU_BC_INFLOW_VELTEMP: for (int k = 0; k < 5; k++) c[k] = v[k]; break;
      case RFLU_BC_INFLOW_TOTANG: {
        double ptot = v[0], ttot = v[1], bh = v[2], bv = v[3];
        double mach = md->bcOpt[p] != RFLU_BCOPT_SUBSONIC ? v[4] : 0.0;
        c[0] = ptot; c[1] = ttot; c[2] = cos(bh); c[3] = sin(bh); c[4] = cos(bv); c[5] = sin(bv);
        c[6] = mach; c[7] = v[5] != 0.0 ? 1.0 : 0.0;
        // supersonic / mixed state: MixtPerf_P_GMaPo, _T_GMaTo, _D_PRT, _C_GRT
        double pr = ptot / pow(1.0 + 0.5 * (g - 1.0) * mach * mach, g / (g - 1.0));
        double tr = ttot / (1.0 + 0.5 * (g - 1.0) * mach * mach);
        double rr = pr / (rgas * tr), ar = sqrt(g * rgas * tr);
        double ur = mach * ar * cos(bh) * cos(bv), vr = mach * ar * sin(bh), wr = mach * ar * cos(bh) * sin(bv);
        c[PC_TT_SUP] = rr; c[PC_TT_SUP + 1] = rr * ur; c[PC_TT_SUP + 2] = rr * vr; c[PC_TT_SUP + 3] = rr * wr;
        c[PC_TT_SUP + 4] = rr * (pr / (rr * (g - 1.0)) + 0.5 * (ur * ur + vr * vr + wr * wr) + 0.0);
        c[PC_TT_SUP + 5] = pr;
        double rs = ptot / (rgas * ttot);  // stagnation state at rest
        c[PC_TT_STAG] = rs;
        c[PC_TT_STAG + 1] = rs * (ptot / (rs * (g - 1.0)) + 0.5 * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0) + 0.0);
        c[PC_TT_STAG + 2] = ptot;
        c[PC_TT_A02] = g * rgas * ttot;  // MixtPerf_C2_GRT
        break;
      }
      default: break;
    }
  }
}

static int create_impl(const rflu_mesh_desc *md, const rflu_params *prm, rflu_handle **out,
                       const std::vector<int> *groupOff) {
  if (!md || !prm || !out) return set_err(nullptr, RFLU_GPU_EINVAL, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_err(nullptr, RFLU_GPU_ENODEV, "no HIP device available (the library has no CPU path)");
  if (prm->spaceOrder != 1 && prm->spaceOrder != 2) return set_err(nullptr, RFLU_GPU_EINVAL, "spaceOrder must be 1 or 2");
  if (prm->spaceDiscr != RFLU_DISCR_ROE && prm->spaceDiscr != RFLU_DISCR_HLLC && prm->spaceDiscr != RFLU_DISCR_AUSM &&
      prm->spaceDiscr != RFLU_DISCR_AUSMUP)
    return set_err(nullptr, RFLU_GPU_EINVAL, "spaceDiscr must be 1 (Roe), 3 (HLLC), 4 (AUSM+) or 5 (AUSM+up)");
  if (prm->fluxPrecision == 1 && prm->spaceDiscr == RFLU_DISCR_AUSMUP)
    return set_err(nullptr, RFLU_GPU_EINVAL, "fluxPrecision 1 supports spaceDiscr 1, 3 and 4 only");
  if (prm->patchOrder == 2 && prm->spaceOrder != 2)
    return set_err(nullptr, RFLU_GPU_EINVAL, "second-order patches need spaceOrder 2");
  if (prm->reconst != RFLU_RECONST_NONE && prm->reconst != RFLU_RECONST_WENO_XYZ &&
      prm->reconst != RFLU_RECONST_VENKAT)
    return set_err(nullptr, RFLU_GPU_EINVAL, "reconst must be 0, 2 (WENO-XYZ) or 11 (Venkatakrishnan)");
  if (prm->fluxPrecision == 1 && prm->reconst == RFLU_RECONST_WENO_XYZ)
    return set_err(nullptr, RFLU_GPU_EINVAL, "fluxPrecision 1 supports reconst 0 and 11 only");
  if (prm->fluxPrecision < 0 || prm->fluxPrecision > 1)
    return set_err(nullptr, RFLU_GPU_EINVAL, "fluxPrecision must be 0 (FP64) or 1 (FP32 gradients and fluxes)");
  if (prm->fluxPrecision == 1 && prm->spaceOrder != 2)
    return set_err(nullptr, RFLU_GPU_EINVAL, "fluxPrecision 1 needs spaceOrder 2");
  for (int p = 0; p < md->nPatches; p++) {
    int t = md->bcType[p];
    if (t < RFLU_BC_SLIPWALL || t > RFLU_BC_INFLOW_TOTANG)
      return set_err(nullptr, RFLU_GPU_EINVAL, "bcType must be one of RFLU_BC_* (1..7)");
    if ((t == RFLU_BC_OUTFLOW || t == RFLU_BC_INFLOW_VELTEMP || t == RFLU_BC_INFLOW_TOTANG) &&
        (md->bcOpt[p] < 0 || md->bcOpt[p] > 2))
      return set_err(nullptr, RFLU_GPU_EINVAL, "outflow/inflow bcOpt must be 0, 1 or 2");
    if (t >= RFLU_BC_FARFIELD && !md->bcVals)
      return set_err(nullptr, RFLU_GPU_EINVAL, "farfield, injection and inflow patches need bcVals");
  }
  rflu_handle *h = new rflu_handle();
  h->prm = *prm;
  if (const char *gv = getenv("ROCFLUPICL_GRAD_VARIANT")) h->gradVariant = atoi(gv);
  if (getenv("ROCFLUPICL_NO_XCD_REMAP")) {
    API_LOCK;
    int zero = 0;
    (void)hipMemcpyToSymbol(HIP_SYMBOL(g_xcdRemap), &zero, sizeof(int));
  }
  rflu::MeshInput in;
  in.nVert = md->nVert; in.xyz = md->xyz; in.nvc = md->nvc; in.nCells = md->nCells; in.c2v = md->c2v;
  in.nPatches = md->nPatches; in.bcType = md->bcType; in.bcOpt = md->bcOpt; in.pout = md->pout;
  in.nBFaces = md->nBFaces; in.nvf = md->nvf; in.bf2v = md->bf2v;
  std::string e = rflu::build_mesh(in, prm->spaceOrder > 1, h->mesh);
  if (!e.empty()) { delete h; return set_err(nullptr, RFLU_GPU_EMESH, "mesh: " + e); }
  const rflu::MeshHost &m = h->mesh;
  h->xyz.assign(md->xyz, md->xyz + 3 * (size_t)md->nVert);
  h->c2v.assign(md->c2v, md->c2v + (size_t)md->nvc * md->nCells);
  Gas &G = h->gas;
  G.cp = prm->refCp; G.g = prm->refGamma; G.gc = G.cp - G.cp / G.g;
  double mol = 8314.3 / G.gc;          // RFLU_SetGasVars: gv(MOL)
  G.gcD = 8314.3 / mol;                // MixtPerf_R_M
  G.gD = G.cp / (G.cp - G.gcD);        // MixtPerf_G_CpR
  G.epsentr = prm->epsentr; G.dissFact = prm->dissFact; G.accelOn = prm->accelOn;
  G.sigma = prm->sigma; G.kp = prm->kp; G.ku = prm->ku;
  for (int q = 0; q < 3; q++) G.accel[q] = prm->accel[q];
  Dev &d = h->d;
  d.nC = m.nCells; d.nF = m.nFaces; d.nB = m.nBFaces; d.S = m.nFaces + m.nBFaces; d.nBlk = nblk(d.nC);
  // non-blocking: work of other handles (other requests in flight) and
  // null-stream copies do not serialise against this handle's stream
  {
    API_LOCK;
    if (hipGetDevice(&h->device) != hipSuccess ||
        hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
      delete h;
      return set_err(nullptr, RFLU_GPU_EHIP, "hipStreamCreate failed");
    }
  }
  int rc = 0;
  // Device numbering: cells and faces are renumbered along a Morton curve of
  // their centroids so stencil and face neighbours sit close in HBM and L2.
  // Only storage order changes: every cell keeps its face list and stencil
  // list in the reference's order, so all sums round identically.
  const int S = d.S, N = d.nC, F = d.nF;
  morton_order(m.cofg.data(), N, h->perm);
  if (getenv("ROCFLUPICL_NO_MORTON"))
    for (int i = 0; i < N; i++) h->perm[i] = i;
  std::vector<int> grp;
  if (groupOff) {  // instances contiguous on the device, Morton order within each
    grp.assign(N, 0);
    for (size_t k = 0; k + 1 < groupOff->size(); k++)
      for (int o = (*groupOff)[k]; o < (*groupOff)[k + 1]; o++) grp[o] = (int)k;
    std::stable_sort(h->perm.begin(), h->perm.end(), [&](int a, int b) { return grp[a] < grp[b]; });
  }
  h->inv.assign(N, 0);
  for (int i = 0; i < N; i++) h->inv[h->perm[i]] = i;
  std::vector<int> fperm, bperm;
  morton_order(m.fc.data(), F, fperm);
  morton_order(m.fc.data() + 3 * (size_t)F, m.nBFaces, bperm);
  if (getenv("ROCFLUPICL_NO_MORTON")) {
    for (int q = 0; q < F; q++) fperm[q] = q;
    for (int q = 0; q < m.nBFaces; q++) bperm[q] = q;
  }
  for (int &x : bperm) x += F;
  std::vector<int> slotNew2Old(fperm);
  slotNew2Old.insert(slotNew2Old.end(), bperm.begin(), bperm.end());
  std::vector<int> slotOld2New(S);
  for (int q = 0; q < S; q++) slotOld2New[slotNew2Old[q]] = q;
  std::vector<int> f2cD(2 * (size_t)F), bf2cD(m.nBFaces), bkind(m.nBFaces), bkopt(m.nBFaces), bpatch(m.nBFaces);
  std::vector<double> bpout(m.nBFaces);
  for (int q = 0; q < F; q++) {
    int o = slotNew2Old[q];
    f2cD[2 * (size_t)q] = h->inv[m.f2c[2 * (size_t)o]];
    f2cD[2 * (size_t)q + 1] = h->inv[m.f2c[2 * (size_t)o + 1]];
  }
  for (int b = 0; b < m.nBFaces; b++) {
    int o = slotNew2Old[F + b] - F;
    int p = m.bPatch[o];
    bf2cD[b] = h->inv[m.bf2c[o]];
    static const int kindOf[8] = {0, 0, 0, BK_NOSLIP, BK_FARFIELD, BK_INJECTION, BK_VELTEMP, BK_TOTANG};
    int t = md->bcType[p];
    bkind[b] = t == RFLU_BC_SLIPWALL ? 0 : t == RFLU_BC_OUTFLOW ? 1 + md->bcOpt[p] : kindOf[t];
    bkopt[b] = md->bcOpt[p];
    bpatch[b] = p;
    bpout[b] = md->pout[p];
  }
  std::vector<double> pcon(PCON * (size_t)std::max(md->nPatches, 1), 0.0);
  patch_constants(md, G, pcon.data());
  std::vector<double> soa(7 * (size_t)S);
  for (int q = 0; q < S; q++) {
    int o = slotNew2Old[q];
    for (int k = 0; k < 4; k++) soa[k * (size_t)S + q] = m.fn[4 * (size_t)o + k];
    for (int k = 0; k < 3; k++) soa[(4 + k) * (size_t)S + q] = m.fc[3 * (size_t)o + k];
  }
  std::vector<int> cfOffD(N + 1, 0), cfIdxD(m.cfIdx.size());
  for (int i = 0; i < N; i++) {
    int o = h->perm[i];
    cfOffD[i + 1] = cfOffD[i] + (m.cfOff[o + 1] - m.cfOff[o]);
    for (int e = m.cfOff[o], t = cfOffD[i]; e < m.cfOff[o + 1]; e++, t++)
      cfIdxD[t] = (slotOld2New[m.cfIdx[e] >> 1] << 1) | (m.cfIdx[e] & 1);
  }
  std::vector<double> volD(N), epsqD(N), cofgSoa(3 * (size_t)N), rmSoa(6 * (size_t)N, 0.0);
  for (int i = 0; i < N; i++) {
    int o = h->perm[i];
    volD[i] = m.vol[o];
    epsqD[i] = m.epsq[o];
    for (int k = 0; k < 3; k++) cofgSoa[k * (size_t)N + i] = m.cofg[3 * (size_t)o + k];
    if (prm->spaceOrder > 1)
      for (int k = 0; k < 6; k++) rmSoa[k * (size_t)N + i] = m.moms[10 * (size_t)o + k];
  }
  std::vector<int> csOffD, csMemD, cfNIntD(N, 0), cfNbD(cfIdxD.size(), -1);
  std::vector<double> cfd(3 * cfIdxD.size(), 0.0);
  for (int i = 0; i < N; i++) {
    int o = h->perm[i];
    for (int e = m.cfOff[o], t = cfOffD[i]; e < m.cfOff[o + 1]; e++, t++) {
      int slot = m.cfIdx[e] >> 1;
      if (slot >= F) break;
      cfNIntD[i]++;
      cfNbD[t] = h->inv[m.f2c[2 * (size_t)slot + ((m.cfIdx[e] & 1) ? 0 : 1)]];
      for (int k = 0; k < 3; k++)  // cofg(cell) - fc(face), RFLU_ModLimiters.F90:443-448
        cfd[k * cfIdxD.size() + t] = m.cofg[3 * (size_t)o + k] - m.fc[3 * (size_t)slot + k];
    }
  }
  if (prm->spaceOrder > 1) {
    csOffD.assign(N + 1, 0);
    csMemD.resize(m.csMem.size());
    std::vector<int> nb;
    for (int i = 0; i < N; i++) {
      int o = h->perm[i];
      nb.clear();
      for (int e = m.cfOff[o]; e < m.cfOff[o + 1]; e++) {
        int slot = m.cfIdx[e] >> 1;
        if (slot >= F) break;
        nb.push_back(m.f2c[2 * (size_t)slot + ((m.cfIdx[e] & 1) ? 0 : 1)]);
      }
      int flagged = 0;
      csOffD[i + 1] = csOffD[i] + (m.csOff[o + 1] - m.csOff[o]);
      for (int e = m.csOff[o], t = csOffD[i]; e < m.csOff[o + 1]; e++, t++) {
        unsigned v = (unsigned)h->inv[m.csMem[e]];
        if (std::find(nb.begin(), nb.end(), m.csMem[e]) != nb.end()) { v |= 0x80000000u; flagged++; }
        csMemD[t] = (int)v;
      }
      if (flagged != (int)nb.size()) {
        std::string msg = "mesh: a face neighbour is missing from its cell's c2c stencil";
        rflu_gpu_destroy(h);
        return set_err(nullptr, RFLU_GPU_EMESH, msg);
      }
    }
  }
  // gradient tiles (see k_grad_tile): BLK consecutive cells per tile;
  // ROCFLUPICL_TILE=768 asks for TBIG-cell tiles (one 12-wave workgroup per
  // CU when its records fit; measured slower on the tet meshes: a CU then
  // idles while its only workgroup stages records)
  std::vector<int> tOff, tMem;
  std::vector<unsigned short> sSlot, cfSlotD;
  if (prm->spaceOrder > 1) {
    int dev = 0;
    hipDeviceProp_t prop;
    size_t ldsMax = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
                        ? prop.sharedMemPerBlock : 65536;
    h->ldsMax = ldsMax;
    const char *ts = getenv("ROCFLUPICL_TILE");
    int want = ts ? atoi(ts) : BLK;
    for (int T : {TBIG, BLK}) {
      if (T == TBIG && want != TBIG) continue;
      int nUmax = 0;
      if (build_tiles(T, N, csOffD, csMemD, cfOffD, cfNIntD, cfNbD, tOff, tMem, sSlot, cfSlotD, nUmax) &&
          (size_t)nUmax * 64 <= ldsMax) {
        h->tileLds = (size_t)nUmax * 64;
        h->tileCells = T;
        h->nTiles = (N + T - 1) / T;
        break;
      }
    }
  }
  double *faceSoa = nullptr;
  rc |= dalloc(h, &d.f2c, 2 * (size_t)d.nF, f2cD.data());
  rc |= dalloc(h, &d.bf2c, (size_t)d.nB, bf2cD.data());
  rc |= dalloc(h, &d.bkind, (size_t)d.nB, bkind.data());
  rc |= dalloc(h, &d.bkopt, (size_t)d.nB, bkopt.data());
  rc |= dalloc(h, &d.bpatch, (size_t)d.nB, bpatch.data());
  rc |= dalloc(h, &d.pcon, pcon.size(), pcon.data());
  rc |= dalloc(h, &d.bpout, (size_t)d.nB, bpout.data());
  rc |= dalloc(h, &faceSoa, 7 * (size_t)S, soa.data());
  rc |= dalloc(h, &d.cfOff, (size_t)N + 1, cfOffD.data());
  rc |= dalloc(h, &d.cfIdx, cfIdxD.size(), cfIdxD.data());
  rc |= dalloc(h, &d.vol, (size_t)N, volD.data());
  rc |= dalloc(h, &d.epsq, (size_t)N, epsqD.data());
  rc |= dalloc(h, &d.cofg, 3 * (size_t)N, cofgSoa.data());
  rc |= dalloc(h, &d.cfNInt, (size_t)N, cfNIntD.data());
  rc |= dalloc(h, &d.cfdx, cfd.size(), cfd.data());
  if (prm->spaceOrder > 1) {
    rc |= dalloc(h, &d.csOff, (size_t)N + 1, csOffD.data());
    rc |= dalloc(h, &d.csMem, csMemD.size(), csMemD.data());
    rc |= dalloc(h, &d.rm, 6 * (size_t)N, rmSoa.data());
    rc |= dalloc(h, &d.grad, 15 * (size_t)N);
    if (prm->reconst == RFLU_RECONST_WENO_XYZ) rc |= dalloc(h, &d.gradRaw, 15 * (size_t)N);
    if (prm->fluxPrecision == 1) {
      if (!h->tileLds) rc |= set_err(h, RFLU_GPU_EINVAL, "fluxPrecision 1 needs gradient tiles that fit in LDS");
      else {
        rc |= dalloc(h, &d.gradF, 16 * (size_t)N);
      }
    }
    if (h->tileLds) {
      rc |= dalloc(h, &d.tOff, tOff.size(), tOff.data());
      rc |= dalloc(h, &d.tMem, tMem.size(), tMem.data());
      rc |= dalloc(h, &d.sSlot, sSlot.size(), sSlot.data());
      rc |= dalloc(h, &d.cfSlot, cfSlotD.size(), cfSlotD.data());
    }
  }
  rc |= dalloc(h, &d.cv, 5 * (size_t)N);
  rc |= dalloc(h, &d.rhsSum, 5 * (size_t)N);
  rc |= dalloc(h, &d.rhs, 5 * (size_t)N);
  rc |= dalloc(h, &d.rec, 8 * (size_t)N);
  rc |= dalloc(h, &d.flux, 5 * (size_t)S);
  rc |= dalloc(h, &d.srad, (size_t)S);
  rc |= dalloc(h, &d.dtCell, (size_t)N);
  rc |= dalloc(h, &d.dvOut, 3 * (size_t)N);
  rc |= dalloc(h, &d.partial, (size_t)d.nBlk);
  rc |= dalloc(h, &d.scal, 8);
  if (groupOff) {
    int K = (int)groupOff->size() - 1;
    std::vector<int> cI(N);
    for (int i = 0; i < N; i++) cI[i] = grp[h->perm[i]];
    std::vector<double> zero(4 * (size_t)K, 0.0);
    d.nInst = K;
    rc |= dalloc(h, &d.cInst, (size_t)N, cI.data());
    rc |= dalloc(h, &d.iOff, groupOff->size(), groupOff->data());  // same counts, contiguous on the device
    rc |= dalloc(h, &d.iScal, 4 * (size_t)K, zero.data());
  }
  rc |= dalloc(h, &d.err, 1);
  if (rc) { std::string msg = h->err; rflu_gpu_destroy(h); return set_err(nullptr, RFLU_GPU_EHIP, msg); }
  d.cfdy = d.cfdx + cfIdxD.size(); d.cfdz = d.cfdx + 2 * cfIdxD.size();
  d.fnx = faceSoa; d.fny = faceSoa + S; d.fnz = faceSoa + 2 * (size_t)S; d.fnm = faceSoa + 3 * (size_t)S;
  d.fcx = faceSoa + 4 * (size_t)S; d.fcy = faceSoa + 5 * (size_t)S; d.fcz = faceSoa + 6 * (size_t)S;
  API_LOCK;
  (void)hipMemsetAsync(d.scal, 0, 8 * sizeof(double), h->stream);
  (void)hipMemsetAsync(d.err, 0, sizeof(int), h->stream);
  (void)hipMemsetAsync(d.rhsSum, 0, 5 * (size_t)N * sizeof(double), h->stream);
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    rflu_gpu_destroy(h);
    return set_err(nullptr, RFLU_GPU_EHIP, "set-up failed on the device");
  }
  *out = h;
  return 0;
}

int rflu_gpu_create(const rflu_mesh_desc *md, const rflu_params *prm, rflu_handle **out) {
  return create_impl(md, prm, out, nullptr);
}

int rflu_batch_create(int nInst, const rflu_mesh_desc *meshes, const rflu_params *prm, rflu_handle **out) {
  if (nInst < 1 || !meshes || !prm || !out) return set_err(nullptr, RFLU_GPU_EINVAL, "null argument or nInst < 1");
  *out = nullptr;
  const int nvc = meshes[0].nvc, nvf = meshes[0].nvf;
  std::vector<double> xyz;
  std::vector<int> c2v, bcType, bcOpt, nbf, bf2v, cellOff{0};
  std::vector<double> pout, bcVals;
  int nV = 0;
  for (int k = 0; k < nInst; k++) {
    const rflu_mesh_desc &m = meshes[k];
    if (m.nvc != nvc || m.nvf != nvf)
      return set_err(nullptr, RFLU_GPU_EINVAL, "a batch holds one cell type (all tets or all hexes)");
    xyz.insert(xyz.end(), m.xyz, m.xyz + 3 * (size_t)m.nVert);
    for (size_t q = 0; q < (size_t)m.nvc * m.nCells; q++) c2v.push_back(m.c2v[q] < 0 ? -1 : m.c2v[q] + nV);
    size_t nb = 0;
    for (int p = 0; p < m.nPatches; p++) {
      bcType.push_back(m.bcType[p]);
      bcOpt.push_back(m.bcOpt[p]);
      pout.push_back(m.pout[p]);
      for (int k = 0; k < RFLU_BCVALS; k++) bcVals.push_back(m.bcVals ? m.bcVals[RFLU_BCVALS * (size_t)p + k] : 0.0);
      nbf.push_back(m.nBFaces[p]);
      nb += (size_t)m.nBFaces[p];
    }
    for (size_t q = 0; q < nb * (size_t)m.nvf; q++) bf2v.push_back(m.bf2v[q] < 0 ? -1 : m.bf2v[q] + nV);
    nV += m.nVert;
    cellOff.push_back(cellOff.back() + m.nCells);
  }
  rflu_mesh_desc u{nV, xyz.data(), nvc, cellOff.back(), c2v.data(), (int)bcType.size(), bcType.data(),
                   bcOpt.data(), pout.data(), nbf.data(), nvf, bf2v.data(), bcVals.data()};
  if (int e = create_impl(&u, prm, out, &cellOff)) return e;
  (*out)->instCell = cellOff;
  return 0;
}

int rflu_batch_sizes(const rflu_handle *h, int inst, int *nCells, int *nInst) {
  if (!h || h->d.nInst == 0 || inst < 0 || inst >= h->d.nInst) return RFLU_GPU_EINVAL;
  if (nCells) *nCells = h->instCell[inst + 1] - h->instCell[inst];
  if (nInst) *nInst = h->d.nInst;
  return 0;
}

int rflu_batch_set_state(rflu_handle *h, int inst, const double *cv) {
  API_LOCK;
  if (!h || !cv || h->d.nInst == 0 || inst < 0 || inst >= h->d.nInst) return RFLU_GPU_EINVAL;
  const int off = h->instCell[inst], nk = h->instCell[inst + 1] - off, N = h->d.nC;
  std::vector<double> tmp(5 * (size_t)nk);
  for (int i = 0; i < nk; i++) {
    int o = h->perm[off + i] - off;
    for (int v = 0; v < 5; v++) tmp[v * (size_t)nk + i] = cv[v * (size_t)nk + o];
  }
  for (int v = 0; v < 5; v++)
    HIPCHK(h, hipMemcpyAsync(h->d.cv + v * (size_t)N + off, tmp.data() + v * (size_t)nk, nk * sizeof(double),
                             hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_records, dim3(nblk(nk)), dim3(BLK), 0, h->stream, h->d, h->gas, h->prm.spaceOrder, off, nk);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int rflu_batch_get_state(rflu_handle *h, int inst, double *cv, double *dv) {
  API_LOCK;
  if (!h || h->d.nInst == 0 || inst < 0 || inst >= h->d.nInst) return RFLU_GPU_EINVAL;
  const int off = h->instCell[inst], nk = h->instCell[inst + 1] - off, N = h->d.nC;
  std::vector<double> tmp(5 * (size_t)nk);
  auto fetch = [&](double *dst, const double *src, int nv) -> int {
    for (int v = 0; v < nv; v++)
      HIPCHK(h, hipMemcpyAsync(tmp.data() + v * (size_t)nk, src + v * (size_t)N + off, nk * sizeof(double),
                               hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int i = 0; i < nk; i++) {
      int o = h->perm[off + i] - off;
      for (int v = 0; v < nv; v++) dst[v * (size_t)nk + o] = tmp[v * (size_t)nk + i];
    }
    return 0;
  };
  if (cv)
    if (int e = fetch(cv, h->d.cv, 5)) return e;
  if (dv) {
    hipLaunchKernelGGL(k_dv_out, dim3(nblk(nk)), dim3(BLK), 0, h->stream, h->d, h->gas, h->prm.spaceOrder,
                       h->d.dvOut, off, nk);
    HIPCHK(h, hipGetLastError());
    if (int e = fetch(dv, h->d.dvOut, 3)) return e;
  }
  return 0;
}

int rflu_batch_get_time(rflu_handle *h, int inst, double *t, double *dtMin) {
  API_LOCK;
  if (!h || h->d.nInst == 0 || inst < 0 || inst >= h->d.nInst) return RFLU_GPU_EINVAL;
  double sc[4];
  HIPCHK(h, hipMemcpyAsync(sc, h->d.iScal + 4 * (size_t)inst, sizeof sc, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (t) *t = sc[S_TIME];
  if (dtMin) *dtMin = sc[S_DTMIN];
  return 0;
}

int rflu_gpu_destroy(rflu_handle *h) {
  API_LOCK;
  if (!h) return 0;
  Dev &d = h->d;
  void *ptrs[] = {d.f2c, d.bf2c, d.bkind, d.bkopt, d.bpatch, d.pcon, d.bpout, d.fnx, d.cfOff, d.cfIdx, d.cfNInt, d.cfdx, d.vol, d.epsq, d.cofg, d.csOff,
                  d.csMem, d.rm, d.grad, d.cv, d.rhsSum, d.rhs, d.rec, d.flux, d.srad, d.dtCell, d.dvOut, d.partial, d.scal,
                  d.err};
  for (void *p : ptrs) if (p) hipFree(p);
  for (void *p : {(void *)d.phi, (void *)d.rhog, (void *)d.gRaw, (void *)d.tOff, (void *)d.tMem, (void *)d.sSlot,
                  (void *)d.gradF, (void *)d.cInst, (void *)d.iOff, (void *)d.iScal,
                  (void *)d.gradRaw, (void *)d.cfSlot, (void *)d.recAlt})
    if (p) hipFree(p);
  picl_free(h->picl);
  if (h->graph) hipGraphExecDestroy(h->graph);
  for (auto e : h->evPool) hipEventDestroy(e);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int rflu_gpu_sizes(const rflu_handle *h, int *nC, int *nF, int *nB, int *nS) {
  if (!h) return RFLU_GPU_EINVAL;
  if (nC) *nC = h->mesh.nCells;
  if (nF) *nF = h->mesh.nFaces;
  if (nB) *nB = h->mesh.nBFaces;
  if (nS) *nS = h->mesh.csOff.empty() ? 0 : h->mesh.csOff.back();
  return 0;
}

int rflu_gpu_get_mesh(const rflu_handle *h, int *f2c, double *fn, double *fc, int *bf2c, double *vol, double *cofg) {
  if (!h) return RFLU_GPU_EINVAL;
  const rflu::MeshHost &m = h->mesh;
  if (f2c) memcpy(f2c, m.f2c.data(), m.f2c.size() * sizeof(int));
  if (fn) memcpy(fn, m.fn.data(), m.fn.size() * sizeof(double));
  if (fc) memcpy(fc, m.fc.data(), m.fc.size() * sizeof(double));
  if (bf2c) memcpy(bf2c, m.bf2c.data(), m.bf2c.size() * sizeof(int));
  if (vol) memcpy(vol, m.vol.data(), m.vol.size() * sizeof(double));
  if (cofg) memcpy(cofg, m.cofg.data(), m.cofg.size() * sizeof(double));
  return 0;
}

int rflu_gpu_get_stencils(const rflu_handle *h, int *off, int *memb, double *moms) {
  if (!h) return RFLU_GPU_EINVAL;
  const rflu::MeshHost &m = h->mesh;
  if (off && !m.csOff.empty()) memcpy(off, m.csOff.data(), m.csOff.size() * sizeof(int));
  if (memb && !m.csMem.empty()) memcpy(memb, m.csMem.data(), m.csMem.size() * sizeof(int));
  if (moms && !m.moms.empty()) memcpy(moms, m.moms.data(), m.moms.size() * sizeof(double));
  return 0;
}

// Host <-> device copies of SoA cell arrays with the cell renumbering applied.
static int to_device_cells(rflu_handle *h, double *dst, const double *src, int nvar) {
  size_t N = h->d.nC;
  h->stage.resize(nvar * N);
  for (int k = 0; k < nvar; k++)
    for (size_t i = 0; i < N; i++) h->stage[k * N + i] = src[k * N + h->perm[i]];
  HIPCHK(h, hipMemcpyAsync(dst, h->stage.data(), nvar * N * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}
static int from_device_cells(rflu_handle *h, double *dst, const double *src, int nvar, bool aos) {
  size_t N = h->d.nC;
  h->stage.resize(nvar * N);
  HIPCHK(h, hipMemcpyAsync(h->stage.data(), src, nvar * N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t i = 0; i < N; i++) {
    size_t o = h->perm[i];
    for (int k = 0; k < nvar; k++) {
      if (aos) dst[o * nvar + k] = h->stage[i * nvar + k];
      else dst[k * N + o] = h->stage[k * N + i];
    }
  }
  return 0;
}

int rflu_gpu_set_state(rflu_handle *h, const double *cv) {
  API_LOCK;
  if (!h || !cv) return RFLU_GPU_EINVAL;
  if (int e = to_device_cells(h, h->d.cv, cv, 5)) return e;
  HIPCHK(h, hipMemsetAsync(h->d.err, 0, sizeof(int), h->stream));
  hipLaunchKernelGGL(k_records, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, h->gas, h->prm.spaceOrder);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int rflu_gpu_get_state(rflu_handle *h, double *cv, double *dv) {
  API_LOCK;
  if (!h) return RFLU_GPU_EINVAL;
  if (cv)
    if (int e = from_device_cells(h, cv, h->d.cv, 5, false)) return e;
  if (dv) {
    hipLaunchKernelGGL(k_dv_out, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, h->gas, h->prm.spaceOrder, h->d.dvOut);
    if (int e = from_device_cells(h, dv, h->d.dvOut, 3, false)) return e;
  }
  return 0;
}

int rflu_gpu_get_rhs(rflu_handle *h, double *rhs) {
  API_LOCK;
  if (!h || !rhs) return RFLU_GPU_EINVAL;
  return from_device_cells(h, rhs, h->d.rhs, 5, false);
}

int rflu_gpu_get_grad(rflu_handle *h, double *grad) {
  API_LOCK;
  if (!h || !grad) return RFLU_GPU_EINVAL;
  if (h->prm.spaceOrder < 2) return set_err(h, RFLU_GPU_EINVAL, "no gradients at spaceOrder 1");
  if (h->prm.fluxPrecision == 1) {  // FP32 gradients, widened
    size_t N = h->d.nC;
    std::vector<float> gf(16 * N);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(gf.data(), h->d.gradF, gf.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < N; i++)
      for (int k = 0; k < 15; k++) grad[15 * (size_t)h->perm[i] + k] = (double)gf[16 * i + k];
    return 0;
  }
  return from_device_cells(h, grad, h->d.grad, 15, true);
}

int rflu_gpu_time_step_inviscid(rflu_handle *h, double *dtCell, double *dtMinLocal) {
  API_LOCK;
  if (!h) return RFLU_GPU_EINVAL;
  if (h->d.nInst) return set_err(h, RFLU_GPU_EINVAL, "a batch steps with rflu_gpu_step (per-instance dt)");
  {
    ProfScope ps(h, "k_srad");
    hipLaunchKernelGGL(k_srad, dim3(nblk(h->d.S)), dim3(BLK), 0, h->stream, h->d, h->gas, h->prm.spaceOrder);
  }
  launch_dtmin(h, 0, dtCell ? 1 : 0);
  HIPCHK(h, hipGetLastError());
  if (dtCell)
    if (int e = from_device_cells(h, dtCell, h->d.dtCell, 1, false)) return e;
  if (dtMinLocal) HIPCHK(h, hipMemcpyAsync(dtMinLocal, h->d.scal + S_DTLOCAL, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int rflu_gpu_minimum_time_step(rflu_handle *h, double *dtMin) {
  API_LOCK;
  if (!h) return RFLU_GPU_EINVAL;
  if (h->d.nInst) return set_err(h, RFLU_GPU_EINVAL, "a batch steps with rflu_gpu_step (per-instance dt)");
  hipLaunchKernelGGL(k_min_time_step, dim3(1), dim3(64), 0, h->stream, h->d, dtpar(h, 1));
  HIPCHK(h, hipGetLastError());
  if (dtMin) HIPCHK(h, hipMemcpyAsync(dtMin, h->d.scal + S_DTMIN, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int rflu_gpu_rk_init(rflu_handle *h, int istage) {
  if (!h || istage < 1 || istage > 3) return RFLU_GPU_EINVAL;
  return 0;
}

int rflu_gpu_cell_gradients(rflu_handle *h) {
  if (!h) return RFLU_GPU_EINVAL;
  launch_grad(h);
  if (h->picl) launch_rhog(h);
  HIPCHK(h, hipGetLastError());
  return 0;
}

int rflu_gpu_convective_fluxes(rflu_handle *h) {
  if (!h) return RFLU_GPU_EINVAL;
  launch_flux(h, 0);
  {
    ProfScope ps(h, "k_gather");
    if (h->prm.fluxPrecision == 1)
      hipLaunchKernelGGL((k_gather<1>), dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d);
    else
      hipLaunchKernelGGL((k_gather<0>), dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d);
  }
  HIPCHK(h, hipGetLastError());
  return 0;
}

int rflu_gpu_source_terms(rflu_handle *h) {
  if (!h) return RFLU_GPU_EINVAL;
  if (h->gas.accelOn) hipLaunchKernelGGL(k_source, dim3(h->d.nBlk), dim3(BLK), 0, h->stream, h->d, h->gas, h->prm.spaceOrder);
  HIPCHK(h, hipGetLastError());
  return 0;
}

int rflu_gpu_integrate_source_terms(rflu_handle *h) { return h ? 0 : RFLU_GPU_EINVAL; }

int rflu_gpu_rk_update(rflu_handle *h, int istage) {
  if (!h || istage < 1 || istage > 3) return RFLU_GPU_EINVAL;
  launch_update(h, istage, 0);
  HIPCHK(h, hipGetLastError());
  return check_err(h);
}

int rflu_gpu_rk_stage(rflu_handle *h, int istage) {
  if (!h || istage < 1 || istage > 3) return RFLU_GPU_EINVAL;
  enqueue_stage(h, istage, false);
  HIPCHK(h, hipGetLastError());
  return check_err(h);
}

static int capture_step(rflu_handle *h) {
  API_LOCK;
  if (h->graph) return 0;
  hipGraph_t g;
  HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  enqueue_step(h);
  HIPCHK(h, hipStreamEndCapture(h->stream, &g));
  HIPCHK(h, hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0));
  hipGraphDestroy(g);
  return 0;
}

int rflu_gpu_step_async(rflu_handle *h, int nSteps) {
  if (!h || nSteps < 0) return RFLU_GPU_EINVAL;
  if (int e = capture_step(h)) return e;
  for (int n = 0; n < nSteps; n++) HIPCHK(h, hipGraphLaunch(h->graph, h->stream));
  return 0;
}

int rflu_gpu_query(rflu_handle *h) {
  API_LOCK;
  if (!h) return RFLU_GPU_EINVAL;
  hipError_t e = hipStreamQuery(h->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  h->err = std::string("hipStreamQuery: ") + hipGetErrorString(e);
  return RFLU_GPU_EHIP;
}

int rflu_gpu_wait(rflu_handle *h) {
  if (!h) return RFLU_GPU_EINVAL;
  return check_err(h);
}

int rflu_gpu_device_bytes(const rflu_handle *h, size_t *bytes) {
  if (!h || !bytes) return RFLU_GPU_EINVAL;
  *bytes = h->devBytes;
  return 0;
}

int rflu_gpu_mem_info(size_t *freeBytes, size_t *totalBytes) {
  API_LOCK;
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) return set_err(nullptr, RFLU_GPU_EHIP, "hipMemGetInfo failed");
  if (freeBytes) *freeBytes = f;
  if (totalBytes) *totalBytes = t;
  return 0;
}

int rflu_gpu_step(rflu_handle *h, int nSteps, int useGraph) {
  if (!h || nSteps < 0) return RFLU_GPU_EINVAL;
  if (useGraph && !h->profiling) {
    if (int e = capture_step(h)) return e;
    for (int n = 0; n < nSteps; n++) HIPCHK(h, hipGraphLaunch(h->graph, h->stream));
  } else {
    for (int n = 0; n < nSteps; n++) enqueue_step(h);
    HIPCHK(h, hipGetLastError());
  }
  int rc = check_err(h);
  if (h->profiling) prof_collect(h);
  return rc;
}

int rflu_gpu_set_time(rflu_handle *h, double t) {
  API_LOCK;
  if (!h) return RFLU_GPU_EINVAL;
  if (h->d.nInst) return set_err(h, RFLU_GPU_EINVAL, "set the time of a batch per instance (not supported)");
  double s[8];
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(s, h->d.scal, sizeof s, hipMemcpyDeviceToHost));
  s[S_TIME] = t;
  s[S_TSR] = 0.0;
  HIPCHK(h, hipMemcpy(h->d.scal, s, sizeof s, hipMemcpyHostToDevice));
  return 0;
}

int rflu_gpu_get_time(rflu_handle *h, double *t, double *dtMin) {
  API_LOCK;
  if (!h) return RFLU_GPU_EINVAL;
  double s[8];
  HIPCHK(h, hipMemcpyAsync(s, h->d.scal, sizeof s, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (t) *t = s[S_TIME];
  if (dtMin) *dtMin = s[S_DTMIN];
  return 0;
}

int rflu_gpu_sync(rflu_handle *h) {
  API_LOCK;
  if (!h) return RFLU_GPU_EINVAL;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int rflu_gpu_bandwidth_probe(size_t bytes, int reps, double *gbps) {
  API_LOCK;
  if (bytes < 16 || reps < 1 || !gbps) return set_err(nullptr, RFLU_GPU_EINVAL, "bad probe arguments");
  size_t n = bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
    if (a) (void)hipFree(a);
    return set_err(nullptr, RFLU_GPU_EHIP, "probe allocation failed");
  }
  (void)hipMemset(a, 0, n * 16);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  int grid = 8192;
  hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(BLK), 0, 0, b, a, n);
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(BLK), 0, 0, b, a, n);
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *gbps = 2.0 * (double)(n * 16) * reps / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

int rflu_gpu_autotune(rflu_handle *h, int reps, int *variant, double *msPerLaunch) {
  API_LOCK;
  if (!h || reps < 1) return RFLU_GPU_EINVAL;
  if (variant) *variant = h->gradVariant;
  if (msPerLaunch) *msPerLaunch = 0.0;
  // gradient kernels that give the same bits: LDS tiles (4), one thread per
  // cell at 1, 4, 5, 6 waves per SIMD (2, 24, 25, 26), cooperative (3);
  // with particles the tile kernel (4) or the per-cell one (2); FP32 mode
  // and first order have a single gradient kernel (or none)
  std::vector<int> cand;
  if (h->prm.spaceOrder >= 2 && h->prm.fluxPrecision == 0) {
    if (h->tileLds) cand.push_back(4);
    cand.push_back(2);
    if (!h->picl) { cand.push_back(24); cand.push_back(25); cand.push_back(26); cand.push_back(3); }
  }
  bool prof = h->profiling;
  h->profiling = false;
  hipEvent_t a, b;
  HIPCHK(h, hipEventCreate(&a));
  HIPCHK(h, hipEventCreate(&b));
  int best = h->gradVariant;
  float bestMs = 1e30f;
  for (int c : cand) {
    h->gradVariant = c;
    launch_grad(h);  // warm-up (and first-touch of the code object)
    HIPCHK(h, hipEventRecord(a, h->stream));
    for (int r = 0; r < reps; r++) launch_grad(h);
    HIPCHK(h, hipEventRecord(b, h->stream));
    HIPCHK(h, hipEventSynchronize(b));
    float ms = 0.0f;
    HIPCHK(h, hipEventElapsedTime(&ms, a, b));
    ms /= (float)reps;
    if (ms < bestMs) { bestMs = ms; best = c; }
  }
  HIPCHK(h, hipGetLastError());
  h->gradVariant = best;
  // the stage after the gradients: k_flux + k_update against k_fu (same
  // bits), timed as stage 2 on this mesh; the state is saved first and
  // put back afterwards
  if (!h->picl) {
    int rc = 0;
    const size_t nC = h->d.nC;
    double *save = nullptr;
    int errSave = 0;
    if (!h->d.recAlt) rc = dalloc(h, &h->d.recAlt, 8 * nC);
    if (!rc) rc = dalloc(h, &save, 18 * nC);
    if (rc) return rc;
    HIPCHK(h, hipMemcpyAsync(save, h->d.cv, 5 * nC * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(save + 5 * nC, h->d.rhsSum, 5 * nC * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(save + 10 * nC, h->d.rec, 8 * nC * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(&errSave, h->d.err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    float ms2[2] = {0.0f, 0.0f};
    for (int f = 0; f < 2; f++) {
      for (int r = 0; r <= reps; r++) {
        if (r == 1) HIPCHK(h, hipEventRecord(a, h->stream));
        if (f) {
          launch_fu(h, 2);
        } else {
          launch_flux(h, 0);
          launch_update(h, 2, 1);
        }
      }
      HIPCHK(h, hipEventRecord(b, h->stream));
      HIPCHK(h, hipEventSynchronize(b));
      HIPCHK(h, hipEventElapsedTime(&ms2[f], a, b));
    }
    HIPCHK(h, hipMemcpyAsync(h->d.cv, save, 5 * nC * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d.rhsSum, save + 5 * nC, 5 * nC * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d.rec, save + 10 * nC, 8 * nC * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d.err, &errSave, sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipFree(save));
    h->devBytes -= 18 * nC * sizeof(double);
    h->fuse = ms2[1] < ms2[0] ? 1 : 0;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  HIPCHK(h, hipGetLastError());
  h->profiling = prof;
  if (h->graph) { (void)hipGraphExecDestroy(h->graph); h->graph = nullptr; }  // re-captured with the choice
  if (variant) *variant = best;
  if (msPerLaunch) *msPerLaunch = cand.empty() ? 0.0 : bestMs;
  return 0;
}

int rflu_gpu_set_fusion(rflu_handle *h, int on) {
  API_LOCK;
  if (!h || (on != 0 && on != 1)) return RFLU_GPU_EINVAL;
  if (on && h->picl) return set_err(h, RFLU_GPU_EINVAL, "fused stages need a region without particles");
  if (on && !h->d.recAlt) {
    int rc = dalloc(h, &h->d.recAlt, 8 * (size_t)h->d.nC);
    if (rc) return rc;
  }
  h->fuse = on;
  if (h->graph) { (void)hipGraphExecDestroy(h->graph); h->graph = nullptr; }
  return 0;
}

int rflu_gpu_get_fusion(const rflu_handle *h, int *on) {
  if (!h || !on) return RFLU_GPU_EINVAL;
  *on = h->fuse;
  return 0;
}

int rflu_gpu_set_profiling(rflu_handle *h, int on) {
  if (!h) return RFLU_GPU_EINVAL;
  h->profiling = on != 0;
  h->prof.clear();
  return 0;
}

int rflu_gpu_kernel_time(rflu_handle *h, const char *name, double *ms, int *n) {
  if (!h || !name) return RFLU_GPU_EINVAL;
  auto it = h->prof.find(name);
  if (it == h->prof.end() || it->second.n == 0) { if (ms) *ms = 0; if (n) *n = 0; return 0; }
  if (ms) *ms = it->second.ms / it->second.n;
  if (n) *n = it->second.n;
  return 0;
}

}  // extern "C"

#include "picl_gpu.inl"
