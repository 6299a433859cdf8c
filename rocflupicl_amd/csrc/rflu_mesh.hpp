// Host-side mesh pipeline of the Rocflu hot path (init-time, integer-exact).
//
// Builds, from a cell-to-vertex grid and per-patch boundary-face lists, the
// arrays the residual kernels consume: the interior face list in Rocflu's
// first-insertion order, boundary-face-to-cell maps, face normals/centroids,
// cell volumes/centroids, cell-to-cell LSQ stencils with their R-factor
// moments, and the cell-to-face gather lists that fix the per-cell residual
// summation order.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace rflu {

struct MeshInput {
  int nVert = 0;
  const double *xyz = nullptr;  // 3*nVert
  int nvc = 0;                  // 4 (tets) or 8 (hexes, or mixed cells padded with -1)
  int nCells = 0;
  const int *c2v = nullptr;     // nvc*nCells, 0-based
  int nPatches = 0;
  const int *bcType = nullptr;
  const int *bcOpt = nullptr;
  const double *pout = nullptr;
  const int *nBFaces = nullptr;
  int nvf = 0;
  const int *bf2v = nullptr;    // nvf per boundary face, patches concatenated
};

struct MeshHost {
  int nCells = 0, nFaces = 0, nBFaces = 0, nvc = 0;
  // interior faces (reference order), then boundary faces in patch order
  std::vector<int> f2c;          // 2*nFaces
  std::vector<int> f2v;          // 4*nFaces (-1 = none)
  std::vector<int> bf2c;         // nBFaces
  std::vector<int> bPatch;       // patch of each boundary face
  std::vector<double> fn;        // 4*(nFaces+nBFaces): nx,ny,nz,|S|
  std::vector<double> fc;        // 3*(nFaces+nBFaces)
  std::vector<double> vol;       // nCells
  std::vector<double> cofg;      // 3*nCells
  // cell -> face-slot gather list, ascending slot; entry = (slot<<1)|isC2
  std::vector<int> cfOff, cfIdx;
  // c2c stencils (order 2 only)
  std::vector<int> csOff, csMem;
  std::vector<double> moms;      // 10 per cell (11,12,22,13,23,33,14,24,34,44)
  std::vector<double> epsq;      // Venkatakrishnan eps^2 = (vol^(1/3))^3
  double stencilMinSvRatio = 0.0;
  int stencilMax = 0;
};

// Returns an empty string on success, else an error message.
std::string build_mesh(const MeshInput &in, bool needStencils, MeshHost &out);

}  // namespace rflu
