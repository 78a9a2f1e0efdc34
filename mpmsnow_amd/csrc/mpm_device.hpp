// mpm_device.hpp -- data layout shared by the HIP kernels and the host runtime.
//
// Particle store: structure-of-arrays, one float column per scalar component
// (column c of slot s at base[c * stride + s]); matrices are 9 columns in
// column-major element order. Slots are a permutation of the caller's
// particle indices that the runtime keeps body- and tile-sorted.
//
// Grid: the reference's dense per-body grid (n0 x n1 x n2 nodes, reference
// node index i + n0 (j + n1 k)) stored tile-major: the grid is cut into
// 4x4x4-node tiles, tile t = tx + T0 (ty + T1 tz) owns 64 consecutive nodes
// (local l = (i&3) + 4 (j&3) + 16 (k&3)); 3-vectors store each tile as three
// 64-float component runs ((t*3 + c) * 64 + l). A particle whose stencil base
// node lies in tile t only touches nodes in tiles t + {0,1}^3, so a workgroup
// that owns the particles of one tile stages an 8-tile (512-node) footprint
// in LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_math.hpp"

namespace mpm {

enum Column : int {
  C_X = 0,     // position (3)
  C_V = 3,     // velocity (3)
  C_F = 6,     // deformation gradient (9)
  C_FP = 15,   // plastic deformation gradient Fp (9)
  C_FI = 24,   // FinvTrans (9)
  C_R = 33,    // R (9)
  C_GI = 42,   // Ginv (9)
  C_M = 51,    // mass
  C_VOL = 52,  // volume
  C_J = 53,
  C_MU = 54,
  C_LAM = 55,
  C_MU0 = 56,  // per-particle Lame parameters before hardening
  C_LAM0 = 57,
  C_XI = 58,   // per-particle hardening coefficient
  NCOL = 59
};

struct PView {
  float* base;
  int64_t stride;
  __device__ __forceinline__ float* col(int c) const { return base + (int64_t)c * stride; }
  __device__ __forceinline__ float& at(int c, int s) const { return base[(int64_t)c * stride + s]; }
};

constexpr int kTile = 4;          // cells per tile edge
constexpr int kTileNodes = 64;    // nodes per tile
constexpr int kFoot = 512;        // nodes in a particle tile's 2x2x2-tile footprint
constexpr int kChunk = 256;       // max particles per workgroup chunk
constexpr int kBlock = 256;

struct GridView {
  int n[3];         // reference node counts
  int T[3];         // tiles per axis
  int NT;           // number of tiles
  float gmin[3];    // m_min
  float h;
  float fv[3];      // frame velocity
  const int* perm;  // tile-sorted particle slots
  const int4* chunks;  // (tile, start in perm, count, 0)
  int nchunks;
};

// error bits raised by kernels (reference exception messages are mapped by the host)
enum ErrorBits : int {
  ERR_NAN_MASS = 1,        // "nans in splatted masses!"
  ERR_NAN_VEL = 2,         // "nans in splatted velocities!"
  ERR_NAN_STRESS = 4,      // "nans in dEnergyDensitydF matrix!"
  ERR_NAN_MOMENTA = 8,     // "nan in explicit momenta!"
  ERR_HARDNESS = 16,       // "infinite hardness!!!"
  ERR_STENCIL = 32,        // internal: particle outside its tile's footprint
  ERR_NAN_PVEL = 64        // "nans in particle velocity data!"
};

// Conjugate-residual scalars, resident in device memory. Reductions are FP64,
// deterministic (fixed-order combination of per-block partials).
struct CRScalars {
  double bnorm2, threshold;
  double rz, ApdPAp, alpha, beta, rnorm2;
  double sums[8];          // totalMass, totalMomentum[3], collidedMass, collidedMomentum[3]
  float guess[3];
  int done;                // 1 = solve finished
  int zero_x;              // rhs was zero: x = 0 (ConjugateResiduals.cpp:34-38)
  int iter;                // current iteration index i
  int iterations;          // x updates performed
  int early_exit;          // zero denominator
  int max_iters;
  int errors;              // ErrorBits
  int active_nodes;        // nodes with m > 0 (measurement only)
  unsigned int ticket[4];  // last-block-done counters
};

}  // namespace mpm
