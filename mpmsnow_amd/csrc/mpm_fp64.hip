// mpm_fp64.hip -- the full-FP64 implicit MPM snow step on gfx950.
//
// Sim::advance of davidsminor/mpmsnow (src/MpmSim/Sim.cpp:57-134, paths relative to the reference
// checkout) with every scalar in double: particle state, grid quantities, shape weights, stresses,
// the conjugate-residual solve and its reductions. The library's step keeps its bulk in FP32; this
// one is the FP64 run its precision is measured against on one GPU at full size. It shares the
// device math of mpm_math.hpp, instantiated for double, and keeps the reference's structure:
//   * bodies: the flood fill of Sim::calculateBodies (Sim.cpp:136-203, |x_i - x_j|^2 < h^2) as a
//     union-find over a cell hash, bodies numbered by their smallest particle index, a particle
//     with no neighbour ballistic;
//   * one dense grid per body over its bounding box (Grid.cpp:156-213), all bodies of a step in
//     one node space so each launch serves every body, node (i, j, k) of a body at
//     offset + i + n0 (j + n1 k);
//   * splats (Grid.cpp:62-150, 387-463, 466-546, 549-632) as FP64 global atomic adds from
//     particles sorted by (body, cell);
//   * explicit momenta, collisions, collision velocities, initial guess, rhs and diagonal
//     preconditioner of Grid::updateGridVelocities (Grid.cpp:634-1016);
//   * the preconditioned conjugate-residual recurrences of ConjugateResiduals.cpp:19-160 with
//     SquareMagnitudeTermination (SquareMagnitudeTermination.cpp:13-28), one set of scalars per
//     body on the device;
//   * G2P, deformation gradients, plasticity and advection (Grid.cpp:1018-1081,
//     SnowConstitutiveModel.cpp:27-121, Sim.cpp:123-131).
// Sums by atomic adds take an order that varies from run to run: results agree run to run to
// the rounding of double (about 1e-16 relative), not bitwise. As in the library's step, every
// body's grid is built from the state at the start of the step, and the particles outside body 0
// first get the plasticity update that the reference's updateParticleData after body 0 gives them
// (from their unchanged F); repeated applications of that update are idempotent up to rounding.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpmsnow_gpu.h"
#include "mpm_host.hpp"
#include "mpm_math.hpp"

using namespace mpmh;

namespace m64 {

using mpm::ColliderSetT;
using mpm::ColliderT;

constexpr int kB = 256;
constexpr int kMaxGravity = 8;
constexpr int kHistCap = 4096;  // residual history entries kept per body and step
constexpr int kSupport = 2;     // B-spline support radius in cells (Grid.cpp:182)

// particle columns of the store, slot s of column c at base[c * n + s]; matrices column-major
enum : int {
  C_X = 0, C_V = 3, C_F = 6, C_FP = 15, C_FI = 24, C_R = 33, C_GI = 42,
  C_M = 51, C_VOL = 52, C_J = 53, C_MU = 54, C_LAM = 55, C_MU0 = 56, C_LAM0 = 57, C_XI = 58, NCOL = 59
};
enum : int { E_MASS = 1, E_VEL = 2, E_STRESS = 4, E_MOMENTA = 8, E_HARD = 16, E_PVEL = 32 };

struct P64 {
  double* base;
  int64_t n;
  __device__ __forceinline__ double& at(int c, int64_t s) const { return base[c * n + s]; }
  __device__ __forceinline__ void load(int c0, int k, int64_t s, double* out) const {
    for (int i = 0; i < k; ++i) out[i] = base[(c0 + i) * n + s];
  }
  __device__ __forceinline__ void store(int c0, int k, int64_t s, const double* in) const {
    for (int i = 0; i < k; ++i) base[(c0 + i) * n + s] = in[i];
  }
};

// one body's dense grid (Grid.cpp:171-213)
struct G64 {
  double gmin[3];
  double h;
  double fv[3];  // frame velocity (Sim.cpp:84-96)
  int n[3];
  int pad;
  int64_t off;   // first node of the body in the batched node space
};

// per-body conjugate-residual scalars and reduction accumulators
struct CR64 {
  double bnorm2, threshold, rz, alpha, beta, rn;
  double acc[3];   // <Ap, M^-1 Ap>, |r|^2, <rp, A rp>
  double sums[8];  // initial guess: total mass, total momentum, collided mass, collided momentum
  int iterations, done, early, histLen;
};

struct Grav {
  int n;
  double g[kMaxGravity][3];
};

__device__ __forceinline__ double wave_sum(double x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// Adds the K values of each thread into the accumulators of its body (dst + body * stride); a
// thread with body < 0 contributes nothing. A block whose contributing threads share one body
// adds its block sum with one atomic per value. Every thread of the block must call it.
template <int K>
__device__ void add_to_body(const double (&v)[K], int body, double* dst, int stride) {
  __shared__ int sb;
  __shared__ double part[K][kB / 64];
  __syncthreads();
  if (threadIdx.x == 0) sb = body;
  __syncthreads();
  const int b0 = sb;
  const int uniform = __syncthreads_and(body == b0 || body < 0);
  if (uniform) {
    if (b0 < 0) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double s = wave_sum(body >= 0 ? v[k] : 0.0);
      if (lane == 0) part[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int w = 0; w < kB / 64; ++w) s += part[k][w];
        atomicAdd(dst + (int64_t)b0 * stride + k, s);
      }
  } else if (body >= 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) atomicAdd(dst + (int64_t)body * stride + k, v[k]);
  }
}

// orderable encoding of a double for unsigned atomic min/max
__device__ __forceinline__ unsigned long long ord_enc(double d) {
  unsigned long long u = (unsigned long long)__double_as_longlong(d);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
double ord_dec(unsigned long long u) {
  u = (u & 0x8000000000000000ull) ? (u & 0x7FFFFFFFFFFFFFFFull) : ~u;
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}

// SquareMagnitude of a 3-vector as the reference's finite test sees it (product of the three
// values, CollisionObject.cpp / Grid.cpp "nan" checks)
__device__ __forceinline__ bool finite3(double a, double b, double c) { return isfinite(a * b * c); }

// The 64 nodes of a particle's cubic B-spline stencil (Grid.cpp:1105-1193): base node
// floor((x - min) / h) - 1 per axis, weights at j - frac for j = -1..2, in the reference's node
// order (x fastest). Calls fn(node, w, grad w) with the products of mpm::st_w / mpm::st_grad;
// only the x axis's weights stay in registers (the y and z ones are evaluated per row), which
// keeps the FP64 kernels out of scratch.
template <class Fn>
__device__ __forceinline__ void for_stencil(const G64& g, const double x[3], bool derivs, Fn&& fn) {
  int base[3];
  double fr[3];
  for (int d = 0; d < 3; ++d) {
    const double f = (x[d] - g.gmin[d]) / g.h;
    const int dp = (int)floor(f);
    fr[d] = f - (double)dp;
    base[d] = dp - 1;
  }
  double wx[4], dwx[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double t = (double)(a - 1) - fr[0];
    wx[a] = mpm::bspline_w(t);
    dwx[a] = derivs ? mpm::bspline_dw(t) / g.h : 0.0;
  }
#pragma unroll 1
  for (int c = 0; c < 4; ++c) {
    const double tz = (double)(c - 1) - fr[2];
    const double wz = mpm::bspline_w(tz), dwz = derivs ? mpm::bspline_dw(tz) / g.h : 0.0;
#pragma unroll 1
    for (int b = 0; b < 4; ++b) {
      const double ty = (double)(b - 1) - fr[1];
      const double wy = mpm::bspline_w(ty), dwy = derivs ? mpm::bspline_dw(ty) / g.h : 0.0;
      const int64_t row = g.off + base[0] + (int64_t)g.n[0] * ((base[1] + b) + (int64_t)g.n[1] * (base[2] + c));
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const double gw[3] = {-dwx[a] * wy * wz, -wx[a] * dwy * wz, -wx[a] * wy * dwz};
        fn(row + a, wx[a] * wy * wz, gw);
      }
    }
  }
}

// ---------------------------------------------------------------- particle kernels
// Sim.cpp:64-73: ballistic particles take the force fields' acceleration
__global__ void k_ballistic(P64 P, const int* label, int64_t n, Grav gv, double dt) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= n || label[s] != -1) return;
  const double m = P.at(C_M, s);
  double f[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < gv.n; ++i)
    for (int c = 0; c < 3; ++c) f[c] += gv.g[i][c] * m;
  for (int c = 0; c < 3; ++c) P.at(C_V + c, s) += dt * (f[c] / m);
}

// SnowConstitutiveModel::updateParticleData (SnowConstitutiveModel.cpp:27-121) for the slots with
// label != skip (mode 0) or label >= 0 (mode 1)
__global__ void k_plasticity(P64 P, const int* label, int64_t n, int mode, double thc, double ths, int* err) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= n) return;
  const int l = label[s];
  if (mode == 0 ? l == 0 : l < 0) return;
  double F[9], Fp[9], Fi[9], R[9], Gi[9], J, mu, lam;
  P.load(C_F, 9, s, F);
  P.load(C_FP, 9, s, Fp);
  const bool ok = mpm::snow_update(F, Fp, Fi, R, Gi, J, mu, lam, P.at(C_MU0, s), P.at(C_LAM0, s), P.at(C_XI, s), thc, ths);
  if (!ok) atomicOr(err, E_HARD);
  P.store(C_F, 9, s, F);
  P.store(C_FP, 9, s, Fp);
  P.store(C_FI, 9, s, Fi);
  P.store(C_R, 9, s, R);
  P.store(C_GI, 9, s, Gi);
  P.at(C_J, s) = J;
  P.at(C_MU, s) = mu;
  P.at(C_LAM, s) = lam;
}

// Folds each thread's position into the bounding box of its body (box + 6 body: orderable min[3],
// max[3]); like add_to_body, a block whose threads share one body folds its own box first.
__device__ void box_to_body(const double (&x)[3], int body, unsigned long long* box) {
  __shared__ int sb;
  __shared__ double part[6][kB / 64];
  __syncthreads();
  if (threadIdx.x == 0) sb = body;
  __syncthreads();
  const int b0 = sb;
  const int uniform = __syncthreads_and(body == b0 || body < 0);
  if (uniform) {
    if (b0 < 0) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double lo = body >= 0 ? x[c] : INFINITY, hi = body >= 0 ? x[c] : -INFINITY;
      for (int o = 32; o > 0; o >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, o, 64));
        hi = fmax(hi, __shfl_xor(hi, o, 64));
      }
      if (lane == 0) {
        part[c][wv] = lo;
        part[3 + c][wv] = hi;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int c = 0; c < 3; ++c) {
        double lo = part[c][0], hi = part[3 + c][0];
        for (int w = 1; w < kB / 64; ++w) {
          lo = fmin(lo, part[c][w]);
          hi = fmax(hi, part[3 + c][w]);
        }
        atomicMin(box + 6 * b0 + c, ord_enc(lo));
        atomicMax(box + 6 * b0 + 3 + c, ord_enc(hi));
      }
  } else if (body >= 0) {
    for (int c = 0; c < 3; ++c) {
      atomicMin(box + 6 * body + c, ord_enc(x[c]));
      atomicMax(box + 6 * body + 3 + c, ord_enc(x[c]));
    }
  }
}

// per body: bounding box and the momentum / mass sums of the frame velocity
__global__ void k_body_reduce(P64 P, const int* label, int64_t n, unsigned long long* bbox, double* mom) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  const int body = s < n ? label[s] : -1;
  double v[4] = {0.0, 0.0, 0.0, 0.0}, x[3] = {0.0, 0.0, 0.0};
  if (body >= 0) {
    const double m = P.at(C_M, s);
    for (int c = 0; c < 3; ++c) {
      x[c] = P.at(C_X + c, s);
      v[c] = P.at(C_V + c, s) * m;
    }
    v[3] = m;
  }
  box_to_body(x, body, bbox);
  add_to_body<4>(v, body, mom, 4);
}

// sort keys: (body, cell of the stencil base) for body particles, ballistic ones last
__global__ void k_grid_keys(P64 P, const int* label, int64_t n, const G64* grids, unsigned long long* key, int* val) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= n) return;
  const int b = label[s];
  val[s] = (int)s;
  if (b < 0) {
    key[s] = ~0ull;
    return;
  }
  const G64 g = grids[b];
  int c[3];
  for (int d = 0; d < 3; ++d) c[d] = (int)floor((P.at(C_X + d, s) - g.gmin[d]) / g.h);
  const unsigned long long cell =
      (unsigned long long)c[0] + (unsigned long long)g.n[0] * ((unsigned long long)c[1] + (unsigned long long)g.n[1] * c[2]);
  key[s] = ((unsigned long long)b << 32) | (cell & 0xFFFFFFFFull);
}

// new slot k takes old slot perm[k]
__global__ void k_permute(P64 src, P64 dst, const int* perm, const int* labelIn, int* labelOut, const int* origIn,
                          int* origOut, int* slotOf) {
  const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (k >= src.n) return;
  const int s = perm[k];
  for (int c = 0; c < NCOL; ++c) dst.at(c, k) = src.at(c, s);
  labelOut[k] = labelIn[s];
  const int o = origIn[s];
  origOut[k] = o;
  slotOf[o] = (int)k;
}

// first slot of every body in the (body, cell) order
__global__ void k_body_starts(const unsigned long long* key, int64_t nbp, int64_t* pstart) {
  const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (k >= nbp) return;
  const int b = (int)(key[k] >> 32);
  if (k == 0 || (int)(key[k - 1] >> 32) != b) pstart[b] = k;
}

__global__ void k_node_body(const G64* grids, const int64_t* nodeOff, int nb, int64_t NN, int* nodeBody) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= NN) return;
  int lo = 0, hi = nb - 1;  // last body with off <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (nodeOff[mid] <= i) lo = mid; else hi = mid - 1;
  }
  nodeBody[i] = lo;
}

// MassSplatter (Grid.cpp:62-99)
__global__ __launch_bounds__(kB) void k_p2g_mass(P64 P, const int* label, int64_t nbp, const G64* grids, double* mass) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= nbp) return;
  const G64 g = grids[label[s]];
  double x[3];
  P.load(C_X, 3, s, x);
  const double m = P.at(C_M, s);
  for_stencil(g, x, false, [&](int64_t idx, double w, const double*) { atomicAdd(mass + idx, m * w); });
}

// Grid.cpp:90-98: negative masses clamp to 0, non-finite ones are an error; counts each body's
// nodes with m > 0
__global__ void k_mass_fix(int64_t NN, const int* nodeBody, double* mass, double* active, int* err) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  const int body = i < NN ? nodeBody[i] : -1;
  double a[1] = {0.0};
  if (body >= 0) {
    double m = mass[i];
    if (m < 0.0) mass[i] = m = 0.0;
    if (!isfinite(m)) atomicOr(err, E_MASS);
    a[0] = m > 0.0 ? 1.0 : 0.0;
  }
  add_to_body<1>(a, body, active, 1);
}

// VelocitySplatter (Grid.cpp:102-150): velocities relative to the frame, each contribution
// divided by the node's mass
__global__ __launch_bounds__(kB) void k_p2g_vel(P64 P, const int* label, int64_t nbp, const G64* grids,
                                                const double* mass, double* vel) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= nbp) return;
  const G64 g = grids[label[s]];
  double x[3];
  P.load(C_X, 3, s, x);
  const double m = P.at(C_M, s);
  double dv[3];
  for (int c = 0; c < 3; ++c) dv[c] = P.at(C_V + c, s) - g.fv[c];
  for_stencil(g, x, false, [&](int64_t idx, double w, const double*) {
    const double cm = mass[idx];
    if (cm > 0.0) {
      const double sw = w * (m / cm);
      for (int k = 0; k < 3; ++k) atomicAdd(vel + 3 * idx + k, sw * dv[k]);
    }
  });
}

__global__ void k_vel_check(int64_t NN, const double* vel, double* prev, int* err) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= 3 * NN) return;
  const double v = vel[i];
  if (!isfinite(v)) atomicOr(err, E_VEL);
  prev[i] = v;
}

// Grid.cpp:265-292
__global__ __launch_bounds__(kB) void k_volumes(P64 P, const int* label, int64_t nbp, const G64* grids,
                                                const double* mass) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= nbp) return;
  const G64 g = grids[label[s]];
  double x[3];
  P.load(C_X, 3, s, x);
  const double cellVolume = g.h * g.h * g.h;
  double density = 0.0;
  for_stencil(g, x, false, [&](int64_t idx, double w, const double*) { density += w * mass[idx] / cellVolume; });
  P.at(C_VOL, s) = P.at(C_M, s) / density;
}

// the force fields' part of the grid forces (Grid.cpp:438-450), before the particles' stresses
__global__ void k_force_init(int64_t NN, const double* mass, Grav gv, double* frc) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= NN) return;
  double f[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < gv.n; ++k)
    for (int c = 0; c < 3; ++c) f[c] += gv.g[k][c] * mass[i];
  for (int c = 0; c < 3; ++c) frc[3 * i + c] = f[c];
}

// ForceSplatter (Grid.cpp:387-435): f_i -= V P F^T grad w_i
__global__ __launch_bounds__(kB) void k_forces(P64 P, const int* label, int64_t nbp, const G64* grids, double* frc,
                                               int* err) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= nbp) return;
  const G64 g = grids[label[s]];
  double F[9], R[9], Fi[9], Pm[9];
  P.load(C_F, 9, s, F);
  P.load(C_R, 9, s, R);
  P.load(C_FI, 9, s, Fi);
  mpm::snow_stress(F, R, Fi, P.at(C_J, s), P.at(C_MU, s), P.at(C_LAM, s), Pm);
  double n2 = 0.0;
  for (int i = 0; i < 9; ++i) n2 += Pm[i] * Pm[i];
  if (!isfinite(sqrt(n2))) atomicOr(err, E_STRESS);
  const double vol = P.at(C_VOL, s);
  double VP[9], fm[9];
  for (int i = 0; i < 9; ++i) VP[i] = vol * Pm[i];
  mpm::mat_mul_bt(VP, F, fm);
  double x[3];
  P.load(C_X, 3, s, x);
  for_stencil(g, x, true, [&](int64_t idx, double, const double* wg) {
    double t[3];
    mpm::mat_vec(fm, wg, t);
    for (int k = 0; k < 3; ++k) atomicAdd(frc + 3 * idx + k, -t[k]);
  });
}

// ForceDifferentialSplatter (Grid.cpp:549-618): df_i -= V dP(dF) F^T grad w_i, dF = sum_j dx_j
// grad w_j^T F; only the bodies still iterating when cr is given
__global__ __launch_bounds__(kB) void k_force_diff(P64 P, const int* label, int64_t nbp, const G64* grids,
                                                   const CR64* cr, const double* dx, double* df) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= nbp) return;
  const int body = label[s];
  if (cr && cr[body].done) return;
  const G64 g = grids[body];
  double x[3];
  P.load(C_X, 3, s, x);
  double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for_stencil(g, x, true, [&](int64_t idx, double, const double* wg) {
    const double* d = dx + 3 * idx;
    for (int col = 0; col < 3; ++col)
      for (int r = 0; r < 3; ++r) G[r + 3 * col] += d[r] * wg[col];
  });
  double F[9], dF[9];
  P.load(C_F, 9, s, F);
  mpm::mat_mul(G, F, dF);
  double R[9], Fi[9], Gi[9], dP[9];
  P.load(C_R, 9, s, R);
  P.load(C_FI, 9, s, Fi);
  P.load(C_GI, 9, s, Gi);
  mpm::snow_stress_differential(dF, R, Fi, Gi, P.at(C_J, s), P.at(C_MU, s), P.at(C_LAM, s), dP);
  const double vol = P.at(C_VOL, s);
  double VdP[9], fm[9];
  for (int i = 0; i < 9; ++i) VdP[i] = vol * dP[i];
  mpm::mat_mul_bt(VdP, F, fm);
  for_stencil(g, x, true, [&](int64_t idx, double, const double* wg) {
    double t[3];
    mpm::mat_vec(fm, wg, t);
    for (int k = 0; k < 3; ++k) atomicAdd(df + 3 * idx + k, -t[k]);
  });
}

// dFidXiSplatter (Grid.cpp:466-537): out_i[a] -= (V dP(e_a grad w_i^T F) F^T grad w_i)[a]
__global__ __launch_bounds__(kB) void k_dfidxi(P64 P, const int* label, int64_t nbp, const G64* grids, double* out) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= nbp) return;
  const G64 g = grids[label[s]];
  double x[3];
  P.load(C_X, 3, s, x);
  double F[9], R[9], Fi[9], Gi[9];
  P.load(C_F, 9, s, F);
  P.load(C_R, 9, s, R);
  P.load(C_FI, 9, s, Fi);
  P.load(C_GI, 9, s, Gi);
  const double J = P.at(C_J, s), mu = P.at(C_MU, s), lam = P.at(C_LAM, s), vol = P.at(C_VOL, s);
  for_stencil(g, x, true, [&](int64_t idx, double, const double* wg) {
#pragma unroll 1
    for (int e = 0; e < 3; ++e) {
      double E[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dF[9], dP[9], VdP[9], fm[9], t[3];
      for (int col = 0; col < 3; ++col) E[e + 3 * col] = wg[col];  // outer(e_a, grad w)
      mpm::mat_mul(E, F, dF);
      mpm::snow_stress_differential(dF, R, Fi, Gi, J, mu, lam, dP);
      for (int i = 0; i < 9; ++i) VdP[i] = vol * dP[i];
      mpm::mat_mul_bt(VdP, F, fm);
      mpm::mat_vec(fm, wg, t);
      atomicAdd(out + 3 * idx + e, -t[e]);
    }
  });
}

// Grid.cpp:1018-1081: PIC/FLIP velocities (alpha 0.95) and F <- (I + dt grad v) F
__global__ __launch_bounds__(kB) void k_g2p(P64 P, const int* label, int64_t nbp, const G64* grids, const double* vel,
                                            const double* prev, double dt) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= nbp) return;
  const G64 g = grids[label[s]];
  double x[3];
  P.load(C_X, 3, s, x);
  double vFlip[3], vPic[3] = {0.0, 0.0, 0.0}, dV[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  P.load(C_V, 3, s, vFlip);
  for_stencil(g, x, true, [&](int64_t idx, double w, const double* wg) {
    const double* vi = vel + 3 * idx;
    const double* vo = prev + 3 * idx;
    for (int k = 0; k < 3; ++k) {
      vFlip[k] += w * (vi[k] - vo[k]);
      vPic[k] += w * (vi[k] + g.fv[k]);
    }
    for (int col = 0; col < 3; ++col)
      for (int r = 0; r < 3; ++r) dV[r + 3 * col] += vi[r] * wg[col];
  });
  const double alpha = 0.95;
  for (int k = 0; k < 3; ++k) P.at(C_V + k, s) = alpha * vFlip[k] + (1.0 - alpha) * vPic[k];
  double A[9], F[9], Fn[9];
  for (int i = 0; i < 9; ++i) A[i] = dt * dV[i];
  A[0] += 1.0;
  A[4] += 1.0;
  A[8] += 1.0;
  P.load(C_F, 9, s, F);
  mpm::mat_mul(A, F, Fn);
  P.store(C_F, 9, s, Fn);
}

// Sim.cpp:123-131: collide with the colliders' velocities added, then x += v dt (every particle)
__global__ void k_advect(P64 P, int64_t n, const ColliderSetT<double>* cols, double dt) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= n) return;
  double x[3], v[3];
  P.load(C_X, 3, s, x);
  P.load(C_V, 3, s, v);
  const double zero[3] = {0.0, 0.0, 0.0};
  mpm::collide(*cols, v, x, zero, true);
  P.store(C_V, 3, s, v);
  for (int c = 0; c < 3; ++c) P.at(C_X + c, s) = x[c] + v[c] * dt;
}

// ---------------------------------------------------------------- node kernels
__device__ __forceinline__ void node_pos(const G64& g, int64_t i, double x[3]) {
  const int64_t l = i - g.off;
  const int a = (int)(l % g.n[0]);
  const int64_t r = l / g.n[0];
  const int b = (int)(r % g.n[1]);
  const int c = (int)(r / g.n[1]);
  x[0] = g.h * (double)a + g.gmin[0];
  x[1] = g.h * (double)b + g.gmin[1];
  x[2] = g.h * (double)c + g.gmin[2];
}

// Grid.cpp:778-818: remove the normal component at collided nodes, zero at multi-contact ones
__device__ __forceinline__ void project(int8_t flag, const double* nrm, double v[3]) {
  if (flag == -1) return;
  if (flag == -2) {
    v[0] = v[1] = v[2] = 0.0;
    return;
  }
  const double nd = nrm[0] * v[0] + nrm[1] * v[1] + nrm[2] * v[2];
  for (int c = 0; c < 3; ++c) v[c] -= nd * nrm[c];
}

// Grid.cpp:634-690 and 692-724: explicit momenta with collisions; collision velocities and
// the collider normals the projection uses
__global__ void k_explicit(int64_t NN, const int* nodeBody, const G64* grids, const double* mass, const double* vel,
                           const double* frc, const ColliderSetT<double>* cols, double dt, double* em, int8_t* flags,
                           double* vc, double* nrm, int* err) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= NN) return;
  const G64 g = grids[nodeBody[i]];
  const double M = mass[i];
  double ev[3], x[3];
  for (int c = 0; c < 3; ++c) ev[c] = vel[3 * i + c];
  if (M > 0.0)
    for (int c = 0; c < 3; ++c) ev[c] += dt * frc[3 * i + c] / M;
  node_pos(g, i, x);
  const int o = mpm::collide(*cols, ev, x, g.fv, false);
  flags[i] = (int8_t)o;
  double e[3];
  for (int c = 0; c < 3; ++c) em[3 * i + c] = e[c] = ev[c] * M;
  if (!finite3(e[0], e[1], e[2])) atomicOr(err, E_MOMENTA);
  double v[3] = {0.0, 0.0, 0.0}, n[3] = {0.0, 0.0, 0.0};
  if (o >= 0) {
    mpm::col_velocity(cols->o[o], x, v);
    for (int c = 0; c < 3; ++c) v[c] -= g.fv[c];
    mpm::col_normal(cols->o[o], x, n);
  }
  for (int c = 0; c < 3; ++c) {
    vc[3 * i + c] = v[c];
    nrm[3 * i + c] = n[c];
  }
}

// Grid.cpp:911-947: sums for the initial guess
__global__ void k_guess_sums(int64_t NN, const int* nodeBody, const double* mass, const double* em, const int8_t* flags,
                             const double* vc, CR64* cr) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  const int body = i < NN ? nodeBody[i] : -1;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (body >= 0) {
    const double M = mass[i];
    v[0] = M;
    for (int c = 0; c < 3; ++c) v[1 + c] = em[3 * i + c];
    if (flags[i] >= 0) {
      v[4] = M;
      for (int c = 0; c < 3; ++c) v[5 + c] = M * vc[3 * i + c];
    }
  }
  add_to_body<8>(v, body, reinterpret_cast<double*>(cr) + offsetof(CR64, sums) / sizeof(double),
                 sizeof(CR64) / sizeof(double));
}

__global__ void k_guess_fill(int64_t NN, const int* nodeBody, const CR64* cr, double* x) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= NN) return;
  const double* s = cr[nodeBody[i]].sums;
  double g[3] = {0.0, 0.0, 0.0};
  if (s[0] != 0.0) {
    if (s[4] == 0.0)
      for (int c = 0; c < 3; ++c) g[c] = s[1 + c] / s[0];
    else
      for (int c = 0; c < 3; ++c) g[c] = s[5 + c] / s[4];
  }
  for (int c = 0; c < 3; ++c) x[3 * i + c] = g[c];
}

// Grid.cpp:982-1002: b = momenta + dt^2 P(dF(vc))
__global__ void k_rhs(int64_t NN, const int8_t* flags, const double* nrm, const double* df, double dt, double* em,
                      int* err) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= NN) return;
  double d[3] = {df[3 * i], df[3 * i + 1], df[3 * i + 2]};
  project(flags[i], nrm + 3 * i, d);
  double e[3];
  for (int c = 0; c < 3; ++c) em[3 * i + c] = e[c] = em[3 * i + c] + dt * dt * d[c];
  if (!finite3(e[0], e[1], e[2])) atomicOr(err, E_MOMENTA);
}

// Grid.cpp:836-886: diag = m - dt^2 dF_i/dx_i, zero entries replaced by 1
__global__ void k_diag_fin(int64_t NN, const double* mass, double dt, double* diag) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= NN) return;
  const double s = -dt * dt;
  for (int c = 0; c < 3; ++c) {
    double d = diag[3 * i + c] * s + mass[i];
    if (d == 0.0) d = 1.0;
    diag[3 * i + c] = d;
  }
}

// ---- conjugate residuals (ConjugateResiduals.cpp:19-160), every body with its own scalars
// A v (Grid.cpp:749-776): first the projected copy of v (the particle pass reads it)
__global__ void k_a_pre(int64_t NN, const int* nodeBody, const CR64* cr, const int8_t* flags, const double* nrm,
                        const double* v, double* tmp, double* df) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= NN || (cr && cr[nodeBody[i]].done)) return;
  double t[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
  project(flags[i], nrm + 3 * i, t);
  for (int c = 0; c < 3; ++c) {
    tmp[3 * i + c] = t[c];
    df[3 * i + c] = 0.0;
  }
}

// then out = P(M tmp - dt^2 df); mode 1: also out2 = out and acc[2] += <rp, out>; mode 2:
// acc[2] += <rp, out>
__global__ void k_a_post(int64_t NN, const int* nodeBody, CR64* cr, const int8_t* flags, const double* nrm,
                         const double* mass, const double* tmp, const double* df, double dt, double* out, double* out2,
                         const double* rp, int mode) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  int body = i < NN ? nodeBody[i] : -1;
  if (body >= 0 && cr[body].done) body = -1;
  double acc[1] = {0.0};
  if (body >= 0) {
    double o[3];
    for (int c = 0; c < 3; ++c) o[c] = mass[i] * tmp[3 * i + c] - dt * dt * df[3 * i + c];
    project(flags[i], nrm + 3 * i, o);
    for (int c = 0; c < 3; ++c) {
      out[3 * i + c] = o[c];
      if (mode == 1) out2[3 * i + c] = o[c];
      if (mode != 0) acc[0] += rp[3 * i + c] * o[c];
    }
  }
  if (mode != 0) add_to_body<1>(acc, body, reinterpret_cast<double*>(cr) + offsetof(CR64, acc) / sizeof(double) + 2,
                                sizeof(CR64) / sizeof(double));
}

__global__ void k_b_norm(int64_t NN, const int* nodeBody, CR64* cr, const double* b) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  const int body = i < NN ? nodeBody[i] : -1;
  double acc[1] = {0.0};
  if (body >= 0)
    for (int c = 0; c < 3; ++c) acc[0] += b[3 * i + c] * b[3 * i + c];
  add_to_body<1>(acc, body, reinterpret_cast<double*>(cr) + offsetof(CR64, acc) / sizeof(double),
                 sizeof(CR64) / sizeof(double));
}

// r = b - A x, rp = M^-1 r, p = rp (ConjugateResiduals.cpp:50-62); x = 0 when |b| = 0
__global__ void k_cr_start(int64_t NN, const int* nodeBody, const CR64* cr, const double* b, const double* Ax,
                           const double* diag, double* x, double* r, double* rp, double* p) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= NN) return;
  const CR64& s = cr[nodeBody[i]];
  if (s.bnorm2 == 0.0) {
    for (int c = 0; c < 3; ++c) x[3 * i + c] = 0.0;
    return;
  }
  for (int c = 0; c < 3; ++c) {
    const double rr = b[3 * i + c] - Ax[3 * i + c];
    r[3 * i + c] = rr;
    rp[3 * i + c] = p[3 * i + c] = rr / diag[3 * i + c];
  }
}

// iteration start: (it > 0) p = rp + beta p, Ap = Ar + beta Ap; then acc[0] += <Ap, M^-1 Ap>
__global__ void k_cr_dir(int64_t NN, const int* nodeBody, CR64* cr, int it, const double* rp, const double* Ar,
                         const double* diag, double* p, double* Ap) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  int body = i < NN ? nodeBody[i] : -1;
  if (body >= 0 && cr[body].done) body = -1;
  double acc[1] = {0.0};
  if (body >= 0) {
    const double beta = cr[body].beta;
    for (int c = 0; c < 3; ++c) {
      double ap = Ap[3 * i + c];
      if (it > 0) {
        p[3 * i + c] = rp[3 * i + c] + beta * p[3 * i + c];
        Ap[3 * i + c] = ap = Ar[3 * i + c] + beta * ap;
      }
      acc[0] += ap * (ap / diag[3 * i + c]);
    }
  }
  add_to_body<1>(acc, body, reinterpret_cast<double*>(cr) + offsetof(CR64, acc) / sizeof(double),
                 sizeof(CR64) / sizeof(double));
}

// x = P(x + alpha p), rp -= alpha M^-1 Ap, r = M rp; acc[1] += |r|^2
__global__ void k_cr_x(int64_t NN, const int* nodeBody, CR64* cr, const int8_t* flags, const double* nrm,
                       const double* p, const double* Ap, const double* diag, double* x, double* rp, double* r) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  int body = i < NN ? nodeBody[i] : -1;
  if (body >= 0 && cr[body].done) body = -1;
  double acc[1] = {0.0};
  if (body >= 0) {
    const double alpha = cr[body].alpha;
    double xx[3];
    for (int c = 0; c < 3; ++c) xx[c] = x[3 * i + c] + alpha * p[3 * i + c];
    project(flags[i], nrm + 3 * i, xx);
    for (int c = 0; c < 3; ++c) {
      x[3 * i + c] = xx[c];
      const double d = diag[3 * i + c];
      const double q = rp[3 * i + c] - alpha * (Ap[3 * i + c] / d);
      rp[3 * i + c] = q;
      const double rr = q * d;
      r[3 * i + c] = rr;
      acc[0] += rr * rr;
    }
  }
  add_to_body<1>(acc, body, reinterpret_cast<double*>(cr) + offsetof(CR64, acc) / sizeof(double) + 1,
                 sizeof(CR64) / sizeof(double));
}

// per-body scalar steps: phase 0 |b|^2 -> threshold; 1 rz; 2 alpha; 3 termination; 4 beta
__global__ void k_cr_scalar(CR64* cr, int nb, int phase, int it, int maxIters, double tol, double* hist,
                            int* ndone) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  CR64& s = cr[b];
  if (phase == 0) {
    s.bnorm2 = s.acc[0];
    s.acc[0] = 0.0;
    s.threshold = tol * tol * s.bnorm2;
    if (s.bnorm2 == 0.0) {
      s.done = 1;
      atomicAdd(ndone, 1);
    }
    return;
  }
  if (s.done) return;
  if (phase == 1) {
    s.rz = s.acc[2];
    s.acc[2] = 0.0;
  } else if (phase == 2) {
    const double d = s.acc[0];
    s.acc[0] = 0.0;
    if (d == 0.0) {
      s.early = 1;
      s.done = 1;
      atomicAdd(ndone, 1);
      return;
    }
    s.alpha = s.rz / d;
  } else if (phase == 3) {
    s.iterations = it + 1;
    const double rn = s.acc[1];
    s.acc[1] = 0.0;
    if ((float)it >= (float)maxIters) {
      s.done = 1;
      atomicAdd(ndone, 1);
      return;
    }
    s.rn = rn;
    if (s.histLen < kHistCap) hist[(int64_t)b * kHistCap + s.histLen] = rn;
    ++s.histLen;
    if (rn < s.threshold) {
      s.done = 1;
      atomicAdd(ndone, 1);
    }
  } else {
    const double rzOld = s.rz;
    s.rz = s.acc[2];
    s.acc[2] = 0.0;
    if (rzOld == 0.0) {
      s.early = 1;
      s.done = 1;
      atomicAdd(ndone, 1);
      return;
    }
    s.beta = s.rz / rzOld;
  }
}

__global__ void k_finalize(int64_t n3, const double* x, const double* vc, double* vel) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n3) vel[i] = x[i] + vc[i];
}

// ---------------------------------------------------------------- bodies (Sim.cpp:136-203)
__global__ void k_check_pvel(P64 P, int64_t n, int* err) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= n) return;
  if (!finite3(P.at(C_V, s), P.at(C_V + 1, s), P.at(C_V + 2, s))) atomicOr(err, E_PVEL);
}

constexpr int kCellBits = 21;
constexpr int kCellOff = 1 << (kCellBits - 1);
__device__ __forceinline__ unsigned long long cell_key(int cx, int cy, int cz) {
  const unsigned long long m = (1ull << kCellBits) - 1;
  return (((unsigned long long)(cx + kCellOff) & m) << (2 * kCellBits)) |
         (((unsigned long long)(cy + kCellOff) & m) << kCellBits) | ((unsigned long long)(cz + kCellOff) & m);
}
__device__ __forceinline__ int cell_coord(double x, double cell) {
  double q = floor(x / cell);
  q = fmin(fmax(q, (double)(1 - kCellOff)), (double)(kCellOff - 2));  // keep neighbour keys in range
  return (int)q;
}

__global__ void k_cell_keys(P64 P, int64_t n, double cell, unsigned long long* key, int* val) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= n) return;
  key[s] = cell_key(cell_coord(P.at(C_X, s), cell), cell_coord(P.at(C_X + 1, s), cell),
                    cell_coord(P.at(C_X + 2, s), cell));
  val[s] = (int)s;
}

__device__ __forceinline__ int64_t lower_bound(const unsigned long long* a, int64_t n, unsigned long long k) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < k) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t prio(int i) {
  uint32_t x = (uint32_t)i * 0x9E3779B1u;
  x ^= x >> 15;
  x *= 0x85EBCA77u;
  x ^= x >> 13;
  return x;
}
// total order used for hooking: a root hooks under the root of lower priority
__device__ __forceinline__ bool before(int a, int b) {
  const uint32_t pa = prio(a), pb = prio(b);
  return pa < pb || (pa == pb && a < b);
}
__device__ __forceinline__ int ld(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ int find_root(int* parent, int x) {
  int p = ld(parent + x);
  while (p != x) {
    const int gp = ld(parent + p);
    if (gp != p) __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // path halving
    x = p;
    p = gp;
  }
  return x;
}

// the squared distance in the reference's rounding (separate products and sums, left to right)
__device__ __forceinline__ double dist2(const double* a, const double* b) {
#pragma clang fp contract(off)
  const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
  return d0 * d0 + d1 * d1 + d2 * d2;
}

// every pair closer than h is joined; hasNb marks particles with a neighbour other than themselves
__global__ void k_link(P64 P, int64_t n, double cell, double h2, const unsigned long long* skey, const int* sval,
                       int* parent, int* hasNb) {
  const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (k >= n) return;
  const int i = sval[k];
  double xi[3];
  P.load(C_X, 3, i, xi);
  const int cx = cell_coord(xi[0], cell), cy = cell_coord(xi[1], cell), cz = cell_coord(xi[2], cell);
  bool any = false;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const unsigned long long nk = cell_key(cx + dx, cy + dy, cz + dz);
        for (int64_t m = lower_bound(skey, n, nk); m < n && skey[m] == nk; ++m) {
          const int j = sval[m];
          if (j == i) continue;
          double xj[3];
          P.load(C_X, 3, j, xj);
          if (!(dist2(xi, xj) < h2)) continue;
          any = true;
          if (i > j) continue;  // each pair hooked once
          int a = find_root(parent, i), b = find_root(parent, j);
          while (a != b) {
            if (before(a, b)) {
              const int t = a;
              a = b;
              b = t;
            }
            // hook root a under b
            const int old = atomicCAS(parent + a, a, b);
            if (old == a) break;
            a = find_root(parent, old);
            b = find_root(parent, b);
          }
        }
      }
  if (any) hasNb[i] = 1;
}

__global__ void k_compress(int64_t n, int* parent, const int* hasNb, const int* orig, int* minOrig) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= n) return;
  const int r = find_root(parent, (int)s);
  parent[s] = r;
  if (hasNb[s]) atomicMin(minOrig + r, orig[s]);
}
__global__ void k_mark(int64_t n, const int* parent, const int* hasNb, const int* minOrig, int* mark) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= n) return;
  if (hasNb[s] && parent[s] == (int)s) mark[minOrig[s]] = 1;
}
// bodies numbered by their smallest caller index (the reference's flood order meets them so)
__global__ void k_assign(int64_t n, const int* parent, const int* hasNb, const int* minOrig, const int* markPos,
                         int* label, int* nballistic) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= n) return;
  if (hasNb[s]) {
    label[s] = markPos[minOrig[parent[s]]];
  } else {
    label[s] = -1;
    atomicAdd(nballistic, 1);
  }
}

template <class T> __global__ void k_fill(T* p, T v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_iota(int* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) p[i] = (int)i;
}
__global__ void k_store_init(P64 P) {
  // v = 0, F = Fp = FinvTrans = R = Ginv = I, volume = 0, J = 1
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s >= P.n) return;
  for (int c = 0; c < 3; ++c) P.at(C_V + c, s) = 0.0;
  for (int m = 0; m < 5; ++m)
    for (int k = 0; k < 9; ++k) P.at(C_F + 9 * m + k, s) = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
  P.at(C_VOL, s) = 0.0;
  P.at(C_J, s) = 1.0;
}
__global__ void k_set_col(P64 P, int c0, int ncomp, const double* aos, const int* slotOf) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= P.n) return;
  const int s = slotOf[i];
  for (int c = 0; c < ncomp; ++c) P.at(c0 + c, s) = aos[i * ncomp + c];
}
__global__ void k_get_col(P64 P, int c0, int ncomp, double* aos, const int* slotOf) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= P.n) return;
  const int s = slotOf[i];
  for (int c = 0; c < ncomp; ++c) aos[i * ncomp + c] = P.at(c0 + c, s);
}

inline unsigned blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kB - 1) / kB); }

struct Var {
  const char* name;
  int col, ncomp;
};
constexpr Var kVars[] = {{"p", C_X, 3},    {"v", C_V, 3},         {"F", C_F, 9},    {"Fp", C_FP, 9},
                         {"FinvTrans", C_FI, 9}, {"R", C_R, 9},   {"Ginv", C_GI, 9}, {"m", C_M, 1},
                         {"volume", C_VOL, 1},   {"J", C_J, 1},   {"mu", C_MU, 1},  {"lambda", C_LAM, 1},
                         {"mu0", C_MU0, 1},      {"lambda0", C_LAM0, 1}, {"xi", C_XI, 1}};
const Var& find_var(const char* name) {
  if (name)
    for (const Var& v : kVars)
      if (std::strcmp(v.name, name) == 0) return v;
  throw InvalidArg(std::string("unknown particle variable '") + (name ? name : "") + "'");
}

struct BodySolve {
  int iterations = 0, early = 0, particles = 0, active = 0;
  double bnorm2 = 0.0, rn = 0.0;
  std::vector<double> history;
};

}  // namespace m64

using namespace m64;

struct mpm_sim64 {
  int device = 0;
  hipStream_t s = nullptr;
  std::string error;
  int64_t n = 0;
  DevBuf store, store2, label, label2, orig, orig2, slotOf;
  bool haveModel = false, initialised = false;
  double h = 0.0, thc = 0.0, ths = 0.0;
  ColliderSetT<double> cols{};
  std::vector<std::unique_ptr<DevBuf>> colData;
  DevBuf colsDev;
  Grav grav{};
  // bodies
  int nb = 0;
  int64_t nballistic = 0;
  DevBuf keys, keys2, vals, vals2, parent, hasNb, minOrig, mark, markPos, cubTmp;
  // grids of the current step
  std::vector<G64> hg;
  std::vector<int64_t> hNodeOff;
  DevBuf grids, nodeOffDev, pstart, bodyBox, bodyMom, nodeBody, activeNodes;
  int64_t NN = 0, nbp = 0;
  DevBuf mass, vel, prev, frc, em, vc, nrm, flags, diag, x, r, rp, p, Ap, Ar, tmp, df;
  DevBuf cr, hist, errDev, counters;
  PinnedBuf pinned;
  std::vector<BodySolve> lastSolves;

  P64 view() { return P64{store.as<double>(), n}; }
  void sync() { HIPCHECK(hipStreamSynchronize(s)); }
  int read_err() {
    int e = 0;
    HIPCHECK(hipMemcpyAsync(&e, errDev.p, sizeof(int), hipMemcpyDeviceToHost, s));
    sync();
    return e;
  }
  void check_err() {
    const int e = read_err();
    if (!e) return;
    HIPCHECK(hipMemsetAsync(errDev.p, 0, sizeof(int), s));
    if (e & E_PVEL) throw std::runtime_error("nans in particle velocity data!");
    if (e & E_MASS) throw std::runtime_error("nans in splatted masses!");
    if (e & E_VEL) throw std::runtime_error("nans in splatted velocities!");
    if (e & E_STRESS) throw std::runtime_error("nans in dEnergyDensitydF matrix!");
    if (e & E_MOMENTA) throw std::runtime_error("nan in explicit momenta!");
    if (e & E_HARD) throw std::runtime_error("infinite hardness!!!");
  }

  template <class T> void sort_pairs(DevBuf& kin, DevBuf& kout, DevBuf& vin, DevBuf& vout, int64_t count, int endBit) {
    size_t bytes = 0;
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin.as<T>(), kout.as<T>(), vin.as<int>(),
                                                vout.as<int>(), (int)count, 0, endBit, s));
    cubTmp.ensure(bytes);
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(cubTmp.p, bytes, kin.as<T>(), kout.as<T>(), vin.as<int>(),
                                                vout.as<int>(), (int)count, 0, endBit, s));
  }

  // Sim::calculateBodies (Sim.cpp:136-203)
  void calculate_bodies() {
    const P64 P = view();
    k_check_pvel<<<blocks(n), kB, 0, s>>>(P, n, errDev.as<int>());
    check_err();
    for (DevBuf* b : {&keys, &keys2}) b->ensure(sizeof(unsigned long long) * n);
    for (DevBuf* b : {&vals, &vals2, &parent, &hasNb, &minOrig, &mark, &markPos}) b->ensure(sizeof(int) * n);
    const double cell = 1.01 * h;  // neighbours closer than h lie in adjacent cells
    k_cell_keys<<<blocks(n), kB, 0, s>>>(P, n, cell, keys.as<unsigned long long>(), vals.as<int>());
    sort_pairs<unsigned long long>(keys, keys2, vals, vals2, n, 3 * kCellBits);
    k_iota<<<blocks(n), kB, 0, s>>>(parent.as<int>(), n);
    k_fill<int><<<blocks(n), kB, 0, s>>>(hasNb.as<int>(), 0, n);
    k_fill<int><<<blocks(n), kB, 0, s>>>(minOrig.as<int>(), INT_MAX, n);
    k_fill<int><<<blocks(n), kB, 0, s>>>(mark.as<int>(), 0, n);
    k_link<<<blocks(n), kB, 0, s>>>(P, n, cell, h * h, keys2.as<unsigned long long>(), vals2.as<int>(),
                                    parent.as<int>(), hasNb.as<int>());
    k_compress<<<blocks(n), kB, 0, s>>>(n, parent.as<int>(), hasNb.as<int>(), orig.as<int>(), minOrig.as<int>());
    k_mark<<<blocks(n), kB, 0, s>>>(n, parent.as<int>(), hasNb.as<int>(), minOrig.as<int>(), mark.as<int>());
    size_t bytes = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, mark.as<int>(), markPos.as<int>(), (int)n, s));
    cubTmp.ensure(bytes);
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(cubTmp.p, bytes, mark.as<int>(), markPos.as<int>(), (int)n, s));
    counters.ensure(4 * sizeof(int));
    HIPCHECK(hipMemsetAsync(counters.p, 0, 4 * sizeof(int), s));
    k_assign<<<blocks(n), kB, 0, s>>>(n, parent.as<int>(), hasNb.as<int>(), minOrig.as<int>(), markPos.as<int>(),
                                      label.as<int>(), counters.as<int>());
    int h2[3] = {0, 0, 0};
    HIPCHECK(hipMemcpyAsync(&h2[0], markPos.as<int>() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(&h2[1], mark.as<int>() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(&h2[2], counters.p, sizeof(int), hipMemcpyDeviceToHost, s));
    sync();
    nb = h2[0] + h2[1];
    nballistic = h2[2];
  }

  // per-body grids of the state at the start of the step (Grid.cpp:156-213), the store sorted by
  // (body, cell), node arrays sized for every body
  void build_grids(bool frame) {
    const P64 P = view();
    bodyBox.ensure(sizeof(unsigned long long) * 6 * nb);
    bodyMom.ensure(sizeof(double) * 4 * nb);
    std::vector<unsigned long long> initBox(6 * (size_t)nb);
    for (int b = 0; b < nb; ++b)
      for (int c = 0; c < 3; ++c) {
        initBox[6 * b + c] = ~0ull;
        initBox[6 * b + 3 + c] = 0ull;
      }
    HIPCHECK(hipMemcpyAsync(bodyBox.p, initBox.data(), sizeof(unsigned long long) * initBox.size(),
                            hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemsetAsync(bodyMom.p, 0, sizeof(double) * 4 * nb, s));
    k_body_reduce<<<blocks(n), kB, 0, s>>>(P, label.as<int>(), n, bodyBox.as<unsigned long long>(),
                                           bodyMom.as<double>());
    std::vector<unsigned long long> box(6 * (size_t)nb);
    std::vector<double> mom(4 * (size_t)nb);
    HIPCHECK(hipMemcpyAsync(box.data(), bodyBox.p, sizeof(unsigned long long) * box.size(), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(mom.data(), bodyMom.p, sizeof(double) * mom.size(), hipMemcpyDeviceToHost, s));
    sync();
    hg.assign(nb, G64{});
    hNodeOff.assign(nb + 1, 0);
    int64_t off = 0;
    for (int b = 0; b < nb; ++b) {
      G64& g = hg[b];
      g.h = h;
      for (int c = 0; c < 3; ++c) {
        g.fv[c] = frame ? mom[4 * b + c] / mom[4 * b + 3] : 0.0;
        const double lo = ord_dec(box[6 * b + c]), hi = ord_dec(box[6 * b + 3 + c]);
        const int cellMin = (int)std::floor(lo / h) - kSupport - 1;
        const int cellMax = (int)std::ceil(hi / h) + kSupport + 1;
        g.gmin[c] = (double)cellMin * h;
        g.n[c] = cellMax - cellMin;
      }
      const long long cells = (long long)g.n[0] * g.n[1] * g.n[2];
      if (cells > 4000000000LL || cells <= 0) throw std::runtime_error("grid is too big");
      g.off = off;
      hNodeOff[b] = off;
      off += cells;
    }
    hNodeOff[nb] = off;
    NN = off;
    grids.ensure(sizeof(G64) * nb);
    HIPCHECK(hipMemcpyAsync(grids.p, hg.data(), sizeof(G64) * nb, hipMemcpyHostToDevice, s));
    nodeOffDev.ensure(sizeof(int64_t) * (nb + 1));
    HIPCHECK(hipMemcpyAsync(nodeOffDev.p, hNodeOff.data(), sizeof(int64_t) * (nb + 1), hipMemcpyHostToDevice, s));
    // sort by (body, cell) and permute the store
    k_grid_keys<<<blocks(n), kB, 0, s>>>(P, label.as<int>(), n, grids.as<G64>(), keys.as<unsigned long long>(),
                                         vals.as<int>());
    sort_pairs<unsigned long long>(keys, keys2, vals, vals2, n, 64);
    store2.ensure(sizeof(double) * NCOL * n);
    label2.ensure(sizeof(int) * n);
    orig2.ensure(sizeof(int) * n);
    k_permute<<<blocks(n), kB, 0, s>>>(P, P64{store2.as<double>(), n}, vals2.as<int>(), label.as<int>(),
                                       label2.as<int>(), orig.as<int>(), orig2.as<int>(), slotOf.as<int>());
    store.swap(store2);
    label.swap(label2);
    orig.swap(orig2);
    nbp = n - nballistic;
    pstart.ensure(sizeof(int64_t) * (nb + 1));
    k_body_starts<<<blocks(nbp), kB, 0, s>>>(keys2.as<unsigned long long>(), nbp, pstart.as<int64_t>());
    // node arrays
    const size_t v3 = sizeof(double) * 3 * NN;
    for (DevBuf* b : {&vel, &prev, &frc, &em, &vc, &nrm, &diag, &x, &r, &rp, &p, &Ap, &Ar, &tmp, &df}) b->ensure(v3);
    mass.ensure(sizeof(double) * NN);
    flags.ensure(NN);
    nodeBody.ensure(sizeof(int) * NN);
    k_node_body<<<blocks(NN), kB, 0, s>>>(grids.as<G64>(), nodeOffDev.as<int64_t>(), nb, NN, nodeBody.as<int>());
    HIPCHECK(hipMemsetAsync(mass.p, 0, sizeof(double) * NN, s));
    const P64 Q = view();
    k_p2g_mass<<<blocks(nbp), kB, 0, s>>>(Q, label.as<int>(), nbp, grids.as<G64>(), mass.as<double>());
    activeNodes.ensure(sizeof(double) * nb);
    HIPCHECK(hipMemsetAsync(activeNodes.p, 0, sizeof(double) * nb, s));
    k_mass_fix<<<blocks(NN), kB, 0, s>>>(NN, nodeBody.as<int>(), mass.as<double>(), activeNodes.as<double>(),
                                         errDev.as<int>());
  }

  // A v for every body still iterating (or all bodies when cr is null); see k_a_post for mode
  void apply_A(const double* v, double* out, double dt, bool iterating, int mode = 0, double* out2 = nullptr) {
    const P64 P = view();
    CR64* c = iterating ? cr.as<CR64>() : nullptr;
    k_a_pre<<<blocks(NN), kB, 0, s>>>(NN, nodeBody.as<int>(), c, flags.as<int8_t>(), nrm.as<double>(), v,
                                      tmp.as<double>(), df.as<double>());
    k_force_diff<<<blocks(nbp), kB, 0, s>>>(P, label.as<int>(), nbp, grids.as<G64>(), c, tmp.as<double>(),
                                            df.as<double>());
    k_a_post<<<blocks(NN), kB, 0, s>>>(NN, nodeBody.as<int>(), cr.as<CR64>(), flags.as<int8_t>(), nrm.as<double>(),
                                       mass.as<double>(), tmp.as<double>(), df.as<double>(), dt, out, out2,
                                       rp.as<double>(), mode);
  }

  void solve(double dt, int maxIters, double tol) {
    const int64_t nbk = blocks(nb);
    cr.ensure(sizeof(CR64) * nb);
    HIPCHECK(hipMemsetAsync(cr.p, 0, sizeof(CR64) * nb, s));
    hist.ensure(sizeof(double) * (size_t)kHistCap * nb);
    counters.ensure(4 * sizeof(int));
    HIPCHECK(hipMemsetAsync(counters.p, 0, 4 * sizeof(int), s));
    int* ndone = counters.as<int>();
    CR64* c = cr.as<CR64>();
    const int* nbd = nodeBody.as<int>();
    double* b = em.as<double>();
    k_b_norm<<<blocks(NN), kB, 0, s>>>(NN, nbd, c, b);
    k_cr_scalar<<<(unsigned)nbk, kB, 0, s>>>(c, nb, 0, 0, maxIters, tol, hist.as<double>(), ndone);
    // r = b - A x, rp = M^-1 r, p = rp, Ap = A p, Ar = A rp (= Ap), rz = <rp, Ar>
    apply_A(x.as<double>(), Ap.as<double>(), dt, true);
    k_cr_start<<<blocks(NN), kB, 0, s>>>(NN, nbd, c, b, Ap.as<double>(), diag.as<double>(), x.as<double>(),
                                         r.as<double>(), rp.as<double>(), p.as<double>());
    apply_A(p.as<double>(), Ap.as<double>(), dt, true, 1, Ar.as<double>());
    k_cr_scalar<<<(unsigned)nbk, kB, 0, s>>>(c, nb, 1, 0, maxIters, tol, hist.as<double>(), ndone);
    int done = 0;
    for (int it = 0;; ++it) {
      k_cr_dir<<<blocks(NN), kB, 0, s>>>(NN, nbd, c, it, rp.as<double>(), Ar.as<double>(), diag.as<double>(),
                                         p.as<double>(), Ap.as<double>());
      k_cr_scalar<<<(unsigned)nbk, kB, 0, s>>>(c, nb, 2, it, maxIters, tol, hist.as<double>(), ndone);
      k_cr_x<<<blocks(NN), kB, 0, s>>>(NN, nbd, c, flags.as<int8_t>(), nrm.as<double>(), p.as<double>(),
                                       Ap.as<double>(), diag.as<double>(), x.as<double>(), rp.as<double>(),
                                       r.as<double>());
      k_cr_scalar<<<(unsigned)nbk, kB, 0, s>>>(c, nb, 3, it, maxIters, tol, hist.as<double>(), ndone);
      apply_A(rp.as<double>(), Ar.as<double>(), dt, true, 2);
      k_cr_scalar<<<(unsigned)nbk, kB, 0, s>>>(c, nb, 4, it, maxIters, tol, hist.as<double>(), ndone);
      if ((it & 7) == 7 || it > (1 << 20)) {
        HIPCHECK(hipMemcpyAsync(&done, ndone, sizeof(int), hipMemcpyDeviceToHost, s));
        sync();
        if (done >= nb) break;
        if (it > (1 << 20)) throw std::runtime_error("conjugate residual did not terminate");
      }
    }
  }

  void advance(double dt, int maxIters, double tol, mpm_step_stats* stats) {
    lastSolves.clear();
    const P64 P0 = view();
    if (nballistic > 0) k_ballistic<<<blocks(n), kB, 0, s>>>(P0, label.as<int>(), n, grav, dt);
    int totalIt = 0, maxIt = 0;
    if (nb > 0) {
      if (nb > 1 || nballistic > 0)
        k_plasticity<<<blocks(n), kB, 0, s>>>(P0, label.as<int>(), n, 0, thc, ths, errDev.as<int>());
      build_grids(true);
      const P64 P = view();
      const int* lab = label.as<int>();
      const G64* G = grids.as<G64>();
      HIPCHECK(hipMemsetAsync(vel.p, 0, sizeof(double) * 3 * NN, s));
      k_p2g_vel<<<blocks(nbp), kB, 0, s>>>(P, lab, nbp, G, mass.as<double>(), vel.as<double>());
      k_vel_check<<<blocks(3 * NN), kB, 0, s>>>(NN, vel.as<double>(), prev.as<double>(), errDev.as<int>());
      check_err();
      k_force_init<<<blocks(NN), kB, 0, s>>>(NN, mass.as<double>(), grav, frc.as<double>());
      k_forces<<<blocks(nbp), kB, 0, s>>>(P, lab, nbp, G, frc.as<double>(), errDev.as<int>());
      k_explicit<<<blocks(NN), kB, 0, s>>>(NN, nodeBody.as<int>(), G, mass.as<double>(), vel.as<double>(),
                                           frc.as<double>(), colsDev.as<ColliderSetT<double>>(), dt, em.as<double>(),
                                           flags.as<int8_t>(), vc.as<double>(), nrm.as<double>(), errDev.as<int>());
      check_err();
      cr.ensure(sizeof(CR64) * nb);
      HIPCHECK(hipMemsetAsync(cr.p, 0, sizeof(CR64) * nb, s));
      k_guess_sums<<<blocks(NN), kB, 0, s>>>(NN, nodeBody.as<int>(), mass.as<double>(), em.as<double>(),
                                             flags.as<int8_t>(), vc.as<double>(), cr.as<CR64>());
      k_guess_fill<<<blocks(NN), kB, 0, s>>>(NN, nodeBody.as<int>(), cr.as<CR64>(), x.as<double>());
      // rhs: b = momenta + dt^2 P(dF(vc))
      HIPCHECK(hipMemsetAsync(df.p, 0, sizeof(double) * 3 * NN, s));
      k_force_diff<<<blocks(nbp), kB, 0, s>>>(P, lab, nbp, G, nullptr, vc.as<double>(), df.as<double>());
      k_rhs<<<blocks(NN), kB, 0, s>>>(NN, flags.as<int8_t>(), nrm.as<double>(), df.as<double>(), dt, em.as<double>(),
                                      errDev.as<int>());
      HIPCHECK(hipMemsetAsync(diag.p, 0, sizeof(double) * 3 * NN, s));
      k_dfidxi<<<blocks(nbp), kB, 0, s>>>(P, lab, nbp, G, diag.as<double>());
      k_diag_fin<<<blocks(NN), kB, 0, s>>>(NN, mass.as<double>(), dt, diag.as<double>());
      check_err();
      solve(dt, maxIters, tol);
      k_finalize<<<blocks(3 * NN), kB, 0, s>>>(3 * NN, x.as<double>(), vc.as<double>(), vel.as<double>());
      k_g2p<<<blocks(nbp), kB, 0, s>>>(P, lab, nbp, G, vel.as<double>(), prev.as<double>(), dt);
      k_plasticity<<<blocks(n), kB, 0, s>>>(P, lab, n, 1, thc, ths, errDev.as<int>());
      // solve statistics
      std::vector<CR64> hc(nb);
      std::vector<int64_t> ps(nb + 1);
      HIPCHECK(hipMemcpyAsync(hc.data(), cr.p, sizeof(CR64) * nb, hipMemcpyDeviceToHost, s));
      HIPCHECK(hipMemcpyAsync(ps.data(), pstart.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, s));
      std::vector<double> hh((size_t)kHistCap * nb), act(nb);
      HIPCHECK(hipMemcpyAsync(hh.data(), hist.p, sizeof(double) * hh.size(), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipMemcpyAsync(act.data(), activeNodes.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
      sync();
      ps[nb] = nbp;
      lastSolves.resize(nb);
      for (int bi = 0; bi < nb; ++bi) {
        BodySolve& o = lastSolves[bi];
        o.iterations = hc[bi].iterations;
        o.early = hc[bi].early;
        o.bnorm2 = hc[bi].bnorm2;
        o.rn = hc[bi].rn;
        o.particles = (int)(ps[bi + 1] - ps[bi]);
        o.active = (int)act[bi];
        const int L = std::min(hc[bi].histLen, kHistCap);
        o.history.assign(hh.begin() + (size_t)bi * kHistCap, hh.begin() + (size_t)bi * kHistCap + L);
        totalIt += o.iterations;
        maxIt = std::max(maxIt, o.iterations);
      }
    }
    const P64 P = view();
    k_advect<<<blocks(n), kB, 0, s>>>(P, n, colsDev.as<ColliderSetT<double>>(), dt);
    check_err();
    calculate_bodies();
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->num_bodies = nb;
      stats->num_ballistic = (int)nballistic;
      stats->total_iterations = totalIt;
      stats->max_iterations = maxIt;
      stats->num_solved = (int)lastSolves.size();
    }
  }
};

namespace {
template <class F> int guarded64(mpm_sim64* sim, F fn) {
  if (!sim) return MPM_ERR_INVALID;
  try {
    HIPCHECK(hipSetDevice(sim->device));
    fn();
    sim->error.clear();
    return MPM_OK;
  } catch (const InvalidArg& e) {
    sim->error = e.what();
    return MPM_ERR_INVALID;
  } catch (const HipError& e) {
    sim->error = e.what();
    return MPM_ERR_HIP;
  } catch (const std::exception& e) {
    sim->error = e.what();
    return MPM_ERR_RUNTIME;
  }
}
}  // namespace

extern "C" {

int mpm_sim64_create(mpm_sim64** out, int device, int64_t n) {
  if (!out) return MPM_ERR_INVALID;
  *out = nullptr;
  if (n <= 0 || n >= INT_MAX) return MPM_ERR_INVALID;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return MPM_ERR_NO_DEVICE;
  if (device < 0 || device >= count) return MPM_ERR_INVALID;
  std::unique_ptr<mpm_sim64> sim(new mpm_sim64());
  sim->device = device;
  sim->n = n;
  const int rc = guarded64(sim.get(), [&] {
    mpm_sim64& S = *sim;
    HIPCHECK(hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking));
    S.store.ensure(sizeof(double) * NCOL * n);
    HIPCHECK(hipMemsetAsync(S.store.p, 0, sizeof(double) * NCOL * n, S.s));
    k_store_init<<<blocks(n), kB, 0, S.s>>>(S.view());
    for (DevBuf* b : {&S.label, &S.orig, &S.slotOf}) b->ensure(sizeof(int) * n);
    k_fill<int><<<blocks(n), kB, 0, S.s>>>(S.label.as<int>(), -1, n);
    k_iota<<<blocks(n), kB, 0, S.s>>>(S.orig.as<int>(), n);
    k_iota<<<blocks(n), kB, 0, S.s>>>(S.slotOf.as<int>(), n);
    S.errDev.ensure(sizeof(int));
    HIPCHECK(hipMemsetAsync(S.errDev.p, 0, sizeof(int), S.s));
    S.colsDev.ensure(sizeof(ColliderSetT<double>));
    HIPCHECK(hipMemcpyAsync(S.colsDev.p, &S.cols, sizeof(S.cols), hipMemcpyHostToDevice, S.s));
    S.sync();
  });
  if (rc != MPM_OK) return rc;
  *out = sim.release();
  return MPM_OK;
}

int mpm_sim64_destroy(mpm_sim64* sim) {
  if (!sim) return MPM_ERR_INVALID;
  (void)hipSetDevice(sim->device);
  if (sim->s) (void)hipStreamSynchronize(sim->s);
  hipStream_t st = sim->s;
  delete sim;
  if (st) (void)hipStreamDestroy(st);
  return MPM_OK;
}

const char* mpm_sim64_last_error(const mpm_sim64* sim) { return sim ? sim->error.c_str() : "null simulation"; }

int mpm_sim64_set(mpm_sim64* sim, const char* name, const double* data, int64_t count) {
  return guarded64(sim, [&] {
    const Var& v = find_var(name);
    if (!data || count != sim->n * v.ncomp) throw InvalidArg("size mismatch for '" + std::string(name) + "'");
    DevBuf tmp;
    tmp.ensure(sizeof(double) * count);
    HIPCHECK(hipMemcpyAsync(tmp.p, data, sizeof(double) * count, hipMemcpyHostToDevice, sim->s));
    k_set_col<<<blocks(sim->n), kB, 0, sim->s>>>(sim->view(), v.col, v.ncomp, tmp.as<double>(), sim->slotOf.as<int>());
    sim->sync();
  });
}

int mpm_sim64_get(mpm_sim64* sim, const char* name, double* data, int64_t count) {
  return guarded64(sim, [&] {
    const Var& v = find_var(name);
    if (!data || count != sim->n * v.ncomp) throw InvalidArg("size mismatch for '" + std::string(name) + "'");
    DevBuf tmp;
    tmp.ensure(sizeof(double) * count);
    k_get_col<<<blocks(sim->n), kB, 0, sim->s>>>(sim->view(), v.col, v.ncomp, tmp.as<double>(), sim->slotOf.as<int>());
    HIPCHECK(hipMemcpyAsync(data, tmp.p, sizeof(double) * count, hipMemcpyDeviceToHost, sim->s));
    sim->sync();
  });
}

int mpm_sim64_model_snow(mpm_sim64* sim, double E, double nu, double xi, double theta_c, double theta_s) {
  return guarded64(sim, [&] {
    if (sim->haveModel) throw InvalidArg("Material point data already has a variable named 'Fp'");
    const double mu0 = E / (2.0 * (1.0 + nu));
    const double lambda0 = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const P64 P = sim->view();
    const int64_t n = sim->n;
    k_fill<double><<<blocks(n), kB, 0, sim->s>>>(P.base + C_MU * n, mu0, n);
    k_fill<double><<<blocks(n), kB, 0, sim->s>>>(P.base + C_LAM * n, lambda0, n);
    k_fill<double><<<blocks(n), kB, 0, sim->s>>>(P.base + C_MU0 * n, mu0, n);
    k_fill<double><<<blocks(n), kB, 0, sim->s>>>(P.base + C_LAM0 * n, lambda0, n);
    k_fill<double><<<blocks(n), kB, 0, sim->s>>>(P.base + C_XI * n, xi, n);
    sim->thc = theta_c;
    sim->ths = theta_s;
    sim->haveModel = true;
    sim->sync();
  });
}

int mpm_sim64_set_colliders(mpm_sim64* sim, const mpm_collider64* c, int n) {
  return guarded64(sim, [&] {
    if (n < 0 || n > mpm::kMaxColliders || (n > 0 && !c)) throw InvalidArg("0 to 16 colliders");
    for (int i = 0; i < n; ++i) {
      if (c[i].type < 0 || c[i].type > 2) throw InvalidArg("unknown collider type");
      if (c[i].type == 2 && (!c[i].phi || c[i].dims[0] < 2 || c[i].dims[1] < 2 || c[i].dims[2] < 2 || !(c[i].voxel > 0.0)))
        throw InvalidArg("volume collider needs phi, dims >= 2 and voxel > 0");
    }
    ColliderSetT<double> cs{};
    std::vector<std::unique_ptr<DevBuf>> data;
    cs.n = n;
    for (int i = 0; i < n; ++i) {
      ColliderT<double>& o = cs.o[i];
      o.type = c[i].type;
      for (int k = 0; k < 4; ++k) o.c[k] = c[i].coeffs[k];
      for (int k = 0; k < 3; ++k) {
        o.center[k] = c[i].center[k];
        o.v[k] = c[i].velocity[k];
        o.origin[k] = c[i].origin[k];
        o.dims[k] = c[i].dims[k];
      }
      o.radius = c[i].radius;
      o.friction = c[i].friction;
      o.sticky = c[i].sticky;
      o.voxel = c[i].voxel;
      o.phi = o.vel = nullptr;
      if (o.type == 2) {
        const size_t pts = (size_t)o.dims[0] * o.dims[1] * o.dims[2];
        data.emplace_back(new DevBuf());
        data.back()->ensure(sizeof(double) * pts);
        HIPCHECK(hipMemcpyAsync(data.back()->p, c[i].phi, sizeof(double) * pts, hipMemcpyHostToDevice, sim->s));
        o.phi = data.back()->as<double>();
        if (c[i].phi_velocity) {
          data.emplace_back(new DevBuf());
          data.back()->ensure(sizeof(double) * 3 * pts);
          HIPCHECK(hipMemcpyAsync(data.back()->p, c[i].phi_velocity, sizeof(double) * 3 * pts, hipMemcpyHostToDevice,
                                  sim->s));
          o.vel = data.back()->as<double>();
        }
      }
    }
    HIPCHECK(hipMemcpyAsync(sim->colsDev.p, &cs, sizeof(cs), hipMemcpyHostToDevice, sim->s));
    sim->sync();
    sim->cols = cs;
    sim->colData.swap(data);
  });
}

int mpm_sim64_set_gravity(mpm_sim64* sim, const double* g, int n) {
  return guarded64(sim, [&] {
    if (n < 0 || n > kMaxGravity || (n > 0 && !g)) throw InvalidArg("0 to 8 gravity fields");
    sim->grav.n = n;
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) sim->grav.g[i][c] = g[3 * i + c];
  });
}

int mpm_sim64_init(mpm_sim64* sim, double grid_size) {
  return guarded64(sim, [&] {
    if (!sim->haveModel) throw InvalidArg("no constitutive model");
    if (!(grid_size > 0.0)) throw InvalidArg("grid size must be positive");
    sim->h = grid_size;
    sim->calculate_bodies();
    if (sim->nb > 0) {
      // Sim.cpp:41-49: one grid per body (no frame velocity) and the particle volumes
      sim->build_grids(false);
      sim->check_err();
      k_volumes<<<blocks(sim->nbp), kB, 0, sim->s>>>(sim->view(), sim->label.as<int>(), sim->nbp,
                                                     sim->grids.as<G64>(), sim->mass.as<double>());
      sim->sync();
    }
    sim->initialised = true;
  });
}

int mpm_sim64_advance(mpm_sim64* sim, double dt, int max_iters, double tolerance, mpm_step_stats* stats) {
  return guarded64(sim, [&] {
    if (!sim->initialised) throw InvalidArg("simulation not initialised (call mpm_sim64_init)");
    if (max_iters < 0) throw InvalidArg("max_iters must be non-negative");
    sim->advance(dt, max_iters, tolerance, stats);
  });
}

int mpm_sim64_num_bodies(const mpm_sim64* sim) { return sim ? sim->nb : -1; }

int mpm_sim64_body_solve_stats(const mpm_sim64* sim, int body, mpm_solve_stats* st) {
  if (!sim || !st || body < 0 || body >= (int)sim->lastSolves.size()) return MPM_ERR_INVALID;
  const BodySolve& o = sim->lastSolves[body];
  st->iterations = o.iterations;
  st->early_exit = o.early;
  st->rhs_norm2 = o.bnorm2;
  st->residual_norm2 = o.rn;
  st->particles = o.particles;
  st->active_nodes = o.active;
  st->fallbacks = 0;
  st->restarts = 0;
  const int L = (int)o.history.size();
  st->history_length = st->history ? std::min(L, st->history_capacity) : 0;
  if (st->history)
    for (int i = 0; i < st->history_length; ++i) st->history[i] = o.history[i];
  return MPM_OK;
}

}  // extern "C"
