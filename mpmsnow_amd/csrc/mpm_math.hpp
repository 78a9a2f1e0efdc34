// mpm_math.hpp -- per-particle / per-node math of the implicit MPM snow step,
// written for gfx950 device code (also compiles for the host).
//
// Semantics follow davidsminor/mpmsnow (paths relative to the reference
// checkout): B-spline src/MpmSim/CubicBsplineShapeFunction.cpp:11-48, shape
// iterator src/MpmSim/Grid.cpp:1105-1193, snow model
// src/MpmSim/SnowConstitutiveModel.cpp:27-287, collisions
// src/CollisionObject.cpp:19-96 and src/MpmSim/CollisionPlane.cpp:23-46.
// Matrices are 3x3 column-major float[9] (element (r,c) at r + 3c), the
// memory order of the reference's Eigen::Matrix3f.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#define MPM_HD __host__ __device__ __forceinline__

namespace mpm {

constexpr int kMaxColliders = 16;

// Scalar helpers for the two precisions the step is built in: float (the library's FP32 bulk)
// and double (the full-FP64 step of the precision study). Every templated function below
// evaluates, for T = float, exactly the float expression it did as a float-only function.
MPM_HD float msqrt(float x) { return sqrtf(x); }
MPM_HD double msqrt(double x) { return sqrt(x); }
MPM_HD float mabs(float x) { return fabsf(x); }
MPM_HD double mabs(double x) { return fabs(x); }
MPM_HD float mfloor(float x) { return floorf(x); }
MPM_HD double mfloor(double x) { return floor(x); }
MPM_HD float mexp(float x) { return expf(x); }
MPM_HD double mexp(double x) { return exp(x); }

// ------------------------------------------------------------ B-spline
template <class T> MPM_HD T bspline_w(T x) {
  T ax = mabs(x);
  if (ax < T(1)) return (T(0.5) * ax - T(1)) * ax * ax + T(2) / T(3);
  if (ax < T(2)) return ((-ax / T(6) + T(1)) * ax - T(2)) * ax + T(4) / T(3);
  return T(0);
}
template <class T> MPM_HD T bspline_dw(T x) {
  T s = T(1);
  if (x < T(0)) { x = -x; s = T(-1); }
  if (x < T(1)) return s * (x * (T(1.5) * x - T(2)));
  if (x < T(2)) { x -= T(2); return s * (T(-0.5) * x * x); }
  return T(0);
}

// Per-particle stencil: base node and the 4 weights (and derivatives) per axis.
template <class T> struct StencilT {
  int base[3];
  T w[3][4];
  T dw[3][4];  // already divided by h
};
using Stencil = StencilT<float>;

// Grid.cpp:1105-1136 for dim = 3: base = floor((p - min)/h) - 1; weights at j - frac, j = -1..2.
template <class T> MPM_HD void make_stencil(const T p[3], const T gmin[3], T h, bool derivs, StencilT<T>& s) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    T frac = (p[d] - gmin[d]) / h;
    int dimPos = (int)mfloor(frac);
    frac -= (T)dimPos;
    s.base[d] = dimPos - 1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      T t = (T)(i - 1) - frac;
      s.w[d][i] = bspline_w(t);
      if (derivs) s.dw[d][i] = bspline_dw(t) / h;
    }
  }
}
// The same stencil with the piecewise cubics written out for the four nodes of each axis
// (frac f in [0, 1), g = 1 - f): w = (g^3/6, (f/2 - 1) f^2 + 2/3, (g/2 - 1) g^2 + 2/3, f^3/6),
// dw = (g^2/2, f (2 - 3f/2), g (3g/2 - 2), -f^2/2) / h, and no divisions besides the base's
// (kept identical so the base matches the binning). Values agree with make_stencil to a few
// float ulps; the solver's Hessian-vector product uses it.
MPM_HD void make_stencil_fast(const float p[3], const float gmin[3], float h, float invh, Stencil& s) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float frac = (p[d] - gmin[d]) / h;
    int dimPos = (int)floorf(frac);
    const float f = frac - (float)dimPos;
    const float g = 1.0f - f;
    s.base[d] = dimPos - 1;
    const float f2 = f * f, g2 = g * g;
    s.w[d][0] = g2 * g * (1.0f / 6.0f);
    s.w[d][1] = (0.5f * f - 1.0f) * f2 + 2.0f / 3.0f;
    s.w[d][2] = (0.5f * g - 1.0f) * g2 + 2.0f / 3.0f;
    s.w[d][3] = f2 * f * (1.0f / 6.0f);
    s.dw[d][0] = 0.5f * g2 * invh;
    s.dw[d][1] = f * (2.0f - 1.5f * f) * invh;
    s.dw[d][2] = g * (1.5f * g - 2.0f) * invh;
    s.dw[d][3] = -0.5f * f2 * invh;
  }
}
// weight of stencil node (a,b,c): w_x * w_y * w_z in the reference's product order
template <class T> MPM_HD T st_w(const StencilT<T>& s, int a, int b, int c) { return s.w[0][a] * s.w[1][b] * s.w[2][c]; }
template <class T> MPM_HD void st_grad(const StencilT<T>& s, int a, int b, int c, T g[3]) {
  g[0] = -s.dw[0][a] * s.w[1][b] * s.w[2][c];
  g[1] = -s.w[0][a] * s.dw[1][b] * s.w[2][c];
  g[2] = -s.w[0][a] * s.w[1][b] * s.dw[2][c];
}

// ------------------------------------------------------------ 3x3 algebra
template <class T> MPM_HD T& at(T* m, int r, int c) { return m[r + 3 * c]; }
template <class T> MPM_HD T at(const T* m, int r, int c) { return m[r + 3 * c]; }

template <class T> MPM_HD void mat_mul(const T* A, const T* B, T* C) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) C[r + 3 * c] = A[r] * B[3 * c] + A[r + 3] * B[3 * c + 1] + A[r + 6] * B[3 * c + 2];
}
// C = A * B^T
template <class T> MPM_HD void mat_mul_bt(const T* A, const T* B, T* C) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) C[r + 3 * c] = A[r] * B[c] + A[r + 3] * B[c + 3] + A[r + 6] * B[c + 6];
}
// C = A^T * B
template <class T> MPM_HD void mat_mul_at(const T* A, const T* B, T* C) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      C[r + 3 * c] = A[3 * r] * B[3 * c] + A[3 * r + 1] * B[3 * c + 1] + A[3 * r + 2] * B[3 * c + 2];
}
template <class T> MPM_HD void mat_vec(const T* A, const T* v, T* o) {
#pragma unroll
  for (int r = 0; r < 3; ++r) o[r] = A[r] * v[0] + A[r + 3] * v[1] + A[r + 6] * v[2];
}
// o = A^T v
template <class T> MPM_HD void mat_tvec(const T* A, const T* v, T* o) {
#pragma unroll
  for (int r = 0; r < 3; ++r) o[r] = A[3 * r] * v[0] + A[3 * r + 1] * v[1] + A[3 * r + 2] * v[2];
}
template <class T> MPM_HD T det3(const T* m) {
  return at(m, 0, 0) * (at(m, 1, 1) * at(m, 2, 2) - at(m, 2, 1) * at(m, 1, 2)) -
         at(m, 1, 0) * (at(m, 0, 1) * at(m, 2, 2) - at(m, 2, 1) * at(m, 0, 2)) +
         at(m, 2, 0) * (at(m, 0, 1) * at(m, 1, 2) - at(m, 1, 1) * at(m, 0, 2));
}
template <class T> MPM_HD void inv3(const T* m, T* inv) {
  T c00 = at(m, 1, 1) * at(m, 2, 2) - at(m, 1, 2) * at(m, 2, 1);
  T c10 = at(m, 1, 2) * at(m, 2, 0) - at(m, 1, 0) * at(m, 2, 2);
  T c20 = at(m, 1, 0) * at(m, 2, 1) - at(m, 1, 1) * at(m, 2, 0);
  T det = at(m, 0, 0) * c00 + at(m, 0, 1) * c10 + at(m, 0, 2) * c20;
  T id = T(1) / det;
  at(inv, 0, 0) = c00 * id;
  at(inv, 1, 0) = c10 * id;
  at(inv, 2, 0) = c20 * id;
  at(inv, 0, 1) = (at(m, 0, 2) * at(m, 2, 1) - at(m, 0, 1) * at(m, 2, 2)) * id;
  at(inv, 1, 1) = (at(m, 0, 0) * at(m, 2, 2) - at(m, 0, 2) * at(m, 2, 0)) * id;
  at(inv, 2, 1) = (at(m, 0, 1) * at(m, 2, 0) - at(m, 0, 0) * at(m, 2, 1)) * id;
  at(inv, 0, 2) = (at(m, 0, 1) * at(m, 1, 2) - at(m, 0, 2) * at(m, 1, 1)) * id;
  at(inv, 1, 2) = (at(m, 0, 2) * at(m, 1, 0) - at(m, 0, 0) * at(m, 1, 2)) * id;
  at(inv, 2, 2) = (at(m, 0, 0) * at(m, 1, 1) - at(m, 0, 1) * at(m, 1, 0)) * id;
}
// reference SnowConstitutiveModel.cpp:258-264
template <class T> MPM_HD T ddot(const T* a, const T* b) {
  return at(a, 0, 0) * at(b, 0, 0) + at(a, 0, 1) * at(b, 0, 1) + at(a, 0, 2) * at(b, 0, 2) + at(a, 1, 0) * at(b, 1, 0) +
         at(a, 1, 1) * at(b, 1, 1) + at(a, 1, 2) * at(b, 1, 2) + at(a, 2, 0) * at(b, 2, 0) + at(a, 2, 1) * at(b, 2, 1) +
         at(a, 2, 2) * at(b, 2, 2);
}

// ------------------------------------------------------------ 3x3 SVD
// One-sided (Hestenes) Jacobi: right rotations orthogonalise the columns of
// B = A V; sigma_i = |B_i|, U_i = B_i / sigma_i. sigma >= 0, sorted
// descending; U, V orthogonal (the conventions of Eigen's JacobiSVD). All the
// model derives from it (R, F^-T, S, J, the clamp) is independent of the
// sign/order freedom of the decomposition for nonsingular F.
// A pair of columns is rotated only while their inner product exceeds twice the unit roundoff of
// T relative to their norms (|g| > 2 eps sqrt(a b), the threshold of Eigen's JacobiSVD): below it
// they are orthogonal to working precision. Without it an F that is the identity up to rounding
// (equal column norms, off-diagonal entries of 1e-16) takes 45-degree rotations whose own
// rounding leaves R off by ~1e-7, i.e. a spurious stress on undeformed snow.
template <class T> MPM_HD T jacobi_skip_tol() { return sizeof(T) == 4 ? T(1.1920929e-7) : T(2.220446049250313e-16); }
template <class T, int SWEEPS> MPM_HD void svd3_jacobi(const T* A, T* U, T* s, T* V) {
  T B[9], W[9] = {T(1), T(0), T(0), T(0), T(1), T(0), T(0), T(0), T(1)};
#pragma unroll
  for (int i = 0; i < 9; ++i) B[i] = A[i];
#pragma unroll
  for (int sweep = 0; sweep < SWEEPS; ++sweep) {
#pragma unroll
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0;
      const int q = pq == 0 ? 1 : 2;
      T a = B[3 * p] * B[3 * p] + B[3 * p + 1] * B[3 * p + 1] + B[3 * p + 2] * B[3 * p + 2];
      T b = B[3 * q] * B[3 * q] + B[3 * q + 1] * B[3 * q + 1] + B[3 * q + 2] * B[3 * q + 2];
      T g = B[3 * p] * B[3 * q] + B[3 * p + 1] * B[3 * q + 1] + B[3 * p + 2] * B[3 * q + 2];
      if (mabs(g) > jacobi_skip_tol<T>() * msqrt(a * b)) {
        T zeta = (b - a) / (T(2) * g);
        T t = (zeta >= T(0) ? T(1) : T(-1)) / (mabs(zeta) + msqrt(T(1) + zeta * zeta));
        T cs = T(1) / msqrt(T(1) + t * t);
        T sn = cs * t;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          T bp = B[r + 3 * p], bq = B[r + 3 * q];
          B[r + 3 * p] = cs * bp - sn * bq;
          B[r + 3 * q] = sn * bp + cs * bq;
          T wp = W[r + 3 * p], wq = W[r + 3 * q];
          W[r + 3 * p] = cs * wp - sn * wq;
          W[r + 3 * q] = sn * wp + cs * wq;
        }
      }
    }
  }
  T n[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) n[c] = msqrt(B[3 * c] * B[3 * c] + B[3 * c + 1] * B[3 * c + 1] + B[3 * c + 2] * B[3 * c + 2]);
  // sort descending (3-element network), permuting columns of B and W
  int o0 = 0, o1 = 1, o2 = 2;
  if (n[o0] < n[o1]) { int t = o0; o0 = o1; o1 = t; }
  if (n[o1] < n[o2]) { int t = o1; o1 = o2; o2 = t; }
  if (n[o0] < n[o1]) { int t = o0; o0 = o1; o1 = t; }
  const int ord[3] = {o0, o1, o2};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s[k] = n[ord[k]];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      V[r + 3 * k] = W[r + 3 * ord[k]];
      U[r + 3 * k] = B[r + 3 * ord[k]];
    }
  }
  // U columns: normalise; complete a degenerate column orthonormally
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (s[k] > T(1e-30) && s[k] > T(1e-12) * s[0]) {
      T inv = T(1) / s[k];
#pragma unroll
      for (int r = 0; r < 3; ++r) U[r + 3 * k] *= inv;
    } else if (k == 2) {
      U[6] = U[1] * U[5] - U[2] * U[4];
      U[7] = U[2] * U[3] - U[0] * U[5];
      U[8] = U[0] * U[4] - U[1] * U[3];
    } else {
      // rank <= 1: pick e_m (smallest |u0| component) orthogonalised against u0 for k = 1,
      // or e_x for k = 0
      T c0 = T(1), c1 = T(0), c2 = T(0);
      if (k == 1) {
        int m = 0;
        if (mabs(U[1]) < mabs(U[m])) m = 1;
        if (mabs(U[2]) < mabs(U[m])) m = 2;
        T e[3] = {T(0), T(0), T(0)};
        e[m] = T(1);
        T d = e[0] * U[0] + e[1] * U[1] + e[2] * U[2];
        c0 = e[0] - d * U[0];
        c1 = e[1] - d * U[1];
        c2 = e[2] - d * U[2];
        T nn = T(1) / msqrt(c0 * c0 + c1 * c1 + c2 * c2);
        c0 *= nn; c1 *= nn; c2 *= nn;
      }
      U[3 * k] = c0; U[3 * k + 1] = c1; U[3 * k + 2] = c2;
    }
  }
}

// ------------------------------------------------------- snow model
struct ParticleCache {  // the model's per-particle state (SnowConstitutiveModel.h:48-57)
  float F[9], Fp[9], Finv[9], R[9], Ginv[9];
  float J, mu, lambda;
};

// updateParticleData for one particle (SnowConstitutiveModel.cpp:27-121).
// T = float: FP32 Jacobi SVD, redone in FP64 when sigma_min < 1e-3; T = double: FP64 SVD.
// Returns false when the hardening factor is not finite ("infinite hardness!!!").
template <class T>
MPM_HD bool snow_update(T F[9], T Fp[9], T Finv[9], T R[9], T Ginv[9], T& J, T& mu, T& lambda, T mu0, T lambda0, T xi,
                        T thc, T ths) {
  T U[9], V[9], sv[3];
  if constexpr (sizeof(T) == 4) {
    svd3_jacobi<float, 5>(F, U, sv, V);
    if (!(sv[2] >= 1e-3f)) {
      double Fd[9], Ud[9], Vd[9], sd[3];
#pragma unroll
      for (int i = 0; i < 9; ++i) Fd[i] = (double)F[i];
      svd3_jacobi<double, 8>(Fd, Ud, sd, Vd);
#pragma unroll
      for (int i = 0; i < 9; ++i) { U[i] = (float)Ud[i]; V[i] = (float)Vd[i]; }
#pragma unroll
      for (int i = 0; i < 3; ++i) sv[i] = (float)sd[i];
    }
  } else {
    svd3_jacobi<double, 8>(F, U, sv, V);
  }
  T d[3], di[3];
  bool modified = false;
  const T hi = T(1) + ths, lo = T(1) - thc;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    T x = sv[i];
    if (x > hi) { modified = true; x = hi; }
    if (x < lo) { modified = true; x = lo; }
    d[i] = x;
    di[i] = T(1) / x;
  }
  // U * diag(d): scale columns
  T UD[9], UDi[9], VD[9];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      UD[r + 3 * c] = U[r + 3 * c] * d[c];
      UDi[r + 3 * c] = U[r + 3 * c] * di[c];
      VD[r + 3 * c] = V[r + 3 * c] * d[c];
    }
  if (modified) {
    T FN[9], VDi[9], T1[9];
    mat_mul(F, Fp, FN);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) VDi[r + 3 * c] = V[r + 3 * c] * di[c];
    mat_mul_bt(VDi, U, T1);  // V D^-1 U^T
    mat_mul(T1, FN, Fp);
    mat_mul_bt(UD, V, F);    // U D V^T
  }
  mat_mul_bt(UDi, V, Finv);  // U D^-1 V^T
  mat_mul_bt(U, V, R);       // U V^T
  T S[9];
  mat_mul_bt(VD, V, S);      // V D V^T
  T G[9];
  at(G, 0, 0) = at(S, 0, 0) + at(S, 1, 1);
  at(G, 1, 1) = at(S, 0, 0) + at(S, 2, 2);
  at(G, 2, 2) = at(S, 1, 1) + at(S, 2, 2);
  at(G, 0, 1) = at(G, 1, 0) = at(S, 1, 2);
  at(G, 0, 2) = at(G, 2, 0) = -at(S, 0, 2);
  at(G, 1, 2) = at(G, 2, 1) = at(S, 0, 1);
  inv3(G, Ginv);
  J = d[0] * d[1] * d[2];
  T hf = xi * (T(1) - det3(Fp));
  if (hf > T(2)) hf = T(2);
  hf = mexp(hf);
  mu = mu0 * hf;
  lambda = lambda0 * hf;
  return isfinite(hf);
}

// dEnergyDensitydF (SnowConstitutiveModel.cpp:187-216): 2 mu (F - R) + lambda (J-1) J F^-T
template <class T> MPM_HD void snow_stress(const T* F, const T* R, const T* Finv, T J, T mu, T lambda, T* P) {
  T a = T(2) * mu;
  T b = lambda * (J - T(1)) * J;
#pragma unroll
  for (int i = 0; i < 9; ++i) P[i] = a * (F[i] - R[i]) + b * Finv[i];
}
// energyDensity (SnowConstitutiveModel.cpp:171-185)
template <class T> MPM_HD T snow_energy(const T* F, const T* R, T J, T mu, T lambda) {
  T rd[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) rd[i] = F[i] - R[i];
  T jm1 = J - T(1);
  return mu * ddot(rd, rd) + T(0.5) * lambda * jm1 * jm1;
}
// dEdFDifferential (SnowConstitutiveModel.cpp:218-256) with computeRdifferential (:266-287)
template <class T>
MPM_HD void snow_stress_differential(const T* dF, const T* R, const T* Finv, const T* Ginv, T J, T mu, T lambda, T* dP) {
  T dJ = J * ddot(Finv, dF);
  // dFinvT = -F^-T dF^T F^-T
  T T1[9], dFi[9];
  mat_mul_bt(Finv, dF, T1);
  mat_mul(T1, Finv, dFi);
  // dR = R * skew(Ginv * w(R^T dF))
  T M[9];
  mat_mul_at(R, dF, M);
  T w0 = at(M, 0, 1) - at(M, 1, 0);
  T w1 = at(M, 0, 2) - at(M, 2, 0);
  T w2 = at(M, 1, 2) - at(M, 2, 1);
  T g0 = at(Ginv, 0, 0) * w0 + at(Ginv, 0, 1) * w1 + at(Ginv, 0, 2) * w2;
  T g1 = at(Ginv, 1, 0) * w0 + at(Ginv, 1, 1) * w1 + at(Ginv, 1, 2) * w2;
  T g2 = at(Ginv, 2, 0) * w0 + at(Ginv, 2, 1) * w1 + at(Ginv, 2, 2) * w2;
  // R * [[0, g0, g1], [-g0, 0, g2], [-g1, -g2, 0]]
  T dR[9];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    T r0 = at(R, r, 0), r1 = at(R, r, 1), r2 = at(R, r, 2);
    at(dR, r, 0) = -r1 * g0 - r2 * g1;
    at(dR, r, 1) = r0 * g0 - r2 * g2;
    at(dR, r, 2) = r0 * g1 + r1 * g2;
  }
  T twoMu = T(2) * mu;
  T dJJ = dJ * J;
  T jm1 = J - T(1);
#pragma unroll
  for (int i = 0; i < 9; ++i)
    dP[i] = twoMu * (dF[i] - dR[i]) + lambda * (dJJ * Finv[i] + jm1 * (dJ * Finv[i] - J * dFi[i]));
}

// ------------------------------------------------------------ colliders
template <class T> struct ColliderT {
  int type;  // 0 plane, 1 sphere, 2 level-set volume
  T c[4];
  T center[3];
  T radius;
  T v[3];
  T friction;
  int sticky;
  // type 2: phi at the points origin + (i, j, k) voxel (point (i, j, k) at i + dims0 (j + dims1 k)),
  // and optionally a velocity per point (3 values; null: the uniform v), in device memory
  int dims[3];
  T origin[3];
  T voxel;
  const T* phi;
  const T* vel;
};
template <class T> struct ColliderSetT {
  int n;
  ColliderT<T> o[kMaxColliders];
};
using Collider = ColliderT<float>;
using ColliderSet = ColliderSetT<float>;

// Collider tests decide the collided flag of a node, which must match the reference bit for bit:
// they keep its separately rounded products and sums (hipcc would otherwise fuse a*b + c
// into one FMA, which rounds once).
// Level-set volume sampling (the reference's VDBCollisionObject, houdiniPlugin/VDBCollisionObject.cpp:
// 11-37, reads phi, its gradient and a velocity field from OpenVDB grids): trilinear
// interpolation of the point values, clamped to the volume like a 3-D texture with clamp
// addressing; each lerp is a + t (b - a) with separately rounded operations.
template <class T> MPM_HD T vol_lerp(T a, T b, T t) {
#pragma clang fp contract(off)
  return a + t * (b - a);
}
// nc components per point, component c of point p at f[p * nc + c]
template <class T> MPM_HD void vol_sample(const ColliderT<T>& o, const T* f, int nc, const T x[3], T* out) {
#pragma clang fp contract(off)
  int i0[3];
  T t[3];
  for (int a = 0; a < 3; ++a) {
    T u = (x[a] - o.origin[a]) / o.voxel;
    const T hi = (T)(o.dims[a] - 1);
    if (!(u >= T(0))) u = T(0);  // below the volume, or NaN
    if (u > hi) u = hi;
    int i = (int)mfloor(u);
    if (i > o.dims[a] - 2) i = o.dims[a] - 2;
    i0[a] = i;
    t[a] = u - (T)i;
  }
  const int64_t sx = nc, sy = (int64_t)nc * o.dims[0], sz = sy * o.dims[1];
  const T* b = f + i0[0] * sx + i0[1] * sy + i0[2] * sz;
  for (int c = 0; c < nc; ++c) {
    const T* q = b + c;
    const T e00 = vol_lerp(q[0], q[sx], t[0]);
    const T e10 = vol_lerp(q[sy], q[sy + sx], t[0]);
    const T e01 = vol_lerp(q[sz], q[sz + sx], t[0]);
    const T e11 = vol_lerp(q[sz + sy], q[sz + sy + sx], t[0]);
    out[c] = vol_lerp(vol_lerp(e00, e10, t[1]), vol_lerp(e01, e11, t[1]), t[2]);
  }
}
template <class T> MPM_HD T col_phi(const ColliderT<T>& o, const T x[3]) {
#pragma clang fp contract(off)
  if (o.type == 0) return o.c[0] * x[0] + o.c[1] * x[1] + o.c[2] * x[2] + o.c[3];
  if (o.type == 2) {
    T p;
    vol_sample(o, o.phi, 1, x, &p);
    return p;
  }
  T d0 = x[0] - o.center[0], d1 = x[1] - o.center[1], d2 = x[2] - o.center[2];
  return msqrt(d0 * d0 + d1 * d1 + d2 * d2) - o.radius;
}
// normalised gradient (CollisionObject::grad + Vector3f::normalize); a volume's gradient is the
// central difference of the sampled phi over one voxel on each axis
template <class T> MPM_HD void col_normal(const ColliderT<T>& o, const T x[3], T n[3]) {
#pragma clang fp contract(off)
  if (o.type == 0) {
    n[0] = o.c[0]; n[1] = o.c[1]; n[2] = o.c[2];
  } else if (o.type == 2) {
    for (int a = 0; a < 3; ++a) {
      T xp[3] = {x[0], x[1], x[2]}, xm[3] = {x[0], x[1], x[2]};
      xp[a] = x[a] + o.voxel;
      xm[a] = x[a] - o.voxel;
      T pp, pm;
      vol_sample(o, o.phi, 1, xp, &pp);
      vol_sample(o, o.phi, 1, xm, &pm);
      n[a] = (pp - pm) / (T(2) * o.voxel);
    }
  } else {
    n[0] = x[0] - o.center[0]; n[1] = x[1] - o.center[1]; n[2] = x[2] - o.center[2];
  }
  T nn = msqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  n[0] /= nn; n[1] /= nn; n[2] /= nn;
}
// CollisionObject::velocity: uniform, or sampled from a volume's velocity field
template <class T> MPM_HD void col_velocity(const ColliderT<T>& o, const T x[3], T v[3]) {
  if (o.type == 2 && o.vel) {
    vol_sample(o, o.vel, 3, x, v);
  } else {
    v[0] = o.v[0]; v[1] = o.v[1]; v[2] = o.v[2];
  }
}

// CollisionObjectSet::collide (src/CollisionObject.cpp:19-96)
template <class T> MPM_HD int collide(const ColliderSetT<T>& cs, T v[3], const T x[3], const T fv[3], bool addVel) {
  int obj = -1;
  bool already = false;
  for (int i = 0; i < cs.n; ++i) {
    const ColliderT<T>& o = cs.o[i];
    if (col_phi(o, x) <= T(0)) {
      if (already) {
        obj = -2;
        v[0] = v[1] = v[2] = T(0);
        break;
      }
      already = true;
      T vo[3];
      col_velocity(o, x, vo);
      vo[0] -= fv[0]; vo[1] -= fv[1]; vo[2] -= fv[2];
      v[0] -= vo[0]; v[1] -= vo[1]; v[2] -= vo[2];
      T n[3];
      col_normal(o, x, n);
      T nd = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
      if (nd < T(0)) {
        obj = i;
        if (o.sticky) {
          v[0] = v[1] = v[2] = T(0);
        } else {
          T t0 = v[0] - nd * n[0], t1 = v[1] - nd * n[1], t2 = v[2] - nd * n[2];
          T tn = msqrt(t0 * t0 + t1 * t1 + t2 * t2);
          if (tn >= -nd * o.friction) {
            T s = T(1) + o.friction * nd / tn;
            v[0] = t0 * s; v[1] = t1 * s; v[2] = t2 * s;
          } else {
            v[0] = v[1] = v[2] = T(0);
          }
        }
      }
      if (addVel) { v[0] += vo[0]; v[1] += vo[1]; v[2] += vo[2]; }
    }
  }
  return obj;
}

}  // namespace mpm
