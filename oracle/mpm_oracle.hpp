// mpm_oracle.hpp -- CPU oracle for the implicit MPM snow step.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker, never the product: it is
// compiled into oracle/_build/libmpm_oracle.so and loaded only by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg. The shipped GPU
// path (mpmsnow_amd/csrc) never includes, links or calls it.
//
// It restates, in plain C++17 templated on the scalar type (float = the
// reference's own precision, double = the FP64 oracle), the algorithm of
// davidsminor/mpmsnow's MpmSim library. Every function cites the reference
// file:line it follows (paths relative to the reference checkout). The
// reference itself cannot be compiled here (it needs Eigen 3.2.0 and TBB 4.1
// headers that the image lacks), so this restatement is the oracle of record.
//
// Third-party arithmetic the reference delegates to Eigen 3.2.0 is restated:
//  * JacobiSVD<Matrix3f>: replaced by a one-sided (Hestenes) Jacobi SVD run in
//    double and rounded to S. All quantities the model derives from it
//    (R = U V^T, F^-T, S = V Sigma V^T, J, clamped F, Fp) are unique functions
//    of F for nonsingular F, so the choice of SVD algorithm only changes
//    rounding. Pinned by invariants and by an independent LAPACK SVD in tests.
//  * VectorXf::dot / squaredNorm: restated as Eigen 3.2's SSE2 linear
//    vectorised reduction (two 16-byte packet accumulators, tail added
//    serially, horizontal add (l0+l2)+(l1+l3)). The exact summation order of
//    the reference binary is unpinned (it depends on its compiler flags).
//  * Matrix3f::inverse / determinant: cofactor formulas.
// TBB's parallel_for over the 8-colour voxel partition becomes an OpenMP
// parallel for over the same partition (Grid.inl:10-18); as in the reference
// the partition guarantees one writer per node, so results do not depend on
// the thread count.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <queue>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mpmo {

// ---------------------------------------------------------------- 3-vectors
template <class S> struct V3 {
  S c[3];
  V3() : c{S(0), S(0), S(0)} {}
  V3(S a, S b, S d) : c{a, b, d} {}
  S& operator[](int i) { return c[i]; }
  const S& operator[](int i) const { return c[i]; }
};
template <class S> inline V3<S> operator+(const V3<S>& a, const V3<S>& b) { return V3<S>(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
template <class S> inline V3<S> operator-(const V3<S>& a, const V3<S>& b) { return V3<S>(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
template <class S> inline V3<S> operator*(S s, const V3<S>& a) { return V3<S>(s * a[0], s * a[1], s * a[2]); }
template <class S> inline V3<S> operator*(const V3<S>& a, S s) { return V3<S>(a[0] * s, a[1] * s, a[2] * s); }
template <class S> inline V3<S> operator/(const V3<S>& a, S s) { return V3<S>(a[0] / s, a[1] / s, a[2] / s); }
template <class S> inline V3<S>& operator+=(V3<S>& a, const V3<S>& b) { a[0] += b[0]; a[1] += b[1]; a[2] += b[2]; return a; }
template <class S> inline V3<S>& operator-=(V3<S>& a, const V3<S>& b) { a[0] -= b[0]; a[1] -= b[1]; a[2] -= b[2]; return a; }
template <class S> inline S dot(const V3<S>& a, const V3<S>& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <class S> inline S sqnorm(const V3<S>& a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
template <class S> inline S norm(const V3<S>& a) { return std::sqrt(sqnorm(a)); }
template <class S> inline void normalize(V3<S>& a) { S n = norm(a); a[0] /= n; a[1] /= n; a[2] /= n; }

// --------------------------------------------------- 3x3 matrices (col-major)
template <class S> struct M3 {
  S a[9];  // a[r + 3*c], the memory order of Eigen::Matrix3f
  M3() { for (int i = 0; i < 9; ++i) a[i] = S(0); }
  S& operator()(int r, int c) { return a[r + 3 * c]; }
  const S& operator()(int r, int c) const { return a[r + 3 * c]; }
  static M3 identity() { M3 m; m(0, 0) = m(1, 1) = m(2, 2) = S(1); return m; }
  static M3 zero() { return M3(); }
};
template <class S> inline M3<S> operator*(const M3<S>& A, const M3<S>& B) {
  M3<S> R;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) R(r, c) = A(r, 0) * B(0, c) + A(r, 1) * B(1, c) + A(r, 2) * B(2, c);
  return R;
}
template <class S> inline V3<S> operator*(const M3<S>& A, const V3<S>& v) {
  return V3<S>(A(0, 0) * v[0] + A(0, 1) * v[1] + A(0, 2) * v[2], A(1, 0) * v[0] + A(1, 1) * v[1] + A(1, 2) * v[2],
               A(2, 0) * v[0] + A(2, 1) * v[1] + A(2, 2) * v[2]);
}
template <class S> inline M3<S> operator*(S s, const M3<S>& A) { M3<S> R; for (int i = 0; i < 9; ++i) R.a[i] = s * A.a[i]; return R; }
template <class S> inline M3<S> operator+(const M3<S>& A, const M3<S>& B) { M3<S> R; for (int i = 0; i < 9; ++i) R.a[i] = A.a[i] + B.a[i]; return R; }
template <class S> inline M3<S> operator-(const M3<S>& A, const M3<S>& B) { M3<S> R; for (int i = 0; i < 9; ++i) R.a[i] = A.a[i] - B.a[i]; return R; }
template <class S> inline M3<S>& operator+=(M3<S>& A, const M3<S>& B) { for (int i = 0; i < 9; ++i) A.a[i] += B.a[i]; return A; }
template <class S> inline M3<S> transpose(const M3<S>& A) { M3<S> R; for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) R(r, c) = A(c, r); return R; }
// u v^T
template <class S> inline M3<S> outer(const V3<S>& u, const V3<S>& v) { M3<S> R; for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) R(r, c) = u[r] * v[c]; return R; }
// first-column cofactor expansion
template <class S> inline S determinant(const M3<S>& m) {
  return m(0, 0) * (m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2)) - m(1, 0) * (m(0, 1) * m(2, 2) - m(2, 1) * m(0, 2)) +
         m(2, 0) * (m(0, 1) * m(1, 2) - m(1, 1) * m(0, 2));
}
template <class S> inline M3<S> inverse(const M3<S>& m) {
  M3<S> cof;  // cofactor of m(r,c) stored transposed (the adjugate)
  cof(0, 0) = m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1);
  cof(1, 0) = m(1, 2) * m(2, 0) - m(1, 0) * m(2, 2);
  cof(2, 0) = m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0);
  S det = m(0, 0) * cof(0, 0) + m(0, 1) * cof(1, 0) + m(0, 2) * cof(2, 0);
  cof(0, 1) = m(0, 2) * m(2, 1) - m(0, 1) * m(2, 2);
  cof(1, 1) = m(0, 0) * m(2, 2) - m(0, 2) * m(2, 0);
  cof(2, 1) = m(0, 1) * m(2, 0) - m(0, 0) * m(2, 1);
  cof(0, 2) = m(0, 1) * m(1, 2) - m(0, 2) * m(1, 1);
  cof(1, 2) = m(0, 2) * m(1, 0) - m(0, 0) * m(1, 2);
  cof(2, 2) = m(0, 0) * m(1, 1) - m(0, 1) * m(1, 0);
  M3<S> inv;
  for (int i = 0; i < 9; ++i) inv.a[i] = cof.a[i] / det;
  return inv;
}
// reference SnowConstitutiveModel.cpp:258-264: row-by-row double dot
template <class S> inline S matrixDoubleDot(const M3<S>& a, const M3<S>& b) {
  return a(0, 0) * b(0, 0) + a(0, 1) * b(0, 1) + a(0, 2) * b(0, 2) + a(1, 0) * b(1, 0) + a(1, 1) * b(1, 1) +
         a(1, 2) * b(1, 2) + a(2, 0) * b(2, 0) + a(2, 1) * b(2, 1) + a(2, 2) * b(2, 2);
}
template <class S> inline bool finite3(S a, S b, S c) { S p = a * b * c; return std::isfinite(p); }

// ------------------------------------------------------------ 3x3 SVD (FP64)
// One-sided Jacobi (Hestenes) on the columns of A: A V = U Sigma. Returns
// sigma >= 0 sorted descending, U and V orthogonal (Eigen JacobiSVD
// conventions: sigma non-negative and descending; det U, det V may be -1).
inline void svd3_double(const double A[9], double U[9], double sig[3], double V[9]) {
  double B[9];
  std::memcpy(B, A, sizeof(B));
  double W[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double a = 0, b = 0, g = 0;
        for (int r = 0; r < 3; ++r) {
          a += B[r + 3 * p] * B[r + 3 * p];
          b += B[r + 3 * q] * B[r + 3 * q];
          g += B[r + 3 * p] * B[r + 3 * q];
        }
        if (g == 0.0) continue;
        double rel = std::fabs(g) / std::sqrt(a * b);
        if (!(rel > 1e-17)) continue;
        off = std::max(off, rel);
        double zeta = (b - a) / (2.0 * g);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (int r = 0; r < 3; ++r) {
          double bp = B[r + 3 * p], bq = B[r + 3 * q];
          B[r + 3 * p] = cs * bp - sn * bq;
          B[r + 3 * q] = sn * bp + cs * bq;
          double wp = W[r + 3 * p], wq = W[r + 3 * q];
          W[r + 3 * p] = cs * wp - sn * wq;
          W[r + 3 * q] = sn * wp + cs * wq;
        }
      }
    if (off < 1e-16) break;
  }
  double s[3];
  for (int c = 0; c < 3; ++c) s[c] = std::sqrt(B[3 * c] * B[3 * c] + B[3 * c + 1] * B[3 * c + 1] + B[3 * c + 2] * B[3 * c + 2]);
  int ord[3] = {0, 1, 2};
  std::sort(ord, ord + 3, [&](int i, int j) { return s[i] > s[j]; });
  for (int k = 0; k < 3; ++k) {
    sig[k] = s[ord[k]];
    for (int r = 0; r < 3; ++r) {
      V[r + 3 * k] = W[r + 3 * ord[k]];
      U[r + 3 * k] = B[r + 3 * ord[k]];
    }
  }
  // normalise U columns; complete degenerate ones orthonormally
  const double tiny = 1e-300;
  for (int k = 0; k < 3; ++k) {
    if (sig[k] > tiny && sig[k] > 1e-14 * sig[0]) {
      for (int r = 0; r < 3; ++r) U[r + 3 * k] /= sig[k];
    } else {
      // pick a unit vector orthogonal to the previous columns
      double cand[3];
      if (k == 0) {
        cand[0] = 1; cand[1] = 0; cand[2] = 0;
      } else if (k == 1) {
        const double* u0 = U;
        int m = 0;
        for (int r = 1; r < 3; ++r) if (std::fabs(u0[r]) < std::fabs(u0[m])) m = r;
        double e[3] = {0, 0, 0};
        e[m] = 1;
        double d = e[0] * u0[0] + e[1] * u0[1] + e[2] * u0[2];
        for (int r = 0; r < 3; ++r) cand[r] = e[r] - d * u0[r];
        double n = std::sqrt(cand[0] * cand[0] + cand[1] * cand[1] + cand[2] * cand[2]);
        for (int r = 0; r < 3; ++r) cand[r] /= n;
      } else {
        const double *u0 = U, *u1 = U + 3;
        cand[0] = u0[1] * u1[2] - u0[2] * u1[1];
        cand[1] = u0[2] * u1[0] - u0[0] * u1[2];
        cand[2] = u0[0] * u1[1] - u0[1] * u1[0];
      }
      for (int r = 0; r < 3; ++r) U[r + 3 * k] = cand[r];
    }
  }
}

// ------------------------------------------------------- cubic B-spline
// reference src/MpmSim/CubicBsplineShapeFunction.cpp:6-48
template <class S> inline S bspline_w(S x) {
  S ax = std::fabs(x);
  if (ax < 1) return (S(0.5) * ax - S(1)) * ax * ax + S(2) / S(3);
  if (ax < 2) return ((-ax / S(6) + S(1)) * ax - S(2)) * ax + S(4) / S(3);
  return S(0);
}
template <class S> inline S bspline_dw(S x) {
  if (x < 0) return -bspline_dw(-x);
  if (x < 1) return x * (S(1.5) * x - S(2));
  if (x < 2) { x -= S(2); return S(-0.5) * x * x; }
  return S(0);
}
constexpr int kSupportRadius = 2;  // CubicBsplineShapeFunction.cpp:6-9

// -------------------------------------------------- material point data
// reference include/MpmSim/MaterialPointData.h, src/MpmSim/MaterialPointData.cpp:8-16
template <class S> struct Particles {
  std::vector<V3<S>> p, v;
  std::vector<M3<S>> F;
  std::vector<S> m, volume;
  // created by SnowConstitutiveModel::setParticles (SnowConstitutiveModel.cpp:123-169)
  bool modelVars = false;
  std::vector<M3<S>> Fp, FinvTrans, R, Ginv;
  std::vector<S> J, mu, lambda;
  // optional per-particle material (extension: per-particle E0 and xi); empty = scalar model values
  std::vector<S> mu0, lambda0, xi;
  size_t size() const { return m.size(); }
  size_t numVariables() const { return 5 + (modelVars ? 7 : 0) + (mu0.empty() ? 0 : 3); }
};

// ------------------------------------------------------ constitutive model
// reference include/MpmSim/ConstitutiveModel.h:15-31
template <class S> struct ConstitutiveModel {
  virtual ~ConstitutiveModel() {}
  virtual void setParticles(Particles<S>& p) = 0;
  virtual void updateParticleData() = 0;
  virtual S energyDensity(size_t p) const = 0;
  virtual M3<S> dEnergyDensitydF(size_t p) const = 0;
  virtual M3<S> dEdFDifferential(const M3<S>& dFp, size_t p) const = 0;
};

// zero-stress model (the reference tests' DummyModel, TestSimClass.cpp:20-43)
template <class S> struct DummyModel : ConstitutiveModel<S> {
  void setParticles(Particles<S>&) override {}
  void updateParticleData() override {}
  S energyDensity(size_t) const override { return 0; }
  M3<S> dEnergyDensitydF(size_t) const override { return M3<S>(); }
  M3<S> dEdFDifferential(const M3<S>&, size_t) const override { return M3<S>(); }
};

template <class S> struct SnowModel : ConstitutiveModel<S> {
  S E, nu, hardening, thc, ths, mu0, lambda0;
  Particles<S>* P = nullptr;
  // SnowConstitutiveModel.cpp:9-25
  SnowModel(S youngsModulus, S poissonRatio, S hard, S compressive, S tensile)
      : E(youngsModulus), nu(poissonRatio), hardening(hard), thc(compressive), ths(tensile) {
    mu0 = (youngsModulus / (S(2) * (S(1) + poissonRatio)));
    lambda0 = (youngsModulus * poissonRatio / ((S(1) + poissonRatio) * (S(1) - S(2) * poissonRatio)));
  }
  // SnowConstitutiveModel.cpp:123-169
  void setParticles(Particles<S>& p) override {
    if (p.modelVars) throw std::runtime_error("Material point data already has a variable named 'Fp'");
    P = &p;
    p.modelVars = true;
    size_t n = p.size();
    p.Fp.assign(n, M3<S>::identity());
    p.FinvTrans.assign(n, M3<S>::identity());
    p.R.assign(n, M3<S>::identity());
    p.Ginv.assign(n, M3<S>::identity());
    p.J.assign(n, S(1));
    p.mu.assign(n, mu0);
    p.lambda.assign(n, lambda0);
  }
  S pmu0(size_t p) const { return P->mu0.empty() ? mu0 : P->mu0[p]; }
  S plambda0(size_t p) const { return P->lambda0.empty() ? lambda0 : P->lambda0[p]; }
  S pxi(size_t p) const { return P->xi.empty() ? hardening : P->xi[p]; }

  // SnowConstitutiveModel.cpp:27-121 for one particle
  void updateOne(size_t p) {
    Particles<S>& d = *P;
    double Fd[9], Ud[9], sd[3], Vd[9];
    for (int i = 0; i < 9; ++i) Fd[i] = double(d.F[p].a[i]);
    svd3_double(Fd, Ud, sd, Vd);
    M3<S> U, V;
    for (int i = 0; i < 9; ++i) { U.a[i] = S(Ud[i]); V.a[i] = S(Vd[i]); }
    S sv[3] = {S(sd[0]), S(sd[1]), S(sd[2])};
    M3<S> D, Dinv;
    bool modified = false;
    for (int i = 0; i < 3; ++i) {
      if (sv[i] > S(1) + ths) { modified = true; sv[i] = S(1) + ths; }
      if (sv[i] < S(1) - thc) { modified = true; sv[i] = S(1) - thc; }
      D(i, i) = sv[i];
      Dinv(i, i) = S(1) / sv[i];
    }
    M3<S> Ut = transpose(U), Vt = transpose(V);
    if (modified) {
      M3<S> FN = d.F[p] * d.Fp[p];
      d.Fp[p] = ((V * Dinv) * Ut) * FN;
      d.F[p] = (U * D) * Vt;
    }
    d.FinvTrans[p] = (U * Dinv) * Vt;
    d.R[p] = U * Vt;
    M3<S> Sm = (V * D) * Vt;
    M3<S> G;
    G(0, 0) = Sm(0, 0) + Sm(1, 1);
    G(1, 1) = Sm(0, 0) + Sm(2, 2);
    G(2, 2) = Sm(1, 1) + Sm(2, 2);
    G(0, 1) = G(1, 0) = Sm(1, 2);
    G(0, 2) = G(2, 0) = -Sm(0, 2);
    G(1, 2) = G(2, 1) = Sm(0, 1);
    d.Ginv[p] = inverse(G);
    d.J[p] = D(0, 0) * D(1, 1) * D(2, 2);
    S hf = pxi(p) * (S(1) - determinant(d.Fp[p]));
    if (hf > S(2)) hf = S(2);
    hf = std::exp(hf);
    if (!std::isfinite(hf)) throw std::runtime_error("infinite hardness!!!");
    d.mu[p] = pmu0(p) * hf;
    d.lambda[p] = plambda0(p) * hf;
  }
  void updateParticleData() override {
    for (size_t p = 0; p < P->F.size(); ++p) updateOne(p);
  }
  // SnowConstitutiveModel.cpp:171-185
  S energyDensity(size_t p) const override {
    const Particles<S>& d = *P;
    M3<S> rd = d.F[p] - d.R[p];
    S Jm1 = d.J[p] - S(1);
    S dd = matrixDoubleDot(rd, rd);
    return (d.mu[p] * dd + S(0.5) * d.lambda[p] * Jm1 * Jm1);
  }
  // SnowConstitutiveModel.cpp:187-216
  M3<S> dEnergyDensitydF(size_t p) const override {
    const Particles<S>& d = *P;
    M3<S> rd = d.F[p] - d.R[p];
    S a = S(2) * d.mu[p];
    S b = d.lambda[p] * (d.J[p] - S(1)) * d.J[p];
    M3<S> ret;
    for (int i = 0; i < 9; ++i) ret.a[i] = a * rd.a[i] + b * d.FinvTrans[p].a[i];
    S n2 = 0;
    for (int i = 0; i < 9; ++i) n2 += ret.a[i] * ret.a[i];
    if (!std::isfinite(std::sqrt(n2))) throw std::runtime_error("nans in dEnergyDensitydF matrix!");
    return ret;
  }
  // SnowConstitutiveModel.cpp:266-287
  static M3<S> computeRdifferential(const M3<S>& dF, const M3<S>& R, const M3<S>& Ginv) {
    M3<S> M = transpose(R) * dF;
    V3<S> w(M(0, 1) - M(1, 0), M(0, 2) - M(2, 0), M(1, 2) - M(2, 1));
    w = Ginv * w;
    M3<S> RtdR;
    RtdR(0, 1) = w[0];
    RtdR(1, 0) = -w[0];
    RtdR(0, 2) = w[1];
    RtdR(2, 0) = -w[1];
    RtdR(1, 2) = w[2];
    RtdR(2, 1) = -w[2];
    return R * RtdR;
  }
  // SnowConstitutiveModel.cpp:218-256
  M3<S> dEdFDifferential(const M3<S>& dFp, size_t p) const override {
    const Particles<S>& d = *P;
    const M3<S>& Fi = d.FinvTrans[p];
    S J = d.J[p];
    S dJ = J * matrixDoubleDot(Fi, dFp);
    M3<S> nFi = S(-1) * Fi;
    M3<S> dFInvTrans = (nFi * transpose(dFp)) * Fi;
    M3<S> dR = computeRdifferential(dFp, d.R[p], d.Ginv[p]);
    S twoMu = S(2) * d.mu[p];
    S lam = d.lambda[p];
    S dJJ = dJ * J;
    S Jm1 = J - S(1);
    M3<S> ret;
    for (int i = 0; i < 9; ++i)
      ret.a[i] = twoMu * (dFp.a[i] - dR.a[i]) + lam * (dJJ * Fi.a[i] + Jm1 * (dJ * Fi.a[i] + J * dFInvTrans.a[i]));
    return ret;
  }
};

// ------------------------------------------------------------ colliders
// reference include/MpmSim/CollisionObject.h:10-45, CollisionPlane.cpp:23-56;
// the sphere is the analytic sphere collider extension (phi = |x-c| - r).
// Type 2 is a level-set volume standing in for the reference's VDBCollisionObject
// (src/houdiniPlugin/VDBCollisionObject.cpp:11-37: phi, gradient and velocity read from OpenVDB
// grids through the Houdini HDK, which is absent here, so the sampling rule is this build's own):
// phi and an optional velocity per point at origin + (i, j, k) voxel, trilinear interpolation
// clamped to the volume, each lerp a + t (b - a); gradient = central difference of the sampled
// phi over one voxel per axis.
template <class S> struct Collider {
  int type = 0;  // 0 plane, 1 sphere, 2 level-set volume
  S coeffs[4] = {0, 0, 0, 0};
  V3<S> center;
  S radius = 0;
  V3<S> vel;
  S friction = S(0.5);
  bool sticky = false;
  int dims[3] = {0, 0, 0};
  V3<S> origin;
  S voxel = S(1);
  std::vector<S> phiv, velv;  // velv: 3 values per point, or empty (uniform vel)
  void sample(const std::vector<S>& f, int nc, const V3<S>& x, S* out) const {
    int i0[3];
    S t[3];
    for (int a = 0; a < 3; ++a) {
      S u = (x[a] - origin[a]) / voxel;
      const S hi = S(dims[a] - 1);
      if (!(u >= S(0))) u = S(0);  // below the volume, or NaN
      if (u > hi) u = hi;
      int i = int(std::floor(u));
      if (i > dims[a] - 2) i = dims[a] - 2;
      i0[a] = i;
      t[a] = u - S(i);
    }
    auto lerp = [](S a, S b, S tt) { return a + tt * (b - a); };
    const size_t sx = size_t(nc), sy = sx * size_t(dims[0]), sz = sy * size_t(dims[1]);
    const size_t b = size_t(i0[0]) * sx + size_t(i0[1]) * sy + size_t(i0[2]) * sz;
    for (int c = 0; c < nc; ++c) {
      const S* q = f.data() + b + size_t(c);
      const S e00 = lerp(q[0], q[sx], t[0]);
      const S e10 = lerp(q[sy], q[sy + sx], t[0]);
      const S e01 = lerp(q[sz], q[sz + sx], t[0]);
      const S e11 = lerp(q[sz + sy], q[sz + sy + sx], t[0]);
      out[c] = lerp(lerp(e00, e10, t[1]), lerp(e01, e11, t[1]), t[2]);
    }
  }
  S phi(const V3<S>& x) const {
    if (type == 0) return coeffs[0] * x[0] + coeffs[1] * x[1] + coeffs[2] * x[2] + coeffs[3];
    if (type == 2) {
      S p;
      sample(phiv, 1, x, &p);
      return p;
    }
    return norm(x - center) - radius;
  }
  void grad(const V3<S>& x, V3<S>& g) const {
    if (type == 0) { g = V3<S>(coeffs[0], coeffs[1], coeffs[2]); return; }
    if (type == 2) {
      for (int a = 0; a < 3; ++a) {
        V3<S> xp = x, xm = x;
        xp[a] = x[a] + voxel;
        xm[a] = x[a] - voxel;
        S pp, pm;
        sample(phiv, 1, xp, &pp);
        sample(phiv, 1, xm, &pm);
        g[a] = (pp - pm) / (S(2) * voxel);
      }
      return;
    }
    g = x - center;
  }
  void velocity(const V3<S>& x, V3<S>& v) const {
    if (type == 2 && !velv.empty()) {
      S o[3];
      sample(velv, 3, x, o);
      v = V3<S>(o[0], o[1], o[2]);
      return;
    }
    v = vel;
  }
};

template <class S> struct CollisionSet {
  std::vector<Collider<S>> objects;
  // src/CollisionObject.cpp:19-96
  int collide(V3<S>& v, const V3<S>& x, const V3<S>& frameVelocity, bool addCollisionVelocity = false) const {
    int collisionObject = -1;
    bool intersectedAlready = false;
    for (size_t objIdx = 0; objIdx < objects.size(); ++objIdx) {
      S phi = objects[objIdx].phi(x);
      if (phi <= 0) {
        if (intersectedAlready) {
          collisionObject = -2;
          v = V3<S>();
          break;
        }
        intersectedAlready = true;
        V3<S> vObj;
        objects[objIdx].velocity(x, vObj);
        vObj -= frameVelocity;
        v -= vObj;
        V3<S> n;
        objects[objIdx].grad(x, n);
        normalize(n);
        S nDotV = dot(n, v);
        if (nDotV < 0) {
          collisionObject = int(objIdx);
          if (objects[objIdx].sticky) {
            v = V3<S>();
          } else {
            V3<S> vPerp = nDotV * n;
            V3<S> vTangent = v - vPerp;
            S vtNorm = norm(vTangent);
            S cf = objects[objIdx].friction;
            if (vtNorm >= -nDotV * cf) {
              v = vTangent * (S(1) + cf * nDotV / norm(vTangent));
            } else {
              v = V3<S>();
            }
          }
        }
        if (addCollisionVelocity) v += vObj;
      }
    }
    return collisionObject;
  }
};

// ------------------------------------------------------------ force fields
// reference src/ForceField.cpp:19-25, GravityField.cpp:9-17
template <class S> struct ForceSet {
  std::vector<V3<S>> gravity;
  void force(V3<S>& f, const V3<S>& /*x*/, S m) const {
    for (size_t i = 0; i < gravity.size(); ++i) f += gravity[i] * m;
  }
};

// ------------------------------------------- Eigen-3.2-style vector reductions
// Restates Eigen 3.2's LinearVectorizedTraversal redux under SSE2 (16-byte
// packets): two packet accumulators over the aligned body, combined, then the
// SSE2 horizontal add, then the tail added serially. Summation order of the
// reference binary itself is unpinned.
template <class S, class Term> inline S eigen_redux(size_t n, Term term) {
  const size_t P = 16 / sizeof(S);
  if (n == 0) return S(0);
  size_t aligned = (n / P) * P;
  if (aligned == 0) {
    S r = term(0);
    for (size_t i = 1; i < n; ++i) r += term(i);
    return r;
  }
  S p0[8], p1[8];
  for (size_t k = 0; k < P; ++k) p0[k] = term(k);
  size_t aligned2 = (n / (2 * P)) * (2 * P);
  if (aligned > P) {
    for (size_t k = 0; k < P; ++k) p1[k] = term(P + k);
    for (size_t i = 2 * P; i < aligned2; i += 2 * P)
      for (size_t k = 0; k < P; ++k) { p0[k] += term(i + k); p1[k] += term(i + P + k); }
    for (size_t k = 0; k < P; ++k) p0[k] += p1[k];
    if (aligned > aligned2)
      for (size_t k = 0; k < P; ++k) p0[k] += term(aligned2 + k);
  }
  S r;
  if (P == 4) r = (p0[0] + p0[2]) + (p0[1] + p0[3]);
  else if (P == 2) r = p0[0] + p0[1];
  else { r = p0[0]; for (size_t k = 1; k < P; ++k) r += p0[k]; }
  for (size_t i = aligned; i < n; ++i) r += term(i);
  return r;
}
template <class S> using Vec = std::vector<S>;
template <class S> inline S vdot(const Vec<S>& a, const Vec<S>& b) {
  return eigen_redux<S>(a.size(), [&](size_t i) { return a[i] * b[i]; });
}
template <class S> inline S vsqnorm(const Vec<S>& a) {
  return eigen_redux<S>(a.size(), [&](size_t i) { return a[i] * a[i]; });
}

// ------------------------------------------------------------- linear solver
// reference include/MpmSim/ProceduralMatrix.h:8-19, LinearSolver.h:12-29,
// TerminationCriterion.h:10-16
template <class S> struct ProceduralMatrix {
  virtual ~ProceduralMatrix() {}
  virtual void multVector(const Vec<S>& x, Vec<S>& result) const = 0;
  virtual void multInverseVector(const Vec<S>& x, Vec<S>& result) const = 0;
  virtual void subspaceProject(Vec<S>& x) const = 0;
};

template <class S> struct Termination {
  // SquareMagnitudeTermination.cpp:8-33 (maxIters is stored as float there)
  float maxIters;
  S tol;
  S threshold = 0;
  bool cancelled = false;
  // TerminationCriterion::cancelled() as a user interrupt arriving during a solve: from its
  // cancelOnCall-th call on it returns true (Sim::advance calls it once per body, Sim.cpp:107)
  int cancelOnCall = -1;
  int cancelCalls = 0;
  bool pollCancelled() {
    ++cancelCalls;
    if (cancelOnCall > 0 && cancelCalls >= cancelOnCall) cancelled = true;
    return cancelled;
  }
  std::vector<S> history;  // ||r||^2 per call (the reference prints sqrt of it)
  Termination(int maxIt, S tolError) : maxIters(float(maxIt)), tol(tolError) {}
  void init(const Vec<S>& b) { S bn = vsqnorm(b); threshold = tol * tol * bn; history.clear(); }
  bool operator()(const Vec<S>& r, int it) {
    if (float(it) >= maxIters) return true;
    S rn = vsqnorm(r);
    history.push_back(rn);
    return rn < threshold;
  }
};

template <class S> struct SolveStats {
  int iterations = 0;        // number of x updates
  S bNorm2 = 0;
  bool earlyExit = false;    // zero denominator
  std::vector<S> residualNorm2;
};

// ConjugateResiduals.cpp:19-160
template <class S>
void conjugateResiduals(const ProceduralMatrix<S>& A, const Vec<S>& b, Vec<S>& x, Termination<S>& term,
                        const ProceduralMatrix<S>* precond, SolveStats<S>* stats,
                        std::vector<Vec<S>>* logResiduals = nullptr, std::vector<Vec<S>>* logDirections = nullptr,
                        const std::function<void(Vec<S>&)>* debug = nullptr) {
  const size_t N = b.size();
  S bNorm2 = vsqnorm(b);
  if (stats) stats->bNorm2 = bNorm2;
  if (bNorm2 == 0) {
    std::fill(x.begin(), x.end(), S(0));
    return;
  }
  term.init(b);
  Vec<S> r(N), rp(N), Ar(N), p(N), Ap(N), pAp(N), Ax(N);
  A.multVector(x, Ax);
  for (size_t i = 0; i < N; ++i) r[i] = b[i] - Ax[i];
  if (precond) precond->multInverseVector(r, rp); else rp = r;
  p = rp;
  A.multVector(p, Ap);
  A.multVector(rp, Ar);
  S rz = vdot(rp, Ar);
  if (logResiduals) { logResiduals->push_back(r); logDirections->push_back(p); }
  for (int i = 0;; ++i) {
    if (precond) precond->multInverseVector(Ap, pAp); else pAp = Ap;
    S ApdPAp = vdot(Ap, pAp);
    if (ApdPAp == 0) { if (stats) stats->earlyExit = true; return; }
    S alpha = rz / ApdPAp;
    for (size_t k = 0; k < N; ++k) x[k] = x[k] + alpha * p[k];
    A.subspaceProject(x);
    if (stats) stats->iterations = i + 1;
    if (debug && *debug) (*debug)(x);
    for (size_t k = 0; k < N; ++k) rp[k] = rp[k] - alpha * pAp[k];
    if (precond) precond->multVector(rp, r); else r = rp;
    if (term(r, i)) break;
    A.multVector(rp, Ar);
    S rz_old = rz;
    rz = vdot(rp, Ar);
    if (rz_old == 0) { if (stats) stats->earlyExit = true; return; }
    S beta = rz / rz_old;
    for (size_t k = 0; k < N; ++k) p[k] = rp[k] + beta * p[k];
    for (size_t k = 0; k < N; ++k) Ap[k] = Ar[k] + beta * Ap[k];
    if (logResiduals) { logResiduals->push_back(r); logDirections->push_back(p); }
  }
  if (stats) stats->residualNorm2 = term.history;
}

// ------------------------------------------------------------------ grid
typedef std::vector<int> IndexList;

template <class S> struct ShapeIt {
  // reference Grid.cpp:1094-1193 (Grid::ShapeFunctionIterator)
  S w[3][4], dw[3][4];
  int base[3];
  int pos[3];
  int dim;
  S h;
  void initialize(const V3<S>& p, const V3<S>& gmin, S gridSize, int dimension, bool derivs) {
    dim = dimension;
    h = gridSize;
    for (int d = 0; d < 3; ++d) {
      pos[d] = 0;
      base[d] = 0;
      for (int i = 0; i < 4; ++i) { w[d][i] = S(1); dw[d][i] = S(0); }
    }
    const int r = kSupportRadius;
    for (int d = 0; d < dim; ++d) {
      S frac = (p[d] - gmin[d]) / gridSize;
      int dimPos = int(std::floor(frac));
      frac -= S(dimPos);
      base[d] = dimPos + 1 - r;
      int j = 1 - r;
      for (int i = 0; i < 4; ++i, ++j) w[d][i] = bspline_w<S>(S(j) - frac);
      if (derivs) {
        j = 1 - r;
        for (int i = 0; i < 4; ++i, ++j) dw[d][i] = bspline_dw<S>(S(j) - frac) / gridSize;
      }
    }
  }
  bool next() {
    ++pos[0];
    if (pos[0] >= 4) {
      if (dim == 1) return false;
      pos[0] = 0;
      ++pos[1];
      if (pos[1] >= 4) {
        if (dim == 2) return false;
        pos[1] = 0;
        ++pos[2];
      }
      if (pos[2] >= 4) return false;
    }
    return true;
  }
  void gridPos(int g[3]) const { for (int d = 0; d < 3; ++d) g[d] = pos[d] + base[d]; }
  S wt() const { return w[0][pos[0]] * w[1][pos[1]] * w[2][pos[2]]; }
  V3<S> grad() const {
    return V3<S>(-dw[0][pos[0]] * w[1][pos[1]] * w[2][pos[2]], -w[0][pos[0]] * dw[1][pos[1]] * w[2][pos[2]],
                 -w[0][pos[0]] * w[1][pos[1]] * dw[2][pos[2]]);
  }
};

template <class S> struct SpatialLess {
  int d;
  const std::vector<V3<S>>* x;
  bool operator()(int a, int b) const { return (*x)[a][d] < (*x)[b][d]; }
};

// Grid.cpp:318-352
template <class S>
void voxelSort(IndexList::iterator begin, IndexList::iterator end, S voxelSize, const std::vector<V3<S>>& x, int dim = 0) {
  std::sort(begin, end, SpatialLess<S>{dim, &x});
  if (dim == 2) return;
  int currentSlice = int(std::floor(x[*begin][dim] / voxelSize));
  IndexList::iterator sliceBegin = begin, sliceEnd = begin + 1;
  for (;; ++sliceEnd) {
    int particleSlice = sliceEnd == end ? currentSlice + 1 : int(std::floor(x[*sliceEnd][dim] / voxelSize));
    if (particleSlice != currentSlice) {
      voxelSort(sliceBegin, sliceEnd, voxelSize, x, dim + 1);
      if (sliceEnd == end) break;
      sliceBegin = sliceEnd;
      currentSlice = particleSlice;
    }
  }
}

template <class S> struct Grid {
  Particles<S>& d;
  IndexList inds;
  S h;
  V3<S> gmin, gmax, frameV;
  int n[3];
  int dim;
  Vec<S> masses, velocities, prevVelocities;
  std::vector<signed char> nodeCollided;
  // 8 colour lists of [begin,end) offsets into inds (Grid.h:134-136)
  std::vector<std::pair<size_t, size_t>> parts[8];

  int coordsToIndex(int i, int j, int k) const { return i + n[0] * (j + n[1] * k); }  // Grid.cpp:1083-1086
  size_t numNodes() const { return size_t(n[0]) * n[1] * n[2]; }
  V3<S> nodePos(int i, int j, int k) const {
    return V3<S>(h * S(i) + gmin[0], h * S(j) + gmin[1], h * S(k) + gmin[2]);
  }
  void shapeInit(ShapeIt<S>& it, size_t p, bool derivs) const { it.initialize(d.p[p], gmin, h, dim, derivs); }

  // Grid.inl:10-18 with the partition of Grid.cpp:354-384
  template <class F> void splat(F f) const {
    for (int c = 0; c < 8; ++c) {
      const int np = int(parts[c].size());
#pragma omp parallel for schedule(dynamic, 1)
      for (int v = 0; v < np; ++v) {
        ShapeIt<S> it;
        for (size_t k = parts[c][v].first; k < parts[c][v].second; ++k) f(inds[k], it);
      }
    }
  }

  // Grid.cpp:354-384
  void computeProcessingPartitions() {
    S voxelSize = S(2 * kSupportRadius) * h;
    voxelSort<S>(inds.begin(), inds.end(), voxelSize, d.p);
    int cur[3] = {0, 0, 0};
    int partIdx = -1;
    for (size_t k = 0; k < inds.size(); ++k) {
      V3<S> x = d.p[inds[k]] / voxelSize;
      int vox[3] = {int(std::floor(x[0])), int(std::floor(x[1])), int(std::floor(x[2]))};
      if (vox[0] != cur[0] || vox[1] != cur[1] || vox[2] != cur[2] || k == 0) {
        cur[0] = vox[0]; cur[1] = vox[1]; cur[2] = vox[2];
        partIdx = (vox[0] & 1) + 2 * (vox[1] & 1) + 4 * (vox[2] & 1);
        parts[partIdx].push_back(std::make_pair(k, k));
      }
      parts[partIdx].back().second += 1;
    }
  }

  // Grid.cpp:156-263
  Grid(Particles<S>& data, const IndexList& particleInds, S gridSize, V3<S> frameVelocity = V3<S>(), int dimension = 3)
      : d(data), inds(particleInds), h(gridSize), frameV(frameVelocity), dim(dimension) {
    gmin = V3<S>(S(1e10), S(1e10), S(1e10));
    gmax = V3<S>(S(-1e10), S(-1e10), S(-1e10));
    for (size_t k = 0; k < inds.size(); ++k) {
      int p = inds[k];
      for (int j = 0; j < dim; ++j) {
        if (d.p[p][j] < gmin[j]) gmin[j] = d.p[p][j];
        if (d.p[p][j] > gmax[j]) gmax[j] = d.p[p][j];
      }
    }
    for (int j = 0; j < dim; ++j) {
      int cellMin = int(std::floor(gmin[j] / h)) - kSupportRadius - 1;
      int cellMax = int(std::ceil(gmax[j] / h)) + kSupportRadius + 1;
      gmin[j] = S(cellMin) * h;
      gmax[j] = S(cellMax) * h;
      n[j] = cellMax - cellMin;
    }
    for (int j = dim; j < 3; ++j) { gmin[j] = gmax[j] = 0; n[j] = 1; }
    long long ncells = (long long)n[0] * (long long)n[1] * (long long)n[2];
    if (ncells > 4000000000LL || ncells <= 0) throw std::runtime_error("grid is too big");
    computeProcessingPartitions();
    masses.assign(size_t(ncells), S(0));
    // MassSplatter, Grid.cpp:62-99
    splat([&](int p, ShapeIt<S>& it) {
      shapeInit(it, p, false);
      int g[3];
      do {
        it.gridPos(g);
        int idx = coordsToIndex(g[0], g[1], g[2]);
        masses[idx] += d.m[p] * it.wt();
      } while (it.next());
    });
    for (size_t i = 0; i < masses.size(); ++i) {
      if (masses[i] < 0) masses[i] = 0;
      if (!std::isfinite(masses[i])) throw std::runtime_error("nans in splatted masses!");
    }
    velocities.assign(size_t(ncells) * 3, S(0));
    // VelocitySplatter, Grid.cpp:102-150
    splat([&](int p, ShapeIt<S>& it) {
      shapeInit(it, p, false);
      int g[3];
      do {
        it.gridPos(g);
        int idx = coordsToIndex(g[0], g[1], g[2]);
        S cm = masses[idx];
        if (cm > 0) {
          S s = it.wt() * (d.m[p] / cm);
          V3<S> dv = d.v[p] - frameV;
          for (int c = 0; c < 3; ++c) velocities[3 * idx + c] += s * dv[c];
        }
      } while (it.next());
    });
    for (size_t i = 0; i < velocities.size(); ++i)
      if (!std::isfinite(velocities[i])) throw std::runtime_error("nans in splatted velocities!");
    prevVelocities = velocities;
  }

  // Grid.cpp:265-292
  void computeParticleVolumes() const {
    ShapeIt<S> it;
    S cellVolume = h * h * h;
    for (size_t k = 0; k < inds.size(); ++k) {
      int p = inds[k];
      S density(0);
      shapeInit(it, p, false);
      int g[3];
      do {
        it.gridPos(g);
        int idx = coordsToIndex(g[0], g[1], g[2]);
        density += it.wt() * masses[idx] / cellVolume;
      } while (it.next());
      d.volume[p] = d.m[p] / density;
    }
  }

  // Grid.cpp:438-463 with ForceSplatter Grid.cpp:387-435
  void calculateForces(Vec<S>& forces, const ConstitutiveModel<S>& model, const ForceSet<S>& fields) const {
    forces.assign(velocities.size(), S(0));
    for (int i = 0; i < n[0]; ++i)
      for (int j = 0; j < n[1]; ++j)
        for (int k = 0; k < n[2]; ++k) {
          int idx = coordsToIndex(i, j, k);
          V3<S> f(forces[3 * idx], forces[3 * idx + 1], forces[3 * idx + 2]);
          fields.force(f, V3<S>(S(i), S(j), S(k)) * h + gmin, masses[idx]);
          for (int c = 0; c < 3; ++c) forces[3 * idx + c] = f[c];
        }
    splat([&](int p, ShapeIt<S>& it) {
      M3<S> fm = (d.volume[p] * model.dEnergyDensitydF(p)) * transpose(d.F[p]);
      shapeInit(it, p, true);
      int g[3];
      do {
        it.gridPos(g);
        V3<S> wg = it.grad();
        int idx = coordsToIndex(g[0], g[1], g[2]);
        V3<S> t = fm * wg;
        for (int c = 0; c < 3; ++c) forces[3 * idx + c] -= t[c];
      } while (it.next());
    });
  }

  // Grid.cpp:620-632 with ForceDifferentialSplatter Grid.cpp:549-618
  void calculateForceDifferentials(Vec<S>& df, const Vec<S>& dx, const ConstitutiveModel<S>& model) const {
    df.assign(velocities.size(), S(0));
    splat([&](int p, ShapeIt<S>& it) {
      M3<S> dFp;
      M3<S> Fp = d.F[p];
      shapeInit(it, p, true);
      int g[3];
      do {
        it.gridPos(g);
        V3<S> wg = it.grad();
        int idx = coordsToIndex(g[0], g[1], g[2]);
        V3<S> dxi(dx[3 * idx], dx[3 * idx + 1], dx[3 * idx + 2]);
        dFp += outer(dxi, wg) * Fp;
      } while (it.next());
      M3<S> fm = (d.volume[p] * model.dEdFDifferential(dFp, p)) * transpose(Fp);
      shapeInit(it, p, true);
      do {
        it.gridPos(g);
        V3<S> wg = it.grad();
        int idx = coordsToIndex(g[0], g[1], g[2]);
        V3<S> t = fm * wg;
        for (int c = 0; c < 3; ++c) df[3 * idx + c] -= t[c];
      } while (it.next());
    });
  }

  // Grid.cpp:539-546 with dFidXiSplatter Grid.cpp:466-537
  void dForceidXi(Vec<S>& out, const ConstitutiveModel<S>& model) const {
    out.assign(velocities.size(), S(0));
    splat([&](int p, ShapeIt<S>& it) {
      M3<S> Fp = d.F[p];
      M3<S> Ft = transpose(Fp);
      shapeInit(it, p, true);
      int g[3];
      do {
        it.gridPos(g);
        V3<S> wg = it.grad();
        int idx = coordsToIndex(g[0], g[1], g[2]);
        for (int a = 0; a < 3; ++a) {
          V3<S> e;
          e[a] = S(1);
          M3<S> dF = outer(e, wg) * Fp;
          M3<S> fm = (d.volume[p] * model.dEdFDifferential(dF, p)) * Ft;
          V3<S> t = fm * wg;
          out[3 * idx + a] += -t[a];
        }
      } while (it.next());
    });
  }

  // Grid.cpp:634-690
  void calculateExplicitMomenta(Vec<S>& em, std::vector<signed char>& collided, S dt, const ConstitutiveModel<S>& model,
                                const CollisionSet<S>& cols, const ForceSet<S>& fields) const {
    Vec<S> forces;
    calculateForces(forces, model, fields);
    em.assign(velocities.size(), S(0));
    collided.assign(masses.size(), -1);
    for (int i = 0; i < n[0]; ++i)
      for (int j = 0; j < n[1]; ++j)
        for (int k = 0; k < n[2]; ++k) {
          int idx = coordsToIndex(i, j, k);
          V3<S> f(forces[3 * idx], forces[3 * idx + 1], forces[3 * idx + 2]);
          V3<S> ev(velocities[3 * idx], velocities[3 * idx + 1], velocities[3 * idx + 2]);
          if (masses[idx] > 0) ev += dt * f / masses[idx];
          V3<S> x = nodePos(i, j, k);
          collided[idx] = (signed char)cols.collide(ev, x, frameV);
          for (int c = 0; c < 3; ++c) em[3 * idx + c] = ev[c] * masses[idx];
          if (!finite3(em[3 * idx], em[3 * idx + 1], em[3 * idx + 2])) throw std::runtime_error("nan in explicit momenta!");
        }
  }

  // Grid.cpp:692-724
  void collisionVelocities(Vec<S>& vc, const CollisionSet<S>& cols, const std::vector<signed char>& collided) const {
    vc.assign(3 * numNodes(), S(0));
    for (int i = 0; i < n[0]; ++i)
      for (int j = 0; j < n[1]; ++j)
        for (int k = 0; k < n[2]; ++k) {
          int idx = coordsToIndex(i, j, k);
          if (collided[idx] < 0) continue;
          V3<S> vObj;
          cols.objects[collided[idx]].velocity(nodePos(i, j, k), vObj);
          V3<S> r = vObj - frameV;
          for (int c = 0; c < 3; ++c) vc[3 * idx + c] = r[c];
        }
  }

  // Grid.cpp:778-818
  void subspaceProject(Vec<S>& v, const CollisionSet<S>& cols) const {
    for (int i = 0; i < n[0]; ++i)
      for (int j = 0; j < n[1]; ++j)
        for (int k = 0; k < n[2]; ++k) {
          int idx = coordsToIndex(i, j, k);
          int o = nodeCollided[idx];
          if (o == -1) continue;
          if (o == -2) { v[3 * idx] = v[3 * idx + 1] = v[3 * idx + 2] = 0; continue; }
          V3<S> vv(v[3 * idx], v[3 * idx + 1], v[3 * idx + 2]);
          V3<S> nrm;
          cols.objects[o].grad(nodePos(i, j, k), nrm);
          normalize(nrm);
          S nd = dot(nrm, vv);
          vv -= nd * nrm;
          for (int c = 0; c < 3; ++c) v[3 * idx + c] = vv[c];
        }
  }

  struct ImplicitMatrix : ProceduralMatrix<S> {
    // Grid.cpp:727-832
    const Grid& g;
    const ConstitutiveModel<S>& model;
    const CollisionSet<S>& cols;
    S dt;
    ImplicitMatrix(const Grid& gg, const ConstitutiveModel<S>& m, const CollisionSet<S>& c, S t) : g(gg), model(m), cols(c), dt(t) {}
    void multVector(const Vec<S>& v, Vec<S>& result) const override {
      result = v;
      subspaceProject(result);
      Vec<S> df;
      g.calculateForceDifferentials(df, result, model);
      for (size_t idx = 0; idx < g.masses.size(); ++idx)
        for (int c = 0; c < 3; ++c)
          result[3 * idx + c] = g.masses[idx] * result[3 * idx + c] - dt * dt * df[3 * idx + c];
      subspaceProject(result);
    }
    void multInverseVector(const Vec<S>&, Vec<S>&) const override {}
    void subspaceProject(Vec<S>& x) const override { g.subspaceProject(x, cols); }
  };

  struct DiagonalPreconditioner : ProceduralMatrix<S> {
    // Grid.cpp:836-886
    Vec<S> diag;
    DiagonalPreconditioner(const Grid& g, const ConstitutiveModel<S>& model, S dt) {
      g.dForceidXi(diag, model);
      S s = -dt * dt;
      for (size_t i = 0; i < diag.size(); ++i) diag[i] *= s;
      for (size_t i = 0; i < g.masses.size(); ++i)
        for (int c = 0; c < 3; ++c) {
          diag[3 * i + c] += g.masses[i];
          if (diag[3 * i + c] == 0) diag[3 * i + c] = 1;
        }
    }
    void multVector(const Vec<S>& x, Vec<S>& r) const override {
      r.resize(x.size());
      for (size_t i = 0; i < x.size(); ++i) r[i] = x[i] * diag[i];
    }
    void multInverseVector(const Vec<S>& x, Vec<S>& r) const override {
      r.resize(x.size());
      for (size_t i = 0; i < x.size(); ++i) r[i] = x[i] / diag[i];
    }
    void subspaceProject(Vec<S>&) const override {}
  };

  // Grid.cpp:889-1016
  void updateGridVelocities(S dt, const ConstitutiveModel<S>& model, const CollisionSet<S>& cols, const ForceSet<S>& fields,
                            Termination<S>& term, SolveStats<S>* stats = nullptr,
                            const std::function<void(Vec<S>&)>* debug = nullptr, Vec<S>* rhsOut = nullptr,
                            V3<S>* guessOut = nullptr) {
    Vec<S> em;
    calculateExplicitMomenta(em, nodeCollided, dt, model, cols, fields);
    Vec<S> vc;
    collisionVelocities(vc, cols, nodeCollided);
    S totalMass = 0, collidedMass = 0;
    V3<S> totalMomentum, collidedMomentum;
    for (size_t idx = 0; idx < masses.size(); ++idx) {
      totalMass += masses[idx];
      totalMomentum += V3<S>(em[3 * idx], em[3 * idx + 1], em[3 * idx + 2]);
      if (nodeCollided[idx] >= 0) {
        collidedMass += masses[idx];
        collidedMomentum += masses[idx] * V3<S>(vc[3 * idx], vc[3 * idx + 1], vc[3 * idx + 2]);
      }
    }
    V3<S> guess;
    if (totalMass == 0) guess = V3<S>();
    else if (collidedMass == 0) guess = totalMomentum / totalMass;
    else guess = collidedMomentum / collidedMass;
    if (guessOut) *guessOut = guess;
    for (size_t idx = 0; idx < masses.size(); ++idx)
      for (int c = 0; c < 3; ++c) velocities[3 * idx + c] = guess[c];
    ImplicitMatrix A(*this, model, cols, dt);
    Vec<S> df;
    calculateForceDifferentials(df, vc, model);
    A.subspaceProject(df);
    for (size_t idx = 0; idx < masses.size(); ++idx) {
      for (int c = 0; c < 3; ++c) em[3 * idx + c] += dt * dt * df[3 * idx + c];
      if (!finite3(em[3 * idx], em[3 * idx + 1], em[3 * idx + 2])) throw std::runtime_error("nan in explicit momenta!");
    }
    if (rhsOut) *rhsOut = em;
    DiagonalPreconditioner pc(*this, model, dt);
    conjugateResiduals<S>(A, em, velocities, term, &pc, stats, nullptr, nullptr, debug);
    for (size_t i = 0; i < velocities.size(); ++i) velocities[i] += vc[i];
  }

  // Grid.cpp:1018-1047
  void updateDeformationGradients(S dt) {
    ShapeIt<S> it;
    for (size_t k = 0; k < inds.size(); ++k) {
      int p = inds[k];
      M3<S> delV;
      shapeInit(it, p, true);
      int g[3];
      do {
        it.gridPos(g);
        V3<S> wg = it.grad();
        int idx = coordsToIndex(g[0], g[1], g[2]);
        delV += outer(V3<S>(velocities[3 * idx], velocities[3 * idx + 1], velocities[3 * idx + 2]), wg);
      } while (it.next());
      d.F[p] = (M3<S>::identity() + dt * delV) * d.F[p];
    }
  }

  // Grid.cpp:1049-1081
  void updateParticleVelocities() {
    ShapeIt<S> it;
    const S alpha = S(0.95);
    for (size_t k = 0; k < inds.size(); ++k) {
      int p = inds[k];
      V3<S> vFlip = d.v[p], vPic;
      shapeInit(it, p, false);
      int g[3];
      do {
        it.gridPos(g);
        int idx = coordsToIndex(g[0], g[1], g[2]);
        S w = it.wt();
        V3<S> vi(velocities[3 * idx], velocities[3 * idx + 1], velocities[3 * idx + 2]);
        V3<S> vo(prevVelocities[3 * idx], prevVelocities[3 * idx + 1], prevVelocities[3 * idx + 2]);
        vFlip += w * (vi - vo);
        vPic += w * (vi + frameV);
      } while (it.next());
      d.v[p] = alpha * vFlip + (S(1) - alpha) * vPic;
    }
  }
};

// ---------------------------------------------------------------- KD-tree
// reference include/MpmSim/KdTree.h, KdTree.inl:89-216 (median split on the
// widest axis via nth_element, leaf <= 4, recursive radius query d^2 < r^2,
// root at index 1, children 2i and 2i+1).
template <class S> struct KDTree {
  struct Node {
    unsigned char axisOrLeaf = 255;
    S cut = 0;
    size_t first = 0, last = 0;  // offsets into perm for leaves
  };
  const std::vector<V3<S>>* pts = nullptr;
  std::vector<int> perm;
  std::vector<Node> nodes;
  int maxLeaf = 4;
  void init(const std::vector<V3<S>>& x, int maxLeafSize = 4) {
    pts = &x;
    maxLeaf = maxLeafSize;
    perm.resize(x.size());
    for (size_t i = 0; i < x.size(); ++i) perm[i] = int(i);
    nodes.clear();
    build(1, 0, perm.size());
  }
  unsigned char majorAxis(size_t f, size_t l) const {
    V3<S> mn(S(1e10), S(1e10), S(1e10)), mx(S(-1e10), S(-1e10), S(-1e10));
    for (size_t i = f; i < l; ++i)
      for (int a = 0; a < 3; ++a) {
        S c = (*pts)[perm[i]][a];
        if (c < mn[a]) mn[a] = c;
        if (c > mx[a]) mx[a] = c;
      }
    unsigned char major = 0;
    V3<S> size = mx - mn;
    for (unsigned char a = 1; a < 3; ++a)
      if (size[a] > size[major]) major = a;
    return major;
  }
  void build(size_t ni, size_t f, size_t l) {
    if (ni >= nodes.size()) nodes.resize(ni + 1);
    if (l - f > size_t(maxLeaf)) {
      unsigned char ax = majorAxis(f, l);
      size_t mid = f + (l - f) / 2;
      const std::vector<V3<S>>& P = *pts;
      std::nth_element(perm.begin() + f, perm.begin() + mid, perm.begin() + l,
                       [&](int a, int b) { return P[a][ax] < P[b][ax]; });
      S cut = P[perm[mid]][ax];
      nodes[ni].axisOrLeaf = ax;
      nodes[ni].cut = cut;
      build(2 * ni, f, mid);
      build(2 * ni + 1, mid, l);
    } else {
      nodes[ni].axisOrLeaf = 255;
      nodes[ni].first = f;
      nodes[ni].last = l;
    }
  }
  void query(const V3<S>& p, S r, std::vector<int>& out) const {
    out.clear();
    walk(1, p, r * r, out);
  }
  void walk(size_t ni, const V3<S>& p, S r2, std::vector<int>& out) const {
    const Node& nd = nodes[ni];
    if (nd.axisOrLeaf == 255) {
      for (size_t i = nd.first; i < nd.last; ++i) {
        const V3<S>& q = (*pts)[perm[i]];
        S d2 = sqnorm(p - q);
        if (d2 < r2) out.push_back(perm[i]);
      }
    } else {
      S dd = p[nd.axisOrLeaf] - nd.cut;
      size_t a, b;
      if (dd > 0) { a = 2 * ni + 1; b = 2 * ni; } else { a = 2 * ni; b = 2 * ni + 1; }
      walk(a, p, r2, out);
      if (dd * dd < r2) walk(b, p, r2, out);
    }
  }
};

// ------------------------------------------------------------------- Sim
template <class S> struct BodyStats {
  SolveStats<S> solve;
  V3<S> frameVelocity;
  int n[3] = {0, 0, 0};
  size_t particles = 0;
};

template <class S> struct Sim {
  Particles<S> d;
  IndexList ballistic;
  std::vector<IndexList> bodies;
  S h;
  ConstitutiveModel<S>& model;
  const CollisionSet<S>& cols;
  const ForceSet<S>& fields;
  int dim;
  std::vector<BodyStats<S>> lastStats;

  // Sim.cpp:17-50
  Sim(const std::vector<V3<S>>& x, const std::vector<S>& masses, S gridSize, ConstitutiveModel<S>& m,
      const CollisionSet<S>& c, const ForceSet<S>& f, int dimension = 3,
      const std::vector<S>* mu0 = nullptr, const std::vector<S>* lambda0 = nullptr, const std::vector<S>* xi = nullptr)
      : h(gridSize), model(m), cols(c), fields(f), dim(dimension) {
    d.p = x;
    d.v.assign(x.size(), V3<S>());
    d.F.assign(x.size(), M3<S>::identity());
    d.m = masses;
    d.volume.assign(x.size(), S(0));
    if (mu0) { d.mu0 = *mu0; d.lambda0 = *lambda0; d.xi = *xi; }
    model.setParticles(d);
    calculateBodies();
    for (size_t b = 0; b < bodies.size(); ++b) {
      Grid<S> g(d, bodies[b], h, V3<S>(), dim);
      g.computeParticleVolumes();
    }
  }

  // Sim.cpp:57-134
  void advance(S dt, Termination<S>& term) {
    lastStats.clear();
    for (size_t k = 0; k < ballistic.size(); ++k) {
      int p = ballistic[k];
      V3<S> accn;
      fields.force(accn, d.p[p], d.m[p]);
      accn = accn / d.m[p];
      d.v[p] += dt * accn;
    }
    for (size_t b = 0; b < bodies.size(); ++b) {
      V3<S> cmv;
      S mass = 0;
      for (size_t k = 0; k < bodies[b].size(); ++k) {
        int p = bodies[b][k];
        cmv += d.v[p] * d.m[p];
        mass += d.m[p];
      }
      cmv = cmv / mass;
      Grid<S> g(d, bodies[b], h, cmv, dim);
      BodyStats<S> bs;
      bs.frameVelocity = cmv;
      bs.particles = bodies[b].size();
      for (int a = 0; a < 3; ++a) bs.n[a] = g.n[a];
      g.updateGridVelocities(dt, model, cols, fields, term, &bs.solve);
      bs.solve.residualNorm2 = term.history;
      lastStats.push_back(bs);
      if (term.pollCancelled()) { calculateBodies(); return; }
      g.updateParticleVelocities();
      g.updateDeformationGradients(dt);
      model.updateParticleData();
    }
    for (size_t p = 0; p < d.p.size(); ++p) {
      cols.collide(d.v[p], d.p[p], V3<S>(), true);
      d.p[p] += d.v[p] * dt;
    }
    calculateBodies();
  }

  // Sim.cpp:136-203
  void calculateBodies() {
    bodies.clear();
    ballistic.clear();
    KDTree<S> tree;
    tree.init(d.p);
    std::vector<char> processed(d.p.size(), 0);
    std::vector<int> nb;
    for (size_t i = 0; i < d.p.size(); ++i) {
      if (!finite3(d.v[i][0], d.v[i][1], d.v[i][2])) throw std::runtime_error("nans in particle velocity data!");
      if (processed[i]) continue;
      tree.query(d.p[i], h, nb);
      if (nb.size() == 1) {
        processed[i] = 1;
        ballistic.push_back(int(i));
        continue;
      }
      bodies.resize(bodies.size() + 1);
      IndexList& b = bodies.back();
      std::queue<int> flood;
      for (size_t j = 0; j < nb.size(); ++j) flood.push(nb[j]);
      while (!flood.empty()) {
        int cur = flood.front();
        flood.pop();
        if (processed[cur]) continue;
        b.push_back(cur);
        processed[cur] = 1;
        tree.query(d.p[cur], h, nb);
        for (size_t j = 0; j < nb.size(); ++j) flood.push(nb[j]);
      }
    }
  }
};

}  // namespace mpmo
