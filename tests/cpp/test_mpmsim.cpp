// Tests of the MpmSim C++ class layer (mpmsnow_amd/cpp/MpmSim.hpp) over the GPU library,
// written the way the reference's own unit tests use those classes (src/tests/TestSimClass.cpp,
// TestGrid.cpp, TestSnowConstitutiveModel.cpp).
//
//   test_mpmsim --host   checks that need no GPU (containers, shape function, termination, colliders)
//   test_mpmsim          all checks; the GPU ones run on device 0
//
// Prints "ok <test>" per passing test; exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "MpmSim.hpp"

using namespace MpmSim;

static void fail(const char* expr, const char* file, int line) {
  std::fprintf(stderr, "FAILED %s at %s:%d\n", expr, file, line);
  std::exit(1);
}
#define CHECK(e) ((e) ? (void)0 : fail(#e, __FILE__, __LINE__))

template <class F>
static bool throws(F f, const char* substring = nullptr) {
  try {
    f();
  } catch (const std::runtime_error& e) {
    return !substring || std::strstr(e.what(), substring) != nullptr;
  }
  return false;
}

// ---------------------------------------------------------------- host-only tests

static void testMaterialPointData() {
  MaterialPointData d;
  CHECK(d.numVariables() == 5);
  d.variable<float>("m");
  d.variable<float>("volume");
  d.variable<Vector3f>("v");
  d.variable<Vector3f>("p");
  d.variable<Matrix3f>("F");
  CHECK(throws([&] { d.variable<float>("Fiddlesticks"); }, "no variable named 'Fiddlesticks'"));
  CHECK(throws([&] { d.variable<float>("p"); }, "of the specified type"));
  CHECK(throws([&] { d.createVariable<float>("m"); }, "already has a variable named 'm'"));
  d.createVariable<float>("temperature");
  CHECK(d.numVariables() == 6);
  const MaterialPointData& cd = d;
  CHECK(cd.variable<float>("temperature").empty());
}

static void testShapeFunction() {
  CubicBsplineShapeFunction sf;
  CHECK(sf.supportRadius() == 2);
  // partition of unity and of the first moment, the derivative sums to zero
  for (int s = 0; s <= 20; ++s) {
    const float x = 0.05f * s;
    double w = 0, xw = 0, dw = 0;
    for (int j = -2; j <= 2; ++j) {
      w += sf.w(x - j);
      xw += (double)j * sf.w(x - j);
      dw += sf.dw(x - j);
    }
    CHECK(std::fabs(w - 1.0) < 1e-6);
    CHECK(std::fabs(xw - x) < 1e-6);
    CHECK(std::fabs(dw) < 1e-6);
  }
  // dw is the derivative of w
  for (float x = -2.3f; x < 2.3f; x += 0.0731f) {
    const float eps = 1e-3f;
    const float fd = (sf.w(x + eps) - sf.w(x - eps)) / (2 * eps);
    CHECK(std::fabs(fd - sf.dw(x)) < 2e-3f);
  }
  CHECK(std::fabs(sf.w(0.0f) - 2.0f / 3.0f) < 1e-7f && sf.w(2.0f) == 0.0f);
}

static void testSquareMagnitudeTermination() {
  SquareMagnitudeTermination t(10, 0.1f);
  VectorXf b = {3.0f, 4.0f};  // |b|^2 = 25: threshold 0.25
  t.init(b);
  VectorXf r = {0.4f, 0.2f};  // 0.2
  CHECK(t(r, 0));
  r = {0.4f, 0.4f};  // 0.32
  CHECK(!t(r, 0));
  CHECK(t(r, 10));
  CHECK(!t.cancelled());
}

static void testCollisionObjects() {
  CollisionPlane plane(Vector4f(0, 1, 0, -0.5f), 0.3f, true);
  CHECK(std::fabs(plane.phi(Vector3f(1, 2, 3)) - 1.5f) < 1e-7f);
  Vector3f g;
  plane.grad(Vector3f(0, 0, 0), g);
  CHECK(g[0] == 0 && g[1] == 1 && g[2] == 0);
  plane.setV(Vector3f(1, 0, 0));
  Vector3f v;
  plane.velocity(Vector3f(), v);
  CHECK(v[0] == 1 && plane.sticky() && plane.coulombFriction() == 0.3f);

  // a level set sampled from a linear function reproduces it exactly inside the lattice
  const int dims[3] = {5, 6, 7};
  std::vector<float> phi(5 * 6 * 7);
  for (int k = 0; k < 7; ++k)
    for (int j = 0; j < 6; ++j)
      for (int i = 0; i < 5; ++i) phi[i + 5 * (j + 6 * k)] = 0.25f * (0.5f * j - 1.0f);
  LevelSetCollisionObject ls(dims, Vector3f(0, 0, 0), 0.5f, phi);
  CHECK(std::fabs(ls.phi(Vector3f(1.1f, 1.3f, 2.2f)) - 0.25f * (1.3f - 1.0f)) < 1e-6f);
  ls.grad(Vector3f(1.1f, 1.3f, 2.2f), g);
  CHECK(std::fabs(g[0]) < 1e-6f && std::fabs(g[1] - 0.25f) < 1e-5f && std::fabs(g[2]) < 1e-6f);
  CHECK(throws([&] { LevelSetCollisionObject bad(dims, Vector3f(), 0.5f, std::vector<float>(3)); }));

  CollisionObject::CollisionObjectSet set;
  set.add(new CollisionPlane(Vector4f(0, 1, 0, 0)));
  set.add(new LevelSetCollisionObject(dims, Vector3f(), 0.5f, phi));
  std::vector<mpm_collider> cs = set.gpuColliders();
  CHECK(cs.size() == 2 && cs[0].type == 0 && cs[1].type == 2 && cs[1].dims[2] == 7 && cs[1].phi != nullptr);
}

// ---------------------------------------------------------------- GPU tests

// a model the device cannot run (TestSimClass.cpp:21-40 builds one the same way)
class DummyModel : public ConstitutiveModel {
 public:
  void setParticles(MaterialPointData&) override {}
  void updateParticleData() override {}
  float energyDensity(size_t) const override { return 0; }
  Matrix3f dEnergyDensitydF(size_t) const override { return Matrix3f::Zero(); }
  Matrix3f dEdFDifferential(const Matrix3f&, size_t) const override { return Matrix3f::Zero(); }
};

static void twoCubesAndBallistics(int side, std::vector<Vector3f>& positions, std::vector<float>& masses,
                                  float gridSize) {
  for (int i = 0; i < side; ++i)
    for (int j = 0; j < side; ++j)
      for (int k = 0; k < side; ++k) {
        positions.push_back(Vector3f(0.5f * gridSize * (i + 0.5f), 0.5f * gridSize * (j + 0.5f), 0.5f * gridSize * (k + 0.5f)));
        masses.push_back(1.0f);
        positions.push_back(Vector3f(0.5f * gridSize * (i + 0.5f) + 6, 0.5f * gridSize * (j + 0.5f), 0.5f * gridSize * (k + 0.5f)));
        masses.push_back(1.0f);
      }
  for (int i = 0; i < side; ++i)
    for (int j = 0; j < side; ++j)
      for (int k = 0; k < side; ++k) {
        positions.push_back(Vector3f(1.1f * gridSize * i + 20, 1.1f * gridSize * j, 1.1f * gridSize * k));
        masses.push_back(1.0f);
      }
}

// TestSimClass::testInitialization (TestSimClass.cpp:44-118)
static void testInitialization() {
  std::vector<Vector3f> positions;
  std::vector<float> masses;
  const float gridSize = 0.1f;
  twoCubesAndBallistics(16, positions, masses, gridSize);
  CubicBsplineShapeFunction shapeFunction;
  DummyModel constitutiveModel;
  CollisionObject::CollisionObjectSet collisionObjects;
  ForceField::ForceFieldSet forceFields;
  Sim sim(positions, masses, gridSize, shapeFunction, constitutiveModel, collisionObjects, forceFields);
  CHECK(sim.particleData().numVariables() == 5);
  sim.particleData().variable<float>("m");
  sim.particleData().variable<float>("volume");
  sim.particleData().variable<Vector3f>("v");
  sim.particleData().variable<Vector3f>("p");
  sim.particleData().variable<Matrix3f>("F");
  CHECK(throws([&] { sim.particleData().variable<float>("Fiddlesticks"); }));
  CHECK(sim.ballisticParticles().size() == 16 * 16 * 16);
  CHECK(sim.bodies().size() == 2);
  CHECK(sim.bodies()[0].size() == 16 * 16 * 16);
  CHECK(sim.bodies()[1].size() == 16 * 16 * 16);
  // particle volumes were computed for the body particles only
  const MaterialPointData& cd = sim.particleData();
  const std::vector<float>& vol = cd.variable<float>("volume");
  CHECK(vol[0] > 0.0f && vol[sim.ballisticParticles()[0]] == 0.0f);
  // the device cannot run this model
  SquareMagnitudeTermination t(10, 0.0f);
  CHECK(throws([&] { sim.advance(0.01f, t); }, "no GPU implementation"));
}

// TestSimClass::testTimestepAdvance (TestSimClass.cpp:120-190)
static void testTimestepAdvance() {
  std::vector<Vector3f> positions;
  std::vector<float> masses;
  const float gridSize = 0.1f;
  twoCubesAndBallistics(4, positions, masses, gridSize);
  CubicBsplineShapeFunction shapeFunction;
  SnowConstitutiveModel constitutiveModel(1.4e5f, 0.2f, 0, 100000.0f, 100000.0f);
  CollisionObject::CollisionObjectSet collisionObjects;
  ForceField::ForceFieldSet forceFields;
  forceFields.add(new GravityField(Vector3f(0, 1.f, 0)));
  Sim sim(positions, masses, gridSize, shapeFunction, constitutiveModel, collisionObjects, forceFields);
  CHECK(sim.particleData().numVariables() == 12);
  SquareMagnitudeTermination t(10, 0.0f);
  sim.advance(0.01f, t);
  sim.advance(0.01f, t);
  sim.advance(0.01f, t);
  const std::vector<Vector3f>& velocities = sim.particleData().variable<Vector3f>("v");
  Vector3f v = Vector3f::Zero();
  for (const Vector3f& vi : velocities) v += vi;
  v /= (float)velocities.size();
  CHECK(std::fabs(v[0]) < 1.e-6);
  CHECK(std::fabs(v[2]) < 1.e-6);
  CHECK(std::fabs(v[1] - 0.03f) < 1.e-5);
}

// TestGrid::testSplatting (TestGrid.cpp:125-211): the grid's masses against a serial splat with
// the host shape function, and mass and momentum conservation
static void testSplatting() {
  MaterialPointData d;
  std::vector<Vector3f>& positions = d.variable<Vector3f>("p");
  std::vector<Vector3f>& velocities = d.variable<Vector3f>("v");
  std::vector<float>& masses = d.variable<float>("m");
  Sim::IndexList particleInds;
  std::srand(7);
  auto rnd = [] { return (float)std::rand() / (float)RAND_MAX; };
  positions.push_back(Vector3f::Zero());
  velocities.push_back(Vector3f(1 + 2 * rnd() - 1, 2 * rnd() - 1, 2 * rnd() - 1));
  masses.push_back(1.0f);
  particleInds.push_back(0);
  for (int i = 0; i < 500; ++i) {
    positions.push_back(Vector3f(rnd(), rnd(), rnd()));
    velocities.push_back(Vector3f(1 + 2 * rnd() - 1, 2 * rnd() - 1, 2 * rnd() - 1));
    masses.push_back(1.0f);
    particleInds.push_back((int)particleInds.size());
  }
  d.variable<Matrix3f>("F").assign(positions.size(), Matrix3f::Identity());
  d.variable<float>("volume").assign(positions.size(), 0.0f);
  CubicBsplineShapeFunction shapeFunction;
  const float gridSize = 0.01f;
  Grid g(d, particleInds, gridSize, shapeFunction);
  const VectorXf gm = g.masses();
  const VectorXf gv = g.velocities();
  const int* n = g.n();
  const std::vector<Vector3f>& px = static_cast<const MaterialPointData&>(d).variable<Vector3f>("p");
  const std::vector<Vector3f>& pv = static_cast<const MaterialPointData&>(d).variable<Vector3f>("v");
  const std::vector<float>& pm = static_cast<const MaterialPointData&>(d).variable<float>("m");
  std::vector<double> serial(gm.size(), 0.0);
  for (size_t p = 0; p < px.size(); ++p) {
    float u[3];
    int base[3];
    for (int a = 0; a < 3; ++a) {
      u[a] = (px[p][a] - g.min()[a]) / gridSize;
      base[a] = (int)std::floor(u[a]) - 1;
    }
    for (int k = 0; k < 4; ++k)
      for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
          const int c[3] = {base[0] + i, base[1] + j, base[2] + k};
          const double w = (double)shapeFunction.w(u[0] - c[0]) * shapeFunction.w(u[1] - c[1]) * shapeFunction.w(u[2] - c[2]);
          serial[c[0] + n[0] * (c[1] + n[1] * c[2])] += pm[p] * w;
        }
  }
  double maxDiff = 0, gridMass = 0, particleMass = 0;
  for (size_t i = 0; i < gm.size(); ++i) {
    maxDiff = std::fmax(maxDiff, std::fabs(serial[i] - gm[i]));
    gridMass += gm[i];
  }
  for (float m : pm) particleMass += m;
  CHECK(maxDiff < 1e-6);
  CHECK(std::fabs(particleMass - gridMass) / particleMass < 1e-4);
  double gp[3] = {0, 0, 0}, pp[3] = {0, 0, 0};
  for (size_t i = 0; i < gm.size(); ++i)
    for (int a = 0; a < 3; ++a) gp[a] += (double)gm[i] * gv[3 * i + a];
  for (size_t p = 0; p < pm.size(); ++p)
    for (int a = 0; a < 3; ++a) pp[a] += (double)pm[p] * pv[p][a];
  const double dn = std::sqrt((gp[0] - pp[0]) * (gp[0] - pp[0]) + (gp[1] - pp[1]) * (gp[1] - pp[1]) + (gp[2] - pp[2]) * (gp[2] - pp[2]));
  const double pn = std::sqrt(pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]);
  CHECK(dn / pn < 1e-4);
}

// TestSnowConstitutiveModel (TestSnowConstitutiveModel.cpp:54-92): zero energy and stress at a
// rotation, positive energy under stretch, dEdFDifferential linear in dF
static void testSnowModel() {
  MaterialPointData d;
  const int np = 3;
  d.variable<Vector3f>("p").assign(np, Vector3f());
  d.variable<float>("m").assign(np, 1.0f);
  Matrix3f rot = Matrix3f::Identity();
  const float c = std::cos(0.7f), s = std::sin(0.7f);
  rot(0, 0) = c, rot(0, 1) = -s, rot(1, 0) = s, rot(1, 1) = c;
  Matrix3f stretch = Matrix3f::Identity();
  stretch(0, 0) = 1.02f, stretch(1, 2) = 0.01f;
  std::vector<Matrix3f>& F = d.variable<Matrix3f>("F");
  F = {rot, stretch, Matrix3f::Identity()};
  SnowConstitutiveModel model(1.4e5f, 0.2f, 10.0f, 0.025f, 0.0075f);
  model.setParticles(d);
  CHECK(d.numVariables() == 12);
  model.updateParticleData();
  CHECK(std::fabs(model.energyDensity(0)) < 1e-2f);  // E0 = 1.4e5: relative 1e-7
  const Matrix3f P0 = model.dEnergyDensitydF(0);
  for (int i = 0; i < 9; ++i) CHECK(std::fabs(P0.m[i]) < 1.0f);
  CHECK(model.energyDensity(1) > 0.0f);
  Matrix3f dF;
  for (int i = 0; i < 9; ++i) dF.m[i] = 0.001f * (i - 4);
  Matrix3f dF2 = dF;
  for (int i = 0; i < 9; ++i) dF2.m[i] *= 2.0f;
  const Matrix3f a = model.dEdFDifferential(dF, 1), b = model.dEdFDifferential(dF2, 1);
  double amax = 0;
  for (int i = 0; i < 9; ++i) amax = std::fmax(amax, std::fabs(a.m[i]));
  CHECK(amax > 0);
  for (int i = 0; i < 9; ++i) CHECK(std::fabs(b.m[i] - 2 * a.m[i]) < 1e-4 * amax + 1e-3);
  // the plasticity clamp ran on the device: the stretched particle's Fp moved off identity
  const std::vector<Matrix3f>& Fp = static_cast<const MaterialPointData&>(d).variable<Matrix3f>("Fp");
  CHECK(std::fabs(Fp[1](0, 0) - 1.0f) > 1e-4f);
}

// host hooks: a custom termination criterion and a LinearSolver::Debug, through a Grid's solve
class StopAfter : public TerminationCriterion {
 public:
  explicit StopAfter(int n) : m_n(n) {}
  bool operator()(VectorXf& r, int i) const override {
    ++calls;
    lastSize = r.size();
    return i >= m_n;
  }
  bool cancelled() const override { return false; }
  mutable int calls = 0;
  mutable size_t lastSize = 0;

 private:
  int m_n;
};
class CountIterates : public LinearSolver::Debug {
 public:
  void operator()(VectorXf& x) override {
    ++count;
    size = x.size();
  }
  int count = 0;
  size_t size = 0;
};

static void snowball(std::vector<Vector3f>& x, std::vector<float>& m, float gridSize, float y0) {
  const float r = 0.25f, sp = 0.5f * gridSize;
  for (float i = -r; i <= r; i += sp)
    for (float j = -r; j <= r; j += sp)
      for (float k = -r; k <= r; k += sp)
        if (i * i + j * j + k * k <= r * r) {
          x.push_back(Vector3f(i, y0 + j, k));
          m.push_back(0.05f);
        }
}

static void testSolverHooks() {
  MaterialPointData d;
  std::vector<Vector3f> x;
  std::vector<float> m;
  const float gridSize = 0.05f;
  snowball(x, m, gridSize, 0.0f);
  d.variable<Vector3f>("p") = x;
  d.variable<float>("m") = m;
  d.variable<Vector3f>("v").assign(x.size(), Vector3f(0, -1, 0));
  d.variable<Matrix3f>("F").assign(x.size(), Matrix3f::Identity());
  d.variable<float>("volume").assign(x.size(), 0.0f);
  SnowConstitutiveModel model(1.4e5f, 0.2f, 10.0f, 0.025f, 0.0075f);
  model.setParticles(d);
  CubicBsplineShapeFunction sf;
  Sim::IndexList inds(x.size());
  for (size_t i = 0; i < inds.size(); ++i) inds[i] = (int)i;
  CollisionObject::CollisionObjectSet cols;
  cols.add(new CollisionPlane(Vector4f(0, 1, 0, 0.2f)));
  ForceField::ForceFieldSet fields;
  fields.add(new GravityField(Vector3f(0, -9.8f, 0)));
  {
    Grid g0(d, inds, gridSize, sf);
    g0.computeParticleVolumes();
  }
  Grid g(d, inds, gridSize, sf);
  StopAfter term(3);
  CountIterates dbg;
  g.updateGridVelocities(0.004f, model, cols, fields, term, &dbg);
  CHECK(g.solveStats().iterations == 4);  // x updates at i = 0..3 (ConjugateResiduals.cpp:84-128)
  CHECK(dbg.count == 4 && dbg.size == 3 * g.nodes());
  CHECK(term.calls >= 3 && term.lastSize == 3 * g.nodes());
  // the plane cuts the ball: some nodes are flagged and carry the plane's (zero) velocity
  VectorXf em;
  std::vector<char> collided;
  g.calculateExplicitMomenta(em, collided, 0.004f, model, cols, fields);
  int flagged = 0;
  for (char f : collided) flagged += f == 0;
  CHECK(flagged > 0);
  VectorXf vc;
  g.collisionVelocities(vc, cols, collided);
  for (float v : vc) CHECK(v == 0.0f);
  // forces: gravity plus elastic forces of an undeformed body = m g per node
  VectorXf f;
  g.calculateForces(f, model, fields);
  const VectorXf gm = g.masses();
  double maxDev = 0;
  for (size_t i = 0; i < gm.size(); ++i) maxDev = std::fmax(maxDev, std::fabs(f[3 * i + 1] - gm[i] * -9.8f));
  CHECK(maxDev < 1e-3);
  // the diagonal of dF/dx is at most zero (the elastic energy is convex near the rest state)
  VectorXf dfdx;
  g.dForceidXi(dfdx, model);
  for (float v : dfdx) CHECK(v <= 1e-3f);
  g.updateParticleVelocities();
  g.updateDeformationGradients(0.004f);
  const std::vector<Vector3f>& v = static_cast<const MaterialPointData&>(d).variable<Vector3f>("v");
  CHECK(v.size() == x.size());
}

// a snowball dropped on a plane through Sim, with a level set floor beside it
static void testSnowballOnGround() {
  std::vector<Vector3f> x;
  std::vector<float> m;
  const float gridSize = 0.05f;
  snowball(x, m, gridSize, 0.3f);
  CubicBsplineShapeFunction sf;
  SnowConstitutiveModel model(1.4e5f, 0.2f, 10.0f, 0.025f, 0.0075f);
  CollisionObject::CollisionObjectSet cols;
  cols.add(new CollisionPlane(Vector4f(0, 1, 0, 0.0f)));
  ForceField::ForceFieldSet fields;
  fields.add(new GravityField(Vector3f(0, -9.8f, 0)));
  Sim sim(x, m, gridSize, sf, model, cols, fields);
  CHECK(sim.bodies().size() == 1 && sim.ballisticParticles().empty());
  SquareMagnitudeTermination t(60, 1e-2f);
  for (int s = 0; s < 60; ++s) sim.advance(0.004f, t);
  const MaterialPointData& cd = sim.particleData();
  float ymin = 1e9f;
  for (const Vector3f& p : cd.variable<Vector3f>("p")) ymin = std::fmin(ymin, p[1]);
  CHECK(ymin > -0.5f * gridSize);  // resting on the plane, not through it
  CHECK(ymin < 0.1f);             // it fell
  CHECK(sim.lastStepStats().total_iterations > 0);
  // a host edit of the velocities reaches the device on the next step
  std::vector<Vector3f>& v = sim.particleData().variable<Vector3f>("v");
  for (Vector3f& vi : v) vi = Vector3f(0, 5, 0);
  sim.advance(0.004f, t);
  const std::vector<Vector3f>& v2 = cd.variable<Vector3f>("v");
  double mean = 0;
  for (const Vector3f& vi : v2) mean += vi[1];
  CHECK(mean / v2.size() > 4.0);
}

// ---------------------------------------------------------------- ConjugateResiduals

// a dense symmetric matrix and its inverse (TestConjugateResiduals.cpp:14-38), row-major
class DenseMatrix : public ProceduralMatrix {
 public:
  DenseMatrix(const std::vector<double>& a, int n) : m_a(a), m_n(n) {}
  void multVector(const VectorXf& x, VectorXf& result) const override {
    result.assign(m_n, 0.0f);
    for (int i = 0; i < m_n; ++i) {
      double s = 0;
      for (int j = 0; j < m_n; ++j) s += m_a[i * m_n + j] * x[j];
      result[i] = (float)s;
    }
  }
  // no preconditioner in the test: the inverse is never asked for
  void multInverseVector(const VectorXf& x, VectorXf& result) const override { result = x; }
  void subspaceProject(VectorXf&) const override {}

 private:
  std::vector<double> m_a;
  int m_n;
};

static double dot(const VectorXf& a, const VectorXf& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += (double)a[i] * b[i];
  return s;
}

// TestConjugateResiduals::testSolve (TestConjugateResiduals.cpp:43-102): a random symmetric
// indefinite 6x6 system, solved in 6 iterations with logging; residual norms decrease, residuals
// are A-orthogonal and search directions A^2-orthogonal, with the reference's bounds
static void testConjugateResidualsSolve() {
  const int n = 6;
  unsigned seed = 12345;
  auto rnd = [&]() {  // uniform in [-1, 1], as Eigen's Random
    seed = seed * 1664525u + 1013904223u;
    return (float)((seed >> 8) * (1.0 / 16777216.0) * 2.0 - 1.0);
  };
  std::vector<float> r0(n * n);
  for (float& v : r0) v = rnd();
  std::vector<double> a(n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) a[i * n + j] = (double)(r0[i * n + j] + r0[j * n + i]);
  DenseMatrix A(a, n);
  SquareMagnitudeTermination t(n, 0.0f);
  ConjugateResiduals solver(t, 0, true);
  VectorXf v(n), result(n, 0.0f);
  for (float& x : v) x = rnd();
  solver(A, v, result);
  VectorXf Ax;
  A.multVector(result, Ax);
  double err = 0;
  for (int i = 0; i < n; ++i) err += (double)(Ax[i] - v[i]) * (Ax[i] - v[i]);
  std::printf("  |A x - v| = %.3e after %d iterations, %zu logged\n", std::sqrt(err), solver.iterations(),
              solver.residuals().size());
  CHECK(std::sqrt(err) < 1e-3);
  const std::vector<VectorXf>& R = solver.residuals();
  const std::vector<VectorXf>& P = solver.searchDirections();
  CHECK((int)R.size() >= n && (int)P.size() >= n);
  for (int i = 1; i < n; ++i) CHECK(dot(R[i], R[i]) < dot(R[i - 1], R[i - 1]));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      if (i == j) continue;
      VectorXf ARj, APi, APj;
      A.multVector(R[j], ARj);
      CHECK(std::fabs(dot(R[i], ARj)) < 1e-4);
      A.multVector(P[i], APi);
      A.multVector(P[j], APj);
      CHECK(std::fabs(dot(APi, APj)) < 1e-4);
    }
}

int main(int argc, char** argv) {
  const bool hostOnly = argc > 1 && std::string(argv[1]) == "--host";
  struct T {
    const char* name;
    void (*fn)();
    bool gpu;
  } tests[] = {
      {"testMaterialPointData", testMaterialPointData, false},
      {"testShapeFunction", testShapeFunction, false},
      {"testSquareMagnitudeTermination", testSquareMagnitudeTermination, false},
      {"testCollisionObjects", testCollisionObjects, false},
      {"testInitialization", testInitialization, true},
      {"testTimestepAdvance", testTimestepAdvance, true},
      {"testSplatting", testSplatting, true},
      {"testSnowModel", testSnowModel, true},
      {"testSolverHooks", testSolverHooks, true},
      {"testSnowballOnGround", testSnowballOnGround, true},
      {"testConjugateResidualsSolve", testConjugateResidualsSolve, true},
  };
  for (const T& t : tests) {
    if (t.gpu && hostOnly) continue;
    std::printf("run %s\n", t.name);
    std::fflush(stdout);
    try {
      t.fn();
    } catch (const std::exception& e) {
      std::fprintf(stderr, "FAILED %s: exception %s\n", t.name, e.what());
      return 1;
    }
    std::printf("ok %s\n", t.name);
    std::fflush(stdout);
  }
  return 0;
}
