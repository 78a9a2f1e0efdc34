// MpmSim.hpp -- the MpmSim C++ class interface over the MI355X C-ABI (mpmsnow_gpu.h).
//
// Header-only host layer with the reference library's class names, constructor arguments,
// member functions and exception texts (reference include/MpmSim/*.h), so that code written
// against MpmSim::Sim / Grid / MaterialPointData / SnowConstitutiveModel / CollisionPlane /
// GravityField / SquareMagnitudeTermination compiles against it with its includes changed.
// Every piece of simulation work runs in the GPU library; this file holds containers,
// argument plumbing and the host<->device transfers only.
//
// Differences a caller sees:
//   * Eigen is not required: Vector3f, Vector4f and Matrix3f below have Eigen's memory layout
//     (3, 4 and 9 contiguous floats, matrices column-major) and the members the reference's
//     callers use; VectorXf is std::vector<float>.
//   * MaterialPointData mirrors its variables on the device. A non-const variable<T>() access
//     schedules an upload of every device-known variable before the next GPU call; reads
//     through a const MaterialPointData& do not. Vectors handed out stay valid (they are
//     refreshed in place, never reallocated while the particle count is unchanged).
//   * Only the models the GPU implements run on it: SnowConstitutiveModel, GravityField,
//     CollisionPlane, CollisionSphere and LevelSetCollisionObject. Sim::advance throws for
//     any other ConstitutiveModel, ForceField or CollisionObject.
//   * A TerminationCriterion that is exactly SquareMagnitudeTermination is evaluated on the
//     device. Any other one, and any LinearSolver::Debug, is called from the host once per
//     CR iteration (the solve then synchronises every iteration); init() is not called.
//   * Sim exposes its partition through bodies() and ballisticParticles() (the reference
//     keeps them private and lets its tests read them as a friend class).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <typeinfo>
#include <vector>

#include "mpmsnow_gpu.h"

namespace MpmSim {

// ---- small fixed-size types with Eigen's layout ------------------------------------------
struct Vector3f {
  float v[3];
  Vector3f() : v{0.0f, 0.0f, 0.0f} {}
  Vector3f(float x, float y, float z) : v{x, y, z} {}
  static Vector3f Zero() { return Vector3f(); }
  float& operator[](int i) { return v[i]; }
  float operator[](int i) const { return v[i]; }
  float& operator()(int i) { return v[i]; }
  float operator()(int i) const { return v[i]; }
  float* data() { return v; }
  const float* data() const { return v; }
  Vector3f& operator+=(const Vector3f& o) { for (int i = 0; i < 3; ++i) v[i] += o.v[i]; return *this; }
  Vector3f& operator-=(const Vector3f& o) { for (int i = 0; i < 3; ++i) v[i] -= o.v[i]; return *this; }
  Vector3f& operator*=(float s) { for (int i = 0; i < 3; ++i) v[i] *= s; return *this; }
  Vector3f& operator/=(float s) { for (int i = 0; i < 3; ++i) v[i] /= s; return *this; }
  Vector3f operator+(const Vector3f& o) const { Vector3f r(*this); return r += o; }
  Vector3f operator-(const Vector3f& o) const { Vector3f r(*this); return r -= o; }
  Vector3f operator*(float s) const { Vector3f r(*this); return r *= s; }
  float dot(const Vector3f& o) const { return v[0] * o.v[0] + v[1] * o.v[1] + v[2] * o.v[2]; }
  float squaredNorm() const { return dot(*this); }
  float norm() const { return std::sqrt(squaredNorm()); }
};

struct Vector4f {
  float v[4];
  Vector4f() : v{0.0f, 0.0f, 0.0f, 0.0f} {}
  Vector4f(float a, float b, float c, float d) : v{a, b, c, d} {}
  float& operator[](int i) { return v[i]; }
  float operator[](int i) const { return v[i]; }
  const float* data() const { return v; }
};

struct Matrix3f {
  float m[9];  // column-major: (r, c) at m[r + 3 c]
  Matrix3f() : m{} {}
  static Matrix3f Zero() { return Matrix3f(); }
  static Matrix3f Identity() {
    Matrix3f a;
    a.m[0] = a.m[4] = a.m[8] = 1.0f;
    return a;
  }
  float& operator()(int r, int c) { return m[r + 3 * c]; }
  float operator()(int r, int c) const { return m[r + 3 * c]; }
  float* data() { return m; }
  const float* data() const { return m; }
};

static_assert(sizeof(Vector3f) == 12 && sizeof(Matrix3f) == 36, "layout must match the C-ABI");

typedef std::vector<float> VectorXf;

namespace detail {
inline void check(const mpm_context* ctx, int rc) {
  if (rc != MPM_OK) throw std::runtime_error(mpm_last_error(ctx));
}
template <class T> struct Comps;
template <> struct Comps<float> { static const int n = 1; };
template <> struct Comps<Vector3f> { static const int n = 3; };
template <> struct Comps<Matrix3f> { static const int n = 9; };
// variables the device holds (mpm_particles_set/get); the second group exists once a model is bound
inline int device_variable(const std::string& name) {
  static const char* base[] = {"p", "v", "F", "m", "volume"};
  static const char* model[] = {"Fp", "FinvTrans", "R", "Ginv", "J", "mu", "lambda", "mu0", "lambda0", "xi"};
  for (const char* b : base)
    if (name == b) return 1;
  for (const char* b : model)
    if (name == b) return 2;
  return 0;
}
}  // namespace detail

// ---- MaterialPointData (MaterialPointData.h:14-44, MaterialPointData.inl) ------------------
class MaterialPointData {
 public:
  // default variables p, v, F, m, volume, created empty (MaterialPointData.cpp:8-16)
  MaterialPointData() {
    createVariable<Vector3f>("p");
    createVariable<Vector3f>("v");
    createVariable<Matrix3f>("F");
    createVariable<float>("m");
    createVariable<float>("volume");
  }
  ~MaterialPointData() {
    if (m_ctx) mpm_destroy(m_ctx);
  }
  MaterialPointData(const MaterialPointData&) = delete;
  MaterialPointData& operator=(const MaterialPointData&) = delete;

  template <typename T>
  void createVariable(const std::string& name) {
    if (m_vars.count(name)) throw std::runtime_error("Material point data already has a variable named '" + name + "'");
    m_vars[name].reset(new Var<T>());
  }
  template <typename T>
  std::vector<T>& variable(const std::string& name) {
    pull();
    m_hostDirty = true;
    return find<T>(name);
  }
  template <typename T>
  const std::vector<T>& variable(const std::string& name) const {
    const_cast<MaterialPointData*>(this)->pull();
    return const_cast<MaterialPointData*>(this)->find<T>(name);
  }
  size_t numVariables() const { return m_vars.size(); }
  std::vector<std::string> variableNames() const {
    std::vector<std::string> names;
    for (const auto& kv : m_vars) names.push_back(kv.first);
    return names;
  }

  // ---- device mirror (no reference counterpart) ----
  // The context holding these particles, created on first use; uploads host changes.
  mpm_context* device() {
    if (!m_ctx && mpm_create(&m_ctx, 0) != MPM_OK) {
      m_ctx = nullptr;
      throw std::runtime_error("mpm_create failed: no usable HIP device");
    }
    const int64_t n = (int64_t)find<float>("m").size();
    if (n == 0) throw std::runtime_error("material point data has no particles ('m' is empty)");
    if (n != m_deviceCount) {
      detail::check(m_ctx, mpm_particles_create(m_ctx, n));
      m_deviceCount = n;
      m_modelBound = false;
      m_hostDirty = true;
    }
    if (m_hasModel && !m_modelBound) {
      detail::check(m_ctx, mpm_model_snow(m_ctx, m_model[0], m_model[1], m_model[2], m_model[3], m_model[4]));
      m_modelBound = true;
      m_hostDirty = true;
    }
    if (m_hostDirty) {
      for (int group = 1; group <= 2; ++group)
        for (auto& kv : m_vars) {
          if (detail::device_variable(kv.first) != group || (group == 2 && !m_modelBound)) continue;
          const int64_t count = kv.second->floats();
          if (count == 0) continue;  // never filled: the device keeps its default
          if (count != n * kv.second->comps())
            throw std::runtime_error("Material point data variable '" + kv.first + "' has the wrong size");
          detail::check(m_ctx, mpm_particles_set(m_ctx, kv.first.c_str(), kv.second->ptr(), count));
        }
      m_hostDirty = false;
    }
    return m_ctx;
  }
  // the device holds newer values than the host vectors (after any GPU step on them)
  void markDeviceNewer() { m_deviceNewer = true; }
  // SnowConstitutiveModel::setParticles records its parameters here; the device model is
  // created with them on the next device() call
  void bindSnowModel(const float params[5]) {
    std::memcpy(m_model, params, sizeof(m_model));
    m_hasModel = true;
    m_modelBound = false;
  }
  bool hasGpuModel() const { return m_hasModel; }

 private:
  struct VarBase {
    virtual ~VarBase() {}
    virtual int comps() const = 0;
    virtual int64_t floats() const = 0;
    virtual float* ptr() = 0;
    virtual void resize(size_t n) = 0;
  };
  template <typename T>
  struct Var : VarBase {
    std::vector<T> data;
    int comps() const override { return detail::Comps<T>::n; }
    int64_t floats() const override { return (int64_t)data.size() * detail::Comps<T>::n; }
    float* ptr() override { return reinterpret_cast<float*>(data.data()); }
    void resize(size_t n) override { data.resize(n); }
  };

  template <typename T>
  std::vector<T>& find(const std::string& name) {
    auto it = m_vars.find(name);
    if (it == m_vars.end()) throw std::runtime_error("Material point data has no variable named '" + name + "'");
    Var<T>* v = dynamic_cast<Var<T>*>(it->second.get());
    if (!v) throw std::runtime_error("Material point data has no variable named '" + name + "' of the specified type");
    return v->data;
  }
  void pull() {
    if (!m_deviceNewer || !m_ctx) return;
    m_deviceNewer = false;
    for (auto& kv : m_vars) {
      const int g = detail::device_variable(kv.first);
      if (g == 0 || (g == 2 && !m_modelBound)) continue;
      kv.second->resize((size_t)m_deviceCount);
      detail::check(m_ctx, mpm_particles_get(m_ctx, kv.first.c_str(), kv.second->ptr(), kv.second->floats()));
    }
  }

  std::map<std::string, std::unique_ptr<VarBase>> m_vars;
  mpm_context* m_ctx = nullptr;
  int64_t m_deviceCount = 0;
  bool m_hostDirty = true;
  bool m_deviceNewer = false;
  bool m_hasModel = false;
  bool m_modelBound = false;
  float m_model[5] = {};
};

// ---- shape function (ShapeFunction.h, CubicBsplineShapeFunction.h) --------------------------
class ShapeFunction {
 public:
  virtual ~ShapeFunction() {}
  virtual int supportRadius() const = 0;
  virtual float w(float x) const = 0;
  virtual float dw(float x) const = 0;
};

// The GPU kernels evaluate this stencil themselves; the host methods are the same function
// (cubic B-spline, support radius 2) for callers that use it directly.
class CubicBsplineShapeFunction : public ShapeFunction {
 public:
  int supportRadius() const override { return 2; }
  float w(float x) const override {
    const float a = std::fabs(x);
    if (a < 1.0f) return 0.5f * a * a * a - a * a + 2.0f / 3.0f;
    if (a < 2.0f) {
      const float b = 2.0f - a;
      return b * b * b / 6.0f;
    }
    return 0.0f;
  }
  float dw(float x) const override {
    const float a = std::fabs(x), s = x < 0.0f ? -1.0f : 1.0f;
    if (a < 1.0f) return s * (1.5f * a * a - 2.0f * a);
    if (a < 2.0f) {
      const float b = 2.0f - a;
      return -s * 0.5f * b * b;
    }
    return 0.0f;
  }
};

// ---- constitutive model (ConstitutiveModel.h, SnowConstitutiveModel.h) ---------------------
class ConstitutiveModel {
 public:
  virtual ~ConstitutiveModel() {}
  virtual void setParticles(MaterialPointData& p) = 0;
  virtual void updateParticleData() = 0;
  virtual float energyDensity(size_t p) const = 0;
  virtual Matrix3f dEnergyDensitydF(size_t p) const = 0;
  virtual Matrix3f dEdFDifferential(const Matrix3f& dFp, size_t p) const = 0;
};

class SnowConstitutiveModel : public ConstitutiveModel {
 public:
  SnowConstitutiveModel(float youngsModulus, float poissonRatio, float hardening, float compressiveStrength,
                        float tensileStrength)
      : m_params{youngsModulus, poissonRatio, hardening, compressiveStrength, tensileStrength} {}

  // creates Fp, FinvTrans, J, R, Ginv, mu, lambda (SnowConstitutiveModel.cpp:123-169); the
  // device initialises them the same way (identity, 1, the Lame parameters)
  void setParticles(MaterialPointData& p) override {
    m_p = &p;
    for (const char* name : {"Fp", "FinvTrans", "R", "Ginv"}) p.createVariable<Matrix3f>(name);
    for (const char* name : {"J", "mu", "lambda"}) p.createVariable<float>(name);
    const size_t n = p.variable<float>("m").size();
    const float E = m_params[0], nu = m_params[1];
    const float mu = E / (2 * (1 + nu)), lambda = E * nu / ((1 + nu) * (1 - 2 * nu));
    for (const char* name : {"Fp", "FinvTrans", "R", "Ginv"}) p.variable<Matrix3f>(name).assign(n, Matrix3f::Identity());
    p.variable<float>("J").assign(n, 1.0f);
    p.variable<float>("mu").assign(n, mu);
    p.variable<float>("lambda").assign(n, lambda);
    p.bindSnowModel(m_params);
  }
  // plasticity, polar decomposition and hardening over every particle (:27-121)
  void updateParticleData() override {
    mpm_context* ctx = bound().device();
    detail::check(ctx, mpm_model_update_particle_data(ctx));
    m_p->markDeviceNewer();
  }
  float energyDensity(size_t p) const override {
    float out = 0.0f;
    const int32_t idx = (int32_t)p;
    mpm_context* ctx = bound().device();
    detail::check(ctx, mpm_model_energy_density(ctx, &idx, 1, &out));
    return out;
  }
  Matrix3f dEnergyDensitydF(size_t p) const override {
    Matrix3f out;
    const int32_t idx = (int32_t)p;
    mpm_context* ctx = bound().device();
    detail::check(ctx, mpm_model_stress(ctx, &idx, 1, out.data()));
    return out;
  }
  Matrix3f dEdFDifferential(const Matrix3f& dFp, size_t p) const override {
    Matrix3f out;
    const int32_t idx = (int32_t)p;
    mpm_context* ctx = bound().device();
    detail::check(ctx, mpm_model_stress_differential(ctx, &idx, dFp.data(), 1, out.data()));
    return out;
  }

 private:
  MaterialPointData& bound() const {
    if (!m_p) throw std::runtime_error("SnowConstitutiveModel: setParticles has not been called");
    return *m_p;
  }
  float m_params[5];
  MaterialPointData* m_p = nullptr;
};

// ---- force fields (ForceField.h, GravityField.h) ------------------------------------------
class ForceField {
 public:
  virtual ~ForceField() {}
  virtual Vector3f force(const Vector3f& x, float m) const = 0;
  virtual Matrix3f dFdx(const Vector3f& x, float m) const = 0;
  // the uniform acceleration this field applies, when it is one the device runs
  virtual bool gpuGravity(Vector3f& /*g*/) const { return false; }

  // owns the fields added to it (ForceField.cpp)
  class ForceFieldSet {
   public:
    ForceFieldSet() {}
    ForceFieldSet(const ForceFieldSet&) = delete;
    ForceFieldSet& operator=(const ForceFieldSet&) = delete;
    ~ForceFieldSet() {
      for (const ForceField* f : m_fields) delete f;
    }
    void add(ForceField* f) { m_fields.push_back(f); }
    void force(Vector3f& f, const Vector3f& x, float m) const {
      for (const ForceField* ff : m_fields) f += ff->force(x, m);
    }
    void dFdx(Matrix3f& df, const Vector3f& x, float m) const {
      for (const ForceField* ff : m_fields) {
        const Matrix3f d = ff->dFdx(x, m);
        for (int i = 0; i < 9; ++i) df.m[i] += d.m[i];
      }
    }
    size_t numFields() const { return m_fields.size(); }
    // upload to a context: every field must be one the device runs
    void toDevice(mpm_context* ctx) const {
      std::vector<float> g;
      for (const ForceField* f : m_fields) {
        Vector3f a;
        if (!f->gpuGravity(a)) throw std::runtime_error("force field has no GPU implementation");
        g.insert(g.end(), a.v, a.v + 3);
      }
      detail::check(ctx, mpm_set_gravity(ctx, g.empty() ? nullptr : g.data(), (int)(g.size() / 3)));
    }

   private:
    std::vector<const ForceField*> m_fields;
  };
};

class GravityField : public ForceField {
 public:
  explicit GravityField(const Vector3f& gravity) : m_gravity(gravity) {}
  Vector3f force(const Vector3f& /*x*/, float m) const override { return m_gravity * m; }
  Matrix3f dFdx(const Vector3f& /*x*/, float /*m*/) const override { return Matrix3f::Zero(); }
  bool gpuGravity(Vector3f& g) const override {
    g = m_gravity;
    return true;
  }

 private:
  Vector3f m_gravity;
};

// ---- collision objects (CollisionObject.h, CollisionPlane.h) -------------------------------
class CollisionObject {
 public:
  virtual ~CollisionObject() {}
  virtual float phi(const Vector3f& x) const = 0;
  virtual void grad(const Vector3f& x, Vector3f& dPhi) const = 0;
  virtual void velocity(const Vector3f& x, Vector3f& v) const = 0;
  virtual float coulombFriction() const = 0;
  virtual bool sticky() const = 0;
  // the device's description of this object; false when the device has no such object
  virtual bool gpuCollider(mpm_collider& /*c*/) const { return false; }

  // owns the objects added to it (CollisionObject.cpp)
  class CollisionObjectSet {
   public:
    CollisionObjectSet() {}
    CollisionObjectSet(const CollisionObjectSet&) = delete;
    CollisionObjectSet& operator=(const CollisionObjectSet&) = delete;
    ~CollisionObjectSet() {
      for (const CollisionObject* o : m_objects) delete o;
    }
    void add(CollisionObject* o) { m_objects.push_back(o); }
    // deletes the objects (the reference's SOP rebuilds its set at every cook by assigning a fresh one)
    void clear() {
      for (const CollisionObject* o : m_objects) delete o;
      m_objects.clear();
    }
    const CollisionObject* object(size_t i) const { return m_objects[i]; }
    size_t numObjects() const { return m_objects.size(); }
    // the objects' device descriptions, in order
    std::vector<mpm_collider> gpuColliders() const {
      std::vector<mpm_collider> cs(m_objects.size());
      for (size_t i = 0; i < m_objects.size(); ++i) {
        std::memset(&cs[i], 0, sizeof(mpm_collider));
        if (!m_objects[i]->gpuCollider(cs[i])) throw std::runtime_error("collision object has no GPU implementation");
      }
      return cs;
    }

   private:
    std::vector<const CollisionObject*> m_objects;
  };
};

// phi = c0 x + c1 y + c2 z + c3, uniform velocity (CollisionPlane.cpp)
class CollisionPlane : public CollisionObject {
 public:
  CollisionPlane(const Vector4f& c, float coulombFriction = 0.5f, bool sticky = false)
      : m_coeffs(c), m_coulombFriction(coulombFriction), m_sticky(sticky) {}
  float phi(const Vector3f& x) const override {
    return m_coeffs[0] * x[0] + m_coeffs[1] * x[1] + m_coeffs[2] * x[2] + m_coeffs[3];
  }
  void grad(const Vector3f& /*x*/, Vector3f& dPhi) const override { dPhi = Vector3f(m_coeffs[0], m_coeffs[1], m_coeffs[2]); }
  void velocity(const Vector3f& /*x*/, Vector3f& v) const override { v = m_v; }
  float coulombFriction() const override { return m_coulombFriction; }
  bool sticky() const override { return m_sticky; }
  void setCoeffs(const Vector4f& c) { m_coeffs = c; }
  void setV(const Vector3f& v) { m_v = v; }
  bool gpuCollider(mpm_collider& c) const override {
    c.type = 0;
    for (int i = 0; i < 4; ++i) c.coeffs[i] = m_coeffs[i];
    for (int i = 0; i < 3; ++i) c.velocity[i] = m_v[i];
    c.friction = m_coulombFriction;
    c.sticky = m_sticky ? 1 : 0;
    return true;
  }

 private:
  Vector4f m_coeffs;
  Vector3f m_v;
  float m_coulombFriction;
  bool m_sticky;
};

// phi = |x - center| - radius (no reference counterpart: an analytic object the device runs)
class CollisionSphere : public CollisionObject {
 public:
  CollisionSphere(const Vector3f& center, float radius, float coulombFriction = 0.5f, bool sticky = false)
      : m_center(center), m_radius(radius), m_coulombFriction(coulombFriction), m_sticky(sticky) {}
  float phi(const Vector3f& x) const override { return (x - m_center).norm() - m_radius; }
  void grad(const Vector3f& x, Vector3f& dPhi) const override {
    const Vector3f d = x - m_center;
    const float r = d.norm();
    dPhi = r > 0.0f ? d * (1.0f / r) : Vector3f(0.0f, 1.0f, 0.0f);
  }
  void velocity(const Vector3f& /*x*/, Vector3f& v) const override { v = m_v; }
  float coulombFriction() const override { return m_coulombFriction; }
  bool sticky() const override { return m_sticky; }
  void setV(const Vector3f& v) { m_v = v; }
  bool gpuCollider(mpm_collider& c) const override {
    c.type = 1;
    for (int i = 0; i < 3; ++i) c.center[i] = m_center[i], c.velocity[i] = m_v[i];
    c.radius = m_radius;
    c.friction = m_coulombFriction;
    c.sticky = m_sticky ? 1 : 0;
    return true;
  }

 private:
  Vector3f m_center, m_v;
  float m_radius, m_coulombFriction;
  bool m_sticky;
};

// A level set sampled on a regular lattice: the device form of the reference's
// VDBCollisionObject (houdiniPlugin/VDBCollisionObject.cpp:11-37; friction 0.5, not sticky),
// whose caller samples its VDB grids into these arrays. Point (i, j, k) sits at
// origin + voxel (i, j, k) and at phi[i + dims0 (j + dims1 k)]; velocities are 3 floats per
// point or empty (then the uniform setV velocity applies). The device samples trilinearly,
// clamped to the lattice; the host phi() below is the same interpolation.
class LevelSetCollisionObject : public CollisionObject {
 public:
  LevelSetCollisionObject(const int dims[3], const Vector3f& origin, float voxel, std::vector<float> phi,
                          std::vector<float> velocities = std::vector<float>(), float coulombFriction = 0.5f,
                          bool sticky = false)
      : m_origin(origin), m_voxel(voxel), m_phi(std::move(phi)), m_vel(std::move(velocities)),
        m_coulombFriction(coulombFriction), m_sticky(sticky) {
    for (int i = 0; i < 3; ++i) m_dims[i] = dims[i];
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    if (m_phi.size() != n || (!m_vel.empty() && m_vel.size() != 3 * n))
      throw std::runtime_error("level set arrays do not match the lattice dimensions");
  }
  float phi(const Vector3f& x) const override { return sample(m_phi.data(), 1, x, 0); }
  // central difference of the sampled phi over one voxel per axis
  void grad(const Vector3f& x, Vector3f& dPhi) const override {
    for (int a = 0; a < 3; ++a) {
      Vector3f lo = x, hi = x;
      lo[a] = x[a] - m_voxel;
      hi[a] = x[a] + m_voxel;
      dPhi[a] = (phi(hi) - phi(lo)) / (2.0f * m_voxel);
    }
  }
  void velocity(const Vector3f& x, Vector3f& v) const override {
    if (m_vel.empty()) {
      v = m_v;
      return;
    }
    for (int c = 0; c < 3; ++c) v[c] = sample(m_vel.data(), 3, x, c);
  }
  float coulombFriction() const override { return m_coulombFriction; }
  bool sticky() const override { return m_sticky; }
  void setV(const Vector3f& v) { m_v = v; }
  const Vector3f& origin() const { return m_origin; }
  float voxel() const { return m_voxel; }
  const int* dims() const { return m_dims; }
  bool hasVelocityField() const { return !m_vel.empty(); }
  bool gpuCollider(mpm_collider& c) const override {
    c.type = 2;
    for (int i = 0; i < 3; ++i) c.dims[i] = m_dims[i], c.origin[i] = m_origin[i], c.velocity[i] = m_v[i];
    c.voxel = m_voxel;
    c.phi = m_phi.data();
    c.phi_velocity = m_vel.empty() ? nullptr : m_vel.data();
    c.friction = m_coulombFriction;
    c.sticky = m_sticky ? 1 : 0;
    return true;
  }

 private:
  float sample(const float* a, int stride, const Vector3f& x, int comp) const {
    int i0[3];
    float t[3];
    for (int ax = 0; ax < 3; ++ax) {
      float u = (x[ax] - m_origin[ax]) / m_voxel;
      const float hi = (float)(m_dims[ax] - 1);
      if (!(u >= 0.0f)) u = 0.0f;
      if (u > hi) u = hi;
      int i = (int)std::floor(u);
      if (i > m_dims[ax] - 2) i = m_dims[ax] - 2;
      i0[ax] = i;
      t[ax] = u - (float)i;
    }
    const size_t sx = (size_t)stride, sy = sx * m_dims[0], sz = sy * m_dims[1];
    const float* q = a + i0[0] * sx + i0[1] * sy + i0[2] * sz + comp;
    const float e00 = lerp(q[0], q[sx], t[0]), e10 = lerp(q[sy], q[sy + sx], t[0]);
    const float e01 = lerp(q[sz], q[sz + sx], t[0]), e11 = lerp(q[sz + sy], q[sz + sy + sx], t[0]);
    return lerp(lerp(e00, e10, t[1]), lerp(e01, e11, t[1]), t[2]);
  }
  static float lerp(float a, float b, float t) { return a + t * (b - a); }
  int m_dims[3];
  Vector3f m_origin, m_v;
  float m_voxel;
  std::vector<float> m_phi, m_vel;
  float m_coulombFriction;
  bool m_sticky;
};

// ---- solver hooks (TerminationCriterion.h, SquareMagnitudeTermination.h, LinearSolver.h) ---
class TerminationCriterion {
 public:
  virtual ~TerminationCriterion() {}
  virtual bool operator()(VectorXf& r, int iterationNumber) const = 0;
  virtual bool cancelled() const = 0;
};

// stop when |r|^2 < tolError^2 |b|^2 or after maxIters iterations (SquareMagnitudeTermination.cpp)
class SquareMagnitudeTermination : public TerminationCriterion {
 public:
  SquareMagnitudeTermination(int maxIters, float tolError) : m_maxIters((float)maxIters), m_tolError(tolError) {}
  // threshold from the right hand side, for host use of operator() (the device solve sets its own)
  void init(const VectorXf& b) {
    double s = 0.0;
    for (float x : b) s += (double)x * x;
    m_threshold = (float)(m_tolError * m_tolError * s);
  }
  bool operator()(VectorXf& r, int iterationNumber) const override {
    if ((float)iterationNumber >= m_maxIters) return true;
    double s = 0.0;
    for (float x : r) s += (double)x * x;
    return s < m_threshold;
  }
  bool cancelled() const override { return false; }
  // true when this object's residual test is the base class's, which the device then runs; a
  // subclass that keeps the test and only adds cancelled() overrides this to return true
  virtual bool deviceTest() const { return typeid(*this) == typeid(SquareMagnitudeTermination); }
  int maxIters() const { return (int)m_maxIters; }
  float tolError() const { return m_tolError; }

 private:
  float m_maxIters, m_tolError, m_threshold = 0.0f;
};

class LinearSolver {
 public:
  class Debug {
   public:
    virtual ~Debug() {}
    virtual void operator()(VectorXf& x) = 0;
  };
};

namespace detail {
struct Hooks {
  TerminationCriterion* term;
  LinearSolver::Debug* debug;
  VectorXf buf;
};
inline int term_cb(void* u, const float* r, int64_t n, int i) {
  Hooks* h = static_cast<Hooks*>(u);
  h->buf.assign(r, r + n);
  return (*h->term)(h->buf, i) ? 1 : 0;
}
inline void debug_cb(void* u, const float* x, int64_t n) {
  Hooks* h = static_cast<Hooks*>(u);
  h->buf.assign(x, x + n);
  (*h->debug)(h->buf);
}
inline int cancel_cb(void* u) { return static_cast<Hooks*>(u)->term->cancelled() ? 1 : 0; }

// device settings for a termination criterion and debug hook; hooks only where needed
inline mpm_solver_params solver_params(TerminationCriterion& t, LinearSolver::Debug* d, Hooks& h) {
  mpm_solver_params p;
  std::memset(&p, 0, sizeof(p));
  h.term = &t;
  h.debug = d;
  p.user = &h;
  const SquareMagnitudeTermination* sq = dynamic_cast<const SquareMagnitudeTermination*>(&t);
  if (sq && sq->deviceTest()) {
    p.max_iters = sq->maxIters();
    p.tolerance = sq->tolError();
    // a subclass may be cancelled: poll it after each body's solve (the base never is, and a
    // cancellation hook keeps the batch in body order, which a plain solve has no need of)
    if (typeid(t) != typeid(SquareMagnitudeTermination)) p.cancelled = cancel_cb;
  } else {
    p.termination = term_cb;
    p.cancelled = cancel_cb;
  }
  if (d) p.debug = debug_cb;
  return p;
}

}  // namespace detail

// ---- ProceduralMatrix (ProceduralMatrix.h:8-19) ------------------------------------------------
class ProceduralMatrix {
 public:
  virtual ~ProceduralMatrix() {}
  virtual void multVector(const VectorXf& x, VectorXf& result) const = 0;
  virtual void multInverseVector(const VectorXf& x, VectorXf& result) const = 0;
  virtual void subspaceProject(VectorXf& x) const = 0;
};

namespace detail {
struct ProcHooks {
  const ProceduralMatrix* mat;
  VectorXf in, out;
};
inline void proc_mult(void* u, const float* x, float* r, int64_t n) {
  ProcHooks* h = static_cast<ProcHooks*>(u);
  h->in.assign(x, x + n);
  h->out.assign(n, 0.0f);
  h->mat->multVector(h->in, h->out);
  if ((int64_t)h->out.size() != n) throw std::runtime_error("multVector returned a vector of the wrong size");
  std::memcpy(r, h->out.data(), sizeof(float) * n);
}
inline void proc_mult_inverse(void* u, const float* x, float* r, int64_t n) {
  ProcHooks* h = static_cast<ProcHooks*>(u);
  h->in.assign(x, x + n);
  h->out.assign(n, 0.0f);
  h->mat->multInverseVector(h->in, h->out);
  if ((int64_t)h->out.size() != n) throw std::runtime_error("multInverseVector returned a vector of the wrong size");
  std::memcpy(r, h->out.data(), sizeof(float) * n);
}
inline void proc_project(void* u, float* x, int64_t n) {
  ProcHooks* h = static_cast<ProcHooks*>(u);
  h->in.assign(x, x + n);
  h->mat->subspaceProject(h->in);
  std::memcpy(x, h->in.data(), sizeof(float) * n);
}
inline mpm_procedural_matrix procedural(ProcHooks& h) {
  mpm_procedural_matrix m;
  std::memset(&m, 0, sizeof(m));
  m.mult = proc_mult;
  m.mult_inverse = proc_mult_inverse;
  m.project = proc_project;
  m.user = &h;
  m.device_pointers = 0;
  return m;
}

// uploads colliders when their descriptions changed since the last upload
struct ColliderCache {
  std::vector<mpm_collider> last;
  bool valid = false;
  void upload(mpm_context* ctx, const CollisionObject::CollisionObjectSet& set) {
    std::vector<mpm_collider> cs = set.gpuColliders();
    if (valid && cs.size() == last.size() &&
        (cs.empty() || std::memcmp(cs.data(), last.data(), cs.size() * sizeof(mpm_collider)) == 0))
      return;
    detail::check(ctx, mpm_set_colliders(ctx, cs.empty() ? nullptr : cs.data(), (int)cs.size()));
    last = cs;
    valid = true;
  }
};
}  // namespace detail

// ---- ConjugateResiduals (ConjugateResiduals.h:20-60, ConjugateResiduals.cpp:19-160) --------------
// The recurrences run on the GPU (device 0, a context of its own created on the first solve); the
// matrix and the preconditioner are called on host vectors once per product. With log set, the
// residuals and search directions of the setup and of every continuing iteration are kept
// (the reference's m_residuals / m_searchDirections, public here through residuals() and
// searchDirections()).
class ConjugateResiduals : public LinearSolver {
 public:
  ConjugateResiduals(TerminationCriterion& terminationCriterion, const ProceduralMatrix* preconditioner = 0,
                     bool log = false)
      : m_terminationCriterion(terminationCriterion), m_preconditioner(preconditioner), m_log(log) {}
  ~ConjugateResiduals() {
    if (m_ctx) mpm_destroy(m_ctx);
  }
  ConjugateResiduals(const ConjugateResiduals&) = delete;
  ConjugateResiduals& operator=(const ConjugateResiduals&) = delete;

  void operator()(const ProceduralMatrix& mat, const VectorXf& rhs, VectorXf& x, Debug* d = 0) const {
    const int64_t n = (int64_t)rhs.size();
    if ((int64_t)x.size() != n) throw std::runtime_error("x and rhs differ in size");
    if (!m_ctx && mpm_create(&m_ctx, 0) != MPM_OK) {
      m_ctx = nullptr;
      throw std::runtime_error("mpm_create failed: no usable HIP device");
    }
    detail::ProcHooks hm{&mat, {}, {}};
    detail::ProcHooks hp{m_preconditioner, {}, {}};
    mpm_procedural_matrix A = detail::procedural(hm);
    mpm_procedural_matrix P = detail::procedural(hp);
    detail::Hooks hooks;
    mpm_solver_params prm = detail::solver_params(m_terminationCriterion, d, hooks);
    prm.cancelled = nullptr;
    mpm_solve_stats st;
    std::memset(&st, 0, sizeof(st));
    // the setup's entry and one per iteration that continues
    const int cap = m_log ? (prm.termination ? 4096 : prm.max_iters + 2) : 0;
    std::vector<float> res((size_t)cap * n), dirs((size_t)cap * n);
    int logged = 0;
    detail::check(m_ctx, mpm_conjugate_residuals(m_ctx, &A, m_preconditioner ? &P : nullptr, rhs.data(), x.data(), n,
                                                 &prm, &st, cap ? res.data() : nullptr,
                                                 cap ? dirs.data() : nullptr, cap, &logged));
    m_iterations = st.iterations;
    for (int k = 0; k < logged; ++k) {
      m_residuals.emplace_back(res.begin() + (size_t)k * n, res.begin() + (size_t)(k + 1) * n);
      m_searchDirections.emplace_back(dirs.begin() + (size_t)k * n, dirs.begin() + (size_t)(k + 1) * n);
    }
  }

  const std::vector<VectorXf>& residuals() const { return m_residuals; }
  const std::vector<VectorXf>& searchDirections() const { return m_searchDirections; }
  int iterations() const { return m_iterations; }

 private:
  TerminationCriterion& m_terminationCriterion;
  const ProceduralMatrix* m_preconditioner;
  bool m_log;
  mutable mpm_context* m_ctx = nullptr;
  mutable int m_iterations = 0;
  mutable std::vector<VectorXf> m_residuals;
  mutable std::vector<VectorXf> m_searchDirections;
};

inline void require_cubic_bspline(const ShapeFunction& sf) {
  if (!dynamic_cast<const CubicBsplineShapeFunction*>(&sf))
    throw std::runtime_error("shape function has no GPU implementation (the device stencil is the cubic B-spline)");
}

// ---- Sim (Sim.h:32-55, Sim.cpp:17-203) ------------------------------------------------------
class Sim {
 public:
  typedef std::vector<int> IndexList;
  typedef IndexList::iterator IndexIterator;
  typedef IndexList::const_iterator ConstIndexIterator;
  typedef std::vector<IndexList> BodyList;
  typedef BodyList::iterator BodyIterator;
  typedef BodyList::const_iterator ConstBodyIterator;

  // particle variables, model variables, bodies and particle volumes (Sim.cpp:17-50)
  Sim(const std::vector<Vector3f>& x, const std::vector<float>& masses, float gridSize,
      const ShapeFunction& shapeFunction, ConstitutiveModel& model,
      const CollisionObject::CollisionObjectSet& collisionObjects, const ForceField::ForceFieldSet& forceFields,
      int dimension = 3)
      : m_gridSize(gridSize), m_shapeFunction(shapeFunction), m_constitutiveModel(model),
        m_collisionObjects(collisionObjects), m_forceFields(forceFields), m_dimension(dimension) {
    require_cubic_bspline(shapeFunction);
    if (masses.size() != x.size()) throw std::runtime_error("positions and masses differ in count");
    m_particleData.variable<Vector3f>("p") = x;
    m_particleData.variable<Vector3f>("v").assign(x.size(), Vector3f::Zero());
    m_particleData.variable<Matrix3f>("F").assign(x.size(), Matrix3f::Identity());
    m_particleData.variable<float>("m") = masses;
    m_particleData.variable<float>("volume").assign(x.size(), 0.0f);
    m_constitutiveModel.setParticles(m_particleData);
    mpm_context* ctx = m_particleData.device();
    detail::check(ctx, mpm_sim_init(ctx, m_gridSize, m_dimension));
    m_particleData.markDeviceNewer();
  }

  MaterialPointData& particleData() { return m_particleData; }
  const MaterialPointData& particleData() const { return m_particleData; }

  // one time step of every body and the ballistic particles (Sim.cpp:57-134)
  void advance(float timeStep, TerminationCriterion& terminationCriterion, LinearSolver::Debug* d = 0) {
    if (!m_particleData.hasGpuModel()) throw std::runtime_error("constitutive model has no GPU implementation");
    mpm_context* ctx = m_particleData.device();
    m_forceFields.toDevice(ctx);
    m_colliders.upload(ctx, m_collisionObjects);
    detail::Hooks hooks;
    mpm_solver_params prm = detail::solver_params(terminationCriterion, d, hooks);
    mpm_step_stats st;
    const int rc = mpm_sim_advance(ctx, timeStep, &prm, &st);
    m_particleData.markDeviceNewer();
    m_partitionValid = false;
    detail::check(ctx, rc);
    m_lastStats = st;
  }

  // the partition the next step uses (Sim.cpp:136-203); bodies in the device's body order,
  // particle indices ascending within a body
  const BodyList& bodies() const {
    refresh();
    return m_bodies;
  }
  const IndexList& ballisticParticles() const {
    refresh();
    return m_ballistic;
  }
  const mpm_step_stats& lastStepStats() const { return m_lastStats; }

 private:
  void refresh() const {
    if (m_partitionValid) return;
    mpm_context* ctx = const_cast<MaterialPointData&>(m_particleData).device();
    const int nb = mpm_sim_num_bodies(ctx);
    if (nb < 0) throw std::runtime_error(mpm_last_error(ctx));
    m_bodies.assign((size_t)nb, IndexList());
    for (int b = 0; b < nb; ++b) {
      const int64_t cnt = mpm_sim_body(ctx, b, nullptr);
      if (cnt < 0) throw std::runtime_error(mpm_last_error(ctx));
      m_bodies[b].resize((size_t)cnt);
      if (cnt) mpm_sim_body(ctx, b, m_bodies[b].data());
    }
    const int64_t nballistic = mpm_sim_ballistic(ctx, nullptr);
    if (nballistic < 0) throw std::runtime_error(mpm_last_error(ctx));
    m_ballistic.resize((size_t)nballistic);
    if (nballistic) mpm_sim_ballistic(ctx, m_ballistic.data());
    m_partitionValid = true;
  }

  MaterialPointData m_particleData;
  float m_gridSize;
  const ShapeFunction& m_shapeFunction;
  ConstitutiveModel& m_constitutiveModel;
  const CollisionObject::CollisionObjectSet& m_collisionObjects;
  const ForceField::ForceFieldSet& m_forceFields;
  int m_dimension;
  detail::ColliderCache m_colliders;
  mpm_step_stats m_lastStats = {};
  mutable BodyList m_bodies;
  mutable IndexList m_ballistic;
  mutable bool m_partitionValid = false;
};

// ---- Grid of one body (Grid.h:23-217, Grid.cpp) ---------------------------------------------
// The public members are the reference's; the force, momentum and collision members the
// reference keeps private for its tests are public here, with the same arguments.
class Grid {
 public:
  // bounding box, binning, mass and velocity splat (Grid.cpp:156-263)
  Grid(MaterialPointData& d, const Sim::IndexList& particleInds, float gridSize, const ShapeFunction& shapeFunction,
       Vector3f frameVelocity = Vector3f::Zero(), int dimension = 3)
      : m_d(d) {
    require_cubic_bspline(shapeFunction);
    m_ctx = d.device();
    detail::check(m_ctx, mpm_grid_create(m_ctx, particleInds.data(), (int64_t)particleInds.size(), gridSize,
                                         frameVelocity.data(), dimension, &m_g));
    detail::check(m_ctx, mpm_grid_dims(m_g, m_n, m_min.data()));
  }
  ~Grid() {
    if (m_g) mpm_grid_destroy(m_g);
  }
  Grid(const Grid&) = delete;
  Grid& operator=(const Grid&) = delete;

  void computeParticleVolumes() const {
    detail::check(m_ctx, mpm_grid_compute_particle_volumes(m_g));
    m_d.markDeviceNewer();
  }
  // implicit velocity update by diagonal-preconditioned conjugate residuals (Grid.cpp:889-1016)
  void updateGridVelocities(float timeStep, const ConstitutiveModel& /*constitutiveModel*/,
                            const CollisionObject::CollisionObjectSet& collisionObjects,
                            const ForceField::ForceFieldSet& fields, TerminationCriterion& termination,
                            LinearSolver::Debug* d = 0) {
    upload(collisionObjects, fields);
    detail::Hooks hooks;
    mpm_solver_params prm = detail::solver_params(termination, d, hooks);
    std::memset(&m_solveStats, 0, sizeof(m_solveStats));
    detail::check(m_ctx, mpm_grid_update_grid_velocities(m_g, timeStep, &prm, &m_solveStats));
  }
  void updateDeformationGradients(float timeStep) {
    detail::check(m_ctx, mpm_grid_update_deformation_gradients(m_g, timeStep));
    m_d.markDeviceNewer();
  }
  void updateParticleVelocities() {
    detail::check(m_ctx, mpm_grid_update_particle_velocities(m_g));
    m_d.markDeviceNewer();
  }

  // (Grid.cpp:438-463)
  void calculateForces(VectorXf& forces, const ConstitutiveModel& /*model*/, const ForceField::ForceFieldSet& fields) const {
    fields.toDevice(m_ctx);
    forces.resize(3 * nodes());
    detail::check(m_ctx, mpm_grid_calculate_forces(m_g, forces.data()));
  }
  // (Grid.cpp:620-632)
  void calculateForceDifferentials(VectorXf& df, const VectorXf& dx, const ConstitutiveModel& /*model*/,
                                   const ForceField::ForceFieldSet& /*fields*/) const {
    if (dx.size() != 3 * nodes()) throw std::runtime_error("dx does not match the grid");
    df.resize(3 * nodes());
    detail::check(m_ctx, mpm_grid_force_differentials(m_g, dx.data(), df.data()));
  }
  // (Grid.cpp:634-690)
  void calculateExplicitMomenta(VectorXf& explicitMomenta, std::vector<char>& nodeCollided, float timeStep,
                                const ConstitutiveModel& /*model*/,
                                const CollisionObject::CollisionObjectSet& collisionObjects,
                                const ForceField::ForceFieldSet& fields) {
    upload(collisionObjects, fields);
    explicitMomenta.resize(3 * nodes());
    nodeCollided.resize(nodes());
    detail::check(m_ctx, mpm_grid_explicit_momenta(m_g, timeStep, explicitMomenta.data(),
                                                   reinterpret_cast<int8_t*>(nodeCollided.data())));
  }
  // (Grid.cpp:692-724)
  void collisionVelocities(VectorXf& vc, const CollisionObject::CollisionObjectSet& collisionObjects,
                           const std::vector<char>& nodeCollided) const {
    m_colliders.upload(m_ctx, collisionObjects);
    if (nodeCollided.size() != nodes()) throw std::runtime_error("collision flags do not match the grid");
    vc.resize(3 * nodes());
    detail::check(m_ctx, mpm_grid_collision_velocities(m_g, reinterpret_cast<const int8_t*>(nodeCollided.data()),
                                                       vc.data()));
  }
  // (Grid.cpp:539-546)
  void dForceidXi(VectorXf& dfdxi, const ConstitutiveModel& /*model*/) const {
    dfdxi.resize(3 * nodes());
    detail::check(m_ctx, mpm_grid_dforce_dxi(m_g, dfdxi.data()));
  }

  // grid state (the reference's m_masses, m_velocities, m_prevVelocities, m_n, m_min)
  VectorXf masses() const { return get("masses", 1); }
  VectorXf velocities() const { return get("velocities", 3); }
  VectorXf prevVelocities() const { return get("prev_velocities", 3); }
  const int* n() const { return m_n; }
  const Vector3f& min() const { return m_min; }
  size_t nodes() const { return (size_t)m_n[0] * m_n[1] * m_n[2]; }
  const mpm_solve_stats& solveStats() const { return m_solveStats; }

 private:
  VectorXf get(const char* name, int comps) const {
    VectorXf out(comps * nodes());
    detail::check(m_ctx, mpm_grid_get(m_g, name, out.data()));
    return out;
  }
  void upload(const CollisionObject::CollisionObjectSet& cols, const ForceField::ForceFieldSet& fields) const {
    fields.toDevice(m_ctx);
    m_colliders.upload(m_ctx, cols);
  }
  MaterialPointData& m_d;
  mpm_context* m_ctx = nullptr;
  mpm_grid* m_g = nullptr;
  int m_n[3] = {0, 0, 0};
  Vector3f m_min;
  mutable detail::ColliderCache m_colliders;
  mpm_solve_stats m_solveStats = {};
};

}  // namespace MpmSim
