// mpm_oracle_capi.cpp -- flat C API over the templated CPU oracle.
//
// TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, smoke() and the
// cpu_baseline leg of bench.py through ctypes. Arrays cross this API as
// double and are converted to the instantiation's scalar type (float: the
// reference's precision; double: the FP64 oracle).
#include "mpm_oracle.hpp"

#include <chrono>
#include <cstdio>
#include <memory>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace mpmo;

namespace {
thread_local std::string g_err;

struct Handle {
  int prec;  // 0 = float, 1 = double
  explicit Handle(int p) : prec(p) {}
  virtual ~Handle() {}
};
template <class S> struct PDH : Handle {
  Particles<S> own;
  Particles<S>* pd;
  PDH() : Handle(sizeof(S) == 8), pd(&own) {}
};
template <class S> struct ModelH : Handle {
  std::unique_ptr<ConstitutiveModel<S>> m;
  ModelH() : Handle(sizeof(S) == 8) {}
};
template <class S> struct ColsH : Handle {
  CollisionSet<S> c;
  ColsH() : Handle(sizeof(S) == 8) {}
};
template <class S> struct FieldsH : Handle {
  ForceSet<S> f;
  FieldsH() : Handle(sizeof(S) == 8) {}
};
template <class S> struct GridH : Handle {
  std::unique_ptr<Grid<S>> g;
  GridH() : Handle(sizeof(S) == 8) {}
};
template <class S> struct SimH : Handle {
  std::unique_ptr<Sim<S>> s;
  PDH<S> view;
  SimH() : Handle(sizeof(S) == 8) {}
};

template <class F> int guard(F f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

template <class S> void setVar(Particles<S>& d, const std::string& name, const double* a, long count) {
  size_t n = d.size();
  auto v3 = [&](std::vector<V3<S>>& dst) {
    if (count != long(3 * n)) throw std::runtime_error("size mismatch for '" + name + "'");
    dst.resize(n);
    for (size_t i = 0; i < n; ++i) for (int c = 0; c < 3; ++c) dst[i][c] = S(a[3 * i + c]);
  };
  auto m3 = [&](std::vector<M3<S>>& dst) {
    if (count != long(9 * n)) throw std::runtime_error("size mismatch for '" + name + "'");
    dst.resize(n);
    for (size_t i = 0; i < n; ++i) for (int c = 0; c < 9; ++c) dst[i].a[c] = S(a[9 * i + c]);
  };
  auto s1 = [&](std::vector<S>& dst) {
    if (count != long(n)) throw std::runtime_error("size mismatch for '" + name + "'");
    dst.resize(n);
    for (size_t i = 0; i < n; ++i) dst[i] = S(a[i]);
  };
  if (name == "p") v3(d.p);
  else if (name == "v") v3(d.v);
  else if (name == "F") m3(d.F);
  else if (name == "m") s1(d.m);
  else if (name == "volume") s1(d.volume);
  else if (name == "Fp") m3(d.Fp);
  else if (name == "FinvTrans") m3(d.FinvTrans);
  else if (name == "R") m3(d.R);
  else if (name == "Ginv") m3(d.Ginv);
  else if (name == "J") s1(d.J);
  else if (name == "mu") s1(d.mu);
  else if (name == "lambda") s1(d.lambda);
  else if (name == "mu0") s1(d.mu0);
  else if (name == "lambda0") s1(d.lambda0);
  else if (name == "xi") s1(d.xi);
  else throw std::runtime_error("Material point data has no variable named '" + name + "'");
}
template <class S> void getVar(const Particles<S>& d, const std::string& name, double* a, long count) {
  size_t n = d.size();
  auto v3 = [&](const std::vector<V3<S>>& src) {
    if (count != long(3 * n) || src.size() != n) throw std::runtime_error("size mismatch for '" + name + "'");
    for (size_t i = 0; i < n; ++i) for (int c = 0; c < 3; ++c) a[3 * i + c] = double(src[i][c]);
  };
  auto m3 = [&](const std::vector<M3<S>>& src) {
    if (count != long(9 * n) || src.size() != n) throw std::runtime_error("size mismatch for '" + name + "'");
    for (size_t i = 0; i < n; ++i) for (int c = 0; c < 9; ++c) a[9 * i + c] = double(src[i].a[c]);
  };
  auto s1 = [&](const std::vector<S>& src) {
    if (count != long(n) || src.size() != n) throw std::runtime_error("size mismatch for '" + name + "'");
    for (size_t i = 0; i < n; ++i) a[i] = double(src[i]);
  };
  if (name == "p") v3(d.p);
  else if (name == "v") v3(d.v);
  else if (name == "F") m3(d.F);
  else if (name == "m") s1(d.m);
  else if (name == "volume") s1(d.volume);
  else if (name == "Fp") m3(d.Fp);
  else if (name == "FinvTrans") m3(d.FinvTrans);
  else if (name == "R") m3(d.R);
  else if (name == "Ginv") m3(d.Ginv);
  else if (name == "J") s1(d.J);
  else if (name == "mu") s1(d.mu);
  else if (name == "lambda") s1(d.lambda);
  else if (name == "mu0") s1(d.mu0);
  else if (name == "lambda0") s1(d.lambda0);
  else if (name == "xi") s1(d.xi);
  else throw std::runtime_error("Material point data has no variable named '" + name + "'");
}

template <class S> std::vector<S> toS(const double* a, size_t n) {
  std::vector<S> r(n);
  for (size_t i = 0; i < n; ++i) r[i] = S(a[i]);
  return r;
}
template <class S> void fromS(const std::vector<S>& v, double* out) {
  for (size_t i = 0; i < v.size(); ++i) out[i] = double(v[i]);
}
template <class S> V3<S> v3of(const double* a) { return a ? V3<S>(S(a[0]), S(a[1]), S(a[2])) : V3<S>(); }

// Dense matrix + optional diagonal preconditioner for CR unit tests
template <class S> struct DenseMat : ProceduralMatrix<S> {
  int N;
  std::vector<S> A;  // column-major
  void multVector(const Vec<S>& x, Vec<S>& r) const override {
    r.assign(N, S(0));
    for (int i = 0; i < N; ++i) {
      S acc = 0;
      for (int j = 0; j < N; ++j) acc += A[i + size_t(N) * j] * x[j];
      r[i] = acc;
    }
  }
  void multInverseVector(const Vec<S>&, Vec<S>&) const override {}
  void subspaceProject(Vec<S>&) const override {}
};
template <class S> struct DiagMat : ProceduralMatrix<S> {
  std::vector<S> d;
  void multVector(const Vec<S>& x, Vec<S>& r) const override { r.resize(x.size()); for (size_t i = 0; i < x.size(); ++i) r[i] = x[i] * d[i]; }
  void multInverseVector(const Vec<S>& x, Vec<S>& r) const override { r.resize(x.size()); for (size_t i = 0; i < x.size(); ++i) r[i] = x[i] / d[i]; }
  void subspaceProject(Vec<S>&) const override {}
};

#define DISPATCH(h, ...)                          \
  do {                                            \
    if (static_cast<Handle*>(h)->prec == 0) {     \
      typedef float S;                            \
      __VA_ARGS__;                                \
    } else {                                      \
      typedef double S;                           \
      __VA_ARGS__;                                \
    }                                             \
  } while (0)

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
void orc_free(void* h) { delete static_cast<Handle*>(h); }

// ---------------------------------------------------------------- particles
void* orc_pd_create(int prec, long n) {
  Handle* r = nullptr;
  auto mk = [&](auto tag) {
    typedef decltype(tag) S;
    auto* h = new PDH<S>();
    h->own.p.assign(n, V3<S>());
    h->own.v.assign(n, V3<S>());
    h->own.F.assign(n, M3<S>::identity());
    h->own.m.assign(n, S(0));
    h->own.volume.assign(n, S(0));
    r = h;
  };
  if (prec == 0) mk(float()); else mk(double());
  return r;
}
int orc_pd_set(void* h, const char* name, const double* a, long count) {
  return guard([&] { DISPATCH(h, setVar<S>(*static_cast<PDH<S>*>(h)->pd, name, a, count)); });
}
int orc_pd_get(void* h, const char* name, double* a, long count) {
  return guard([&] { DISPATCH(h, getVar<S>(*static_cast<PDH<S>*>(h)->pd, name, a, count)); });
}
long orc_pd_size(void* h) {
  long r = 0;
  DISPATCH(h, r = long(static_cast<PDH<S>*>(h)->pd->size()));
  return r;
}
int orc_pd_num_vars(void* h) {
  int r = 0;
  DISPATCH(h, r = int(static_cast<PDH<S>*>(h)->pd->numVariables()));
  return r;
}

// -------------------------------------------------------------------- model
void* orc_model_snow(int prec, double E, double nu, double xi, double thc, double ths) {
  if (prec == 0) {
    auto* h = new ModelH<float>();
    h->m.reset(new SnowModel<float>(float(E), float(nu), float(xi), float(thc), float(ths)));
    return h;
  }
  auto* h = new ModelH<double>();
  h->m.reset(new SnowModel<double>(E, nu, xi, thc, ths));
  return h;
}
void* orc_model_dummy(int prec) {
  if (prec == 0) { auto* h = new ModelH<float>(); h->m.reset(new DummyModel<float>()); return h; }
  auto* h = new ModelH<double>();
  h->m.reset(new DummyModel<double>());
  return h;
}
int orc_model_lame(void* h, double* mu_lambda) {
  return guard([&] {
    DISPATCH(h, {
      auto* sm = dynamic_cast<SnowModel<S>*>(static_cast<ModelH<S>*>(h)->m.get());
      if (!sm) throw std::runtime_error("not a snow model");
      mu_lambda[0] = sm->mu0;
      mu_lambda[1] = sm->lambda0;
    });
  });
}
int orc_model_set_particles(void* mh, void* pdh) {
  return guard([&] { DISPATCH(mh, static_cast<ModelH<S>*>(mh)->m->setParticles(*static_cast<PDH<S>*>(pdh)->pd)); });
}
int orc_model_update(void* mh) {
  return guard([&] { DISPATCH(mh, static_cast<ModelH<S>*>(mh)->m->updateParticleData()); });
}
int orc_model_energy(void* mh, long p, double* out) {
  return guard([&] { DISPATCH(mh, *out = double(static_cast<ModelH<S>*>(mh)->m->energyDensity(p))); });
}
int orc_model_dEdF(void* mh, long p, double* out) {
  return guard([&] {
    DISPATCH(mh, {
      M3<S> r = static_cast<ModelH<S>*>(mh)->m->dEnergyDensitydF(p);
      for (int i = 0; i < 9; ++i) out[i] = double(r.a[i]);
    });
  });
}
int orc_model_dEdF_diff(void* mh, const double* dF, long p, double* out) {
  return guard([&] {
    DISPATCH(mh, {
      M3<S> d;
      for (int i = 0; i < 9; ++i) d.a[i] = S(dF[i]);
      M3<S> r = static_cast<ModelH<S>*>(mh)->m->dEdFDifferential(d, p);
      for (int i = 0; i < 9; ++i) out[i] = double(r.a[i]);
    });
  });
}
// one SVD (for tests of the FP64 decomposition itself)
void orc_svd3(const double* A, double* U, double* s, double* V) { svd3_double(A, U, s, V); }
double orc_bspline_w(int prec, double x) { return prec == 0 ? double(bspline_w<float>(float(x))) : bspline_w<double>(x); }
double orc_bspline_dw(int prec, double x) { return prec == 0 ? double(bspline_dw<float>(float(x))) : bspline_dw<double>(x); }

// ---------------------------------------------------------- colliders/fields
void* orc_cols_create(int prec) { return prec == 0 ? (Handle*)new ColsH<float>() : (Handle*)new ColsH<double>(); }
int orc_cols_add_plane(void* h, const double* coeffs, const double* v, double mu, int sticky) {
  return guard([&] {
    DISPATCH(h, {
      Collider<S> c;
      c.type = 0;
      for (int i = 0; i < 4; ++i) c.coeffs[i] = S(coeffs[i]);
      c.vel = v3of<S>(v);
      c.friction = S(mu);
      c.sticky = sticky != 0;
      static_cast<ColsH<S>*>(h)->c.objects.push_back(c);
    });
  });
}
int orc_cols_add_sphere(void* h, const double* center, double radius, const double* v, double mu, int sticky) {
  return guard([&] {
    DISPATCH(h, {
      Collider<S> c;
      c.type = 1;
      c.center = v3of<S>(center);
      c.radius = S(radius);
      c.vel = v3of<S>(v);
      c.friction = S(mu);
      c.sticky = sticky != 0;
      static_cast<ColsH<S>*>(h)->c.objects.push_back(c);
    });
  });
}
// phi: dims0 dims1 dims2 values, point (i, j, k) at i + dims0 (j + dims1 k); pvel: 3 per point or null
int orc_cols_add_volume(void* h, const double* phi, const int* dims, const double* origin, double voxel,
                        const double* pvel, const double* v, double mu, int sticky) {
  return guard([&] {
    DISPATCH(h, {
      Collider<S> c;
      c.type = 2;
      size_t np = 1;
      for (int a = 0; a < 3; ++a) {
        if (dims[a] < 2) throw std::runtime_error("level-set dims must be >= 2");
        c.dims[a] = dims[a];
        np *= size_t(dims[a]);
      }
      c.origin = v3of<S>(origin);
      c.voxel = S(voxel);
      c.phiv.assign(phi, phi + np);
      if (pvel) c.velv.assign(pvel, pvel + 3 * np);
      c.vel = v3of<S>(v);
      c.friction = S(mu);
      c.sticky = sticky != 0;
      static_cast<ColsH<S>*>(h)->c.objects.push_back(c);
    });
  });
}
int orc_cols_collide(void* h, double* v, const double* x, const double* fv, int addVel) {
  int r = 0;
  DISPATCH(h, {
    V3<S> vv = v3of<S>(v);
    r = static_cast<ColsH<S>*>(h)->c.collide(vv, v3of<S>(x), v3of<S>(fv), addVel != 0);
    for (int c = 0; c < 3; ++c) v[c] = double(vv[c]);
  });
  return r;
}
void* orc_fields_create(int prec) { return prec == 0 ? (Handle*)new FieldsH<float>() : (Handle*)new FieldsH<double>(); }
int orc_fields_add_gravity(void* h, const double* g) {
  return guard([&] { DISPATCH(h, static_cast<FieldsH<S>*>(h)->f.gravity.push_back(v3of<S>(g))); });
}

// --------------------------------------------------------------------- grid
void* orc_grid_create(void* pdh, const int* inds, long count, double h, const double* fv, int dim) {
  Handle* out = nullptr;
  int rc = guard([&] {
    DISPATCH(pdh, {
      auto* gh = new GridH<S>();
      IndexList il(inds, inds + count);
      try {
        gh->g.reset(new Grid<S>(*static_cast<PDH<S>*>(pdh)->pd, il, S(h), v3of<S>(fv), dim));
      } catch (...) {
        delete gh;
        throw;
      }
      out = gh;
    });
  });
  return rc == 0 ? out : nullptr;
}
int orc_grid_info(void* gh, int* n, double* mn) {
  return guard([&] {
    DISPATCH(gh, {
      Grid<S>& g = *static_cast<GridH<S>*>(gh)->g;
      for (int a = 0; a < 3; ++a) { n[a] = g.n[a]; mn[a] = double(g.gmin[a]); }
    });
  });
}
int orc_grid_get(void* gh, const char* name, double* out) {
  return guard([&] {
    DISPATCH(gh, {
      Grid<S>& g = *static_cast<GridH<S>*>(gh)->g;
      std::string nm(name);
      if (nm == "masses") fromS(g.masses, out);
      else if (nm == "velocities") fromS(g.velocities, out);
      else if (nm == "prevVelocities") fromS(g.prevVelocities, out);
      else if (nm == "nodeCollided") { for (size_t i = 0; i < g.nodeCollided.size(); ++i) out[i] = g.nodeCollided[i]; }
      else throw std::runtime_error("unknown grid variable");
    });
  });
}
int orc_grid_set(void* gh, const char* name, const double* in) {
  return guard([&] {
    DISPATCH(gh, {
      Grid<S>& g = *static_cast<GridH<S>*>(gh)->g;
      std::string nm(name);
      if (nm == "masses") g.masses = toS<S>(in, g.masses.size());
      else if (nm == "velocities") g.velocities = toS<S>(in, g.velocities.size());
      else if (nm == "prevVelocities") g.prevVelocities = toS<S>(in, g.prevVelocities.size());
      else if (nm == "nodeCollided") { g.nodeCollided.resize(g.numNodes()); for (size_t i = 0; i < g.nodeCollided.size(); ++i) g.nodeCollided[i] = (signed char)in[i]; }
      else throw std::runtime_error("unknown grid variable");
    });
  });
}
long orc_grid_inds(void* gh, int* out) {
  long r = 0;
  DISPATCH(gh, {
    Grid<S>& g = *static_cast<GridH<S>*>(gh)->g;
    if (out) std::copy(g.inds.begin(), g.inds.end(), out);
    r = long(g.inds.size());
  });
  return r;
}
long orc_grid_partition(void* gh, int colour, long* beginEnd) {
  long r = 0;
  DISPATCH(gh, {
    Grid<S>& g = *static_cast<GridH<S>*>(gh)->g;
    const auto& P = g.parts[colour];
    if (beginEnd)
      for (size_t i = 0; i < P.size(); ++i) { beginEnd[2 * i] = long(P[i].first); beginEnd[2 * i + 1] = long(P[i].second); }
    r = long(P.size());
  });
  return r;
}
// the 64-node stencil of the shape-function iterator at point x (Grid.cpp:1105-1193):
// nodes[64*3] grid coordinates, w[64], dw[64*3] weight gradients, in iteration order
int orc_grid_stencil(void* gh, const double* x, int* nodes, double* w, double* dw) {
  return guard([&] {
    DISPATCH(gh, {
      Grid<S>& g = *static_cast<GridH<S>*>(gh)->g;
      ShapeIt<S> it;
      it.initialize(v3of<S>(x), g.gmin, g.h, g.dim, true);
      int k = 0, gp[3];
      do {
        it.gridPos(gp);
        V3<S> gr = it.grad();
        for (int a = 0; a < 3; ++a) { nodes[3 * k + a] = gp[a]; dw[3 * k + a] = double(gr[a]); }
        w[k] = double(it.wt());
        ++k;
      } while (it.next() && k < 64);
    });
  });
}
int orc_grid_volumes(void* gh) {
  return guard([&] { DISPATCH(gh, static_cast<GridH<S>*>(gh)->g->computeParticleVolumes()); });
}
int orc_grid_forces(void* gh, void* mh, void* fh, double* out) {
  return guard([&] {
    DISPATCH(gh, {
      Vec<S> f;
      static_cast<GridH<S>*>(gh)->g->calculateForces(f, *static_cast<ModelH<S>*>(mh)->m, static_cast<FieldsH<S>*>(fh)->f);
      fromS(f, out);
    });
  });
}
int orc_grid_force_diff(void* gh, void* mh, const double* dx, double* out) {
  return guard([&] {
    DISPATCH(gh, {
      Grid<S>& g = *static_cast<GridH<S>*>(gh)->g;
      Vec<S> d = toS<S>(dx, g.velocities.size()), df;
      g.calculateForceDifferentials(df, d, *static_cast<ModelH<S>*>(mh)->m);
      fromS(df, out);
    });
  });
}
int orc_grid_dfidxi(void* gh, void* mh, double* out) {
  return guard([&] {
    DISPATCH(gh, {
      Vec<S> r;
      static_cast<GridH<S>*>(gh)->g->dForceidXi(r, *static_cast<ModelH<S>*>(mh)->m);
      fromS(r, out);
    });
  });
}
int orc_grid_explicit_momenta(void* gh, double dt, void* mh, void* ch, void* fh, double* em, signed char* collided) {
  return guard([&] {
    DISPATCH(gh, {
      Grid<S>& g = *static_cast<GridH<S>*>(gh)->g;
      Vec<S> e;
      std::vector<signed char> c;
      g.calculateExplicitMomenta(e, c, S(dt), *static_cast<ModelH<S>*>(mh)->m, static_cast<ColsH<S>*>(ch)->c,
                                 static_cast<FieldsH<S>*>(fh)->f);
      fromS(e, em);
      std::copy(c.begin(), c.end(), collided);
    });
  });
}
int orc_grid_collision_velocities(void* gh, void* ch, const signed char* collided, double* vc) {
  return guard([&] {
    DISPATCH(gh, {
      Grid<S>& g = *static_cast<GridH<S>*>(gh)->g;
      std::vector<signed char> c(collided, collided + g.numNodes());
      Vec<S> v;
      g.collisionVelocities(v, static_cast<ColsH<S>*>(ch)->c, c);
      fromS(v, vc);
    });
  });
}
// stats: [iterations, bNorm2, earlyExit]; hist: ||r||^2 per termination call
int orc_grid_update_velocities(void* gh, double dt, void* mh, void* ch, void* fh, int maxIters, double tol,
                               double* stats, double* hist, int histCap, int* histLen, double* rhsOut,
                               double* guessOut) {
  return guard([&] {
    DISPATCH(gh, {
      Grid<S>& g = *static_cast<GridH<S>*>(gh)->g;
      Termination<S> t(maxIters, S(tol));
      SolveStats<S> st;
      Vec<S> rhs;
      V3<S> guess;
      g.updateGridVelocities(S(dt), *static_cast<ModelH<S>*>(mh)->m, static_cast<ColsH<S>*>(ch)->c,
                             static_cast<FieldsH<S>*>(fh)->f, t, &st, nullptr, rhsOut ? &rhs : nullptr, &guess);
      if (guessOut) for (int c = 0; c < 3; ++c) guessOut[c] = double(guess[c]);
      if (stats) { stats[0] = st.iterations; stats[1] = double(st.bNorm2); stats[2] = st.earlyExit; }
      int L = int(t.history.size());
      if (histLen) *histLen = L;
      if (hist) for (int i = 0; i < L && i < histCap; ++i) hist[i] = double(t.history[i]);
      if (rhsOut) fromS(rhs, rhsOut);
    });
  });
}
// updateGridVelocities with the reference's LinearSolver::Debug hook (TestGrid.cpp:398-453):
// every CR iterate x (velocities relative to the colliders, after each x update) is kept, the
// first iterCap of them in iters (3 n doubles each); *count = the number of hook calls
int orc_grid_update_velocities_iterates(void* gh, double dt, void* mh, void* ch, void* fh, int maxIters, double tol,
                                        double* iters, int iterCap, int* count) {
  return guard([&] {
    DISPATCH(gh, {
      Grid<S>& g = *static_cast<GridH<S>*>(gh)->g;
      Termination<S> t(maxIters, S(tol));
      SolveStats<S> st;
      const size_t nv = g.velocities.size();
      int calls = 0;
      std::function<void(Vec<S>&)> hook = [&](Vec<S>& x) {
        if (calls < iterCap)
          for (size_t i = 0; i < nv; ++i) iters[(size_t)calls * nv + i] = double(x[i]);
        ++calls;
      };
      g.updateGridVelocities(S(dt), *static_cast<ModelH<S>*>(mh)->m, static_cast<ColsH<S>*>(ch)->c,
                             static_cast<FieldsH<S>*>(fh)->f, t, &st, &hook, nullptr);
      if (count) *count = calls;
    });
  });
}
int orc_grid_update_F(void* gh, double dt) {
  return guard([&] { DISPATCH(gh, static_cast<GridH<S>*>(gh)->g->updateDeformationGradients(S(dt))); });
}
int orc_grid_update_pv(void* gh) {
  return guard([&] { DISPATCH(gh, static_cast<GridH<S>*>(gh)->g->updateParticleVelocities()); });
}

int orc_voxel_sort(int prec, int* inds, long n, double voxelSize, const double* x) {
  return guard([&] {
    auto run = [&](auto tag) {
      typedef decltype(tag) S;
      std::vector<V3<S>> pts(size_t(3 * n) / 3);
      long np = 0;
      for (long i = 0; i < n; ++i) np = std::max<long>(np, inds[i] + 1);
      pts.resize(np);
      for (long i = 0; i < np; ++i) pts[i] = v3of<S>(x + 3 * i);
      IndexList il(inds, inds + n);
      voxelSort<S>(il.begin(), il.end(), S(voxelSize), pts);
      std::copy(il.begin(), il.end(), inds);
    };
    if (prec == 0) run(float()); else run(double());
  });
}

// CR on a dense column-major matrix; diag (optional) is a diagonal preconditioner.
int orc_cr_dense(int prec, int N, const double* A, const double* b, double* x, const double* diag, int maxIters,
                 double tol, double* stats, double* logR, double* logP, int logCap, int* logLen) {
  return guard([&] {
    auto run = [&](auto tag) {
      typedef decltype(tag) S;
      DenseMat<S> M;
      M.N = N;
      M.A = toS<S>(A, size_t(N) * N);
      DiagMat<S> D;
      if (diag) D.d = toS<S>(diag, N);
      Vec<S> bb = toS<S>(b, N), xx = toS<S>(x, N);
      Termination<S> t(maxIters, S(tol));
      SolveStats<S> st;
      std::vector<Vec<S>> lr, lp;
      conjugateResiduals<S>(M, bb, xx, t, diag ? &D : nullptr, &st, &lr, &lp);
      fromS(xx, x);
      if (stats) { stats[0] = st.iterations; stats[1] = double(st.bNorm2); stats[2] = st.earlyExit; }
      if (logLen) *logLen = int(lr.size());
      for (int i = 0; i < int(lr.size()) && i < logCap; ++i) {
        if (logR) fromS(lr[i], logR + size_t(i) * N);
        if (logP) fromS(lp[i], logP + size_t(i) * N);
      }
    };
    if (prec == 0) run(float()); else run(double());
  });
}

// ---------------------------------------------------------------------- sim
void* orc_sim_create(int prec, long n, const double* x, const double* m, double h, void* mh, void* ch, void* fh, int dim,
                     const double* mu0, const double* lambda0, const double* xi) {
  Handle* out = nullptr;
  guard([&] {
    auto run = [&](auto tag) {
      typedef decltype(tag) S;
      std::vector<V3<S>> pts(n);
      for (long i = 0; i < n; ++i) pts[i] = v3of<S>(x + 3 * i);
      std::vector<S> ms = toS<S>(m, n);
      std::vector<S> a, b, c;
      if (mu0) { a = toS<S>(mu0, n); b = toS<S>(lambda0, n); c = toS<S>(xi, n); }
      auto* sh = new SimH<S>();
      try {
        sh->s.reset(new Sim<S>(pts, ms, S(h), *static_cast<ModelH<S>*>(mh)->m, static_cast<ColsH<S>*>(ch)->c,
                               static_cast<FieldsH<S>*>(fh)->f, dim, mu0 ? &a : nullptr, mu0 ? &b : nullptr,
                               mu0 ? &c : nullptr));
      } catch (...) {
        delete sh;
        throw;
      }
      sh->view.pd = &sh->s->d;
      out = sh;
    };
    if (prec == 0) run(float()); else run(double());
  });
  return out;
}
void* orc_sim_pd(void* sh) {
  void* r = nullptr;
  DISPATCH(sh, r = &static_cast<SimH<S>*>(sh)->view);
  return r;
}
int orc_sim_advance(void* sh, double dt, int maxIters, double tol) {
  return guard([&] {
    DISPATCH(sh, {
      Termination<S> t(maxIters, S(tol));
      static_cast<SimH<S>*>(sh)->s->advance(S(dt), t);
    });
  });
}
// the same with a termination whose cancelled() turns true on its cancelOnCall-th call; returns 1
// if the step was cancelled
int orc_sim_advance_cancel(void* sh, double dt, int maxIters, double tol, int cancelOnCall, int* wasCancelled) {
  return guard([&] {
    DISPATCH(sh, {
      Termination<S> t(maxIters, S(tol));
      t.cancelOnCall = cancelOnCall;
      static_cast<SimH<S>*>(sh)->s->advance(S(dt), t);
      *wasCancelled = t.cancelled ? 1 : 0;
    });
  });
}
// out: [iterations, bNorm2, earlyExit, fv0, fv1, fv2, n0, n1, n2, particles]; returns residual history length
int orc_sim_body_stats(void* sh, long body, double* out, double* hist, int cap) {
  int L = -1;
  guard([&] {
    DISPATCH(sh, {
      const auto& st = static_cast<SimH<S>*>(sh)->s->lastStats.at(body);
      out[0] = st.solve.iterations;
      out[1] = double(st.solve.bNorm2);
      out[2] = st.solve.earlyExit;
      for (int a = 0; a < 3; ++a) { out[3 + a] = double(st.frameVelocity[a]); out[6 + a] = st.n[a]; }
      out[9] = double(st.particles);
      L = int(st.solve.residualNorm2.size());
      if (hist) for (int i = 0; i < L && i < cap; ++i) hist[i] = double(st.solve.residualNorm2[i]);
    });
  });
  return L;
}
long orc_sim_num_stats(void* sh) {
  long r = 0;
  DISPATCH(sh, r = long(static_cast<SimH<S>*>(sh)->s->lastStats.size()));
  return r;
}
long orc_sim_num_bodies(void* sh) {
  long r = 0;
  DISPATCH(sh, r = long(static_cast<SimH<S>*>(sh)->s->bodies.size()));
  return r;
}
long orc_sim_body(void* sh, long b, int* out) {
  long r = 0;
  DISPATCH(sh, {
    const auto& B = static_cast<SimH<S>*>(sh)->s->bodies.at(b);
    if (out) std::copy(B.begin(), B.end(), out);
    r = long(B.size());
  });
  return r;
}
long orc_sim_ballistic(void* sh, int* out) {
  long r = 0;
  DISPATCH(sh, {
    const auto& B = static_cast<SimH<S>*>(sh)->s->ballistic;
    if (out) std::copy(B.begin(), B.end(), out);
    r = long(B.size());
  });
  return r;
}
int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
