// mpm_runtime.hip -- host runtime behind include/mpmsnow_gpu.h.
//
// Owns the device particle store (double-buffered SoA columns kept in body-
// and tile-sorted slot order), per-body tiled grids, and the orchestration of
// Sim::advance (davidsminor/mpmsnow src/MpmSim/Sim.cpp:57-134) and
// Grid::updateGridVelocities (src/MpmSim/Grid.cpp:889-1016) as sequences of
// gfx950 kernel launches on one HIP stream. The conjugate-residual loop runs
// device-resident: each kernel skips itself once the solve is done, and the
// host checks the 4-byte done flag once per batch of iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mpmsnow_gpu.h"
#include "mpm_kernels.hpp"

using namespace mpm;

namespace {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};
struct InvalidArg : std::runtime_error {
  explicit InvalidArg(const std::string& s) : std::runtime_error(s) {}
};

#define HIPCHECK(x)                                                                   \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) throw HipError(std::string("hip: ") + hipGetErrorName(e_)); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    if (p) HIPCHECK(hipFree(p));
    p = nullptr;
    size_t c = std::max<size_t>(bytes, 256);
    HIPCHECK(hipMalloc(&p, c));
    cap = c;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

struct PinnedBuf {
  void* p = nullptr;
  size_t cap = 0;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    if (p) HIPCHECK(hipHostFree(p));
    p = nullptr;
    HIPCHECK(hipHostMalloc(&p, bytes));
    cap = bytes;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

struct VarInfo {
  int col;
  int ncomp;
  bool model;
};
const std::unordered_map<std::string, VarInfo>& var_table() {
  static const std::unordered_map<std::string, VarInfo> t = {
      {"p", {C_X, 3, false}},      {"v", {C_V, 3, false}},     {"F", {C_F, 9, false}},
      {"m", {C_M, 1, false}},      {"volume", {C_VOL, 1, false}}, {"Fp", {C_FP, 9, true}},
      {"FinvTrans", {C_FI, 9, true}}, {"R", {C_R, 9, true}},   {"Ginv", {C_GI, 9, true}},
      {"J", {C_J, 1, true}},       {"mu", {C_MU, 1, true}},    {"lambda", {C_LAM, 1, true}},
      {"mu0", {C_MU0, 1, true}},   {"lambda0", {C_LAM0, 1, true}}, {"xi", {C_XI, 1, true}}};
  return t;
}

float decode_ord(unsigned int u) {
  unsigned int v = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
  float f;
  std::memcpy(&f, &v, 4);
  return f;
}

}  // namespace

// Kernel timing for measurement: HIP events recorded on the library's stream around
// each instrumented launch, resolved at the next synchronisation point. Algorithmic
// bytes per launch = bytes per particle * particles + bytes per node * nodes with m > 0.
enum ProfKind { PK_P2G_MASS, PK_P2G_VEL, PK_FORCES, PK_HVP, PK_DIAG, PK_G2P, PK_CR_X, PK_CR_AR, PK_CR_P, PK_COUNT };
const char* const kProfNames[PK_COUNT] = {"p2g_mass", "p2g_vel", "forces", "hvp", "diag",
                                          "g2p", "cr_update_x", "cr_ar", "cr_update_p"};
const double kProfBytesPerParticle[PK_COUNT] = {16, 28, 136, 172, 172, 324, 0, 0, 0};
const double kProfBytesPerNode[PK_COUNT] = {4, 16, 12, 29, 12, 24, 85, 41, 84};

struct SolveRecord {
  int iterations = 0;
  int early_exit = 0;
  double rhs_norm2 = 0, residual_norm2 = 0;
  std::vector<double> history;
};

struct mpm_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int64_t n = 0;
  bool model = false;
  float E = 0, nu = 0, xi = 0, thc = 0, ths = 0, mu0 = 0, lambda0 = 0;
  bool perParticleMaterial = false;
  DevBuf store[2];
  int cur = 0;
  DevBuf orig_of[2], slot_of;
  DevBuf cols_dev, grav_dev, errflag, staging, idx_tmp, slots_tmp;
  ColliderSet hcols{};
  std::vector<float> hgrav;
  PinnedBuf pinned;
  // simulation state
  bool simInit = false;
  float h = 0;
  int dim = 3;
  std::vector<std::vector<int>> bodies;  // original indices, ascending
  std::vector<int> ballistic;
  std::vector<SolveRecord> lastSolves;
  std::vector<std::unique_ptr<mpm_grid>> gridPool;
  DevBuf new_to_old, ball_slots, other_slots;
  int iterHint = 8;
  // measurement
  struct ProfRec {
    int kind;
    hipEvent_t a, b;
    const mpm_grid* grid;
  };
  bool profOn = false;
  std::vector<ProfRec> profPending;
  std::vector<hipEvent_t> eventPool;
  double profMs[PK_COUNT] = {};
  int64_t profCount[PK_COUNT] = {};
  double profBytes[PK_COUNT] = {};
  hipEvent_t new_event() {
    if (!eventPool.empty()) {
      hipEvent_t e = eventPool.back();
      eventPool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return e;
  }
  int prof_begin(int kind, const mpm_grid* g) {
    if (!profOn) return -1;
    ProfRec r{kind, new_event(), new_event(), g};
    HIPCHECK(hipEventRecord(r.a, stream));
    profPending.push_back(r);
    return (int)profPending.size() - 1;
  }
  void prof_end(int idx) {
    if (idx < 0) return;
    HIPCHECK(hipEventRecord(profPending[idx].b, stream));
  }
  void prof_resolve();
  ~mpm_context() {
    for (auto& r : profPending) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : eventPool) (void)hipEventDestroy(e);
  }

  PView view(int which = -1) const {
    int w = which < 0 ? cur : which;
    PView v;
    v.base = store[w].as<float>();
    v.stride = n;
    return v;
  }
  void sync() { HIPCHECK(hipStreamSynchronize(stream)); }
  const ColliderSet* cols() const { return cols_dev.as<ColliderSet>(); }
  int* errp() const { return errflag.as<int>(); }
  void clear_err() { HIPCHECK(hipMemsetAsync(errflag.p, 0, sizeof(int), stream)); }
  void check_errors(int mask = ~0) {
    int e = 0;
    HIPCHECK(hipMemcpyAsync(&e, errflag.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    sync();
    raise(e & mask);
  }
  static void raise(int e) {
    if (e & ERR_NAN_MASS) throw std::runtime_error("nans in splatted masses!");
    if (e & ERR_NAN_VEL) throw std::runtime_error("nans in splatted velocities!");
    if (e & ERR_NAN_STRESS) throw std::runtime_error("nans in dEnergyDensitydF matrix!");
    if (e & ERR_NAN_MOMENTA) throw std::runtime_error("nan in explicit momenta!");
    if (e & ERR_HARDNESS) throw std::runtime_error("infinite hardness!!!");
    if (e & ERR_NAN_PVEL) throw std::runtime_error("nans in particle velocity data!");
    if (e & ERR_STENCIL) throw std::runtime_error("particle outside its grid tile (internal binning error)");
  }
  // upload original indices and map them to slots (device int array of count entries)
  int* slots_from_orig(const int32_t* orig, int64_t count, DevBuf& out) {
    out.ensure(sizeof(int) * std::max<int64_t>(count, 1));
    idx_tmp.ensure(sizeof(int) * std::max<int64_t>(count, 1));
    for (int64_t i = 0; i < count; ++i)
      if (orig[i] < 0 || orig[i] >= n) throw InvalidArg("particle index out of range");
    if (count > 0) {
      HIPCHECK(hipMemcpyAsync(idx_tmp.p, orig, sizeof(int) * count, hipMemcpyHostToDevice, stream));
      launch_gather_slots(slot_of.as<int>(), idx_tmp.as<int>(), (int)count, out.as<int>(), stream);
    }
    return out.as<int>();
  }
};

struct mpm_grid {
  mpm_context* ctx = nullptr;
  GridView g{};
  int np = 0;
  int activeNodes = 0;
  float fv[3] = {0, 0, 0};
  DevBuf slots, perm, key, rank, count, offset, nch, chunk_off, chunks, scantmp, totals;
  DevBuf mass, vel, prev, flags, vc, b, x, z, p, Ap, Ar, acc, acc2, diag, draw, sc, partials, hist, dense, dense2;
  DevBuf bbox, mom;
  int64_t NN = 0;

  float* f(DevBuf& d) { return d.as<float>(); }
  size_t vecBytes() const { return sizeof(float) * 3 * (size_t)NN; }
  CRScalars* scal() { return sc.as<CRScalars>(); }

  // Grid::Grid (Grid.cpp:156-263) on a device slot list of np particles
  void build(const int* slotList, int npart, float h, const float* frameV, bool computeFrame) {
    mpm_context& c = *ctx;
    hipStream_t s = c.stream;
    np = npart;
    if (npart <= 0) throw InvalidArg("grid needs at least one particle");
    // body reductions: bbox and comoving frame (Sim.cpp:84-93)
    bbox.ensure(6 * sizeof(unsigned int));
    mom.ensure(4 * sizeof(double));
    unsigned int init[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    HIPCHECK(hipMemcpyAsync(bbox.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemsetAsync(mom.p, 0, 4 * sizeof(double), s));
    launch_body_reduce(c.view(), slotList, 0, npart, bbox.as<unsigned int>(), mom.as<double>(), s);
    struct {
      unsigned int bb[6];
      unsigned int pad[2];
      double mm[4];
    } hb;
    HIPCHECK(hipMemcpyAsync(hb.bb, bbox.p, 6 * sizeof(unsigned int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(hb.mm, mom.p, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    c.sync();
    if (computeFrame) {
      for (int a = 0; a < 3; ++a) fv[a] = (float)(hb.mm[a] / hb.mm[3]);
    } else {
      for (int a = 0; a < 3; ++a) fv[a] = frameV ? frameV[a] : 0.0f;
    }
    // grid geometry (Grid.cpp:193-213), float arithmetic as the reference
    GridView G{};
    G.h = h;
    for (int a = 0; a < 3; ++a) {
      float mn = decode_ord(hb.bb[a]), mx = decode_ord(hb.bb[3 + a]);
      int cellMin = (int)std::floor(mn / h) - 2 - 1;
      int cellMax = (int)std::ceil(mx / h) + 2 + 1;
      G.gmin[a] = (float)cellMin * h;
      G.n[a] = cellMax - cellMin;
      G.fv[a] = fv[a];
    }
    long long ncells = (long long)G.n[0] * G.n[1] * G.n[2];
    if (ncells > 4000000000LL || ncells <= 0) throw std::runtime_error("grid is too big");
    for (int a = 0; a < 3; ++a) G.T[a] = (G.n[a] + 3) / 4;
    long long nt = (long long)G.T[0] * G.T[1] * G.T[2];
    if (nt * 64 * 3 > (long long)INT32_MAX * 2) throw std::runtime_error("grid is too big");
    G.NT = (int)nt;
    NN = nt * 64;
    // binning + counting sort by tile
    slots.ensure(sizeof(int) * npart);
    HIPCHECK(hipMemcpyAsync(slots.p, slotList, sizeof(int) * npart, hipMemcpyDeviceToDevice, s));
    key.ensure(sizeof(int) * npart);
    rank.ensure(sizeof(int) * npart);
    perm.ensure(sizeof(int) * npart);
    count.ensure(sizeof(int) * G.NT);
    offset.ensure(sizeof(int) * G.NT);
    nch.ensure(sizeof(int) * G.NT);
    chunk_off.ensure(sizeof(int) * G.NT);
    scantmp.ensure(sizeof(int) * scan_tmp_ints(G.NT));
    totals.ensure(sizeof(int) * 4);
    HIPCHECK(hipMemsetAsync(count.p, 0, sizeof(int) * G.NT, s));
    c.clear_err();
    launch_bin(c.view(), slots.as<int>(), 0, npart, G, key.as<int>(), rank.as<int>(), count.as<int>(), c.errp(), s);
    scan_exclusive(count.as<int>(), offset.as<int>(), G.NT, totals.as<int>(), scantmp.as<int>(), s);
    launch_chunk_counts(count.as<int>(), G.NT, nch.as<int>(), s);
    scan_exclusive(nch.as<int>(), chunk_off.as<int>(), G.NT, totals.as<int>() + 1, scantmp.as<int>(), s);
    launch_scatter(slots.as<int>(), 0, key.as<int>(), rank.as<int>(), offset.as<int>(), npart, perm.as<int>(), s);
    int tot[2] = {0, 0};
    int e = 0;
    HIPCHECK(hipMemcpyAsync(tot, totals.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(&e, c.errflag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    c.sync();
    mpm_context::raise(e);
    if (tot[0] != npart) throw std::runtime_error("particle binning lost particles");
    chunks.ensure(sizeof(int4) * std::max(tot[1], 1));
    launch_make_chunks(count.as<int>(), offset.as<int>(), chunk_off.as<int>(), G.NT, chunks.as<int4>(), s);
    G.perm = perm.as<int>();
    G.chunks = chunks.as<int4>();
    G.nchunks = tot[1];
    g = G;
    // node arrays
    mass.ensure(sizeof(float) * NN);
    flags.ensure(NN);
    for (DevBuf* d : {&vel, &prev, &vc, &b, &x, &z, &p, &Ap, &Ar, &acc, &acc2, &diag, &draw}) d->ensure(vecBytes());
    sc.ensure(sizeof(CRScalars));
    partials.ensure(sizeof(double) * 8 * 4096);
    HIPCHECK(hipMemsetAsync(mass.p, 0, sizeof(float) * NN, s));
    HIPCHECK(hipMemsetAsync(vel.p, 0, vecBytes(), s));
    HIPCHECK(hipMemsetAsync(flags.p, 0xFF, NN, s));  // -1: no collision until updateGridVelocities
    HIPCHECK(hipMemsetAsync(sc.p, 0, sizeof(CRScalars), s));
    // MassSplatter, clamp/NaN, VelocitySplatter, NaN, prevVelocities
    int pr = c.prof_begin(PK_P2G_MASS, this);
    launch_p2g_mass(c.view(), g, f(mass), c.errp(), s);
    c.prof_end(pr);
    launch_mass_fix(f(mass), g, c.errp(), s);
    pr = c.prof_begin(PK_P2G_VEL, this);
    launch_p2g_vel(c.view(), g, f(mass), f(vel), c.errp(), s);
    c.prof_end(pr);
    launch_check_vel(f(vel), f(prev), g, c.errp(), s);
  }

  void to_dense(const float* tiled, int ncomp, float* host) {
    int64_t nd = (int64_t)g.n[0] * g.n[1] * g.n[2];
    dense.ensure(sizeof(float) * ncomp * nd);
    launch_to_dense_f(tiled, ncomp, dense.as<float>(), g, ctx->stream);
    HIPCHECK(hipMemcpyAsync(host, dense.p, sizeof(float) * ncomp * nd, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  }
  void from_dense(const float* host, int ncomp, float* tiled) {
    int64_t nd = (int64_t)g.n[0] * g.n[1] * g.n[2];
    dense.ensure(sizeof(float) * ncomp * nd);
    HIPCHECK(hipMemcpyAsync(dense.p, host, sizeof(float) * ncomp * nd, hipMemcpyHostToDevice, ctx->stream));
    if (ncomp == 3) HIPCHECK(hipMemsetAsync(tiled, 0, vecBytes(), ctx->stream));
    launch_from_dense_f(dense.as<float>(), ncomp, tiled, g, ctx->stream);
  }

  const float* grav() const { return ctx->grav_dev.as<float>(); }
  int ngrav() const { return (int)(ctx->hgrav.size() / 3); }

  // internal forces into acc (zeroed first)
  void forces_into_acc() {
    HIPCHECK(hipMemsetAsync(acc.p, 0, vecBytes(), ctx->stream));
    int pr = ctx->prof_begin(PK_FORCES, this);
    launch_forces(ctx->view(), g, f(acc), ctx->errp(), ctx->stream);
    ctx->prof_end(pr);
  }
  void hvp(const float* d, bool project, float* out, const CRScalars* scp) {
    int pr = ctx->prof_begin(PK_HVP, this);
    launch_hvp(ctx->view(), g, d, project, flags.as<int8_t>(), ctx->cols(), out, scp, ctx->errp(), ctx->stream);
    ctx->prof_end(pr);
  }

  // Grid::updateGridVelocities (Grid.cpp:889-1016)
  void updateGridVelocities(float dt, const mpm_solver_params* prm, SolveRecord& rec) {
    mpm_context& c = *ctx;
    hipStream_t s = c.stream;
    const bool hooks = prm && (prm->termination || prm->debug);
    int maxIters = prm ? prm->max_iters : 0;
    float tol = prm ? prm->tolerance : 0.0f;
    if (prm && prm->termination) {
      maxIters = INT32_MAX;
      tol = 0.0f;
    }
    // scalars: zero everything, set max_iters
    CRScalars init{};
    init.max_iters = maxIters;
    c.pinned.ensure(4096);
    std::memcpy(c.pinned.p, &init, sizeof(init));
    HIPCHECK(hipMemcpyAsync(sc.p, c.pinned.p, sizeof(init), hipMemcpyHostToDevice, s));
    c.sync();
    hist.ensure(sizeof(double) * 4096);
    c.clear_err();
    // explicit momenta, flags, vc, guess sums
    forces_into_acc();
    launch_node_explicit(g, f(mass), f(vel), f(acc), grav(), ngrav(), dt, c.cols(), f(b), flags.as<int8_t>(), f(vc),
                         scal(), partials.as<double>(), s);
    // DF(vc) (no input projection) and the diagonal
    HIPCHECK(hipMemsetAsync(acc2.p, 0, vecBytes(), s));
    hvp(f(vc), false, f(acc2), nullptr);
    HIPCHECK(hipMemsetAsync(draw.p, 0, vecBytes(), s));
    int pr = c.prof_begin(PK_DIAG, this);
    launch_diag(c.view(), g, f(draw), c.errp(), s);
    c.prof_end(pr);
    launch_rhs(g, f(mass), flags.as<int8_t>(), c.cols(), f(acc2), f(draw), dt, tol, f(b), f(x), f(diag), scal(),
               partials.as<double>(), s);
    // CR setup: A x, A z (ConjugateResiduals.cpp:50-76)
    HIPCHECK(hipMemsetAsync(acc.p, 0, vecBytes(), s));
    hvp(f(x), true, f(acc), scal());
    launch_cr_setup1(g, f(mass), flags.as<int8_t>(), c.cols(), f(acc), f(x), f(b), f(diag), dt, f(z), f(p), scal(), s);
    HIPCHECK(hipMemsetAsync(acc.p, 0, vecBytes(), s));
    hvp(f(z), true, f(acc), scal());
    launch_cr_setup2(g, f(mass), flags.as<int8_t>(), c.cols(), f(acc), f(z), f(diag), dt, f(Ap), scal(),
                     partials.as<double>(), s);
    CRScalars* hsc = c.pinned.as<CRScalars>();
    auto fetch = [&]() {
      HIPCHECK(hipMemcpyAsync(hsc, sc.p, sizeof(CRScalars), hipMemcpyDeviceToHost, s));
      c.sync();
    };
    fetch();
    activeNodes = hsc->active_nodes;
    mpm_context::raise(hsc->errors);
    int e0 = 0;
    HIPCHECK(hipMemcpyAsync(&e0, c.errflag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    c.sync();
    mpm_context::raise(e0);
    std::vector<float> hostvec;
    int enqueued = 0;
    const int histCap = 4096;
    const int64_t iterCap = (prm && prm->termination) ? 1000000 : (int64_t)maxIters + 2;
    while (!hsc->done) {
      if (enqueued > iterCap) throw std::runtime_error("conjugate residual did not terminate");
      int batch = hooks ? 1 : (enqueued == 0 ? std::max(1, c.iterHint) : 8);
      for (int k = 0; k < batch; ++k) {
        int px = c.prof_begin(PK_CR_X, this);
        launch_cr_update_x(g, flags.as<int8_t>(), c.cols(), f(p), f(Ap), f(diag), f(x), f(z), scal(),
                           partials.as<double>(), hist.as<double>(), histCap, s);
        c.prof_end(px);
        if (hooks) {
          fetch();
          int it = hsc->iterations - 1;
          if (prm->debug) {
            hostvec.resize((size_t)3 * g.n[0] * g.n[1] * g.n[2]);
            to_dense(f(x), 3, hostvec.data());
            prm->debug(prm->user, hostvec.data(), (int64_t)hostvec.size());
          }
          if (prm->termination && !hsc->done) {
            launch_residual_vector(g, f(z), f(diag), f(acc2), s);
            hostvec.resize((size_t)3 * g.n[0] * g.n[1] * g.n[2]);
            to_dense(f(acc2), 3, hostvec.data());
            if (prm->termination(prm->user, hostvec.data(), (int64_t)hostvec.size(), it)) {
              int one = 1;
              HIPCHECK(hipMemcpyAsync(&scal()->done, &one, sizeof(int), hipMemcpyHostToDevice, s));
              c.sync();
            }
          }
        }
        HIPCHECK(hipMemsetAsync(acc.p, 0, vecBytes(), s));
        hvp(f(z), true, f(acc), scal());
        int pa = c.prof_begin(PK_CR_AR, this);
        launch_cr_ar(g, f(mass), flags.as<int8_t>(), c.cols(), f(acc), f(z), dt, f(Ar), scal(), partials.as<double>(),
                     s);
        c.prof_end(pa);
        int pp = c.prof_begin(PK_CR_P, this);
        launch_cr_update_p(g, f(z), f(Ar), f(diag), f(p), f(Ap), scal(), partials.as<double>(), s);
        c.prof_end(pp);
      }
      enqueued += batch;
      fetch();
    }
    launch_finalize(g, f(x), f(vc), f(vel), scal(), s);
    fetch();
    rec.iterations = hsc->iterations;
    rec.early_exit = hsc->early_exit;
    rec.rhs_norm2 = hsc->bnorm2;
    rec.residual_norm2 = hsc->rnorm2;
    int L = std::min(hsc->iterations, histCap);
    rec.history.resize(L);
    if (L > 0) {
      HIPCHECK(hipMemcpyAsync(rec.history.data(), hist.p, sizeof(double) * L, hipMemcpyDeviceToHost, s));
      c.sync();
    }
    if (rec.iterations > 0) c.iterHint = rec.iterations;
  }
};

void mpm_context::prof_resolve() {
  if (profPending.empty()) return;
  sync();
  for (auto& r : profPending) {
    float ms = 0.0f;
    hipError_t e = hipEventElapsedTime(&ms, r.a, r.b);
    // launches skipped after the solve finished still count: they are part of the solve's cost
    if (e == hipSuccess) {
      profMs[r.kind] += ms;
      profCount[r.kind] += 1;
      profBytes[r.kind] += kProfBytesPerParticle[r.kind] * (r.grid ? r.grid->np : 0) +
                           kProfBytesPerNode[r.kind] * (r.grid ? r.grid->activeNodes : 0);
    }
    eventPool.push_back(r.a);
    eventPool.push_back(r.b);
  }
  profPending.clear();
}

namespace {

template <class F> int guarded(mpm_context* ctx, F fn) {
  if (!ctx) return MPM_ERR_INVALID;
  try {
    HIPCHECK(hipSetDevice(ctx->device));
    fn();
    ctx->err.clear();
    return MPM_OK;
  } catch (const HipError& e) {
    ctx->err = e.what();
    return MPM_ERR_HIP;
  } catch (const InvalidArg& e) {
    ctx->err = e.what();
    return MPM_ERR_INVALID;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return MPM_ERR_RUNTIME;
  }
}

void require_particles(mpm_context* c) {
  if (c->n <= 0) throw InvalidArg("no particles (call mpm_particles_create first)");
}

// Sim::calculateBodies (Sim.cpp:136-203): connected components of the relation
// |x_i - x_j|^2 < h^2 (float arithmetic of KdTree.inl:187-190); isolated particles are
// ballistic; bodies are ordered by their smallest particle index.
void calculate_bodies(mpm_context& c) {
  int n = (int)c.n;
  std::vector<float> x((size_t)3 * n);
  {
    DevBuf tmp;
    tmp.ensure(sizeof(float) * 3 * n);
    // slot_of maps original -> slot; get_column gathers in original order
    launch_get_column(c.view(), C_X, 3, tmp.as<float>(), c.slot_of.as<int>(), n, c.stream);
    HIPCHECK(hipMemcpyAsync(x.data(), tmp.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c.stream));
    c.clear_err();
    launch_check_pvel(c.view(), n, c.errp(), c.stream);
    c.check_errors();
  }
  const float h = c.h;
  const float r2 = h * h;
  // hash grid with cell size h
  struct Key {
    int64_t k;
  };
  auto cellOf = [&](int i, int a) { return (int64_t)std::floor(x[3 * i + a] / h); };
  auto pack = [](int64_t i, int64_t j, int64_t k) {
    return (uint64_t)((i & 0x1FFFFF) | ((j & 0x1FFFFF) << 21) | ((k & 0x1FFFFF) << 42));
  };
  std::vector<std::pair<uint64_t, int>> keys(n);
  for (int i = 0; i < n; ++i) keys[i] = {pack(cellOf(i, 0), cellOf(i, 1), cellOf(i, 2)), i};
  std::sort(keys.begin(), keys.end());
  std::unordered_map<uint64_t, std::pair<int, int>> cells;
  cells.reserve(keys.size() * 2);
  for (int a = 0; a < n;) {
    int b = a;
    while (b < n && keys[b].first == keys[a].first) ++b;
    cells[keys[a].first] = {a, b};
    a = b;
  }
  std::vector<int> parent(n);
  for (int i = 0; i < n; ++i) parent[i] = i;
  auto find = [&](int i) {
    while (parent[i] != i) {
      parent[i] = parent[parent[i]];
      i = parent[i];
    }
    return i;
  };
  std::vector<char> hasNb(n, 0);
  for (int i = 0; i < n; ++i) {
    int64_t ci = cellOf(i, 0), cj = cellOf(i, 1), ck = cellOf(i, 2);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          auto it = cells.find(pack(ci + dx, cj + dy, ck + dz));
          if (it == cells.end()) continue;
          for (int q = it->second.first; q < it->second.second; ++q) {
            int j = keys[q].second;
            if (j <= i) continue;
            float d0 = x[3 * i] - x[3 * j], d1 = x[3 * i + 1] - x[3 * j + 1], d2 = x[3 * i + 2] - x[3 * j + 2];
            float d = d0 * d0 + d1 * d1 + d2 * d2;
            if (d < r2) {
              hasNb[i] = hasNb[j] = 1;
              int a = find(i), b = find(j);
              if (a != b) {
                if (a < b) parent[b] = a;
                else parent[a] = b;
              }
            }
          }
        }
  }
  c.bodies.clear();
  c.ballistic.clear();
  std::unordered_map<int, int> rootBody;
  for (int i = 0; i < n; ++i) {
    if (!hasNb[i]) {
      c.ballistic.push_back(i);
      continue;
    }
    int r = find(i);
    auto it = rootBody.find(r);
    if (it == rootBody.end()) {
      rootBody[r] = (int)c.bodies.size();
      c.bodies.push_back({i});
    } else {
      c.bodies[it->second].push_back(i);
    }
  }
}

mpm_grid* pooled_grid(mpm_context& c, size_t b) {
  while (c.gridPool.size() <= b) {
    c.gridPool.emplace_back(new mpm_grid());
    c.gridPool.back()->ctx = &c;
  }
  return c.gridPool[b].get();
}

}  // namespace

extern "C" {

int mpm_version(void) { return 1; }

int mpm_create(mpm_context** out, int device) {
  if (!out) return MPM_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return MPM_ERR_NO_DEVICE;
  std::unique_ptr<mpm_context> c(new mpm_context());
  c->device = device;
  int rc = guarded(c.get(), [&] {
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->errflag.ensure(sizeof(int));
    c->cols_dev.ensure(sizeof(ColliderSet));
    c->grav_dev.ensure(sizeof(float) * 3 * 64);
    ColliderSet empty{};
    HIPCHECK(hipMemcpy(c->cols_dev.p, &empty, sizeof(empty), hipMemcpyHostToDevice));
  });
  if (rc != MPM_OK) return rc;
  *out = c.release();
  return MPM_OK;
}

int mpm_destroy(mpm_context* ctx) {
  if (!ctx) return MPM_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->gridPool.clear();
  hipStream_t s = ctx->stream;
  delete ctx;
  if (s) (void)hipStreamDestroy(s);
  return MPM_OK;
}

const char* mpm_last_error(const mpm_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mpm_particles_create(mpm_context* ctx, int64_t n) {
  return guarded(ctx, [&] {
    if (n <= 0 || n > INT32_MAX / 2) throw InvalidArg("particle count out of range");
    ctx->n = n;
    ctx->model = false;
    ctx->simInit = false;
    ctx->perParticleMaterial = false;
    for (int w = 0; w < 2; ++w) {
      ctx->store[w].ensure(sizeof(float) * NCOL * n);
      ctx->orig_of[w].ensure(sizeof(int) * n);
    }
    ctx->slot_of.ensure(sizeof(int) * n);
    ctx->cur = 0;
    HIPCHECK(hipMemsetAsync(ctx->store[0].p, 0, sizeof(float) * NCOL * n, ctx->stream));
    std::vector<int> iota(n);
    for (int64_t i = 0; i < n; ++i) iota[i] = (int)i;
    HIPCHECK(hipMemcpyAsync(ctx->orig_of[0].p, iota.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipMemcpyAsync(ctx->slot_of.p, iota.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    launch_identity_columns(ctx->view(), C_F, (int)n, ctx->stream);
    ctx->sync();
  });
}

int64_t mpm_particles_count(const mpm_context* ctx) { return ctx ? ctx->n : 0; }

int mpm_particles_num_variables(const mpm_context* ctx) {
  if (!ctx) return 0;
  return 5 + (ctx->model ? 7 : 0) + (ctx->perParticleMaterial ? 3 : 0);
}

int mpm_particles_set(mpm_context* ctx, const char* name, const float* data, int64_t count) {
  return guarded(ctx, [&] {
    require_particles(ctx);
    auto it = var_table().find(name ? name : "");
    if (it == var_table().end() || (it->second.model && !ctx->model))
      throw InvalidArg(std::string("Material point data has no variable named '") + (name ? name : "") + "'");
    const VarInfo& vi = it->second;
    if (count != ctx->n * vi.ncomp) throw InvalidArg("size mismatch");
    ctx->staging.ensure(sizeof(float) * count);
    HIPCHECK(hipMemcpyAsync(ctx->staging.p, data, sizeof(float) * count, hipMemcpyHostToDevice, ctx->stream));
    launch_set_column(ctx->view(), vi.col, vi.ncomp, ctx->staging.as<float>(), ctx->slot_of.as<int>(), (int)ctx->n,
                      ctx->stream);
    if (vi.col == C_MU0 || vi.col == C_LAM0 || vi.col == C_XI) ctx->perParticleMaterial = true;
    ctx->sync();
  });
}

int mpm_particles_get(mpm_context* ctx, const char* name, float* data, int64_t count) {
  return guarded(ctx, [&] {
    require_particles(ctx);
    auto it = var_table().find(name ? name : "");
    if (it == var_table().end() || (it->second.model && !ctx->model))
      throw InvalidArg(std::string("Material point data has no variable named '") + (name ? name : "") + "'");
    const VarInfo& vi = it->second;
    if (count != ctx->n * vi.ncomp) throw InvalidArg("size mismatch");
    ctx->staging.ensure(sizeof(float) * count);
    launch_get_column(ctx->view(), vi.col, vi.ncomp, ctx->staging.as<float>(), ctx->slot_of.as<int>(), (int)ctx->n,
                      ctx->stream);
    HIPCHECK(hipMemcpyAsync(data, ctx->staging.p, sizeof(float) * count, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  });
}

int mpm_model_snow(mpm_context* ctx, float E, float nu, float hardening, float compressive, float tensile) {
  return guarded(ctx, [&] {
    require_particles(ctx);
    if (ctx->model) throw std::runtime_error("Material point data already has a variable named 'Fp'");
    ctx->E = E;
    ctx->nu = nu;
    ctx->xi = hardening;
    ctx->thc = compressive;
    ctx->ths = tensile;
    ctx->mu0 = (E / (2 * (1 + nu)));
    ctx->lambda0 = (E * nu / ((1 + nu) * (1 - 2 * nu)));
    int n = (int)ctx->n;
    PView P = ctx->view();
    for (int col : {C_FP, C_FI, C_R, C_GI}) launch_identity_columns(P, col, n, ctx->stream);
    launch_fill_column(P, C_J, 1.0f, n, ctx->stream);
    launch_fill_column(P, C_MU, ctx->mu0, n, ctx->stream);
    launch_fill_column(P, C_LAM, ctx->lambda0, n, ctx->stream);
    launch_fill_column(P, C_MU0, ctx->mu0, n, ctx->stream);
    launch_fill_column(P, C_LAM0, ctx->lambda0, n, ctx->stream);
    launch_fill_column(P, C_XI, ctx->xi, n, ctx->stream);
    ctx->model = true;
    ctx->sync();
  });
}

int mpm_model_update_particle_data(mpm_context* ctx) {
  return guarded(ctx, [&] {
    require_particles(ctx);
    if (!ctx->model) throw InvalidArg("no constitutive model");
    ctx->clear_err();
    launch_plasticity(ctx->view(), nullptr, 0, (int)ctx->n, ctx->thc, ctx->ths, ctx->errp(), ctx->stream);
    ctx->check_errors();
  });
}

static int model_probe(mpm_context* ctx, const int32_t* particles, const float* dF, int64_t n, int mode, float* out) {
  return guarded(ctx, [&] {
    require_particles(ctx);
    if (!ctx->model) throw InvalidArg("no constitutive model");
    if (n <= 0) return;
    DevBuf slots, din, dout;
    ctx->slots_from_orig(particles, n, slots);
    int w = mode == 0 ? 1 : 9;
    dout.ensure(sizeof(float) * w * n);
    if (dF) {
      din.ensure(sizeof(float) * 9 * n);
      HIPCHECK(hipMemcpyAsync(din.p, dF, sizeof(float) * 9 * n, hipMemcpyHostToDevice, ctx->stream));
    }
    launch_probe(ctx->view(), slots.as<int>(), (int)n, mode, din.as<float>(), dout.as<float>(), ctx->stream);
    HIPCHECK(hipMemcpyAsync(out, dout.p, sizeof(float) * w * n, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  });
}
int mpm_model_energy_density(mpm_context* ctx, const int32_t* p, int64_t n, float* out) {
  return model_probe(ctx, p, nullptr, n, 0, out);
}
int mpm_model_stress(mpm_context* ctx, const int32_t* p, int64_t n, float* out) {
  return model_probe(ctx, p, nullptr, n, 1, out);
}
int mpm_model_stress_differential(mpm_context* ctx, const int32_t* p, const float* dF, int64_t n, float* out) {
  if (!dF) return MPM_ERR_INVALID;
  return model_probe(ctx, p, dF, n, 2, out);
}

int mpm_set_colliders(mpm_context* ctx, const mpm_collider* cols, int n) {
  return guarded(ctx, [&] {
    if (n < 0 || n > kMaxColliders) throw InvalidArg("at most 16 colliders are supported");
    ColliderSet cs{};
    cs.n = n;
    for (int i = 0; i < n; ++i) {
      Collider& o = cs.o[i];
      o.type = cols[i].type;
      if (o.type != 0 && o.type != 1) throw InvalidArg("unknown collider type");
      for (int a = 0; a < 4; ++a) o.c[a] = cols[i].coeffs[a];
      for (int a = 0; a < 3; ++a) {
        o.center[a] = cols[i].center[a];
        o.v[a] = cols[i].velocity[a];
      }
      o.radius = cols[i].radius;
      o.friction = cols[i].friction;
      o.sticky = cols[i].sticky;
    }
    ctx->hcols = cs;
    HIPCHECK(hipMemcpyAsync(ctx->cols_dev.p, &ctx->hcols, sizeof(ColliderSet), hipMemcpyHostToDevice, ctx->stream));
    ctx->sync();
  });
}

int mpm_set_gravity(mpm_context* ctx, const float* g, int n) {
  return guarded(ctx, [&] {
    if (n < 0 || n > 64) throw InvalidArg("at most 64 gravity fields");
    ctx->hgrav.assign(g, g + 3 * n);
    if (n > 0)
      HIPCHECK(hipMemcpyAsync(ctx->grav_dev.p, ctx->hgrav.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice,
                              ctx->stream));
    ctx->sync();
  });
}

// ------------------------------------------------------------------ grid API
int mpm_grid_create(mpm_context* ctx, const int32_t* particles, int64_t count, float h, const float fv[3], int dim,
                    mpm_grid** out) {
  if (!out) return MPM_ERR_INVALID;
  *out = nullptr;
  std::unique_ptr<mpm_grid> g(new mpm_grid());
  g->ctx = ctx;
  int rc = guarded(ctx, [&] {
    require_particles(ctx);
    if (dim != 3) throw InvalidArg("only dimension 3 is supported on the GPU");
    if (!(h > 0)) throw InvalidArg("grid size must be positive");
    DevBuf slots;
    ctx->slots_from_orig(particles, count, slots);
    g->build(slots.as<int>(), (int)count, h, fv, false);
    ctx->check_errors();
  });
  if (rc == MPM_OK) *out = g.release();
  return rc;
}

int mpm_grid_destroy(mpm_grid* grid) {
  if (grid && grid->ctx) {
    (void)hipSetDevice(grid->ctx->device);
    (void)hipStreamSynchronize(grid->ctx->stream);
  }
  delete grid;
  return MPM_OK;
}

int mpm_grid_dims(const mpm_grid* grid, int n[3], float mn[3]) {
  if (!grid) return MPM_ERR_INVALID;
  for (int a = 0; a < 3; ++a) {
    if (n) n[a] = grid->g.n[a];
    if (mn) mn[a] = grid->g.gmin[a];
  }
  return MPM_OK;
}

int mpm_grid_get(mpm_grid* grid, const char* name, void* out) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    std::string nm(name ? name : "");
    if (nm == "masses") grid->to_dense(grid->f(grid->mass), 1, static_cast<float*>(out));
    else if (nm == "velocities") grid->to_dense(grid->f(grid->vel), 3, static_cast<float*>(out));
    else if (nm == "prev_velocities") grid->to_dense(grid->f(grid->prev), 3, static_cast<float*>(out));
    else if (nm == "node_collided") {
      int64_t nd = (int64_t)grid->g.n[0] * grid->g.n[1] * grid->g.n[2];
      grid->dense2.ensure(nd);
      launch_to_dense_i8(grid->flags.as<int8_t>(), grid->dense2.as<int8_t>(), grid->g, grid->ctx->stream);
      HIPCHECK(hipMemcpyAsync(out, grid->dense2.p, nd, hipMemcpyDeviceToHost, grid->ctx->stream));
      grid->ctx->sync();
    } else
      throw InvalidArg("unknown grid variable '" + nm + "'");
  });
}

int mpm_grid_set(mpm_grid* grid, const char* name, const void* in) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    std::string nm(name ? name : "");
    if (nm == "masses") grid->from_dense(static_cast<const float*>(in), 1, grid->f(grid->mass));
    else if (nm == "velocities") grid->from_dense(static_cast<const float*>(in), 3, grid->f(grid->vel));
    else if (nm == "prev_velocities") grid->from_dense(static_cast<const float*>(in), 3, grid->f(grid->prev));
    else if (nm == "node_collided") {
      int64_t nd = (int64_t)grid->g.n[0] * grid->g.n[1] * grid->g.n[2];
      grid->dense2.ensure(nd);
      HIPCHECK(hipMemcpyAsync(grid->dense2.p, in, nd, hipMemcpyHostToDevice, grid->ctx->stream));
      launch_from_dense_i8(grid->dense2.as<int8_t>(), grid->flags.as<int8_t>(), grid->g, grid->ctx->stream);
    } else
      throw InvalidArg("unknown grid variable '" + nm + "'");
    grid->ctx->sync();
  });
}

int mpm_grid_compute_particle_volumes(mpm_grid* grid) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    grid->ctx->clear_err();
    launch_volumes(grid->ctx->view(), grid->g, grid->f(grid->mass), grid->ctx->errp(), grid->ctx->stream);
    grid->ctx->check_errors();
  });
}

int mpm_grid_calculate_forces(mpm_grid* grid, float* forces) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    if (!c.model) throw InvalidArg("no constitutive model");
    c.clear_err();
    grid->forces_into_acc();
    launch_node_forces(grid->g, grid->f(grid->mass), grid->f(grid->acc), grid->grav(), grid->ngrav(),
                       grid->f(grid->acc2), c.stream);
    c.check_errors();
    grid->to_dense(grid->f(grid->acc2), 3, forces);
  });
}

int mpm_grid_force_differentials(mpm_grid* grid, const float* dx, float* df) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    if (!c.model) throw InvalidArg("no constitutive model");
    c.clear_err();
    grid->from_dense(dx, 3, grid->f(grid->acc2));
    HIPCHECK(hipMemsetAsync(grid->acc.p, 0, grid->vecBytes(), c.stream));
    launch_hvp(c.view(), grid->g, grid->f(grid->acc2), false, grid->flags.as<int8_t>(), c.cols(), grid->f(grid->acc),
               nullptr, c.errp(), c.stream);
    c.check_errors();
    grid->to_dense(grid->f(grid->acc), 3, df);
  });
}

int mpm_grid_dforce_dxi(mpm_grid* grid, float* out) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    if (!c.model) throw InvalidArg("no constitutive model");
    c.clear_err();
    HIPCHECK(hipMemsetAsync(grid->acc.p, 0, grid->vecBytes(), c.stream));
    launch_diag(c.view(), grid->g, grid->f(grid->acc), c.errp(), c.stream);
    c.check_errors();
    grid->to_dense(grid->f(grid->acc), 3, out);
  });
}

int mpm_grid_explicit_momenta(mpm_grid* grid, float dt, float* momenta, int8_t* collided) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    if (!c.model) throw InvalidArg("no constitutive model");
    c.clear_err();
    HIPCHECK(hipMemsetAsync(grid->sc.p, 0, sizeof(CRScalars), c.stream));
    grid->forces_into_acc();
    // write into scratch buffers so the grid's own node flags (set only by
    // updateGridVelocities in the reference) stay untouched
    DevBuf fl;
    fl.ensure(grid->NN);
    launch_node_explicit(grid->g, grid->f(grid->mass), grid->f(grid->vel), grid->f(grid->acc), grid->grav(),
                         grid->ngrav(), dt, c.cols(), grid->f(grid->acc2), fl.as<int8_t>(), grid->f(grid->draw),
                         grid->scal(), grid->partials.as<double>(), c.stream);
    CRScalars hs;
    HIPCHECK(hipMemcpyAsync(&hs, grid->sc.p, sizeof(hs), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    c.check_errors();
    mpm_context::raise(hs.errors);
    grid->to_dense(grid->f(grid->acc2), 3, momenta);
    int64_t nd = (int64_t)grid->g.n[0] * grid->g.n[1] * grid->g.n[2];
    grid->dense2.ensure(nd);
    launch_to_dense_i8(fl.as<int8_t>(), grid->dense2.as<int8_t>(), grid->g, c.stream);
    HIPCHECK(hipMemcpyAsync(collided, grid->dense2.p, nd, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

int mpm_grid_collision_velocities(mpm_grid* grid, const int8_t* collided, float* vc) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    int64_t nd = (int64_t)grid->g.n[0] * grid->g.n[1] * grid->g.n[2];
    for (int64_t i = 0; i < nd; ++i)
      if (collided[i] >= c.hcols.n) throw InvalidArg("collision object index out of range");
    DevBuf fl;
    fl.ensure(grid->NN);
    HIPCHECK(hipMemsetAsync(fl.p, 0xFF, grid->NN, c.stream));
    grid->dense2.ensure(nd);
    HIPCHECK(hipMemcpyAsync(grid->dense2.p, collided, nd, hipMemcpyHostToDevice, c.stream));
    launch_from_dense_i8(grid->dense2.as<int8_t>(), fl.as<int8_t>(), grid->g, c.stream);
    launch_collision_velocities(grid->g, fl.as<int8_t>(), c.cols(), grid->f(grid->acc2), c.stream);
    grid->to_dense(grid->f(grid->acc2), 3, vc);
  });
}

static void fill_stats(const SolveRecord& r, mpm_solve_stats* st) {
  if (!st) return;
  st->iterations = r.iterations;
  st->early_exit = r.early_exit;
  st->rhs_norm2 = r.rhs_norm2;
  st->residual_norm2 = r.residual_norm2;
  int L = (int)r.history.size();
  if (st->history && st->history_capacity > 0) {
    int k = std::min(L, st->history_capacity);
    std::memcpy(st->history, r.history.data(), sizeof(double) * k);
    st->history_length = k;
  } else {
    st->history_length = 0;
  }
}

int mpm_grid_update_grid_velocities(mpm_grid* grid, float dt, const mpm_solver_params* params,
                                    mpm_solve_stats* stats) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    if (!grid->ctx->model) throw InvalidArg("no constitutive model");
    if (!params) throw InvalidArg("solver parameters required");
    SolveRecord rec;
    grid->updateGridVelocities(dt, params, rec);
    fill_stats(rec, stats);
  });
}

int mpm_grid_update_particle_velocities(mpm_grid* grid) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    c.clear_err();
    launch_g2p(c.view(), grid->g, grid->f(grid->vel), grid->f(grid->prev), 0, 0.0f, 0.0f, 0.0f, c.cols(), c.errp(),
               c.stream);
    c.check_errors();
  });
}

int mpm_grid_update_deformation_gradients(mpm_grid* grid, float dt) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    c.clear_err();
    launch_g2p(c.view(), grid->g, grid->f(grid->vel), grid->f(grid->prev), 1, dt, 0.0f, 0.0f, c.cols(), c.errp(),
               c.stream);
    c.check_errors();
  });
}

// ------------------------------------------------------------------ Sim API
int mpm_sim_init(mpm_context* ctx, float h, int dim) {
  return guarded(ctx, [&] {
    require_particles(ctx);
    if (dim != 3) throw InvalidArg("only dimension 3 is supported on the GPU");
    if (!(h > 0)) throw InvalidArg("grid size must be positive");
    ctx->h = h;
    ctx->dim = dim;
    calculate_bodies(*ctx);
    // Sim::Sim: volumes of the initial bodies (Sim.cpp:44-49)
    for (size_t b = 0; b < ctx->bodies.size(); ++b) {
      const auto& B = ctx->bodies[b];
      DevBuf slots;
      ctx->slots_from_orig(B.data(), (int64_t)B.size(), slots);
      mpm_grid* g = pooled_grid(*ctx, b);
      float zero[3] = {0, 0, 0};
      g->build(slots.as<int>(), (int)B.size(), h, zero, false);
      launch_volumes(ctx->view(), g->g, g->f(g->mass), ctx->errp(), ctx->stream);
      ctx->check_errors();
    }
    ctx->simInit = true;
    ctx->lastSolves.clear();
  });
}

int mpm_sim_advance(mpm_context* ctx, float dt, const mpm_solver_params* params, mpm_step_stats* stats) {
  return guarded(ctx, [&] {
    if (!ctx->simInit) throw InvalidArg("simulation not initialised (call mpm_sim_init)");
    if (!ctx->model) throw InvalidArg("no constitutive model");
    if (!params) throw InvalidArg("solver parameters required");
    mpm_context& c = *ctx;
    hipStream_t s = c.stream;
    const int n = (int)c.n;
    c.lastSolves.clear();
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const bool cancelHook = params->cancelled != nullptr;
    // body slot lists
    std::vector<int> bodyStart;
    size_t nb = c.bodies.size();
    c.clear_err();
    bool cancelled = false;
    for (size_t b = 0; b < nb; ++b) {
      const auto& B = c.bodies[b];
      DevBuf slots;
      c.slots_from_orig(B.data(), (int64_t)B.size(), slots);
      mpm_grid* g = pooled_grid(c, b);
      g->build(slots.as<int>(), (int)B.size(), c.h, nullptr, true);
      c.check_errors();
      SolveRecord rec;
      g->updateGridVelocities(dt, params, rec);
      c.lastSolves.push_back(rec);
      if (stats) {
        stats->total_iterations += rec.iterations;
        stats->max_iterations = std::max(stats->max_iterations, rec.iterations);
      }
      if (cancelHook && params->cancelled(params->user)) {
        cancelled = true;
        break;
      }
      // G2P + F + plasticity (+ advection when no cancellation can intervene)
      c.clear_err();
      int pg = c.prof_begin(PK_G2P, g);
      launch_g2p(c.view(), g->g, g->f(g->vel), g->f(g->prev), cancelHook ? 3 : 2, dt, c.thc, c.ths, c.cols(), c.errp(),
                 s);
      c.prof_end(pg);
      if (b == 0 && (int64_t)B.size() < c.n) {
        // updateParticleData runs over all particles after each body (Sim.cpp:118); the other
        // particles are brought up to date here once, from their unchanged F
        std::vector<char> inBody(n, 0);
        for (int i : B) inBody[i] = 1;
        std::vector<int> others;
        others.reserve(n - B.size());
        for (int i = 0; i < n; ++i)
          if (!inBody[i]) others.push_back(i);
        c.slots_from_orig(others.data(), (int64_t)others.size(), c.other_slots);
        launch_plasticity(c.view(), c.other_slots.as<int>(), 0, (int)others.size(), c.thc, c.ths, c.errp(), s);
      }
      c.check_errors();
    }
    // ballistic particles: velocity (Sim.cpp:64-73) and advection
    if (!c.ballistic.empty()) {
      c.slots_from_orig(c.ballistic.data(), (int64_t)c.ballistic.size(), c.ball_slots);
      launch_ballistic(c.view(), c.ball_slots.as<int>(), (int)c.ballistic.size(), c.grav_dev.as<float>(),
                       (int)(c.hgrav.size() / 3), dt, c.cols(), cancelled ? 1 : (cancelHook ? 1 : 3), s);
    }
    if (cancelHook && !cancelled) {
      // advection of all body particles (Sim.cpp:123-131)
      for (size_t b = 0; b < nb; ++b) {
        mpm_grid* g = pooled_grid(c, b);
        launch_advect(c.view(), g->perm.as<int>(), g->np, dt, c.cols(), s);
      }
      if (!c.ballistic.empty())
        launch_ballistic(c.view(), c.ball_slots.as<int>(), (int)c.ballistic.size(), c.grav_dev.as<float>(),
                         (int)(c.hgrav.size() / 3), dt, c.cols(), 2, s);
    }
    if (!cancelled && nb > 0) {
      // keep the store body- and tile-sorted: new slot order = body perms, then the rest
      std::vector<char> placed;
      int64_t off = 0;
      c.new_to_old.ensure(sizeof(int) * n);
      for (size_t b = 0; b < nb; ++b) {
        mpm_grid* g = pooled_grid(c, b);
        HIPCHECK(hipMemcpyAsync(c.new_to_old.as<int>() + off, g->perm.p, sizeof(int) * g->np, hipMemcpyDeviceToDevice,
                                s));
        off += g->np;
      }
      if (off + (int64_t)c.ballistic.size() != n)
        throw std::runtime_error("body partition does not cover every particle (internal error)");
      if (off < n) {
        HIPCHECK(hipMemcpyAsync(c.new_to_old.as<int>() + off, c.ball_slots.p, sizeof(int) * (n - off),
                                hipMemcpyDeviceToDevice, s));
      }
      int nxt = 1 - c.cur;
      launch_permute(c.view(c.cur), c.view(nxt), c.new_to_old.as<int>(), n, c.orig_of[c.cur].as<int>(),
                     c.orig_of[nxt].as<int>(), c.slot_of.as<int>(), s);
      c.cur = nxt;
    }
    c.sync();
    c.prof_resolve();
    calculate_bodies(c);
    if (stats) {
      stats->num_bodies = (int)c.bodies.size();
      stats->num_ballistic = (int)c.ballistic.size();
      stats->cancelled = cancelled ? 1 : 0;
    }
  });
}

int mpm_profile_enable(mpm_context* ctx, int on) {
  return guarded(ctx, [&] {
    ctx->prof_resolve();
    ctx->profOn = on != 0;
    for (int k = 0; k < PK_COUNT; ++k) {
      ctx->profMs[k] = 0;
      ctx->profCount[k] = 0;
      ctx->profBytes[k] = 0;
    }
  });
}

int mpm_profile_query(mpm_context* ctx, const char* kernel, double* total_ms, int64_t* launches,
                      double* algorithmic_bytes) {
  return guarded(ctx, [&] {
    ctx->prof_resolve();
    std::string k(kernel ? kernel : "");
    for (int i = 0; i < PK_COUNT; ++i)
      if (k == kProfNames[i]) {
        if (total_ms) *total_ms = ctx->profMs[i];
        if (launches) *launches = ctx->profCount[i];
        if (algorithmic_bytes) *algorithmic_bytes = ctx->profBytes[i];
        return;
      }
    throw InvalidArg("unknown kernel class '" + k + "'");
  });
}

int mpm_sim_num_bodies(const mpm_context* ctx) { return ctx ? (int)ctx->bodies.size() : 0; }

int64_t mpm_sim_body(const mpm_context* ctx, int b, int32_t* out) {
  if (!ctx || b < 0 || b >= (int)ctx->bodies.size()) return -1;
  const auto& B = ctx->bodies[b];
  if (out) std::copy(B.begin(), B.end(), out);
  return (int64_t)B.size();
}

int64_t mpm_sim_ballistic(const mpm_context* ctx, int32_t* out) {
  if (!ctx) return -1;
  if (out) std::copy(ctx->ballistic.begin(), ctx->ballistic.end(), out);
  return (int64_t)ctx->ballistic.size();
}

int mpm_sim_body_solve_stats(const mpm_context* ctx, int b, mpm_solve_stats* st) {
  if (!ctx || b < 0 || b >= (int)ctx->lastSolves.size()) return MPM_ERR_INVALID;
  fill_stats(ctx->lastSolves[b], st);
  return MPM_OK;
}

}  // extern "C"
