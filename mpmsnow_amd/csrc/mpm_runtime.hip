// mpm_runtime.hip -- host runtime behind include/mpmsnow_gpu.h.
//
// Owns the device particle store (double-buffered SoA columns kept in body-
// and tile-sorted slot order) and a batch of per-body grids, and sequences
// Sim::advance (davidsminor/mpmsnow src/MpmSim/Sim.cpp:57-134) and
// Grid::updateGridVelocities (src/MpmSim/Grid.cpp:889-1016) as gfx950 kernel
// launches on one HIP stream. All bodies of a step share one batch, so every
// kernel launch serves every body; each body keeps its own conjugate-residual
// scalars and termination, and the host only polls a done-counter once per
// batch of iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mpmsnow_gpu.h"
#include "lpc_net.hpp"
#include "mpm_host.hpp"
#include "mpm_kernels.hpp"

using namespace mpm;
using namespace mpmh;

// Scratch of the allocation-free entry point mpm_model_stress_differential_noalloc: the slot list
// of one call's particles. Each GPU holds its own copy of the array (4 MiB) until the process
// ends; host code reaches it only through the per-device address that hipGetSymbolAddress returns
// (taken once per device, at the first mpm_create on it). Every call on a GPU shares it, so calls
// that overlap on one GPU race.
constexpr int64_t kNoAllocMaxParticles = 1 << 20;
static __device__ __align__(256) int scratch[kNoAllocMaxParticles];

namespace {

struct VarInfo {
  int col;
  int ncomp;
  bool model;
};
const std::unordered_map<std::string, VarInfo>& var_table() {
  static const std::unordered_map<std::string, VarInfo> t = {
      {"p", {C_X, 3, false}},         {"v", {C_V, 3, false}},       {"F", {C_F, 9, false}},
      {"m", {C_M, 1, false}},         {"volume", {C_VOL, 1, false}}, {"Fp", {C_FP, 9, true}},
      {"FinvTrans", {C_FI, 9, true}}, {"R", {C_R, 9, true}},         {"Ginv", {C_GI, 9, true}},
      {"J", {C_J, 1, true}},          {"mu", {C_MU, 1, true}},       {"lambda", {C_LAM, 1, true}},
      {"mu0", {C_MU0, 1, true}},      {"lambda0", {C_LAM0, 1, true}}, {"xi", {C_XI, 1, true}}};
  return t;
}

float decode_ord(unsigned int u) {
  unsigned int v = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
  float f;
  std::memcpy(&f, &v, 4);
  return f;
}

void raise_errors(int e) {
  if (e & ERR_NAN_MASS) throw std::runtime_error("nans in splatted masses!");
  if (e & ERR_NAN_VEL) throw std::runtime_error("nans in splatted velocities!");
  if (e & ERR_NAN_STRESS) throw std::runtime_error("nans in dEnergyDensitydF matrix!");
  if (e & ERR_NAN_MOMENTA) throw std::runtime_error("nan in explicit momenta!");
  if (e & ERR_HARDNESS) throw std::runtime_error("infinite hardness!!!");
  if (e & ERR_NAN_PVEL) throw std::runtime_error("nans in particle velocity data!");
  if (e & ERR_STENCIL) throw std::runtime_error("particle outside its grid tile (internal binning error)");
  if (e & ERR_INTERNAL) throw std::runtime_error("launch does not fit the batch layout (internal error)");
  if (e & ERR_INDEX) throw InvalidArg("particle index out of range");
}

// Kernel timing for measurement: HIP events recorded on the library's stream around each
// instrumented launch, resolved at the next synchronisation point. Algorithmic bytes of a
// launch = bytes/particle x particles + bytes/node x nodes with m > 0, over the bodies the
// launch worked on (bodies whose solve had finished skip the launch). The one-workgroup solver
// of small bodies is one launch for all their iterations: its bytes are those of an HVP and
// the three node kernels per iteration (plus the setup's two).
enum ProfKind {
  PK_P2G_MASS, PK_P2G_VEL, PK_FORCES, PK_HVP, PK_DIAG, PK_G2P, PK_CR_X, PK_CR_AR, PK_CR_PX, PK_LPC, PK_CR_SMALL,
  PK_COUNT
};
const char* const kProfNames[PK_COUNT] = {"p2g_mass", "p2g_vel", "forces", "hvp", "diag", "g2p",
                                          "cr_update_x", "cr_ar", "cr_update_px", "lpc_forward", "cr_small"};
// the network is MFMA work: its launches are timed but carry no HBM byte count
const double kProfBytesPerParticle[PK_COUNT] = {16, 28, 136, 172, 172, 324, 0, 0, 0, 0, 172};
const double kProfBytesPerNode[PK_COUNT] = {4, 16, 12, 29, 12, 24, 85, 65, 121, 0, 121 + 65 + 29};
// ProfRec::iter values other than CR-loop iterations
constexpr int kProfOnce = -1;     // once per solve, every body of the batch
constexpr int kProfSetup = -2;    // CR setup of the bodies the batched loop solves
constexpr int kProfSmall = -3;    // the small-body solver

}  // namespace

namespace {
constexpr int kMaxDevices = 256;
std::mutex gScratchMutex;
int* gScratch[kMaxDevices] = {};  // device address of `scratch` on each device (null: not taken yet)
// the calling thread's current device must be `device`
int* scratch_on(int device) {
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (!gScratch[device]) {
    void* p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    gScratch[device] = static_cast<int*>(p);
  }
  return gScratch[device];
}
}  // namespace

struct SolveRecord {
  bool small = false;  // solved by the one-workgroup small-body solver
  int iterations = 0;
  int early_exit = 0;
  int fallbacks = 0;  // learned preconditioner: switched to the diagonal one
  int restarts = 0;   // CR recurrence restarts after a cancelled denominator expansion
  double rhs_norm2 = 0, residual_norm2 = 0;
  int np = 0, active = 0;
  std::vector<double> history;
};

struct mpm_context;

// A batch of per-body grids with all their device state.
struct GridBatch {
  mpm_context* ctx = nullptr;
  std::vector<BodyGrid> hb;
  int nb = 0, NT = 0, nblocks = 0, nchunks = 0, np = 0;
  int64_t NN = 0;
  DevBuf bodies, bodyOrder, tileBody, blockBody, perm, pcode, key, rank, count, offset, nch, chunkOff, chunks, scantmp, totals;
  DevBuf stageA, stageB;
  DevBuf mass, vel, prev, flags, vc, b, x, z, p, Ap, Ar, diag, tmp3, flags2, scal, status, partials, tickets, hist, dense,
      dense2;
  DevBuf bbox, mom, cnt, bboxR, momR, cntR;
  DevBuf tileAct;  // per tile: active in the CR loop (set by k_node_explicit)
  DevBuf pairCnt, pairPos, groups;  // chunk pairs of the CR loop's HVP
  DevBuf nrm;      // collider unit normal at nodes with a flag >= 0 (set with the flags)
  DevBuf fr, fp0, fp1, fq0, fq1;  // flexible CR residual; training and grid-API scratch vectors
  DevBuf fP, fQ;                   // flexible CR: the ring of directions p_j and A p_j (fcr_m vectors each)
  DevBuf smallList;               // bodies handed to the small-body solver
  std::vector<char> smallLayout;  // per body: laid out after the prefix (small-body solver candidate)
  std::vector<int> layoutOrder;   // bodies by ascending tileBase
  int prefixTiles = 0, prefixChunks = 0;  // tiles / chunks of the bodies laid out first
  DevBuf* keepKey = nullptr;  // when set, build() keeps a copy of the binning's keys here (caller order)
  LpcLevels levels;
  bool levelsBuilt = false;
  uint64_t version = 0;  // bumped by every build
  std::vector<SolveRecord> rec;

  Batch view() const {
    Batch B;
    B.bodies = bodies.as<BodyGrid>();
    B.nbodies = nb;
    B.blockBody = blockBody.as<int>();
    B.tileBody = tileBody.as<int>();
    B.perm = perm.as<int>();
    B.chunks = chunks.as<int4>();
    B.nchunks = nchunks;
    B.tileChunk0 = chunkOff.as<int>();
    B.tileChunkN = nch.as<int>();
    B.NT = NT;
    B.nblocks = nblocks;
    B.tickets = tickets.as<unsigned int>();
    B.tileActive = tileAct.as<unsigned char>();
    B.nrm = nrm.as<float>();
    return B;
  }
  float* f(DevBuf& d) { return d.as<float>(); }
  size_t vecBytes() const { return sizeof(float) * 3 * (size_t)NN; }
  CRScalars* sc() { return scal.as<CRScalars>(); }
  BatchStatus* st() { return status.as<BatchStatus>(); }

  // indexOrder: lay the bodies' tiles out in body order (no small-body suffix)
  void build(const int* slots, const int* labels, int n, int nbodies, float h, const float* frameV, bool computeFrame,
             bool indexOrder = false);
  void forces_stage();
  void hvp(const float* d, bool project, float* stage, const CRScalars* scp, int iter);
  // paired: over the chunk pairs of the last pair_chunks (built for this Bv)
  void hvp(const Batch& Bv, const float* d, bool project, float* stage, const CRScalars* scp, int iter,
           bool paired = false);
  void pair_chunks(const Batch& Bv);
  void solve(float dt, const mpm_solver_params* prm);
  void train_lpc(float dt, CRScalars* hsc, size_t scb);
  void ensure_levels() {
    if (!levelsBuilt) levels.build(hb, layoutOrder, bodyOrder.as<int>(), ctx_stream());
    levelsBuilt = true;
  }
  hipStream_t ctx_stream() const;
  void to_dense(int body, const float* batch, int ncomp, float* host);
  void from_dense(int body, const float* host, int ncomp, float* batch);
  void fetch_history(int body, SolveRecord& r);
};

struct mpm_grid {
  mpm_context* ctx = nullptr;
  GridBatch batch;
  DevBuf slots;
  DevBuf binKey, cells;  // the binning's key per particle of the grid (caller order) and its decoding
  int64_t count = 0;
};

// Residuals of one solve for training the learned preconditioner, and the training settings
struct LpcCapture {
  int want = 0;                  // samples the next solve keeps (0: no training)
  int epochs = 0;
  float lr = 0.0f;
  int n = 0;                     // samples kept
  std::vector<int> iters;        // CR iteration of each sample
  std::vector<int> schedule = {0, 1, 2, 3, 4, 6, 8, 11, 15, 20, 26, 33, 41, 50, 60, 72};  // iterations to keep
  DevBuf data;                   // [sample][3 NN floats]: residuals r_k
  DevBuf xdata;                  // [sample][3 NN floats]: iterates x_k (exact-correction loss)
  DevBuf scale, weight, sums, loss, energy;  // [sample][body]
  int64_t NN = 0;
  int lossMode = 0;              // 0: optimal-step residual loss; 1: exact-correction energy-norm loss
  double lossFirst = 0, lossLast = 0, lossDiag = 0;  // mean per-body loss: first, last epoch, diagonal only
  int pairs = 0;                 // (body, sample) pairs the last training used
  float* r(int k) { return data.as<float>() + (int64_t)k * 3 * NN; }
  float* xk(int k) { return xdata.as<float>() + (int64_t)k * 3 * NN; }
};

struct mpm_context {
  int device = 0;
  int* scratch = nullptr;  // this device's copy of the file-scope `scratch` array
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // small-body solver, concurrent with the batched CR loop
  hipEvent_t smallReady = nullptr, smallDone = nullptr;  // stream -> stream2 -> stream ordering
  int smallMaxParticles = 2048;   // bodies up to this size and kSmallBlocks node blocks (0: off)
  std::string err;
  int64_t n = 0;
  bool model = false;
  float E = 0, nu = 0, xi = 0, thc = 0, ths = 0, mu0 = 0, lambda0 = 0;
  bool perParticleMaterial = false;
  DevBuf store[2];
  int cur = 0;
  DevBuf orig_of[2], slot_of;
  DevBuf cols_dev, grav_dev, errflag, staging, idx_tmp;
  ColliderSet hcols{};
  DevBuf sdfPhi[kMaxColliders], sdfVel[kMaxColliders];  // level-set collider volumes
  std::vector<float> hgrav;
  PinnedBuf pinned;
  // simulation state
  bool simInit = false;
  float h = 0;
  int dim = 3;
  int nbodies = 0;
  int nballistic = 0;
  DevBuf labels;  // body of each slot (-1 ballistic)
  GridBatch simBatch;
  std::vector<SolveRecord> lastSolves;
  bool bodyListsValid = false;
  std::vector<std::vector<int>> bodyLists;
  std::vector<int> ballisticList;
  // body-partition scratch
  DevBuf ccBucket, ccRank, ccCount, ccOffset, ccSorted, ccParent, ccHasNb, ccIsRoot, ccRootIdx, ccMinOrig, ccTmp,
      ccTot, ccSkip, newToOld, ballPos;
  int iterHint = 8;
  int iterHintFcr = 8;
  int iterHintHandoff = 8;  // CR iterations after a learned solve's handoff (last solve's count)
  // learned preconditioner
  LpcNet net;
  int lpcMode = 0;  // 0: diagonal-preconditioned CR; 1: learned flexible CR
  double lpcFallback = 0.95;  // fallback ratio of a network iteration's |r|^2 to the previous one
  int lpcDirs = 2;            // directions the flexible CR keeps (1..kFcrMax)
  int lpcHandoff = 1;         // 1: hand the solve to the CR recurrences once the network is off
  int lpcStrikes = kLpcStrikes;  // fallbacks after which a body keeps the diagonal preconditioner
  LpcCapture cap;
  // measurement
  struct ProfRec {
    int kind;
    int iter;  // CR loop iteration, or kProfOnce / kProfSetup / kProfSmall
    hipEvent_t a, b;
    const GridBatch* batch;
    hipStream_t stream;
  };
  unsigned int profMask = 0;  // kernel classes being timed
  int profPeriod = 1;         // time one launch in profPeriod of each class
  int64_t profSeen[PK_COUNT] = {};
  std::vector<ProfRec> profPending;
  std::vector<hipEvent_t> eventPool;
  double profMs[PK_COUNT] = {};
  int64_t profCount[PK_COUNT] = {};
  double profBytes[PK_COUNT] = {};

  PView view(int which = -1) const {
    int w = which < 0 ? cur : which;
    PView v;
    v.base = store[w].as<float>();
    v.stride = n;
    return v;
  }
  void sync() { HIPCHECK(hipStreamSynchronize(stream)); }
  const ColliderSet* cols() const { return cols_dev.as<ColliderSet>(); }
  const float* grav() const { return grav_dev.as<float>(); }
  int ngrav() const { return (int)(hgrav.size() / 3); }
  int* errp() const { return errflag.as<int>(); }
  void clear_err() { HIPCHECK(hipMemsetAsync(errflag.p, 0, sizeof(int), stream)); }
  void check_errors() {
    int e = 0;
    HIPCHECK(hipMemcpyAsync(&e, errflag.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    sync();
    raise_errors(e);
  }
  // upload caller indices and map them to slots
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
  int prof_begin(int kind, const GridBatch* g, int iter = kProfOnce, hipStream_t on = nullptr) {
    if (!(profMask & (1u << kind))) return -1;
    if (profPeriod > 1 && profSeen[kind]++ % profPeriod != 0) return -1;
    ProfRec r{kind, iter, new_event(), new_event(), g, on ? on : stream};
    HIPCHECK(hipEventRecord(r.a, r.stream));
    profPending.push_back(r);
    return (int)profPending.size() - 1;
  }
  void prof_end(int idx) {
    if (idx < 0) return;
    HIPCHECK(hipEventRecord(profPending[idx].b, profPending[idx].stream));
  }
  void prof_resolve() {
    if (profPending.empty()) return;
    sync();
    // per batch: bodies of the batched loop by iteration count, with suffix sums of their
    // particles and active nodes, so that a CR-loop launch's bytes cost a binary search
    struct LoopSums {
      std::vector<int> its;              // ascending
      std::vector<double> np, active;    // sums over bodies k.. of the ascending order
      double npAll = 0, activeAll = 0;   // every body (kProfOnce)
      double npSetup = 0, activeSetup = 0, npSmall = 0, activeSmall = 0;  // batched / small bodies
    };
    std::vector<std::pair<const GridBatch*, LoopSums>> sums;
    auto sums_of = [&](const GridBatch* g) -> const LoopSums& {
      for (auto& e : sums)
        if (e.first == g) return e.second;
      LoopSums L;
      std::vector<std::pair<int, int>> order;  // (iterations, body)
      for (int bi = 0; bi < (int)g->rec.size(); ++bi) {
        const SolveRecord& r = g->rec[bi];
        L.npAll += r.np;
        L.activeAll += r.active;
        if (r.small) {
          L.npSmall += (double)r.np * (r.iterations + 2);
          L.activeSmall += (double)r.active * (r.iterations + 2);
        } else {
          L.npSetup += r.np;
          L.activeSetup += r.active;
          order.push_back({r.iterations, bi});
        }
      }
      std::sort(order.begin(), order.end());
      const size_t m = order.size();
      L.its.resize(m);
      L.np.assign(m + 1, 0.0);
      L.active.assign(m + 1, 0.0);
      for (size_t k = m; k-- > 0;) {
        const SolveRecord& r = g->rec[order[k].second];
        L.its[k] = order[k].first;
        L.np[k] = L.np[k + 1] + r.np;
        L.active[k] = L.active[k + 1] + r.active;
      }
      sums.push_back({g, std::move(L)});
      return sums.back().second;
    };
    for (auto& r : profPending) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
        double np = 0, active = 0;
        if (r.batch) {
          const LoopSums& L = sums_of(r.batch);
          if (r.iter == kProfOnce) {
            np = L.npAll;
            active = L.activeAll;
          } else if (r.iter == kProfSmall) {
            np = L.npSmall;
            active = L.activeSmall;
          } else if (r.iter == kProfSetup) {
            np = L.npSetup;
            active = L.activeSetup;
          } else {
            // a batched body takes part in CR-loop launch `iter` if its solve had not finished:
            // iter < iterations for the x update, iter + 1 < iterations for the others
            const int lim = r.kind == PK_CR_X ? r.iter : r.iter + 1;
            const size_t k = std::upper_bound(L.its.begin(), L.its.end(), lim) - L.its.begin();
            np = L.np[k];
            active = L.active[k];
          }
        }
        profMs[r.kind] += ms;
        profCount[r.kind] += 1;
        profBytes[r.kind] += kProfBytesPerParticle[r.kind] * np + kProfBytesPerNode[r.kind] * active;
      }
      eventPool.push_back(r.a);
      eventPool.push_back(r.b);
    }
    profPending.clear();
  }
  ~mpm_context() {
    for (auto& r : profPending) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : eventPool) (void)hipEventDestroy(e);
  }
};

hipStream_t GridBatch::ctx_stream() const { return ctx->stream; }

// ------------------------------------------------------------------ GridBatch
// Grid::Grid (Grid.cpp:156-263) for every body of the batch: bbox and frame per body,
// geometry, binning into the batch's tiles, mass and velocity splat.
void GridBatch::build(const int* slotList, const int* labelList, int n, int nbodies, float h, const float* frameV,
                      bool computeFrame, bool indexOrder) {
  mpm_context& c = *ctx;
  hipStream_t s = c.stream;
  nb = nbodies;
  if (nb <= 0) throw InvalidArg("a grid needs at least one body");
  if (n <= 0) throw InvalidArg("grid needs at least one particle");
  // per-body reductions
  bbox.ensure(sizeof(unsigned int) * 6 * nb);
  mom.ensure(sizeof(double) * 4 * nb);
  cnt.ensure(sizeof(int) * nb);
  const size_t nrep = (size_t)body_replicas() * nb;
  bboxR.ensure(sizeof(unsigned int) * 6 * nrep);
  momR.ensure(sizeof(double) * 4 * nrep);
  cntR.ensure(sizeof(int) * nrep);
  launch_body_reduce(c.view(), slotList, labelList, n, nb, bbox.as<unsigned int>(), mom.as<double>(), cnt.as<int>(),
                     bboxR.as<unsigned int>(), momR.as<double>(), cntR.as<int>(), s);
  std::vector<unsigned int> hbox(6 * nb);
  std::vector<double> hmom(4 * nb);
  std::vector<int> hcnt(nb);
  HIPCHECK(hipMemcpyAsync(hbox.data(), bbox.p, sizeof(unsigned int) * 6 * nb, hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(hmom.data(), mom.p, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(hcnt.data(), cnt.p, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
  c.sync();
  // geometry (Grid.cpp:193-213) in the reference's float arithmetic
  hb.assign(nb, BodyGrid{});
  int tileBase = 0;
  np = 0;
  for (int b = 0; b < nb; ++b) {
    BodyGrid& G = hb[b];
    G.h = h;
    G.np = hcnt[b];
    np += hcnt[b];
    if (hcnt[b] <= 0) throw std::runtime_error("empty body (internal error)");
    for (int a = 0; a < 3; ++a) {
      float mn = decode_ord(hbox[6 * b + a]), mx = decode_ord(hbox[6 * b + 3 + a]);
      int cellMin = (int)std::floor(mn / h) - 2 - 1;
      int cellMax = (int)std::ceil(mx / h) + 2 + 1;
      G.gmin[a] = (float)cellMin * h;
      G.n[a] = cellMax - cellMin;
      G.fv[a] = computeFrame ? (float)(hmom[4 * b + a] / hmom[4 * b + 3]) : (frameV ? frameV[3 * b + a] : 0.0f);
    }
    long long ncells = (long long)G.n[0] * G.n[1] * G.n[2];
    if (ncells > 4000000000LL || ncells <= 0) throw std::runtime_error("grid is too big");
    for (int a = 0; a < 3; ++a) G.T[a] = (G.n[a] + 3) / 4;
    long long nt = (long long)G.T[0] * G.T[1] * G.T[2];
    nt = (nt + 3) / 4 * 4;
    if ((long long)tileBase + nt > INT32_MAX / 256) throw std::runtime_error("grid is too big");
    G.NT = (int)nt;
    G.nblocks = G.NT / 4;
    tileBase += G.NT;
  }
  // tile layout: the bodies the batched CR loop iterates first, the small-body solver's after, so
  // that the loop's launches cover a prefix of the chunks and node blocks
  smallLayout.assign(nb, 0);
  std::vector<int>& order = layoutOrder;
  order.clear();
  order.reserve(nb);
  tileBase = 0;
  for (int pass = 0; pass < 2; ++pass) {
    for (int b = 0; b < nb; ++b) {
      BodyGrid& G = hb[b];
      const int small = !indexOrder && G.nblocks <= kSmallBlocks && G.np <= c.smallMaxParticles ? 1 : 0;
      if (small != pass) continue;
      smallLayout[b] = (char)small;
      order.push_back(b);
      G.tileBase = tileBase;
      G.blockBase = tileBase / 4;
      tileBase += G.NT;
    }
    if (pass == 0) prefixTiles = tileBase;
  }
  NT = tileBase;
  if (NT >= (1 << kKeyCodeShift)) throw std::runtime_error("grid is too big");
  nblocks = NT / 4;
  NN = (int64_t)NT * 64;
  levelsBuilt = false;
  ++version;
  bodies.ensure(sizeof(BodyGrid) * nb);
  HIPCHECK(hipMemcpyAsync(bodies.p, hb.data(), sizeof(BodyGrid) * nb, hipMemcpyHostToDevice, s));
  tileBody.ensure(sizeof(int) * NT);
  blockBody.ensure(sizeof(int) * nblocks);
  bodyOrder.ensure(sizeof(int) * nb);
  HIPCHECK(hipMemcpyAsync(bodyOrder.p, order.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
  launch_tile_body(bodies.as<BodyGrid>(), bodyOrder.as<int>(), nb, NT, tileBody.as<int>(), blockBody.as<int>(), s);
  // binning + counting sort by global tile
  key.ensure(sizeof(int) * n);
  rank.ensure(sizeof(int) * n);
  perm.ensure(sizeof(int) * np);
  pcode.ensure(np);
  count.ensure(sizeof(int) * NT);
  offset.ensure(sizeof(int) * NT);
  nch.ensure(sizeof(int) * NT);
  chunkOff.ensure(sizeof(int) * NT);
  scantmp.ensure(sizeof(int) * scan_tmp_ints(NT));
  totals.ensure(sizeof(int) * 4);
  HIPCHECK(hipMemsetAsync(count.p, 0, sizeof(int) * NT, s));
  c.clear_err();
  launch_bin(c.view(), slotList, labelList, n, bodies.as<BodyGrid>(), key.as<int>(), rank.as<int>(), count.as<int>(),
             c.errp(), s);
  if (keepKey) {  // k_sort_tiles reuses key as scratch below
    keepKey->ensure(sizeof(int) * n);
    HIPCHECK(hipMemcpyAsync(keepKey->p, key.p, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
  }
  scan_exclusive(count.as<int>(), offset.as<int>(), NT, totals.as<int>(), scantmp.as<int>(), s);
  launch_chunk_counts(count.as<int>(), NT, nch.as<int>(), s);
  scan_exclusive(nch.as<int>(), chunkOff.as<int>(), NT, totals.as<int>() + 1, scantmp.as<int>(), s);
  launch_scatter(slotList, key.as<int>(), rank.as<int>(), offset.as<int>(), n, perm.as<int>(), pcode.as<int8_t>(), s);
  launch_sort_tiles(count.as<int>(), offset.as<int>(), perm.as<int>(), pcode.as<int8_t>(), key.as<int>(),
                    rank.as<int>(), NT, s);
  int tot[2] = {0, 0};
  int e = 0, pc = 0;
  HIPCHECK(hipMemcpyAsync(tot, totals.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(&e, c.errflag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  if (prefixTiles < NT)
    HIPCHECK(hipMemcpyAsync(&pc, chunkOff.as<int>() + prefixTiles, sizeof(int), hipMemcpyDeviceToHost, s));
  c.sync();
  raise_errors(e);
  if (tot[0] != np) throw std::runtime_error("particle binning lost particles (internal error)");
  nchunks = tot[1];
  prefixChunks = prefixTiles < NT ? pc : nchunks;
  chunks.ensure(sizeof(int4) * std::max(nchunks, 1));
  launch_make_chunks(count.as<int>(), offset.as<int>(), chunkOff.as<int>(), tileBody.as<int>(), NT, chunks.as<int4>(),
                     s);
  // node arrays and staging
  stageA.ensure(sizeof(float) * (size_t)kStage * std::max(nchunks, 1));
  stageB.ensure(sizeof(float) * (size_t)kStage * std::max(nchunks, 1));
  mass.ensure(sizeof(float) * NN);
  flags.ensure(NN);
  flags2.ensure(NN);
  for (DevBuf* d : {&vel, &prev, &vc, &b, &x, &z, &p, &Ap, &Ar, &diag, &tmp3, &nrm}) d->ensure(vecBytes());
  scal.ensure(sizeof(CRScalars) * nb);
  status.ensure(sizeof(BatchStatus));
  partials.ensure(sizeof(double) * kPartialStride * std::max(nblocks, 1));
  tileAct.ensure(std::max(NT, 1));
  HIPCHECK(hipMemsetAsync(tileAct.p, 1, std::max(NT, 1), s));
  // reduction tickets return to zero after every use (the last arrival resets each counter), so
  // they are cleared only when the buffer is (re)allocated
  const void* oldTickets = tickets.p;
  tickets.ensure(sizeof(unsigned int) * kTicketsPerBody * nb);
  if (tickets.p != oldTickets) HIPCHECK(hipMemsetAsync(tickets.p, 0, tickets.cap, s));
  HIPCHECK(hipMemsetAsync(flags.p, 0xFF, NN, s));  // -1: no collision until updateGridVelocities
  rec.assign(nb, SolveRecord{});
  for (int bi = 0; bi < nb; ++bi) rec[bi].np = hb[bi].np;
  Batch B = view();
  int pr = c.prof_begin(PK_P2G_MASS, this);
  launch_p2g_mass(c.view(), B, f(stageA), c.errp(), s);
  c.prof_end(pr);
  launch_gather_mass(B, f(stageA), f(mass), c.errp(), s);
  pr = c.prof_begin(PK_P2G_VEL, this);
  launch_p2g_vel(c.view(), B, f(mass), f(stageA), c.errp(), s);
  c.prof_end(pr);
  launch_gather_vel(B, f(stageA), f(vel), f(prev), c.errp(), s);
}

void GridBatch::to_dense(int body, const float* batch, int ncomp, float* host) {
  const BodyGrid& G = hb.at(body);
  int64_t nd = (int64_t)G.n[0] * G.n[1] * G.n[2];
  dense.ensure(sizeof(float) * ncomp * nd);
  launch_to_dense_f(batch, ncomp, dense.as<float>(), G, ctx->stream);
  HIPCHECK(hipMemcpyAsync(host, dense.p, sizeof(float) * ncomp * nd, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
}
void GridBatch::from_dense(int body, const float* host, int ncomp, float* batch) {
  const BodyGrid& G = hb.at(body);
  int64_t nd = (int64_t)G.n[0] * G.n[1] * G.n[2];
  dense.ensure(sizeof(float) * ncomp * nd);
  HIPCHECK(hipMemcpyAsync(dense.p, host, sizeof(float) * ncomp * nd, hipMemcpyHostToDevice, ctx->stream));
  HIPCHECK(hipMemsetAsync(batch, 0, ncomp == 3 ? vecBytes() : sizeof(float) * NN, ctx->stream));
  launch_from_dense_f(dense.as<float>(), ncomp, batch, G, ctx->stream);
}

void GridBatch::forces_stage() {
  int pr = ctx->prof_begin(PK_FORCES, this);
  launch_forces(ctx->view(), view(), f(stageA), ctx->errp(), ctx->stream);
  ctx->prof_end(pr);
}
void GridBatch::hvp(const float* d, bool project, float* stage, const CRScalars* scp, int iter) {
  hvp(view(), d, project, stage, scp, iter);
}
void GridBatch::hvp(const Batch& Bv, const float* d, bool project, float* stage, const CRScalars* scp, int iter,
                    bool paired) {
  int pr = ctx->prof_begin(PK_HVP, this, iter);
  if (paired)
    launch_hvp_pair(ctx->view(), Bv, groups.as<int2>(), totals.as<int>() + 2, d, project, flags.as<int8_t>(),
                    ctx->cols(), stage, scp, ctx->errp(), ctx->stream);
  else
    launch_hvp(ctx->view(), Bv, d, project, flags.as<int8_t>(), ctx->cols(), stage, scp, ctx->errp(), ctx->stream);
  ctx->prof_end(pr);
}
void GridBatch::pair_chunks(const Batch& Bv) {
  const int npairs = std::max(1, (Bv.nchunks + 1) / 2);
  pairCnt.ensure(sizeof(int) * npairs);
  pairPos.ensure(sizeof(int) * npairs);
  groups.ensure(sizeof(int2) * std::max(1, Bv.nchunks));
  scantmp.ensure(sizeof(int) * scan_tmp_ints(npairs));
  launch_pair_chunks(Bv, pairCnt.as<int>(), pairPos.as<int>(), groups.as<int2>(), totals.as<int>() + 2,
                     scantmp.as<int>(), ctx->stream);
}

void GridBatch::fetch_history(int body, SolveRecord& r) {
  int L = std::min(r.iterations, kHistCap);
  r.history.resize(L);
  if (L > 0) {
    HIPCHECK(hipMemcpyAsync(r.history.data(), hist.as<double>() + (int64_t)body * kHistCap, sizeof(double) * L,
                            hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  }
}

// Grid::updateGridVelocities (Grid.cpp:889-1016) for every body of the batch
void GridBatch::solve(float dt, const mpm_solver_params* prm) {
  mpm_context& c = *ctx;
  hipStream_t s = c.stream;
  const bool hooks = prm && (prm->termination || prm->debug);
  int maxIters = prm ? prm->max_iters : 0;
  float tol = prm ? prm->tolerance : 0.0f;
  if (prm && prm->termination) {
    maxIters = INT32_MAX;
    tol = 0.0f;
  }
  Batch B = view();
  // per-body scalars
  size_t scb = sizeof(CRScalars) * nb;
  c.pinned.ensure(scb + 4096 + sizeof(int) * nb);
  CRScalars* hsc = c.pinned.as<CRScalars>();
  std::memset(hsc, 0, scb);
  for (int bi = 0; bi < nb; ++bi) {
    hsc[bi].max_iters = maxIters;
    rec[bi].small = false;
  }
  HIPCHECK(hipMemcpyAsync(scal.p, hsc, scb, hipMemcpyHostToDevice, s));
  HIPCHECK(hipMemsetAsync(status.p, 0, sizeof(BatchStatus), s));
  hist.ensure(sizeof(double) * kHistCap * nb);
  c.clear_err();
  // explicit momenta, flags, vc, guess sums
  forces_stage();
  launch_node_explicit(B, f(mass), f(vel), f(stageA), c.grav(), c.ngrav(), dt, c.cols(), f(b), flags.as<int8_t>(),
                       f(vc), sc(), st(), partials.as<double>(), tileAct.as<unsigned char>(), f(nrm), s);
  // DF(vc) (no input projection) and the diagonal, then rhs / guess / diagonal finish
  hvp(f(vc), false, f(stageA), nullptr, -1);
  int pr = c.prof_begin(PK_DIAG, this);
  launch_diag(c.view(), B, f(stageB), c.errp(), s);
  c.prof_end(pr);
  launch_rhs(B, f(mass), flags.as<int8_t>(), c.cols(), f(stageA), f(stageB), dt, tol, f(b), f(x), f(diag), sc(), st(),
             partials.as<double>(), s);
  BatchStatus* hst = reinterpret_cast<BatchStatus*>(reinterpret_cast<char*>(c.pinned.p) + scb);
  auto fetch_status = [&]() {
    HIPCHECK(hipMemcpyAsync(hst, status.p, sizeof(BatchStatus), hipMemcpyDeviceToHost, s));
    c.sync();
  };
  // status and error flag in one wait
  auto fetch_status_errors = [&]() {
    int* e = reinterpret_cast<int*>(reinterpret_cast<char*>(hst) + sizeof(BatchStatus));
    HIPCHECK(hipMemcpyAsync(hst, status.p, sizeof(BatchStatus), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(e, c.errflag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    c.sync();
    raise_errors(*e);
  };
  const int64_t iterCap = (prm && prm->termination) ? 1000000 : (int64_t)maxIters + 2;
  const bool learned = c.lpcMode == 1 && !hooks;
  // learned mode: once no unfinished body uses the network any more, the rest of the solve runs
  // the CR recurrences below from the current iterate (two node passes per iteration instead of
  // the flexible CR's four)
  bool handoff = false;
  int iterAtHandoff = 0;
  if (learned) {
    // flexible conjugate residual with the learned preconditioner (mpm_kernels.hip, k_fcr_*)
    for (DevBuf* d : {&fr, &fp0, &fp1, &fq0, &fq1}) d->ensure(vecBytes());
    const int mdirs = c.lpcDirs;
    const int64_t vs = 3 * (int64_t)NN;  // floats between stored direction vectors
    fP.ensure(vecBytes() * mdirs);
    fQ.ensure(vecBytes() * mdirs);
    const int netOn = c.net.initialised ? 1 : 0;
    pair_chunks(B);
    hvp(B, f(x), true, f(stageA), sc(), -1, true);
    launch_fcr_init(B, f(mass), flags.as<int8_t>(), c.cols(), f(stageA), f(x), f(b), f(diag), dt, f(fr), netOn, c.lpcFallback, mdirs, c.lpcStrikes, sc(),
                    st(), partials.as<double>(), s);
    fetch_status_errors();
    bool netActive = netOn != 0;
    if (netActive) {
      ensure_levels();
      c.net.alloc_acts(levels, false);
    }
    int it = 0;
    while (hst->bodies_done < nb) {
      if (it > iterCap) throw std::runtime_error("conjugate residual did not terminate");
      int batch = it == 0 ? std::max(1, (c.lpcHandoff && netActive) ? std::min(c.iterHintFcr, 8) : c.iterHintFcr) : 8;
      for (int k = 0; k < batch; ++k, ++it) {
        if (netActive) {
          int pn = c.prof_begin(PK_LPC, this, it);
          c.net.forward(B, levels, f(fr), f(diag), f(mass), flags.as<int8_t>(), nullptr, sc(), nullptr, f(z), s);
          c.prof_end(pn);
        }
        launch_fcr_diag_z(B, f(fr), f(diag), f(z), sc(), s);
        hvp(B, f(z), true, f(stageA), sc(), it, true);
        launch_fcr_az(B, f(mass), flags.as<int8_t>(), c.cols(), f(stageA), f(z), f(diag), dt, f(Ar), f(fQ), vs,
                      sc(), st(), partials.as<double>(), s);
        launch_fcr_orth(B, f(z), f(Ar), f(fr), f(diag), f(fP), f(fQ), vs, sc(), st(), partials.as<double>(), s);
        launch_fcr_update(B, flags.as<int8_t>(), c.cols(), f(mass), f(diag), f(fP), f(fQ), vs, f(x), f(fr), sc(),
                          st(), partials.as<double>(), hist.as<double>(), s);
      }
      fetch_status();
      if (netActive) {
        // the network runs only while some unfinished body still uses it
        HIPCHECK(hipMemcpyAsync(hsc, scal.p, scb, hipMemcpyDeviceToHost, s));
        c.sync();
        // (a body off the network for one iteration only, lpc_pause, still uses it)
        netActive = false;
        for (int bi = 0; bi < nb; ++bi)
          if (!hsc[bi].done && (!hsc[bi].lpc_off || hsc[bi].lpc_pause)) netActive = true;
        if (!netActive && c.lpcHandoff && hst->bodies_done < nb) {
          handoff = true;
          iterAtHandoff = INT32_MAX;
          for (int bi = 0; bi < nb; ++bi)
            if (!hsc[bi].done) iterAtHandoff = std::min(iterAtHandoff, hsc[bi].iter);
          break;
        }
      }
    }
  }
  if (!learned || handoff) {
    // bodies of a few node blocks run their whole solve in one workgroup each, on the second
    // stream, while the batched loop below iterates the rest
    int nsmall = 0;
    // (the iteration cap bounds the persistent kernel's run time)
    // (after a handoff: the small bodies still iterating, whose solves restart from their iterate;
    // hsc holds the scalars read at the handoff)
    if (!hooks && (handoff || c.cap.want == 0) && c.smallMaxParticles > 0 && maxIters <= (1 << 16)) {
      int* list = reinterpret_cast<int*>(reinterpret_cast<char*>(c.pinned.p) + scb + 4096);
      for (int bi = 0; bi < nb; ++bi) {
        rec[bi].small = smallLayout[bi] != 0 && !(handoff && hsc[bi].done);
        if (rec[bi].small) list[nsmall++] = bi;
      }
      if (nsmall > 0) {
        smallList.ensure(sizeof(int) * nsmall);
        HIPCHECK(hipMemcpyAsync(smallList.p, list, sizeof(int) * nsmall, hipMemcpyHostToDevice, s));
        launch_small_claim(smallList.as<int>(), nsmall, sc(), st(), s);
        HIPCHECK(hipEventRecord(c.smallReady, s));
        HIPCHECK(hipStreamWaitEvent(c.stream2, c.smallReady, 0));
        int ps = c.prof_begin(PK_CR_SMALL, this, kProfSmall, c.stream2);
        launch_small_cr(c.view(), B, smallList.as<int>(), nsmall, f(mass), flags.as<int8_t>(), c.cols(), dt, f(stageA),
                        f(b), f(diag), f(x), f(z), f(p), f(Ap), f(Ar), sc(), hist.as<double>(), c.errp(), c.stream2);
        c.prof_end(ps);
      }
    }
    // with the small-body solver running, the batched launches cover the prefix of the layout
    Batch Bb = B;
    if (nsmall > 0) {
      Bb.nblocks = prefixTiles / 4;
      Bb.nchunks = prefixChunks;
    }
    // CR setup (ConjugateResiduals.cpp:50-76); the HVPs of the solve run over chunk pairs. After
    // a handoff it starts from the flexible CR's iterate, r = b - A x recomputed; the bodies'
    // iteration indices, histories and caps continue
    if (!handoff || nsmall > 0) pair_chunks(Bb);
    hvp(Bb, f(x), true, f(stageA), sc(), kProfSetup, true);
    launch_cr_setup1(Bb, f(mass), flags.as<int8_t>(), c.cols(), f(stageA), f(x), f(b), f(diag), dt, f(z), f(p), sc(),
                     s);
    hvp(Bb, f(z), true, f(stageA), sc(), kProfSetup, true);
    launch_cr_setup2(Bb, f(mass), flags.as<int8_t>(), c.cols(), f(stageA), f(z), f(diag), dt, f(Ap), sc(), st(),
                     partials.as<double>(), s);
    fetch_status_errors();
    // training pairs: iterate x_k and residual r_k = M z_k at the iterations of the capture schedule
    LpcCapture& cap = c.cap;
    const std::vector<int>& kCaptureIters = cap.schedule;
    const int capWant = (hooks || handoff) ? 0 : std::min<int>(cap.want, (int)kCaptureIters.size());
    if (capWant > 0) {
      cap.NN = NN;
      cap.data.ensure(sizeof(float) * 3 * (size_t)NN * capWant);
      if (cap.lossMode == 1) cap.xdata.ensure(sizeof(float) * 3 * (size_t)NN * capWant);
      cap.iters.clear();
      cap.n = 0;
    }
    auto snapshot = [&](int iterIndex) {
      if (cap.n >= capWant || kCaptureIters[cap.n] != iterIndex) return;
      launch_residual_vector(B, f(z), f(diag), cap.r(cap.n), s);
      if (cap.lossMode == 1)
        HIPCHECK(hipMemcpyAsync(cap.xk(cap.n), f(x), vecBytes(), hipMemcpyDeviceToDevice, s));
      cap.iters.push_back(iterIndex);
      ++cap.n;
    };
    if (capWant > 0) snapshot(0);
    // one CR iteration (ConjugateResiduals.cpp:84-151) for every body still iterating
    // iteration `it` from its x update on; from the second iteration on, the previous
    // iteration's p / Ap update runs in the same pass (k_cr_update_px)
    auto iterate_x = [&](int it) {
      if (it == 0) {
        int px = c.prof_begin(PK_CR_X, this, it);
        launch_cr_update_x(Bb, flags.as<int8_t>(), c.cols(), f(p), f(Ap), f(diag), f(x), f(z), sc(), st(),
                           partials.as<double>(), hist.as<double>(), s);
        c.prof_end(px);
      } else {
        int pp = c.prof_begin(PK_CR_PX, this, it);
        launch_cr_update_px(Bb, flags.as<int8_t>(), c.cols(), f(Ar), f(diag), f(p), f(Ap), f(x), f(z), sc(), st(),
                            partials.as<double>(), hist.as<double>(), s);
        c.prof_end(pp);
      }
      if (capWant > 0) snapshot(it + 1);
    };
    auto iterate_rest = [&](int it) {
      hvp(Bb, f(z), true, f(stageA), sc(), it, true);
      int pa = c.prof_begin(PK_CR_AR, this, it);
      launch_cr_ar(Bb, f(mass), flags.as<int8_t>(), c.cols(), f(stageA), f(z), f(Ap), f(diag), dt, f(Ar), sc(), st(),
                   partials.as<double>(), s);
      c.prof_end(pa);
    };
    auto set_done_flag = [&](int bi, int v) {
      hsc[bi].done = v;
      HIPCHECK(hipMemcpyAsync(&sc()[bi].done, &hsc[bi].done, sizeof(int), hipMemcpyHostToDevice, s));
      c.sync();
    };
    if (!hooks) {
      // all bodies iterate in lockstep; the host polls the done counter once per batch of iterations
      int it = 0;
      while (hst->bodies_done + hst->small_bodies < nb) {
        if (it > iterCap) throw std::runtime_error("conjugate residual did not terminate");
        int batch = it == 0 ? std::max(1, handoff ? c.iterHintHandoff : c.iterHint) : 8;
        for (int k = 0; k < batch; ++k, ++it) {
          iterate_x(it);
          iterate_rest(it);
        }
        fetch_status();
      }
      if (nsmall > 0) {
        HIPCHECK(hipEventRecord(c.smallDone, c.stream2));
        HIPCHECK(hipStreamWaitEvent(s, c.smallDone, 0));
      }
      if (capWant > 0 && cap.n > 0) train_lpc(dt, hsc, scb);
    } else {
      // host hooks see the bodies one after another, as the reference solves them: bodies still
      // iterating are parked (done = 2, which every kernel skips) and released one at a time
      HIPCHECK(hipMemcpyAsync(hsc, scal.p, scb, hipMemcpyDeviceToHost, s));
      c.sync();
      std::vector<int> pending;
      for (int bi = 0; bi < nb; ++bi)
        if (!hsc[bi].done) {
          pending.push_back(bi);
          set_done_flag(bi, 2);
        }
      std::vector<float> hostvec;
      for (int bi : pending) {
        set_done_flag(bi, 0);
        const BodyGrid& G = hb[bi];
        const int64_t nd = 3LL * G.n[0] * G.n[1] * G.n[2];
        hostvec.resize(nd);
        for (int it = 0;; ++it) {
          if (it > iterCap) throw std::runtime_error("conjugate residual did not terminate");
          iterate_x(it);
          HIPCHECK(hipMemcpyAsync(&hsc[bi], &sc()[bi], sizeof(CRScalars), hipMemcpyDeviceToHost, s));
          c.sync();
          if (hsc[bi].iterations == it + 1) {  // this x update happened
            if (prm->debug) {
              to_dense(bi, f(x), 3, hostvec.data());
              prm->debug(prm->user, hostvec.data(), nd);
            }
            if (prm->termination && !hsc[bi].done) {
              launch_residual_vector(B, f(z), f(diag), f(tmp3), s);
              to_dense(bi, f(tmp3), 3, hostvec.data());
              if (prm->termination(prm->user, hostvec.data(), nd, it)) set_done_flag(bi, 1);
            }
          }
          if (hsc[bi].done) break;
          iterate_rest(it);
        }
      }
    }
  }
  launch_finalize(B, f(x), f(vc), f(vel), sc(), s);
  HIPCHECK(hipMemcpyAsync(hsc, scal.p, scb, hipMemcpyDeviceToHost, s));
  c.sync();
  int maxIt = 0;
  for (int bi = 0; bi < nb; ++bi) {
    SolveRecord& r = rec[bi];
    r.iterations = hsc[bi].iterations;
    r.early_exit = hsc[bi].early_exit;
    r.rhs_norm2 = hsc[bi].bnorm2;
    r.residual_norm2 = hsc[bi].rnorm2;
    r.active = hsc[bi].active_nodes;
    r.fallbacks = learned ? hsc[bi].fallbacks : 0;
    r.restarts = hsc[bi].restarts;
    r.history.clear();
    if (!r.small) maxIt = std::max(maxIt, r.iterations);
    raise_errors(hsc[bi].errors);
  }
  if (handoff) {
    c.iterHintFcr = std::max(1, iterAtHandoff);
    c.iterHintHandoff = std::max(1, maxIt - iterAtHandoff);
  } else if (maxIt > 0) {
    (learned ? c.iterHintFcr : c.iterHint) = maxIt;
  }
}

// Train the learned preconditioner on the residuals captured in this solve, before the particle
// update, while the system's operator is still in place: Adam on the per-body residual left by
// the optimal step along A z, 1 - <r, A z>_D^2 / (|r|_D^2 |A z|_D^2), z = r / diag + s_b net(r)
// (mpm_kernels.hip k_lpc_*).
void GridBatch::train_lpc(float dt, CRScalars* hsc, size_t scb) {
  mpm_context& c = *ctx;
  hipStream_t s = c.stream;
  LpcCapture& cap = c.cap;
  const Batch B = view();
  if (!c.net.initialised) c.net.init(1, s);
  ensure_levels();
  c.net.alloc_acts(levels, true);
  for (DevBuf* d : {&fp0, &fp1}) d->ensure(vecBytes());
  cap.sums.ensure(sizeof(double) * 3 * nb * cap.n);
  HIPCHECK(hipMemsetAsync(cap.sums.p, 0, sizeof(double) * 3 * nb * cap.n, s));
  for (int k = 0; k < cap.n; ++k)
    launch_lpc_sample_stats(B, cap.r(k), f(diag), f(mass), cap.sums.as<double>() + 3 * nb * k, s);
  std::vector<double> sums(3 * nb * cap.n);
  HIPCHECK(hipMemcpyAsync(sums.data(), cap.sums.p, sizeof(double) * sums.size(), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(hsc, scal.p, scb, hipMemcpyDeviceToHost, s));
  c.sync();
  const bool exact = cap.lossMode == 1;
  std::vector<double> energy;
  if (exact) {
    // each sample's error energy e^T A e = (x* - x_k) . r_k, x* the solve's final iterate
    cap.energy.ensure(sizeof(double) * nb * cap.n);
    HIPCHECK(hipMemsetAsync(cap.energy.p, 0, sizeof(double) * nb * cap.n, s));
    for (int k = 0; k < cap.n; ++k)
      launch_lpc_exact_energy(B, f(x), cap.xk(k), cap.r(k), cap.energy.as<double>() + nb * k, s);
    energy.resize(nb * cap.n);
    HIPCHECK(hipMemcpyAsync(energy.data(), cap.energy.p, sizeof(double) * energy.size(), hipMemcpyDeviceToHost, s));
    c.sync();
  }
  std::vector<float> hs(nb * cap.n), hw(nb * cap.n);
  std::vector<int> useful;
  for (int k = 0; k < cap.n; ++k) {
    bool any = false;
    for (int bi = 0; bi < nb; ++bi) {
      const double* q = &sums[3 * (nb * k + bi)];
      // bodies of at least 256 particles still iterating at this sample; for the exact-correction
      // loss, 4 or more iterations before the end, so that x* - x_k is well above x*'s own error
      const int ahead = exact ? 4 : 0;
      const bool use = hb[bi].np >= 256 && hsc[bi].iterations > cap.iters[k] + ahead && q[2] > 0.0 && q[1] > 0.0 &&
                       (!exact || energy[nb * k + bi] > 0.0);
      hs[nb * k + bi] = (float)std::sqrt(q[1] > 0.0 ? q[0] / q[1] : 0.0);
      hw[nb * k + bi] = use ? 1.0f : 0.0f;
      any = any || use;
    }
    if (any) useful.push_back(k);
  }
  cap.scale.ensure(sizeof(float) * hs.size());
  cap.weight.ensure(sizeof(float) * hw.size());
  cap.loss.ensure(sizeof(double) * (1 + 3 * nb));
  HIPCHECK(hipMemcpyAsync(cap.scale.p, hs.data(), sizeof(float) * hs.size(), hipMemcpyHostToDevice, s));
  HIPCHECK(hipMemcpyAsync(cap.weight.p, hw.data(), sizeof(float) * hw.size(), hipMemcpyHostToDevice, s));
  const float* sk = cap.scale.as<float>();
  const float* wk = cap.weight.as<float>();
  const int8_t* fl = flags.as<int8_t>();
  double* dl = cap.loss.as<double>();
  double* dots = dl + 1;
  const double* ek = exact ? cap.energy.as<double>() : nullptr;
  // the loss of step z = r / diag + scale clamp(raw) for sample k; with the exact-correction loss
  // also the network output's gradient into g (when given)
  auto step_loss = [&](int k, const float* raw, __bf16* g) {
    launch_lpc_compose_z(B, cap.r(k), f(diag), raw, sk + nb * k, wk + nb * k, f(fp0), s);
    hvp(f(fp0), true, f(stageA), nullptr, -1);
    if (exact) {
      HIPCHECK(hipMemsetAsync(dots, 0, sizeof(double) * 2 * nb, s));
      launch_lpc_energy_dots(B, f(mass), fl, c.cols(), f(stageA), f(fp0), cap.r(k), dt, wk + nb * k, f(Ar), dots, s);
      launch_lpc_energy_grad(B, f(Ar), cap.r(k), raw, sk + nb * k, wk + nb * k, dots, ek + nb * k, g, dl, s);
      return;
    }
    HIPCHECK(hipMemsetAsync(dots, 0, sizeof(double) * 3 * nb, s));
    launch_lpc_res_dots(B, f(mass), fl, c.cols(), f(stageA), f(fp0), cap.r(k), f(diag), dt, wk + nb * k, f(Ar),
                        dots, s);
    launch_lpc_res_y(B, f(Ar), cap.r(k), f(diag), wk + nb * k, dots, f(fp1), dl, s);
  };
  auto fetch_loss = [&]() {
    double h = 0.0;
    HIPCHECK(hipMemcpyAsync(&h, dl, sizeof(double), hipMemcpyDeviceToHost, s));
    c.sync();
    return h;
  };
  // loss of the diagonal step alone
  HIPCHECK(hipMemsetAsync(dl, 0, sizeof(double), s));
  for (int k : useful) step_loss(k, nullptr, nullptr);
  int nbodies_used = 0;
  for (int k : useful)
    for (int bi = 0; bi < nb; ++bi) nbodies_used += hw[nb * k + bi] > 0.0f ? 1 : 0;
  const double norm = nbodies_used > 0 ? 1.0 / nbodies_used : 0.0;
  cap.pairs = nbodies_used;
  cap.lossDiag = fetch_loss() * norm;
  for (int ep = 0; ep < cap.epochs; ++ep) {
    HIPCHECK(hipMemsetAsync(dl, 0, sizeof(double), s));
    for (int k : useful) {
      float* raw = c.net.raw.as<float>();
      c.net.forward(B, levels, cap.r(k), f(diag), f(mass), fl, sk + nb * k, nullptr, raw, nullptr, s);
      step_loss(k, raw, c.net.g6.as<bf16>());
      if (!exact) {
        hvp(f(fp1), true, f(stageA), nullptr, -1);
        launch_lpc_res_grad(B, f(mass), fl, c.cols(), f(stageA), f(fp1), raw, dt, sk + nb * k, wk + nb * k,
                            c.net.g6.as<bf16>(), s);
      }
      c.net.backward(B, levels, s);
      c.net.adam(cap.lr, s);
    }
    const double l = fetch_loss() * norm;
    if (ep == 0) cap.lossFirst = l;
    cap.lossLast = l;
  }
  cap.want = 0;
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

// Sim::calculateBodies (Sim.cpp:136-203) on the GPU: connected components of the relation
// |x_i - x_j|^2 < h^2 (float arithmetic of KdTree.inl:187-190) by cell hashing and
// sampled union-find; isolated particles are ballistic; bodies are ordered by their smallest
// particle index, as the reference's flood fill numbers them.
// keepErrors: the error flag still holds the step's earlier errors, raised by this wait
void calculate_bodies(mpm_context& c, bool keepErrors = false) {
  const int n = (int)c.n;
  hipStream_t s = c.stream;
  if (!keepErrors) c.clear_err();
  launch_check_pvel(c.view(), n, c.errp(), s);
  uint32_t tsize = 1;
  while (tsize < 2u * (uint32_t)n) tsize <<= 1;
  const uint32_t mask = tsize - 1;
  // any pair closer than h lies in adjacent cells: the 1% margin covers the rounding of
  // x / cell for |x| / h up to ~1e5
  const float cell = c.h * 1.01f;
  c.ccBucket.ensure(sizeof(int) * n);
  c.ccRank.ensure(sizeof(int) * n);
  c.ccCount.ensure(sizeof(int) * tsize);
  c.ccOffset.ensure(sizeof(int) * tsize);
  c.ccSorted.ensure(sizeof(float4) * n);
  c.ccParent.ensure(sizeof(int) * n);
  c.ccHasNb.ensure(sizeof(int) * n);
  c.ccIsRoot.ensure(sizeof(int) * n);
  c.ccRootIdx.ensure(sizeof(int) * n);
  c.ccMinOrig.ensure(sizeof(int) * n);
  c.ccTmp.ensure(sizeof(int) * scan_tmp_ints(std::max<int64_t>(tsize, n)));
  c.ccTot.ensure(sizeof(int) * 4);
  c.ccSkip.ensure(kCcSkipBytes);
  c.labels.ensure(sizeof(int) * n);
  HIPCHECK(hipMemsetAsync(c.ccCount.p, 0, sizeof(int) * tsize, s));
  launch_cc_hash(c.view(), n, cell, mask, c.ccBucket.as<int>(), c.ccRank.as<int>(), c.ccCount.as<int>(), s);
  scan_exclusive(c.ccCount.as<int>(), c.ccOffset.as<int>(), (int)tsize, c.ccTot.as<int>(), c.ccTmp.as<int>(), s);
  launch_cc_scatter(c.view(), c.ccBucket.as<int>(), c.ccRank.as<int>(), c.ccOffset.as<int>(), n,
                    c.ccSorted.as<float4>(), s);
  launch_iota(c.ccParent.as<int>(), n, s);
  // pass 1's outgoing links go to ccBucket and ccRootIdx, free between the scatter and the numbering
  // pass 1's roots go to ccIsRoot, free until the numbering
  launch_cc_link(n, c.h, cell, mask, c.ccOffset.as<int>(), c.ccCount.as<int>(), c.ccSorted.as<float4>(),
                 c.ccParent.as<int>(), c.ccHasNb.as<int>(), c.ccIsRoot.as<int>(), c.ccSkip.p, c.ccBucket.as<int>(),
                 c.ccRootIdx.as<int>(), s);
  HIPCHECK(hipMemsetAsync(c.ccIsRoot.p, 0, sizeof(int) * n, s));
  launch_fill_int(c.ccMinOrig.as<int>(), INT_MAX, n, s);
  launch_cc_roots(n, c.ccParent.as<int>(), c.ccHasNb.as<int>(), c.orig_of[c.cur].as<int>(), c.ccIsRoot.as<int>(),
                  c.ccMinOrig.as<int>(), s);
  // bodies numbered by smallest original particle index (the reference's flood order)
  launch_cc_number(n, c.ccParent.as<int>(), c.ccHasNb.as<int>(), c.ccIsRoot.as<int>(), c.ccMinOrig.as<int>(),
                   c.ccRootIdx.as<int>(), c.ccRank.as<int>(), c.ccTot.as<int>() + 1, c.ccTmp.as<int>(),
                   c.labels.as<int>(), s);
  // ballistic count for statistics
  c.ballPos.ensure(sizeof(int) * n);
  launch_is_ballistic(c.labels.as<int>(), n, c.ccIsRoot.as<int>(), s);
  scan_exclusive(c.ccIsRoot.as<int>(), c.ballPos.as<int>(), n, c.ccTot.as<int>() + 2, c.ccTmp.as<int>(), s);
  // one wait: body and ballistic counts and the velocity check
  int tot[2] = {0, 0}, e = 0;
  HIPCHECK(hipMemcpyAsync(tot, c.ccTot.as<int>() + 1, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(&e, c.errflag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  c.sync();
  raise_errors(e);
  c.nbodies = tot[0];
  c.nballistic = tot[1];
  c.bodyListsValid = false;
}

void ensure_body_lists(const mpm_context& cc) {
  mpm_context& c = const_cast<mpm_context&>(cc);
  if (c.bodyListsValid) return;
  const int n = (int)c.n;
  std::vector<int> lab(n), orig(n);
  HIPCHECK(hipMemcpyAsync(lab.data(), c.labels.p, sizeof(int) * n, hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipMemcpyAsync(orig.data(), c.orig_of[c.cur].p, sizeof(int) * n, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  c.bodyLists.assign(c.nbodies, {});
  c.ballisticList.clear();
  for (int s = 0; s < n; ++s) {
    if (lab[s] >= 0) c.bodyLists[lab[s]].push_back(orig[s]);
    else c.ballisticList.push_back(orig[s]);
  }
  for (auto& b : c.bodyLists) std::sort(b.begin(), b.end());
  std::sort(c.ballisticList.begin(), c.ballisticList.end());
  c.bodyListsValid = true;
}

void fill_stats(const SolveRecord& r, mpm_solve_stats* st) {
  if (!st) return;
  st->fallbacks = r.fallbacks;
  st->restarts = r.restarts;
  st->particles = r.np;
  st->active_nodes = r.active;
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

}  // namespace

extern "C" {

int mpm_version(void) { return 1; }

int mpm_create(mpm_context** out, int device) {
  if (!out) return MPM_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return MPM_ERR_NO_DEVICE;
  if (device >= kMaxDevices) return MPM_ERR_INVALID;
  std::unique_ptr<mpm_context> c(new mpm_context());
  c->device = device;
  c->simBatch.ctx = c.get();
  int rc = guarded(c.get(), [&] {
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&c->smallReady, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&c->smallDone, hipEventDisableTiming));
    c->errflag.ensure(sizeof(int));
    c->cols_dev.ensure(sizeof(ColliderSet));
    c->grav_dev.ensure(sizeof(float) * 3 * 64);
    ColliderSet empty{};
    HIPCHECK(hipMemcpy(c->cols_dev.p, &empty, sizeof(empty), hipMemcpyHostToDevice));
    c->scratch = scratch_on(device);  // guarded() made `device` current
  });
  if (rc != MPM_OK) return rc;
  *out = c.release();
  return MPM_OK;
}

int mpm_destroy(mpm_context* ctx) {
  if (!ctx) return MPM_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
  hipStream_t s = ctx->stream, s2 = ctx->stream2;
  if (ctx->smallReady) (void)hipEventDestroy(ctx->smallReady);
  if (ctx->smallDone) (void)hipEventDestroy(ctx->smallDone);
  delete ctx;
  if (s) (void)hipStreamDestroy(s);
  if (s2) (void)hipStreamDestroy(s2);
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
    std::iota(iota.begin(), iota.end(), 0);
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
    launch_plasticity(ctx->view(), nullptr, INT_MIN, (int)ctx->n, ctx->thc, ctx->ths, ctx->errp(), ctx->stream);
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

// The variant of mpm_model_stress_differential that allocates nothing: device-resident inputs
// and output, the slot list in `scratch` (this device's address, taken at mpm_create). The index
// check runs on the device; the call waits for its result like the others.
int mpm_model_stress_differential_noalloc(mpm_context* ctx, const int32_t* particles, const float* dF, int64_t n,
                                          float* out) {
  return guarded(ctx, [&] {
    require_particles(ctx);
    if (!ctx->model) throw InvalidArg("no constitutive model");
    if (n < 0 || n > kNoAllocMaxParticles)
      throw InvalidArg("mpm_model_stress_differential_noalloc takes at most 1048576 particles per call");
    if (n == 0) return;
    if (!particles || !dF || !out) throw InvalidArg("null device pointer");
    launch_fill_int(ctx->errp(), 0, 1, ctx->stream);  // a fill kernel, not a memset node
    launch_gather_slots_checked(ctx->slot_of.as<int>(), particles, (int)n, (int)ctx->n, ctx->scratch, ctx->errp(),
                                ctx->stream);
    launch_probe(ctx->view(), ctx->scratch, (int)n, 2, dF, out, ctx->stream);
    ctx->check_errors();
  });
}

int mpm_set_colliders(mpm_context* ctx, const mpm_collider* cols, int n) {
  return guarded(ctx, [&] {
    if (n < 0 || n > kMaxColliders) throw InvalidArg("at most 16 colliders are supported");
    if (n > 0 && !cols) throw InvalidArg("null collider array");
    // pass 1: validate every collider before any device memory is touched, so that a rejected
    // call leaves the previous collider set (and the volumes its device copy points into) intact
    ColliderSet cs{};
    cs.n = n;
    int64_t npts[kMaxColliders] = {};
    for (int i = 0; i < n; ++i) {
      Collider& o = cs.o[i];
      o.type = cols[i].type;
      if (o.type != 0 && o.type != 1 && o.type != 2) throw InvalidArg("unknown collider type");
      for (int a = 0; a < 4; ++a) o.c[a] = cols[i].coeffs[a];
      for (int a = 0; a < 3; ++a) {
        o.center[a] = cols[i].center[a];
        o.v[a] = cols[i].velocity[a];
      }
      o.radius = cols[i].radius;
      o.friction = cols[i].friction;
      o.sticky = cols[i].sticky;
      o.phi = nullptr;
      o.vel = nullptr;
      if (o.type == 2) {
        const mpm_collider& q = cols[i];
        if (!q.phi) throw InvalidArg("level-set collider without phi");
        if (!(q.voxel > 0.0f)) throw InvalidArg("level-set voxel size must be positive");
        int64_t np = 1;
        for (int a = 0; a < 3; ++a) {
          if (q.dims[a] < 2 || q.dims[a] > (1 << 12)) throw InvalidArg("level-set dims must be in [2, 4096]");
          np *= q.dims[a];
          o.dims[a] = q.dims[a];
          o.origin[a] = q.origin[a];
        }
        o.voxel = q.voxel;
        npts[i] = np;
      }
    }
    // pass 2: the volumes go to fresh buffers; the context's buffers are replaced only once the
    // new set is on the device (the old ones are freed on leaving this scope, after the wait)
    DevBuf phi[kMaxColliders], vel[kMaxColliders];
    for (int i = 0; i < n; ++i) {
      if (cs.o[i].type != 2) continue;
      const mpm_collider& q = cols[i];
      phi[i].ensure(sizeof(float) * npts[i]);
      HIPCHECK(hipMemcpyAsync(phi[i].p, q.phi, sizeof(float) * npts[i], hipMemcpyHostToDevice, ctx->stream));
      cs.o[i].phi = phi[i].as<float>();
      if (q.phi_velocity) {
        vel[i].ensure(sizeof(float) * 3 * npts[i]);
        HIPCHECK(hipMemcpyAsync(vel[i].p, q.phi_velocity, sizeof(float) * 3 * npts[i], hipMemcpyHostToDevice,
                                ctx->stream));
        cs.o[i].vel = vel[i].as<float>();
      }
    }
    HIPCHECK(hipMemcpyAsync(ctx->cols_dev.p, &cs, sizeof(ColliderSet), hipMemcpyHostToDevice, ctx->stream));
    ctx->sync();
    ctx->hcols = cs;
    for (int i = 0; i < kMaxColliders; ++i) {
      ctx->sdfPhi[i].swap(phi[i]);
      ctx->sdfVel[i].swap(vel[i]);
    }
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

// ------------------------------------------------------------------ grid API (one body)
int mpm_grid_create(mpm_context* ctx, const int32_t* particles, int64_t count, float h, const float fv[3], int dim,
                    mpm_grid** out) {
  if (!out) return MPM_ERR_INVALID;
  *out = nullptr;
  std::unique_ptr<mpm_grid> g(new mpm_grid());
  g->ctx = ctx;
  g->batch.ctx = ctx;
  int rc = guarded(ctx, [&] {
    require_particles(ctx);
    if (dim != 3) throw InvalidArg("only dimension 3 is supported on the GPU");
    if (!(h > 0)) throw InvalidArg("grid size must be positive");
    ctx->slots_from_orig(particles, count, g->slots);
    float z[3] = {0, 0, 0};
    g->count = count;
    g->batch.keepKey = &g->binKey;
    g->batch.build(g->slots.as<int>(), nullptr, (int)count, 1, h, fv ? fv : z, false);
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
  if (!grid || grid->batch.hb.empty()) return MPM_ERR_INVALID;
  for (int a = 0; a < 3; ++a) {
    if (n) n[a] = grid->batch.hb[0].n[a];
    if (mn) mn[a] = grid->batch.hb[0].gmin[a];
  }
  return MPM_OK;
}

int mpm_grid_get(mpm_grid* grid, const char* name, void* out) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    GridBatch& B = grid->batch;
    std::string nm(name ? name : "");
    if (nm == "masses") B.to_dense(0, B.f(B.mass), 1, static_cast<float*>(out));
    else if (nm == "velocities") B.to_dense(0, B.f(B.vel), 3, static_cast<float*>(out));
    else if (nm == "prev_velocities") B.to_dense(0, B.f(B.prev), 3, static_cast<float*>(out));
    else if (nm == "node_collided") {
      const BodyGrid& G = B.hb[0];
      int64_t nd = (int64_t)G.n[0] * G.n[1] * G.n[2];
      B.dense2.ensure(nd);
      launch_to_dense_i8(B.flags.as<int8_t>(), B.dense2.as<int8_t>(), G, grid->ctx->stream);
      HIPCHECK(hipMemcpyAsync(out, B.dense2.p, nd, hipMemcpyDeviceToHost, grid->ctx->stream));
      grid->ctx->sync();
    } else if (nm == "base_cells") {
      // the stencil base node of each particle as the binning kernel computed it
      const int64_t n = grid->count;
      grid->cells.ensure(sizeof(int) * 3 * std::max<int64_t>(n, 1));
      launch_decode_bin(grid->binKey.as<int>(), (int)n, B.bodies.as<BodyGrid>(), B.tileBody.as<int>(),
                        grid->cells.as<int>(), grid->ctx->stream);
      HIPCHECK(hipMemcpyAsync(out, grid->cells.p, sizeof(int) * 3 * n, hipMemcpyDeviceToHost, grid->ctx->stream));
      grid->ctx->sync();
    } else
      throw InvalidArg("unknown grid variable '" + nm + "'");
  });
}

int mpm_grid_set(mpm_grid* grid, const char* name, const void* in) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    GridBatch& B = grid->batch;
    std::string nm(name ? name : "");
    if (nm == "masses") B.from_dense(0, static_cast<const float*>(in), 1, B.f(B.mass));
    else if (nm == "velocities") B.from_dense(0, static_cast<const float*>(in), 3, B.f(B.vel));
    else if (nm == "prev_velocities") B.from_dense(0, static_cast<const float*>(in), 3, B.f(B.prev));
    else if (nm == "node_collided") {
      const BodyGrid& G = B.hb[0];
      int64_t nd = (int64_t)G.n[0] * G.n[1] * G.n[2];
      B.dense2.ensure(nd);
      HIPCHECK(hipMemcpyAsync(B.dense2.p, in, nd, hipMemcpyHostToDevice, grid->ctx->stream));
      launch_from_dense_i8(B.dense2.as<int8_t>(), B.flags.as<int8_t>(), G, grid->ctx->stream);
      launch_node_normals(B.view(), B.flags.as<int8_t>(), grid->ctx->cols(), B.f(B.nrm), grid->ctx->stream);
    } else
      throw InvalidArg("unknown grid variable '" + nm + "'");
    grid->ctx->sync();
  });
}

int mpm_grid_compute_particle_volumes(mpm_grid* grid) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    grid->ctx->clear_err();
    launch_volumes(grid->ctx->view(), grid->batch.view(), grid->batch.f(grid->batch.mass), grid->ctx->errp(),
                   grid->ctx->stream);
    grid->ctx->check_errors();
  });
}

int mpm_grid_calculate_forces(mpm_grid* grid, float* forces) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    GridBatch& B = grid->batch;
    if (!c.model) throw InvalidArg("no constitutive model");
    c.clear_err();
    B.forces_stage();
    launch_node_forces(B.view(), B.f(B.mass), B.f(B.stageA), c.grav(), c.ngrav(), B.f(B.tmp3), c.stream);
    c.check_errors();
    B.to_dense(0, B.f(B.tmp3), 3, forces);
  });
}

int mpm_grid_force_differentials(mpm_grid* grid, const float* dx, float* df) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    GridBatch& B = grid->batch;
    if (!c.model) throw InvalidArg("no constitutive model");
    c.clear_err();
    B.from_dense(0, dx, 3, B.f(B.tmp3));
    B.hvp(B.f(B.tmp3), false, B.f(B.stageA), nullptr, -1);
    launch_gather3(B.view(), B.f(B.stageA), B.f(B.Ar), c.stream);
    c.check_errors();
    B.to_dense(0, B.f(B.Ar), 3, df);
  });
}

int mpm_grid_dforce_dxi(mpm_grid* grid, float* out) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    GridBatch& B = grid->batch;
    if (!c.model) throw InvalidArg("no constitutive model");
    c.clear_err();
    launch_diag(c.view(), B.view(), B.f(B.stageB), c.errp(), c.stream);
    launch_gather3(B.view(), B.f(B.stageB), B.f(B.tmp3), c.stream);
    c.check_errors();
    B.to_dense(0, B.f(B.tmp3), 3, out);
  });
}

int mpm_grid_explicit_momenta(mpm_grid* grid, float dt, float* momenta, int8_t* collided) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    GridBatch& B = grid->batch;
    if (!c.model) throw InvalidArg("no constitutive model");
    c.clear_err();
    HIPCHECK(hipMemsetAsync(B.scal.p, 0, sizeof(CRScalars) * B.nb, c.stream));
    HIPCHECK(hipMemsetAsync(B.status.p, 0, sizeof(BatchStatus), c.stream));
    B.forces_stage();
    // scratch outputs: the grid's own node flags are set only by updateGridVelocities
    launch_node_explicit(B.view(), B.f(B.mass), B.f(B.vel), B.f(B.stageA), c.grav(), c.ngrav(), dt, c.cols(),
                         B.f(B.tmp3), B.flags2.as<int8_t>(), B.f(B.Ar), B.sc(), B.st(), B.partials.as<double>(),
                         B.tileAct.as<unsigned char>(), B.f(B.p), c.stream);  // normals: scratch (p)
    CRScalars hs;
    HIPCHECK(hipMemcpyAsync(&hs, B.scal.p, sizeof(hs), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    c.check_errors();
    raise_errors(hs.errors);
    B.to_dense(0, B.f(B.tmp3), 3, momenta);
    const BodyGrid& G = B.hb[0];
    int64_t nd = (int64_t)G.n[0] * G.n[1] * G.n[2];
    B.dense2.ensure(nd);
    launch_to_dense_i8(B.flags2.as<int8_t>(), B.dense2.as<int8_t>(), G, c.stream);
    HIPCHECK(hipMemcpyAsync(collided, B.dense2.p, nd, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

int mpm_grid_collision_velocities(mpm_grid* grid, const int8_t* collided, float* vc) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    GridBatch& B = grid->batch;
    const BodyGrid& G = B.hb[0];
    int64_t nd = (int64_t)G.n[0] * G.n[1] * G.n[2];
    for (int64_t i = 0; i < nd; ++i)
      if (collided[i] >= c.hcols.n) throw InvalidArg("collision object index out of range");
    HIPCHECK(hipMemsetAsync(B.flags2.p, 0xFF, B.NN, c.stream));
    B.dense2.ensure(nd);
    HIPCHECK(hipMemcpyAsync(B.dense2.p, collided, nd, hipMemcpyHostToDevice, c.stream));
    launch_from_dense_i8(B.dense2.as<int8_t>(), B.flags2.as<int8_t>(), G, c.stream);
    launch_collision_velocities(B.view(), B.flags2.as<int8_t>(), c.cols(), B.f(B.tmp3), c.stream);
    B.to_dense(0, B.f(B.tmp3), 3, vc);
  });
}

int mpm_grid_update_grid_velocities(mpm_grid* grid, float dt, const mpm_solver_params* params,
                                    mpm_solve_stats* stats) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    if (!grid->ctx->model) throw InvalidArg("no constitutive model");
    if (!params) throw InvalidArg("solver parameters required");
    GridBatch& B = grid->batch;
    B.solve(dt, params);
    grid->ctx->prof_resolve();
    B.fetch_history(0, B.rec[0]);
    fill_stats(B.rec[0], stats);
  });
}

// ConjugateResiduals::operator() (ConjugateResiduals.cpp:19-160) over a caller's ProceduralMatrix
// (ProceduralMatrix.h:8-19): the recurrences of the grid solve on n-vectors in device memory (FP32
// vectors, FP64 inner products in a fixed order), the matrix, its preconditioner and the projection
// applied by the caller's callbacks on host or device vectors.
int mpm_conjugate_residuals(mpm_context* ctx, const mpm_procedural_matrix* A, const mpm_procedural_matrix* precond,
                            const float* b, float* x, int64_t n, const mpm_solver_params* params,
                            mpm_solve_stats* stats, float* residuals, float* directions, int log_capacity,
                            int* logged) {
  return guarded(ctx, [&] {
    if (!A || !A->mult) throw InvalidArg("a matrix with a multVector callback is required");
    if (precond && (!precond->mult || !precond->mult_inverse))
      throw InvalidArg("a preconditioner needs multVector and multInverseVector callbacks");
    if (!b || !x || n <= 0) throw InvalidArg("b, x and n > 0 are required");
    if (!params) throw InvalidArg("solver parameters required");
    if (log_capacity < 0 || (log_capacity > 0 && (!residuals || !directions)))
      throw InvalidArg("logs need residual and direction arrays of log_capacity x n floats");
    mpm_context& c = *ctx;
    hipStream_t s = c.stream;
    const size_t bytes = sizeof(float) * (size_t)n;
    DevBuf vb, vx, vr, vrp, vp, vAp, vAr, vpAp, vtmp, red;
    for (DevBuf* d : {&vb, &vx, &vr, &vrp, &vp, &vAp, &vAr, &vpAp, &vtmp}) d->ensure(bytes);
    red.ensure(sizeof(double) * 520);
    std::vector<float> host(n), host2(n);
    HIPCHECK(hipMemcpyAsync(vb.p, b, bytes, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(vx.p, x, bytes, hipMemcpyHostToDevice, s));
    float* db = vb.as<float>();
    float* dx = vx.as<float>();
    float* r = vr.as<float>();
    float* rp = vrp.as<float>();
    float* p = vp.as<float>();
    float* Ap = vAp.as<float>();
    float* Ar = vAr.as<float>();
    float* pAp = vpAp.as<float>();
    // one callback: in -> out (out may equal in only for the projection)
    auto call = [&](const mpm_procedural_matrix* M, int which, const float* in, float* out) {
      c.sync();
      if (M->device_pointers) {
        if (which == 0) M->mult(M->user, in, out, n);
        else if (which == 1) M->mult_inverse(M->user, in, out, n);
        else if (M->project) M->project(M->user, out, n);
        return;
      }
      if (which == 2 && !M->project) return;
      // stream-ordered copies, each waited for: the next kernel on the stream reads `out`
      HIPCHECK(hipMemcpyAsync(host.data(), which == 2 ? out : in, bytes, hipMemcpyDeviceToHost, s));
      c.sync();
      float* res = host2.data();
      if (which == 2) {
        M->project(M->user, host.data(), n);
        res = host.data();
      } else {
        (which == 0 ? M->mult : M->mult_inverse)(M->user, host.data(), host2.data(), n);
      }
      HIPCHECK(hipMemcpyAsync(out, res, bytes, hipMemcpyHostToDevice, s));
      c.sync();
    };
    auto dot = [&](const float* u, const float* v) {
      launch_vec_dot(u, v, n, red.as<double>(), red.as<double>() + 512, s);
      double h = 0.0;
      HIPCHECK(hipMemcpyAsync(&h, red.as<double>() + 512, sizeof(double), hipMemcpyDeviceToHost, s));
      c.sync();
      return h;
    };
    auto to_host = [&](const float* v, float* dst) {
      HIPCHECK(hipMemcpyAsync(dst, v, bytes, hipMemcpyDeviceToHost, s));
      c.sync();
    };
    int nlog = 0;
    auto log = [&]() {
      if (nlog >= log_capacity) return;
      to_host(r, residuals + (int64_t)nlog * n);
      to_host(p, directions + (int64_t)nlog * n);
      ++nlog;
    };
    std::vector<double> hist;
    int iterations = 0, early = 0;
    const double bNorm2 = dot(db, db);
    double rn = 0.0;
    if (bNorm2 == 0.0) {  // ConjugateResiduals.cpp:34-38
      launch_vec_lincomb(dx, 0.0f, db, 0.0f, nullptr, n, s);
    } else {
      const double threshold = (double)params->tolerance * (double)params->tolerance * bNorm2;
      const float maxIters = (float)params->max_iters;
      call(A, 0, dx, Ap);  // Ax, held in Ap until p exists
      launch_vec_lincomb(r, 1.0f, db, -1.0f, Ap, n, s);
      if (precond) call(precond, 1, r, rp);
      else HIPCHECK(hipMemcpyAsync(rp, r, bytes, hipMemcpyDeviceToDevice, s));
      HIPCHECK(hipMemcpyAsync(p, rp, bytes, hipMemcpyDeviceToDevice, s));
      call(A, 0, p, Ap);
      call(A, 0, rp, Ar);
      double rz = dot(rp, Ar);
      log();
      for (int i = 0;; ++i) {
        if (precond) call(precond, 1, Ap, pAp);
        else HIPCHECK(hipMemcpyAsync(pAp, Ap, bytes, hipMemcpyDeviceToDevice, s));
        const double den = dot(Ap, pAp);
        if (den == 0.0) {
          early = 1;
          break;
        }
        const double alpha = rz / den;
        launch_vec_lincomb(dx, 1.0f, dx, (float)alpha, p, n, s);
        call(A, 2, dx, dx);
        iterations = i + 1;
        if (params->debug) {
          to_host(dx, host.data());
          params->debug(params->user, host.data(), n);
        }
        launch_vec_lincomb(rp, 1.0f, rp, (float)-alpha, pAp, n, s);
        if (precond) call(precond, 0, rp, r);
        else HIPCHECK(hipMemcpyAsync(r, rp, bytes, hipMemcpyDeviceToDevice, s));
        bool stop;
        if (params->termination) {
          to_host(r, host.data());
          stop = params->termination(params->user, host.data(), n, i) != 0;
        } else if ((float)i >= maxIters) {
          stop = true;
        } else {
          rn = dot(r, r);
          hist.push_back(rn);
          stop = rn < threshold;
        }
        if (stop) break;
        call(A, 0, rp, Ar);
        const double rzOld = rz;
        rz = dot(rp, Ar);
        if (rzOld == 0.0) {
          early = 1;
          break;
        }
        const double beta = rz / rzOld;
        launch_vec_lincomb(p, 1.0f, rp, (float)beta, p, n, s);
        launch_vec_lincomb(Ap, 1.0f, Ar, (float)beta, Ap, n, s);
        log();
      }
    }
    to_host(dx, x);
    if (logged) *logged = nlog;
    if (stats) {
      stats->iterations = iterations;
      stats->early_exit = early;
      stats->rhs_norm2 = bNorm2;
      stats->residual_norm2 = rn;
      stats->particles = 0;
      stats->active_nodes = 0;
      stats->fallbacks = 0;
      stats->restarts = 0;
      const int L = (int)hist.size();
      stats->history_length = stats->history ? std::min(L, stats->history_capacity) : 0;
      for (int k = 0; k < stats->history_length; ++k) stats->history[k] = hist[k];
    }
  });
}

int mpm_grid_update_particle_velocities(mpm_grid* grid) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    GridBatch& B = grid->batch;
    c.clear_err();
    launch_g2p(c.view(), B.view(), B.f(B.vel), B.f(B.prev), 0, 0.0f, 0.0f, 0.0f, c.cols(), c.errp(), c.stream);
    c.check_errors();
  });
}

int mpm_grid_update_deformation_gradients(mpm_grid* grid, float dt) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    GridBatch& B = grid->batch;
    c.clear_err();
    launch_g2p(c.view(), B.view(), B.f(B.vel), B.f(B.prev), 1, dt, 0.0f, 0.0f, c.cols(), c.errp(), c.stream);
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
    // Sim::Sim: volumes of the initial bodies (Sim.cpp:44-49), frames at rest
    if (ctx->nbodies > 0) {
      GridBatch& B = ctx->simBatch;
      std::vector<float> zero(3 * ctx->nbodies, 0.0f);
      B.build(nullptr, ctx->labels.as<int>(), (int)ctx->n, ctx->nbodies, h, zero.data(), false);
      launch_volumes(ctx->view(), B.view(), B.f(B.mass), ctx->errp(), ctx->stream);
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
    const int nb = c.nbodies;
    const int* labels = c.labels.as<int>();
    bool cancelled = false;
    c.clear_err();
    GridBatch& B = c.simBatch;
    if (nb > 0) {
      // updateParticleData after the first body runs over every particle (Sim.cpp:118): the
      // particles outside body 0 (other bodies and ballistic ones) are brought up to date from
      // their unchanged F before their own grids are built, which is where the reference's
      // sequential loop would have done it
      if (nb > 1 || c.nballistic > 0)
        launch_plasticity(c.view(), labels, 0, n, c.thc, c.ths, c.errp(), s);
      // with a cancellation hook the bodies' chunks stay in body order: a cancelled step keeps the
      // updates of a prefix of the bodies
      B.build(nullptr, labels, n, nb, c.h, nullptr, true, params->cancelled != nullptr);
      c.check_errors();
      B.solve(dt, params);
      // the reference polls cancellation after each body's solve and keeps the particle
      // updates of the bodies before it (Sim.cpp:104-108); chunks are body-ordered, so those
      // bodies are a prefix of the chunk list
      int ncomplete = nb;
      if (cancelHook)
        for (int bi = 0; bi < nb; ++bi)
          if (params->cancelled(params->user)) {
            ncomplete = bi;
            cancelled = true;
            break;
          }
      Batch view = B.view();
      if (cancelled) {
        int prefix = 0;
        if (ncomplete > 0) {
          HIPCHECK(hipMemcpyAsync(&prefix, B.chunkOff.as<int>() + B.hb[ncomplete].tileBase, sizeof(int),
                                  hipMemcpyDeviceToHost, s));
          c.sync();
        }
        view.nchunks = prefix;
      }
      int pg = c.prof_begin(PK_G2P, &B);
      launch_g2p(c.view(), view, B.f(B.vel), B.f(B.prev), 2, dt, c.thc, c.ths, c.cols(), c.errp(), s);
      // advection of the body particles (Sim.cpp:123-131), not after a cancellation (Sim.cpp:107-111)
      if (!cancelled) launch_advect(c.view(), labels, n, dt, c.cols(), s);
      c.prof_end(pg);
      c.lastSolves = B.rec;
      if (stats)
        for (const SolveRecord& r : B.rec) {
          stats->total_iterations += r.iterations;
          stats->max_iterations = std::max(stats->max_iterations, r.iterations);
        }
    }
    // ballistic particles: velocity (Sim.cpp:64-73) and, unless cancelled, advection
    if (c.nballistic > 0)
      launch_ballistic(c.view(), labels, n, c.grav(), c.ngrav(), dt, c.cols(), cancelled ? 1 : 3, s);
    // errors of the particle update are raised with the body partition's wait below
    if (!cancelled && nb > 0) {
      // keep the store body- and tile-sorted: body particles in batch order, then ballistic ones
      c.newToOld.ensure(sizeof(int) * n);
      HIPCHECK(hipMemcpyAsync(c.newToOld.p, B.perm.p, sizeof(int) * B.np, hipMemcpyDeviceToDevice, s));
      if (B.np + c.nballistic != n) throw std::runtime_error("body partition does not cover every particle");
      if (c.nballistic > 0) {
        launch_is_ballistic(labels, n, c.ccIsRoot.as<int>(), s);
        scan_exclusive(c.ccIsRoot.as<int>(), c.ballPos.as<int>(), n, c.ccTot.as<int>() + 3, c.ccTmp.as<int>(), s);
        launch_ballistic_compact(labels, c.ballPos.as<int>(), n, c.newToOld.as<int>(), B.np, s);
      }
      int nxt = 1 - c.cur;
      launch_permute(c.view(c.cur), c.view(nxt), c.newToOld.as<int>(), n, c.orig_of[c.cur].as<int>(),
                     c.orig_of[nxt].as<int>(), c.slot_of.as<int>(), s);
      c.cur = nxt;
    }
    c.prof_resolve();
    calculate_bodies(c, true);
    if (stats) {
      stats->num_bodies = c.nbodies;
      stats->num_ballistic = c.nballistic;
      stats->cancelled = cancelled ? 1 : 0;
      stats->num_solved = (int)c.lastSolves.size();
    }
  });
}

// ------------------------------------------------------------------ learned preconditioner
int64_t mpm_lpc_num_params(void) { return LpcNet().nparams; }

int mpm_lpc_init(mpm_context* ctx, uint64_t seed) {
  return guarded(ctx, [&] { ctx->net.init(seed, ctx->stream); });
}

int mpm_lpc_set_params(mpm_context* ctx, const float* params, int64_t n) {
  return guarded(ctx, [&] {
    if (!params || n != ctx->net.nparams) throw InvalidArg("parameter count mismatch");
    ctx->net.set(params, ctx->stream);
  });
}

int mpm_lpc_get_params(mpm_context* ctx, float* params, int64_t n) {
  return guarded(ctx, [&] {
    if (!params || n != ctx->net.nparams) throw InvalidArg("parameter count mismatch");
    ctx->net.get(params, ctx->stream);
  });
}

int mpm_lpc_set_mode(mpm_context* ctx, int mode) {
  return guarded(ctx, [&] {
    if (mode != 0 && mode != 1) throw InvalidArg("mode must be 0 (diagonal CR) or 1 (learned flexible CR)");
    ctx->lpcMode = mode;
  });
}

int mpm_lpc_set_fallback(mpm_context* ctx, double ratio) {
  return guarded(ctx, [&] {
    if (!(ratio > 0.0 && ratio <= 1.0)) throw InvalidArg("fallback ratio must lie in (0, 1]");
    ctx->lpcFallback = ratio;
  });
}

int mpm_lpc_set_directions(mpm_context* ctx, int m) {
  return guarded(ctx, [&] {
    if (m < 1 || m > kFcrMax) throw InvalidArg("the flexible CR keeps 1 to 8 directions");
    ctx->lpcDirs = m;
  });
}

int mpm_lpc_set_handoff(mpm_context* ctx, int on) {
  return guarded(ctx, [&] {
    if (on != 0 && on != 1) throw InvalidArg("handoff must be 0 or 1");
    ctx->lpcHandoff = on;
  });
}

int mpm_lpc_set_strikes(mpm_context* ctx, int strikes) {
  return guarded(ctx, [&] {
    if (strikes < 1 || strikes > 16) throw InvalidArg("strikes: 1 to 16 fallbacks");
    ctx->lpcStrikes = strikes;
  });
}

int mpm_lpc_set_training_loss(mpm_context* ctx, int mode) {
  return guarded(ctx, [&] {
    if (mode != 0 && mode != 1) throw InvalidArg("training loss: 0 (optimal-step residual) or 1 (exact correction)");
    ctx->cap.lossMode = mode;
  });
}

int mpm_lpc_set_capture_iterations(mpm_context* ctx, const int32_t* iters, int n) {
  return guarded(ctx, [&] {
    if (!iters || n < 1 || n > 16) throw InvalidArg("1 to 16 capture iterations");
    for (int k = 0; k < n; ++k)
      if (iters[k] < 0 || (k > 0 && iters[k] <= iters[k - 1]))
        throw InvalidArg("capture iterations must be non-negative and increasing");
    ctx->cap.schedule.assign(iters, iters + n);
  });
}

int mpm_lpc_train_next(mpm_context* ctx, int samples, int epochs, float learning_rate) {
  return guarded(ctx, [&] {
    if (samples < 0 || samples > 16) throw InvalidArg("at most 16 samples per solve");
    if (epochs < 0) throw InvalidArg("epochs must be >= 0");
    ctx->cap.want = samples;
    ctx->cap.epochs = epochs;
    ctx->cap.lr = learning_rate;
    ctx->cap.lossFirst = ctx->cap.lossLast = ctx->cap.lossDiag = 0.0;
  });
}

int mpm_lpc_last_loss(mpm_context* ctx, double* first, double* last, double* diagonal) {
  return guarded(ctx, [&] {
    if (first) *first = ctx->cap.lossFirst;
    if (last) *last = ctx->cap.lossLast;
    if (diagonal) *diagonal = ctx->cap.lossDiag;
  });
}

int mpm_lpc_last_pairs(const mpm_context* ctx) { return ctx ? ctx->cap.pairs : -1; }

namespace {
// dense single-body inputs of the grid API into the batch layout
void grid_lpc_inputs(mpm_grid* grid, const float* r, const float* diag, const float* e, float scale, float weight) {
  GridBatch& B = grid->batch;
  for (DevBuf* d : {&B.fr, &B.fp0, &B.fp1}) d->ensure(B.vecBytes());
  B.from_dense(0, r, 3, B.f(B.fr));
  B.from_dense(0, diag, 3, B.f(B.fp0));
  if (e) B.from_dense(0, e, 3, B.f(B.fp1));
  float hv[2] = {scale, weight};
  B.fq0.ensure(sizeof(float) * 2);
  HIPCHECK(hipMemcpyAsync(B.fq0.p, hv, sizeof(hv), hipMemcpyHostToDevice, grid->ctx->stream));
  B.ensure_levels();
}
}  // namespace

int mpm_grid_lpc_apply(mpm_grid* grid, const float* r, const float* diag, float scale, float* out) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    GridBatch& B = grid->batch;
    grid_lpc_inputs(grid, r, diag, nullptr, scale, 0.0f);
    c.net.alloc_acts(B.levels, true);
    c.net.forward(B.view(), B.levels, B.f(B.fr), B.f(B.fp0), B.f(B.mass), nullptr, B.fq0.as<float>(), nullptr,
                  c.net.raw.as<float>(), nullptr, c.stream);
    B.to_dense(0, c.net.raw.as<float>(), 3, out);
  });
}

int mpm_grid_lpc_gradients(mpm_grid* grid, const float* r, const float* diag, const float* e, float scale,
                           float weight, float* grad, double* loss) {
  if (!grid) return MPM_ERR_INVALID;
  return guarded(grid->ctx, [&] {
    mpm_context& c = *grid->ctx;
    GridBatch& B = grid->batch;
    grid_lpc_inputs(grid, r, diag, e, scale, weight);
    const double l = c.net.gradients(B.view(), B.levels, B.f(B.fr), B.f(B.fp0), B.f(B.mass), nullptr, B.f(B.fp1),
                                     B.fq0.as<float>(), B.fq0.as<float>() + 1, c.stream);
    if (loss) *loss = l;
    if (grad) {
      HIPCHECK(hipMemcpyAsync(grad, c.net.G.p, sizeof(float) * c.net.nparams, hipMemcpyDeviceToHost, c.stream));
      c.sync();
    }
  });
}

int mpm_set_small_bodies(mpm_context* ctx, int max_particles) {
  return guarded(ctx, [&] {
    if (max_particles < 0) throw InvalidArg("max_particles must be >= 0");
    ctx->smallMaxParticles = max_particles;
  });
}

int mpm_profile_enable(mpm_context* ctx, int mask) {
  return guarded(ctx, [&] {
    ctx->prof_resolve();
    ctx->profMask = (unsigned int)mask;
    for (int k = 0; k < PK_COUNT; ++k) {
      ctx->profMs[k] = 0;
      ctx->profCount[k] = 0;
      ctx->profBytes[k] = 0;
      ctx->profSeen[k] = 0;
    }
  });
}

int mpm_profile_period(mpm_context* ctx, int period) {
  return guarded(ctx, [&] {
    if (period < 1) throw InvalidArg("profile period must be >= 1");
    ctx->prof_resolve();
    ctx->profPeriod = period;
    for (int k = 0; k < PK_COUNT; ++k) ctx->profSeen[k] = 0;
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

int mpm_sim_num_bodies(const mpm_context* ctx) { return ctx ? ctx->nbodies : 0; }

int mpm_sim_grid_layout(const mpm_context* ctx, int64_t* out) {
  if (!ctx || !out) return MPM_ERR_INVALID;
  const GridBatch& B = ctx->simBatch;
  out[0] = B.nb;
  out[1] = B.NT;
  out[2] = B.nchunks;
  out[3] = B.NN;
  out[4] = B.np;
  out[5] = B.prefixTiles;
  out[6] = B.prefixChunks;
  return MPM_OK;
}

int64_t mpm_sim_body(const mpm_context* ctx, int b, int32_t* out) {
  if (!ctx || b < 0 || b >= ctx->nbodies) return -1;
  try {
    ensure_body_lists(*ctx);
  } catch (...) {
    return -1;
  }
  const auto& B = ctx->bodyLists[b];
  if (out) std::copy(B.begin(), B.end(), out);
  return (int64_t)B.size();
}

int64_t mpm_sim_ballistic(const mpm_context* ctx, int32_t* out) {
  if (!ctx) return -1;
  try {
    ensure_body_lists(*ctx);
  } catch (...) {
    return -1;
  }
  if (out) std::copy(ctx->ballisticList.begin(), ctx->ballisticList.end(), out);
  return (int64_t)ctx->ballisticList.size();
}

int mpm_sim_body_solve_stats(const mpm_context* ctx, int b, mpm_solve_stats* st) {
  if (!ctx || b < 0 || b >= (int)ctx->lastSolves.size()) return MPM_ERR_INVALID;
  mpm_context& c = const_cast<mpm_context&>(*ctx);
  try {
    HIPCHECK(hipSetDevice(c.device));
    SolveRecord& r = c.lastSolves[b];
    if (st && st->history && st->history_capacity > 0 && r.history.empty() && r.iterations > 0)
      c.simBatch.fetch_history(b, r);
  } catch (const std::exception& e) {
    c.err = e.what();
    return MPM_ERR_HIP;
  }
  fill_stats(ctx->lastSolves[b], st);
  return MPM_OK;
}

}  // extern "C"
