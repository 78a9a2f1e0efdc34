// mpm_device.hpp -- data layout shared by the HIP kernels and the host runtime.
//
// Particle store: structure-of-arrays, one float column per scalar component
// (column c of slot s at base[c * stride + s]); matrices are 9 columns in
// column-major element order. Slots are a permutation of the caller's
// particle indices that the runtime keeps body- and tile-sorted.
//
// Grids: every body of a step gets the reference's dense grid over its own
// bounding box (n0 x n1 x n2 nodes, reference node index i + n0 (j + n1 k)),
// and all bodies' grids live in one batched node space so that each kernel
// launch serves every body. A body's grid is cut into 4x4x4-node tiles
// (tile t = tx + T0 (ty + T1 tz) owns 64 consecutive nodes, local
// l = (i&3) + 4 (j&3) + 16 (k&3)); its tiles occupy the global tile range
// [tileBase, tileBase + NT) with NT padded to a multiple of 4, so one
// 256-thread node block (4 tiles) never spans two bodies. 3-vectors store each
// tile as three 64-float component runs ((t*3 + c) * 64 + l).
//
// Particle-side kernels run one workgroup per chunk (up to 256 particles whose
// stencil base node lies in one tile). A particle in tile t only touches nodes
// in tiles t + {0,1}^3, so the workgroup stages that 8-tile (512-node)
// footprint in LDS, accumulates its scatter there in fixed point with integer
// LDS atomics (float LDS atomics run ~45x slower per wave-op), and writes the
// footprint to a per-chunk staging slot with plain coalesced stores. Node
// kernels gather, for each node, the staged footprints of the chunks of the
// (up to 8) tiles whose footprint covers it. No device-scope float atomics are
// used: on MI355X those resolve beyond the per-XCD L2s and were the bottleneck.
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
  __device__ __forceinline__ float& at(int c, int s) const { return base[(int64_t)c * stride + s]; }
  // at() for a column c that is the same in every lane of the wave: the column base is held in
  // scalar registers and the slot is a 32-bit byte offset (slots stay below 2^30, which
  // mpm_particles_create checks), the form the loads take as their vector operand
  __device__ __forceinline__ float& atu(int c, int s) const {
    const uint64_t o = (uint64_t)((int64_t)c * stride) * 4u;
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)o);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(o >> 32));
    char* col = reinterpret_cast<char*>(base) + ((hi << 32) | lo);
    return *reinterpret_cast<float*>(col + (uint32_t)s * 4u);
  }
};

constexpr int kFoot = 512;        // nodes in a tile's 2x2x2-tile footprint
constexpr int kStage = 3 * kFoot; // floats per staged chunk footprint
constexpr int kChunk = 256;       // max particles per workgroup chunk
constexpr int kBlock = 256;       // threads per workgroup; node blocks cover 4 tiles
constexpr int kHistCap = 4096;     // residual history entries kept per body
constexpr int kKeyCodeShift = 24;  // binning keys: global tile in the low bits, base-node code above

// Grid geometry of one body (Grid.cpp:193-213) and its place in the batch.
struct BodyGrid {
  int n[3];        // reference node counts
  int T[3];        // tiles per axis
  int NT;          // tiles, padded to a multiple of 4
  int tileBase;    // first global tile
  int blockBase;   // first global node block (= tileBase / 4)
  int nblocks;     // node blocks (= NT / 4)
  int np;          // particles
  float gmin[3];   // m_min
  float h;
  float fv[3];     // comoving frame velocity
};

struct Batch {
  const BodyGrid* bodies;
  int nbodies;
  const int* blockBody;     // body of each node block
  const int* tileBody;      // body of each global tile
  const int* perm;          // tile-sorted particle slots of all bodies
  const int4* chunks;       // (global tile, start in perm, count, body)
  int nchunks;
  const int* tileChunk0;    // first chunk of each global tile
  const int* tileChunkN;    // number of chunks of each global tile
  int NT;                   // global tiles
  int nblocks;              // global node blocks
  unsigned int* tickets;    // reduction tickets: [body][set][1 + kSub][kTicketStride]
  const unsigned char* tileActive;  // per tile: some node has mass or a collider flag
  const float* nrm;         // per node with a collider flag >= 0: the collider's unit normal there
};

// Per-body reductions inside node kernels: the body's blocks form kSub sub-groups (local
// block index mod kSub). Every block stores its partial sums and counts in on its sub-group's
// ticket; the last block of a sub-group adds the sub-group's partials in member order, stores
// that total and counts in on the body's main ticket; the last of those adds the sub-group
// totals in sub-group order. The last block's wait is then two short reads instead of one over
// every block of a large body. Tickets sit on separate 128-byte lines; spreading the arrivals
// keeps same-address atomics short. Per body and set: kSetTickets unsigned ints of tickets,
// then kSub sub-group totals of kPartialStride doubles.
constexpr int kSub = 32;
constexpr int kTicketStride = 32;
constexpr int kTicketSets = 3;
constexpr int kPartialStride = 16;  // doubles per block in the partials buffer
constexpr int kSetTickets = (kSub + 1) * kTicketStride;
constexpr int kSetStride = kSetTickets + kSub * kPartialStride * 2;  // unsigned ints
constexpr int kTicketsPerBody = kTicketSets * kSetStride;

// error bits raised by kernels (reference exception messages are mapped by the host)
enum ErrorBits : int {
  ERR_NAN_MASS = 1,        // "nans in splatted masses!"
  ERR_NAN_VEL = 2,         // "nans in splatted velocities!"
  ERR_NAN_STRESS = 4,      // "nans in dEnergyDensitydF matrix!"
  ERR_NAN_MOMENTA = 8,     // "nan in explicit momenta!"
  ERR_HARDNESS = 16,       // "infinite hardness!!!"
  ERR_STENCIL = 32,        // internal: particle outside its tile's footprint
  ERR_NAN_PVEL = 64,       // "nans in particle velocity data!"
  ERR_INTERNAL = 128,      // internal: a launch that does not fit the batch's layout
  ERR_INDEX = 256          // a caller particle index out of range (device-resident index lists)
};

constexpr int kFcrMax = 8;  // flexible CR: most stored directions (the context's setting, default 2)

// Conjugate-residual scalars of one body, resident in device memory. Reductions are
// FP64 and deterministic (per-block partials combined in a fixed order by the last
// block of the body to finish).
struct CRScalars {
  double bnorm2, threshold;
  double rz, ApdPAp, alpha, beta, rnorm2;
  double sums[8];          // totalMass, totalMomentum[3], collidedMass, collidedMomentum[3]
  float guess[3];
  int done;                // 1 = solve finished; 2 = parked; kDoneSmall = small-body solver
  int zero_x;              // rhs was zero: x = 0 (ConjugateResiduals.cpp:34-38)
  int iter;                // current iteration index i
  int iterations;          // x updates performed
  int early_exit;          // zero denominator
  int restarts;            // CR recurrence restarts (beta = 0) after a cancelled denominator expansion
  int max_iters;
  int errors;              // ErrorBits
  int active_nodes;        // nodes with m > 0 (measurement)
  // flexible conjugate residual with the learned preconditioner (truncated GCR over the last
  // fcr_m directions, fcr_m <= kFcrMax)
  double gbeta[kFcrMax];   // <A z, A p_j> / <A p_j, A p_j>
  double ApAp[kFcrMax];    // <A p_j, A p_j> of the stored directions
  double rAp;              // <r, A p> of the new direction
  double rr_prev;          // |r|^2 before the last update
  double lpc_ratio;        // a network iteration whose |r|^2 exceeds lpc_ratio * rr_prev is a fallback
  double diag_rate;        // |r|^2 / rr_prev of the body's last diagonal iteration (lpc_ratio before one)
  double lpc_s2;           // sum over active nodes of (r / diag)^2, and their count (3 per node)
  double lpc_cnt;
  float lpc_scale;         // RMS of r / diag over active nodes: network input scale
  int lpc_off;             // 1: this body uses the diagonal preconditioner
  int fallbacks;           // iterations that switched this body to the diagonal preconditioner
  int ndirs;               // directions stored (<= fcr_m)
  int fcr_m;               // directions kept
  int slot;                // ring slot of the most recent direction
  int retry;               // 1: the network gave no direction; redo the iteration on the diagonal
  int lpc_pause;           // 1: lpc_off holds for one iteration only (a fallback before the last)
  int lpc_strikes;         // fallbacks after which the body keeps the diagonal preconditioner
};
// default fallbacks after which a body keeps the diagonal preconditioner (4M slab, learned leg:
// 1 strike 0.84x of plain CR, 2 strikes 0.81x, 3 strikes 0.77x)
constexpr int kLpcStrikes = 1;
constexpr float kLpcClamp = 8.0f;  // network inputs, outputs and training targets, in units of the scale

// Counters shared by all bodies of a batch.
struct BatchStatus {
  int bodies_done;
  int errors;
  int small_bodies;  // bodies handed to the one-workgroup solver (done = kDoneSmall)
};
// Small bodies: at most kSmallBlocks node blocks; their CR solve runs in one workgroup.
constexpr int kSmallBlocks = 8;
constexpr int kDoneSmall = 3;  // CRScalars::done of a body that the small-body solver owns

}  // namespace mpm
