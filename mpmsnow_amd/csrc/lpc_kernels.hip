// lpc_kernels.hip -- gfx950 kernels of the learned preconditioner (layout: lpc_kernels.hpp).
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16. For one 4x4x4 output tile
// (M = 64 voxels) the K dimension runs over (offset, input channel) pairs, 27 CIN of them,
// in chunks of 32; a lane's 8 consecutive K values are 8 consecutive channels of one offset
// (CIN is a multiple of 16), so each A fragment is one 16-byte LDS read of the staged 6x6x6
// input halo and each B fragment one 16-byte read of the weights, stored [COUT][27 CIN].
// Operand layout of the instruction (checked on the device): lane l holds A[l % 16][8 (l / 16)
// + v] and B[8 (l / 16) + v][l % 16] for v < 8, and D[4 (l / 16) + v][l % 16] for v < 4.
#include "lpc_kernels.hpp"

namespace mpm {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;

inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
inline int grid1d(int64_t n, int cap = 16384) {
  int b = cdiv(n, kThreads);
  return b < 1 ? 1 : (b > cap ? cap : b);
}
#define LPC_STRIDE(i, n) \
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < (n); i += (int64_t)gridDim.x * kThreads)

__device__ __forceinline__ void lv_coords(const LevelGrid& G, int lt, int& tx, int& ty, int& tz) {
  tx = lt % G.T[0];
  int r = lt / G.T[0];
  ty = r % G.T[1];
  tz = r / G.T[1];
}
__device__ __forceinline__ int lv_tile(const LevelGrid& G, int tx, int ty, int tz) {
  if (tx < 0 || ty < 0 || tz < 0 || tx >= G.T[0] || ty >= G.T[1] || tz >= G.T[2]) return -1;
  return G.tileBase + tx + G.T[0] * (ty + G.T[1] * tz);
}
__device__ __forceinline__ bool lv_real(const LevelGrid& G, int lt) { return lt < G.T[0] * G.T[1] * G.T[2]; }

// Stage the 6x6x6-voxel input halo of output tile (tx, ty, tz) (zeros outside the body).
// Every load of a thread is issued before its first LDS store, so the thread waits for global
// memory once rather than once per 16-byte piece.
template <int CIN>
__device__ __forceinline__ void stage_halo(const bf16* in, const int* nb, bf16* halo) {
  constexpr int CG = CIN / 8;
  constexpr int TOT = 216 * CG;
  constexpr int IT = (TOT + kThreads - 1) / kThreads;
  uint4 v[IT];
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const int q = (int)threadIdx.x + k * kThreads;
    v[k] = make_uint4(0u, 0u, 0u, 0u);
    if (q < TOT) {
      const int hv = q / CG, cg = q - (q / CG) * CG;
      const int hx = hv % 6, hy = (hv / 6) % 6, hz = hv / 36;
      const int o = (hx == 0 ? 0 : (hx == 5 ? 2 : 1)) + 3 * (hy == 0 ? 0 : (hy == 5 ? 2 : 1)) +
                    9 * (hz == 0 ? 0 : (hz == 5 ? 2 : 1));
      const int lv = ((hx + 3) & 3) + 4 * ((hy + 3) & 3) + 16 * ((hz + 3) & 3);
      const int t = nb[o];
      if (t >= 0) v[k] = *reinterpret_cast<const uint4*>(in + ((int64_t)t * 64 + lv) * CIN + cg * 8);
    }
  }
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const int q = (int)threadIdx.x + k * kThreads;
    if (q < TOT) *reinterpret_cast<uint4*>(halo + q * 8) = v[k];
  }
}
// neighbour tiles of an output tile, index (dx + 1) + 3 (dy + 1) + 9 (dz + 1)
__device__ __forceinline__ void neighbour_tiles(const LevelGrid& G, bool real, int tx, int ty, int tz, int* nb) {
  if (threadIdx.x < 27) {
    const int o = threadIdx.x;
    nb[o] = real ? lv_tile(G, tx + o % 3 - 1, ty + (o / 3) % 3 - 1, tz + o / 9 - 1) : -1;
  }
}

// ------------------------------------------------------------ convolution (forward / data gradient)
// WAVES waves per output tile (1 or 4): the tile's 64 voxels are four 16-row M blocks, 4 / WAVES
// of them per wave, and each B fragment (weights) is loaded once per K chunk for a wave's blocks.
// With one wave per tile a workgroup holds only the tile's halo in LDS, so a CU keeps several
// times more tiles in flight than with four: the convolutions wait on the tile, halo and weight
// loads, not on the MFMAs. Levels of few tiles (the coarsest) keep four waves per tile so that
// the device still has work for every SIMD.
constexpr int kConvFewTiles = 1024;
// A slice's halo is 216 voxels x 16 channels = 432 pieces of 16 bytes, loaded in one round of
// ceil(432 / NTHR) pieces per thread into registers and then stored to LDS, so that the loads
// of slice s + 1 are in flight while the MFMAs of slice s run.
template <int NTHR>
constexpr int halo_pieces() { return (432 + NTHR - 1) / NTHR; }
template <int CIN, int NTHR>
__device__ __forceinline__ void load_halo_slice(const bf16* in, const int* nb, int sl,
                                                uint4 (&v)[halo_pieces<NTHR>()]) {
#pragma unroll
  for (int k = 0; k < halo_pieces<NTHR>(); ++k) {
    const int q = (int)threadIdx.x + k * NTHR;
    v[k] = make_uint4(0u, 0u, 0u, 0u);
    if (q < 432) {
      const int hv = q >> 1, cg = q & 1;
      const int hx = hv % 6, hy = (hv / 6) % 6, hz = hv / 36;
      const int o = (hx == 0 ? 0 : (hx == 5 ? 2 : 1)) + 3 * (hy == 0 ? 0 : (hy == 5 ? 2 : 1)) +
                    9 * (hz == 0 ? 0 : (hz == 5 ? 2 : 1));
      const int lv = ((hx + 3) & 3) + 4 * ((hy + 3) & 3) + 16 * ((hz + 3) & 3);
      const int t = nb[o];
      if (t >= 0) v[k] = *reinterpret_cast<const uint4*>(in + ((int64_t)t * 64 + lv) * CIN + 16 * sl + cg * 8);
    }
  }
}
// channels [16 sl, 16 sl + 16) of the halo, 16 per voxel in LDS
template <int NTHR>
__device__ __forceinline__ void store_halo_slice(const uint4 (&v)[halo_pieces<NTHR>()], bf16* halo) {
#pragma unroll
  for (int k = 0; k < halo_pieces<NTHR>(); ++k) {
    const int q = (int)threadIdx.x + k * NTHR;
    if (q < 432) *reinterpret_cast<uint4*>(halo + q * 8) = v[k];
  }
}
template <int CIN, int COUT, int MODE, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_conv(LevelView V, const bf16* __restrict__ in,
                                                    const bf16* __restrict__ W, const float* __restrict__ bias,
                                                    ConvOut o) {
  static_assert(CIN % 16 == 0 && COUT % 16 == 0, "channel counts");
  static_assert(WAVES == 1 || WAVES == 4, "one or four waves per tile");
  constexpr int KTOT = 27 * CIN;
  constexpr int NT16 = COUT / 16;
  constexpr int MB = 4 / WAVES;  // M blocks per wave
  constexpr int NSL = CIN / 16;  // 16-channel slices of the input, staged one at a time
  constexpr int KS = 27 * 16;    // K of one slice: (offset, channel of the slice)
  constexpr int NCHS = (KS + 31) / 32;
  __shared__ __align__(16) bf16 halo[216 * 16];
  __shared__ int nb[27];
  const int tile = blockIdx.x;
  const int b = V.tileBody[tile];
  if (o.sc && (o.sc[b].done || o.sc[b].lpc_off)) return;
  const LevelGrid G = V.bodies[b];
  const int lt = tile - G.tileBase;
  const bool real = lv_real(G, lt);
  int tx = 0, ty = 0, tz = 0;
  if (real) lv_coords(G, lt, tx, ty, tz);
  neighbour_tiles(G, real, tx, ty, tz, nb);
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row = lane & 15, q = lane >> 4;
  floatx4 acc[MB][NT16];
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int t = 0; t < NT16; ++t) acc[m][t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  // the epilogue's other inputs (skip activations; residual and diagonal of the output layer),
  // requested before the K loop so that their latency overlaps it
  float pre[MB][NT16][4];
  float pre2[MB][4];
#pragma unroll
  for (int m = 0; m < MB; ++m) {
#pragma unroll
    for (int t = 0; t < NT16; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) pre[m][t][v] = 0.0f;
#pragma unroll
    for (int v = 0; v < 4; ++v) pre2[m][v] = 1.0f;
  }
  // the output layer's body scale, requested with the other epilogue inputs
  const float lscale = (MODE == CONV_FINAL && o.z) ? o.sc[b].lpc_scale : 0.0f;
  if (MODE == CONV_FINAL) {
    if (row < 3 && o.z) {
#pragma unroll
      for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t zi = ((int64_t)tile * 3 + row) * 64 + 16 * (w * MB + m) + 4 * q + v;
          pre[m][0][v] = o.r[zi];
          pre2[m][v] = o.diag[zi];
        }
    }
  } else if (o.out2) {
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int t = 0; t < NT16; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          pre[m][t][v] = (float)o.skip[((int64_t)tile * 64 + 16 * (w * MB + m) + 4 * q + v) * COUT + 16 * t + row];
  }
  // K runs slice by slice: a slice's halo (16 channels, 6.9 KB) is the workgroup's only LDS, so
  // that a CU holds as many tiles as at 16 input channels; the next slice's halo loads overlap
  // this slice's MFMAs
  uint4 pieces[halo_pieces<64 * WAVES>()];
  load_halo_slice<CIN, 64 * WAVES>(in, nb, 0, pieces);
  for (int sl = 0; sl < NSL; ++sl) {
    if (sl > 0) __syncthreads();  // the previous slice's reads of halo are done
    store_halo_slice<64 * WAVES>(pieces, halo);
    __syncthreads();
    if (sl + 1 < NSL) load_halo_slice<CIN, 64 * WAVES>(in, nb, sl + 1, pieces);
    // one K chunk of 32 (within the slice: offset f / 16, channel f % 16)
    auto kstep = [&](int c) {
      const int f = 32 * c + 8 * q;
      const bool inK = f < KS;
      const int off = inK ? f >> 4 : 0, cl = f & 15;
      bf16x8 bw[NT16];
#pragma unroll
      for (int t = 0; t < NT16; ++t) {
        if (inK) {
          bw[t] = *reinterpret_cast<const bf16x8*>(W + (int64_t)(16 * t + row) * KTOT + off * CIN + 16 * sl + cl);
        } else {
#pragma unroll
          for (int u = 0; u < 8; ++u) bw[t][u] = (bf16)0.0f;
        }
      }
      const int ox = off % 3, oy = (off / 3) % 3, oz = off / 9;
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int va = 16 * (w * MB + m) + row;  // the A row (output voxel) of this M block
        const int vx = va & 3, vy = (va >> 2) & 3, vz = va >> 4;
        bf16x8 a;
        if (inK) {
          a = *reinterpret_cast<const bf16x8*>(halo + (((vz + oz) * 6 + vy + oy) * 6 + vx + ox) * 16 + cl);
        } else {
#pragma unroll
          for (int u = 0; u < 8; ++u) a[u] = (bf16)0.0f;
        }
#pragma unroll
        for (int t = 0; t < NT16; ++t)
          acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[t], acc[m][t], 0, 0, 0);
      }
    };
    // the single slice of 16-channel inputs is unrolled whole; wider inputs by two chunks, which
    // keeps them at 4 waves per SIMD (measured: full unrolling gives the 16-channel convolutions
    // 15% and costs the 64-channel one 13%)
    if constexpr (NSL == 1) {
#pragma unroll
      for (int c = 0; c < NCHS; ++c) kstep(c);
    } else {
#pragma unroll 2
      for (int c = 0; c < NCHS; ++c) kstep(c);
    }
  }
  // epilogue: lane holds D[4 q + v][row] of each (M block, N tile)
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int t = 0; t < NT16; ++t) {
      const int co = 16 * t + row;
      const float bco = bias ? bias[co] : 0.0f;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int vox = 16 * (w * MB + m) + 4 * q + v;
        const bool valid = real && tx * 4 + (vox & 3) < G.n[0] && ty * 4 + ((vox >> 2) & 3) < G.n[1] &&
                           tz * 4 + (vox >> 4) < G.n[2];
        float val = acc[m][t][v] + bco;
        if (MODE == CONV_RELU) val = val > 0.0f ? val : 0.0f;
        if (!valid) val = 0.0f;
        const int64_t e = ((int64_t)tile * 64 + vox) * COUT + co;
        if (MODE == CONV_FINAL) {
          if (co < 3) {
            const int64_t zi = ((int64_t)tile * 3 + co) * 64 + vox;
            if (o.raw) o.raw[zi] = val;
            if (o.z)
              o.z[zi] = pre[m][0][v] / pre2[m][v] + fminf(fmaxf(val, -kLpcClamp), kLpcClamp) * lscale;
          }
        } else {
          const bf16 hv = (bf16)val;
          o.out[e] = hv;
          if (o.out2) o.out2[e] = (bf16)((float)hv + pre[m][t][v]);
        }
      }
    }
}

// ------------------------------------------------------------ weight gradient
// partial[group][co][f] = sum over the group's tiles and voxels v of g[v][co] in(v + off(f))[ci(f)]
// GEMM: M = COUT (co), N = 27 CIN (f), K = voxels. Block (group, N range of 16 N tiles).
template <int CIN, int COUT>
__global__ __launch_bounds__(kThreads) void k_conv_dw(LevelView V, const bf16* __restrict__ in,
                                                     const bf16* __restrict__ g, float* __restrict__ partial,
                                                     int tilesPerGroup) {
  constexpr int KTOT = 27 * CIN;
  constexpr int NTN = KTOT / 16;  // N tiles in total
  constexpr int MT = COUT / 16;
  constexpr int NR = 16 / MT;     // n tiles per block
  constexpr int PAIRS = MT * NR;  // (m tile, n tile) pairs of this block: 16
  constexpr int PER_WAVE = PAIRS / 4;
  __shared__ __align__(16) bf16 halo[216 * CIN];
  __shared__ __align__(16) bf16 gs[64 * COUT];
  __shared__ int nb[27];
  const int group = blockIdx.x;
  const int nt0 = blockIdx.y * NR;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row = lane & 15, q = lane >> 4;
  floatx4 acc[PER_WAVE];
#pragma unroll
  for (int k = 0; k < PER_WAVE; ++k) acc[k] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  const int t0 = group * tilesPerGroup;
  const int t1 = min(V.NT, t0 + tilesPerGroup);
  for (int tile = t0; tile < t1; ++tile) {
    const int b = V.tileBody[tile];
    const LevelGrid G = V.bodies[b];
    const int lt = tile - G.tileBase;
    const bool real = lv_real(G, lt);
    int tx = 0, ty = 0, tz = 0;
    if (real) lv_coords(G, lt, tx, ty, tz);
    __syncthreads();  // previous tile's LDS reads are done
    neighbour_tiles(G, real, tx, ty, tz, nb);
    for (int i = threadIdx.x; i < 64 * COUT / 8; i += kThreads)
      *reinterpret_cast<uint4*>(gs + i * 8) = *reinterpret_cast<const uint4*>(g + (int64_t)tile * 64 * COUT + i * 8);
    __syncthreads();
    stage_halo<CIN>(in, nb, halo);
    __syncthreads();
    if (!real) continue;
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {  // 64 voxels = 2 chunks of 32
#pragma unroll
      for (int k = 0; k < PER_WAVE; ++k) {
        const int pair = w + 4 * k;
        if (pair >= PAIRS) break;
        const int mt = pair / NR, ntl = pair % NR;
        const int nt = nt0 + ntl;
        if (nt >= NTN) continue;
        bf16x8 a, bb;
        const int co = 16 * mt + row;
        const int f = 16 * nt + row;
        const int off = f / CIN, ci = f - off * CIN;
        const int ox = off % 3, oy = (off / 3) % 3, oz = off / 9;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int vox = 32 * kc + 8 * q + u;
          a[u] = gs[vox * COUT + co];
          const int hx = (vox & 3) + ox, hy = ((vox >> 2) & 3) + oy, hz = (vox >> 4) + oz;
          bb[u] = halo[((hz * 6 + hy) * 6 + hx) * CIN + ci];
        }
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[k], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < PER_WAVE; ++k) {
    const int pair = w + 4 * k;
    if (pair >= PAIRS) break;
    const int mt = pair / NR, ntl = pair % NR;
    const int nt = nt0 + ntl;
    if (nt >= NTN) continue;
    const int f = 16 * nt + row;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int co = 16 * mt + 4 * q + v;
      partial[((int64_t)group * COUT + co) * KTOT + f] = acc[k][v];
    }
  }
}

__global__ void k_reduce_groups(const float* partial, int groups, int64_t len, float* out) {
  LPC_STRIDE(i, len) {
    float s = 0.0f;
    for (int gi = 0; gi < groups; ++gi) s += partial[(int64_t)gi * len + i];
    out[i] = s;
  }
}

// bias gradient partials: one block per group, thread c sums channel c
__global__ void k_bias_partial(int C, LevelView V, const bf16* g, float* partial, int tilesPerGroup) {
  const int group = blockIdx.x;
  const int t0 = group * tilesPerGroup;
  const int t1 = min(V.NT, t0 + tilesPerGroup);
  __shared__ float red[kThreads];
  // thread layout: channel c = threadIdx.x % C, voxel stripe = threadIdx.x / C
  const int c = threadIdx.x % C, stripe = threadIdx.x / C, nstripe = kThreads / C;
  float s = 0.0f;
  for (int tile = t0; tile < t1; ++tile)
    for (int v = stripe; v < 64; v += nstripe) s += (float)g[((int64_t)tile * 64 + v) * C + c];
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < C) {
    float t = 0.0f;
    for (int k = 0; k < nstripe; ++k) t += red[k * C + threadIdx.x];
    partial[(int64_t)group * C + threadIdx.x] = t;
  }
}

// ------------------------------------------------------------ pooling
template <int C>
__global__ void k_pool(LevelView F, LevelView Cc, const bf16* in, bf16* out, float scale) {
  constexpr int CG = C / 8;
  const int tile = blockIdx.x;
  const int b = Cc.tileBody[tile];
  const LevelGrid G = Cc.bodies[b];
  const LevelGrid Gf = F.bodies[b];
  const int lt = tile - G.tileBase;
  int tx, ty, tz;
  lv_coords(G, lt, tx, ty, tz);
  for (int idx = threadIdx.x; idx < 64 * CG; idx += kThreads) {
    const int v = idx / CG, cg = idx - (idx / CG) * CG;
    const int X = tx * 4 + (v & 3), Y = ty * 4 + ((v >> 2) & 3), Z = tz * 4 + (v >> 4);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const bool valid = X < G.n[0] && Y < G.n[1] && Z < G.n[2];
    if (valid) {
      for (int c8 = 0; c8 < 8; ++c8) {
        const int fx = 2 * X + (c8 & 1), fy = 2 * Y + ((c8 >> 1) & 1), fz = 2 * Z + (c8 >> 2);
        const int ft = lv_tile(Gf, fx >> 2, fy >> 2, fz >> 2);
        if (ft < 0) continue;
        const int fl = (fx & 3) + 4 * (fy & 3) + 16 * (fz & 3);
        bf16x8 x = *reinterpret_cast<const bf16x8*>(in + ((int64_t)ft * 64 + fl) * C + cg * 8);
#pragma unroll
        for (int u = 0; u < 8; ++u) s[u] += (float)x[u];
      }
    }
    bf16x8 r;
#pragma unroll
    for (int u = 0; u < 8; ++u) r[u] = (bf16)(s[u] * scale);
    *reinterpret_cast<bf16x8*>(out + ((int64_t)tile * 64 + v) * C + cg * 8) = r;
  }
}
// Unpooling runs one wave per tile (each thread copies a few parent values): a CU then holds four
// times the tiles in flight of a four-wave workgroup, which is what hides the chain of table and
// parent loads (26.8 -> 21.2 us at 4M, 32 channels). Pooling, whose threads each sum eight
// children, stays at four waves (at one wave it took twice as long).
constexpr int kPoolThreads = 64;
template <int C>
__global__ __launch_bounds__(kPoolThreads) void k_unpool(LevelView F, LevelView Cc, const bf16* in, bf16* out,
                                                         float scale, int accumulate) {
  constexpr int CG = C / 8;
  const int tile = blockIdx.x;
  const int b = F.tileBody[tile];
  const LevelGrid Gf = F.bodies[b];
  const LevelGrid G = Cc.bodies[b];
  const int lt = tile - Gf.tileBase;
  const bool real = lv_real(Gf, lt);
  int tx = 0, ty = 0, tz = 0;
  if (real) lv_coords(Gf, lt, tx, ty, tz);
  for (int idx = threadIdx.x; idx < 64 * CG; idx += kPoolThreads) {
    const int v = idx / CG, cg = idx - (idx / CG) * CG;
    const int x = tx * 4 + (v & 3), y = ty * 4 + ((v >> 2) & 3), z = tz * 4 + (v >> 4);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const bool valid = real && x < Gf.n[0] && y < Gf.n[1] && z < Gf.n[2];
    if (valid) {
      const int X = x >> 1, Y = y >> 1, Z = z >> 1;
      const int ct = lv_tile(G, X >> 2, Y >> 2, Z >> 2);
      if (ct >= 0) {
        const int cl = (X & 3) + 4 * (Y & 3) + 16 * (Z & 3);
        bf16x8 p = *reinterpret_cast<const bf16x8*>(in + ((int64_t)ct * 64 + cl) * C + cg * 8);
#pragma unroll
        for (int u = 0; u < 8; ++u) s[u] = (float)p[u] * scale;
      }
    }
    bf16x8* dst = reinterpret_cast<bf16x8*>(out + ((int64_t)tile * 64 + v) * C + cg * 8);
    bf16x8 r;
    if (accumulate) {
      bf16x8 old = *dst;
#pragma unroll
      for (int u = 0; u < 8; ++u) r[u] = (bf16)((float)old[u] + s[u]);
    } else {
#pragma unroll
      for (int u = 0; u < 8; ++u) r[u] = (bf16)s[u];
    }
    *dst = r;
  }
}

// ------------------------------------------------------------ elementwise
__global__ void k_relu_mask(bf16* g, const bf16* a, int64_t n) {
  LPC_STRIDE(i, n) {
    if (!((float)a[i] > 0.0f)) g[i] = (bf16)0.0f;
  }
}
__global__ void k_add_bf16(const bf16* a, const bf16* b, bf16* out, int64_t n) {
  LPC_STRIDE(i, n) out[i] = (bf16)((float)a[i] + (float)b[i]);
}
__global__ void k_transpose_flip(const bf16* W, bf16* WT, int cin, int cout) {
  const int64_t n = (int64_t)27 * cin * cout;
  LPC_STRIDE(i, n) {
    // i indexes W: co = i / (27 cin), f = off cin + ci
    const int co = (int)(i / (27 * cin));
    const int f = (int)(i - (int64_t)co * 27 * cin);
    const int off = f / cin, ci = f - off * cin;
    WT[(int64_t)ci * 27 * cout + (26 - off) * cout + co] = W[i];
  }
}
__global__ void k_adam(float* p, const float* g, float* m, float* v, bf16* pb, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2, float gscale) {
  LPC_STRIDE(i, n) {
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    pb[i] = (bf16)pi;
  }
}
__global__ void k_f32_to_bf16(const float* in, bf16* out, int64_t n) {
  LPC_STRIDE(i, n) out[i] = (bf16)in[i];
}
// order: the bodies by ascending tileBase
__global__ void k_level_tile_body(const LevelGrid* bodies, const int* order, int nb, int NT, int* tileBody) {
  LPC_STRIDE(ti, NT) {
    const int t = (int)ti;
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (bodies[order[mid]].tileBase <= t) lo = mid;
      else hi = mid - 1;
    }
    tileBody[t] = order[lo];
  }
}

// ------------------------------------------------------------ level-1 features and training gradient
__device__ __forceinline__ bool node_valid(const BodyGrid& G, int T, int l) {
  const int lt = T - G.tileBase;
  if (lt >= G.T[0] * G.T[1] * G.T[2]) return false;
  const int tx = lt % G.T[0], ty = (lt / G.T[0]) % G.T[1], tz = lt / (G.T[0] * G.T[1]);
  return tx * 4 + (l & 3) < G.n[0] && ty * 4 + ((l >> 2) & 3) < G.n[1] && tz * 4 + (l >> 4) < G.n[2];
}
__global__ void k_lpc_features(Batch B, const float* r, const float* diag, const float* mass, const int8_t* flags,
                               const float* scale, const CRScalars* sc, bf16* in) {
  // one block per node block (4 tiles = 256 nodes); 16 channels per node
  const int b = B.blockBody[blockIdx.x];
  if (sc && (sc[b].done || sc[b].lpc_off)) return;
  const BodyGrid G = B.bodies[b];
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int T = (int)(i >> 6), l = (int)(i & 63);
  const float s = scale ? scale[b] : sc[b].lpc_scale;
  const float inv = s > 0.0f ? 1.0f / s : 0.0f;
  bf16x8 lo, hi;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    lo[u] = (bf16)0.0f;
    hi[u] = (bf16)0.0f;
  }
  if (node_valid(G, T, l)) {
    const int64_t vb = (int64_t)T * 192 + l;
    const float m = mass[i];
    const float col = flags && flags[i] != -1 ? 1.0f : 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float q = fminf(fmaxf(r[vb + 64 * c] / diag[vb + 64 * c] * inv, -kLpcClamp), kLpcClamp);
      lo[c] = (bf16)q;
      lo[3 + c] = (bf16)(q * (m / diag[vb + 64 * c]));
      if (c < 2) lo[6 + c] = (bf16)(q * col);
      else hi[0] = (bf16)(q * col);
    }
  }
  bf16x8* dst = reinterpret_cast<bf16x8*>(in + i * 16);
  dst[0] = lo;
  dst[1] = hi;
}
__global__ void k_lpc_out_grad(Batch B, const float* raw, const float* e, const float* r, const float* diag,
                               const float* scale, const float* weight, bf16* g, double* loss) {
  const int b = B.blockBody[blockIdx.x];
  const BodyGrid G = B.bodies[b];
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int T = (int)(i >> 6), l = (int)(i & 63);
  const float wb = weight[b];
  const float inv = scale[b] > 0.0f ? 1.0f / scale[b] : 0.0f;
  bf16x8 lo, hi;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    lo[u] = (bf16)0.0f;
    hi[u] = (bf16)0.0f;
  }
  double ls = 0.0;
  if (wb > 0.0f && node_valid(G, T, l)) {
    const int64_t vb = (int64_t)T * 192 + l;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float t = fminf(fmaxf((e[vb + 64 * c] - r[vb + 64 * c] / diag[vb + 64 * c]) * inv, -kLpcClamp), kLpcClamp);
      const float d = raw[vb + 64 * c] - t;
      lo[c] = (bf16)(2.0f * wb * d);
      ls += (double)wb * (double)d * (double)d;
    }
  }
  bf16x8* dst = reinterpret_cast<bf16x8*>(g + i * 16);
  dst[0] = lo;
  dst[1] = hi;
  for (int off = 32; off > 0; off >>= 1) ls += __shfl_down(ls, off, 64);
  if ((threadIdx.x & 63) == 0 && ls != 0.0) atomicAdd(loss, ls);
}

__global__ void k_sub_f32(const float* a, const float* b, float* out, int64_t n) {
  LPC_STRIDE(i, n) out[i] = a[i] - b[i];
}
// per-body sums over a captured residual: (r / diag)^2 and 3 x count over nodes with m > 0, and
// r^2 / diag over all nodes of the grid
__global__ void k_lpc_sample_stats(Batch B, const float* r, const float* diag, const float* mass, double* sums) {
  const int b = B.blockBody[blockIdx.x];
  const BodyGrid G = B.bodies[b];
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int T = (int)(i >> 6), l = (int)(i & 63);
  double s[3] = {0.0, 0.0, 0.0};
  if (node_valid(G, T, l)) {
    const int64_t vb = (int64_t)T * 192 + l;
    for (int c = 0; c < 3; ++c) {
      const float qf = r[vb + 64 * c] / diag[vb + 64 * c];
      s[2] += (double)r[vb + 64 * c] * (double)qf;
      if (mass[i] > 0.0f) s[0] += (double)qf * (double)qf;
    }
    if (mass[i] > 0.0f) s[1] = 3.0;
  }
  __shared__ double red[3][kThreads / 64];
  for (int k = 0; k < 3; ++k) {
    double v = s[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[threadIdx.x][w];
    if (t != 0.0) atomicAdd(sums + 3 * b + threadIdx.x, t);
  }
}

}  // namespace

// ============================================================ launchers
#define CONV_CASE(CI, CO, MD)                                                                                  \
  if (cin == CI && cout == CO && mode == MD) {                                                                 \
    if (V.NT < kConvFewTiles)                                                                                  \
      hipLaunchKernelGGL((k_conv<CI, CO, MD, 4>), dim3(V.NT), dim3(256), 0, s, V, in, W, bias, o);            \
    else                                                                                                       \
      hipLaunchKernelGGL((k_conv<CI, CO, MD, 1>), dim3(V.NT), dim3(64), 0, s, V, in, W, bias, o);             \
    return;                                                                                                    \
  }
void launch_conv(int cin, int cout, int mode, const LevelView& V, const bf16* in, const bf16* W, const float* bias,
                 const ConvOut& o, hipStream_t s) {
  if (V.NT <= 0) return;
  CONV_CASE(16, 16, CONV_RELU)
  CONV_CASE(16, 32, CONV_RELU)
  CONV_CASE(32, 64, CONV_RELU)
  CONV_CASE(64, 32, CONV_RELU)
  CONV_CASE(32, 16, CONV_RELU)
  CONV_CASE(16, 16, CONV_FINAL)
  // data-gradient passes (transposed, flipped weights)
  CONV_CASE(16, 16, CONV_LINEAR)
  CONV_CASE(32, 16, CONV_LINEAR)
  CONV_CASE(64, 32, CONV_LINEAR)
  CONV_CASE(32, 64, CONV_LINEAR)
  CONV_CASE(16, 32, CONV_LINEAR)
}
#define DW_CASE(CI, CO)                                                                                          \
  if (cin == CI && cout == CO) {                                                                                 \
    hipLaunchKernelGGL((k_conv_dw<CI, CO>), dim3(groups, cdiv(27 * CI / 16, 16 / (CO / 16))), dim3(kThreads), 0, s, V, in, g, \
                       partial, tilesPerGroup);                                                                  \
    return;                                                                                                      \
  }
void launch_conv_dw(int cin, int cout, const LevelView& V, const bf16* in, const bf16* g, float* partial,
                    int tilesPerGroup, hipStream_t s) {
  if (V.NT <= 0) return;
  const int groups = cdiv(V.NT, tilesPerGroup);
  DW_CASE(16, 16)
  DW_CASE(16, 32)
  DW_CASE(32, 64)
  DW_CASE(64, 32)
  DW_CASE(32, 16)
}
void launch_reduce_groups(const float* partial, int groups, int64_t len, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_groups, dim3(grid1d(len)), dim3(kThreads), 0, s, partial, groups, len, out);
}
void launch_bias_partial(int c, const LevelView& V, const bf16* g, float* partial, int tilesPerGroup, hipStream_t s) {
  if (V.NT <= 0) return;
  hipLaunchKernelGGL(k_bias_partial, dim3(cdiv(V.NT, tilesPerGroup)), dim3(kThreads), 0, s, c, V, g, partial,
                     tilesPerGroup);
}
void launch_pool(int c, const LevelView& fine, const LevelView& coarse, const bf16* in, bf16* out, float scale,
                 hipStream_t s) {
  if (coarse.NT <= 0) return;
  if (c == 16) hipLaunchKernelGGL(k_pool<16>, dim3(coarse.NT), dim3(kThreads), 0, s, fine, coarse, in, out, scale);
  else if (c == 32) hipLaunchKernelGGL(k_pool<32>, dim3(coarse.NT), dim3(kThreads), 0, s, fine, coarse, in, out, scale);
  else hipLaunchKernelGGL(k_pool<64>, dim3(coarse.NT), dim3(kThreads), 0, s, fine, coarse, in, out, scale);
}
void launch_unpool(int c, const LevelView& fine, const LevelView& coarse, const bf16* in, bf16* out, float scale,
                   bool accumulate, hipStream_t s) {
  if (fine.NT <= 0) return;
  const int acc = accumulate ? 1 : 0;
  if (c == 16)
    hipLaunchKernelGGL(k_unpool<16>, dim3(fine.NT), dim3(kPoolThreads), 0, s, fine, coarse, in, out, scale, acc);
  else if (c == 32)
    hipLaunchKernelGGL(k_unpool<32>, dim3(fine.NT), dim3(kPoolThreads), 0, s, fine, coarse, in, out, scale, acc);
  else
    hipLaunchKernelGGL(k_unpool<64>, dim3(fine.NT), dim3(kPoolThreads), 0, s, fine, coarse, in, out, scale, acc);
}
void launch_relu_mask(bf16* g, const bf16* a, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_relu_mask, dim3(grid1d(n)), dim3(kThreads), 0, s, g, a, n);
}
void launch_add_bf16(const bf16* a, const bf16* b, bf16* out, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_add_bf16, dim3(grid1d(n)), dim3(kThreads), 0, s, a, b, out, n);
}
void launch_transpose_flip(const bf16* W, bf16* WT, int cin, int cout, hipStream_t s) {
  hipLaunchKernelGGL(k_transpose_flip, dim3(grid1d((int64_t)27 * cin * cout)), dim3(kThreads), 0, s, W, WT, cin,
                     cout);
}
void launch_adam(float* p, const float* g, float* m, float* v, bf16* pb, int64_t n, float lr, float b1, float b2,
                 float eps, float c1, float c2, float gscale, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_adam, dim3(grid1d(n)), dim3(kThreads), 0, s, p, g, m, v, pb, n, lr, b1, b2, eps, c1, c2,
                       gscale);
}
void launch_f32_to_bf16(const float* in, bf16* out, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid1d(n)), dim3(kThreads), 0, s, in, out, n);
}
void launch_level_tile_body(const LevelGrid* bodies, const int* order, int nb, int NT, int* tileBody, hipStream_t s) {
  if (NT > 0)
    hipLaunchKernelGGL(k_level_tile_body, dim3(grid1d(NT)), dim3(kThreads), 0, s, bodies, order, nb, NT, tileBody);
}
void launch_lpc_features(const Batch& B, const float* r, const float* diag, const float* mass, const int8_t* flags,
                         const float* scale, const CRScalars* sc, bf16* in, hipStream_t s) {
  if (B.nblocks > 0)
    hipLaunchKernelGGL(k_lpc_features, dim3(B.nblocks), dim3(kThreads), 0, s, B, r, diag, mass, flags, scale, sc, in);
}
__global__ void k_sumsq(const float* g, int64_t n, double* out) {
  double s = 0.0;
  LPC_STRIDE(i, n) s += (double)g[i] * (double)g[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0 && s != 0.0) atomicAdd(out, s);
}
void launch_sumsq(const float* g, int64_t n, double* out, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_sumsq, dim3(grid1d(n, 1024)), dim3(kThreads), 0, s, g, n, out);
}
void launch_sub_f32(const float* a, const float* b, float* out, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_sub_f32, dim3(grid1d(n)), dim3(kThreads), 0, s, a, b, out, n);
}
void launch_lpc_sample_stats(const Batch& B, const float* r, const float* diag, const float* mass, double* sums,
                             hipStream_t s) {
  if (B.nblocks > 0)
    hipLaunchKernelGGL(k_lpc_sample_stats, dim3(B.nblocks), dim3(kThreads), 0, s, B, r, diag, mass, sums);
}
void launch_lpc_out_grad(const Batch& B, const float* raw, const float* e, const float* r, const float* diag,
                         const float* scale, const float* weight, bf16* g, double* loss, hipStream_t s) {
  if (B.nblocks > 0)
    hipLaunchKernelGGL(k_lpc_out_grad, dim3(B.nblocks), dim3(kThreads), 0, s, B, raw, e, r, diag, scale, weight, g,
                       loss);
}

}  // namespace mpm
