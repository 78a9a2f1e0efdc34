// mpm_kernels.hip -- gfx950 kernels of the implicit MPM snow step.
//
// Layout and scatter scheme: see mpm_device.hpp. The arithmetic of each step
// follows the reference (davidsminor/mpmsnow, paths relative to its checkout)
// as cited per kernel.
#include <climits>

#include "mpm_kernels.hpp"

namespace mpm {

namespace {

inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
inline int grid_for(int64_t n, int cap = 8192) {
  int b = cdiv(n, kBlock);
  return b < 1 ? 1 : (b > cap ? cap : b);
}

#define GRID_STRIDE(i, n) for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (n); i += (int64_t)gridDim.x * kBlock)

// ------------------------------------------------------------ geometry helpers
__device__ __forceinline__ void tile_xyz(const BodyGrid& G, int lt, int& tx, int& ty, int& tz) {
  tx = lt % G.T[0];
  int r = lt / G.T[0];
  ty = r % G.T[1];
  tz = r / G.T[1];
}
__device__ __forceinline__ int tile_id(const BodyGrid& G, int tx, int ty, int tz) {
  if (tx < 0 || ty < 0 || tz < 0 || tx >= G.T[0] || ty >= G.T[1] || tz >= G.T[2]) return -1;
  return G.tileBase + tx + G.T[0] * (ty + G.T[1] * tz);
}
// LDS footprints are linear boxes: node (rx, ry, rz) in [0, 8) of a chunk's 2x2x2-tile footprint
// sits at rx + 8 ry + kFootSZ rz, so a particle's 64 stencil nodes are constant offsets from its
// base node (immediate LDS offsets, no per-node address arithmetic). The z stride 66 (not 64)
// spreads the bases of a wave's particles over all 32 LDS banks: with 64 the bank ignores rz and
// a wave's scatter and gather run at about twice the bank conflicts.
constexpr int kFootSZ = 66;
constexpr int kFootLds = 7 + 7 * 8 + 7 * kFootSZ + 1;  // LDS words of one footprint component
__device__ __forceinline__ int foot_index(int rx, int ry, int rz) { return rx + 8 * ry + kFootSZ * rz; }
// Loops that move a footprint between LDS and the tile-major global layout run over j = 64 f + l
// (footprint tile f = (x >> 2) | (y >> 2) << 1 | (z >> 2) << 2, local node
// l = (x & 3) | (y & 3) << 2 | (z & 3) << 4), so global accesses stay contiguous; this gives the
// LDS index of j.
__device__ __forceinline__ void foot_xyz(int j, int& x, int& y, int& z) {
  const int f = j >> 6, l = j & 63;
  x = (f & 1) << 2 | (l & 3);
  y = ((f >> 1) & 1) << 2 | ((l >> 2) & 3);
  z = (f >> 2) << 2 | (l >> 4);
}
__device__ __forceinline__ int foot_lds(int j) {
  int x, y, z;
  foot_xyz(j, x, y, z);
  return foot_index(x, y, z);
}
// A chunk's particles have their base node in the chunk's tile (local 0-3; the scatters check
// it), so their stencils reach at most 3 nodes into each next tile: footprint node (f, l) can
// receive a contribution only where every axis shifted into the next tile (bit of f set) has a
// local coordinate below 3. 343 of the 512 staged nodes; the others are never written or read.
__device__ __forceinline__ bool foot_live(int f, int l) {
  return !(((f & 1) && (l & 3) == 3) || ((f & 2) && ((l >> 2) & 3) == 3) || ((f & 4) && (l >> 4) == 3));
}
// grid node position h i + min with the reference's two roundings (no fused multiply-add)
__device__ __forceinline__ void node_pos(const BodyGrid& G, int i, int j, int k, float x[3]) {
#pragma clang fp contract(off)
  x[0] = G.h * (float)i + G.gmin[0];
  x[1] = G.h * (float)j + G.gmin[1];
  x[2] = G.h * (float)k + G.gmin[2];
}
// ImplicitUpdateMatrix::subspaceProject for one node (Grid.cpp:778-818). The collider's normal
// at the node (CollisionObject::grad, normalised) is evaluated once per solve, where the flags are
// set (k_node_explicit, k_node_normals), and read here from nrm at the node's vector offset b:
// the solver's kernels then carry no collider evaluation.
__device__ __forceinline__ void project(const float* nrm, int64_t b, int flag, float v[3]) {
  if (flag == -1) return;
  if (flag == -2) {
    v[0] = v[1] = v[2] = 0.0f;
    return;
  }
  const float n[3] = {nrm[b], nrm[b + 64], nrm[b + 128]};
  float nd = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
  v[0] -= nd * n[0];
  v[1] -= nd * n[1];
  v[2] -= nd * n[2];
}
__device__ __forceinline__ void load_x(const PView& P, int s, float x[3]) {
  x[0] = P.atu(C_X, s);
  x[1] = P.atu(C_X + 1, s);
  x[2] = P.atu(C_X + 2, s);
}
__device__ __forceinline__ void load_m9(const PView& P, int col, int s, float m[9]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) m[i] = P.atu(col + i, s);
}
__device__ __forceinline__ void store_m9(const PView& P, int col, int s, const float m[9]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) P.atu(col + i, s) = m[i];
}

// ------------------------------------------------------------ chunk context
struct ChunkCtx {
  int4 ch;       // (global tile, start in perm, count, body)
  BodyGrid G;
  int ox, oy, oz;  // node origin of the chunk's tile
  const float* nrm;  // Batch::nrm
};
// all threads: the chunk, its body, and (threads 0..7) the footprint's global tiles
__device__ __forceinline__ ChunkCtx chunk_begin_at(const Batch& B, int* ft, int4 ch) {
  ChunkCtx c;
  c.ch = ch;
  c.nrm = B.nrm;
  c.G = B.bodies[c.ch.w];
  int tx, ty, tz;
  tile_xyz(c.G, c.ch.x - c.G.tileBase, tx, ty, tz);
  c.ox = tx * 4;
  c.oy = ty * 4;
  c.oz = tz * 4;
  if (threadIdx.x < 8) {
    int f = threadIdx.x;
    ft[f] = tile_id(c.G, tx + (f & 1), ty + ((f >> 1) & 1), tz + (f >> 2));
  }
  return c;
}
__device__ __forceinline__ ChunkCtx chunk_begin(const Batch& B, int* ft, int chunk) {
  return chunk_begin_at(B, ft, B.chunks[chunk]);
}
__device__ __forceinline__ bool chunk_stencil(const ChunkCtx& c, const float x[3], bool derivs, Stencil& st, int& rx,
                                              int& ry, int& rz) {
  make_stencil(x, c.G.gmin, c.G.h, derivs, st);
  rx = st.base[0] - c.ox;
  ry = st.base[1] - c.oy;
  rz = st.base[2] - c.oz;
  return rx >= 0 && rx < 4 && ry >= 0 && ry < 4 && rz >= 0 && rz < 4;
}
// load a batch 3-vector footprint into LDS (optionally projected); d and flags hold tiles from
// global tile tileOff on (0: the whole batch). Nodes no stencil of the chunk reaches (foot_live)
// are set to zero without a load.
__device__ __forceinline__ void load_foot3(const ChunkCtx& c, const int* ft, const float* d, float (*sd)[kFootLds],
                                           const int8_t* flags, const ColliderSet* cols, bool proj,
                                           int tileOff = 0) {
  for (int j = threadIdx.x; j < kFoot; j += kBlock) {
    const int gt = ft[j >> 6], l = j & 63, i = foot_lds(j);
    float v[3] = {0.0f, 0.0f, 0.0f};
    if (gt >= 0 && foot_live(j >> 6, l)) {
      int64_t b = (int64_t)(gt - tileOff) * 192 + l;
      v[0] = d[b];
      v[1] = d[b + 64];
      v[2] = d[b + 128];
      if (proj) {
        int fl = flags[(int64_t)(gt - tileOff) * 64 + l];
        if (fl != -1) project(c.nrm, (int64_t)gt * 192 + l, fl, v);
      }
    }
    sd[0][i] = v[0];
    sd[1][i] = v[1];
    sd[2][i] = v[2];
  }
}
// ------------------------------------------------------------ fixed-point footprint accumulation
// LDS float atomics (ds_add_f32) run at a small fraction of the rate of integer ones on gfx950,
// so each chunk accumulates its footprint scatter in fixed point with integer LDS atomics. A
// chunk picks a power-of-two scale S from T, the sum over its particles of a bound on any one
// contribution, so that every scaled contribution and every node sum stays below 2^43; each
// scaled value y is split as y = h 2^22 + l (h = trunc(y 2^-22), |l| < 2^22) and the integer
// parts of h and l go to two 32-bit accumulators (node sums of h stay below 2^21, of l below
// 2^30 for at most 256 particles). The resolution is 2^-43 of T, far finer than float
// rounding, and integer addition is order-independent: the staged footprints are bitwise
// reproducible. A chunk with a non-finite T stages NaN.
static_assert(kChunk == kBlock, "one particle per thread in chunk kernels");
template <int NC> struct FxFoot {
  int hi[NC][kFootLds];
  int lo[NC][kFootLds];
};
template <int NC> __device__ __forceinline__ void fx_zero(FxFoot<NC>& a) {
  for (int i = threadIdx.x; i < kFootLds; i += kBlock)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      a.hi[c][i] = 0;
      a.lo[c][i] = 0;
    }
}
__device__ __forceinline__ void fx_add(int* hi, int* lo, float y) {
  const float h = truncf(y * 2.384185791015625e-07f);  // 2^-22
  const float l = fmaf(h, -4194304.0f, y);             // exact: y - h 2^22
  atomicAdd(hi, (int)h);
  atomicAdd(lo, (int)l);
}
struct FxScale {
  float S, invS;
  bool bad;
};
// T: per-thread contribution bounds, summed over the block (every thread calls); sums of scaled
// contributions stay below 2^(top + 1)
__device__ __forceinline__ FxScale fx_scale(float bound, int top = 42) {
  __shared__ float red[kBlock / 64];
  float v = bound;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const float T = ((red[0] + red[1]) + red[2]) + red[3];
  FxScale r;
  r.bad = !(T <= 3.0e38f);  // NaN or infinite
  int k = 126;
  if (T > 0.0f && !r.bad) {
    k = top - ilogbf(T);  // T < 2^(ilogb + 1)
    k = k > 126 ? 126 : (k < -126 ? -126 : k);
  }
  r.S = __int_as_float((k + 127) << 23);
  r.invS = __int_as_float((127 - k) << 23);
  return r;
}
// staged layout per chunk: [footprint tile f][component c][64 local nodes]
template <int NC>
__device__ __forceinline__ void fx_store(const FxFoot<NC>& a, const FxScale& sc, float* stage, int slot = -1) {
  float* dst = stage + (int64_t)(slot < 0 ? (int)blockIdx.x : slot) * kStage;
  const double inv = (double)sc.invS;
  for (int j = threadIdx.x; j < kFoot; j += kBlock) {
    const int f = j >> 6, l = j & 63, i = foot_lds(j);
    if (!foot_live(f, l)) continue;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double v = ((double)a.hi[c][i] * 4194304.0 + (double)a.lo[c][i]) * inv;
      dst[f * 192 + c * 64 + l] = sc.bad ? __int_as_float(0x7fc00000) : (float)v;
    }
  }
}
// Single-word variant for the HVP, the solver's inner kernel: one integer per value and a
// scale per component c from T_c, the chunk's sum of per-particle bounds on that component's
// contributions (tight: the stencil's gradient weights are at most (2/3)^3 / h), so that sums
// stay below 2^30; each contribution is rounded to the nearest unit. Resolution 2^-30 of T_c,
// with half the LDS atomics and a fifth of the conversion work of the split form.
template <int NC> struct FxFoot1 {
  int v[NC][kFootLds];
};
struct FxScale3 {
  float S[3], invS[3];
  bool bad;
};
__device__ __forceinline__ FxScale3 fx_scale3(const float bound[3], int top) {
  __shared__ float red[3][kBlock / 64];
  float v[3] = {bound[0], bound[1], bound[2]};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    for (int off = 32; off > 0; off >>= 1) v[c] += __shfl_xor(v[c], off, 64);
    if ((threadIdx.x & 63) == 0) red[c][threadIdx.x >> 6] = v[c];
  }
  __syncthreads();
  FxScale3 r;
  r.bad = false;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float T = ((red[c][0] + red[c][1]) + red[c][2]) + red[c][3];
    r.bad = r.bad || !(T <= 3.0e38f);
    int k = 126;
    if (T > 0.0f && T <= 3.0e38f) {
      k = top - ilogbf(T);
      k = k > 126 ? 126 : (k < -126 ? -126 : k);
    }
    r.S[c] = __int_as_float((k + 127) << 23);
    r.invS[c] = __int_as_float((127 - k) << 23);
  }
  return r;
}
template <int NC> __device__ __forceinline__ void fx_zero(FxFoot1<NC>& a) {
  for (int i = threadIdx.x; i < kFootLds; i += kBlock)
#pragma unroll
    for (int c = 0; c < NC; ++c) a.v[c][i] = 0;
}
__device__ __forceinline__ void fx_store(const FxFoot1<3>& a, const FxScale3& sc, float* stage, int slot) {
  float* dst = stage + (int64_t)slot * kStage;
  for (int j = threadIdx.x; j < kFoot; j += kBlock) {
    const int f = j >> 6, l = j & 63, i = foot_lds(j);
    if (!foot_live(f, l)) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      dst[f * 192 + c * 64 + l] = sc.bad ? __int_as_float(0x7fc00000) : (float)a.v[c][i] * sc.invS[c];
  }
}
constexpr float kGradWeightMax = 0.2963f;  // (2/3)^3 rounded up: |ghat_c| h over the stencil
constexpr int kFx1Top = 29;

// Separable stencil sums. The gradient weight of stencil node (i, j, k) is
// ghat = (dwx_i wy_j wz_k, wx_i dwy_j wz_k, wx_i wy_j dwz_k) (st_grad returns -ghat); for a
// footprint value v (LDS, linear layout, v[0] = the particle's base node) this returns
// out[c] = sum_ijk v(i, j, k) ghat_c(i, j, k) by contracting one axis at a time (188 FMAs
// instead of 15 operations per node).
__device__ __forceinline__ void sep_gather(const float* v, const Stencil& st, float out[3]) {
  float ex[4], fy[4], hz[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    ex[k] = fy[k] = hz[k] = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float d = v[i + 8 * j + kFootSZ * k];
        a0 = fmaf(d, st.w[0][i], a0);
        a1 = fmaf(d, st.dw[0][i], a1);
      }
      ex[k] = fmaf(a1, st.w[1][j], ex[k]);
      fy[k] = fmaf(a0, st.dw[1][j], fy[k]);
      hz[k] = fmaf(a0, st.w[1][j], hz[k]);
    }
  }
  out[0] = out[1] = out[2] = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    out[0] = fmaf(ex[k], st.w[2][k], out[0]);
    out[1] = fmaf(fy[k], st.w[2][k], out[1]);
    out[2] = fmaf(hz[k], st.dw[2][k], out[2]);
  }
}
// floor(x + 0.5) as an int in one instruction (v_cvt_rpi_i32_f32; the compiler's forms of a
// rounding conversion take two or three)
__device__ __forceinline__ int cvt_round(float x) {
  int r;
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}
// acc(i, j, k) += round(f . ghat(i, j, k)) over the stencil, with f . ghat evaluated as
// P(j, k) dwx_i + Q(j, k) wx_i, P = f0 wy_j wz_k, Q = f1 dwy_j wz_k + f2 wy_j dwz_k. The chunk's
// scale is shared by all its nodes, so a node that few or light particles reach holds a value of
// few fixed-point units: rounding to nearest keeps each contribution's error within half a unit
// and unbiased (truncation toward zero, the same instruction count, erred by up to a unit, always
// toward zero, which left such nodes' HVP values biased by up to 1e-3 of themselves).
__device__ __forceinline__ void sep_scatter(int* acc, const Stencil& st, float f0, float f1, float f2) {
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float P = f0 * (st.w[1][j] * st.w[2][k]);
      const float Q = fmaf(f1, st.dw[1][j] * st.w[2][k], f2 * (st.w[1][j] * st.dw[2][k]));
#pragma unroll
      for (int i = 0; i < 4; ++i)
        atomicAdd(&acc[i + 8 * j + kFootSZ * k], cvt_round(fmaf(P, st.dw[0][i], Q * st.w[0][i])));
    }
}

// One z-plane k (0-3, uniform across a wave) of sep_gather and sep_scatter: the same terms, in the
// same order, so that four waves that take a plane each reproduce the one-thread results bitwise.
// &sd[a][base] as a register the compiler cannot fold into the reads' constant offsets: the
// stencil's reads then lie within ds_read2's 8-bit word offsets of it (at most
// 3 + 3 * 8 + 3 * kFootSZ = 225 words) and take no address arithmetic each, where a folded
// component offset (a * kFootLds words) would put every pair of reads of components 1 and 2 out
// of that range.
__device__ __forceinline__ const float* foot_at(float (*sd)[kFootLds], int a, int base) {
  int off = a * kFootLds + base;
  asm("" : "+v"(off));
  return reinterpret_cast<const float*>(sd) + off;
}
__device__ __forceinline__ float pick4(const float v[4], int k) {
  return k == 0 ? v[0] : (k == 1 ? v[1] : (k == 2 ? v[2] : v[3]));
}
__device__ __forceinline__ void sep_gather_plane(const float* v, const Stencil& st, int k, float e[3]) {
  e[0] = e[1] = e[2] = 0.0f;
  const float* vk = v + kFootSZ * k;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float d = vk[i + 8 * j];
      a0 = fmaf(d, st.w[0][i], a0);
      a1 = fmaf(d, st.dw[0][i], a1);
    }
    e[0] = fmaf(a1, st.w[1][j], e[0]);
    e[1] = fmaf(a0, st.dw[1][j], e[1]);
    e[2] = fmaf(a0, st.w[1][j], e[2]);
  }
}
__device__ __forceinline__ void sep_scatter_plane(int* acc, const Stencil& st, int k, float f0, float f1, float f2) {
  const float wz = pick4(st.w[2], k), dwz = pick4(st.dw[2], k);
  int* ak = acc + kFootSZ * k;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float P = f0 * (st.w[1][j] * wz);
    const float Q = fmaf(f1, st.dw[1][j] * wz, f2 * (st.w[1][j] * dwz));
#pragma unroll
    for (int i = 0; i < 4; ++i) atomicAdd(&ak[i + 8 * j], cvt_round(fmaf(P, st.dw[0][i], Q * st.w[0][i])));
  }
}

__device__ __forceinline__ float abs_sum9(const float m[9]) {
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < 9; ++i) s += fabsf(m[i]);
  return s;
}

// gather of staged footprints for node (tile T, local l) of body G, component c
__device__ __forceinline__ float gather_node(const Batch& B, const BodyGrid& G, const float* stage, int T, int l,
                                             int c) {
  int tx, ty, tz;
  tile_xyz(G, T - G.tileBase, tx, ty, tz);
  float v = 0.0f;
#pragma unroll
  for (int f = 0; f < 8; ++f) {
    int ts = foot_live(f, l) ? tile_id(G, tx - (f & 1), ty - ((f >> 1) & 1), tz - (f >> 2)) : -1;
    if (ts < 0) continue;
    int c0 = B.tileChunk0[ts], cn = B.tileChunkN[ts];
    for (int q = 0; q < cn; ++q) v += stage[(int64_t)(c0 + q) * kStage + f * 192 + c * 64 + l];
  }
  return v;
}
// tc0 / tcn: first chunk and chunk count of each tile from global tile tileOff on; stage holds
// the chunks from the one numbered tc0 = 0 on. Only the source tiles whose footprint reaches node
// l are read (foot_live). The chunk tables of all 8 source tiles and the
// first chunk of each are loaded before any sum (most tiles have one chunk), so a node waits for
// two rounds of memory latency instead of sixteen; the sums still run source tile by source
// tile and chunk by chunk, in the same order as a plain double loop.
__device__ __forceinline__ void gather3_tab(const BodyGrid& G, const int* tc0, const int* tcn, int tileOff,
                                            const float* stage, int T, int l, float o[3]) {
  int tx, ty, tz;
  tile_xyz(G, T - G.tileBase, tx, ty, tz);
  int c0[8], cn[8];
#pragma unroll
  for (int f = 0; f < 8; ++f) {
    const int ts = foot_live(f, l) ? tile_id(G, tx - (f & 1), ty - ((f >> 1) & 1), tz - (f >> 2)) : -1;
    c0[f] = 0;
    cn[f] = 0;
    if (ts >= 0) {
      c0[f] = tc0[ts - tileOff];
      cn[f] = tcn[ts - tileOff];
    }
  }
  float v[8][3];
#pragma unroll
  for (int f = 0; f < 8; ++f) {
    v[f][0] = v[f][1] = v[f][2] = 0.0f;
    if (cn[f] > 0) {
      const float* s = stage + (int64_t)c0[f] * kStage + f * 192 + l;
      v[f][0] = s[0];
      v[f][1] = s[64];
      v[f][2] = s[128];
    }
  }
  float a = 0.0f, b = 0.0f, cc = 0.0f;
#pragma unroll
  for (int f = 0; f < 8; ++f) {
    if (cn[f] > 0) {
      a += v[f][0];
      b += v[f][1];
      cc += v[f][2];
    }
    for (int q = 1; q < cn[f]; ++q) {
      const float* s = stage + (int64_t)(c0[f] + q) * kStage + f * 192 + l;
      a += s[0];
      b += s[64];
      cc += s[128];
    }
  }
  o[0] = a;
  o[1] = b;
  o[2] = cc;
}
__device__ __forceinline__ void gather_node3(const Batch& B, const BodyGrid& G, const float* stage, int T, int l,
                                             float o[3]) {
  gather3_tab(G, B.tileChunk0, B.tileChunkN, 0, stage, T, l, o);
}

// ------------------------------------------------------------ node block context
struct NodeCtx {
  int b;        // body
  BodyGrid G;
  int64_t i;    // global node
  int T, l;     // global tile, local node
  int ni, nj, nk;
  bool valid;   // inside the body's reference grid
  const float* nrm;  // Batch::nrm
};
__device__ __forceinline__ NodeCtx node_in(const BodyGrid& G, int body, int blk, const float* nrm) {
  NodeCtx c;
  c.b = body;
  c.G = G;
  c.nrm = nrm;
  c.i = (int64_t)blk * kBlock + threadIdx.x;
  c.T = (int)(c.i >> 6);
  c.l = (int)(c.i & 63);
  int tx, ty, tz;
  tile_xyz(c.G, c.T - c.G.tileBase, tx, ty, tz);
  c.ni = tx * 4 + (c.l & 3);
  c.nj = ty * 4 + ((c.l >> 2) & 3);
  c.nk = tz * 4 + (c.l >> 4);
  c.valid = c.ni < c.G.n[0] && c.nj < c.G.n[1] && c.nk < c.G.n[2] && (c.T - c.G.tileBase) < c.G.T[0] * c.G.T[1] * c.G.T[2];
  return c;
}
__device__ __forceinline__ NodeCtx node_begin(const Batch& B) {
  const int b = B.blockBody[blockIdx.x];
  return node_in(B.bodies[b], b, blockIdx.x, B.nrm);
}
__device__ __forceinline__ void ld3(const float* v, int64_t i, float o[3]) {
  int64_t b = (i >> 6) * 192 + (i & 63);
  o[0] = v[b];
  o[1] = v[b + 64];
  o[2] = v[b + 128];
}
__device__ __forceinline__ void st3(float* v, int64_t i, const float o[3]) {
  int64_t b = (i >> 6) * 192 + (i & 63);
  v[b] = o[0];
  v[b + 64] = o[1];
  v[b + 128] = o[2];
}
__device__ __forceinline__ void node_project(const NodeCtx& c, const ColliderSet*, int fl, float v[3]) {
  if (fl == -1) return;
  project(c.nrm, (c.i >> 6) * 192 + (c.i & 63), fl, v);
}
// ImplicitUpdateMatrix::multVector epilogue (Grid.cpp:764-770): P(m P d - dt^2 df)
__device__ __forceinline__ void apply_A(const NodeCtx& c, const ColliderSet* cols, int fl, float m, const float d[3],
                                        const float df[3], float dt2, float out[3]) {
  float pd[3] = {d[0], d[1], d[2]};
  node_project(c, cols, fl, pd);
  out[0] = m * pd[0] - dt2 * df[0];
  out[1] = m * pd[1] - dt2 * df[1];
  out[2] = m * pd[2] - dt2 * df[2];
  node_project(c, cols, fl, out);
}

// ------------------------------------------------------------ per-body reductions
template <int NV> __device__ __forceinline__ void block_sum(double (&v)[NV], double (*sm)[NV]) {
#pragma unroll
  for (int k = 0; k < NV; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sm[wid][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = ((sm[0][k] + sm[1][k]) + sm[2][k]) + sm[3][k];
}
__device__ __forceinline__ unsigned int* body_tickets(const Batch& B, int body, int set) {
  return B.tickets + (int64_t)body * kTicketsPerBody + (int64_t)set * kSetStride;
}
// Adds, in thread t, the NV-vectors rows[t], rows[t + kBlock], ... (row r at src + r * stride,
// n rows) in that order, loading U rows at a time, then sums over the block (block_sum); the
// total is in thread 0. Past the end it adds +0.0, which leaves a sum that started at +0.0
// unchanged.
template <int NV>
__device__ __forceinline__ void rows_sum(const double* src, int64_t stride, int n, double (&acc)[NV],
                                         double (*sm)[NV]) {
  constexpr int U = NV <= 2 ? 8 : (NV <= 4 ? 4 : 2);
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int r0 = threadIdx.x; r0 < n; r0 += U * kBlock) {
    double pv[U][NV];
    bool on[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + u * kBlock;
      on[u] = r < n;
      const double* row = src + (int64_t)(on[u] ? r : 0) * stride;
#pragma unroll
      for (int k = 0; k < NV; ++k) pv[u][k] = __hip_atomic_load(row + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int k = 0; k < NV; ++k) acc[k] += on[u] ? pv[u][k] : 0.0;
  }
  block_sum<NV>(acc, sm);
}
// Stores v (thread 0) at dst with agent-scope atomic stores, waits for their completion, then
// takes a ticket: true (thread 0) for the arrival that completes `arrivals` (the counter is
// reset for the next use). A release fence would write back the whole L2 of this XCD instead.
template <int NV>
__device__ __forceinline__ bool store_and_count(const double (&v)[NV], double* dst, unsigned int* ticket,
                                                unsigned int arrivals) {
#pragma unroll
  for (int k = 0; k < NV; ++k) __hip_atomic_store(dst + k, v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  if (__hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == arrivals - 1u) {
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
  }
  return false;
}
// Returns true in every thread of the last block of the body to finish, with out[] = the
// body's total: sub-group totals (members in block order) added in sub-group order. For a
// body of at most kSub blocks every sub-group is one block and this is the block-order sum.
template <int NV>
__device__ bool body_sum(double (&v)[NV], const BodyGrid& G, double* partials, unsigned int* tk,
                         double (&out)[NV]) {
  static_assert(NV <= kPartialStride, "partials stride");
  __shared__ double sm[kBlock / 64][NV];
  __shared__ bool amLast;
  double* subTot = reinterpret_cast<double*>(tk + kSetTickets);
  block_sum<NV>(v, sm);
  const int local = blockIdx.x - G.blockBase;
  const int sub = local % kSub;
  const int members = (G.nblocks - sub + kSub - 1) / kSub;
  const int nsubs = G.nblocks < kSub ? G.nblocks : kSub;
  if (threadIdx.x == 0)
    amLast = store_and_count<NV>(v, partials + (int64_t)blockIdx.x * kPartialStride, tk + (1 + sub) * kTicketStride,
                                 (unsigned int)members);
  __syncthreads();
  if (!amLast) return false;
  double acc[NV];
  rows_sum<NV>(partials + (int64_t)(G.blockBase + sub) * kPartialStride, (int64_t)kSub * kPartialStride, members,
               acc, sm);
  __syncthreads();
  if (threadIdx.x == 0) amLast = store_and_count<NV>(acc, subTot + sub * kPartialStride, tk, (unsigned int)nsubs);
  __syncthreads();
  if (!amLast) return false;
  rows_sum<NV>(subTot, kPartialStride, nsubs, acc, sm);
  __shared__ double res[NV];
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) res[k] = acc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = res[k];
  return true;
}
__device__ __forceinline__ void set_done(CRScalars& s, BatchStatus* st) {
  if (!s.done) {
    s.done = 1;
    atomicAdd(&st->bodies_done, 1);
  }
}

// ------------------------------------------------------------ particle store kernels
__global__ void k_set_column(PView P, int col0, int ncomp, const float* aos, const int* slot_of, int n) {
  GRID_STRIDE(t, (int64_t)n * ncomp) {
    int i = (int)(t / ncomp), c = (int)(t % ncomp);
    P.at(col0 + c, slot_of[i]) = aos[t];
  }
}
__global__ void k_get_column(PView P, int col0, int ncomp, float* aos, const int* slot_of, int n) {
  GRID_STRIDE(t, (int64_t)n * ncomp) {
    int i = (int)(t / ncomp), c = (int)(t % ncomp);
    aos[t] = P.at(col0 + c, slot_of[i]);
  }
}
__global__ void k_fill_column(PView P, int col, float value, int n) {
  GRID_STRIDE(s, n) P.atu(col, (int)s) = value;
}
__global__ void k_fill_int(int* p, int value, int64_t n) {
  GRID_STRIDE(i, n) p[i] = value;
}
__global__ void k_iota(int* p, int64_t n) {
  GRID_STRIDE(i, n) p[i] = (int)i;
}
__global__ void k_identity_columns(PView P, int col0, int n) {
  GRID_STRIDE(s, n) {
#pragma unroll
    for (int i = 0; i < 9; ++i) P.atu(col0 + i, (int)s) = (i == 0 || i == 4 || i == 8) ? 1.0f : 0.0f;
  }
}
__global__ void k_permute(PView src, PView dst, const int* new_to_old, int n, const int* orig_old, int* orig_new,
                          int* slot_of) {
  GRID_STRIDE(si, n) {
    int s = (int)si;
    int o = new_to_old[s];
#pragma unroll 4
    for (int c = 0; c < NCOL; ++c) dst.atu(c, s) = src.atu(c, o);
    int orig = orig_old[o];
    orig_new[s] = orig;
    slot_of[orig] = s;
  }
}
__global__ void k_gather_slots(const int* slot_of, const int* orig, int n, int* slots) {
  GRID_STRIDE(k, n) slots[k] = slot_of[orig[k]];
}
// the same for a device-resident index list that the host has not checked: an index outside
// [0, np) gives slot -1 and raises ERR_INDEX
__global__ void k_gather_slots_checked(const int* slot_of, const int* orig, int n, int np, int* slots, int* err) {
  GRID_STRIDE(k, n) {
    const int o = orig[k];
    const bool ok = o >= 0 && o < np;
    if (!ok) atomicOr(err, ERR_INDEX);
    slots[k] = ok ? slot_of[o] : -1;
  }
}

// ------------------------------------------------------------ body partition (Sim.cpp:136-203)
// Bucket of a cell: blocks of 4x4x4 cells placed by a spatial hash of the block, the 64 cells of
// a block in consecutive buckets. In bucket order the particles of a block are contiguous, so a
// workgroup of the sampling pass holds about one block of neighbouring cells.
__device__ __forceinline__ uint32_t cell_hash(int cx, int cy, int cz, uint32_t mask) {
  const uint32_t block = (uint32_t)(cx >> 2) * 73856093u ^ (uint32_t)(cy >> 2) * 19349663u ^
                         (uint32_t)(cz >> 2) * 83492791u;
  const uint32_t local = (uint32_t)(cx & 3) | ((uint32_t)(cy & 3) << 2) | ((uint32_t)(cz & 3) << 4);
  return ((block << 6) | local) & mask;
}
// |x - y|^2 as (d0 d0 + d1 d1) + d2 d2 with every product and sum rounded separately, like the
// host's Vector3f::squaredNorm: the neighbour test d^2 < h^2 must match it bit for bit (the HIP
// __fmul_rn / __fadd_rn helpers are plain operators that contraction would fuse)
__device__ __forceinline__ float sqdist_ref(const float x[3], float y0, float y1, float y2) {
#pragma clang fp contract(off)
  const float d0 = x[0] - y0, d1 = x[1] - y1, d2 = x[2] - y2;
  return (d0 * d0 + d1 * d1) + d2 * d2;
}
__device__ __forceinline__ void cell_of(const float x[3], float cell, int c[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) c[a] = (int)floorf(x[a] / cell);
}
// count[t]++ returning the old value, one atomic per group of lanes with the same t (particles in
// storage order mostly share a tile with their neighbours, whose lanes would otherwise queue on
// one counter); every active lane calls. The groups are found first (one ballot per distinct t)
// and their leaders then add at once, so a wave waits for one atomic round trip, not one per
// group. Ranks within a tile follow arrival order either way: k_sort_tiles fixes the final order.
__device__ __forceinline__ int wave_aggregated_inc(int* count, int t) {
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(1);
  int leader = lane, below = 0, size = 1;
  while (todo) {
    const int l = __ffsll((long long)todo) - 1;
    const int lt = __shfl(t, l, 64);
    const unsigned long long peers = __ballot(t == lt);
    if (t == lt) {
      leader = l;
      below = __popcll(peers & ((1ull << lane) - 1ull));
      size = __popcll(peers);
    }
    todo &= ~peers;
  }
  int base = 0;
  if (lane == leader) base = atomicAdd(count + t, size);
  base = __shfl(base, leader, 64);
  return base + below;
}
__global__ void k_cc_hash(PView P, int n, float cell, uint32_t mask, int* bucket, int* rank, int* count) {
  GRID_STRIDE(si, n) {
    int s = (int)si;
    float x[3];
    load_x(P, s, x);
    int c[3];
    cell_of(x, cell, c);
    uint32_t h = cell_hash(c[0], c[1], c[2], mask);
    bucket[s] = (int)h;
    rank[s] = wave_aggregated_inc(count, (int)h);  // the partition does not depend on this order
  }
}
// cell-sorted copy of the positions, slot in .w, so candidate scans read contiguous memory
__global__ void k_cc_scatter(PView P, const int* bucket, const int* rank, const int* offset, int n, float4* sorted) {
  GRID_STRIDE(si, n) {
    int s = (int)si;
    sorted[offset[bucket[s]] + rank[s]] = make_float4(P.atu(C_X, s), P.atu(C_X + 1, s), P.atu(C_X + 2, s), __int_as_float(s));
  }
}
// Link priority: a bijective integer mix, so hooking follows a pseudo-random order of the
// slots rather than their (spatially sorted) index order, which would build long chains.
__device__ __forceinline__ uint32_t cc_priority(int x) {
  uint32_t h = (uint32_t)x;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
// Parent pointers only ever point into their own set (hooking joins two sets, path halving
// moves a pointer up its own tree), so any value a pointer ever held names a member of the
// same set, and an ancestor of its node. Every read of the forest therefore goes through the
// cache (workgroup-scope loads); only the hook itself is a device-scope compare-and-swap.
// Device-coherent loads would end every find of a large component on the line of its root, where
// the whole grid queues.
__device__ __forceinline__ int cc_load_cached(const int* parent, int x) {
  return __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// root of x in kernels where no thread hooks roots (compression and root collection)
__device__ __forceinline__ int cc_find_cached(const int* parent, int x) {
  int p = cc_load_cached(parent, x);
  while (p != x) {
    x = p;
    p = cc_load_cached(parent, x);
  }
  return x;
}
// Root search for the hooking loop, through the cache, with path halving. Every value a pointer
// ever held is an ancestor (in the current forest) of its node, of strictly lower priority, so a
// stale read only moves the search to a lower ancestor; and a stale "root" is caught by the
// hook's compare-and-swap, whose returned value continues the search one real level up.
__device__ __forceinline__ int cc_find_halving(int* parent, int x) {
  int p = cc_load_cached(parent, x);
  while (p != x) {
    const int gp = cc_load_cached(parent, p);
    if (gp != p) __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = p;
    p = gp;
  }
  return x;
}
// hook the root of higher priority under the other; priorities strictly fall towards a root
__device__ __forceinline__ void cc_union(int* parent, int a, int b) {
  {
    // cheap tests that catch most pairs inside one already linked component: direct parents,
    // then the roots as the cache shows them (any pointer value ever held names a member of the
    // same set, so equal cached roots prove one set; unequal ones fall through to the hook)
    const int pa = cc_load_cached(parent, a), pb = cc_load_cached(parent, b);
    if (pa == pb || pa == b || pb == a) return;
    if (cc_find_cached(parent, pa) == cc_find_cached(parent, pb)) return;
  }
  while (true) {
    a = cc_find_halving(parent, a);
    b = cc_find_halving(parent, b);
    if (a == b) return;
    if (cc_priority(a) < cc_priority(b)) {
      int t = a;
      a = b;
      b = t;
    }
    const int old = atomicCAS(parent + a, a, b);
    if (old == a) return;
    a = old;  // a was no longer a root: continue from its current parent
  }
}
// union-find over a workgroup's particles in LDS (indices into the workgroup's range), hooking
// by the same priorities as the global forest so that a local root is its set's minimum
__device__ __forceinline__ int cc_local_find(int* lp, int x) {
  int p = __hip_atomic_load(lp + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  while (p != x) {
    x = p;
    p = __hip_atomic_load(lp + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  return x;
}
__device__ __forceinline__ void cc_local_union(int* lp, const uint32_t* prio, int a, int b) {
  while (true) {
    a = cc_local_find(lp, a);
    b = cc_local_find(lp, b);
    if (a == b) return;
    if (prio[a] < prio[b]) {
      const int t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(lp + a, a, b) == a) return;
  }
}
// Neighbour pairs at |x_i - x_j|^2 < h^2 in the reference's float arithmetic (KdTree.inl:187-190),
// linked in two passes after the subgraph-sampling scheme of Sutton et al. (Afforest, IPDPS
// 2018): pass 1 (k_cc_sample, k_cc_join) hooks every particle to its first kCcSampleLinks
// neighbours, which already joins most of each body, and records whether it has any neighbour;
// after compression and a sampled vote for the largest components, pass 2 (k_cc_link) hooks every
// remaining neighbour of the particles outside the skipped components: the largest, and with it
// every other voted component whose box (grown by the link distance) meets none of the boxes
// already skipped. A particle of a skipped component may skip its pairs: for a pair with one end
// outside them, that end links it; a pair with both ends in one is already joined; no pair has
// its ends in two of them, whose boxes are farther apart than the link distance. Membership is
// read from the pointers as compression left them (snap), which pass 2's hooks do not change.
// Threads walk the bucket order so a wave's candidate cells mostly coincide.
constexpr int kCcSampleLinks = 2;
// the components pass 2 skips (the vote's candidates, their boxes, the selection)
struct CcSkip {
  int cand[kCcSkipMax];
  unsigned int box[kCcSkipMax][6];  // min x, y, z, max x, y, z as float_order values
  int nskip;
  int skip[kCcSkipMax];
};
static_assert(sizeof(CcSkip) <= kCcSkipBytes, "skip state size");
// Pass 1 for one workgroup of kBlock consecutive particles in bucket order (about one block of
// cells): the sample links inside the workgroup's range are joined in LDS, every particle's
// global pointer is set to its local root (plain stores: no other workgroup writes pointers in
// this kernel), and the links that leave the range are kept for k_cc_join.
__global__ __launch_bounds__(kBlock) void k_cc_sample(int n, float h, float cell, uint32_t mask, const int* offset,
                                                      const int* count, const float4* sorted, int* parent,
                                                      int* hasNb, int* out0, int* out1) {
  __shared__ int lp[kBlock];
  __shared__ uint32_t prio[kBlock];
  __shared__ int lslot[kBlock];
  const float r2 = h * h;
  const int t = threadIdx.x;
  const int base = blockIdx.x * kBlock;
  const int ti = base + t;
  const bool act = ti < n;
  float4 me = act ? sorted[ti] : make_float4(0.f, 0.f, 0.f, 0.f);
  const int s = act ? __float_as_int(me.w) : -1;
  lp[t] = t;
  prio[t] = act ? cc_priority(s) : 0xffffffffu;
  lslot[t] = s;
  __syncthreads();
  int ext[kCcSampleLinks];
  int next = 0, found = 0;
  for (int k = 0; k < kCcSampleLinks; ++k) ext[k] = -1;
  if (act) {
    const float x[3] = {me.x, me.y, me.z};
    int c[3];
    cell_of(x, cell, c);
    for (int dz = -1; dz <= 1 && found < kCcSampleLinks; ++dz)
      for (int dy = -1; dy <= 1 && found < kCcSampleLinks; ++dy)
        for (int dx = -1; dx <= 1 && found < kCcSampleLinks; ++dx) {
          uint32_t hb = cell_hash(c[0] + dx, c[1] + dy, c[2] + dz, mask);
          int o = offset[hb], e = o + count[hb];
          for (int q = o; q < e && found < kCcSampleLinks; ++q) {
            float4 y = sorted[q];
            int j = __float_as_int(y.w);
            if (j == s) continue;
            if (sqdist_ref(x, y.x, y.y, y.z) < r2) {
              if (q >= base && q < base + kBlock) cc_local_union(lp, prio, t, q - base);
              else ext[next++] = j;
              ++found;
            }
          }
        }
    hasNb[s] = found > 0 ? 1 : 0;
    out0[ti] = ext[0];
    out1[ti] = ext[1];
  }
  __syncthreads();
  if (act) {
    const int r = cc_local_find(lp, t);
    if (r != t) parent[s] = lslot[r];
  }
}
// pass 1's links that left their workgroup's range, joined in the global forest
__global__ void k_cc_join(int n, const float4* sorted, const int* in0, const int* in1, int* parent) {
  GRID_STRIDE(ti, n) {
    const int a = in0[ti], b = in1[ti];
    if (a < 0 && b < 0) continue;
    const int s = __float_as_int(sorted[ti].w);
    if (a >= 0) cc_union(parent, s, a);
    if (b >= 0) cc_union(parent, s, b);
  }
}
// pass 2: every neighbour pair of the particles outside the skipped components
__global__ void k_cc_link(int n, float h, float cell, uint32_t mask, const int* offset, const int* count,
                          const float4* sorted, int* parent, const int* snap, const CcSkip* S) {
  const float r2 = h * h;
  const int ns = S->nskip;
  int sk[kCcSkipMax];
#pragma unroll
  for (int k = 0; k < kCcSkipMax; ++k) sk[k] = k < ns ? S->skip[k] : -1;
  GRID_STRIDE(ti, n) {
    float4 me = sorted[ti];
    const int s = __float_as_int(me.w);
    const int r = snap[s];
    bool skip = false;
#pragma unroll
    for (int k = 0; k < kCcSkipMax; ++k) skip = skip || sk[k] == r;
    if (skip) continue;
    const float x[3] = {me.x, me.y, me.z};
    int c[3];
    cell_of(x, cell, c);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          uint32_t hb = cell_hash(c[0] + dx, c[1] + dy, c[2] + dz, mask);
          int o = offset[hb], e = o + count[hb];
          for (int q = o; q < e; ++q) {
            float4 y = sorted[q];
            int j = __float_as_int(y.w);
            if (j == s) continue;
            if (sqdist_ref(x, y.x, y.y, y.z) < r2) cc_union(parent, s, j);
          }
        }
  }
}
// every pointer straight to its root; snap keeps these pass-1 roots
__global__ void k_cc_compress(int n, int* parent, int* snap) {
  GRID_STRIDE(s, n) {
    const int r = cc_find_cached(parent, (int)s);
    __hip_atomic_store(parent + s, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    snap[s] = r;
  }
}
// The kCcSkipMax most frequent roots among kCcVote evenly spaced particles (of compressed
// pointers), most frequent first (ties to the smaller root), each occurring at least twice; -1
// past the last. Also resets the candidates' boxes. One workgroup of kCcVote threads.
constexpr int kCcVote = 1024;
__global__ __launch_bounds__(kCcVote) void k_cc_vote(int n, const int* parent, CcSkip* S) {
  __shared__ int v[kCcVote];
  __shared__ int bestCount[kCcVote / 64], bestRoot[kCcVote / 64];
  const int t = threadIdx.x;
  const int64_t stride = n / kCcVote > 0 ? n / kCcVote : 1;
  const int64_t sIdx = (int64_t)t * stride;
  v[t] = sIdx < n ? cc_load_cached(parent, (int)sIdx) : INT_MAX;
  __syncthreads();
  for (int k = 2; k <= kCcVote; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int l = t ^ j;
      if (l > t) {
        const int a = v[t], b = v[l];
        if ((a > b) == ((t & k) == 0)) {
          v[t] = b;
          v[l] = a;
        }
      }
      __syncthreads();
    }
  // run length at each run's last element: its first element by binary search
  int cnt = 0;
  const int r = v[t];
  if (r != INT_MAX && (t == kCcVote - 1 || v[t + 1] != r)) {
    int lo = 0, hi = t;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (v[mid] < r) lo = mid + 1; else hi = mid;
    }
    cnt = t - lo + 1;
  }
  if (t < 6 * kCcSkipMax) (&S->box[0][0])[t] = (t % 6) < 3 ? 0xffffffffu : 0u;
  // the longest runs one at a time; ties to the smaller root
  __shared__ int winner;
  for (int k = 0; k < kCcSkipMax; ++k) {
    int bc = cnt, br = cnt > 0 ? r : INT_MAX;
    for (int off = 32; off > 0; off >>= 1) {
      const int oc = __shfl_down(bc, off, 64), orr = __shfl_down(br, off, 64);
      if (oc > bc || (oc == bc && orr < br)) {
        bc = oc;
        br = orr;
      }
    }
    if ((t & 63) == 0) {
      bestCount[t >> 6] = bc;
      bestRoot[t >> 6] = br;
    }
    __syncthreads();
    if (t == 0) {
      int c = bestCount[0], rr = bestRoot[0];
      for (int w = 1; w < kCcVote / 64; ++w)
        if (bestCount[w] > c || (bestCount[w] == c && bestRoot[w] < rr)) {
          c = bestCount[w];
          rr = bestRoot[w];
        }
      S->cand[k] = c >= 2 ? rr : -1;
      winner = c >= 2 ? rr : -1;
    }
    __syncthreads();
    if (cnt > 0 && r == winner) cnt = 0;
  }
}
// floats as unsigned integers of the same order (for atomic min / max)
__device__ __forceinline__ unsigned int float_order(float f) {
  const unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float order_float(unsigned int o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}
// bounding boxes of the voted candidates: a workgroup takes kBoxPerThread x kBlock consecutive
// particles of the bucket order, each wave reduces its particles of each candidate present in it
// into the workgroup's boxes in LDS, and the workgroup adds those to the global boxes
constexpr int kBoxPerThread = 16;
__global__ __launch_bounds__(kBlock) void k_cc_skip_boxes(int n, const float4* sorted, const int* snap, CcSkip* S) {
  __shared__ unsigned int lbox[kCcSkipMax][6];
  __shared__ int cand[kCcSkipMax];
  const int t = threadIdx.x, lane = t & 63;
  if (t < 6 * kCcSkipMax) (&lbox[0][0])[t] = (t % 6) < 3 ? 0xffffffffu : 0u;
  if (t < kCcSkipMax) cand[t] = S->cand[t];
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kBlock * kBoxPerThread;
  for (int it = 0; it < kBoxPerThread; ++it) {
    const int64_t ti = base + (int64_t)it * kBlock + t;
    const bool act = ti < n;
    const float4 me = act ? sorted[ti] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int r = act ? snap[__float_as_int(me.w)] : -1;
    int j = -1;
#pragma unroll
    for (int k = 0; k < kCcSkipMax; ++k)
      if (cand[k] >= 0 && cand[k] == r) j = k;
    unsigned long long todo = __ballot(j >= 0);
    while (todo) {
      const int l = __ffsll((long long)todo) - 1;
      const int lj = __shfl(j, l, 64);
      const bool mine = j == lj;
      const unsigned long long peers = __ballot(mine);
      float v[6] = {mine ? me.x : INFINITY, mine ? me.y : INFINITY, mine ? me.z : INFINITY,
                    mine ? me.x : -INFINITY, mine ? me.y : -INFINITY, mine ? me.z : -INFINITY};
      for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int a = 0; a < 6; ++a) {
          const float o = __shfl_xor(v[a], off, 64);
          v[a] = a < 3 ? fminf(v[a], o) : fmaxf(v[a], o);
        }
      if (lane == l) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          atomicMin(&lbox[lj][a], float_order(v[a]));
          atomicMax(&lbox[lj][a + 3], float_order(v[a + 3]));
        }
      }
      todo &= ~peers;
    }
  }
  __syncthreads();
  if (t < 6 * kCcSkipMax) {
    const int k = t / 6, a = t % 6;
    const unsigned int v = lbox[k][a];
    if (a < 3 && v != 0xffffffffu) atomicMin(&S->box[k][a], v);
    if (a >= 3 && v != 0u) atomicMax(&S->box[k][a], v);
  }
}
// the candidates skipped by pass 2: in vote order, each whose box, grown by the link distance
// (and 0.1% for the float distance test), meets no box already taken. One thread.
__global__ void k_cc_skip_select(float h, CcSkip* S) {
  const float m = 1.001f * h;
  float lo[kCcSkipMax][3], hi[kCcSkipMax][3];
  int ns = 0;
  for (int k = 0; k < kCcSkipMax; ++k) {
    const int r = S->cand[k];
    if (r < 0) break;
    float l[3], u[3];
    for (int a = 0; a < 3; ++a) {
      l[a] = order_float(S->box[k][a]);
      u[a] = order_float(S->box[k][a + 3]);
    }
    bool ok = true;
    for (int q = 0; q < ns && ok; ++q) {
      bool meet = true;
      for (int a = 0; a < 3; ++a) meet = meet && l[a] <= hi[q][a] + m && lo[q][a] <= u[a] + m;
      ok = !meet;
    }
    if (!ok) continue;
    for (int a = 0; a < 3; ++a) {
      lo[ns][a] = l[a];
      hi[ns][a] = u[a];
    }
    S->skip[ns++] = r;
  }
  S->nskip = ns;
}
// every slot points at its root; smallest caller index per component. The store is
// body-sorted, so a wave's slots usually share one root: reduce in registers first.
__global__ void k_cc_roots(int n, int* parent, const int* hasNb, const int* orig_of, int* isRoot, int* minOrig) {
  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += (int64_t)gridDim.x * kBlock) {
    int64_t k = base + threadIdx.x;
    int r = -1, o = INT_MAX;
    if (k < n) {
      int s = (int)k;
      int root = cc_find_cached(parent, s);
      parent[s] = root;
      if (hasNb[s]) {
        r = root;
        o = orig_of[s];
        if (root == s) isRoot[s] = 1;
      }
    }
    unsigned long long have = __ballot(r >= 0);
    if (have == 0ull) continue;
    int first = __ffsll((long long)have) - 1;
    int rr = __shfl(r, first, 64);
    // minOrig only decreases: an atomic that a plain read shows cannot lower it is skipped (a
    // stale read only costs an unneeded atomic), which keeps the thousands of waves of a large
    // body from queueing on its root's word
    if (__all((r == rr) || (r < 0))) {
      for (int off = 32; off > 0; off >>= 1) o = min(o, __shfl_down(o, off, 64));
      if ((threadIdx.x & 63) == 0 && o < __hip_atomic_load(minOrig + rr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMin(minOrig + rr, o);
    } else if (r >= 0) {
      if (o < __hip_atomic_load(minOrig + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(minOrig + r, o);
    }
  }
}
// Bodies are numbered by their smallest original particle index (distinct per body): mark those
// indices, and after an exclusive scan of the marks a body's number is its index's scan value.
__global__ void k_cc_mark(int n, const int* isRoot, const int* minOrig, int* mark) {
  GRID_STRIDE(s, n) if (isRoot[s]) mark[minOrig[s]] = 1;
}
__global__ void k_cc_assign(int n, const int* parent, const int* hasNb, const int* minOrig, const int* markPos,
                            int* labels) {
  GRID_STRIDE(s, n) labels[s] = hasNb[s] ? markPos[minOrig[parent[s]]] : -1;
}
// velocity finiteness (Sim.cpp:151-159)
__global__ void k_check_pvel(PView P, int n, int* err) {
  int bad = 0;
  GRID_STRIDE(si, n) {
    int s = (int)si;
    if (!isfinite(P.atu(C_V, s) * P.atu(C_V + 1, s) * P.atu(C_V + 2, s))) bad = 1;
  }
  if (bad) atomicOr(err, ERR_NAN_PVEL);
}

// ------------------------------------------------------------ body reductions
__device__ __forceinline__ unsigned int ord_enc(float f) {
  unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
// bbox (Grid.cpp:177-191) and comoving-frame sums (Sim.cpp:84-93) per body. Each block
// takes a contiguous range of the (body-sorted) list and every thread keeps running sums for
// its current body, flushing them when the body changes; blocks whose threads all end on one
// body combine in LDS and flush once.
struct BodyAcc {
  unsigned int mn[3], mx[3];
  double m[4];
  int cnt;
};
__device__ __forceinline__ void body_acc_reset(BodyAcc& a) {
  for (int q = 0; q < 3; ++q) {
    a.mn[q] = 0xFFFFFFFFu;
    a.mx[q] = 0u;
  }
  for (int q = 0; q < 4; ++q) a.m[q] = 0.0;
  a.cnt = 0;
}
__device__ __forceinline__ void body_acc_flush(const BodyAcc& a, int b, unsigned int* bbox, double* mom, int* count) {
  for (int q = 0; q < 3; ++q) {
    // skipped when a plain read shows the bound already covers this block's
    if (a.mn[q] < __hip_atomic_load(bbox + 6 * b + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMin(bbox + 6 * b + q, a.mn[q]);
    if (a.mx[q] > __hip_atomic_load(bbox + 6 * b + 3 + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(bbox + 6 * b + 3 + q, a.mx[q]);
  }
  for (int q = 0; q < 4; ++q) atomicAdd(mom + 4 * b + q, a.m[q]);
  atomicAdd(count + b, a.cnt);
}
// Each block adds into replica blockIdx % kBodyReplicas of the accumulators (bbox, mom, count
// point at replica 0, replica r at + r nb bodies): the hundreds of blocks of a large body would
// otherwise queue their atomics on one cache line. k_body_combine folds the replicas in order.
constexpr int kBodyReplicas = 16;
__global__ void k_body_reduce(PView P, const int* slots, const int* labels, int n, int nb, unsigned int* bboxR,
                              double* momR, int* countR) {
  const int64_t rep = (int64_t)(blockIdx.x % kBodyReplicas) * nb;
  unsigned int* bbox = bboxR + 6 * rep;
  double* mom = momR + 4 * rep;
  int* count = countR + rep;
  const int64_t per = ((int64_t)n + gridDim.x - 1) / gridDim.x;
  const int64_t beg = (int64_t)blockIdx.x * per;
  const int64_t end = beg + per < n ? beg + per : n;
  BodyAcc a;
  body_acc_reset(a);
  int cb = -1;
  for (int64_t k = beg + threadIdx.x; k < end; k += kBlock) {
    int s = slots ? slots[k] : (int)k;
    int b = labels ? labels[s] : 0;
    if (b < 0) continue;
    if (b != cb) {
      if (cb >= 0) body_acc_flush(a, cb, bbox, mom, count);
      body_acc_reset(a);
      cb = b;
    }
    float m = P.atu(C_M, s);
    for (int q = 0; q < 3; ++q) {
      unsigned int e = ord_enc(P.atu(C_X + q, s));
      a.mn[q] = min(a.mn[q], e);
      a.mx[q] = max(a.mx[q], e);
      a.m[q] += (double)(P.atu(C_V + q, s) * m);
    }
    a.m[3] += (double)m;
    a.cnt += 1;
  }
  __shared__ int blockBody;
  __shared__ int uniform;
  if (threadIdx.x == 0) {
    blockBody = -1;
    uniform = 1;
  }
  __syncthreads();
  if (cb >= 0) atomicCAS(&blockBody, -1, cb);
  __syncthreads();
  if (cb >= 0 && cb != blockBody) uniform = 0;
  __syncthreads();
  if (!uniform) {
    if (cb >= 0) body_acc_flush(a, cb, bbox, mom, count);
    return;
  }
  if (blockBody < 0) return;
  // one body in the block: combine in registers across the wave, then across waves in LDS
  for (int q = 0; q < 3; ++q)
    for (int off = 32; off > 0; off >>= 1) {
      a.mn[q] = min(a.mn[q], (unsigned int)__shfl_xor((int)a.mn[q], off, 64));
      a.mx[q] = max(a.mx[q], (unsigned int)__shfl_xor((int)a.mx[q], off, 64));
    }
  for (int q = 0; q < 4; ++q)
    for (int off = 32; off > 0; off >>= 1) a.m[q] += __shfl_xor(a.m[q], off, 64);
  for (int off = 32; off > 0; off >>= 1) a.cnt += __shfl_xor(a.cnt, off, 64);
  __shared__ BodyAcc wacc[kBlock / 64];
  if ((threadIdx.x & 63) == 0) wacc[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    BodyAcc t = wacc[0];
    for (int w = 1; w < kBlock / 64; ++w) {
      for (int q = 0; q < 3; ++q) {
        t.mn[q] = min(t.mn[q], wacc[w].mn[q]);
        t.mx[q] = max(t.mx[q], wacc[w].mx[q]);
      }
      for (int q = 0; q < 4; ++q) t.m[q] += wacc[w].m[q];
      t.cnt += wacc[w].cnt;
    }
    body_acc_flush(t, blockBody, bbox, mom, count);
  }
}

__global__ void k_body_init(int nrep, unsigned int* bboxR, double* momR, int* countR) {
  GRID_STRIDE(i, nrep) {
    for (int a = 0; a < 3; ++a) {
      bboxR[6 * i + a] = 0xFFFFFFFFu;
      bboxR[6 * i + 3 + a] = 0u;
    }
    for (int a = 0; a < 4; ++a) momR[4 * i + a] = 0.0;
    countR[i] = 0;
  }
}
__global__ void k_body_combine(int nb, const unsigned int* bboxR, const double* momR, const int* countR,
                               unsigned int* bbox, double* mom, int* count) {
  GRID_STRIDE(bi, nb) {
    const int b = (int)bi;
    unsigned int mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    int c = 0;
    for (int r = 0; r < kBodyReplicas; ++r) {
      const int64_t i = (int64_t)r * nb + b;
      for (int a = 0; a < 3; ++a) {
        mn[a] = min(mn[a], bboxR[6 * i + a]);
        mx[a] = max(mx[a], bboxR[6 * i + 3 + a]);
      }
      for (int a = 0; a < 4; ++a) m[a] += momR[4 * i + a];
      c += countR[i];
    }
    for (int a = 0; a < 3; ++a) {
      bbox[6 * b + a] = mn[a];
      bbox[6 * b + 3 + a] = mx[a];
    }
    for (int a = 0; a < 4; ++a) mom[4 * b + a] = m[a];
    count[b] = c;
  }
}

// ------------------------------------------------------------ binning
// order: the bodies by ascending tileBase (the batch's layout order)
__global__ void k_tile_body(const BodyGrid* bodies, const int* order, int nb, int NT, int* tileBody, int* blockBody) {
  GRID_STRIDE(ti, NT) {
    int t = (int)ti;
    int lo = 0, hi = nb - 1;
    while (lo < hi) {  // last body in layout order with tileBase <= t
      int mid = (lo + hi + 1) >> 1;
      if (bodies[order[mid]].tileBase <= t) lo = mid;
      else hi = mid - 1;
    }
    tileBody[t] = order[lo];
    if ((t & 3) == 0) blockBody[t >> 2] = order[lo];
  }
}
__global__ void k_bin(PView P, const int* slots, const int* labels, int n, const BodyGrid* bodies, int* key, int* rank,
                      int* count, int* err) {
  GRID_STRIDE(k, n) {
    int s = slots ? slots[k] : (int)k;
    int b = labels ? labels[s] : 0;
    if (b < 0) {
      key[k] = -1;
      continue;
    }
    const BodyGrid& G = bodies[b];
    float x[3];
    load_x(P, s, x);
    int t3[3], code = 0;
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float frac = (x[d] - G.gmin[d]) / G.h;
      int bc = (int)floorf(frac) - 1;
      int t = bc >> 2;
      ok = ok && bc >= 0 && t < G.T[d];
      t3[d] = t;
      code |= (bc & 3) << (2 * d);
    }
    if (!ok) {
      atomicOr(err, ERR_STENCIL);
      key[k] = -1;
      continue;
    }
    int t = G.tileBase + t3[0] + G.T[0] * (t3[1] + G.T[1] * t3[2]);
    key[k] = t | (code << kKeyCodeShift);  // the stencil base node within the tile, for k_sort_tiles
    rank[k] = wave_aggregated_inc(count, t);
  }
}
// a binning key back to the reference's stencil base node (i, j, k) in its body's grid:
// 4 x the tile's coordinates plus the base node's position within the tile; -1 for an unbinned key
__global__ void k_decode_bin(const int* key, int n, const BodyGrid* bodies, const int* tileBody, int* cells) {
  GRID_STRIDE(k, n) {
    const int kc = key[k];
    int c[3] = {-1, -1, -1};
    if (kc >= 0) {
      const int t = kc & ((1 << kKeyCodeShift) - 1);
      const int code = kc >> kKeyCodeShift;
      const BodyGrid& G = bodies[tileBody[t]];
      int tx, ty, tz;
      tile_xyz(G, t - G.tileBase, tx, ty, tz);
      c[0] = 4 * tx + (code & 3);
      c[1] = 4 * ty + ((code >> 2) & 3);
      c[2] = 4 * tz + ((code >> 4) & 3);
    }
    cells[3 * k] = c[0];
    cells[3 * k + 1] = c[1];
    cells[3 * k + 2] = c[2];
  }
}
__global__ void k_scatter(const int* slots, const int* key, const int* rank, const int* offset, int n, int* perm,
                          int8_t* pcode) {
  GRID_STRIDE(k, n) {
    const int kc = key[k];
    if (kc < 0) continue;
    const int t = kc & ((1 << kKeyCodeShift) - 1);
    const int pos = offset[t] + rank[k];
    perm[pos] = slots ? slots[k] : (int)k;
    pcode[pos] = (int8_t)(kc >> kKeyCodeShift);
  }
}
// Orders each tile's particles deterministically and so that particles sharing a stencil base
// node (their 64 scatter targets coincide, and same-address LDS atomics serialise) are spread
// over different waves: by (r, b), b the base node within the tile (0..63) and r the particle's
// rank by slot among those with base b. The binning's atomics leave a tile in arrival order,
// and its split into chunks of kChunk particles fixes each chunk's fixed-point scale, so a fixed
// order makes every staged footprint, and the whole step, bitwise reproducible. Positions come
// from counting (each key is unique): tiles up to kSortCap particles count in LDS, larger ones
// in global memory through tmpA and tmpB (the binning's per-particle key and rank arrays, dead
// by now).
constexpr int kSortCap = 4096;
__global__ __launch_bounds__(kBlock) void k_sort_tiles(const int* count, const int* offset, int* perm,
                                                       const int8_t* pcode, int* tmpA, int* tmpB) {
  __shared__ int sSlot[kSortCap], sKey[kSortCap];
  __shared__ int8_t sCode[kSortCap];
  const int t = blockIdx.x;
  const int c = count[t];
  if (c <= 1) return;
  int* seg = perm + offset[t];
  const int8_t* cseg = pcode + offset[t];
  if (c <= kSortCap) {
    for (int i = threadIdx.x; i < c; i += kBlock) {
      sSlot[i] = seg[i];
      sCode[i] = cseg[i];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < c; i += kBlock) {
      const int v = sSlot[i], b = sCode[i];
      int r = 0;
      for (int j = 0; j < c; ++j) r += (sCode[j] == b) & (sSlot[j] < v);
      sKey[i] = r * 64 + b;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < c; i += kBlock) {
      const int kk = sKey[i];
      int pos = 0;
      for (int j = 0; j < c; ++j) pos += sKey[j] < kk;
      seg[pos] = sSlot[i];
    }
  } else {
    int* key = tmpA + offset[t];
    int* pos = tmpB + offset[t];
    for (int i = threadIdx.x; i < c; i += kBlock) {
      const int v = seg[i], b = cseg[i];
      int r = 0;
      for (int j = 0; j < c; ++j) r += (cseg[j] == b) & (seg[j] < v);
      key[i] = r * 64 + b;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < c; i += kBlock) {
      const int kk = key[i];
      int p = 0;
      for (int j = 0; j < c; ++j) p += key[j] < kk;
      pos[i] = p;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < c; i += kBlock) key[pos[i]] = seg[i];  // keys are no longer read
    __syncthreads();
    for (int i = threadIdx.x; i < c; i += kBlock) seg[i] = key[i];
  }
}
__global__ void k_chunk_counts(const int* count, int NT, int* nch) {
  GRID_STRIDE(t, NT) nch[t] = (count[t] + kChunk - 1) / kChunk;
}
__global__ void k_make_chunks(const int* count, const int* offset, const int* chunk_off, const int* tileBody, int NT,
                              int4* chunks) {
  GRID_STRIDE(ti, NT) {
    int t = (int)ti;
    int c = count[t];
    int o = chunk_off[t];
    for (int q = 0; q * kChunk < c; ++q) {
      int cnt = c - q * kChunk;
      chunks[o + q] = make_int4(t, offset[t] + q * kChunk, cnt < kChunk ? cnt : kChunk, tileBody[t]);
    }
  }
}
// exclusive scan: 1024 elements per block
__global__ void k_scan_block(const int* in, int* out, int n, int* bsum) {
  __shared__ int sm[kBlock];
  int64_t b0 = (int64_t)blockIdx.x * 1024;
  int v[4];
  int tot = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int64_t i = b0 + threadIdx.x * 4 + q;
    v[q] = i < n ? in[i] : 0;
    tot += v[q];
  }
  sm[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    int a = threadIdx.x >= off ? sm[threadIdx.x - off] : 0;
    __syncthreads();
    sm[threadIdx.x] += a;
    __syncthreads();
  }
  int run = sm[threadIdx.x] - tot;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int64_t i = b0 + threadIdx.x * 4 + q;
    if (i < n) out[i] = run;
    run += v[q];
  }
  if (threadIdx.x == kBlock - 1) bsum[blockIdx.x] = sm[kBlock - 1];
}
__global__ void k_scan_add(int* out, int n, const int* boff) {
  int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  int add = boff[blockIdx.x];
#pragma unroll
  for (int q = 0; q < 4; ++q, i += kBlock)
    if (i < n) out[i] += add;
}
__global__ void k_copy_int(const int* src, int* dst) { *dst = *src; }

// ------------------------------------------------------------ dense <-> batch
template <class T> __global__ void k_to_dense(const T* src, int ncomp, T* dst, BodyGrid G) {
  int64_t nn = (int64_t)G.n[0] * G.n[1] * G.n[2];
  GRID_STRIDE(idx, nn) {
    int i = (int)(idx % G.n[0]);
    int64_t r = idx / G.n[0];
    int j = (int)(r % G.n[1]);
    int k = (int)(r / G.n[1]);
    int t = G.tileBase + (i >> 2) + G.T[0] * ((j >> 2) + G.T[1] * (k >> 2));
    int l = (i & 3) + 4 * (j & 3) + 16 * (k & 3);
    for (int c = 0; c < ncomp; ++c) {
      int64_t si = ncomp == 1 ? (int64_t)t * 64 + l : ((int64_t)t * 3 + c) * 64 + l;
      dst[idx * ncomp + c] = src[si];
    }
  }
}
template <class T> __global__ void k_from_dense(const T* src, int ncomp, T* dst, BodyGrid G) {
  int64_t nn = (int64_t)G.n[0] * G.n[1] * G.n[2];
  GRID_STRIDE(idx, nn) {
    int i = (int)(idx % G.n[0]);
    int64_t r = idx / G.n[0];
    int j = (int)(r % G.n[1]);
    int k = (int)(r / G.n[1]);
    int t = G.tileBase + (i >> 2) + G.T[0] * ((j >> 2) + G.T[1] * (k >> 2));
    int l = (i & 3) + 4 * (j & 3) + 16 * (k & 3);
    for (int c = 0; c < ncomp; ++c) {
      int64_t si = ncomp == 1 ? (int64_t)t * 64 + l : ((int64_t)t * 3 + c) * 64 + l;
      dst[si] = src[idx * ncomp + c];
    }
  }
}

// ------------------------------------------------------------ P2G
// MassSplatter (Grid.cpp:62-99): m_i += m_p * w_ip
__global__ __launch_bounds__(kBlock) void k_p2g_mass(PView P, Batch B, float* stage, int* err) {
  __shared__ FxFoot<1> acc;
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(B, ft, blockIdx.x);
  fx_zero(acc);
  const int k = threadIdx.x;
  bool have = k < c.ch.z;
  float m = 0.0f;
  Stencil st;
  int rx = 0, ry = 0, rz = 0;
  if (have) {
    int s = B.perm[c.ch.y + k];
    float x[3];
    load_x(P, s, x);
    m = P.atu(C_M, s);
    if (!chunk_stencil(c, x, false, st, rx, ry, rz)) {
      atomicOr(err, ERR_STENCIL);
      have = false;
    }
  }
  const FxScale sc = fx_scale(have ? fabsf(m) : 0.0f);  // w <= 1
  if (have) {
    const float ms = m * sc.S;
#pragma unroll
    for (int cz = 0; cz < 4; ++cz)
#pragma unroll
      for (int cy = 0; cy < 4; ++cy)
#pragma unroll
        for (int cx = 0; cx < 4; ++cx) {
          int idx = foot_index(rx + cx, ry + cy, rz + cz);
          fx_add(&acc.hi[0][idx], &acc.lo[0][idx], ms * st_w(st, cx, cy, cz));
        }
  }
  __syncthreads();
  fx_store(acc, sc, stage);
}
// mass gather, clamp negative masses, NaN check (Grid.cpp:225-241)
__global__ void k_gather_mass(Batch B, const float* stage, float* mass, int* err) {
  NodeCtx c = node_begin(B);
  float m = c.valid ? gather_node(B, c.G, stage, c.T, c.l, 0) : 0.0f;
  if (!isfinite(m)) atomicOr(err, ERR_NAN_MASS);
  mass[c.i] = m < 0.0f ? 0.0f : m;
}
// VelocitySplatter (Grid.cpp:102-150): v_i += w (m_p / m_i) (v_p - v_f) where m_i > 0
__global__ __launch_bounds__(kBlock) void k_p2g_vel(PView P, Batch B, const float* mass, float* stage, int* err) {
  __shared__ float sm[kFootLds];
  __shared__ FxFoot<3> acc;
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(B, ft, blockIdx.x);
  fx_zero(acc);
  __syncthreads();
  for (int j = threadIdx.x; j < kFoot; j += kBlock) {
    int gt = ft[j >> 6];
    sm[foot_lds(j)] = gt >= 0 ? mass[(int64_t)gt * 64 + (j & 63)] : 0.0f;
  }
  const int k = threadIdx.x;
  bool have = k < c.ch.z;
  float m = 0.0f, dv[3] = {0.0f, 0.0f, 0.0f};
  Stencil st;
  int rx = 0, ry = 0, rz = 0;
  if (have) {
    int s = B.perm[c.ch.y + k];
    float x[3];
    load_x(P, s, x);
    m = P.atu(C_M, s);
    dv[0] = P.atu(C_V, s) - c.G.fv[0];
    dv[1] = P.atu(C_V + 1, s) - c.G.fv[1];
    dv[2] = P.atu(C_V + 2, s) - c.G.fv[2];
    if (!chunk_stencil(c, x, false, st, rx, ry, rz)) {
      atomicOr(err, ERR_STENCIL);
      have = false;
    }
  }
  // w m_p / m_i <= 1 wherever m_i >= w m_p
  const FxScale sc = fx_scale(have ? fabsf(dv[0]) + fabsf(dv[1]) + fabsf(dv[2]) : 0.0f);
  if (have) {
#pragma unroll
    for (int cz = 0; cz < 4; ++cz)
#pragma unroll
      for (int cy = 0; cy < 4; ++cy)
#pragma unroll
        for (int cx = 0; cx < 4; ++cx) {
          int idx = foot_index(rx + cx, ry + cy, rz + cz);
          float M = sm[idx];
          if (M > 0.0f) {
            float f = st_w(st, cx, cy, cz) * (m / M);
            fx_add(&acc.hi[0][idx], &acc.lo[0][idx], (f * dv[0]) * sc.S);
            fx_add(&acc.hi[1][idx], &acc.lo[1][idx], (f * dv[1]) * sc.S);
            fx_add(&acc.hi[2][idx], &acc.lo[2][idx], (f * dv[2]) * sc.S);
          }
        }
  }
  __syncthreads();
  fx_store(acc, sc, stage);
}
// velocity gather, NaN check (Grid.cpp:250-260) and m_prevVelocities = m_velocities (Grid.cpp:262)
__global__ void k_gather_vel(Batch B, const float* stage, float* vel, float* prev, int* err) {
  NodeCtx c = node_begin(B);
  float v[3] = {0.0f, 0.0f, 0.0f};
  if (c.valid) gather_node3(B, c.G, stage, c.T, c.l, v);
  if (!isfinite(v[0]) || !isfinite(v[1]) || !isfinite(v[2])) atomicOr(err, ERR_NAN_VEL);
  st3(vel, c.i, v);
  st3(prev, c.i, v);
}
__global__ void k_gather3(Batch B, const float* stage, float* out) {
  NodeCtx c = node_begin(B);
  float v[3] = {0.0f, 0.0f, 0.0f};
  if (c.valid) gather_node3(B, c.G, stage, c.T, c.l, v);
  st3(out, c.i, v);
}
// computeParticleVolumes (Grid.cpp:265-292): V_p = m_p / sum_i w_ip m_i / h^3
__global__ __launch_bounds__(kBlock) void k_volumes(PView P, Batch B, const float* mass, int* err) {
  __shared__ float sm[kFootLds];
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(B, ft, blockIdx.x);
  __syncthreads();
  for (int j = threadIdx.x; j < kFoot; j += kBlock) {
    int gt = ft[j >> 6];
    sm[foot_lds(j)] = gt >= 0 ? mass[(int64_t)gt * 64 + (j & 63)] : 0.0f;
  }
  __syncthreads();
  const float cellVolume = c.G.h * c.G.h * c.G.h;
  for (int k = threadIdx.x; k < c.ch.z; k += kBlock) {
    int s = B.perm[c.ch.y + k];
    float x[3];
    load_x(P, s, x);
    Stencil st;
    int rx, ry, rz;
    if (!chunk_stencil(c, x, false, st, rx, ry, rz)) {
      atomicOr(err, ERR_STENCIL);
      continue;
    }
    float density = 0.0f;
#pragma unroll
    for (int cz = 0; cz < 4; ++cz)
#pragma unroll
      for (int cy = 0; cy < 4; ++cy)
#pragma unroll
        for (int cx = 0; cx < 4; ++cx)
          density += st_w(st, cx, cy, cz) * sm[foot_index(rx + cx, ry + cy, rz + cz)] / cellVolume;
    P.atu(C_VOL, s) = P.atu(C_M, s) / density;
  }
}

// ------------------------------------------------------------ forces
// ForceSplatter (Grid.cpp:387-435): f_i -= V_p P_p F_p^T grad w_ip
__global__ __launch_bounds__(kBlock) void k_forces(PView P, Batch B, float* stage, int* err) {
  __shared__ FxFoot<3> acc;
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(B, ft, blockIdx.x);
  fx_zero(acc);
  const int k = threadIdx.x;
  bool have = k < c.ch.z;
  float fm[9];
  Stencil st;
  int rx = 0, ry = 0, rz = 0;
  if (have) {
    int s = B.perm[c.ch.y + k];
    float x[3], F[9], R[9], Fi[9], Pm[9];
    load_x(P, s, x);
    load_m9(P, C_F, s, F);
    load_m9(P, C_R, s, R);
    load_m9(P, C_FI, s, Fi);
    snow_stress(F, R, Fi, P.atu(C_J, s), P.atu(C_MU, s), P.atu(C_LAM, s), Pm);
    float n2 = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) n2 += Pm[i] * Pm[i];
    if (!isfinite(sqrtf(n2))) atomicOr(err, ERR_NAN_STRESS);
    float vol = P.atu(C_VOL, s);
#pragma unroll
    for (int i = 0; i < 9; ++i) Pm[i] = vol * Pm[i];
    mat_mul_bt(Pm, F, fm);
    if (!chunk_stencil(c, x, true, st, rx, ry, rz)) {
      atomicOr(err, ERR_STENCIL);
      have = false;
    }
  }
  // |(fm grad w)_a| <= sum |fm| / h
  const FxScale sc = fx_scale(have ? abs_sum9(fm) / c.G.h : 0.0f);
  if (have) {
#pragma unroll
    for (int i = 0; i < 9; ++i) fm[i] = -(fm[i] * sc.S);
#pragma unroll
    for (int cz = 0; cz < 4; ++cz)
#pragma unroll
      for (int cy = 0; cy < 4; ++cy)
#pragma unroll
        for (int cx = 0; cx < 4; ++cx) {
          float gw[3], t[3];
          st_grad(st, cx, cy, cz, gw);
          mat_vec(fm, gw, t);
          int idx = foot_index(rx + cx, ry + cy, rz + cz);
          fx_add(&acc.hi[0][idx], &acc.lo[0][idx], t[0]);
          fx_add(&acc.hi[1][idx], &acc.lo[1][idx], t[1]);
          fx_add(&acc.hi[2][idx], &acc.lo[2][idx], t[2]);
        }
  }
  __syncthreads();
  fx_store(acc, sc, stage);
}

__device__ __forceinline__ void field_force(const float* grav, int ngrav, float m, float f[3]) {
  for (int q = 0; q < ngrav; ++q) {
    f[0] += grav[3 * q] * m;
    f[1] += grav[3 * q + 1] * m;
    f[2] += grav[3 * q + 2] * m;
  }
}

// calculateForces output (Grid.cpp:438-463): f = sum_fields g m + internal
__global__ void k_node_forces(Batch B, const float* mass, const float* stage, const float* grav, int ngrav,
                              float* out) {
  NodeCtx c = node_begin(B);
  float f[3] = {0.0f, 0.0f, 0.0f}, fi[3] = {0.0f, 0.0f, 0.0f};
  if (c.valid) {
    field_force(grav, ngrav, mass[c.i], f);
    gather_node3(B, c.G, stage, c.T, c.l, fi);
  }
  f[0] += fi[0];
  f[1] += fi[1];
  f[2] += fi[2];
  st3(out, c.i, f);
}

// calculateExplicitMomenta (Grid.cpp:634-690) + collisionVelocities (Grid.cpp:692-724)
// + the sums of the initial guess (Grid.cpp:911-947)
__global__ void k_node_explicit(Batch B, const float* mass, const float* vel, const float* stage, const float* grav,
                                int ngrav, float dt, const ColliderSet* cols, float* em, int8_t* flags, float* vc,
                                CRScalars* sc, BatchStatus* bst, double* partials, unsigned char* tileActive,
                                float* nrm) {
  NodeCtx c = node_begin(B);
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int bad = 0, active = 0;
  float e[3] = {0.0f, 0.0f, 0.0f}, cv[3] = {0.0f, 0.0f, 0.0f};
  int fl = -1;
  if (c.valid) {
    float m = mass[c.i];
    active = m > 0.0f ? 1 : 0;
    float f[3] = {0.0f, 0.0f, 0.0f}, fi[3];
    field_force(grav, ngrav, m, f);
    gather_node3(B, c.G, stage, c.T, c.l, fi);
    f[0] += fi[0];
    f[1] += fi[1];
    f[2] += fi[2];
    float ev[3];
    ld3(vel, c.i, ev);
    if (m > 0.0f) {
      ev[0] += dt * f[0] / m;
      ev[1] += dt * f[1] / m;
      ev[2] += dt * f[2] / m;
    }
    float x[3];
    node_pos(c.G, c.ni, c.nj, c.nk, x);
    fl = collide(*cols, ev, x, c.G.fv, false);
    e[0] = ev[0] * m;
    e[1] = ev[1] * m;
    e[2] = ev[2] * m;
    if (!isfinite(e[0] * e[1] * e[2])) bad = 1;
    if (fl >= 0) {
      float n[3];
      col_normal(cols->o[fl], x, n);
      st3(nrm, c.i, n);
      col_velocity(cols->o[fl], x, cv);
      cv[0] -= c.G.fv[0];
      cv[1] -= c.G.fv[1];
      cv[2] -= c.G.fv[2];
      s[4] = (double)m;
      s[5] = (double)(m * cv[0]);
      s[6] = (double)(m * cv[1]);
      s[7] = (double)(m * cv[2]);
    }
    s[0] = (double)m;
    s[1] = (double)e[0];
    s[2] = (double)e[1];
    s[3] = (double)e[2];
  }
  st3(em, c.i, e);
  st3(vc, c.i, cv);
  flags[c.i] = (int8_t)fl;
  // A tile outside every chunk's footprint (no particle in it or in its -x/-y/-z neighbour tiles)
  // and with no collider flag keeps x = guess and z = p = Ap = 0 through the whole CR loop: b,
  // A x and every staged HVP contribution are exactly zero there. The loop kernels skip its
  // loads and stores and add zeros in its place. (Zero mass alone is not enough: a tiny edge
  // weight can round to zero in the mass scatter and not in the HVP's.) A tile is one wave of a
  // node block.
  int src = 0;
  {
    int tx, ty, tz;
    tile_xyz(c.G, c.T - c.G.tileBase, tx, ty, tz);
#pragma unroll
    for (int f = 0; f < 8; ++f) {
      const int ts = tile_id(c.G, tx - (f & 1), ty - ((f >> 1) & 1), tz - (f >> 2));
      if (ts >= 0) src += B.tileChunkN[ts];
    }
  }
  const bool act = __ballot(src > 0 || fl != -1) != 0ull;
  if ((threadIdx.x & 63) == 0) tileActive[c.T] = act ? 1 : 0;
  CRScalars& S = sc[c.b];
  if (bad) atomicOr(&S.errors, ERR_NAN_MOMENTA);
  double sa[9] = {s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7], (double)active};
  double tot[9];
  if (body_sum<9>(sa, c.G, partials, body_tickets(B, c.b, 0), tot) && threadIdx.x == 0) {
    for (int k = 0; k < 8; ++k) S.sums[k] = tot[k];
    S.active_nodes = (int)tot[8];
    float gv[3] = {0.0f, 0.0f, 0.0f};
    if (tot[0] == 0.0) {
    } else if (tot[4] == 0.0) {
      for (int a = 0; a < 3; ++a) gv[a] = (float)(tot[1 + a] / tot[0]);
    } else {
      for (int a = 0; a < 3; ++a) gv[a] = (float)(tot[5 + a] / tot[4]);
    }
    S.guess[0] = gv[0];
    S.guess[1] = gv[1];
    S.guess[2] = gv[2];
  }
}

// the collider normals of nodes whose flags the caller set (Grid::m_nodeCollided written directly)
__global__ void k_node_normals(Batch B, const int8_t* flags, const ColliderSet* cols, float* nrm) {
  NodeCtx c = node_begin(B);
  const int fl = flags[c.i];
  if (fl < 0 || fl >= cols->n || !c.valid) return;
  float x[3], n[3];
  node_pos(c.G, c.ni, c.nj, c.nk, x);
  col_normal(cols->o[fl], x, n);
  st3(nrm, c.i, n);
}
__global__ void k_collision_velocities(Batch B, const int8_t* flags, const ColliderSet* cols, float* vc) {
  NodeCtx c = node_begin(B);
  int fl = flags[c.i];
  float v[3] = {0.0f, 0.0f, 0.0f};
  if (fl >= 0 && c.valid) {
    float x[3];
    node_pos(c.G, c.ni, c.nj, c.nk, x);
    col_velocity(cols->o[fl], x, v);
    for (int a = 0; a < 3; ++a) v[a] -= c.G.fv[a];
  }
  st3(vc, c.i, v);
}

// ------------------------------------------------------------ HVP
// ForceDifferentialSplatter (Grid.cpp:549-618), factorised:
//   dF_p = (sum_j d_j grad w_jp^T) F_p ; df_i -= V_p dP(dF_p) F_p^T grad w_ip
// One chunk's staged footprint of the HVP (LDS work areas passed in; every thread of the block calls).
// particle part: footprint of d in sd (loaded), scatter into acc (zeroed); ends on a barrier
// The slot and position of this thread's particle of chunk ch, loaded ahead of the footprint
// phase so that their latency overlaps it (thread k holds particle k of the chunk, if any).
struct HvpPre {
  int s;
  float x[3];
};
__device__ __forceinline__ HvpPre hvp_prefetch(const PView& P, const Batch& B, int4 ch) {
  HvpPre q;
  q.s = -1;
  q.x[0] = q.x[1] = q.x[2] = 0.0f;
  if ((int)threadIdx.x < ch.z) {
    q.s = B.perm[ch.y + threadIdx.x];
    load_x(P, q.s, q.x);
  }
  return q;
}
// One particle's part of the HVP: stencil, gathered gradient G of the footprint sd, and
// fm = V dP(G F) F^T (the 3x3 map whose rows scatter against the gradient weights); have is
// cleared when the particle's stencil leaves the chunk's footprint.
__device__ __forceinline__ void hvp_particle(const PView& P, const ChunkCtx& c, float (*sd)[kFootLds], int* err,
                                             const HvpPre& pre, bool& have, Stencil& st, int& base, float fm[9]) {
  if (have) {
    const int s = pre.s;
    const float x[3] = {pre.x[0], pre.x[1], pre.x[2]};
    make_stencil_fast(x, c.G.gmin, c.G.h, 1.0f / c.G.h, st);
    const int rx = st.base[0] - c.ox, ry = st.base[1] - c.oy, rz = st.base[2] - c.oz;
    if (!(rx >= 0 && rx < 4 && ry >= 0 && ry < 4 && rz >= 0 && rz < 4)) {
      atomicOr(err, ERR_STENCIL);
      have = false;
    } else {
      base = foot_index(rx, ry, rz);
      // G[a + 3 c] = sum_i d_a(i) grad_c w_i = -sum_i d_a(i) ghat_c(i)
      float G[9];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        float g[3];
        sep_gather(foot_at(sd, a, base), st, g);
        G[a] = -g[0];
        G[a + 3] = -g[1];
        G[a + 6] = -g[2];
      }
      float F[9], R[9], Fi[9], Gi[9], dF[9], dP[9];
      load_m9(P, C_F, s, F);
      load_m9(P, C_R, s, R);
      load_m9(P, C_FI, s, Fi);
      load_m9(P, C_GI, s, Gi);
      mat_mul(G, F, dF);
      snow_stress_differential(dF, R, Fi, Gi, P.atu(C_J, s), P.atu(C_MU, s), P.atu(C_LAM, s), dP);
      float vol = P.atu(C_VOL, s);
#pragma unroll
      for (int i = 0; i < 9; ++i) dP[i] = vol * dP[i];
      mat_mul_bt(dP, F, fm);
    }
  }
}
// per-component bound on the particle's scatter contributions (the fixed-point scale's input)
__device__ __forceinline__ void hvp_bound(const ChunkCtx& c, bool have, const float fm[9], float bound[3]) {
  bound[0] = bound[1] = bound[2] = 0.0f;
  if (have) {
    const float gmax = kGradWeightMax / c.G.h;
#pragma unroll
    for (int a = 0; a < 3; ++a) bound[a] = (fabsf(fm[a]) + fabsf(fm[a + 3]) + fabsf(fm[a + 6])) * gmax;
  }
}
__device__ __forceinline__ FxScale3 hvp_particles(const PView& P, const Batch& B, const ChunkCtx& c,
                                                  float (*sd)[kFootLds], FxFoot1<3>& acc, int* err,
                                                  const HvpPre& pre) {
  bool have = (int)threadIdx.x < c.ch.z;
  float fm[9];
  Stencil st;
  int base = 0;
  hvp_particle(P, c, sd, err, pre, have, st, base, fm);
  // df_a(i) -= sum_c fm[a + 3 c] grad_c w_i = sum_c fm[a + 3 c] ghat_c(i)
  float bound[3];
  hvp_bound(c, have, fm, bound);
  const FxScale3 scl = fx_scale3(bound, kFx1Top);
  if (have) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
      sep_scatter(&acc.v[a][base], st, fm[a] * scl.S[a], fm[a + 3] * scl.S[a], fm[a + 6] * scl.S[a]);
  }
  __syncthreads();
  return scl;
}
// LDS work area of hvp_particles_split: each plane's gather terms, and the scaled rows of fm
struct HvpSplitLds {
  float part[4][9][64];
  float fms[9][64];
};
// hvp_particles for a chunk of at most 64 particles (every thread calls): particle k is taken by
// thread k of each of the four waves, wave w doing z-plane w of its stencil's gather and scatter,
// and the first wave the rest, so that a particle's serial work is about a third. The gather
// planes are combined in sep_gather's order and the scatter values are the same expressions, so
// the staged footprint is bitwise that of hvp_particles.
// pre: this lane's particle (slot, position) when the caller holds it (the tiny solver keeps
// them across iterations), else loaded here. The first wave requests the particle's matrices
// before the gather, so that their latency overlaps it.
__device__ __forceinline__ FxScale3 hvp_particles_split(const PView& P, const Batch& B, const ChunkCtx& c,
                                                        float (*sd)[kFootLds], FxFoot1<3>& acc, int* err,
                                                        HvpSplitLds& W, const HvpPre* pre = nullptr) {
  const int lane = threadIdx.x & 63, plane = threadIdx.x >> 6;
  bool have = lane < c.ch.z;
  int s = -1;
  float x[3] = {0.0f, 0.0f, 0.0f};
  if (have) {
    if (pre) {
      s = pre->s;
      x[0] = pre->x[0];
      x[1] = pre->x[1];
      x[2] = pre->x[2];
    } else {
      s = B.perm[c.ch.y + lane];
      load_x(P, s, x);
    }
  }
  float F[9], R[9], Fi[9], Gi[9], Jp = 0.0f, mu = 0.0f, lam = 0.0f, vol = 0.0f;
  if (have && plane == 0) {
    load_m9(P, C_F, s, F);
    load_m9(P, C_R, s, R);
    load_m9(P, C_FI, s, Fi);
    load_m9(P, C_GI, s, Gi);
    Jp = P.atu(C_J, s);
    mu = P.atu(C_MU, s);
    lam = P.atu(C_LAM, s);
    vol = P.atu(C_VOL, s);
  }
  Stencil st;
  int base = 0;
  if (have) {
    make_stencil_fast(x, c.G.gmin, c.G.h, 1.0f / c.G.h, st);
    const int rx = st.base[0] - c.ox, ry = st.base[1] - c.oy, rz = st.base[2] - c.oz;
    if (!(rx >= 0 && rx < 4 && ry >= 0 && ry < 4 && rz >= 0 && rz < 4)) {
      if (plane == 0) atomicOr(err, ERR_STENCIL);
      have = false;
    } else {
      base = foot_index(rx, ry, rz);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        float e[3];
        sep_gather_plane(foot_at(sd, a, base), st, plane, e);
        W.part[plane][3 * a][lane] = e[0];
        W.part[plane][3 * a + 1][lane] = e[1];
        W.part[plane][3 * a + 2][lane] = e[2];
      }
    }
  }
  __syncthreads();
  const bool first = have && plane == 0;
  float fm[9];
  if (first) {
    float G[9];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float o0 = 0.0f, o1 = 0.0f, o2 = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        o0 = fmaf(W.part[k][3 * a][lane], st.w[2][k], o0);
        o1 = fmaf(W.part[k][3 * a + 1][lane], st.w[2][k], o1);
        o2 = fmaf(W.part[k][3 * a + 2][lane], st.dw[2][k], o2);
      }
      G[a] = -o0;
      G[a + 3] = -o1;
      G[a + 6] = -o2;
    }
    float dF[9], dP[9];
    mat_mul(G, F, dF);
    snow_stress_differential(dF, R, Fi, Gi, Jp, mu, lam, dP);
#pragma unroll
    for (int i = 0; i < 9; ++i) dP[i] = vol * dP[i];
    mat_mul_bt(dP, F, fm);
  }
  float bound[3];
  hvp_bound(c, first, fm, bound);
  const FxScale3 scl = fx_scale3(bound, kFx1Top);
  if (first) {
#pragma unroll
    for (int q = 0; q < 9; ++q) W.fms[q][lane] = fm[q] * scl.S[q % 3];
  }
  __syncthreads();
  if (have) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
      sep_scatter_plane(&acc.v[a][base], st, plane, W.fms[a][lane], W.fms[a + 3][lane], W.fms[a + 6][lane]);
  }
  __syncthreads();
  return scl;
}
// ch = B.chunks[chunk], pre = hvp_prefetch(P, B, ch)
// split: the small-body solver's work area; a chunk of at most 64 particles then runs
// hvp_particles_split (same staged bits)
template <bool PROJ>
__device__ __forceinline__ void hvp_chunk(const PView& P, const Batch& B, int chunk, int4 ch, const HvpPre& pre,
                                          const float* d, const int8_t* flags, const ColliderSet* cols, float* stage,
                                          int* err, float (*sd)[kFootLds], FxFoot1<3>& acc, int* ft,
                                          HvpSplitLds* split = nullptr) {
  ChunkCtx c = chunk_begin_at(B, ft, ch);
  fx_zero(acc);
  __syncthreads();
  load_foot3(c, ft, d, sd, flags, cols, PROJ);
  __syncthreads();
  const FxScale3 scl = (split && ch.z <= 64) ? hvp_particles_split(P, B, c, sd, acc, err, *split)
                                             : hvp_particles(P, B, c, sd, acc, err, pre);
  fx_store(acc, scl, stage, chunk);
}
template <bool PROJ>
__device__ __forceinline__ void hvp_chunk(const PView& P, const Batch& B, int chunk, const float* d,
                                          const int8_t* flags, const ColliderSet* cols, float* stage, int* err,
                                          float (*sd)[kFootLds], FxFoot1<3>& acc, int* ft,
                                          HvpSplitLds* split = nullptr) {
  const int4 ch = B.chunks[chunk];
  hvp_chunk<PROJ>(P, B, chunk, ch, split && ch.z <= 64 ? HvpPre{} : hvp_prefetch(P, B, ch), d, flags, cols, stage,
                  err, sd, acc, ft, split);
}
// The chunk's particle slots and positions are requested first, beside the done flag of its
// body: the chain of dependent loads before the stencil work is the chunk, then the slot, body
// and flag together, then position and footprint together.
template <bool PROJ>
__global__ __launch_bounds__(kBlock) void k_hvp(PView P, Batch B, const float* d, const int8_t* flags,
                                                const ColliderSet* cols, float* stage, const CRScalars* sc, int* err) {
  __shared__ float sd[3][kFootLds];
  __shared__ FxFoot1<3> acc;
  __shared__ int ft[8];
  const int4 ch = B.chunks[blockIdx.x];
  const HvpPre pre = hvp_prefetch(P, B, ch);
  if (sc && sc[ch.w].done) return;
  hvp_chunk<PROJ>(P, B, blockIdx.x, ch, pre, d, flags, cols, stage, err, sd, acc, ft);
}

// Paired HVP: a workgroup takes two consecutive chunks whose particles fit in one (at most
// kChunk together), chunk A's particles on threads 0..nA-1 and chunk B's on nA..nA+nB-1, each
// scattering into its own footprint. Sparse chunks (surface tiles, fragments) otherwise run
// their particles in mostly idle waves: the kernel is bound by VALU issue. Each chunk's
// fixed-point scale is computed from its particles' bounds placed at their index within the
// chunk, in the same reduction tree as a chunk of its own, so every staged footprint is bitwise
// that of k_hvp (and of the small-body solver's per-chunk HVP).
// groups[i] = (chunk A, chunk B or -1); blocks from *ngroups on exit.
__global__ void k_pair_chunks(const int4* chunks, int nch, int* cnt) {
  GRID_STRIDE(i, (nch + 1) / 2) {
    const int a = 2 * (int)i, b = a + 1;
    cnt[i] = (b < nch && chunks[a].z + chunks[b].z <= kChunk) || b >= nch ? 1 : 2;
  }
}
__global__ void k_pair_groups(const int4* chunks, int nch, const int* cnt, const int* pos, int2* groups) {
  GRID_STRIDE(i, (nch + 1) / 2) {
    const int a = 2 * (int)i, b = a + 1, o = pos[i];
    if (cnt[i] == 1) {
      groups[o] = make_int2(a, b < nch ? b : -1);
    } else {
      groups[o] = make_int2(a, -1);
      groups[o + 1] = make_int2(b, -1);
    }
  }
}
template <bool PROJ>
__global__ __launch_bounds__(kBlock) void k_hvp_pair(PView P, Batch B, const int2* groups, const int* ngroups,
                                                     const float* d, const int8_t* flags, const ColliderSet* cols,
                                                     float* stage, const CRScalars* sc, int* err) {
  if ((int)blockIdx.x >= *ngroups) return;
  __shared__ float sd[2][3][kFootLds];
  __shared__ FxFoot1<3> acc[2];
  __shared__ int ft[2][8];
  __shared__ float bnd[2][3][kBlock];
  const int2 g = groups[blockIdx.x];
  const int4 chA = B.chunks[g.x];
  const bool pair = g.y >= 0;
  const int4 chB = pair ? B.chunks[g.y] : make_int4(chA.x, chA.y, 0, chA.w);
  // the threads' particles: chunk A on 0..nA-1, chunk B on nA..nA+nB-1
  const int k = threadIdx.x;
  const int cid = k < chA.z ? 0 : 1;
  const int kk = cid ? k - chA.z : k;
  HvpPre pre;
  pre.s = -1;
  pre.x[0] = pre.x[1] = pre.x[2] = 0.0f;
  const int4 chM = cid ? chB : chA;
  if (kk < chM.z) {
    pre.s = B.perm[chM.y + kk];
    load_x(P, pre.s, pre.x);
  }
  const bool doneA = sc && sc[chA.w].done;
  const bool doneB = !pair || (sc && sc[chB.w].done);
  if (doneA && doneB) return;
  const ChunkCtx cA = chunk_begin_at(B, ft[0], chA);
  const ChunkCtx cB = chunk_begin_at(B, ft[1], chB);
  fx_zero(acc[0]);
  if (pair) fx_zero(acc[1]);
  __syncthreads();
  if (!doneA) load_foot3(cA, ft[0], d, sd[0], flags, cols, PROJ);
  if (!doneB) load_foot3(cB, ft[1], d, sd[1], flags, cols, PROJ);
  __syncthreads();
  const ChunkCtx& c = cid ? cB : cA;
  bool have = kk < chM.z && !(cid ? doneB : doneA);
  float fm[9];
  Stencil st;
  int base = 0;
  hvp_particle(P, c, sd[cid], err, pre, have, st, base, fm);
  float bound[3];
  hvp_bound(c, have, fm, bound);
  // bounds at each particle's index within its chunk; the rest zero
  for (int i = k; i < 2 * 3 * kBlock; i += kBlock) (&bnd[0][0][0])[i] = 0.0f;
  __syncthreads();
  if (have)
#pragma unroll
    for (int a = 0; a < 3; ++a) bnd[cid][a][kk] = bound[a];
  __syncthreads();
  float bA[3] = {bnd[0][0][k], bnd[0][1][k], bnd[0][2][k]};
  const FxScale3 sclA = fx_scale3(bA, kFx1Top);
  __syncthreads();  // fx_scale3's reduction buffer is reused
  float bB[3] = {bnd[1][0][k], bnd[1][1][k], bnd[1][2][k]};
  const FxScale3 sclB = fx_scale3(bB, kFx1Top);
  if (have) {
    const FxScale3& scl = cid ? sclB : sclA;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      sep_scatter(&acc[cid].v[a][base], st, fm[a] * scl.S[a], fm[a + 3] * scl.S[a], fm[a + 6] * scl.S[a]);
  }
  __syncthreads();
  if (!doneA) fx_store(acc[0], sclA, stage, g.x);
  if (!doneB) fx_store(acc[1], sclB, stage, g.y);
}

// dFidXiSplatter (Grid.cpp:466-537) through per-particle quadratic forms: with g = F^T grad w,
// the reference's entry [-V dP(e_a g^T) F^T grad w]_a equals -V g^T H_a g where
// H_a(c, d) = dP(e_a e_d^T)(a, c) (dP is linear in dF).
__global__ __launch_bounds__(kBlock) void k_diag(PView P, Batch B, float* stage, int* err) {
  __shared__ FxFoot<3> acc;
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(B, ft, blockIdx.x);
  fx_zero(acc);
  const int k = threadIdx.x;
  bool have = k < c.ch.z;
  float F[9], H[3][9];  // H[a][c + 3 d]
  float bound = 0.0f;
  Stencil st;
  int rx = 0, ry = 0, rz = 0;
  if (have) {
    int s = B.perm[c.ch.y + k];
    float x[3];
    load_x(P, s, x);
    float R[9], Fi[9], Gi[9];
    load_m9(P, C_F, s, F);
    load_m9(P, C_R, s, R);
    load_m9(P, C_FI, s, Fi);
    load_m9(P, C_GI, s, Gi);
    float J = P.atu(C_J, s), mu = P.atu(C_MU, s), lam = P.atu(C_LAM, s), vol = P.atu(C_VOL, s);
    float hs = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int dd = 0; dd < 3; ++dd) {
        float E[9], dP[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) E[i] = 0.0f;
        E[a + 3 * dd] = 1.0f;
        snow_stress_differential(E, R, Fi, Gi, J, mu, lam, dP);
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          H[a][cc + 3 * dd] = -vol * dP[a + 3 * cc];
          hs += fabsf(H[a][cc + 3 * dd]);
        }
      }
    // |g|_inf <= sum |F| / h, so |g^T H_a g| <= sum |H| (sum |F| / h)^2
    const float gb = abs_sum9(F) / c.G.h;
    bound = hs * gb * gb;
    if (!chunk_stencil(c, x, true, st, rx, ry, rz)) {
      atomicOr(err, ERR_STENCIL);
      have = false;
    }
  }
  const FxScale sc = fx_scale(have ? bound : 0.0f);
  if (have) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int i = 0; i < 9; ++i) H[a][i] *= sc.S;
#pragma unroll
    for (int cz = 0; cz < 4; ++cz)
#pragma unroll
      for (int cy = 0; cy < 4; ++cy)
#pragma unroll
        for (int cx = 0; cx < 4; ++cx) {
          float gw[3], gg[3];
          st_grad(st, cx, cy, cz, gw);
          mat_tvec(F, gw, gg);
          int idx = foot_index(rx + cx, ry + cy, rz + cz);
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            float q = 0.0f;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
              q += gg[cc] * (H[a][cc] * gg[0] + H[a][cc + 3] * gg[1] + H[a][cc + 6] * gg[2]);
            fx_add(&acc.hi[a][idx], &acc.lo[a][idx], q);
          }
        }
  }
  __syncthreads();
  fx_store(acc, sc, stage);
}

// ------------------------------------------------------------ solver node kernels
// rhs (Grid.cpp:982-1002), initial guess fill (Grid.cpp:944-947), diagonal (Grid.cpp:847-866), |b|^2
__global__ void k_rhs(Batch B, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* stageVc,
                      const float* stageDiag, float dt, float tol, float* b, float* x, float* diag, CRScalars* sc,
                      BatchStatus* bst, double* partials) {
  NodeCtx c = node_begin(B);
  CRScalars& S = sc[c.b];
  double s[1] = {0.0};
  int bad = 0;
  const float dt2 = dt * dt;
  const float ndt2 = -dt * dt;
  float bb[3], gx[3] = {S.guess[0], S.guess[1], S.guess[2]}, dr[3] = {0.0f, 0.0f, 0.0f};
  ld3(b, c.i, bb);
  float m = mass[c.i];
  if (c.valid) {
    float df[3];
    gather_node3(B, c.G, stageVc, c.T, c.l, df);
    node_project(c, cols, flags[c.i], df);
    bb[0] += dt2 * df[0];
    bb[1] += dt2 * df[1];
    bb[2] += dt2 * df[2];
    if (!isfinite(bb[0] * bb[1] * bb[2])) bad = 1;
    gather_node3(B, c.G, stageDiag, c.T, c.l, dr);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    dr[a] = dr[a] * ndt2 + m;
    if (dr[a] == 0.0f) dr[a] = 1.0f;
  }
  st3(b, c.i, bb);
  st3(x, c.i, gx);
  st3(diag, c.i, dr);
  s[0] = (double)bb[0] * bb[0] + (double)bb[1] * bb[1] + (double)bb[2] * bb[2];
  if (bad) atomicOr(&S.errors, ERR_NAN_MOMENTA);
  double tot[1];
  if (body_sum<1>(s, c.G, partials, body_tickets(B, c.b, 1), tot) && threadIdx.x == 0) {
    S.bnorm2 = tot[0];
    S.threshold = (double)tol * (double)tol * tot[0];
    S.iter = 0;
    S.iterations = 0;
    S.early_exit = 0;
    S.zero_x = 0;
    if (S.errors & ERR_NAN_MOMENTA) {
      atomicOr(&bst->errors, ERR_NAN_MOMENTA);
      set_done(S, bst);
    } else if (tot[0] == 0.0) {
      S.zero_x = 1;
      set_done(S, bst);
    }
  }
}

// The per-node work and the scalar updates of each CR kernel are device functions shared with
// the one-workgroup solver of small bodies (k_small_cr), so that both compute the same bits.

// CR setup part 1 (ConjugateResiduals.cpp:50-67): r = b - A x ; z = M^-1 r ; p = z
__device__ __forceinline__ void cr_setup1_core(const NodeCtx& c, const ColliderSet* cols, int fl, float m,
                                               const float xv[3], const float df[3], const float bb[3],
                                               const float dg[3], float dt2, float zz[3]) {
  float Ax[3];
  apply_A(c, cols, fl, m, xv, df, dt2, Ax);
#pragma unroll
  for (int a = 0; a < 3; ++a) zz[a] = (bb[a] - Ax[a]) / dg[a];
}
__device__ __forceinline__ void cr_setup1_node(const Batch& B, const NodeCtx& c, const float* mass,
                                               const int8_t* flags, const ColliderSet* cols, const float* stage,
                                               const float* x, const float* b, const float* diag, float dt2, float* z,
                                               float* p) {
  float xv[3], df[3] = {0.0f, 0.0f, 0.0f}, bb[3], dg[3], zz[3];
  ld3(x, c.i, xv);
  if (c.valid) gather_node3(B, c.G, stage, c.T, c.l, df);
  ld3(b, c.i, bb);
  ld3(diag, c.i, dg);
  cr_setup1_core(c, cols, flags[c.i], mass[c.i], xv, df, bb, dg, dt2, zz);
  st3(z, c.i, zz);
  st3(p, c.i, zz);
}
__global__ void k_cr_setup1(Batch B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                            const float* stage, const float* x, const float* b, const float* diag, float dt, float* z,
                            float* p, const CRScalars* sc) {
  NodeCtx c = node_begin(B);
  if (sc[c.b].done) return;
  cr_setup1_node(B, c, mass, flags, cols, stage, x, b, diag, dt * dt, z, p);
}
// CR setup part 2 (ConjugateResiduals.cpp:69-76 and the first alpha, :86-102):
// Ap = Ar = A z ; rz = <z, Ar> ; alpha = rz / <Ap, M^-1 Ap>
__device__ __forceinline__ void cr_setup2_core(const NodeCtx& c, const ColliderSet* cols, int fl, float m,
                                               const float zz[3], const float df[3], const float dg[3], float dt2,
                                               float Ar[3], double (&s)[2]) {
  apply_A(c, cols, fl, m, zz, df, dt2, Ar);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    s[0] += (double)zz[a] * Ar[a];
    s[1] += (double)Ar[a] * (Ar[a] / dg[a]);
  }
}
__device__ __forceinline__ void cr_setup2_node(const Batch& B, const NodeCtx& c, const float* mass,
                                               const int8_t* flags, const ColliderSet* cols, const float* stage,
                                               const float* z, const float* diag, float dt2, float* Ap,
                                               double (&s)[2]) {
  float zz[3], df[3] = {0.0f, 0.0f, 0.0f}, Ar[3], dg[3];
  ld3(z, c.i, zz);
  if (c.valid) gather_node3(B, c.G, stage, c.T, c.l, df);
  ld3(diag, c.i, dg);
  cr_setup2_core(c, cols, flags[c.i], mass[c.i], zz, df, dg, dt2, Ar, s);
  st3(Ap, c.i, Ar);
}
// returns true when the solve ends here
__device__ __forceinline__ bool cr_setup2_finish(CRScalars& S, const double (&tot)[2]) {
  S.rz = tot[0];
  S.ApdPAp = tot[1];
  // S.iter is 0 here from the rhs pass, or the flexible CR's count after a handoff
  if (tot[1] == 0.0) {
    S.early_exit = 1;
    return true;
  }
  S.alpha = tot[0] / tot[1];
  return false;
}
__global__ void k_cr_setup2(Batch B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                            const float* stage, const float* z, const float* diag, float dt, float* Ap, CRScalars* sc,
                            BatchStatus* bst, double* partials) {
  NodeCtx c = node_begin(B);
  CRScalars& S = sc[c.b];
  if (S.done) return;
  double s[2] = {0.0, 0.0};
  cr_setup2_node(B, c, mass, flags, cols, stage, z, diag, dt * dt, Ap, s);
  double tot[2];
  if (body_sum<2>(s, c.G, partials, body_tickets(B, c.b, 2), tot) && threadIdx.x == 0)
    if (cr_setup2_finish(S, tot)) set_done(S, bst);
}
// iteration part 1 (ConjugateResiduals.cpp:104-128): x = P(x + alpha p) ; z -= alpha M^-1 Ap ;
// r = M z ; termination on |r|^2 (SquareMagnitudeTermination.cpp:19-28)
__device__ __forceinline__ void cr_update_x_core(const NodeCtx& c, const ColliderSet* cols, int fl, float alpha,
                                                 float xv[3], const float pv[3], const float ap[3],
                                                 const float dg[3], float zz[3], double (&s)[1]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) xv[a] = xv[a] + alpha * pv[a];
  node_project(c, cols, fl, xv);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    zz[a] = zz[a] - alpha * (ap[a] / dg[a]);
    float r = zz[a] * dg[a];
    s[0] += (double)r * r;
  }
}
__device__ __forceinline__ void cr_update_x_node(const NodeCtx& c, const int8_t* flags, const ColliderSet* cols,
                                                 const float* p, const float* Ap, const float* diag, float alpha,
                                                 float* x, float* z, double (&s)[1]) {
  float xv[3], pv[3], ap[3], dg[3], zz[3];
  ld3(x, c.i, xv);
  ld3(p, c.i, pv);
  ld3(Ap, c.i, ap);
  ld3(diag, c.i, dg);
  ld3(z, c.i, zz);
  cr_update_x_core(c, cols, flags[c.i], alpha, xv, pv, ap, dg, zz, s);
  st3(x, c.i, xv);
  st3(z, c.i, zz);
}
template <int NV>
__device__ __forceinline__ bool cr_update_x_finish(CRScalars& S, const double (&tot)[NV], int body, double* hist) {
  int it = S.iter;
  S.rnorm2 = tot[0];
  S.iterations = it + 1;
  // after a p / Ap update (k_cr_update_px) the pass also summed <Ap, M^-1 Ap> of the Ap it
  // stored: the next denominator's expansion starts from the stored vector's own value
  if constexpr (NV > 1) S.ApdPAp = tot[1];
  if (hist && it < kHistCap) hist[(int64_t)body * kHistCap + it] = tot[0];
  return it >= S.max_iters || tot[0] < S.threshold;
}
__global__ void k_cr_update_x(Batch B, const int8_t* flags, const ColliderSet* cols, const float* p, const float* Ap,
                              const float* diag, float* x, float* z, CRScalars* sc, BatchStatus* bst, double* partials,
                              double* hist) {
  NodeCtx c = node_begin(B);
  CRScalars& S = sc[c.b];
  if (S.done) return;
  double s[1] = {0.0};
  if (!B.tileActive || B.tileActive[c.T])
    cr_update_x_node(c, flags, cols, p, Ap, diag, (float)S.alpha, x, z, s);
  double tot[1];
  if (body_sum<1>(s, c.G, partials, body_tickets(B, c.b, 0), tot) && threadIdx.x == 0)
    if (cr_update_x_finish<1>(S, tot, c.b, hist)) set_done(S, bst);
}
// iteration part 2 (ConjugateResiduals.cpp:130-145): Ar = A z ; rz = <z, Ar> ; beta. The
// next iteration's denominator <Ap', M^-1 Ap'> with Ap' = Ar + beta Ap (:147-151, :86-102) is
// taken from the same pass by expanding it: <Ar, M^-1 Ar> + 2 beta <Ar, M^-1 Ap> + beta^2
// <Ap, M^-1 Ap>, the last term summed over the stored Ap by the pass that wrote it. The p / Ap
// update then runs in the next x update's pass (k_cr_update_px) instead of a pass of its own.
// The expansion is a difference of nearly equal terms when Ar is nearly parallel to -beta Ap
// (a stalling or indefinite solve); summing its last term over the Ap actually stored keeps it
// in step with the direct form. If it still comes out below kCrDenKeep of the terms' size
// (rounding noise, or not positive), the recurrence restarts with beta = 0 (p = z, Ap = Ar),
// whose denominator <Ar, M^-1 Ar> is summed directly.
constexpr double kCrDenKeep = 1e-10;
__device__ __forceinline__ void cr_ar_core(const NodeCtx& c, const ColliderSet* cols, int fl, float m,
                                           const float zz[3], const float df[3], const float ap[3],
                                           const float dg[3], float dt2, float ar[3], double (&s)[3]) {
  apply_A(c, cols, fl, m, zz, df, dt2, ar);
  s[0] = (double)zz[0] * ar[0] + (double)zz[1] * ar[1] + (double)zz[2] * ar[2];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    s[1] += (double)ar[a] * (ar[a] / dg[a]);
    s[2] += (double)ar[a] * (ap[a] / dg[a]);
  }
}
__device__ __forceinline__ void cr_ar_node(const Batch& B, const NodeCtx& c, const float* mass, const int8_t* flags,
                                           const ColliderSet* cols, const float* stage, const float* z,
                                           const float* Ap, const float* diag, float dt2, float* Ar,
                                           double (&s)[3]) {
  float zz[3], df[3] = {0.0f, 0.0f, 0.0f}, ap[3], dg[3], ar[3];
  ld3(z, c.i, zz);
  ld3(Ap, c.i, ap);
  ld3(diag, c.i, dg);
  if (c.valid) gather_node3(B, c.G, stage, c.T, c.l, df);
  cr_ar_core(c, cols, flags[c.i], mass[c.i], zz, df, ap, dg, dt2, ar, s);
  st3(Ar, c.i, ar);
}
// beta, and alpha for the next x update; returns true when the solve ends here (rz of the
// previous iteration or the next denominator is zero: ConjugateResiduals.cpp:86-102, :140-145)
__device__ __forceinline__ bool cr_ar_finish(CRScalars& S, const double (&tot)[3]) {
  double rz_old = S.rz;
  S.rz = tot[0];
  if (rz_old == 0.0) {
    S.early_exit = 1;
    return true;
  }
  double beta = tot[0] / rz_old;
  double den = tot[1] + 2.0 * beta * tot[2] + beta * beta * S.ApdPAp;
  if (!(den > kCrDenKeep * (tot[1] + beta * beta * S.ApdPAp))) {
    beta = 0.0;
    den = tot[1];
    S.restarts = S.restarts + 1;
  }
  S.beta = beta;
  S.ApdPAp = den;
  S.iter = S.iter + 1;
  if (den == 0.0) {
    S.early_exit = 1;
    return true;
  }
  S.alpha = S.rz / den;
  return false;
}
__global__ void k_cr_ar(Batch B, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* stage,
                        const float* z, const float* Ap, const float* diag, float dt, float* Ar, CRScalars* sc,
                        BatchStatus* bst, double* partials) {
  NodeCtx c = node_begin(B);
  CRScalars& S = sc[c.b];
  if (S.done) return;
  double s[3] = {0.0, 0.0, 0.0};
  if (!B.tileActive || B.tileActive[c.T])
    cr_ar_node(B, c, mass, flags, cols, stage, z, Ap, diag, dt * dt, Ar, s);
  double tot[3];
  if (body_sum<3>(s, c.G, partials, body_tickets(B, c.b, 1), tot) && threadIdx.x == 0)
    if (cr_ar_finish(S, tot)) set_done(S, bst);
}
// iteration part 3 of the previous iteration (ConjugateResiduals.cpp:147-151: p = z + beta p,
// Ap = Ar + beta Ap) and part 1 of this one (x update, termination), in one pass
// s[1] sums <Ap, M^-1 Ap> of the updated Ap as stored (the next expansion's last term)
__device__ __forceinline__ void cr_update_px_core(const NodeCtx& c, const ColliderSet* cols, int fl, float beta,
                                                  float alpha, float xv[3], float pv[3], float ap[3],
                                                  const float ar[3], const float dg[3], float zz[3],
                                                  double (&s)[2]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    pv[a] = zz[a] + beta * pv[a];
    ap[a] = ar[a] + beta * ap[a];
    s[1] += (double)ap[a] * (ap[a] / dg[a]);
  }
  double sx[1] = {s[0]};
  cr_update_x_core(c, cols, fl, alpha, xv, pv, ap, dg, zz, sx);
  s[0] = sx[0];
}
__device__ __forceinline__ void cr_update_px_node(const NodeCtx& c, const int8_t* flags, const ColliderSet* cols,
                                                  const float* Ar, const float* diag, float beta, float alpha,
                                                  float* p, float* Ap, float* x, float* z, double (&s)[2]) {
  float xv[3], pv[3], ap[3], ar[3], dg[3], zz[3];
  ld3(x, c.i, xv);
  ld3(p, c.i, pv);
  ld3(Ap, c.i, ap);
  ld3(Ar, c.i, ar);
  ld3(diag, c.i, dg);
  ld3(z, c.i, zz);
  cr_update_px_core(c, cols, flags[c.i], beta, alpha, xv, pv, ap, ar, dg, zz, s);
  st3(p, c.i, pv);
  st3(Ap, c.i, ap);
  st3(x, c.i, xv);
  st3(z, c.i, zz);
}
__global__ void k_cr_update_px(Batch B, const int8_t* flags, const ColliderSet* cols, const float* Ar,
                               const float* diag, float* p, float* Ap, float* x, float* z, CRScalars* sc,
                               BatchStatus* bst, double* partials, double* hist) {
  NodeCtx c = node_begin(B);
  CRScalars& S = sc[c.b];
  if (S.done) return;
  double s[2] = {0.0, 0.0};
  if (!B.tileActive || B.tileActive[c.T])
    cr_update_px_node(c, flags, cols, Ar, diag, (float)S.beta, (float)S.alpha, p, Ap, x, z, s);
  double tot[2];
  if (body_sum<2>(s, c.G, partials, body_tickets(B, c.b, 0), tot) && threadIdx.x == 0)
    if (cr_update_x_finish<2>(S, tot, c.b, hist)) set_done(S, bst);
}

// ------------------------------------------------------------ small bodies
// A body of at most kSmallBlocks node blocks runs its whole conjugate residual solve
// (ConjugateResiduals.cpp:50-151) in one workgroup, concurrently with the batched iteration of
// the large bodies: fragments of a few particles otherwise keep the batched loop launching
// for tens of iterations after the large bodies converge. The host marks such bodies
// kDoneSmall after the rhs, which every batched kernel skips, and counts them in
// BatchStatus::small_bodies. The node work, the HVP chunk scatter, the reduction order (block
// tree per node block, then block totals in block order, as body_sum) and the scalar updates
// are those of the batched kernels, so the iterates are bitwise the same.
//
// The smallest bodies (at most kTinyBlocks node blocks and kTinyChunks particle chunks: the
// fragments of a few particles that can take hundreds of iterations) keep their whole state
// on chip: each thread holds its nodes' x, p, Ap, Ar, diag, mass and flags in registers, z and
// the staged HVP footprints live in LDS, and an iteration touches global memory only for the
// particles' constitutive data. The arithmetic is the same per-node cores and HVP pieces as
// the batched kernels, on the same float values.
constexpr int kTinyBlocks = 2;
constexpr int kTinyChunks = 4;
constexpr int kTinyTiles = kTinyBlocks * kBlock / 64;
struct TinyLds {
  float z[kTinyTiles * 192];            // the body's z (or x during setup), tile-major like the batch
  float stage[kTinyChunks * kStage];    // staged footprints of the body's chunks
  int8_t flags[kTinyTiles * 64];
  int tc0[kTinyTiles], tcn[kTinyTiles];  // each tile's first chunk (relative to the body's) and count
  int4 ch[kTinyChunks];
  int ft[kTinyChunks][8];               // footprint tiles of each chunk (global ids, -1 outside)
  HvpSplitLds split;                    // chunks of at most 64 particles: hvp_particles_split
};
template <int NV> struct SmallRed {
  double w[kBlock / 64][kSmallBlocks][NV];
  double sm[kBlock / 64][NV];
  double res[NV];
};
// every thread of the workgroup calls; returns the body total in every thread
template <int NV, class Fn>
__device__ __forceinline__ void small_sweep(const BodyGrid& G, int body, const float* nrm, SmallRed<NV>& R, Fn&& fn,
                                            double (&tot)[NV]) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int j = 0; j < G.nblocks; ++j) {
    NodeCtx c = node_in(G, body, G.blockBase + j, nrm);
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = 0.0;
    fn(c, v);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
      if (lane == 0) R.w[wid][j][k] = v[k];
    }
  }
  __syncthreads();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    acc[k] = 0.0;
    const int t = threadIdx.x;
    if (t < G.nblocks) acc[k] += ((R.w[0][t][k] + R.w[1][t][k]) + R.w[2][t][k]) + R.w[3][t][k];
  }
  block_sum<NV>(acc, R.sm);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) R.res[k] = acc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = R.res[k];
}
// tiny bodies: node j of this thread is G.blockBase + j; fn(j, v) adds this thread's terms
template <int NV, class Fn>
__device__ __forceinline__ void tiny_sweep(const BodyGrid& G, SmallRed<NV>& R, Fn&& fn, double (&tot)[NV]) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kTinyBlocks; ++j) {
    if (j >= G.nblocks) break;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = 0.0;
    fn(j, v);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
      if (lane == 0) R.w[wid][j][k] = v[k];
    }
  }
  __syncthreads();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    acc[k] = 0.0;
    const int t = threadIdx.x;
    if (t < G.nblocks) acc[k] += ((R.w[0][t][k] + R.w[1][t][k]) + R.w[2][t][k]) + R.w[3][t][k];
  }
  block_sum<NV>(acc, R.sm);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) R.res[k] = acc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = R.res[k];
}
// HVP of the LDS vector T.z over the body's chunks into T.stage (every thread calls)
__device__ __forceinline__ void tiny_hvp(const PView& P, const Batch& B, const BodyGrid& G, int nch, TinyLds& T,
                                         const ColliderSet* cols, int* err, float (*sd)[kFootLds],
                                         FxFoot1<3>& acc, const HvpPre (&pre)[kTinyChunks]) {
  for (int q = 0; q < nch; ++q) {
    // the previous chunk's fx_store reads acc in another thread mapping than fx_zero writes it
    __syncthreads();
    ChunkCtx c;
    c.ch = T.ch[q];
    c.G = G;
    c.nrm = B.nrm;
    int tx, ty, tz;
    tile_xyz(G, c.ch.x - G.tileBase, tx, ty, tz);
    c.ox = tx * 4;
    c.oy = ty * 4;
    c.oz = tz * 4;
    fx_zero(acc);
    load_foot3(c, T.ft[q], T.z, sd, T.flags, cols, true, G.tileBase);
    __syncthreads();
    // chunk q's entry by selects (a register array indexed by the loop counter would be moved
    // to memory)
    HvpPre pq = pre[0];
#pragma unroll
    for (int k = 1; k < kTinyChunks; ++k) {
      const bool m = q == k;
      pq.s = m ? pre[k].s : pq.s;
      pq.x[0] = m ? pre[k].x[0] : pq.x[0];
      pq.x[1] = m ? pre[k].x[1] : pq.x[1];
      pq.x[2] = m ? pre[k].x[2] : pq.x[2];
    }
    // a chunk of at most 64 particles spreads each particle over the four waves
    const FxScale3 scl = c.ch.z <= 64 ? hvp_particles_split(P, B, c, sd, acc, err, T.split, &pq)
                                      : hvp_particles(P, B, c, sd, acc, err, pq);
    fx_store(acc, scl, T.stage, q);
  }
  __syncthreads();
}
__device__ __forceinline__ void tiny_st3(TinyLds& T, const BodyGrid& G, const NodeCtx& c, const float v[3]) {
  const int lt = c.T - G.tileBase;
  T.z[lt * 192 + c.l] = v[0];
  T.z[lt * 192 + 64 + c.l] = v[1];
  T.z[lt * 192 + 128 + c.l] = v[2];
}
// the tiny-body solve (every thread of the workgroup calls; L and end as in k_small_cr)
__device__ __forceinline__ void tiny_solve(const PView& P, const Batch& B, int body, const BodyGrid& G, int c0,
                                           int nch, const float* mass, const int8_t* flags, const ColliderSet* cols,
                                           float dt, const float* bvec, const float* diag, float* x, double* hist,
                                           int* err, TinyLds& T, float (*sd)[kFootLds], FxFoot1<3>& acc,
                                           SmallRed<2>& R2, SmallRed<1>& R1, SmallRed<3>& R3, CRScalars& L,
                                           int& end) {
  const float dt2 = dt * dt;
  // chunk table, footprints and per-tile chunk ranges of the body
  if (threadIdx.x < nch * 8) {
    const int q = threadIdx.x >> 3, f = threadIdx.x & 7;
    const int4 ch = B.chunks[c0 + q];
    if (f == 0) T.ch[q] = ch;
    int tx, ty, tz;
    tile_xyz(G, ch.x - G.tileBase, tx, ty, tz);
    T.ft[q][f] = tile_id(G, tx + (f & 1), ty + ((f >> 1) & 1), tz + (f >> 2));
  }
  if (threadIdx.x < G.NT) {
    T.tc0[threadIdx.x] = B.tileChunk0[G.tileBase + threadIdx.x] - c0;
    T.tcn[threadIdx.x] = B.tileChunkN[G.tileBase + threadIdx.x];
  }
  // each thread's particle of every chunk (slot and position; a chunk's particles are fixed for
  // the solve): thread t of a chunk of more than 64 particles, lane t % 64 of one of at most 64
  HvpPre pre[kTinyChunks];
#pragma unroll
  for (int q = 0; q < kTinyChunks; ++q) {
    pre[q].s = -1;
    pre[q].x[0] = pre[q].x[1] = pre[q].x[2] = 0.0f;
    if (q < nch) {
      const int4 ch = B.chunks[c0 + q];
      const int k = ch.z <= 64 ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
      if (k < ch.z) {
        pre[q].s = B.perm[ch.y + k];
        load_x(P, pre[q].s, pre[q].x);
      }
    }
  }
  // this thread's nodes
  NodeCtx cn[kTinyBlocks];
  float xr[kTinyBlocks][3], pr[kTinyBlocks][3], apr[kTinyBlocks][3], arr[kTinyBlocks][3], dgr[kTinyBlocks][3],
      zr[kTinyBlocks][3], mr[kTinyBlocks];
  int flr[kTinyBlocks];
#pragma unroll
  for (int j = 0; j < kTinyBlocks; ++j) {
    if (j >= G.nblocks) break;
    cn[j] = node_in(G, body, G.blockBase + j, B.nrm);
    ld3(x, cn[j].i, xr[j]);
    ld3(diag, cn[j].i, dgr[j]);
    mr[j] = mass[cn[j].i];
    flr[j] = flags[cn[j].i];
    tiny_st3(T, G, cn[j], xr[j]);
    T.flags[(cn[j].T - G.tileBase) * 64 + cn[j].l] = (int8_t)flr[j];
  }
  __syncthreads();
  // setup (ConjugateResiduals.cpp:50-76)
  tiny_hvp(P, B, G, nch, T, cols, err, sd, acc, pre);
#pragma unroll
  for (int j = 0; j < kTinyBlocks; ++j) {
    if (j >= G.nblocks) break;
    float df[3] = {0.0f, 0.0f, 0.0f}, bb[3];
    if (cn[j].valid) gather3_tab(G, T.tc0, T.tcn, G.tileBase, T.stage, cn[j].T, cn[j].l, df);
    ld3(bvec, cn[j].i, bb);
    cr_setup1_core(cn[j], cols, flr[j], mr[j], xr[j], df, bb, dgr[j], dt2, zr[j]);
#pragma unroll
    for (int a = 0; a < 3; ++a) pr[j][a] = zr[j][a];
  }
  __syncthreads();  // every gather of the x footprints is done before z replaces x
#pragma unroll
  for (int j = 0; j < kTinyBlocks; ++j) {
    if (j >= G.nblocks) break;
    tiny_st3(T, G, cn[j], zr[j]);
  }
  __syncthreads();
  tiny_hvp(P, B, G, nch, T, cols, err, sd, acc, pre);
  double t2[2];
  tiny_sweep<2>(G, R2, [&](int j, double (&v)[2]) {
    float df[3] = {0.0f, 0.0f, 0.0f};
    if (cn[j].valid) gather3_tab(G, T.tc0, T.tcn, G.tileBase, T.stage, cn[j].T, cn[j].l, df);
    cr_setup2_core(cn[j], cols, flr[j], mr[j], zr[j], df, dgr[j], dt2, apr[j], v);
  }, t2);
  if (threadIdx.x == 0) end = cr_setup2_finish(L, t2);
  __syncthreads();
  double t1[1], t3[3];
  for (int it = 0; !end; ++it) {
    const float alpha = (float)L.alpha;
    if (it == 0) {
      tiny_sweep<1>(G, R1, [&](int j, double (&v)[1]) {
        cr_update_x_core(cn[j], cols, flr[j], alpha, xr[j], pr[j], apr[j], dgr[j], zr[j], v);
        tiny_st3(T, G, cn[j], zr[j]);
      }, t1);
    } else {
      const float beta = (float)L.beta;
      double t2p[2];
      tiny_sweep<2>(G, R2, [&](int j, double (&v)[2]) {
        cr_update_px_core(cn[j], cols, flr[j], beta, alpha, xr[j], pr[j], apr[j], arr[j], dgr[j], zr[j], v);
        tiny_st3(T, G, cn[j], zr[j]);
      }, t2p);
      if (threadIdx.x == 0) end = cr_update_x_finish<2>(L, t2p, body, hist);
    }
    if (it == 0 && threadIdx.x == 0) end = cr_update_x_finish<1>(L, t1, body, hist);
    __syncthreads();
    if (end) break;
    tiny_hvp(P, B, G, nch, T, cols, err, sd, acc, pre);
    tiny_sweep<3>(G, R3, [&](int j, double (&v)[3]) {
      float df[3] = {0.0f, 0.0f, 0.0f};
      if (cn[j].valid) gather3_tab(G, T.tc0, T.tcn, G.tileBase, T.stage, cn[j].T, cn[j].l, df);
      cr_ar_core(cn[j], cols, flr[j], mr[j], zr[j], df, apr[j], dgr[j], dt2, arr[j], v);
    }, t3);
    if (threadIdx.x == 0) end = cr_ar_finish(L, t3);
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < kTinyBlocks; ++j) {
    if (j >= G.nblocks) break;
    st3(x, cn[j].i, xr[j]);
  }
}
__global__ void k_small_claim(const int* list, int n, CRScalars* sc, BatchStatus* bst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  CRScalars& S = sc[list[i]];
  if (S.done == 0) {
    S.done = kDoneSmall;
    atomicAdd(&bst->small_bodies, 1);
  }
}
__global__ __launch_bounds__(kBlock) void k_small_cr(PView P, Batch B, const int* list, const float* mass,
                                                     const int8_t* flags, const ColliderSet* cols, float dt,
                                                     float* stage, const float* bvec, const float* diag, float* x,
                                                     float* z, float* p, float* Ap, float* Ar, CRScalars* sc,
                                                     double* hist, int* err) {
  __shared__ float sd[3][kFootLds];
  __shared__ FxFoot1<3> acc;
  __shared__ int ft[8];
  __shared__ SmallRed<2> R2;
  __shared__ SmallRed<1> R1;
  __shared__ SmallRed<3> R3;
  __shared__ CRScalars L;
  __shared__ int end;
  __shared__ TinyLds T;
  const int body = list[blockIdx.x];
  if (sc[body].done != kDoneSmall) return;  // finished at the rhs: not claimed
  const BodyGrid G = B.bodies[body];
  if (G.nblocks > kSmallBlocks) {  // the host only lists bodies that fit
    if (threadIdx.x == 0) atomicOr(err, ERR_INTERNAL);
    return;
  }
  if (threadIdx.x == 0) {
    L = sc[body];
    end = 0;
  }
  const float dt2 = dt * dt;
  const int lastTile = G.tileBase + G.NT - 1;
  const int c0 = B.tileChunk0[G.tileBase];
  const int c1 = B.tileChunk0[lastTile] + B.tileChunkN[lastTile];
  if (G.nblocks <= kTinyBlocks && c1 - c0 <= kTinyChunks) {
    __syncthreads();
    tiny_solve(P, B, body, G, c0, c1 - c0, mass, flags, cols, dt, bvec, diag, x, hist, err, T, sd, acc, R2, R1, R3,
               L, end);
    if (threadIdx.x == 0) {
      L.done = 1;
      sc[body] = L;
    }
    return;
  }
  auto hvp = [&](const float* d) {
    for (int ch = c0; ch < c1; ++ch) {
      __syncthreads();  // the previous chunk's fx_store reads acc in another mapping than fx_zero writes it
      hvp_chunk<true>(P, B, ch, d, flags, cols, stage, err, sd, acc, ft, &T.split);
    }
    __syncthreads();
  };
  __syncthreads();
  // setup
  hvp(x);
  for (int j = 0; j < G.nblocks; ++j)
    cr_setup1_node(B, node_in(G, body, G.blockBase + j, B.nrm), mass, flags, cols, stage, x, bvec, diag, dt2, z,
                   p);
  __syncthreads();
  hvp(z);
  double t2[2];
  small_sweep<2>(G, body, B.nrm, R2, [&](const NodeCtx& c, double (&v)[2]) {
    cr_setup2_node(B, c, mass, flags, cols, stage, z, diag, dt2, Ap, v);
  }, t2);
  if (threadIdx.x == 0) end = cr_setup2_finish(L, t2);
  __syncthreads();
  double t1[1], t3[3];
  for (int it = 0; !end; ++it) {
    const float alpha = (float)L.alpha;
    if (it == 0) {
      small_sweep<1>(G, body, B.nrm, R1, [&](const NodeCtx& c, double (&v)[1]) {
        cr_update_x_node(c, flags, cols, p, Ap, diag, alpha, x, z, v);
      }, t1);
    } else {
      const float beta = (float)L.beta;
      double t2p[2];
      small_sweep<2>(G, body, B.nrm, R2, [&](const NodeCtx& c, double (&v)[2]) {
        cr_update_px_node(c, flags, cols, Ar, diag, beta, alpha, p, Ap, x, z, v);
      }, t2p);
      if (threadIdx.x == 0) end = cr_update_x_finish<2>(L, t2p, body, hist);
    }
    if (it == 0 && threadIdx.x == 0) end = cr_update_x_finish<1>(L, t1, body, hist);
    __syncthreads();
    if (end) break;
    hvp(z);
    small_sweep<3>(G, body, B.nrm, R3, [&](const NodeCtx& c, double (&v)[3]) {
      cr_ar_node(B, c, mass, flags, cols, stage, z, Ap, diag, dt2, Ar, v);
    }, t3);
    if (threadIdx.x == 0) end = cr_ar_finish(L, t3);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    L.done = 1;
    sc[body] = L;
  }
}
// ------------------------------------------------------------ flexible conjugate residual
// The learned preconditioner changes from one iteration to the next, which breaks the
// short recurrences of ConjugateResiduals.cpp:84-151. The solve then runs a truncated
// generalised conjugate residual in the inner product <u, v>_D = sum u v / diag: each
// preconditioned vector z is orthogonalised against the last fcr_m directions (A p_j in that
// inner product) and x, r take the step that minimises |r|_D. With the diagonal
// preconditioner z = r / diag this is the preconditioned conjugate residual iteration of
// ConjugateResiduals.cpp (same Krylov space, same minimised norm). The same termination
// (|r|^2 < tol^2 |b|^2 or the iteration cap, SquareMagnitudeTermination.cpp:19-28) judges
// the result. Whenever an iteration that used the network fails to lower |r|^2 by the body's
// fallback ratio (the context's setting: 0.95 by default, 1 = any rise) or by as much as the
// body's last diagonal iteration did, the body takes its next iteration with the diagonal
// preconditioner; after the context's strike count of such fallbacks (kLpcStrikes by default) it
// keeps the diagonal for the rest of the
// solve.

// r = b - A x0 ; network input scale ; iteration state
__global__ void k_fcr_init(Batch B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                           const float* stage, const float* x, const float* b, const float* diag, float dt, float* r,
                           int netOn, double ratio, int mdirs, int strikes, CRScalars* sc, BatchStatus* bst,
                           double* partials) {
  NodeCtx c = node_begin(B);
  CRScalars& S = sc[c.b];
  if (S.done) return;
  const float dt2 = dt * dt;
  float xv[3], df[3] = {0.0f, 0.0f, 0.0f}, Ax[3], bb[3], dg[3], rr[3];
  ld3(x, c.i, xv);
  if (c.valid) gather_node3(B, c.G, stage, c.T, c.l, df);
  const float m = mass[c.i];
  apply_A(c, cols, flags[c.i], m, xv, df, dt2, Ax);
  ld3(b, c.i, bb);
  ld3(diag, c.i, dg);
  double s[3] = {0.0, 0.0, m > 0.0f ? 3.0 : 0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    rr[a] = bb[a] - Ax[a];
    s[0] += (double)rr[a] * rr[a];
    if (m > 0.0f) {
      const double q = (double)(rr[a] / dg[a]);
      s[1] += q * q;
    }
  }
  st3(r, c.i, rr);
  double tot[3];
  if (body_sum<3>(s, c.G, partials, body_tickets(B, c.b, 0), tot) && threadIdx.x == 0) {
    S.rnorm2 = tot[0];
    S.rr_prev = tot[0];
    S.lpc_ratio = ratio;
    S.diag_rate = ratio;
    S.lpc_s2 = tot[1];
    S.lpc_cnt = tot[2];
    S.lpc_scale = tot[2] > 0.0 ? (float)sqrt(tot[1] / tot[2]) : 0.0f;
    S.lpc_off = netOn ? 0 : 1;
    S.fallbacks = 0;
    S.retry = 0;
    S.lpc_pause = 0;
    S.lpc_strikes = strikes;
    S.ndirs = 0;
    S.fcr_m = mdirs;
    S.slot = mdirs - 1;
    for (int j = 0; j < kFcrMax; ++j) {
      S.ApAp[j] = 0.0;
      S.gbeta[j] = 0.0;
    }
    S.iter = 0;
    S.iterations = 0;
    S.early_exit = 0;
  }
}
// z = r / diag for bodies on the diagonal preconditioner
__global__ void k_fcr_diag_z(Batch B, const float* r, const float* diag, float* z, const CRScalars* sc) {
  NodeCtx c = node_begin(B);
  const CRScalars& S = sc[c.b];
  if (S.done || !S.lpc_off) return;
  float rr[3], dg[3], zz[3];
  ld3(r, c.i, rr);
  ld3(diag, c.i, dg);
#pragma unroll
  for (int a = 0; a < 3; ++a) zz[a] = rr[a] / dg[a];
  st3(z, c.i, zz);
}
// A z and its projections on the stored directions' A p_j
__global__ void k_fcr_az(Batch B, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* stage,
                         const float* z, const float* diag, float dt, float* Az, const float* Q, int64_t vs,
                         CRScalars* sc, BatchStatus* bst, double* partials) {
  NodeCtx c = node_begin(B);
  CRScalars& S = sc[c.b];
  if (S.done) return;
  const float dt2 = dt * dt;
  float zz[3], df[3] = {0.0f, 0.0f, 0.0f}, az[3];
  ld3(z, c.i, zz);
  if (c.valid) gather_node3(B, c.G, stage, c.T, c.l, df);
  apply_A(c, cols, flags[c.i], mass[c.i], zz, df, dt2, az);
  st3(Az, c.i, az);
  double s[kFcrMax];
  float dg[3];
  ld3(diag, c.i, dg);
#pragma unroll
  for (int j = 0; j < kFcrMax; ++j) {
    s[j] = 0.0;
    if (j < S.fcr_m && S.ApAp[j] != 0.0) {
      float q[3];
      ld3(Q + j * vs, c.i, q);
      s[j] = (double)az[0] * (q[0] / dg[0]) + (double)az[1] * (q[1] / dg[1]) + (double)az[2] * (q[2] / dg[2]);
    }
  }
  double tot[kFcrMax];
  if (body_sum<kFcrMax>(s, c.G, partials, body_tickets(B, c.b, 1), tot) && threadIdx.x == 0) {
    for (int j = 0; j < kFcrMax; ++j) S.gbeta[j] = S.ApAp[j] != 0.0 ? tot[j] / S.ApAp[j] : 0.0;
  }
}
// new direction p = z - sum beta_j p_j, A p = A z - sum beta_j A p_j into the ring's next slot;
// alpha = <r, A p> / <A p, A p>
__global__ void k_fcr_orth(Batch B, const float* z, const float* Az, const float* r, const float* diag, float* P,
                           float* Q, int64_t vs, CRScalars* sc, BatchStatus* bst, double* partials) {
  NodeCtx c = node_begin(B);
  CRScalars& S = sc[c.b];
  if (S.done) return;
  const int slot = (S.slot + 1) % S.fcr_m;
  float p[3], ap[3], rr[3];
  ld3(z, c.i, p);
  ld3(Az, c.i, ap);
  ld3(r, c.i, rr);
#pragma unroll
  for (int j = 0; j < kFcrMax; ++j) {
    const float bj = j < S.fcr_m ? (float)S.gbeta[j] : 0.0f;
    if (bj != 0.0f) {
      float pj[3], qj[3];
      ld3(P + j * vs, c.i, pj);
      ld3(Q + j * vs, c.i, qj);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        p[a] -= bj * pj[a];
        ap[a] -= bj * qj[a];
      }
    }
  }
  st3(P + slot * vs, c.i, p);
  st3(Q + slot * vs, c.i, ap);
  double s[2] = {0.0, 0.0};
  float dg[3];
  ld3(diag, c.i, dg);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double w = (double)(ap[a] / dg[a]);
    s[0] += (double)ap[a] * w;
    s[1] += (double)rr[a] * w;
  }
  double tot[2];
  if (body_sum<2>(s, c.G, partials, body_tickets(B, c.b, 2), tot) && threadIdx.x == 0) {
    if (tot[0] == 0.0 && !S.lpc_off) {
      // the network returned nothing usable: switch to the diagonal and redo the iteration
      S.lpc_off = 1;
      S.fallbacks += 1;
      S.retry = 1;
      S.ApAp[slot] = 0.0;  // this ring slot now holds no direction
    } else if (tot[0] == 0.0) {
      S.early_exit = 1;
      set_done(S, bst);
    } else {
      S.retry = 0;
      S.ApAp[slot] = tot[0];
      S.rAp = tot[1];
      S.alpha = tot[1] / tot[0];
      S.slot = slot;
      S.ndirs = S.ndirs < S.fcr_m ? S.ndirs + 1 : S.fcr_m;
    }
  }
}
// x = P(x + alpha p), r -= alpha A p; termination, fallback and the next network input scale
__global__ void k_fcr_update(Batch B, const int8_t* flags, const ColliderSet* cols, const float* mass,
                             const float* diag, const float* P, const float* Q, int64_t vs, float* x, float* r,
                             CRScalars* sc, BatchStatus* bst, double* partials, double* hist) {
  NodeCtx c = node_begin(B);
  CRScalars& S = sc[c.b];
  if (S.done || S.retry) return;
  const float alpha = (float)S.alpha;
  float xv[3], p[3], ap[3], rr[3], dg[3];
  ld3(x, c.i, xv);
  ld3(P + S.slot * vs, c.i, p);
  ld3(Q + S.slot * vs, c.i, ap);
  ld3(r, c.i, rr);
  ld3(diag, c.i, dg);
#pragma unroll
  for (int a = 0; a < 3; ++a) xv[a] = xv[a] + alpha * p[a];
  node_project(c, cols, flags[c.i], xv);
  st3(x, c.i, xv);
  const float m = mass[c.i];
  double s[3] = {0.0, 0.0, m > 0.0f ? 3.0 : 0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    rr[a] = rr[a] - alpha * ap[a];
    s[0] += (double)rr[a] * rr[a];
    if (m > 0.0f) {
      const double q = (double)(rr[a] / dg[a]);
      s[1] += q * q;
    }
  }
  st3(r, c.i, rr);
  double tot[3];
  if (body_sum<3>(s, c.G, partials, body_tickets(B, c.b, 0), tot) && threadIdx.x == 0) {
    const int it = S.iter;
    S.rnorm2 = tot[0];
    S.iterations = it + 1;
    if (hist && it < kHistCap) hist[(int64_t)c.b * kHistCap + it] = tot[0];
    S.lpc_s2 = tot[1];
    S.lpc_cnt = tot[2];
    S.lpc_scale = tot[2] > 0.0 ? (float)sqrt(tot[1] / tot[2]) : 0.0f;
    // a network iteration must lower |r|^2 at least as much as the body's last diagonal
    // iteration did (and by the fallback ratio)
    const double bar = fmin(S.lpc_ratio, S.diag_rate);
    if (S.lpc_off && S.rr_prev > 0.0) S.diag_rate = tot[0] / S.rr_prev;
    if (it >= S.max_iters || tot[0] < S.threshold) {
      set_done(S, bst);
    } else if (!S.lpc_off && tot[0] > bar * S.rr_prev) {
      S.lpc_off = 1;
      S.fallbacks += 1;
      S.lpc_pause = S.fallbacks < S.lpc_strikes ? 1 : 0;
    } else if (S.lpc_off && S.lpc_pause) {
      S.lpc_off = 0;  // one diagonal iteration done: back to the network
      S.lpc_pause = 0;
    }
    S.rr_prev = tot[0];
    S.iter = it + 1;
  }
}

// ------------------------------------------------------------ learned preconditioner training
// The network is trained on what one flexible step achieves with its correction: the residual
// left after the optimal step along u = A z, per body L_b = 1 - <r, u>_D^2 / (|r|_D^2 |u|_D^2)
// (<a, b>_D = sum a b / diag), z = r / diag + s_b clamp(out). A is symmetric, so
// dL_b/dz = A y with y = (2 a / (R c)) D^-1 (alpha u - r), a = <r, u>_D, c = |u|_D^2,
// R = |r|_D^2, alpha = a / c.
__global__ void k_lpc_compose_z(Batch B, const float* r, const float* diag, const float* raw, const float* scale,
                                const float* weight, float* z) {
  NodeCtx c = node_begin(B);
  float rr[3], dg[3], zz[3] = {0.0f, 0.0f, 0.0f};
  if (weight[c.b] > 0.0f && c.valid) {
    ld3(r, c.i, rr);
    ld3(diag, c.i, dg);
    float o[3] = {0.0f, 0.0f, 0.0f};
    if (raw) ld3(raw, c.i, o);
#pragma unroll
    for (int a = 0; a < 3; ++a) zz[a] = rr[a] / dg[a] + fminf(fmaxf(o[a], -kLpcClamp), kLpcClamp) * scale[c.b];
  }
  st3(z, c.i, zz);
}
// u = A z from the staged force differential; per-body dots[3 b + {0, 1, 2}] += <r, u>_D, |u|_D^2, |r|_D^2
__global__ void k_lpc_res_dots(Batch B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                               const float* stage, const float* z, const float* r, const float* diag, float dt,
                               const float* weight, float* u, double* dots) {
  NodeCtx c = node_begin(B);
  float uu[3] = {0.0f, 0.0f, 0.0f};
  double s[3] = {0.0, 0.0, 0.0};
  if (weight[c.b] > 0.0f && c.valid) {
    float zz[3], df[3], rr[3], dg[3];
    ld3(z, c.i, zz);
    gather_node3(B, c.G, stage, c.T, c.l, df);
    apply_A(c, cols, flags[c.i], mass[c.i], zz, df, dt * dt, uu);
    ld3(r, c.i, rr);
    ld3(diag, c.i, dg);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      s[0] += (double)rr[a] * (double)(uu[a] / dg[a]);
      s[1] += (double)uu[a] * (double)(uu[a] / dg[a]);
      s[2] += (double)rr[a] * (double)(rr[a] / dg[a]);
    }
  }
  st3(u, c.i, uu);
  __shared__ double red[3][kBlock / 64];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double v = s[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const double t = ((red[threadIdx.x][0] + red[threadIdx.x][1]) + red[threadIdx.x][2]) + red[threadIdx.x][3];
    if (t != 0.0) atomicAdd(dots + 3 * c.b + threadIdx.x, t);
  }
}
// y = (2 a / (R c)) D^-1 (alpha u - r); *loss += 1 - a^2 / (R c) once per body (first node block)
__global__ void k_lpc_res_y(Batch B, const float* u, const float* r, const float* diag, const float* weight,
                            const double* dots, float* y, double* loss) {
  NodeCtx c = node_begin(B);
  float yy[3] = {0.0f, 0.0f, 0.0f};
  const double a = dots[3 * c.b], cc = dots[3 * c.b + 1], R = dots[3 * c.b + 2];
  const bool ok = weight[c.b] > 0.0f && cc > 0.0 && R > 0.0;
  if (ok && c.valid) {
    const float alpha = (float)(a / cc), f = (float)(2.0 * a / (R * cc));
    float uu[3], rr[3], dg[3];
    ld3(u, c.i, uu);
    ld3(r, c.i, rr);
    ld3(diag, c.i, dg);
#pragma unroll
    for (int k = 0; k < 3; ++k) yy[k] = f * (alpha * uu[k] - rr[k]) / dg[k];
  }
  st3(y, c.i, yy);
  if (ok && threadIdx.x == 0 && blockIdx.x == (unsigned)c.G.blockBase) atomicAdd(loss, 1.0 - a * a / (R * cc));
}
// output-side gradient of the network: g = s_b (A y) where |out| < clamp, bf16 [node][16 channels]
__global__ void k_lpc_res_grad(Batch B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                               const float* stage, const float* y, const float* raw, float dt, const float* scale,
                               const float* weight, __bf16* g) {
  NodeCtx c = node_begin(B);
  float gv[3] = {0.0f, 0.0f, 0.0f};
  if (weight[c.b] > 0.0f && c.valid) {
    float yy[3], df[3], ay[3], o[3];
    ld3(y, c.i, yy);
    gather_node3(B, c.G, stage, c.T, c.l, df);
    apply_A(c, cols, flags[c.i], mass[c.i], yy, df, dt * dt, ay);
    ld3(raw, c.i, o);
#pragma unroll
    for (int a = 0; a < 3; ++a) gv[a] = fabsf(o[a]) < kLpcClamp ? scale[c.b] * ay[a] : 0.0f;
  }
  __bf16* dst = g + c.i * 16;
#pragma unroll
  for (int k = 0; k < 16; ++k) dst[k] = (__bf16)(k < 3 ? gv[k] : 0.0f);
}

// ---- exact-correction (energy-norm) loss. With the solve's final iterate x*, sample k's exact
// correction is e = x* - x_k (A e = r_k up to the final residual) and its error energy
// E = e^T A e = e^T r_k. A step alpha z leaves (e - alpha z)^T A (e - alpha z); the best alpha
// leaves the fraction L = 1 - (z^T r)^2 / ((z^T A z) E) of E, the per-body loss.
template <int NV>
__device__ __forceinline__ void block_add_to(double (&s)[NV], double* dst, bool on) {
  __shared__ double red[NV][kBlock / 64];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double v = s[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < NV && on) {
    const double t = ((red[threadIdx.x][0] + red[threadIdx.x][1]) + red[threadIdx.x][2]) + red[threadIdx.x][3];
    if (t != 0.0) atomicAdd(dst + threadIdx.x, t);
  }
}
// energy[b] += (x* - x_k) . r_k over the body's nodes
__global__ void k_lpc_exact_energy(Batch B, const float* x, const float* xk, const float* rk, double* energy) {
  NodeCtx c = node_begin(B);
  double s[1] = {0.0};
  if (c.valid) {
    float a[3], b[3], r[3];
    ld3(x, c.i, a);
    ld3(xk, c.i, b);
    ld3(rk, c.i, r);
#pragma unroll
    for (int k = 0; k < 3; ++k) s[0] += ((double)a[k] - (double)b[k]) * (double)r[k];
  }
  block_add_to<1>(s, energy + c.b, true);
}
// u = A z from the staged force differential; dots[2 b + {0, 1}] += z . r, z . u
__global__ void k_lpc_energy_dots(Batch B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                                  const float* stage, const float* z, const float* r, float dt, const float* weight,
                                  float* u, double* dots) {
  NodeCtx c = node_begin(B);
  float uu[3] = {0.0f, 0.0f, 0.0f};
  double s[2] = {0.0, 0.0};
  const bool on = weight[c.b] > 0.0f;
  if (on && c.valid) {
    float zz[3], df[3], rr[3];
    ld3(z, c.i, zz);
    gather_node3(B, c.G, stage, c.T, c.l, df);
    apply_A(c, cols, flags[c.i], mass[c.i], zz, df, dt * dt, uu);
    ld3(r, c.i, rr);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      s[0] += (double)zz[a] * rr[a];
      s[1] += (double)zz[a] * uu[a];
    }
  }
  st3(u, c.i, uu);
  block_add_to<2>(s, dots + 2 * c.b, on);
}
// the network output's gradient g = scale dL/dz with dL/dz = (2 a / (c E)) ((a / c) u - r),
// a = z . r, c = z . u (where the output is not clamped; bf16 [node][16 channels]); *loss += the
// body's L once (first node block of the body)
__global__ void k_lpc_energy_grad(Batch B, const float* u, const float* r, const float* raw, const float* scale,
                                  const float* weight, const double* dots, const double* energy, __bf16* g,
                                  double* loss) {
  NodeCtx c = node_begin(B);
  const double a = dots[2 * c.b], cc = dots[2 * c.b + 1], E = energy[c.b];
  const bool ok = weight[c.b] > 0.0f && cc > 0.0 && E > 0.0;
  float gv[3] = {0.0f, 0.0f, 0.0f};
  if (ok && c.valid && raw) {
    const float alpha = (float)(a / cc), f = (float)(2.0 * a / (cc * E));
    float uu[3], rr[3], o[3];
    ld3(u, c.i, uu);
    ld3(r, c.i, rr);
    ld3(raw, c.i, o);
#pragma unroll
    for (int k = 0; k < 3; ++k) gv[k] = fabsf(o[k]) < kLpcClamp ? scale[c.b] * f * (alpha * uu[k] - rr[k]) : 0.0f;
  }
  if (g) {
    __bf16* dst = g + c.i * 16;
#pragma unroll
    for (int k = 0; k < 16; ++k) dst[k] = (__bf16)(k < 3 ? gv[k] : 0.0f);
  }
  if (ok && threadIdx.x == 0 && blockIdx.x == (unsigned)c.G.blockBase) atomicAdd(loss, 1.0 - a * a / (cc * E));
}

// ---- vector kernels of the conjugate-residual solve over a caller's matrix (mpm_conjugate_residuals)
constexpr int kDotBlocks = 512;
// partials[block] = sum of a[i] b[i] over the block's grid-stride share, in FP64, fixed order
__global__ __launch_bounds__(kBlock) void k_vec_partial_dots(const float* a, const float* b, int64_t n,
                                                             double* partials) {
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)kDotBlocks * kBlock)
    v += (double)a[i] * (double)b[i];
  __shared__ double sm[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = ((sm[0] + sm[1]) + sm[2]) + sm[3];
}
__global__ __launch_bounds__(kBlock) void k_vec_sum_partials(const double* partials, double* out) {
  double v = 0.0;
  for (int i = threadIdx.x; i < kDotBlocks; i += kBlock) v += partials[i];
  __shared__ double sm[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) *out = ((sm[0] + sm[1]) + sm[2]) + sm[3];
}
// out = a u + c v (v may be null: out = a u)
__global__ void k_vec_lincomb(float* out, float a, const float* u, float c, const float* v, int64_t n) {
  GRID_STRIDE(i, n) out[i] = v ? a * u[i] + c * v[i] : a * u[i];
}

__global__ void k_residual_vector(Batch B, const float* z, const float* diag, float* r) {
  NodeCtx c = node_begin(B);
  float zz[3], dg[3];
  ld3(z, c.i, zz);
  ld3(diag, c.i, dg);
  zz[0] *= dg[0];
  zz[1] *= dg[1];
  zz[2] *= dg[2];
  st3(r, c.i, zz);
}
// m_velocities += vc (Grid.cpp:1014); x = 0 when the rhs was zero
__global__ void k_finalize(Batch B, const float* x, const float* vc, float* vel, const CRScalars* sc) {
  NodeCtx c = node_begin(B);
  const bool zx = sc[c.b].zero_x != 0;
  float xv[3], cv[3];
  ld3(x, c.i, xv);
  ld3(vc, c.i, cv);
  if (zx) xv[0] = xv[1] = xv[2] = 0.0f;
  xv[0] += cv[0];
  xv[1] += cv[1];
  xv[2] += cv[2];
  st3(vel, c.i, xv);
}

// ------------------------------------------------------------ G2P
// updateParticleVelocities (Grid.cpp:1049-1081), updateDeformationGradients (Grid.cpp:1018-1047),
// and for the full step (MODE 2) SnowConstitutiveModel::updateParticleData
// (SnowConstitutiveModel.cpp:27-121). The step's advection (Sim.cpp:123-131) is k_advect: fused
// here it raised the kernel to 256 VGPRs with spills.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_g2p(PView P, Batch B, const float* vel, const float* prev, float dt,
                                                float thc, float ths, const ColliderSet* cols, int* err) {
  constexpr bool DO_V = MODE != 1;
  constexpr bool DO_F = MODE != 0;
  __shared__ float sv[3][kFootLds];
  __shared__ float sp[DO_V ? 3 : 1][kFootLds];
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(B, ft, blockIdx.x);
  __syncthreads();
  for (int j = threadIdx.x; j < kFoot; j += kBlock) {
    int gt = ft[j >> 6], i = foot_lds(j);
    int64_t b = (int64_t)gt * 192 + (j & 63);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      sv[a][i] = gt >= 0 ? vel[b + 64 * a] : 0.0f;
      if (DO_V) sp[a][i] = gt >= 0 ? prev[b + 64 * a] : 0.0f;
    }
  }
  __syncthreads();
  const float alpha = 0.95f;
  const float beta = 1.0f - alpha;
  for (int k = threadIdx.x; k < c.ch.z; k += kBlock) {
    int s = B.perm[c.ch.y + k];
    float x[3];
    load_x(P, s, x);
    Stencil st;
    int rx, ry, rz;
    if (!chunk_stencil(c, x, DO_F, st, rx, ry, rz)) {
      atomicOr(err, ERR_STENCIL);
      continue;
    }
    float vf[3] = {0.0f, 0.0f, 0.0f}, vp[3] = {0.0f, 0.0f, 0.0f};
    if (DO_V) {
      vf[0] = P.atu(C_V, s);
      vf[1] = P.atu(C_V + 1, s);
      vf[2] = P.atu(C_V + 2, s);
    }
    float dV[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) dV[i] = 0.0f;
#pragma unroll
    for (int cz = 0; cz < 4; ++cz)
#pragma unroll
      for (int cy = 0; cy < 4; ++cy)
#pragma unroll
        for (int cx = 0; cx < 4; ++cx) {
          int idx = foot_index(rx + cx, ry + cy, rz + cz);
          float vi[3] = {sv[0][idx], sv[1][idx], sv[2][idx]};
          if (DO_V) {
            float w = st_w(st, cx, cy, cz);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              vf[a] += w * (vi[a] - sp[a][idx]);
              vp[a] += w * (vi[a] + c.G.fv[a]);
            }
          }
          if (DO_F) {
            float gw[3];
            st_grad(st, cx, cy, cz, gw);
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
              dV[0 + 3 * cc] += vi[0] * gw[cc];
              dV[1 + 3 * cc] += vi[1] * gw[cc];
              dV[2 + 3 * cc] += vi[2] * gw[cc];
            }
          }
        }
    float v[3];
    if (DO_V) {
#pragma unroll
      for (int a = 0; a < 3; ++a) v[a] = alpha * vf[a] + beta * vp[a];
#pragma unroll
      for (int a = 0; a < 3; ++a) P.atu(C_V + a, s) = v[a];
    }
    if (DO_F) {
      float F[9], A[9], Fn[9];
      load_m9(P, C_F, s, F);
#pragma unroll
      for (int i = 0; i < 9; ++i) A[i] = ((i == 0 || i == 4 || i == 8) ? 1.0f : 0.0f) + dt * dV[i];
      mat_mul(A, F, Fn);
      if (MODE == 1) {
        store_m9(P, C_F, s, Fn);
      } else {
        float Fp[9], Fi[9], R[9], Gi[9], J, mu, lam;
        load_m9(P, C_FP, s, Fp);
        bool ok = snow_update(Fn, Fp, Fi, R, Gi, J, mu, lam, P.atu(C_MU0, s), P.atu(C_LAM0, s), P.atu(C_XI, s), thc, ths);
        if (!ok) atomicOr(err, ERR_HARDNESS);
        store_m9(P, C_F, s, Fn);
        store_m9(P, C_FP, s, Fp);
        store_m9(P, C_FI, s, Fi);
        store_m9(P, C_R, s, R);
        store_m9(P, C_GI, s, Gi);
        P.atu(C_J, s) = J;
        P.atu(C_MU, s) = mu;
        P.atu(C_LAM, s) = lam;
      }
    }
  }
}

__global__ void k_plasticity(PView P, const int* labels, int skipLabel, int n, float thc, float ths, int* err) {
  GRID_STRIDE(si, n) {
    int s = (int)si;
    if (labels && labels[s] == skipLabel) continue;
    float F[9], Fp[9], Fi[9], R[9], Gi[9], J, mu, lam;
    load_m9(P, C_F, s, F);
    load_m9(P, C_FP, s, Fp);
    bool ok = snow_update(F, Fp, Fi, R, Gi, J, mu, lam, P.atu(C_MU0, s), P.atu(C_LAM0, s), P.atu(C_XI, s), thc, ths);
    if (!ok) atomicOr(err, ERR_HARDNESS);
    store_m9(P, C_F, s, F);
    store_m9(P, C_FP, s, Fp);
    store_m9(P, C_FI, s, Fi);
    store_m9(P, C_R, s, R);
    store_m9(P, C_GI, s, Gi);
    P.atu(C_J, s) = J;
    P.atu(C_MU, s) = mu;
    P.atu(C_LAM, s) = lam;
  }
}

// ballistic particles (Sim.cpp:64-73) and/or their advection (Sim.cpp:123-131)
__global__ void k_ballistic(PView P, const int* labels, int n, const float* grav, int ngrav, float dt,
                            const ColliderSet* cols, int mode) {
  GRID_STRIDE(si, n) {
    int s = (int)si;
    if (labels[s] != -1) continue;
    float v[3], x[3];
    load_x(P, s, x);
#pragma unroll
    for (int q = 0; q < 3; ++q) v[q] = P.atu(C_V + q, s);
    if (mode & 1) {
      float m = P.atu(C_M, s);
      float a[3] = {0.0f, 0.0f, 0.0f};
      field_force(grav, ngrav, m, a);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        a[q] = a[q] / m;
        v[q] = v[q] + dt * a[q];
      }
    }
    if (mode & 2) {
      const float zero[3] = {0.0f, 0.0f, 0.0f};
      collide(*cols, v, x, zero, true);
#pragma unroll
      for (int q = 0; q < 3; ++q) P.atu(C_X + q, s) = x[q] + v[q] * dt;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) P.atu(C_V + q, s) = v[q];
  }
}
__global__ void k_advect(PView P, const int* labels, int n, float dt, const ColliderSet* cols) {
  GRID_STRIDE(si, n) {
    int s = (int)si;
    if (labels[s] < 0) continue;
    float v[3], x[3];
    load_x(P, s, x);
#pragma unroll
    for (int q = 0; q < 3; ++q) v[q] = P.atu(C_V + q, s);
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    collide(*cols, v, x, zero, true);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      P.atu(C_V + q, s) = v[q];
      P.atu(C_X + q, s) = x[q] + v[q] * dt;
    }
  }
}
__global__ void k_probe(PView P, const int* slots, int n, int mode, const float* dF, float* out) {
  GRID_STRIDE(k, n) {
    int s = slots[k];
    if (s < 0) {  // rejected index (the call fails): the output stays defined
      const int w = mode == 0 ? 1 : 9;
      for (int c = 0; c < w; ++c) out[w * (int64_t)k + c] = __int_as_float(0x7fc00000);
      continue;
    }
    float F[9], R[9], Fi[9], Gi[9];
    load_m9(P, C_F, s, F);
    load_m9(P, C_R, s, R);
    load_m9(P, C_FI, s, Fi);
    load_m9(P, C_GI, s, Gi);
    float J = P.atu(C_J, s), mu = P.atu(C_MU, s), lam = P.atu(C_LAM, s);
    if (mode == 0) {
      out[k] = snow_energy(F, R, J, mu, lam);
    } else if (mode == 1) {
      snow_stress(F, R, Fi, J, mu, lam, out + 9 * (int64_t)k);
    } else {
      snow_stress_differential(dF + 9 * (int64_t)k, R, Fi, Gi, J, mu, lam, out + 9 * (int64_t)k);
    }
  }
}
__global__ void k_is_ballistic(const int* labels, int n, int* flag) {
  GRID_STRIDE(s, n) flag[s] = labels[s] == -1 ? 1 : 0;
}
__global__ void k_ballistic_compact(const int* labels, const int* pos, int n, int* out, int offset) {
  GRID_STRIDE(s, n) if (labels[s] == -1) out[offset + pos[s]] = (int)s;
}

}  // namespace

// ============================================================ launchers
#define LAUNCH(kernel, blocks, ...) hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), 0, s, __VA_ARGS__)

void launch_set_column(PView P, int col0, int ncomp, const float* aos, const int* slot_of, int n, hipStream_t s) {
  if (n > 0) LAUNCH(k_set_column, grid_for((int64_t)n * ncomp), P, col0, ncomp, aos, slot_of, n);
}
void launch_get_column(PView P, int col0, int ncomp, float* aos, const int* slot_of, int n, hipStream_t s) {
  if (n > 0) LAUNCH(k_get_column, grid_for((int64_t)n * ncomp), P, col0, ncomp, aos, slot_of, n);
}
void launch_fill_column(PView P, int col, float value, int n, hipStream_t s) {
  if (n > 0) LAUNCH(k_fill_column, grid_for(n), P, col, value, n);
}
void launch_fill_int(int* p, int value, int64_t n, hipStream_t s) {
  if (n > 0) LAUNCH(k_fill_int, grid_for(n), p, value, n);
}
void launch_iota(int* p, int64_t n, hipStream_t s) {
  if (n > 0) LAUNCH(k_iota, grid_for(n), p, n);
}
void launch_identity_columns(PView P, int col0, int n, hipStream_t s) {
  if (n > 0) LAUNCH(k_identity_columns, grid_for(n), P, col0, n);
}
void launch_permute(PView src, PView dst, const int* new_to_old, int n, const int* orig_of_old, int* orig_of_new,
                    int* slot_of, hipStream_t s) {
  if (n > 0) LAUNCH(k_permute, grid_for(n, 16384), src, dst, new_to_old, n, orig_of_old, orig_of_new, slot_of);
}
void launch_gather_slots(const int* slot_of, const int* orig, int n, int* slots, hipStream_t s) {
  if (n > 0) LAUNCH(k_gather_slots, grid_for(n), slot_of, orig, n, slots);
}
void launch_gather_slots_checked(const int* slot_of, const int* orig, int n, int np, int* slots, int* err,
                                 hipStream_t s) {
  if (n > 0) LAUNCH(k_gather_slots_checked, grid_for(n), slot_of, orig, n, np, slots, err);
}
void launch_cc_hash(PView P, int n, float cell, uint32_t mask, int* bucket, int* rank, int* count, hipStream_t s) {
  if (n > 0) LAUNCH(k_cc_hash, grid_for(n, 16384), P, n, cell, mask, bucket, rank, count);
}
void launch_cc_scatter(PView P, const int* bucket, const int* rank, const int* offset, int n, float4* sorted,
                       hipStream_t s) {
  if (n > 0) LAUNCH(k_cc_scatter, grid_for(n, 16384), P, bucket, rank, offset, n, sorted);
}
void launch_cc_link(int n, float h, float cell, uint32_t mask, const int* offset, const int* count,
                    const float4* sorted, int* parent, int* hasNb, int* snap, void* skip, int* links0, int* links1,
                    hipStream_t s) {
  if (n <= 0) return;
  CcSkip* S = static_cast<CcSkip*>(skip);
  const dim3 blocks((unsigned)((n + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_cc_sample, blocks, dim3(kBlock), 0, s, n, h, cell, mask, offset, count, sorted, parent, hasNb,
                     links0, links1);
  LAUNCH(k_cc_join, grid_for(n, 16384), n, sorted, (const int*)links0, (const int*)links1, parent);
  LAUNCH(k_cc_compress, grid_for(n, 16384), n, parent, snap);
  hipLaunchKernelGGL(k_cc_vote, dim3(1), dim3(kCcVote), 0, s, n, (const int*)parent, S);
  const int64_t perBlock = (int64_t)kBlock * kBoxPerThread;
  hipLaunchKernelGGL(k_cc_skip_boxes, dim3((unsigned)((n + perBlock - 1) / perBlock)), dim3(kBlock), 0, s, n, sorted,
                     (const int*)snap, S);
  hipLaunchKernelGGL(k_cc_skip_select, dim3(1), dim3(1), 0, s, h, S);
  LAUNCH(k_cc_link, grid_for(n, 16384), n, h, cell, mask, offset, count, sorted, parent, (const int*)snap,
         (const CcSkip*)S);
}
void launch_cc_roots(int n, int* parent, const int* hasNb, const int* orig_of, int* isRoot, int* minOrig,
                     hipStream_t s) {
  if (n > 0) LAUNCH(k_cc_roots, grid_for(n, 16384), n, parent, hasNb, orig_of, isRoot, minOrig);
}
void launch_cc_number(int n, const int* parent, const int* hasNb, const int* isRoot, const int* minOrig, int* mark,
                      int* markPos, int* nbodies, int* tmp, int* labels, hipStream_t s) {
  if (n <= 0) return;
  LAUNCH(k_fill_int, grid_for(n, 16384), mark, 0, (int64_t)n);
  LAUNCH(k_cc_mark, grid_for(n, 16384), n, isRoot, minOrig, mark);
  scan_exclusive(mark, markPos, n, nbodies, tmp, s);
  LAUNCH(k_cc_assign, grid_for(n, 16384), n, parent, hasNb, minOrig, (const int*)markPos, labels);
}
void launch_check_pvel(PView P, int n, int* err, hipStream_t s) {
  if (n > 0) LAUNCH(k_check_pvel, grid_for(n), P, n, err);
}
void launch_body_reduce(PView P, const int* slots, const int* labels, int n, int nb, unsigned int* bbox, double* mom,
                        int* count, unsigned int* bboxR, double* momR, int* countR, hipStream_t s) {
  if (nb <= 0) return;
  const int nrep = kBodyReplicas * nb;
  LAUNCH(k_body_init, grid_for(nrep), nrep, bboxR, momR, countR);
  if (n > 0) LAUNCH(k_body_reduce, grid_for(n, 512), P, slots, labels, n, nb, bboxR, momR, countR);
  LAUNCH(k_body_combine, grid_for(nb), nb, (const unsigned int*)bboxR, (const double*)momR, (const int*)countR, bbox,
         mom, count);
}
int body_replicas() { return kBodyReplicas; }
void launch_tile_body(const BodyGrid* bodies, const int* order, int nb, int NT, int* tileBody, int* blockBody,
                      hipStream_t s) {
  if (NT > 0) LAUNCH(k_tile_body, grid_for(NT), bodies, order, nb, NT, tileBody, blockBody);
}
void launch_bin(PView P, const int* slots, const int* labels, int n, const BodyGrid* bodies, int* key, int* rank,
                int* count, int* err, hipStream_t s) {
  if (n > 0) LAUNCH(k_bin, grid_for(n, 16384), P, slots, labels, n, bodies, key, rank, count, err);
}
void launch_decode_bin(const int* key, int n, const BodyGrid* bodies, const int* tileBody, int* cells, hipStream_t s) {
  if (n > 0) LAUNCH(k_decode_bin, grid_for(n), key, n, bodies, tileBody, cells);
}
void launch_scatter(const int* slots, const int* key, const int* rank, const int* offset, int n, int* perm,
                    int8_t* pcode, hipStream_t s) {
  if (n > 0) LAUNCH(k_scatter, grid_for(n, 16384), slots, key, rank, offset, n, perm, pcode);
}
void launch_chunk_counts(const int* count, int NT, int* nch, hipStream_t s) {
  if (NT > 0) LAUNCH(k_chunk_counts, grid_for(NT), count, NT, nch);
}
void launch_make_chunks(const int* count, const int* offset, const int* chunk_off, const int* tileBody, int NT,
                        int4* chunks, hipStream_t s) {
  if (NT > 0) LAUNCH(k_make_chunks, grid_for(NT), count, offset, chunk_off, tileBody, NT, chunks);
}
// per level: block sums (nb) + 8 pad + block offsets (nb) + 8 pad, then the next level
int64_t scan_tmp_ints(int64_t n) {
  int64_t tot = 0;
  while (n > 1) {
    n = cdiv(n, 1024);
    tot += 2 * n + 16;
  }
  return tot + 16;
}
void scan_exclusive(const int* in, int* out, int n, int* total, int* tmp, hipStream_t s) {
  int nb = cdiv(n, 1024);
  if (nb < 1) nb = 1;
  int* bsum = tmp;
  hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(kBlock), 0, s, in, out, n, bsum);
  if (nb == 1) {
    hipLaunchKernelGGL(k_copy_int, dim3(1), dim3(1), 0, s, (const int*)bsum, total);
    return;
  }
  int* boff = bsum + nb + 8;
  scan_exclusive(bsum, boff, nb, total, boff + nb + 8, s);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kBlock), 0, s, out, n, (const int*)boff);
}
void launch_to_dense_f(const float* src, int ncomp, float* dst, const BodyGrid& G, hipStream_t s) {
  LAUNCH(k_to_dense<float>, grid_for((int64_t)G.n[0] * G.n[1] * G.n[2]), src, ncomp, dst, G);
}
void launch_from_dense_f(const float* src, int ncomp, float* dst, const BodyGrid& G, hipStream_t s) {
  LAUNCH(k_from_dense<float>, grid_for((int64_t)G.n[0] * G.n[1] * G.n[2]), src, ncomp, dst, G);
}
void launch_to_dense_i8(const int8_t* src, int8_t* dst, const BodyGrid& G, hipStream_t s) {
  LAUNCH(k_to_dense<int8_t>, grid_for((int64_t)G.n[0] * G.n[1] * G.n[2]), src, 1, dst, G);
}
void launch_from_dense_i8(const int8_t* src, int8_t* dst, const BodyGrid& G, hipStream_t s) {
  LAUNCH(k_from_dense<int8_t>, grid_for((int64_t)G.n[0] * G.n[1] * G.n[2]), src, 1, dst, G);
}
void launch_p2g_mass(PView P, const Batch& B, float* stage, int* err, hipStream_t s) {
  if (B.nchunks > 0) LAUNCH(k_p2g_mass, B.nchunks, P, B, stage, err);
}
void launch_p2g_vel(PView P, const Batch& B, const float* mass, float* stage, int* err, hipStream_t s) {
  if (B.nchunks > 0) LAUNCH(k_p2g_vel, B.nchunks, P, B, mass, stage, err);
}
void launch_volumes(PView P, const Batch& B, const float* mass, int* err, hipStream_t s) {
  if (B.nchunks > 0) LAUNCH(k_volumes, B.nchunks, P, B, mass, err);
}
void launch_forces(PView P, const Batch& B, float* stage, int* err, hipStream_t s) {
  if (B.nchunks > 0) LAUNCH(k_forces, B.nchunks, P, B, stage, err);
}
void launch_hvp(PView P, const Batch& B, const float* d, bool project, const int8_t* flags, const ColliderSet* cols,
                float* stage, const CRScalars* sc, int* err, hipStream_t s) {
  if (B.nchunks <= 0) return;
  if (project)
    LAUNCH(k_hvp<true>, B.nchunks, P, B, d, flags, cols, stage, sc, err);
  else
    LAUNCH(k_hvp<false>, B.nchunks, P, B, d, flags, cols, stage, sc, err);
}
void launch_pair_chunks(const Batch& B, int* cnt, int* pos, int2* groups, int* ngroups, int* tmp, hipStream_t s) {
  const int npairs = (B.nchunks + 1) / 2;
  if (npairs <= 0) return;
  LAUNCH(k_pair_chunks, grid_for(npairs), B.chunks, B.nchunks, cnt);
  scan_exclusive(cnt, pos, npairs, ngroups, tmp, s);
  LAUNCH(k_pair_groups, grid_for(npairs), B.chunks, B.nchunks, (const int*)cnt, (const int*)pos, groups);
}
void launch_hvp_pair(PView P, const Batch& B, const int2* groups, const int* ngroups, const float* d, bool project,
                     const int8_t* flags, const ColliderSet* cols, float* stage, const CRScalars* sc, int* err,
                     hipStream_t s) {
  if (B.nchunks <= 0) return;
  if (project)
    LAUNCH(k_hvp_pair<true>, B.nchunks, P, B, groups, ngroups, d, flags, cols, stage, sc, err);
  else
    LAUNCH(k_hvp_pair<false>, B.nchunks, P, B, groups, ngroups, d, flags, cols, stage, sc, err);
}
void launch_diag(PView P, const Batch& B, float* stage, int* err, hipStream_t s) {
  if (B.nchunks > 0) LAUNCH(k_diag, B.nchunks, P, B, stage, err);
}
void launch_g2p(PView P, const Batch& B, const float* vel, const float* prev, int mode, float dt, float thc, float ths,
                const ColliderSet* cols, int* err, hipStream_t s) {
  if (B.nchunks <= 0) return;
  if (mode == 0)
    LAUNCH(k_g2p<0>, B.nchunks, P, B, vel, prev, dt, thc, ths, cols, err);
  else if (mode == 1)
    LAUNCH(k_g2p<1>, B.nchunks, P, B, vel, prev, dt, thc, ths, cols, err);
  else
    LAUNCH(k_g2p<2>, B.nchunks, P, B, vel, prev, dt, thc, ths, cols, err);
}
#define NODE_LAUNCH(kernel, ...) \
  if (B.nblocks > 0) hipLaunchKernelGGL(kernel, dim3(B.nblocks), dim3(kBlock), 0, s, __VA_ARGS__)
void launch_gather_mass(const Batch& B, const float* stage, float* mass, int* err, hipStream_t s) {
  NODE_LAUNCH(k_gather_mass, B, stage, mass, err);
}
void launch_gather_vel(const Batch& B, const float* stage, float* vel, float* prev, int* err, hipStream_t s) {
  NODE_LAUNCH(k_gather_vel, B, stage, vel, prev, err);
}
void launch_gather3(const Batch& B, const float* stage, float* out, hipStream_t s) {
  NODE_LAUNCH(k_gather3, B, stage, out);
}
void launch_node_forces(const Batch& B, const float* mass, const float* stage, const float* grav, int ngrav,
                        float* out, hipStream_t s) {
  NODE_LAUNCH(k_node_forces, B, mass, stage, grav, ngrav, out);
}
void launch_node_explicit(const Batch& B, const float* mass, const float* vel, const float* stage, const float* grav,
                          int ngrav, float dt, const ColliderSet* cols, float* em, int8_t* flags, float* vc,
                          CRScalars* sc, BatchStatus* st, double* partials, unsigned char* tileActive, float* nrm,
                          hipStream_t s) {
  NODE_LAUNCH(k_node_explicit, B, mass, vel, stage, grav, ngrav, dt, cols, em, flags, vc, sc, st, partials,
              tileActive, nrm);
}
void launch_node_normals(const Batch& B, const int8_t* flags, const ColliderSet* cols, float* nrm, hipStream_t s) {
  NODE_LAUNCH(k_node_normals, B, flags, cols, nrm);
}
void launch_collision_velocities(const Batch& B, const int8_t* flags, const ColliderSet* cols, float* vc,
                                 hipStream_t s) {
  NODE_LAUNCH(k_collision_velocities, B, flags, cols, vc);
}
void launch_rhs(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* stageVc,
                const float* stageDiag, float dt, float tol, float* b, float* x, float* diag, CRScalars* sc,
                BatchStatus* st, double* partials, hipStream_t s) {
  NODE_LAUNCH(k_rhs, B, mass, flags, cols, stageVc, stageDiag, dt, tol, b, x, diag, sc, st, partials);
}
void launch_cr_setup1(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                      const float* stage, const float* x, const float* b, const float* diag, float dt, float* z,
                      float* p, const CRScalars* sc, hipStream_t s) {
  NODE_LAUNCH(k_cr_setup1, B, mass, flags, cols, stage, x, b, diag, dt, z, p, sc);
}
void launch_cr_setup2(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                      const float* stage, const float* z, const float* diag, float dt, float* Ap, CRScalars* sc,
                      BatchStatus* st, double* partials, hipStream_t s) {
  NODE_LAUNCH(k_cr_setup2, B, mass, flags, cols, stage, z, diag, dt, Ap, sc, st, partials);
}
void launch_cr_update_x(const Batch& B, const int8_t* flags, const ColliderSet* cols, const float* p, const float* Ap,
                        const float* diag, float* x, float* z, CRScalars* sc, BatchStatus* st, double* partials,
                        double* hist, hipStream_t s) {
  NODE_LAUNCH(k_cr_update_x, B, flags, cols, p, Ap, diag, x, z, sc, st, partials, hist);
}
void launch_cr_ar(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* stage,
                  const float* z, const float* Ap, const float* diag, float dt, float* Ar, CRScalars* sc,
                  BatchStatus* st, double* partials, hipStream_t s) {
  NODE_LAUNCH(k_cr_ar, B, mass, flags, cols, stage, z, Ap, diag, dt, Ar, sc, st, partials);
}
void launch_cr_update_px(const Batch& B, const int8_t* flags, const ColliderSet* cols, const float* Ar,
                         const float* diag, float* p, float* Ap, float* x, float* z, CRScalars* sc, BatchStatus* st,
                         double* partials, double* hist, hipStream_t s) {
  NODE_LAUNCH(k_cr_update_px, B, flags, cols, Ar, diag, p, Ap, x, z, sc, st, partials, hist);
}
void launch_sort_tiles(const int* count, const int* offset, int* perm, const int8_t* pcode, int* tmpA, int* tmpB,
                       int NT, hipStream_t s) {
  if (NT > 0) hipLaunchKernelGGL(k_sort_tiles, dim3(NT), dim3(kBlock), 0, s, count, offset, perm, pcode, tmpA, tmpB);
}
void launch_small_claim(const int* list, int n, CRScalars* sc, BatchStatus* st, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_small_claim, dim3((n + 255) / 256), dim3(256), 0, s, list, n, sc, st);
}
void launch_small_cr(PView P, const Batch& B, const int* list, int n, const float* mass, const int8_t* flags,
                     const ColliderSet* cols, float dt, float* stage, const float* b, const float* diag, float* x,
                     float* z, float* p, float* Ap, float* Ar, CRScalars* sc, double* hist, int* err, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_small_cr, dim3(n), dim3(kBlock), 0, s, P, B, list, mass, flags, cols, dt, stage, b, diag, x,
                       z, p, Ap, Ar, sc, hist, err);
}
void launch_fcr_init(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                     const float* stage, const float* x, const float* b, const float* diag, float dt, float* r,
                     int netOn, double ratio, int mdirs, int strikes, CRScalars* sc, BatchStatus* st,
                     double* partials, hipStream_t s) {
  NODE_LAUNCH(k_fcr_init, B, mass, flags, cols, stage, x, b, diag, dt, r, netOn, ratio, mdirs, strikes, sc, st,
              partials);
}
void launch_fcr_diag_z(const Batch& B, const float* r, const float* diag, float* z, const CRScalars* sc,
                       hipStream_t s) {
  NODE_LAUNCH(k_fcr_diag_z, B, r, diag, z, sc);
}
void launch_fcr_az(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                   const float* stage, const float* z, const float* diag, float dt, float* Az, const float* Q,
                   int64_t vs, CRScalars* sc, BatchStatus* st, double* partials, hipStream_t s) {
  NODE_LAUNCH(k_fcr_az, B, mass, flags, cols, stage, z, diag, dt, Az, Q, vs, sc, st, partials);
}
void launch_fcr_orth(const Batch& B, const float* z, const float* Az, const float* r, const float* diag, float* P,
                     float* Q, int64_t vs, CRScalars* sc, BatchStatus* st, double* partials, hipStream_t s) {
  NODE_LAUNCH(k_fcr_orth, B, z, Az, r, diag, P, Q, vs, sc, st, partials);
}
void launch_fcr_update(const Batch& B, const int8_t* flags, const ColliderSet* cols, const float* mass,
                       const float* diag, const float* P, const float* Q, int64_t vs, float* x, float* r,
                       CRScalars* sc, BatchStatus* st, double* partials, double* hist, hipStream_t s) {
  NODE_LAUNCH(k_fcr_update, B, flags, cols, mass, diag, P, Q, vs, x, r, sc, st, partials, hist);
}
void launch_lpc_compose_z(const Batch& B, const float* r, const float* diag, const float* raw, const float* scale,
                          const float* weight, float* z, hipStream_t s) {
  NODE_LAUNCH(k_lpc_compose_z, B, r, diag, raw, scale, weight, z);
}
void launch_lpc_res_dots(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                         const float* stage, const float* z, const float* r, const float* diag, float dt,
                         const float* weight, float* u, double* dots, hipStream_t s) {
  NODE_LAUNCH(k_lpc_res_dots, B, mass, flags, cols, stage, z, r, diag, dt, weight, u, dots);
}
void launch_lpc_res_y(const Batch& B, const float* u, const float* r, const float* diag, const float* weight,
                      const double* dots, float* y, double* loss, hipStream_t s) {
  NODE_LAUNCH(k_lpc_res_y, B, u, r, diag, weight, dots, y, loss);
}
void launch_lpc_res_grad(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                         const float* stage, const float* y, const float* raw, float dt, const float* scale,
                         const float* weight, __bf16* g, hipStream_t s) {
  NODE_LAUNCH(k_lpc_res_grad, B, mass, flags, cols, stage, y, raw, dt, scale, weight, g);
}
void launch_lpc_exact_energy(const Batch& B, const float* x, const float* xk, const float* rk, double* energy,
                             hipStream_t s) {
  NODE_LAUNCH(k_lpc_exact_energy, B, x, xk, rk, energy);
}
void launch_lpc_energy_dots(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                            const float* stage, const float* z, const float* r, float dt, const float* weight,
                            float* u, double* dots, hipStream_t s) {
  NODE_LAUNCH(k_lpc_energy_dots, B, mass, flags, cols, stage, z, r, dt, weight, u, dots);
}
void launch_lpc_energy_grad(const Batch& B, const float* u, const float* r, const float* raw, const float* scale,
                            const float* weight, const double* dots, const double* energy, __bf16* g, double* loss,
                            hipStream_t s) {
  NODE_LAUNCH(k_lpc_energy_grad, B, u, r, raw, scale, weight, dots, energy, g, loss);
}
void launch_vec_dot(const float* a, const float* b, int64_t n, double* partials, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_vec_partial_dots, dim3(kDotBlocks), dim3(kBlock), 0, s, a, b, n, partials);
  hipLaunchKernelGGL(k_vec_sum_partials, dim3(1), dim3(kBlock), 0, s, partials, out);
}
void launch_vec_lincomb(float* out, float a, const float* u, float c, const float* v, int64_t n, hipStream_t s) {
  int64_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_vec_lincomb, dim3((unsigned)blocks), dim3(kBlock), 0, s, out, a, u, c, v, n);
}
void launch_residual_vector(const Batch& B, const float* z, const float* diag, float* r, hipStream_t s) {
  NODE_LAUNCH(k_residual_vector, B, z, diag, r);
}
void launch_finalize(const Batch& B, const float* x, const float* vc, float* vel, const CRScalars* sc,
                     hipStream_t s) {
  NODE_LAUNCH(k_finalize, B, x, vc, vel, sc);
}
void launch_plasticity(PView P, const int* labels, int skipLabel, int n, float thc, float ths, int* err,
                       hipStream_t s) {
  if (n > 0) LAUNCH(k_plasticity, grid_for(n, 16384), P, labels, skipLabel, n, thc, ths, err);
}
void launch_ballistic(PView P, const int* labels, int n, const float* grav, int ngrav, float dt,
                      const ColliderSet* cols, int mode, hipStream_t s) {
  if (n > 0) LAUNCH(k_ballistic, grid_for(n, 16384), P, labels, n, grav, ngrav, dt, cols, mode);
}
void launch_advect(PView P, const int* labels, int n, float dt, const ColliderSet* cols, hipStream_t s) {
  if (n > 0) LAUNCH(k_advect, grid_for(n, 16384), P, labels, n, dt, cols);
}
void launch_probe(PView P, const int* slots, int n, int mode, const float* dF, float* out, hipStream_t s) {
  if (n > 0) LAUNCH(k_probe, grid_for(n), P, slots, n, mode, dF, out);
}
void launch_is_ballistic(const int* labels, int n, int* flag, hipStream_t s) {
  if (n > 0) LAUNCH(k_is_ballistic, grid_for(n, 16384), labels, n, flag);
}
void launch_ballistic_compact(const int* labels, const int* pos, int n, int* out, int offset, hipStream_t s) {
  if (n > 0) LAUNCH(k_ballistic_compact, grid_for(n, 16384), labels, pos, n, out, offset);
}

}  // namespace mpm
