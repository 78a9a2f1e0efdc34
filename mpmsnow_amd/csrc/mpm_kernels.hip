// mpm_kernels.hip -- gfx950 kernels of the implicit MPM snow step.
//
// Particle-side kernels (P2G, volumes, forces, HVP, diagonal, G2P) run one
// workgroup per "chunk" (up to 256 particles whose stencil base node lies in
// one 4x4x4-node tile). The workgroup stages the 8-tile footprint of the grid
// vectors it reads in LDS, accumulates its scatter into an LDS footprint with
// ds_add_f32, and flushes the footprint to HBM with one global f32 atomic per
// touched node component (instead of 64 x 3 atomics per particle).
// Node-side kernels are grid-stride loops over the tiled node arrays; their
// reductions are FP64 and deterministic (per-block partials combined in a
// fixed order by the last block to finish).
//
// The arithmetic of each step follows the reference (davidsminor/mpmsnow,
// paths relative to its checkout) as cited per kernel.
#include "mpm_kernels.hpp"

namespace mpm {

namespace {

inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
inline int grid_for(int64_t n, int cap = 4096) {
  int b = cdiv(n, kBlock);
  return b < 1 ? 1 : (b > cap ? cap : b);
}

// ------------------------------------------------------------ footprints
__device__ __forceinline__ void tile_coords(const GridView& g, int t, int& tx, int& ty, int& tz) {
  tx = t % g.T[0];
  int r = t / g.T[0];
  ty = r % g.T[1];
  tz = r / g.T[1];
}
__device__ __forceinline__ int foot_tile(const GridView& g, int t, int f) {
  int tx, ty, tz;
  tile_coords(g, t, tx, ty, tz);
  tx += f & 1;
  ty += (f >> 1) & 1;
  tz += f >> 2;
  if (tx >= g.T[0] || ty >= g.T[1] || tz >= g.T[2]) return -1;
  return tx + g.T[0] * (ty + g.T[1] * tz);
}
// footprint-relative node (rx, ry, rz) in [0, 8) -> LDS index in [0, 512)
__device__ __forceinline__ int foot_index(int rx, int ry, int rz) {
  return ((((rz >> 2) << 2) | ((ry >> 2) << 1) | (rx >> 2)) << 6) | ((rz & 3) << 4) | ((ry & 3) << 2) | (rx & 3);
}
// node coordinates of tiled node index (tile, local)
__device__ __forceinline__ void node_coords(const GridView& g, int t, int l, int& i, int& j, int& k) {
  int tx, ty, tz;
  tile_coords(g, t, tx, ty, tz);
  i = tx * 4 + (l & 3);
  j = ty * 4 + ((l >> 2) & 3);
  k = tz * 4 + (l >> 4);
}
__device__ __forceinline__ bool node_valid(const GridView& g, int i, int j, int k) {
  return i < g.n[0] && j < g.n[1] && k < g.n[2];
}
__device__ __forceinline__ void node_pos(const GridView& g, int i, int j, int k, float x[3]) {
  x[0] = g.h * (float)i + g.gmin[0];
  x[1] = g.h * (float)j + g.gmin[1];
  x[2] = g.h * (float)k + g.gmin[2];
}
// ImplicitUpdateMatrix::subspaceProject for one node (Grid.cpp:778-818)
__device__ __forceinline__ void project(const ColliderSet* cols, int flag, const float x[3], float v[3]) {
  if (flag == -1) return;
  if (flag == -2) {
    v[0] = v[1] = v[2] = 0.0f;
    return;
  }
  float n[3];
  col_normal(cols->o[flag], x, n);
  float nd = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
  v[0] -= nd * n[0];
  v[1] -= nd * n[1];
  v[2] -= nd * n[2];
}

struct ChunkCtx {
  int4 ch;
  int ox, oy, oz;  // node origin of the chunk's tile
};
__device__ __forceinline__ ChunkCtx chunk_begin(const GridView& g, int* ft) {
  ChunkCtx c;
  c.ch = g.chunks[blockIdx.x];
  int tx, ty, tz;
  tile_coords(g, c.ch.x, tx, ty, tz);
  c.ox = tx * 4;
  c.oy = ty * 4;
  c.oz = tz * 4;
  if (threadIdx.x < 8) ft[threadIdx.x] = foot_tile(g, c.ch.x, threadIdx.x);
  return c;
}
__device__ __forceinline__ void load_x(const PView& P, int s, float x[3]) {
  x[0] = P.at(C_X, s);
  x[1] = P.at(C_X + 1, s);
  x[2] = P.at(C_X + 2, s);
}
__device__ __forceinline__ void load_m9(const PView& P, int col, int s, float m[9]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) m[i] = P.at(col + i, s);
}
__device__ __forceinline__ void store_m9(const PView& P, int col, int s, const float m[9]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) P.at(col + i, s) = m[i];
}
// the stencil of particle s relative to the chunk footprint; false if it lies outside
__device__ __forceinline__ bool chunk_stencil(const GridView& g, const ChunkCtx& c, const float x[3], bool derivs,
                                              Stencil& st, int& rx, int& ry, int& rz) {
  make_stencil(x, g.gmin, g.h, derivs, st);
  rx = st.base[0] - c.ox;
  ry = st.base[1] - c.oy;
  rz = st.base[2] - c.oz;
  return rx >= 0 && rx < 4 && ry >= 0 && ry < 4 && rz >= 0 && rz < 4;
}
// load a tiled 3-vector footprint into LDS (optionally projected)
__device__ __forceinline__ void load_foot3(const GridView& g, const ChunkCtx& c, const int* ft, const float* d,
                                           float (*sd)[kFoot], const int8_t* flags, const ColliderSet* cols,
                                           bool proj) {
  for (int i = threadIdx.x; i < kFoot; i += kBlock) {
    int f = i >> 6, l = i & 63;
    int gt = ft[f];
    float v[3] = {0.0f, 0.0f, 0.0f};
    if (gt >= 0) {
      int64_t b = (int64_t)gt * 192 + l;
      v[0] = d[b];
      v[1] = d[b + 64];
      v[2] = d[b + 128];
      if (proj) {
        int fl = flags[(int64_t)gt * 64 + l];
        if (fl != -1) {
          int ni = c.ox + (f & 1) * 4 + (l & 3), nj = c.oy + ((f >> 1) & 1) * 4 + ((l >> 2) & 3),
              nk = c.oz + (f >> 2) * 4 + (l >> 4);
          float xn[3];
          node_pos(g, ni, nj, nk, xn);
          project(cols, fl, xn, v);
        }
      }
    }
    sd[0][i] = v[0];
    sd[1][i] = v[1];
    sd[2][i] = v[2];
  }
}
__device__ __forceinline__ void flush_foot3(const int* ft, float (*so)[kFoot], float* out) {
  for (int i = threadIdx.x; i < kFoot; i += kBlock) {
    int gt = ft[i >> 6];
    if (gt < 0) continue;
    int64_t b = (int64_t)gt * 192 + (i & 63);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float v = so[c][i];
      if (v != 0.0f) atomicAdd(out + b + 64 * c, v);
    }
  }
}

// ------------------------------------------------------------ reductions
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
// Writes this block's partials; returns true in every thread of the last block to finish,
// with out[] = the fixed-order total over all blocks.
template <int NV>
__device__ bool grid_sum(double (&v)[NV], double* partials, unsigned int* ticket, double (&out)[NV]) {
  __shared__ double sm[kBlock / 64][NV];
  __shared__ bool amLast;
  block_sum<NV>(v, sm);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(int64_t)blockIdx.x * NV + k] = v[k];
    __threadfence();
    unsigned int t = atomicAdd(ticket, 1u);
    amLast = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!amLast) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock)
#pragma unroll
    for (int k = 0; k < NV; ++k)
      acc[k] += __hip_atomic_load(partials + (int64_t)b * NV + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  block_sum<NV>(acc, sm);
  __shared__ double res[NV];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) res[k] = acc[k];
    *ticket = 0u;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = res[k];
  return true;
}

// ------------------------------------------------------------ particles
__global__ void k_set_column(PView P, int col0, int ncomp, const float* aos, const int* slot_of, int n) {
  int64_t total = (int64_t)n * ncomp;
  for (int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
    int i = (int)(t / ncomp), c = (int)(t % ncomp);
    P.at(col0 + c, slot_of[i]) = aos[t];
  }
}
__global__ void k_get_column(PView P, int col0, int ncomp, float* aos, const int* slot_of, int n) {
  int64_t total = (int64_t)n * ncomp;
  for (int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
    int i = (int)(t / ncomp), c = (int)(t % ncomp);
    aos[t] = P.at(col0 + c, slot_of[i]);
  }
}
__global__ void k_fill_column(PView P, int col, float value, int n) {
  for (int s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) P.at(col, s) = value;
}
__global__ void k_identity_columns(PView P, int col0, int n) {
  for (int s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock)
#pragma unroll
    for (int i = 0; i < 9; ++i) P.at(col0 + i, s) = (i == 0 || i == 4 || i == 8) ? 1.0f : 0.0f;
}
__global__ void k_permute(PView src, PView dst, const int* new_to_old, int n, const int* orig_old, int* orig_new,
                          int* slot_of) {
  for (int s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) {
    int o = new_to_old[s];
#pragma unroll 4
    for (int c = 0; c < NCOL; ++c) dst.at(c, s) = src.at(c, o);
    int orig = orig_old[o];
    orig_new[s] = orig;
    slot_of[orig] = s;
  }
}
__global__ void k_gather_slots(const int* slot_of, const int* orig, int n, int* slots) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) slots[k] = slot_of[orig[k]];
}

__device__ __forceinline__ unsigned int ord_enc(float f) {
  unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
// bbox (Grid.cpp:177-191) and comoving-frame sums (Sim.cpp:84-93) of one body
__global__ void k_body_reduce(PView P, const int* slots, int base, int n, unsigned int* bbox, double* mom) {
  unsigned int mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    int s = slots ? slots[k] : base + k;
    float m = P.at(C_M, s);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      unsigned int e = ord_enc(P.at(C_X + a, s));
      mn[a] = min(mn[a], e);
      mx[a] = max(mx[a], e);
      acc[a] += (double)(P.at(C_V + a, s) * m);
    }
    acc[3] += (double)m;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
    for (int off = 32; off > 0; off >>= 1) {
      mn[a] = min(mn[a], (unsigned int)__shfl_down((int)mn[a], off, 64));
      mx[a] = max(mx[a], (unsigned int)__shfl_down((int)mx[a], off, 64));
    }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_down(acc[k], off, 64);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      atomicMin(bbox + a, mn[a]);
      atomicMax(bbox + 3 + a, mx[a]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) atomicAdd(mom + k, acc[k]);
  }
}

// ------------------------------------------------------------ binning
__global__ void k_bin(PView P, const int* slots, int base, int n, GridView g, int* key, int* rank, int* count,
                      int* err) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    int s = slots ? slots[k] : base + k;
    float x[3];
    load_x(P, s, x);
    int t3[3];
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float frac = (x[d] - g.gmin[d]) / g.h;
      int b = (int)floorf(frac) - 1;
      int t = b >> 2;
      ok = ok && b >= 0 && t < g.T[d];
      t3[d] = t;
    }
    if (!ok) {
      atomicOr(err, ERR_STENCIL);
      key[k] = -1;
      continue;
    }
    int t = t3[0] + g.T[0] * (t3[1] + g.T[1] * t3[2]);
    key[k] = t;
    rank[k] = atomicAdd(count + t, 1);
  }
}
__global__ void k_scatter(const int* slots, int base, const int* key, const int* rank, const int* offset, int n,
                          int* perm) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    int t = key[k];
    if (t < 0) continue;
    perm[offset[t] + rank[k]] = slots ? slots[k] : base + k;
  }
}
__global__ void k_chunk_counts(const int* count, int NT, int* nch) {
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < NT; t += gridDim.x * kBlock) nch[t] = (count[t] + kChunk - 1) / kChunk;
}
__global__ void k_make_chunks(const int* count, const int* offset, const int* chunk_off, int NT, int4* chunks) {
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < NT; t += gridDim.x * kBlock) {
    int c = count[t];
    int o = chunk_off[t];
    for (int q = 0; q * kChunk < c; ++q) {
      int cnt = c - q * kChunk;
      chunks[o + q] = make_int4(t, offset[t] + q * kChunk, cnt < kChunk ? cnt : kChunk, 0);
    }
  }
}

// exclusive scan: 1024 elements per block
__global__ void k_scan_block(const int* in, int* out, int n, int* bsum) {
  __shared__ int sm[kBlock];
  int b0 = blockIdx.x * 1024;
  int v[4];
  int tot = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int i = b0 + threadIdx.x * 4 + q;
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
    int i = b0 + threadIdx.x * 4 + q;
    if (i < n) out[i] = run;
    run += v[q];
  }
  if (threadIdx.x == kBlock - 1) bsum[blockIdx.x] = sm[kBlock - 1];
}
__global__ void k_scan_add(int* out, int n, const int* boff) {
  int i = blockIdx.x * 1024 + threadIdx.x;
  int add = boff[blockIdx.x];
#pragma unroll
  for (int q = 0; q < 4; ++q, i += kBlock)
    if (i < n) out[i] += add;
}
__global__ void k_copy_int(const int* src, int* dst) { *dst = *src; }

// ------------------------------------------------------------ dense <-> tiled
template <class T> __global__ void k_to_dense(const T* src, int ncomp, T* dst, GridView g) {
  int64_t nn = (int64_t)g.n[0] * g.n[1] * g.n[2];
  for (int64_t idx = blockIdx.x * (int64_t)kBlock + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * kBlock) {
    int i = (int)(idx % g.n[0]);
    int64_t r = idx / g.n[0];
    int j = (int)(r % g.n[1]);
    int k = (int)(r / g.n[1]);
    int t = (i >> 2) + g.T[0] * ((j >> 2) + g.T[1] * (k >> 2));
    int l = (i & 3) + 4 * (j & 3) + 16 * (k & 3);
    for (int c = 0; c < ncomp; ++c) {
      int64_t si = ncomp == 1 ? (int64_t)t * 64 + l : ((int64_t)t * 3 + c) * 64 + l;
      dst[idx * ncomp + c] = src[si];
    }
  }
}
template <class T> __global__ void k_from_dense(const T* src, int ncomp, T* dst, GridView g) {
  int64_t nn = (int64_t)g.n[0] * g.n[1] * g.n[2];
  for (int64_t idx = blockIdx.x * (int64_t)kBlock + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * kBlock) {
    int i = (int)(idx % g.n[0]);
    int64_t r = idx / g.n[0];
    int j = (int)(r % g.n[1]);
    int k = (int)(r / g.n[1]);
    int t = (i >> 2) + g.T[0] * ((j >> 2) + g.T[1] * (k >> 2));
    int l = (i & 3) + 4 * (j & 3) + 16 * (k & 3);
    for (int c = 0; c < ncomp; ++c) {
      int64_t si = ncomp == 1 ? (int64_t)t * 64 + l : ((int64_t)t * 3 + c) * 64 + l;
      dst[si] = src[idx * ncomp + c];
    }
  }
}

// ------------------------------------------------------------ P2G
// MassSplatter (Grid.cpp:62-99): m_i += m_p * w_ip
__global__ __launch_bounds__(kBlock) void k_p2g_mass(PView P, GridView g, float* mass, int* err) {
  __shared__ float acc[kFoot];
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(g, ft);
  for (int i = threadIdx.x; i < kFoot; i += kBlock) acc[i] = 0.0f;
  __syncthreads();
  for (int k = threadIdx.x; k < c.ch.z; k += kBlock) {
    int s = g.perm[c.ch.y + k];
    float x[3];
    load_x(P, s, x);
    float m = P.at(C_M, s);
    Stencil st;
    int rx, ry, rz;
    if (!chunk_stencil(g, c, x, false, st, rx, ry, rz)) {
      atomicOr(err, ERR_STENCIL);
      continue;
    }
#pragma unroll
    for (int cz = 0; cz < 4; ++cz)
#pragma unroll
      for (int cy = 0; cy < 4; ++cy)
#pragma unroll
        for (int cx = 0; cx < 4; ++cx) atomicAdd(&acc[foot_index(rx + cx, ry + cy, rz + cz)], m * st_w(st, cx, cy, cz));
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kFoot; i += kBlock) {
    float v = acc[i];
    int gt = ft[i >> 6];
    if (v != 0.0f && gt >= 0) atomicAdd(mass + (int64_t)gt * 64 + (i & 63), v);
  }
}
// clamp negative masses, NaN check (Grid.cpp:225-241)
__global__ void k_mass_fix(float* mass, int64_t nn, int* err) {
  int bad = 0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nn; i += (int64_t)gridDim.x * kBlock) {
    float m = mass[i];
    if (m < 0.0f) mass[i] = 0.0f;
    if (!isfinite(m)) bad = 1;
  }
  if (bad) atomicOr(err, ERR_NAN_MASS);
}
// VelocitySplatter (Grid.cpp:102-150): v_i += w (m_p / m_i) (v_p - v_f) where m_i > 0
__global__ __launch_bounds__(kBlock) void k_p2g_vel(PView P, GridView g, const float* mass, float* vel, int* err) {
  __shared__ float sm[kFoot];
  __shared__ float acc[3][kFoot];
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(g, ft);
  __syncthreads();
  for (int i = threadIdx.x; i < kFoot; i += kBlock) {
    int gt = ft[i >> 6];
    sm[i] = gt >= 0 ? mass[(int64_t)gt * 64 + (i & 63)] : 0.0f;
    acc[0][i] = acc[1][i] = acc[2][i] = 0.0f;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < c.ch.z; k += kBlock) {
    int s = g.perm[c.ch.y + k];
    float x[3];
    load_x(P, s, x);
    float m = P.at(C_M, s);
    float dv[3] = {P.at(C_V, s) - g.fv[0], P.at(C_V + 1, s) - g.fv[1], P.at(C_V + 2, s) - g.fv[2]};
    Stencil st;
    int rx, ry, rz;
    if (!chunk_stencil(g, c, x, false, st, rx, ry, rz)) {
      atomicOr(err, ERR_STENCIL);
      continue;
    }
#pragma unroll
    for (int cz = 0; cz < 4; ++cz)
#pragma unroll
      for (int cy = 0; cy < 4; ++cy)
#pragma unroll
        for (int cx = 0; cx < 4; ++cx) {
          int idx = foot_index(rx + cx, ry + cy, rz + cz);
          float M = sm[idx];
          if (M > 0.0f) {
            float sc = st_w(st, cx, cy, cz) * (m / M);
            atomicAdd(&acc[0][idx], sc * dv[0]);
            atomicAdd(&acc[1][idx], sc * dv[1]);
            atomicAdd(&acc[2][idx], sc * dv[2]);
          }
        }
  }
  __syncthreads();
  flush_foot3(ft, acc, vel);
}
// NaN check (Grid.cpp:250-260) and m_prevVelocities = m_velocities (Grid.cpp:262)
__global__ void k_check_vel(const float* vel, float* prev, int64_t n3, int* err) {
  int bad = 0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n3; i += (int64_t)gridDim.x * kBlock) {
    float v = vel[i];
    prev[i] = v;
    if (!isfinite(v)) bad = 1;
  }
  if (bad) atomicOr(err, ERR_NAN_VEL);
}
// computeParticleVolumes (Grid.cpp:265-292): V_p = m_p / sum_i w_ip m_i / h^3
__global__ __launch_bounds__(kBlock) void k_volumes(PView P, GridView g, const float* mass, int* err) {
  __shared__ float sm[kFoot];
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(g, ft);
  __syncthreads();
  for (int i = threadIdx.x; i < kFoot; i += kBlock) {
    int gt = ft[i >> 6];
    sm[i] = gt >= 0 ? mass[(int64_t)gt * 64 + (i & 63)] : 0.0f;
  }
  __syncthreads();
  const float cellVolume = g.h * g.h * g.h;
  for (int k = threadIdx.x; k < c.ch.z; k += kBlock) {
    int s = g.perm[c.ch.y + k];
    float x[3];
    load_x(P, s, x);
    Stencil st;
    int rx, ry, rz;
    if (!chunk_stencil(g, c, x, false, st, rx, ry, rz)) {
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
    P.at(C_VOL, s) = P.at(C_M, s) / density;
  }
}

// ------------------------------------------------------------ forces
// ForceSplatter (Grid.cpp:387-435): f_i -= V_p P_p F_p^T grad w_ip
__global__ __launch_bounds__(kBlock) void k_forces(PView P, GridView g, float* out, int* err) {
  __shared__ float acc[3][kFoot];
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(g, ft);
  for (int i = threadIdx.x; i < kFoot; i += kBlock) acc[0][i] = acc[1][i] = acc[2][i] = 0.0f;
  __syncthreads();
  for (int k = threadIdx.x; k < c.ch.z; k += kBlock) {
    int s = g.perm[c.ch.y + k];
    float x[3], F[9], R[9], Fi[9], Pm[9];
    load_x(P, s, x);
    load_m9(P, C_F, s, F);
    load_m9(P, C_R, s, R);
    load_m9(P, C_FI, s, Fi);
    snow_stress(F, R, Fi, P.at(C_J, s), P.at(C_MU, s), P.at(C_LAM, s), Pm);
    float n2 = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) n2 += Pm[i] * Pm[i];
    if (!isfinite(sqrtf(n2))) atomicOr(err, ERR_NAN_STRESS);
    float vol = P.at(C_VOL, s);
#pragma unroll
    for (int i = 0; i < 9; ++i) Pm[i] = vol * Pm[i];
    float fm[9];
    mat_mul_bt(Pm, F, fm);
    Stencil st;
    int rx, ry, rz;
    if (!chunk_stencil(g, c, x, true, st, rx, ry, rz)) {
      atomicOr(err, ERR_STENCIL);
      continue;
    }
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
          atomicAdd(&acc[0][idx], -t[0]);
          atomicAdd(&acc[1][idx], -t[1]);
          atomicAdd(&acc[2][idx], -t[2]);
        }
  }
  __syncthreads();
  flush_foot3(ft, acc, out);
}

__device__ __forceinline__ void field_force(const float* grav, int ngrav, float m, float f[3]) {
  for (int q = 0; q < ngrav; ++q) {
    f[0] += grav[3 * q] * m;
    f[1] += grav[3 * q + 1] * m;
    f[2] += grav[3 * q + 2] * m;
  }
}

// calculateForces output (Grid.cpp:438-463): f = sum_fields g m + internal
__global__ void k_node_forces(GridView g, const float* mass, const float* acc, const float* grav, int ngrav,
                              float* out) {
  int64_t nn = (int64_t)g.NT * 64;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nn; i += (int64_t)gridDim.x * kBlock) {
    int t = (int)(i >> 6), l = (int)(i & 63);
    int64_t b = (int64_t)t * 192 + l;
    float f[3] = {0.0f, 0.0f, 0.0f};
    field_force(grav, ngrav, mass[i], f);
    out[b] = f[0] + acc[b];
    out[b + 64] = f[1] + acc[b + 64];
    out[b + 128] = f[2] + acc[b + 128];
  }
}

// calculateExplicitMomenta (Grid.cpp:634-690) + collisionVelocities (Grid.cpp:692-724)
// + the sums of the initial guess (Grid.cpp:911-947)
__global__ void k_node_explicit(GridView g, const float* mass, const float* vel, const float* acc, const float* grav,
                                int ngrav, float dt, const ColliderSet* cols, float* em, int8_t* flags, float* vc,
                                CRScalars* sc, double* partials) {
  int64_t nn = (int64_t)g.NT * 64;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int bad = 0;
  int active = 0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nn; i += (int64_t)gridDim.x * kBlock) {
    int t = (int)(i >> 6), l = (int)(i & 63);
    int64_t b = (int64_t)t * 192 + l;
    int ni, nj, nk;
    node_coords(g, t, l, ni, nj, nk);
    if (!node_valid(g, ni, nj, nk)) {
      em[b] = em[b + 64] = em[b + 128] = 0.0f;
      vc[b] = vc[b + 64] = vc[b + 128] = 0.0f;
      flags[i] = -1;
      continue;
    }
    float m = mass[i];
    active += m > 0.0f ? 1 : 0;
    float f[3] = {0.0f, 0.0f, 0.0f};
    field_force(grav, ngrav, m, f);
    f[0] += acc[b];
    f[1] += acc[b + 64];
    f[2] += acc[b + 128];
    float ev[3] = {vel[b], vel[b + 64], vel[b + 128]};
    if (m > 0.0f) {
      ev[0] += dt * f[0] / m;
      ev[1] += dt * f[1] / m;
      ev[2] += dt * f[2] / m;
    }
    float x[3];
    node_pos(g, ni, nj, nk, x);
    int fl = collide(*cols, ev, x, g.fv, false);
    flags[i] = (int8_t)fl;
    float e0 = ev[0] * m, e1 = ev[1] * m, e2 = ev[2] * m;
    em[b] = e0;
    em[b + 64] = e1;
    em[b + 128] = e2;
    if (!isfinite(e0 * e1 * e2)) bad = 1;
    float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
    if (fl >= 0) {
      const Collider& o = cols->o[fl];
      c0 = o.v[0] - g.fv[0];
      c1 = o.v[1] - g.fv[1];
      c2 = o.v[2] - g.fv[2];
      s[4] += (double)m;
      s[5] += (double)(m * c0);
      s[6] += (double)(m * c1);
      s[7] += (double)(m * c2);
    }
    vc[b] = c0;
    vc[b + 64] = c1;
    vc[b + 128] = c2;
    s[0] += (double)m;
    s[1] += (double)e0;
    s[2] += (double)e1;
    s[3] += (double)e2;
  }
  if (bad) atomicOr(&sc->errors, ERR_NAN_MOMENTA);
  for (int off = 32; off > 0; off >>= 1) active += __shfl_down(active, off, 64);
  if ((threadIdx.x & 63) == 0 && active) atomicAdd(&sc->active_nodes, active);
  double tot[8];
  if (grid_sum<8>(s, partials, &sc->ticket[0], tot) && threadIdx.x == 0) {
    for (int k = 0; k < 8; ++k) sc->sums[k] = tot[k];
    float gv[3] = {0.0f, 0.0f, 0.0f};
    if (tot[0] == 0.0) {
    } else if (tot[4] == 0.0) {
      for (int a = 0; a < 3; ++a) gv[a] = (float)(tot[1 + a] / tot[0]);
    } else {
      for (int a = 0; a < 3; ++a) gv[a] = (float)(tot[5 + a] / tot[4]);
    }
    sc->guess[0] = gv[0];
    sc->guess[1] = gv[1];
    sc->guess[2] = gv[2];
  }
}

__global__ void k_collision_velocities(GridView g, const int8_t* flags, const ColliderSet* cols, float* vc) {
  int64_t nn = (int64_t)g.NT * 64;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nn; i += (int64_t)gridDim.x * kBlock) {
    int t = (int)(i >> 6), l = (int)(i & 63);
    int64_t b = (int64_t)t * 192 + l;
    int fl = flags[i];
    float c[3] = {0.0f, 0.0f, 0.0f};
    if (fl >= 0) {
      const Collider& o = cols->o[fl];
      for (int a = 0; a < 3; ++a) c[a] = o.v[a] - g.fv[a];
    }
    vc[b] = c[0];
    vc[b + 64] = c[1];
    vc[b + 128] = c[2];
  }
}

// ------------------------------------------------------------ HVP
// ForceDifferentialSplatter (Grid.cpp:549-618), factorised:
//   dF_p = (sum_j d_j grad w_jp^T) F_p ; df_i -= V_p dP(dF_p) F_p^T grad w_ip
template <bool PROJ>
__global__ __launch_bounds__(kBlock) void k_hvp(PView P, GridView g, const float* d, const int8_t* flags,
                                                const ColliderSet* cols, float* out, const CRScalars* sc, int* err) {
  if (sc && sc->done) return;
  __shared__ float sd[3][kFoot];
  __shared__ float so[3][kFoot];
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(g, ft);
  __syncthreads();
  load_foot3(g, c, ft, d, sd, flags, cols, PROJ);
  for (int i = threadIdx.x; i < kFoot; i += kBlock) so[0][i] = so[1][i] = so[2][i] = 0.0f;
  __syncthreads();
  for (int k = threadIdx.x; k < c.ch.z; k += kBlock) {
    int s = g.perm[c.ch.y + k];
    float x[3];
    load_x(P, s, x);
    Stencil st;
    int rx, ry, rz;
    if (!chunk_stencil(g, c, x, true, st, rx, ry, rz)) {
      atomicOr(err, ERR_STENCIL);
      continue;
    }
    float G[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) G[i] = 0.0f;
#pragma unroll
    for (int cz = 0; cz < 4; ++cz)
#pragma unroll
      for (int cy = 0; cy < 4; ++cy)
#pragma unroll
        for (int cx = 0; cx < 4; ++cx) {
          float gw[3];
          st_grad(st, cx, cy, cz, gw);
          int idx = foot_index(rx + cx, ry + cy, rz + cz);
          float d0 = sd[0][idx], d1 = sd[1][idx], d2 = sd[2][idx];
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            G[0 + 3 * cc] += d0 * gw[cc];
            G[1 + 3 * cc] += d1 * gw[cc];
            G[2 + 3 * cc] += d2 * gw[cc];
          }
        }
    float F[9], R[9], Fi[9], Gi[9], dF[9], dP[9], fm[9];
    load_m9(P, C_F, s, F);
    load_m9(P, C_R, s, R);
    load_m9(P, C_FI, s, Fi);
    load_m9(P, C_GI, s, Gi);
    mat_mul(G, F, dF);
    snow_stress_differential(dF, R, Fi, Gi, P.at(C_J, s), P.at(C_MU, s), P.at(C_LAM, s), dP);
    float vol = P.at(C_VOL, s);
#pragma unroll
    for (int i = 0; i < 9; ++i) dP[i] = vol * dP[i];
    mat_mul_bt(dP, F, fm);
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
          atomicAdd(&so[0][idx], -t[0]);
          atomicAdd(&so[1][idx], -t[1]);
          atomicAdd(&so[2][idx], -t[2]);
        }
  }
  __syncthreads();
  flush_foot3(ft, so, out);
}

// dFidXiSplatter (Grid.cpp:466-537) through the per-particle quadratic forms:
// with g = F^T grad w, the reference's entry [-V dP(e_a g^T) F^T grad w]_a equals
// -V g^T H_a g where H_a(c, d) = dP(e_a e_d^T)(a, c) (dP is linear in dF).
__global__ __launch_bounds__(kBlock) void k_diag(PView P, GridView g, float* out, int* err) {
  __shared__ float so[3][kFoot];
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(g, ft);
  for (int i = threadIdx.x; i < kFoot; i += kBlock) so[0][i] = so[1][i] = so[2][i] = 0.0f;
  __syncthreads();
  for (int k = threadIdx.x; k < c.ch.z; k += kBlock) {
    int s = g.perm[c.ch.y + k];
    float x[3];
    load_x(P, s, x);
    float F[9], R[9], Fi[9], Gi[9];
    load_m9(P, C_F, s, F);
    load_m9(P, C_R, s, R);
    load_m9(P, C_FI, s, Fi);
    load_m9(P, C_GI, s, Gi);
    float J = P.at(C_J, s), mu = P.at(C_MU, s), lam = P.at(C_LAM, s), vol = P.at(C_VOL, s);
    float H[3][9];  // H[a][c + 3 d]
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
        for (int cc = 0; cc < 3; ++cc) H[a][cc + 3 * dd] = dP[a + 3 * cc];
      }
    Stencil st;
    int rx, ry, rz;
    if (!chunk_stencil(g, c, x, true, st, rx, ry, rz)) {
      atomicOr(err, ERR_STENCIL);
      continue;
    }
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
            atomicAdd(&so[a][idx], -vol * q);
          }
        }
  }
  __syncthreads();
  flush_foot3(ft, so, out);
}

// ------------------------------------------------------------ solver node kernels
#define NODE_LOOP                                                                                   \
  int64_t nn = (int64_t)g.NT * 64;                                                                  \
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nn; i += (int64_t)gridDim.x * kBlock)

__device__ __forceinline__ void node_project(const GridView& g, const ColliderSet* cols, const int8_t* flags,
                                             int64_t i, float v[3]) {
  int fl = flags[i];
  if (fl == -1) return;
  int ni, nj, nk;
  node_coords(g, (int)(i >> 6), (int)(i & 63), ni, nj, nk);
  float x[3];
  node_pos(g, ni, nj, nk, x);
  project(cols, fl, x, v);
}
// ImplicitUpdateMatrix::multVector epilogue (Grid.cpp:764-770): P(m P d - dt^2 df)
__device__ __forceinline__ void apply_A(const GridView& g, const ColliderSet* cols, const int8_t* flags, int64_t i,
                                        float m, const float d[3], const float df[3], float dt2, float out[3]) {
  float pd[3] = {d[0], d[1], d[2]};
  node_project(g, cols, flags, i, pd);
  out[0] = m * pd[0] - dt2 * df[0];
  out[1] = m * pd[1] - dt2 * df[1];
  out[2] = m * pd[2] - dt2 * df[2];
  node_project(g, cols, flags, i, out);
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

// rhs (Grid.cpp:982-1002), initial guess fill (Grid.cpp:944-947), diagonal (Grid.cpp:847-866), |b|^2
__global__ void k_rhs(GridView g, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* accvc,
                      const float* draw, float dt, float tol, float* b, float* x, float* diag, CRScalars* sc,
                      double* partials) {
  double s[1] = {0.0};
  int bad = 0;
  const float dt2 = dt * dt;
  const float ndt2 = -dt * dt;
  const float g0 = sc->guess[0], g1 = sc->guess[1], g2 = sc->guess[2];
  NODE_LOOP {
    float df[3], bb[3], dr[3];
    ld3(accvc, i, df);
    node_project(g, cols, flags, i, df);
    ld3(b, i, bb);
    bb[0] += dt2 * df[0];
    bb[1] += dt2 * df[1];
    bb[2] += dt2 * df[2];
    if (!isfinite(bb[0] * bb[1] * bb[2])) bad = 1;
    st3(b, i, bb);
    float gx[3] = {g0, g1, g2};
    st3(x, i, gx);
    ld3(draw, i, dr);
    float m = mass[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      dr[a] = dr[a] * ndt2 + m;
      if (dr[a] == 0.0f) dr[a] = 1.0f;
    }
    st3(diag, i, dr);
    s[0] += (double)bb[0] * bb[0] + (double)bb[1] * bb[1] + (double)bb[2] * bb[2];
  }
  if (bad) atomicOr(&sc->errors, ERR_NAN_MOMENTA);
  double tot[1];
  if (grid_sum<1>(s, partials, &sc->ticket[1], tot) && threadIdx.x == 0) {
    sc->bnorm2 = tot[0];
    sc->threshold = (double)tol * (double)tol * tot[0];
    sc->iter = 0;
    sc->iterations = 0;
    sc->early_exit = 0;
    sc->zero_x = 0;
    sc->done = 0;
    if (sc->errors & ERR_NAN_MOMENTA) sc->done = 1;
    if (tot[0] == 0.0) {
      sc->done = 1;
      sc->zero_x = 1;
    }
  }
}

// CR setup part 1 (ConjugateResiduals.cpp:50-67): r = b - A x ; z = M^-1 r ; p = z
__global__ void k_cr_setup1(GridView g, const float* mass, const int8_t* flags, const ColliderSet* cols,
                            const float* acc, const float* x, const float* b, const float* diag, float dt, float* z,
                            float* p, const CRScalars* sc) {
  if (sc->done) return;
  const float dt2 = dt * dt;
  NODE_LOOP {
    float xv[3], df[3], Ax[3], bb[3], dg[3], zz[3];
    ld3(x, i, xv);
    ld3(acc, i, df);
    apply_A(g, cols, flags, i, mass[i], xv, df, dt2, Ax);
    ld3(b, i, bb);
    ld3(diag, i, dg);
#pragma unroll
    for (int a = 0; a < 3; ++a) zz[a] = (bb[a] - Ax[a]) / dg[a];
    st3(z, i, zz);
    st3(p, i, zz);
  }
}
// CR setup part 2 (ConjugateResiduals.cpp:69-76 and the first alpha, :86-102):
// Ap = Ar = A z ; rz = <z, Ar> ; alpha = rz / <Ap, M^-1 Ap>
__global__ void k_cr_setup2(GridView g, const float* mass, const int8_t* flags, const ColliderSet* cols,
                            const float* acc, const float* z, const float* diag, float dt, float* Ap, CRScalars* sc,
                            double* partials) {
  if (sc->done) return;
  const float dt2 = dt * dt;
  double s[2] = {0.0, 0.0};
  NODE_LOOP {
    float zz[3], df[3], Ar[3], dg[3];
    ld3(z, i, zz);
    ld3(acc, i, df);
    apply_A(g, cols, flags, i, mass[i], zz, df, dt2, Ar);
    st3(Ap, i, Ar);
    ld3(diag, i, dg);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      s[0] += (double)zz[a] * Ar[a];
      s[1] += (double)Ar[a] * (Ar[a] / dg[a]);
    }
  }
  double tot[2];
  if (grid_sum<2>(s, partials, &sc->ticket[2], tot) && threadIdx.x == 0) {
    sc->rz = tot[0];
    sc->ApdPAp = tot[1];
    if (tot[1] == 0.0) {
      sc->done = 1;
      sc->early_exit = 1;
    } else {
      sc->alpha = tot[0] / tot[1];
    }
    sc->iter = 0;
  }
}
// iteration part 1 (ConjugateResiduals.cpp:104-128): x = P(x + alpha p) ; z -= alpha M^-1 Ap ;
// r = M z ; termination on |r|^2 (SquareMagnitudeTermination.cpp:19-28)
__global__ void k_cr_update_x(GridView g, const int8_t* flags, const ColliderSet* cols, const float* p,
                              const float* Ap, const float* diag, float* x, float* z, CRScalars* sc, double* partials,
                              double* hist, int hist_cap) {
  if (sc->done) return;
  const float alpha = (float)sc->alpha;
  double s[1] = {0.0};
  NODE_LOOP {
    float xv[3], pv[3], ap[3], dg[3], zz[3];
    ld3(x, i, xv);
    ld3(p, i, pv);
#pragma unroll
    for (int a = 0; a < 3; ++a) xv[a] = xv[a] + alpha * pv[a];
    node_project(g, cols, flags, i, xv);
    st3(x, i, xv);
    ld3(Ap, i, ap);
    ld3(diag, i, dg);
    ld3(z, i, zz);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      zz[a] = zz[a] - alpha * (ap[a] / dg[a]);
      float r = zz[a] * dg[a];
      s[0] += (double)r * r;
    }
    st3(z, i, zz);
  }
  double tot[1];
  if (grid_sum<1>(s, partials, &sc->ticket[0], tot) && threadIdx.x == 0) {
    int it = sc->iter;
    sc->rnorm2 = tot[0];
    sc->iterations = it + 1;
    if (hist && it < hist_cap) hist[it] = tot[0];
    if (it >= sc->max_iters || tot[0] < sc->threshold) sc->done = 1;
  }
}
// iteration part 2 (ConjugateResiduals.cpp:130-145): Ar = A z ; rz = <z, Ar> ; beta
__global__ void k_cr_ar(GridView g, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* acc,
                        const float* z, float dt, float* Ar, CRScalars* sc, double* partials) {
  if (sc->done) return;
  const float dt2 = dt * dt;
  double s[1] = {0.0};
  NODE_LOOP {
    float zz[3], df[3], ar[3];
    ld3(z, i, zz);
    ld3(acc, i, df);
    apply_A(g, cols, flags, i, mass[i], zz, df, dt2, ar);
    st3(Ar, i, ar);
    s[0] += (double)zz[0] * ar[0] + (double)zz[1] * ar[1] + (double)zz[2] * ar[2];
  }
  double tot[1];
  if (grid_sum<1>(s, partials, &sc->ticket[1], tot) && threadIdx.x == 0) {
    double rz_old = sc->rz;
    sc->rz = tot[0];
    if (rz_old == 0.0) {
      sc->done = 1;
      sc->early_exit = 1;
    } else {
      sc->beta = tot[0] / rz_old;
    }
  }
}
// iteration part 3 (ConjugateResiduals.cpp:147-151 and the next alpha, :86-102)
__global__ void k_cr_update_p(GridView g, const float* z, const float* Ar, const float* diag, float* p, float* Ap,
                              CRScalars* sc, double* partials) {
  if (sc->done) return;
  const float beta = (float)sc->beta;
  double s[1] = {0.0};
  NODE_LOOP {
    float zz[3], pv[3], ar[3], ap[3], dg[3];
    ld3(z, i, zz);
    ld3(p, i, pv);
    ld3(Ar, i, ar);
    ld3(Ap, i, ap);
    ld3(diag, i, dg);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      pv[a] = zz[a] + beta * pv[a];
      ap[a] = ar[a] + beta * ap[a];
      s[0] += (double)ap[a] * (ap[a] / dg[a]);
    }
    st3(p, i, pv);
    st3(Ap, i, ap);
  }
  double tot[1];
  if (grid_sum<1>(s, partials, &sc->ticket[2], tot) && threadIdx.x == 0) {
    sc->ApdPAp = tot[0];
    if (tot[0] == 0.0) {
      sc->done = 1;
      sc->early_exit = 1;
    } else {
      sc->alpha = sc->rz / tot[0];
    }
    sc->iter = sc->iter + 1;
  }
}
__global__ void k_residual_vector(GridView g, const float* z, const float* diag, float* r) {
  NODE_LOOP {
    float zz[3], dg[3];
    ld3(z, i, zz);
    ld3(diag, i, dg);
    zz[0] *= dg[0];
    zz[1] *= dg[1];
    zz[2] *= dg[2];
    st3(r, i, zz);
  }
}
// m_velocities += vc (Grid.cpp:1014); x = 0 when the rhs was zero
__global__ void k_finalize(GridView g, const float* x, const float* vc, float* vel, const CRScalars* sc) {
  const bool zx = sc->zero_x != 0;
  NODE_LOOP {
    float xv[3], cv[3];
    ld3(x, i, xv);
    ld3(vc, i, cv);
    if (zx) xv[0] = xv[1] = xv[2] = 0.0f;
    xv[0] += cv[0];
    xv[1] += cv[1];
    xv[2] += cv[2];
    st3(vel, i, xv);
  }
}

// ------------------------------------------------------------ G2P
// updateParticleVelocities (Grid.cpp:1049-1081), updateDeformationGradients (Grid.cpp:1018-1047),
// and for the full step SnowConstitutiveModel::updateParticleData (SnowConstitutiveModel.cpp:27-121)
// plus the advection of Sim.cpp:123-131.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_g2p(PView P, GridView g, const float* vel, const float* prev, float dt,
                                                float thc, float ths, const ColliderSet* cols, int* err) {
  constexpr bool DO_V = MODE != 1;
  constexpr bool DO_F = MODE != 0;
  __shared__ float sv[3][kFoot];
  __shared__ float sp[DO_V ? 3 : 1][kFoot];
  __shared__ int ft[8];
  ChunkCtx c = chunk_begin(g, ft);
  __syncthreads();
  for (int i = threadIdx.x; i < kFoot; i += kBlock) {
    int gt = ft[i >> 6];
    int64_t b = (int64_t)gt * 192 + (i & 63);
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
    int s = g.perm[c.ch.y + k];
    float x[3];
    load_x(P, s, x);
    Stencil st;
    int rx, ry, rz;
    if (!chunk_stencil(g, c, x, DO_F, st, rx, ry, rz)) {
      atomicOr(err, ERR_STENCIL);
      continue;
    }
    float vf[3] = {0.0f, 0.0f, 0.0f}, vp[3] = {0.0f, 0.0f, 0.0f};
    if (DO_V) {
      vf[0] = P.at(C_V, s);
      vf[1] = P.at(C_V + 1, s);
      vf[2] = P.at(C_V + 2, s);
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
              vp[a] += w * (vi[a] + g.fv[a]);
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
      if (MODE != 2) {
#pragma unroll
        for (int a = 0; a < 3; ++a) P.at(C_V + a, s) = v[a];
      }
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
        bool ok = snow_update(Fn, Fp, Fi, R, Gi, J, mu, lam, P.at(C_MU0, s), P.at(C_LAM0, s), P.at(C_XI, s), thc, ths);
        if (!ok) atomicOr(err, ERR_HARDNESS);
        store_m9(P, C_F, s, Fn);
        store_m9(P, C_FP, s, Fp);
        store_m9(P, C_FI, s, Fi);
        store_m9(P, C_R, s, R);
        store_m9(P, C_GI, s, Gi);
        P.at(C_J, s) = J;
        P.at(C_MU, s) = mu;
        P.at(C_LAM, s) = lam;
        if (MODE == 2) {
          // advection (Sim.cpp:123-131)
          const float zero[3] = {0.0f, 0.0f, 0.0f};
          collide(*cols, v, x, zero, true);
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            P.at(C_V + a, s) = v[a];
            P.at(C_X + a, s) = x[a] + v[a] * dt;
          }
        }
      }
    }
  }
}

__global__ void k_plasticity(PView P, const int* slots, int base, int n, float thc, float ths, int* err) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    int s = slots ? slots[k] : base + k;
    float F[9], Fp[9], Fi[9], R[9], Gi[9], J, mu, lam;
    load_m9(P, C_F, s, F);
    load_m9(P, C_FP, s, Fp);
    bool ok = snow_update(F, Fp, Fi, R, Gi, J, mu, lam, P.at(C_MU0, s), P.at(C_LAM0, s), P.at(C_XI, s), thc, ths);
    if (!ok) atomicOr(err, ERR_HARDNESS);
    store_m9(P, C_F, s, F);
    store_m9(P, C_FP, s, Fp);
    store_m9(P, C_FI, s, Fi);
    store_m9(P, C_R, s, R);
    store_m9(P, C_GI, s, Gi);
    P.at(C_J, s) = J;
    P.at(C_MU, s) = mu;
    P.at(C_LAM, s) = lam;
  }
}

// ballistic particles (Sim.cpp:64-73) and/or their advection (Sim.cpp:123-131)
__global__ void k_ballistic(PView P, const int* slots, int n, const float* grav, int ngrav, float dt,
                            const ColliderSet* cols, int mode) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    int s = slots[k];
    float v[3], x[3];
    load_x(P, s, x);
#pragma unroll
    for (int q = 0; q < 3; ++q) v[q] = P.at(C_V + q, s);
    if (mode & 1) {
      float m = P.at(C_M, s);
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
      for (int q = 0; q < 3; ++q) P.at(C_X + q, s) = x[q] + v[q] * dt;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) P.at(C_V + q, s) = v[q];
  }
}
__global__ void k_advect(PView P, const int* slots, int n, float dt, const ColliderSet* cols) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    int s = slots[k];
    float v[3], x[3];
    load_x(P, s, x);
#pragma unroll
    for (int q = 0; q < 3; ++q) v[q] = P.at(C_V + q, s);
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    collide(*cols, v, x, zero, true);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      P.at(C_V + q, s) = v[q];
      P.at(C_X + q, s) = x[q] + v[q] * dt;
    }
  }
}

__global__ void k_probe(PView P, const int* slots, int n, int mode, const float* dF, float* out) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    int s = slots[k];
    float F[9], R[9], Fi[9], Gi[9];
    load_m9(P, C_F, s, F);
    load_m9(P, C_R, s, R);
    load_m9(P, C_FI, s, Fi);
    load_m9(P, C_GI, s, Gi);
    float J = P.at(C_J, s), mu = P.at(C_MU, s), lam = P.at(C_LAM, s);
    if (mode == 0) {
      out[k] = snow_energy(F, R, J, mu, lam);
    } else if (mode == 1) {
      snow_stress(F, R, Fi, J, mu, lam, out + 9 * (int64_t)k);
    } else {
      snow_stress_differential(dF + 9 * (int64_t)k, R, Fi, Gi, J, mu, lam, out + 9 * (int64_t)k);
    }
  }
}
// velocity finiteness (Sim.cpp:151-159)
__global__ void k_check_pvel(PView P, int n, int* err) {
  int bad = 0;
  for (int s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock)
    if (!isfinite(P.at(C_V, s) * P.at(C_V + 1, s) * P.at(C_V + 2, s))) bad = 1;
  if (bad) atomicOr(err, ERR_NAN_PVEL);
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
void launch_identity_columns(PView P, int col0, int n, hipStream_t s) {
  if (n > 0) LAUNCH(k_identity_columns, grid_for(n), P, col0, n);
}
void launch_permute(PView src, PView dst, const int* new_to_old, int n, const int* orig_of_old, int* orig_of_new,
                    int* slot_of, hipStream_t s) {
  if (n > 0) LAUNCH(k_permute, grid_for(n, 8192), src, dst, new_to_old, n, orig_of_old, orig_of_new, slot_of);
}
void launch_gather_slots(const int* slot_of, const int* orig, int n, int* slots, hipStream_t s) {
  if (n > 0) LAUNCH(k_gather_slots, grid_for(n), slot_of, orig, n, slots);
}
void launch_body_reduce(PView P, const int* slots, int base, int n, unsigned int* bbox6, double* mom4, hipStream_t s) {
  if (n > 0) LAUNCH(k_body_reduce, grid_for(n, 2048), P, slots, base, n, bbox6, mom4);
}
void launch_bin(PView P, const int* slots, int base, int n, const GridView& g, int* key, int* rank, int* count, int* err,
                hipStream_t s) {
  if (n > 0) LAUNCH(k_bin, grid_for(n, 8192), P, slots, base, n, g, key, rank, count, err);
}
void launch_scatter(const int* slots, int base, const int* key, const int* rank, const int* offset, int n, int* perm,
                    hipStream_t s) {
  if (n > 0) LAUNCH(k_scatter, grid_for(n, 8192), slots, base, key, rank, offset, n, perm);
}
void launch_chunk_counts(const int* count, int NT, int* nch, hipStream_t s) {
  LAUNCH(k_chunk_counts, grid_for(NT), count, NT, nch);
}
void launch_make_chunks(const int* count, const int* offset, const int* chunk_off, int NT, int4* chunks, hipStream_t s) {
  LAUNCH(k_make_chunks, grid_for(NT), count, offset, chunk_off, NT, chunks);
}
// per level: block sums (nb) + 8 pad + block offsets (nb) + 8 pad, then the next level
int64_t scan_tmp_ints(int n) {
  int64_t tot = 0;
  while (n > 1) {
    n = cdiv(n, 1024);
    tot += 2 * (int64_t)n + 16;
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
void launch_to_dense_f(const float* src, int ncomp, float* dst, const GridView& g, hipStream_t s) {
  LAUNCH(k_to_dense<float>, grid_for((int64_t)g.n[0] * g.n[1] * g.n[2]), src, ncomp, dst, g);
}
void launch_from_dense_f(const float* src, int ncomp, float* dst, const GridView& g, hipStream_t s) {
  LAUNCH(k_from_dense<float>, grid_for((int64_t)g.n[0] * g.n[1] * g.n[2]), src, ncomp, dst, g);
}
void launch_to_dense_i8(const int8_t* src, int8_t* dst, const GridView& g, hipStream_t s) {
  LAUNCH(k_to_dense<int8_t>, grid_for((int64_t)g.n[0] * g.n[1] * g.n[2]), src, 1, dst, g);
}
void launch_from_dense_i8(const int8_t* src, int8_t* dst, const GridView& g, hipStream_t s) {
  LAUNCH(k_from_dense<int8_t>, grid_for((int64_t)g.n[0] * g.n[1] * g.n[2]), src, 1, dst, g);
}
void launch_p2g_mass(PView P, const GridView& g, float* mass, int* err, hipStream_t s) {
  if (g.nchunks > 0) LAUNCH(k_p2g_mass, g.nchunks, P, g, mass, err);
}
void launch_mass_fix(float* mass, const GridView& g, int* err, hipStream_t s) {
  LAUNCH(k_mass_fix, grid_for((int64_t)g.NT * 64), mass, (int64_t)g.NT * 64, err);
}
void launch_p2g_vel(PView P, const GridView& g, const float* mass, float* vel, int* err, hipStream_t s) {
  if (g.nchunks > 0) LAUNCH(k_p2g_vel, g.nchunks, P, g, mass, vel, err);
}
void launch_check_vel(const float* vel, float* prev, const GridView& g, int* err, hipStream_t s) {
  LAUNCH(k_check_vel, grid_for((int64_t)g.NT * 192), vel, prev, (int64_t)g.NT * 192, err);
}
void launch_volumes(PView P, const GridView& g, const float* mass, int* err, hipStream_t s) {
  if (g.nchunks > 0) LAUNCH(k_volumes, g.nchunks, P, g, mass, err);
}
void launch_forces(PView P, const GridView& g, float* acc, int* err, hipStream_t s) {
  if (g.nchunks > 0) LAUNCH(k_forces, g.nchunks, P, g, acc, err);
}
void launch_node_forces(const GridView& g, const float* mass, const float* acc, const float* grav, int ngrav,
                        float* out, hipStream_t s) {
  LAUNCH(k_node_forces, grid_for((int64_t)g.NT * 64), g, mass, acc, grav, ngrav, out);
}
void launch_node_explicit(const GridView& g, const float* mass, const float* vel, const float* acc, const float* grav,
                          int ngrav, float dt, const ColliderSet* cols, float* em, int8_t* flags, float* vc,
                          CRScalars* sc, double* partials, hipStream_t s) {
  LAUNCH(k_node_explicit, grid_for((int64_t)g.NT * 64), g, mass, vel, acc, grav, ngrav, dt, cols, em, flags, vc, sc,
         partials);
}
void launch_collision_velocities(const GridView& g, const int8_t* flags, const ColliderSet* cols, float* vc,
                                 hipStream_t s) {
  LAUNCH(k_collision_velocities, grid_for((int64_t)g.NT * 64), g, flags, cols, vc);
}
void launch_hvp(PView P, const GridView& g, const float* d, bool project, const int8_t* flags, const ColliderSet* cols,
                float* out, const CRScalars* sc, int* err, hipStream_t s) {
  if (g.nchunks <= 0) return;
  if (project)
    LAUNCH(k_hvp<true>, g.nchunks, P, g, d, flags, cols, out, sc, err);
  else
    LAUNCH(k_hvp<false>, g.nchunks, P, g, d, flags, cols, out, sc, err);
}
void launch_diag(PView P, const GridView& g, float* out, int* err, hipStream_t s) {
  if (g.nchunks > 0) LAUNCH(k_diag, g.nchunks, P, g, out, err);
}
void launch_rhs(const GridView& g, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* accvc,
                const float* draw, float dt, float tol, float* b, float* x, float* diag, CRScalars* sc,
                double* partials, hipStream_t s) {
  LAUNCH(k_rhs, grid_for((int64_t)g.NT * 64), g, mass, flags, cols, accvc, draw, dt, tol, b, x, diag, sc, partials);
}
void launch_cr_setup1(const GridView& g, const float* mass, const int8_t* flags, const ColliderSet* cols,
                      const float* acc, const float* x, const float* b, const float* diag, float dt, float* z, float* p,
                      const CRScalars* sc, hipStream_t s) {
  LAUNCH(k_cr_setup1, grid_for((int64_t)g.NT * 64), g, mass, flags, cols, acc, x, b, diag, dt, z, p, sc);
}
void launch_cr_setup2(const GridView& g, const float* mass, const int8_t* flags, const ColliderSet* cols,
                      const float* acc, const float* z, const float* diag, float dt, float* Ap, CRScalars* sc,
                      double* partials, hipStream_t s) {
  LAUNCH(k_cr_setup2, grid_for((int64_t)g.NT * 64), g, mass, flags, cols, acc, z, diag, dt, Ap, sc, partials);
}
void launch_cr_update_x(const GridView& g, const int8_t* flags, const ColliderSet* cols, const float* p,
                        const float* Ap, const float* diag, float* x, float* z, CRScalars* sc, double* partials,
                        double* hist, int hist_cap, hipStream_t s) {
  LAUNCH(k_cr_update_x, grid_for((int64_t)g.NT * 64), g, flags, cols, p, Ap, diag, x, z, sc, partials, hist, hist_cap);
}
void launch_cr_ar(const GridView& g, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* acc,
                  const float* z, float dt, float* Ar, CRScalars* sc, double* partials, hipStream_t s) {
  LAUNCH(k_cr_ar, grid_for((int64_t)g.NT * 64), g, mass, flags, cols, acc, z, dt, Ar, sc, partials);
}
void launch_cr_update_p(const GridView& g, const float* z, const float* Ar, const float* diag, float* p, float* Ap,
                        CRScalars* sc, double* partials, hipStream_t s) {
  LAUNCH(k_cr_update_p, grid_for((int64_t)g.NT * 64), g, z, Ar, diag, p, Ap, sc, partials);
}
void launch_residual_vector(const GridView& g, const float* z, const float* diag, float* r, hipStream_t s) {
  LAUNCH(k_residual_vector, grid_for((int64_t)g.NT * 64), g, z, diag, r);
}
void launch_finalize(const GridView& g, const float* x, const float* vc, float* vel, const CRScalars* sc,
                     hipStream_t s) {
  LAUNCH(k_finalize, grid_for((int64_t)g.NT * 64), g, x, vc, vel, sc);
}
void launch_g2p(PView P, const GridView& g, const float* vel, const float* prev, int mode, float dt, float thc,
                float ths, const ColliderSet* cols, int* err, hipStream_t s) {
  if (g.nchunks <= 0) return;
  if (mode == 0)
    LAUNCH(k_g2p<0>, g.nchunks, P, g, vel, prev, dt, thc, ths, cols, err);
  else if (mode == 1)
    LAUNCH(k_g2p<1>, g.nchunks, P, g, vel, prev, dt, thc, ths, cols, err);
  else if (mode == 2)
    LAUNCH(k_g2p<2>, g.nchunks, P, g, vel, prev, dt, thc, ths, cols, err);
  else
    LAUNCH(k_g2p<3>, g.nchunks, P, g, vel, prev, dt, thc, ths, cols, err);
}
void launch_plasticity(PView P, const int* slots, int base, int n, float thc, float ths, int* err, hipStream_t s) {
  if (n > 0) LAUNCH(k_plasticity, grid_for(n, 8192), P, slots, base, n, thc, ths, err);
}
void launch_ballistic(PView P, const int* slots, int n, const float* grav, int ngrav, float dt, const ColliderSet* cols,
                      int mode, hipStream_t s) {
  if (n > 0) LAUNCH(k_ballistic, grid_for(n), P, slots, n, grav, ngrav, dt, cols, mode);
}
void launch_advect(PView P, const int* slots, int n, float dt, const ColliderSet* cols, hipStream_t s) {
  if (n > 0) LAUNCH(k_advect, grid_for(n), P, slots, n, dt, cols);
}
void launch_probe(PView P, const int* slots, int n, int mode, const float* dF, float* out, hipStream_t s) {
  if (n > 0) LAUNCH(k_probe, grid_for(n), P, slots, n, mode, dF, out);
}
void launch_check_pvel(PView P, int n, int* err, hipStream_t s) {
  if (n > 0) LAUNCH(k_check_pvel, grid_for(n), P, n, err);
}

}  // namespace mpm
