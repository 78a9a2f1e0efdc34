// lpc_kernels.hpp -- kernels of the learned preconditioner (a small BF16 3-D convolutional
// network on MFMA, FP32 accumulation) that maps a grid residual to a velocity correction
// inside the flexible conjugate-residual solve. No counterpart in the reference: it attaches
// where ConjugateResiduals takes a preconditioner (ConjugateResiduals.h:20,
// ProceduralMatrix.h:16).
//
// Grid hierarchy. Level 1 is the grid-node space of a batch (4x4x4-node tiles, mpm_device.hpp).
// Level l+1 halves each axis of level l (voxel X averages voxels 2X, 2X+1 per axis). Every
// level is stored in 4x4x4-voxel tiles per body; activations are bf16 [tile][64 voxels][C]
// with the channel index fastest, local voxel (x&3) + 4 (y&3) + 16 (z&3). Voxels outside a
// body's extent at that level hold zero.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_device.hpp"

namespace mpm {

typedef __bf16 bf16;

// one body at one level of the hierarchy
struct LevelGrid {
  int n[3];      // voxels per axis
  int T[3];      // tiles per axis
  int tileBase;  // first tile of the body at this level
  int NT;        // tiles of the body (level 1: the batch's padded count)
};
struct LevelView {
  const LevelGrid* bodies;
  const int* tileBody;
  int NT;  // tiles of the level over all bodies
};

enum ConvMode : int {
  CONV_RELU = 0,    // out = relu(W * in + b) [+ skip -> out2]
  CONV_LINEAR = 1,  // out = W * in (+ b if given), no activation (backward data pass)
  CONV_FINAL = 2    // channels 0..2 -> fp32 solver-layout output
};

struct ConvOut {
  bf16* out;           // [tiles][64][COUT] (RELU, LINEAR)
  const bf16* skip;    // added after the activation when out2 is set
  bf16* out2;          // out + skip
  float* raw;          // FINAL: network output channels 0..2, [tile][3][64] (may be null)
  float* z;            // FINAL: preconditioned vector r / diag + scale out, [tile][3][64] (may be null)
  const float* r;      // FINAL: residual and diagonal (with z)
  const float* diag;
  const CRScalars* sc;  // per-body scalars: done bodies are skipped, scale and fallback read
};

// conv layers: (CIN, COUT) in {(16,16), (16,32), (32,64), (64,32), (32,16)}, FINAL (16,16)
void launch_conv(int cin, int cout, int mode, const LevelView& V, const bf16* in, const bf16* W, const float* bias,
                 const ConvOut& o, hipStream_t s);
// dW partials: partial[group][COUT][27 CIN] += sum over the group's voxels of g (x) in-shifted;
// `groups` = ceil(NT / tilesPerGroup)
void launch_conv_dw(int cin, int cout, const LevelView& V, const bf16* in, const bf16* g, float* partial,
                    int tilesPerGroup, hipStream_t s);
// out[k] = sum_g partial[g][k] over `groups` groups (fixed order), optional accumulate
void launch_reduce_groups(const float* partial, int groups, int64_t len, float* out, hipStream_t s);
// bias gradient partials: partial[group][c] = sum over the group's voxels of g[v][c]
void launch_bias_partial(int c, const LevelView& V, const bf16* g, float* partial, int tilesPerGroup, hipStream_t s);
// coarse = scale * sum of the 8 children (mean: scale = 1/8)
void launch_pool(int c, const LevelView& fine, const LevelView& coarse, const bf16* in, bf16* out, float scale,
                 hipStream_t s);
// fine (=, or += when accumulate) scale * parent value
void launch_unpool(int c, const LevelView& fine, const LevelView& coarse, const bf16* in, bf16* out, float scale,
                   bool accumulate, hipStream_t s);
// g = g * (a > 0), elementwise over n values
void launch_relu_mask(bf16* g, const bf16* a, int64_t n, hipStream_t s);
// out = a + b, elementwise
void launch_add_bf16(const bf16* a, const bf16* b, bf16* out, int64_t n, hipStream_t s);
// WT[ci][(26 - off) COUT + co] = W[co][off CIN + ci]  (weights of the transposed, flipped conv)
void launch_transpose_flip(const bf16* W, bf16* WT, int cin, int cout, hipStream_t s);
// Adam on fp32 parameters (gradient multiplied by gscale first); writes the bf16 copy
void launch_adam(float* p, const float* g, float* m, float* v, bf16* pb, int64_t n, float lr, float b1, float b2,
                 float eps, float c1, float c2, float gscale, hipStream_t s);
// *out += sum g^2
void launch_sumsq(const float* g, int64_t n, double* out, hipStream_t s);
void launch_f32_to_bf16(const float* in, bf16* out, int64_t n, hipStream_t s);
// per-level tile -> body table
void launch_level_tile_body(const LevelGrid* bodies, const int* order, int nb, int NT, int* tileBody, hipStream_t s);
// network input at level 0, every channel linear in q = clamp((r / diag) / scale): q (3),
// q m / diag (3), q [collision flag set] (3; flags may be null), zeros (7). With zero biases the
// network is then positively homogeneous in r. The per-body scale comes from `scale`
// when given, else from sc[b].lpc_scale (bodies that are done or use the diagonal
// preconditioner are skipped when sc is given)
void launch_lpc_features(const Batch& B, const float* r, const float* diag, const float* mass, const int8_t* flags,
                         const float* scale, const CRScalars* sc, bf16* in, hipStream_t s);
void launch_sub_f32(const float* a, const float* b, float* out, int64_t n, hipStream_t s);
// per-body sums of a captured residual: sums[3 b + {0, 1}] += (r / diag)^2, 3 x nodes over nodes
// with m > 0; sums[3 b + 2] += r^2 / diag over all nodes
void launch_lpc_sample_stats(const Batch& B, const float* r, const float* diag, const float* mass, double* sums,
                             hipStream_t s);
// training output gradient: g[v][c] = 2 w_b (raw - t) for c < 3 (0 elsewhere), target
// t = clamp((e - r / diag) / scale, +-kLpcClamp): the network learns the part of the exact
// correction e that the diagonal step r / diag misses
void launch_lpc_out_grad(const Batch& B, const float* raw, const float* e, const float* r, const float* diag,
                         const float* scale, const float* weight, bf16* g, double* loss, hipStream_t s);

}  // namespace mpm
