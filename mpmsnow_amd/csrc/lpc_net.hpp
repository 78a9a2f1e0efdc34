// lpc_net.hpp -- the learned preconditioner's network: parameters, activations, forward pass
// (inference and training), backward pass and Adam step, all as launches of lpc_kernels.hip
// on the caller's stream. Included by the runtime.
//
// Network (a 3-level U-Net on the grid hierarchy of lpc_kernels.hpp; 3x3x3 convolutions, bf16
// weights and activations, fp32 accumulation):
//   in  (level 0, 16 ch, linear in q = r/diag/s: q (3), q m/diag (3), q collided (3), zeros)
//   e1 = relu(conv0(in))            level 0, 16 ch
//   e2 = relu(conv1(pool(e1)))      level 1, 32 ch
//   e3 = relu(conv2(pool(e2)))      level 2, 64 ch
//   d2 = relu(conv3(unpool(e3))) + e2   level 1, 32 ch
//   d1 = relu(conv4(unpool(d2))) + e1   level 0, 16 ch
//   out = conv5(d1), channels 0..2       level 0: correction / s
// The solver uses z = r / diag + s out: the network adds to the diagonal step what it misses.
// Training minimises sum_b w_b |out - (e - r / diag) / s|^2 over captured (residual, exact
// correction) pairs, w_b = 1 / |(e - r / diag) / s|^2 of body b. The output layer starts at zero,
// so an untrained network is the diagonal preconditioner. Training keeps the biases at zero: with
// inputs linear in r the network is then positively homogeneous in r, so its correction follows
// the residual from one iteration to the next instead of repeating a fixed shape.

#pragma once

#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "lpc_kernels.hpp"
#include "mpm_host.hpp"

namespace mpmh {

struct LpcLayer {
  int cin, cout, level, mode;
  int64_t w_off, b_off;  // offsets into the parameter vector
};

struct LpcLevels {
  std::vector<mpm::LevelGrid> host[3];
  DevBuf bodies[3], tileBody[3];
  int NT[3] = {0, 0, 0};
  mpm::LevelView view(int l, const mpm::Batch& B) const {
    mpm::LevelView v;
    v.bodies = bodies[l].as<mpm::LevelGrid>();
    v.tileBody = l == 0 ? B.tileBody : tileBody[l].as<int>();
    v.NT = NT[l];
    return v;
  }
  // level 0 = the batch's node tiles; levels 1, 2 halve each axis. order: the bodies in the
  // batch's tile layout order (ascending tileBase), also on the device (orderDev); levels 1, 2 are
  // laid out in the same order.
  void build(const std::vector<mpm::BodyGrid>& hb, const std::vector<int>& order, const int* orderDev,
             hipStream_t s) {
    const int nb = (int)hb.size();
    for (int l = 0; l < 3; ++l) host[l].assign(nb, mpm::LevelGrid{});
    int base[3] = {0, 0, 0};
    for (int b : order) {
      int n[3] = {hb[b].n[0], hb[b].n[1], hb[b].n[2]};
      for (int l = 0; l < 3; ++l) {
        mpm::LevelGrid& G = host[l][b];
        for (int a = 0; a < 3; ++a) {
          G.n[a] = n[a];
          G.T[a] = l == 0 ? hb[b].T[a] : (n[a] + 3) / 4;
        }
        G.tileBase = l == 0 ? hb[b].tileBase : base[l];
        G.NT = l == 0 ? hb[b].NT : G.T[0] * G.T[1] * G.T[2];
        base[l] = G.tileBase + G.NT;
        for (int a = 0; a < 3; ++a) n[a] = (n[a] + 1) / 2;
      }
    }
    for (int l = 0; l < 3; ++l) {
      NT[l] = base[l];
      bodies[l].ensure(sizeof(mpm::LevelGrid) * std::max(nb, 1));
      HIPCHECK(hipMemcpyAsync(bodies[l].p, host[l].data(), sizeof(mpm::LevelGrid) * nb, hipMemcpyHostToDevice, s));
      if (l > 0) {
        tileBody[l].ensure(sizeof(int) * std::max(NT[l], 1));
        mpm::launch_level_tile_body(bodies[l].as<mpm::LevelGrid>(), orderDev, nb, NT[l], tileBody[l].as<int>(), s);
      }
    }
    HIPCHECK(hipStreamSynchronize(s));
  }
};

class LpcNet {
 public:
  static constexpr int kLayers = 6;
  LpcLayer layer[kLayers];
  int64_t nparams = 0;
  bool initialised = false;
  int64_t adamStep = 0;

  DevBuf P, Pb, G, M, V, WT, partial;
  // activations (bf16) and training gradients
  DevBuf in, e1, p1, e2, p2, e3, u2, a2, d2, u1, a1, d1, raw;
  DevBuf g6, gd1, ga1, gu1, gd2, ga2, gu2, ge3, gp2, gp1, ge1, ge2;
  DevBuf loss;

  LpcNet() {
    const int spec[kLayers][4] = {{16, 16, 0, mpm::CONV_RELU}, {16, 32, 1, mpm::CONV_RELU},
                                  {32, 64, 2, mpm::CONV_RELU}, {64, 32, 1, mpm::CONV_RELU},
                                  {32, 16, 0, mpm::CONV_RELU}, {16, 16, 0, mpm::CONV_FINAL}};
    int64_t off = 0;
    for (int k = 0; k < kLayers; ++k) {
      LpcLayer& L = layer[k];
      L.cin = spec[k][0];
      L.cout = spec[k][1];
      L.level = spec[k][2];
      L.mode = spec[k][3];
      L.w_off = off;
      off += (int64_t)27 * L.cin * L.cout;
      L.b_off = off;
      off += L.cout;
    }
    nparams = off;
  }

  // He-normal initialisation (fan-in 27 cin); the output layer starts at zero, so an untrained
  // network leaves the diagonal step unchanged
  void init(uint64_t seed, hipStream_t s) {
    std::vector<float> h(nparams, 0.0f);
    std::mt19937_64 rng(seed);
    for (int k = 0; k < kLayers - 1; ++k) {
      const LpcLayer& L = layer[k];
      std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (27.0f * L.cin)));
      for (int64_t i = 0; i < (int64_t)27 * L.cin * L.cout; ++i) h[L.w_off + i] = nd(rng);
    }
    set(h.data(), s);
  }
  void set(const float* h, hipStream_t s) {
    alloc_params();
    HIPCHECK(hipMemcpyAsync(P.p, h, sizeof(float) * nparams, hipMemcpyHostToDevice, s));
    mpm::launch_f32_to_bf16(P.as<float>(), Pb.as<mpm::bf16>(), nparams, s);
    HIPCHECK(hipMemsetAsync(M.p, 0, sizeof(float) * nparams, s));
    HIPCHECK(hipMemsetAsync(V.p, 0, sizeof(float) * nparams, s));
    HIPCHECK(hipStreamSynchronize(s));
    adamStep = 0;
    initialised = true;
  }
  void get(float* h, hipStream_t s) {
    if (!initialised) throw InvalidArg("learned preconditioner has no weights");
    HIPCHECK(hipMemcpyAsync(h, P.p, sizeof(float) * nparams, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
  }

  void alloc_acts(const LpcLevels& Lv, bool training) {
    auto act = [&](DevBuf& d, int level, int c) { d.ensure(sizeof(mpm::bf16) * 64 * (size_t)std::max(Lv.NT[level], 1) * c); };
    act(in, 0, 16);
    act(e1, 0, 16);
    act(p1, 1, 16);
    act(e2, 1, 32);
    act(p2, 2, 32);
    act(e3, 2, 64);
    act(u2, 1, 64);
    act(a2, 1, 32);
    act(d2, 1, 32);
    act(u1, 0, 32);
    act(a1, 0, 16);
    act(d1, 0, 16);
    if (training) {
      raw.ensure(sizeof(float) * 192 * (size_t)std::max(Lv.NT[0], 1));
      act(g6, 0, 16);
      act(gd1, 0, 16);
      act(ga1, 0, 16);
      act(gu1, 0, 32);
      act(gd2, 1, 32);
      act(ga2, 1, 32);
      act(gu2, 1, 64);
      act(ge3, 2, 64);
      act(gp2, 2, 32);
      act(gp1, 1, 16);
      act(ge1, 0, 16);
      act(ge2, 1, 32);
      loss.ensure(2 * sizeof(double));
    }
  }

  // Forward pass. features: r, diag, mass; scale per body from `scale` (training) or
  // sc[b].lpc_scale (solver, which also skips done bodies and bodies on the diagonal
  // preconditioner). Output: raw (fp32 [tile][3][64], may be null) and/or z = s out.
  void forward(const mpm::Batch& B, const LpcLevels& Lv, const float* r, const float* diag, const float* mass,
               const int8_t* flags, const float* scale, const mpm::CRScalars* sc, float* rawOut, float* z,
               hipStream_t s) {
    using namespace mpm;
    if (!initialised) throw InvalidArg("learned preconditioner has no weights");
    const bf16* W = Pb.as<bf16>();
    const float* Pf = P.as<float>();
    auto Wl = [&](int k) { return W + layer[k].w_off; };
    auto bl = [&](int k) { return Pf + layer[k].b_off; };
    auto bfp = [](DevBuf& d) { return d.as<bf16>(); };
    const LevelView L0 = Lv.view(0, B), L1 = Lv.view(1, B), L2 = Lv.view(2, B);
    launch_lpc_features(B, r, diag, mass, flags, scale, sc, bfp(in), s);
    ConvOut o{};
    o.sc = sc;
    o.out = bfp(e1);
    launch_conv(16, 16, CONV_RELU, L0, bfp(in), Wl(0), bl(0), o, s);
    launch_pool(16, L0, L1, bfp(e1), bfp(p1), 0.125f, s);
    o.out = bfp(e2);
    launch_conv(16, 32, CONV_RELU, L1, bfp(p1), Wl(1), bl(1), o, s);
    launch_pool(32, L1, L2, bfp(e2), bfp(p2), 0.125f, s);
    o.out = bfp(e3);
    launch_conv(32, 64, CONV_RELU, L2, bfp(p2), Wl(2), bl(2), o, s);
    launch_unpool(64, L1, L2, bfp(e3), bfp(u2), 1.0f, false, s);
    o.out = bfp(a2);
    o.skip = bfp(e2);
    o.out2 = bfp(d2);
    launch_conv(64, 32, CONV_RELU, L1, bfp(u2), Wl(3), bl(3), o, s);
    launch_unpool(32, L0, L1, bfp(d2), bfp(u1), 1.0f, false, s);
    o.out = bfp(a1);
    o.skip = bfp(e1);
    o.out2 = bfp(d1);
    launch_conv(32, 16, CONV_RELU, L0, bfp(u1), Wl(4), bl(4), o, s);
    ConvOut f{};
    f.sc = sc;
    f.raw = rawOut;
    f.z = z;
    f.r = r;
    f.diag = diag;
    launch_conv(16, 16, CONV_FINAL, L0, bfp(d1), Wl(5), bl(5), f, s);
  }

  // weight and bias gradients of layer k from its input activation and output-side gradient
  void layer_grads(int k, const LpcLevels& Lv, const mpm::Batch& B, const mpm::bf16* input, const mpm::bf16* g,
                   hipStream_t s) {
    using namespace mpm;
    const LpcLayer& L = layer[k];
    const LevelView V = Lv.view(L.level, B);
    const int tpg = 64;
    const int groups = (V.NT + tpg - 1) / tpg;
    const int64_t wlen = (int64_t)27 * L.cin * L.cout;
    partial.ensure(sizeof(float) * (size_t)groups * wlen);
    launch_conv_dw(L.cin, L.cout, V, input, g, partial.as<float>(), tpg, s);
    launch_reduce_groups(partial.as<float>(), groups, wlen, G.as<float>() + L.w_off, s);
    launch_bias_partial(L.cout, V, g, partial.as<float>(), tpg, s);
    launch_reduce_groups(partial.as<float>(), groups, L.cout, G.as<float>() + L.b_off, s);
  }
  // data gradient through layer k: out = conv_k^T(g)
  void layer_dx(int k, const LpcLevels& Lv, const mpm::Batch& B, const mpm::bf16* g, mpm::bf16* out,
                hipStream_t s) {
    using namespace mpm;
    const LpcLayer& L = layer[k];
    bf16* wt = WT.as<bf16>() + L.w_off;
    launch_transpose_flip(Pb.as<bf16>() + L.w_off, wt, L.cin, L.cout, s);
    ConvOut o{};
    o.out = out;
    launch_conv(L.cout, L.cin, CONV_LINEAR, Lv.view(L.level, B), g, wt, nullptr, o, s);
  }

  // One gradient evaluation: forward on (r, diag, mass, scale), loss against target e, and the
  // parameter gradient in G. Returns the loss.
  double gradients(const mpm::Batch& B, const LpcLevels& Lv, const float* r, const float* diag, const float* mass,
                   const int8_t* flags, const float* e, const float* scale, const float* weight, hipStream_t s) {
    using namespace mpm;
    alloc_acts(Lv, true);
    G.ensure(sizeof(float) * nparams);
    WT.ensure(sizeof(bf16) * nparams);
    forward(B, Lv, r, diag, mass, flags, scale, nullptr, raw.as<float>(), nullptr, s);
    HIPCHECK(hipMemsetAsync(loss.p, 0, sizeof(double), s));
    launch_lpc_out_grad(B, raw.as<float>(), e, r, diag, scale, weight, g6.as<bf16>(), loss.as<double>(), s);
    backward(B, Lv, s);
    double h = 0.0;
    HIPCHECK(hipMemcpyAsync(&h, loss.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    return h;
  }

  // Parameter gradient in G from the output-side gradient in g6 and the activations of the last
  // training forward pass.
  void backward(const mpm::Batch& B, const LpcLevels& Lv, hipStream_t s) {
    using namespace mpm;
    G.ensure(sizeof(float) * nparams);
    WT.ensure(sizeof(bf16) * nparams);
    auto bfp = [](DevBuf& d) { return d.as<bf16>(); };
    const LevelView L0 = Lv.view(0, B), L1 = Lv.view(1, B), L2 = Lv.view(2, B);
    const int64_t n0 = (int64_t)64 * L0.NT, n1 = (int64_t)64 * L1.NT, n2 = (int64_t)64 * L2.NT;
    // layer 5: out = conv5(d1)
    layer_grads(5, Lv, B, bfp(d1), bfp(g6), s);
    layer_dx(5, Lv, B, bfp(g6), bfp(gd1), s);
    // d1 = a1 + e1
    HIPCHECK(hipMemcpyAsync(ge1.p, gd1.p, sizeof(bf16) * n0 * 16, hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipMemcpyAsync(ga1.p, gd1.p, sizeof(bf16) * n0 * 16, hipMemcpyDeviceToDevice, s));
    launch_relu_mask(bfp(ga1), bfp(a1), n0 * 16, s);
    // layer 4: a1 = relu(conv4(u1))
    layer_grads(4, Lv, B, bfp(u1), bfp(ga1), s);
    layer_dx(4, Lv, B, bfp(ga1), bfp(gu1), s);
    // u1 = unpool(d2): gradient sums the 8 children
    launch_pool(32, L0, L1, bfp(gu1), bfp(gd2), 1.0f, s);
    // d2 = a2 + e2
    HIPCHECK(hipMemcpyAsync(ge2.p, gd2.p, sizeof(bf16) * n1 * 32, hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipMemcpyAsync(ga2.p, gd2.p, sizeof(bf16) * n1 * 32, hipMemcpyDeviceToDevice, s));
    launch_relu_mask(bfp(ga2), bfp(a2), n1 * 32, s);
    // layer 3: a2 = relu(conv3(u2))
    layer_grads(3, Lv, B, bfp(u2), bfp(ga2), s);
    layer_dx(3, Lv, B, bfp(ga2), bfp(gu2), s);
    // u2 = unpool(e3)
    launch_pool(64, L1, L2, bfp(gu2), bfp(ge3), 1.0f, s);
    launch_relu_mask(bfp(ge3), bfp(e3), n2 * 64, s);
    // layer 2: e3 = relu(conv2(p2))
    layer_grads(2, Lv, B, bfp(p2), bfp(ge3), s);
    layer_dx(2, Lv, B, bfp(ge3), bfp(gp2), s);
    // p2 = pool(e2) / 8
    launch_unpool(32, L1, L2, bfp(gp2), bfp(ge2), 0.125f, true, s);
    launch_relu_mask(bfp(ge2), bfp(e2), n1 * 32, s);
    // layer 1: e2 = relu(conv1(p1))
    layer_grads(1, Lv, B, bfp(p1), bfp(ge2), s);
    layer_dx(1, Lv, B, bfp(ge2), bfp(gp1), s);
    // p1 = pool(e1) / 8
    launch_unpool(16, L0, L1, bfp(gp1), bfp(ge1), 0.125f, true, s);
    launch_relu_mask(bfp(ge1), bfp(e1), n0 * 16, s);
    // layer 0: e1 = relu(conv0(in))
    layer_grads(0, Lv, B, bfp(in), bfp(ge1), s);
  }

  // Adam step; the gradient is clipped to clipFactor times the running mean of its norm. Bias
  // gradients are dropped (biases stay at their value, zero after init).
  double gnormMean = 0.0;
  void adam(float lr, hipStream_t s, double clipFactor = 3.0) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int k = 0; k < kLayers; ++k)
      HIPCHECK(hipMemsetAsync(G.as<float>() + layer[k].b_off, 0, sizeof(float) * layer[k].cout, s));
    loss.ensure(2 * sizeof(double));
    double* gn = loss.as<double>() + 1;
    HIPCHECK(hipMemsetAsync(gn, 0, sizeof(double), s));
    mpm::launch_sumsq(G.as<float>(), nparams, gn, s);
    double h = 0.0;
    HIPCHECK(hipMemcpyAsync(&h, gn, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    const double norm = std::sqrt(h);
    if (!std::isfinite(norm)) return;  // skip a step with a non-finite gradient
    float gscale = 1.0f;
    if (gnormMean > 0.0 && norm > clipFactor * gnormMean) gscale = (float)(clipFactor * gnormMean / norm);
    gnormMean = gnormMean > 0.0 ? 0.95 * gnormMean + 0.05 * std::min(norm, clipFactor * gnormMean) : norm;
    ++adamStep;
    const float c1 = 1.0f - std::pow(b1, (float)adamStep), c2 = 1.0f - std::pow(b2, (float)adamStep);
    mpm::launch_adam(P.as<float>(), G.as<float>(), M.as<float>(), V.as<float>(), Pb.as<mpm::bf16>(), nparams, lr, b1,
                     b2, eps, c1, c2, gscale, s);
  }

 private:
  void alloc_params() {
    P.ensure(sizeof(float) * nparams);
    Pb.ensure(sizeof(mpm::bf16) * nparams);
    G.ensure(sizeof(float) * nparams);
    M.ensure(sizeof(float) * nparams);
    V.ensure(sizeof(float) * nparams);
    WT.ensure(sizeof(mpm::bf16) * nparams);
  }
};

}  // namespace mpmh
