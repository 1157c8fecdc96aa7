// capi_mpnn.hip -- the warm-start network's C ABI (os_mpnn_*): parameters,
// forward pass (marginals), loss and gradient, and Adam training steps, all
// on the GPU over instances of a dense batch (os_dense).
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/onesolver.h"
#include "os_ctx.h"
#include "os_dense.h"
#include "os_kernels.h"

namespace os {

void mp_launch_features(hipStream_t st, int B, int n, int Np, const double *q, double *rowstat,
                        double *scale, uint16_t *A, float *X0, uint16_t *XT0);
void mp_launch_agg(hipStream_t st, int B, int Np, const uint16_t *A, const uint16_t *XT, float *Y,
                   bool accumulate);
void mp_launch_node_fwd(hipStream_t st, int B, int Np, const float *Hin, const float *M,
                        const float *wl, float *Z, float *H, uint16_t *HT);
void mp_launch_node_bwd(hipStream_t st, int B, int Np, const float *dH, const float *Z,
                        const float *wl, float *dZ, float *dHprev, uint16_t *dMsgT);
void mp_launch_head(hipStream_t st, int B, int n, int Np, const float *H3, const float *wo,
                    const uint8_t *labels, float inv_count, float *p, double *lossterm,
                    float *dlogit, float *dH3);
void mp_launch_wgrad(hipStream_t st, int rows, int W, const float *X, const float *G,
                     float *partial, float *grad_w, float *grad_b);
void mp_launch_loss_sum(hipStream_t st, size_t count, const double *terms, double *out);
void mp_launch_adam(hipStream_t st, int P, const float *grad, const uint8_t *kind, float *w,
                    float *m, float *v, float *wc, float lr, float b1, float b2, float eps,
                    int step);
void mp_launch_csr_features(hipStream_t st, int I, int max_n, const CsrInst *inst,
                            const int32_t *row_ptr, const double *val, const double *diag,
                            const int32_t *scale_rows, double *scale, float *X0);
void mp_launch_csr_agg(hipStream_t st, int I, int max_n, const CsrInst *inst,
                       const int32_t *row_ptr, const int32_t *col, const double *val,
                       const double *scale, const float *Hin, float *M);
CsrView csr_view(os_csr_batch *b, Ctx **ctx);
void mp_launch_compute_copy(hipStream_t st, int P, const uint8_t *kind, const float *w,
                            float *wc);

namespace {

constexpr int kF0 = 4, kHid = 64, kLayers = 3;
constexpr int kLayerStride = 2 * 64 * 64 + 64;         // Ws, Wm, b of one layer (padded)
constexpr int kWo = kLayers * kLayerStride;            // head weights
constexpr int kBo = kWo + 64;                          // head bias
constexpr int kP = kBo + 1;                            // padded parameter count
constexpr int kReal = 2 * kF0 * kHid + kHid + (kLayers - 1) * (2 * kHid * kHid + kHid) + kHid + 1;
constexpr uint32_t TAG_MPNN_INIT = 0x20;

// Visits (padded index, real index) of every real parameter in the packed
// order: per layer Ws [fin x 64], Wm [fin x 64], b [64]; then wo [64], bo.
template <typename F>
void for_each_param(F &&f) {
  int r = 0;
  for (int l = 0; l < kLayers; ++l) {
    const int fin = l == 0 ? kF0 : kHid;
    for (int part = 0; part < 2; ++part)
      for (int a = 0; a < fin; ++a)
        for (int o = 0; o < kHid; ++o) f(l * kLayerStride + part * 4096 + a * 64 + o, r++, 1, fin);
    for (int o = 0; o < kHid; ++o) f(l * kLayerStride + 8192 + o, r++, 2, fin);
  }
  for (int a = 0; a < kHid; ++a) f(kWo + a, r++, 1, kHid + 1);
  f(kBo, r++, 2, 0);
}

}  // namespace

struct Mpnn {
  Ctx *ctx = nullptr;
  int step = 0;
  DevBuf w, wc, grad, m, v, kind;
  int B = 0, Np = 0;
  DevBuf rowstat, scale, A, X0, XT0, HT[kLayers], M[kLayers], Z[kLayers], H[kLayers];
  DevBuf p, lossterm, dlogit, dH, dHprev, dZ, dMsgT, partial, lossout, labels;
  DevBuf srows;  // sparse batches: the rows of each instance's scale
  ~Mpnn() {
    for (auto *b : {&w, &wc, &grad, &m, &v, &kind, &rowstat, &scale, &A, &X0, &XT0, &p, &lossterm,
                    &dlogit, &dH, &dHprev, &dZ, &dMsgT, &partial, &lossout, &labels, &srows})
      b->release();
    for (int l = 0; l < kLayers; ++l)
      for (auto *b : {&HT[l], &M[l], &Z[l], &H[l]}) b->release();
  }

  void shape(int b, int np) {
    const size_t rows = (size_t)b * np;
    rowstat.ensure(sizeof(double) * 2 * rows);
    scale.ensure(sizeof(double) * b);
    A.ensure(sizeof(uint16_t) * rows * np);
    X0.ensure(sizeof(float) * rows * 64);
    XT0.ensure(sizeof(uint16_t) * rows * 64);
    for (int l = 0; l < kLayers; ++l) {
      HT[l].ensure(sizeof(uint16_t) * rows * 64);
      M[l].ensure(sizeof(float) * rows * 64);
      Z[l].ensure(sizeof(float) * rows * 64);
      H[l].ensure(sizeof(float) * rows * 64);
    }
    p.ensure(sizeof(float) * rows);
    lossterm.ensure(sizeof(double) * rows);
    dlogit.ensure(sizeof(float) * rows);
    dH.ensure(sizeof(float) * rows * 64);
    dHprev.ensure(sizeof(float) * rows * 64);
    dZ.ensure(sizeof(float) * rows * 64);
    dMsgT.ensure(sizeof(uint16_t) * rows * 64);
    partial.ensure(sizeof(float) * (rows / 64) * 65 * 64);
    lossout.ensure(sizeof(double));
    B = b;
    Np = np;
  }

  // forward pass over instances [first, first+count) of d; with labels, the
  // loss terms, dlogit and dH3 as well
  void forward(const Dense &d, int first, int count, const uint8_t *dev_labels) {
    const int n = d.n, np = (n + 63) / 64 * 64;
    shape(count, np);
    hipStream_t st = ctx->stream;
    const double *q = d.q64.as<double>() + (size_t)first * n * n;
    mp_launch_features(st, count, n, np, q, rowstat.as<double>(), scale.as<double>(),
                       A.as<uint16_t>(), X0.as<float>(), XT0.as<uint16_t>());
    for (int l = 0; l < kLayers; ++l) {
      const float *hin = l == 0 ? X0.as<float>() : H[l - 1].as<float>();
      const uint16_t *hint = l == 0 ? XT0.as<uint16_t>() : HT[l - 1].as<uint16_t>();
      mp_launch_agg(st, count, np, A.as<uint16_t>(), hint, M[l].as<float>(), false);
      mp_launch_node_fwd(st, count, np, hin, M[l].as<float>(), wc.as<float>() + l * kLayerStride,
                         Z[l].as<float>(), H[l].as<float>(),
                         l + 1 < kLayers ? HT[l].as<uint16_t>() : nullptr);
    }
    mp_launch_head(st, count, n, np, H[kLayers - 1].as<float>(), wc.as<float>() + kWo, dev_labels,
                   1.0f / ((float)count * (float)n), p.as<float>(), lossterm.as<double>(),
                   dlogit.as<float>(), dH.as<float>());
  }

  // forward pass over a sparse batch (marginals in p, one row per spin);
  // buffers for the rows only (no dense adjacency)
  void forward_csr(const CsrView &v, const int32_t *dev_scale_rows) {
    const size_t rows = (size_t)(v.total_spins + 63) / 64 * 64;
    hipStream_t st = ctx->stream;
    scale.ensure(sizeof(double) * std::max(v.num_instances, 1));
    X0.ensure(sizeof(float) * rows * 64);
    for (int l = 0; l < kLayers; ++l) {
      M[l].ensure(sizeof(float) * rows * 64);
      Z[l].ensure(sizeof(float) * rows * 64);
      H[l].ensure(sizeof(float) * rows * 64);
    }
    p.ensure(sizeof(float) * rows);
    // padding rows: zero features and messages, so every row the node
    // kernels read is defined
    OS_HIP_CHECK(hipMemsetAsync(X0.ptr, 0, sizeof(float) * rows * 64, st));
    for (int l = 0; l < kLayers; ++l)
      OS_HIP_CHECK(hipMemsetAsync(M[l].ptr, 0, sizeof(float) * rows * 64, st));
    mp_launch_csr_features(st, v.num_instances, v.max_n, v.inst, v.row_ptr, v.val, v.diag,
                           dev_scale_rows, scale.as<double>(), X0.as<float>());
    for (int l = 0; l < kLayers; ++l) {
      const float *hin = l == 0 ? X0.as<float>() : H[l - 1].as<float>();
      mp_launch_csr_agg(st, v.num_instances, v.max_n, v.inst, v.row_ptr, v.col, v.val,
                        scale.as<double>(), hin, M[l].as<float>());
      mp_launch_node_fwd(st, 1, (int)rows, hin, M[l].as<float>(),
                         wc.as<float>() + l * kLayerStride, Z[l].as<float>(), H[l].as<float>(),
                         nullptr);
    }
    mp_launch_head(st, 1, (int)rows, (int)rows, H[kLayers - 1].as<float>(),
                   wc.as<float>() + kWo, nullptr, 0.0f, p.as<float>(), nullptr, nullptr,
                   nullptr);
    // the next dense forward re-shapes its buffers from scratch
    B = 0;
    Np = 0;
  }

  // gradient of the mean loss into `grad` (padded layout); forward() with
  // labels must have run
  void backward() {
    hipStream_t st = ctx->stream;
    const int rows = B * Np;
    float *g = grad.as<float>();
    mp_launch_wgrad(st, rows, 1, H[kLayers - 1].as<float>(), dlogit.as<float>(),
                    partial.as<float>(), g + kWo, g + kBo);
    for (int l = kLayers - 1; l >= 0; --l) {
      const float *wl = wc.as<float>() + l * kLayerStride;
      mp_launch_node_bwd(st, B, Np, dH.as<float>(), Z[l].as<float>(), wl, dZ.as<float>(),
                         l > 0 ? dHprev.as<float>() : nullptr, dMsgT.as<uint16_t>());
      const float *hin = l == 0 ? X0.as<float>() : H[l - 1].as<float>();
      mp_launch_wgrad(st, rows, 64, hin, dZ.as<float>(), partial.as<float>(),
                      g + l * kLayerStride, g + l * kLayerStride + 8192);
      mp_launch_wgrad(st, rows, 64, M[l].as<float>(), dZ.as<float>(), partial.as<float>(),
                      g + l * kLayerStride + 4096, nullptr);
      if (l > 0) {
        mp_launch_agg(st, B, Np, A.as<uint16_t>(), dMsgT.as<uint16_t>(), dHprev.as<float>(), true);
        std::swap(dH, dHprev);
      }
    }
  }

  double loss_value(int count, int n) {
    mp_launch_loss_sum(ctx->stream, (size_t)B * Np, lossterm.as<double>(), lossout.as<double>());
    double s = 0.0;
    OS_HIP_CHECK(hipMemcpyAsync(&s, lossout.ptr, sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    ctx->sync();
    return s / ((double)count * n);
  }

  const uint8_t *upload_labels(const uint8_t *host, int count, int n) {
    labels.ensure((size_t)count * n);
    OS_HIP_CHECK(hipMemcpyAsync(labels.ptr, host, (size_t)count * n, hipMemcpyHostToDevice,
                                ctx->stream));
    return labels.as<uint8_t>();
  }
};

}  // namespace os

using os::Ctx;
using os::Dense;
using os::Mpnn;

namespace {

Mpnn &net_of(os_mpnn *h) {
  OS_REQUIRE(h, "null network");
  return *reinterpret_cast<Mpnn *>(h);
}

void check_range(const Dense &d, int first, int count) {
  OS_REQUIRE(count >= 1 && first >= 0 && first + count <= d.I, "instance range out of bounds");
  OS_REQUIRE(d.n >= 2, "instances need at least 2 spins");
}

}  // namespace

extern "C" {

int os_mpnn_num_params(void) { return os::kReal; }

int os_mpnn_create(os_ctx *ctx_, uint64_t seed, os_mpnn **out) {
  return os::guarded([&] {
    OS_REQUIRE(ctx_ && out, "null argument");
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    auto net = std::make_unique<Mpnn>();
    net->ctx = &c;
    std::vector<float> w(os::kP, 0.0f);
    std::vector<uint8_t> kind(os::kP, 0);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    os::for_each_param([&](int pi, int ri, int kd, int fan_in) {
      kind[pi] = (uint8_t)kd;
      if (kd == 1) {
        // Glorot-uniform from Philox counter (real index, 0, 0, init tag)
        const uint4 o = os::philox4x32_10(make_uint4((uint32_t)ri, 0u, 0u, os::TAG_MPNN_INIT), k0, k1);
        const double u = ((double)(o.x >> 8) + 0.5) * 0x1p-24;
        const int fan_out = fan_in == os::kHid + 1 ? 1 : os::kHid;
        const double lim = std::sqrt(6.0 / ((fan_in == os::kHid + 1 ? os::kHid : fan_in) + fan_out));
        w[pi] = (float)((2.0 * u - 1.0) * lim);
      }
    });
    for (auto *b : {&net->w, &net->wc, &net->grad, &net->m, &net->v})
      b->ensure(sizeof(float) * os::kP);
    net->kind.ensure(os::kP);
    hipStream_t st = c.stream;
    OS_HIP_CHECK(hipMemcpyAsync(net->w.ptr, w.data(), sizeof(float) * os::kP,
                                hipMemcpyHostToDevice, st));
    OS_HIP_CHECK(hipMemcpyAsync(net->kind.ptr, kind.data(), os::kP, hipMemcpyHostToDevice, st));
    OS_HIP_CHECK(hipMemsetAsync(net->m.ptr, 0, sizeof(float) * os::kP, st));
    OS_HIP_CHECK(hipMemsetAsync(net->v.ptr, 0, sizeof(float) * os::kP, st));
    OS_HIP_CHECK(hipMemsetAsync(net->grad.ptr, 0, sizeof(float) * os::kP, st));
    os::mp_launch_compute_copy(st, os::kP, net->kind.as<uint8_t>(), net->w.as<float>(),
                               net->wc.as<float>());
    c.sync();
    *out = reinterpret_cast<os_mpnn *>(net.release());
  });
}

int os_mpnn_destroy(os_mpnn *h) {
  return os::guarded([&] {
    if (!h) return;
    Mpnn *net = reinterpret_cast<Mpnn *>(h);
    net->ctx->activate();
    (void)hipStreamSynchronize(net->ctx->stream);
    delete net;
  });
}

int os_mpnn_get_params(os_mpnn *h, float *params) {
  return os::guarded([&] {
    Mpnn &net = net_of(h);
    OS_REQUIRE(params, "null argument");
    net.ctx->activate();
    std::vector<float> w(os::kP);
    OS_HIP_CHECK(hipMemcpyAsync(w.data(), net.w.ptr, sizeof(float) * os::kP,
                                hipMemcpyDeviceToHost, net.ctx->stream));
    net.ctx->sync();
    os::for_each_param([&](int pi, int ri, int, int) { params[ri] = w[pi]; });
  });
}

int os_mpnn_set_params(os_mpnn *h, const float *params) {
  return os::guarded([&] {
    Mpnn &net = net_of(h);
    OS_REQUIRE(params, "null argument");
    net.ctx->activate();
    std::vector<float> w(os::kP, 0.0f);
    os::for_each_param([&](int pi, int ri, int, int) { w[pi] = params[ri]; });
    hipStream_t st = net.ctx->stream;
    OS_HIP_CHECK(hipMemcpyAsync(net.w.ptr, w.data(), sizeof(float) * os::kP,
                                hipMemcpyHostToDevice, st));
    OS_HIP_CHECK(hipMemsetAsync(net.m.ptr, 0, sizeof(float) * os::kP, st));
    OS_HIP_CHECK(hipMemsetAsync(net.v.ptr, 0, sizeof(float) * os::kP, st));
    os::mp_launch_compute_copy(st, os::kP, net.kind.as<uint8_t>(), net.w.as<float>(),
                               net.wc.as<float>());
    net.step = 0;
    net.ctx->sync();
  });
}

int os_mpnn_predict(os_mpnn *h, os_dense *data, int first, int count, float *marginals) {
  return os::guarded([&] {
    Mpnn &net = net_of(h);
    OS_REQUIRE(data && marginals, "null argument");
    const Dense &d = *reinterpret_cast<Dense *>(data);
    check_range(d, first, count);
    net.ctx->activate();
    net.forward(d, first, count, nullptr);
    std::vector<float> p((size_t)count * net.Np);
    OS_HIP_CHECK(hipMemcpyAsync(p.data(), net.p.ptr, sizeof(float) * p.size(),
                                hipMemcpyDeviceToHost, net.ctx->stream));
    net.ctx->sync();
    for (int b = 0; b < count; ++b)
      std::memcpy(marginals + (size_t)b * d.n, p.data() + (size_t)b * net.Np,
                  sizeof(float) * d.n);
  });
}

int os_mpnn_predict_csr(os_mpnn *h, os_csr_batch *batch, const int32_t *scale_rows,
                        float *marginals) {
  return os::guarded([&] {
    Mpnn &net = net_of(h);
    OS_REQUIRE(batch && marginals, "null argument");
    Ctx *bctx = nullptr;
    const os::CsrView v = os::csr_view(batch, &bctx);
    OS_REQUIRE(bctx == net.ctx, "the network and the batch belong to different contexts");
    OS_REQUIRE(v.inst && v.row_ptr && v.diag && (v.col && v.val || v.total_spins == 0),
               "the batch's device arrays are missing");
    net.ctx->activate();
    const int32_t *dev_rows = nullptr;
    if (scale_rows) {
      for (int i = 0; i < v.num_instances; ++i)
        OS_REQUIRE(scale_rows[i] >= 1, "scale_rows must be positive");
      net.srows.ensure(sizeof(int32_t) * std::max(v.num_instances, 1));
      OS_HIP_CHECK(hipMemcpyAsync(net.srows.ptr, scale_rows, sizeof(int32_t) * v.num_instances,
                                  hipMemcpyHostToDevice, net.ctx->stream));
      dev_rows = net.srows.as<int32_t>();
    }
    net.forward_csr(v, dev_rows);
    OS_HIP_CHECK(hipMemcpyAsync(marginals, net.p.ptr, sizeof(float) * (size_t)v.total_spins,
                                hipMemcpyDeviceToHost, net.ctx->stream));
    net.ctx->sync();
  });
}

int os_mpnn_loss_grad(os_mpnn *h, os_dense *data, int first, int count, const uint8_t *labels,
                      double *loss, float *grad) {
  return os::guarded([&] {
    Mpnn &net = net_of(h);
    OS_REQUIRE(data && labels, "null argument");
    const Dense &d = *reinterpret_cast<Dense *>(data);
    check_range(d, first, count);
    net.ctx->activate();
    net.forward(d, first, count, net.upload_labels(labels, count, d.n));
    net.backward();
    const double l = net.loss_value(count, d.n);
    if (loss) *loss = l;
    if (grad) {
      std::vector<float> g(os::kP);
      OS_HIP_CHECK(hipMemcpyAsync(g.data(), net.grad.ptr, sizeof(float) * os::kP,
                                  hipMemcpyDeviceToHost, net.ctx->stream));
      net.ctx->sync();
      os::for_each_param([&](int pi, int ri, int, int) { grad[ri] = g[pi]; });
    }
  });
}

int os_mpnn_train_step(os_mpnn *h, os_dense *data, int first, int count, const uint8_t *labels,
                       float lr, double *loss) {
  return os::guarded([&] {
    Mpnn &net = net_of(h);
    OS_REQUIRE(data && labels, "null argument");
    OS_REQUIRE(lr > 0.0f, "learning rate must be positive");
    const Dense &d = *reinterpret_cast<Dense *>(data);
    check_range(d, first, count);
    net.ctx->activate();
    net.forward(d, first, count, net.upload_labels(labels, count, d.n));
    net.backward();
    ++net.step;
    os::mp_launch_adam(net.ctx->stream, os::kP, net.grad.as<float>(), net.kind.as<uint8_t>(),
                       net.w.as<float>(), net.m.as<float>(), net.v.as<float>(),
                       net.wc.as<float>(), lr, 0.9f, 0.999f, 1e-8f, net.step);
    const double l = net.loss_value(count, d.n);
    if (loss) *loss = l;
  });
}

}  // extern "C"
