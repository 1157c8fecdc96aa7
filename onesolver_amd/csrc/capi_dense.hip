// capi_dense.hip -- dense sweep-mode annealing entry points (os_dense_*).
//
// A dense problem keeps Q resident on the GPU in three forms:
//   q64    FP64 N x N (field refresh, energy verification),
//   qpack  bf16 [N/16 blocks][Np columns][32]: for block b and column j the
//          16 values hi(Q[16b+k][j]) then the 16 values lo(Q[16b+k][j]),
//          Q = Qhi + Qlo + O(2^-17 |Q|), diagonal zeroed (the MFMA B operand),
//   qdb    fp32 [N/16][16][16] diagonal sub-blocks (in-block corrections).
// A run keeps per-replica fields (fp32 R x Np), states and best states
// (bit-packed), running and best energies (FP64) on the GPU; sweeps run in
// segments of `refresh_interval`, after each of which the fields and energies
// are recomputed in FP64 and the largest fp32 field error is recorded.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/onesolver.h"
#include "os_ctx.h"
#include "os_kernels.h"

namespace os {

uint16_t f32_to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

float bf16_to_f32(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

struct Dense {
  Ctx *ctx = nullptr;
  int n = 0, np = 0, npw = 0;
  DevBuf q64, qpack, qdb;
  DevBuf fields, xbits, bestbits, energy, ebest, tscale, scratch, drift, bytes, marg, status;
  ~Dense() {
    for (auto *b : {&q64, &qpack, &qdb, &fields, &xbits, &bestbits, &energy, &ebest, &tscale,
                    &scratch, &drift, &bytes, &marg, &status})
      b->release();
  }
};

}  // namespace os

using os::Ctx;
using os::Dense;

#define dguard os::guarded

extern "C" {

int os_dense_create(os_ctx *ctx_, int n, const double *q_sym, os_dense **out) {
  return dguard([&] {
    OS_REQUIRE(ctx_ && q_sym && out, "null argument");
    OS_REQUIRE(n >= 1, "n must be positive");
    if (n > os::dense_max_n())
      throw os::Error(OS_ERR_UNSUPPORTED, "dense sweep path supports n <= " +
                                              std::to_string(os::dense_max_n()));
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    auto d = new Dense();
    d->ctx = &c;
    d->n = n;
    d->np = os::dense_padded_n(n);
    d->npw = d->np / 32;
    const int np = d->np, nb = np / 16;
    std::vector<uint16_t> pack((size_t)nb * np * 32, 0);
    std::vector<float> qdb((size_t)nb * 256, 0.0f);
    for (int b = 0; b < nb; ++b) {
      for (int k = 0; k < 16; ++k) {
        const int i = b * 16 + k;
        if (i >= n) continue;
        for (int j = 0; j < n; ++j) {
          if (j == i) continue;
          const float qf = (float)q_sym[(size_t)i * n + j];
          const uint16_t hi = os::f32_to_bf16(qf);
          const uint16_t lo = os::f32_to_bf16(qf - os::bf16_to_f32(hi));
          uint16_t *dst = &pack[((size_t)b * np + j) * 32];
          dst[k] = hi;
          dst[16 + k] = lo;
          if (j / 16 == b) qdb[(size_t)b * 256 + k * 16 + (j % 16)] = qf;
        }
      }
    }
    try {
      d->q64.ensure(sizeof(double) * (size_t)n * n);
      d->qpack.ensure(pack.size() * sizeof(uint16_t));
      d->qdb.ensure(qdb.size() * sizeof(float));
      OS_HIP_CHECK(hipMemcpyAsync(d->q64.ptr, q_sym, sizeof(double) * (size_t)n * n,
                                  hipMemcpyHostToDevice, c.stream));
      OS_HIP_CHECK(hipMemcpyAsync(d->qpack.ptr, pack.data(), pack.size() * sizeof(uint16_t),
                                  hipMemcpyHostToDevice, c.stream));
      OS_HIP_CHECK(hipMemcpyAsync(d->qdb.ptr, qdb.data(), qdb.size() * sizeof(float),
                                  hipMemcpyHostToDevice, c.stream));
      c.sync();
    } catch (...) {
      delete d;
      throw;
    }
    *out = reinterpret_cast<os_dense *>(d);
  });
}

int os_dense_destroy(os_dense *d) {
  return dguard([&] { delete reinterpret_cast<Dense *>(d); });
}

int os_dense_anneal(os_dense *d_, const os_sweep_params *p, double *best_energies,
                    uint8_t *best_states, uint8_t *final_states, uint8_t *best_state,
                    double *best_energy, os_sweep_stats *stats) {
  return dguard([&] {
    OS_REQUIRE(d_ && p, "null argument");
    OS_REQUIRE(p->num_replicas >= 1, "num_replicas must be positive");
    OS_REQUIRE(p->num_sweeps >= 0, "num_sweeps must be non-negative");
    OS_REQUIRE(p->num_sweeps == 0 || p->tscale, "tscale is null");
    OS_REQUIRE(!(p->init_states && p->init_marginals), "give init_states or init_marginals, not both");
    Dense &d = *reinterpret_cast<Dense *>(d_);
    Ctx &c = *d.ctx;
    c.activate();
    const int n = d.n, np = d.np, npw = d.npw;
    const int R = (int)p->num_replicas;
    const int Rp = (R + 63) / 64 * 64;
    const int S = p->num_sweeps;
    const int seg = p->refresh_interval > 0 ? p->refresh_interval : std::max(S, 1);
    d.fields.ensure(sizeof(float) * (size_t)Rp * np);
    d.xbits.ensure(sizeof(uint32_t) * (size_t)Rp * npw);
    d.bestbits.ensure(sizeof(uint32_t) * (size_t)Rp * npw);
    d.energy.ensure(sizeof(double) * Rp);
    d.ebest.ensure(sizeof(double) * Rp);
    d.tscale.ensure(sizeof(float) * std::max(S, 1));
    d.scratch.ensure(os::energy_scratch_bytes(n, Rp));
    d.drift.ensure(sizeof(unsigned int));
    hipEvent_t t0, t1;
    OS_HIP_CHECK(hipEventCreate(&t0));
    OS_HIP_CHECK(hipEventCreate(&t1));
    std::vector<hipEvent_t> ev;
    struct EvGuard {
      std::vector<hipEvent_t> &v;
      hipEvent_t a, b;
      ~EvGuard() {
        for (auto e : v) (void)hipEventDestroy(e);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
      }
    } guard{ev, t0, t1};
    d.status.ensure(sizeof(int));
    OS_HIP_CHECK(hipEventRecord(t0, c.stream));
    OS_HIP_CHECK(hipMemsetAsync(d.status.ptr, 0, sizeof(int), c.stream));
    if (S > 0)
      OS_HIP_CHECK(hipMemcpyAsync(d.tscale.ptr, p->tscale, sizeof(float) * S,
                                  hipMemcpyHostToDevice, c.stream));
    // initial states
    if (p->init_states) {
      d.bytes.ensure((size_t)Rp * n);
      OS_HIP_CHECK(hipMemsetAsync(d.bytes.ptr, 0, (size_t)Rp * n, c.stream));
      OS_HIP_CHECK(hipMemcpyAsync(d.bytes.ptr, p->init_states, (size_t)R * n,
                                  hipMemcpyHostToDevice, c.stream));
      os::launch_pack_bytes(c.stream, d.bytes.as<uint8_t>(), n, Rp, npw, d.xbits.as<uint32_t>());
    } else if (p->init_marginals) {
      d.marg.ensure(sizeof(float) * n);
      OS_HIP_CHECK(hipMemcpyAsync(d.marg.ptr, p->init_marginals, sizeof(float) * n,
                                  hipMemcpyHostToDevice, c.stream));
      os::launch_warm_init(c.stream, n, npw, Rp, p->seed, p->instance, d.marg.as<float>(),
                           d.xbits.as<uint32_t>());
    } else {
      os::launch_sweep_init(c.stream, n, npw, Rp, p->seed, p->instance, d.xbits.as<uint32_t>());
    }
    // fields and energies of the start states (FP64), best = start
    OS_HIP_CHECK(hipMemsetAsync(d.fields.ptr, 0, sizeof(float) * (size_t)Rp * np, c.stream));
    OS_HIP_CHECK(hipMemsetAsync(d.drift.ptr, 0, sizeof(unsigned int), c.stream));
    os::launch_energy_fp64(c.stream, n, npw, d.q64.as<double>(), Rp, d.xbits.as<uint32_t>(),
                           d.energy.as<double>(), d.scratch.as<double>(), d.fields.as<float>(), np,
                           d.drift.as<unsigned int>());
    OS_HIP_CHECK(hipMemcpyAsync(d.ebest.ptr, d.energy.ptr, sizeof(double) * Rp,
                                hipMemcpyDeviceToDevice, c.stream));
    OS_HIP_CHECK(hipMemcpyAsync(d.bestbits.ptr, d.xbits.ptr, sizeof(uint32_t) * (size_t)Rp * npw,
                                hipMemcpyDeviceToDevice, c.stream));
    OS_HIP_CHECK(hipMemsetAsync(d.drift.ptr, 0, sizeof(unsigned int), c.stream));
    int launches = 0;
    for (int s0 = 0; s0 < S; s0 += seg) {
      const int ns = std::min(seg, S - s0);
      hipEvent_t a, b;
      OS_HIP_CHECK(hipEventCreate(&a));
      ev.push_back(a);
      OS_HIP_CHECK(hipEventCreate(&b));
      ev.push_back(b);
      OS_HIP_CHECK(hipEventRecord(a, c.stream));
      os::launch_dense_sweeps(c.stream, n, np, d.qpack.as<uint4>(), d.qdb.as<float>(),
                              d.tscale.as<float>() + s0, s0, ns, p->seed, p->instance, Rp,
                              d.fields.as<float>(), d.xbits.as<uint32_t>(),
                              d.bestbits.as<uint32_t>(), d.energy.as<double>(),
                              d.ebest.as<double>(), d.status.as<int>());
      OS_HIP_CHECK(hipEventRecord(b, c.stream));
      ++launches;
      // FP64 refresh of fields and running energies; records the fp32 drift
      os::launch_energy_fp64(c.stream, n, npw, d.q64.as<double>(), Rp, d.xbits.as<uint32_t>(),
                             d.energy.as<double>(), d.scratch.as<double>(), d.fields.as<float>(),
                             np, d.drift.as<unsigned int>());
    }
    // FP64 energies of the best states (the returned energies)
    os::launch_energy_fp64(c.stream, n, npw, d.q64.as<double>(), Rp, d.bestbits.as<uint32_t>(),
                           d.ebest.as<double>(), d.scratch.as<double>(), nullptr, 0, nullptr);
    std::vector<double> e(Rp);
    OS_HIP_CHECK(hipMemcpyAsync(e.data(), d.ebest.ptr, sizeof(double) * Rp, hipMemcpyDeviceToHost,
                                c.stream));
    int kstatus = 0;
    OS_HIP_CHECK(hipMemcpyAsync(&kstatus, d.status.ptr, sizeof(int), hipMemcpyDeviceToHost,
                                c.stream));
    unsigned int drift_bits = 0;
    OS_HIP_CHECK(hipMemcpyAsync(&drift_bits, d.drift.ptr, sizeof(unsigned int),
                                hipMemcpyDeviceToHost, c.stream));
    if (best_states || final_states) d.bytes.ensure((size_t)Rp * n);
    if (best_states) {
      os::launch_unpack_bits(c.stream, d.bestbits.as<uint32_t>(), n, R, npw, d.bytes.as<uint8_t>());
      OS_HIP_CHECK(hipMemcpyAsync(best_states, d.bytes.ptr, (size_t)R * n, hipMemcpyDeviceToHost,
                                  c.stream));
    }
    if (final_states) {
      c.sync();
      os::launch_unpack_bits(c.stream, d.xbits.as<uint32_t>(), n, R, npw, d.bytes.as<uint8_t>());
      OS_HIP_CHECK(hipMemcpyAsync(final_states, d.bytes.ptr, (size_t)R * n, hipMemcpyDeviceToHost,
                                  c.stream));
    }
    OS_HIP_CHECK(hipEventRecord(t1, c.stream));
    c.sync();
    if (kstatus != 0)
      throw os::Error(OS_ERR_INTERNAL, "dense sweep kernel aborted: a work-group wait expired");
    const size_t bi = (size_t)(std::min_element(e.begin(), e.begin() + R) - e.begin());
    if (best_energies) std::memcpy(best_energies, e.data(), sizeof(double) * R);
    if (best_energy) *best_energy = e[bi];
    if (best_state) {
      std::vector<uint32_t> wbits(npw);
      OS_HIP_CHECK(hipMemcpy(wbits.data(), d.bestbits.as<uint32_t>() + bi * npw,
                             sizeof(uint32_t) * npw, hipMemcpyDeviceToHost));
      for (int j = 0; j < n; ++j) best_state[j] = (uint8_t)((wbits[j >> 5] >> (j & 31)) & 1u);
    }
    if (stats) {
      double kms = 0.0;
      for (size_t i = 0; i + 1 < ev.size(); i += 2) {
        float ms = 0.0f;
        OS_HIP_CHECK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
        kms += ms;
      }
      float tms = 0.0f;
      OS_HIP_CHECK(hipEventElapsedTime(&tms, t0, t1));
      float drift;
      std::memcpy(&drift, &drift_bits, 4);
      stats->kernel_ms = kms;
      stats->total_ms = tms;
      stats->kernel_launches = launches;
      stats->max_field_drift = drift;
      stats->attempts = (uint64_t)R * (uint64_t)S * (uint64_t)n;
    }
  });
}

}  // extern "C"
