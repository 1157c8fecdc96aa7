// capi_dense.hip -- dense sweep-mode annealing entry points (os_dense_*).
//
// A dense problem keeps Q resident on the GPU in three forms:
//   q64    FP64 N x N (field refresh, energy verification),
//   qpack  bf16 [N/16 blocks][Np columns][32]: for block b and column j the
//          16 values hi(Q[16b+k][j]) then the 16 values lo(Q[16b+k][j]),
//          Q = Qhi + Qlo + O(2^-17 |Q|), diagonal zeroed (the MFMA B operand),
//   qdb    fp32 [N/16][16][16] diagonal sub-blocks (in-block corrections),
// packed on the GPU from q64 (gen.hip).  A batch stacks I same-size
// instances in every array.
// A run keeps per-replica fields (fp32 R x Np), states and best states
// (bit-packed), running and best energies (FP64) on the GPU; sweeps run in
// segments of `refresh_interval`, after each of which the fields and energies
// are recomputed in FP64 and the largest fp32 field error is recorded.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/onesolver.h"
#include "os_ctx.h"
#include "os_dense.h"
#include "os_kernels.h"

using os::Ctx;
using os::Dense;

#define dguard os::guarded

extern "C" {

namespace {

constexpr int kMaxStreamN = 16384;

// Allocates a batch of I instances of size n; pack() then fills the MFMA
// operand layout from the FP64 matrices in d->q64.
os::Dense *dense_alloc(Ctx &c, int I, int n) {
  if (n > kMaxStreamN)
    throw os::Error(OS_ERR_UNSUPPORTED,
                    "dense sweep path supports n <= " + std::to_string(kMaxStreamN));
  auto d = std::make_unique<Dense>();
  d->ctx = &c;
  d->I = I;
  d->n = n;
  d->stream = n > os::dense_max_n();
  const int kb = d->stream ? os::stream_block() : 16;
  d->np = d->stream ? (n + kb - 1) / kb * kb : os::dense_padded_n(n);
  d->npw = d->np / 32;
  const int np = d->np, nb = np / kb;
  d->q64.ensure(sizeof(double) * (size_t)n * n * I);
  d->qpack.ensure(sizeof(uint16_t) * (size_t)nb * np * 2 * kb * I);
  d->qdb.ensure(sizeof(float) * (size_t)nb * kb * kb * I);
  return d.release();
}

void pack(Dense &d) {
  if (d.stream)
    os::launch_pack_q64(d.ctx->stream, d.I, d.n, d.np, d.q64.as<double>(), d.qpack.as<uint16_t>(),
                        d.qdb.as<float>());
  else
    os::launch_pack_q(d.ctx->stream, d.I, d.n, d.np, d.q64.as<double>(), d.qpack.as<uint16_t>(),
                      d.qdb.as<float>());
}

}  // namespace

int os_dense_batch_create(os_ctx *ctx_, int num_instances, int n, const double *q_sym,
                          os_dense **out) {
  return dguard([&] {
    OS_REQUIRE(ctx_ && q_sym && out, "null argument");
    OS_REQUIRE(n >= 1, "n must be positive");
    OS_REQUIRE(num_instances >= 1, "num_instances must be positive");
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    std::unique_ptr<Dense> d(dense_alloc(c, num_instances, n));
    OS_HIP_CHECK(hipMemcpyAsync(d->q64.ptr, q_sym, sizeof(double) * (size_t)n * n * num_instances,
                                hipMemcpyHostToDevice, c.stream));
    pack(*d);
    c.sync();
    *out = reinterpret_cast<os_dense *>(d.release());
  });
}

int os_dense_batch_create_sk(os_ctx *ctx_, int num_instances, int n, uint64_t seed,
                             uint32_t instance_base, os_dense **out) {
  return dguard([&] {
    OS_REQUIRE(ctx_ && out, "null argument");
    OS_REQUIRE(n >= 2, "n must be at least 2");
    OS_REQUIRE(num_instances >= 1, "num_instances must be positive");
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    std::unique_ptr<Dense> d(dense_alloc(c, num_instances, n));
    os::launch_gen_sk(c.stream, num_instances, n, seed, instance_base, d->q64.as<double>());
    pack(*d);
    c.sync();
    *out = reinterpret_cast<os_dense *>(d.release());
  });
}

int os_dense_batch_create_uniform(os_ctx *ctx_, int num_instances, int n, uint64_t seed,
                                  uint32_t instance_base, os_dense **out) {
  return dguard([&] {
    OS_REQUIRE(ctx_ && out, "null argument");
    OS_REQUIRE(n >= 1, "n must be positive");
    OS_REQUIRE(num_instances >= 1, "num_instances must be positive");
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    std::unique_ptr<Dense> d(dense_alloc(c, num_instances, n));
    os::launch_gen_uniform(c.stream, num_instances, n, seed, instance_base, d->q64.as<double>());
    pack(*d);
    c.sync();
    *out = reinterpret_cast<os_dense *>(d.release());
  });
}

int os_dense_copy_q(os_dense *d_, double *q_out) {
  return dguard([&] {
    OS_REQUIRE(d_ && q_out, "null argument");
    Dense &d = *reinterpret_cast<Dense *>(d_);
    d.ctx->activate();
    OS_HIP_CHECK(hipMemcpyAsync(q_out, d.q64.ptr, sizeof(double) * (size_t)d.n * d.n * d.I,
                                hipMemcpyDeviceToHost, d.ctx->stream));
    d.ctx->sync();
  });
}

}  // extern "C"

os::Dense *os::dense_create_device(Ctx &c, int n, void (*fill)(void *arg, double *q64),
                                   void *arg) {
  c.activate();
  std::unique_ptr<Dense> d(dense_alloc(c, 1, n));
  OS_HIP_CHECK(hipMemsetAsync(d->q64.ptr, 0, sizeof(double) * (size_t)n * n, c.stream));
  fill(arg, d->q64.as<double>());
  pack(*d);
  c.sync();
  return d.release();
}

extern "C" {

int os_dense_create(os_ctx *ctx_, int n, const double *q_sym, os_dense **out) {
  return os_dense_batch_create(ctx_, 1, n, q_sym, out);
}

int os_dense_destroy(os_dense *d) {
  return dguard([&] { delete reinterpret_cast<Dense *>(d); });
}

}  // extern "C"

namespace {

// Replica exchange: ladders of K temperature slots with scales tscale[K];
// after every `interval` sweeps, adjacent slots (k, k+1), k of the round's
// parity, swap their replicas with probability
// min(1, exp((beta_k - beta_{k+1}) (E_k - E_{k+1}))).
struct Exchange {
  int K = 0;
  const float *tscale = nullptr;
  int interval = 0;
  uint32_t *final_slots = nullptr;  // optional [I][R] slot of each replica at the end
};

// Population annealing: segments of `theta` sweeps; after each but the last,
// systematic resampling with dbeta = 1/ts[k+1] - 1/ts[k] on the running
// energies; rows take the state words, fields and energy of their ancestor.
struct Population {
  int K = 0;
  int theta = 0;
  const float *tscale = nullptr;  // [K]
  uint32_t *ancestors = nullptr;  // optional [I][K-1][R]
};

int dense_run(os_dense *d_, const os_sweep_params *p, const Exchange *xc, const Population *pa,
              double *best_energies,
              uint8_t *best_states, uint8_t *final_states, uint8_t *best_state,
              double *best_energy, os_sweep_stats *stats) {
  return dguard([&] {
    OS_REQUIRE(d_ && p, "null argument");
    OS_REQUIRE(p->num_replicas >= 1, "num_replicas must be positive");
    OS_REQUIRE(p->num_sweeps >= 0, "num_sweeps must be non-negative");
    OS_REQUIRE(p->num_sweeps == 0 || p->tscale || xc, "tscale is null");
    OS_REQUIRE(!(p->init_states && p->init_marginals), "give init_states or init_marginals, not both");
    Dense &d = *reinterpret_cast<Dense *>(d_);
    OS_REQUIRE((uint64_t)p->instance + (uint64_t)d.I <= (1u << 24),
               "instance ids must stay below 2^24");
    Ctx &c = *d.ctx;
    c.activate();
    const int I = d.I, n = d.n, np = d.np, npw = d.npw;
    const int R = (int)p->num_replicas;
    const int Rp = (R + 63) / 64 * 64;
    const size_t T = (size_t)I * Rp;  // state rows
    const size_t nn = (size_t)n * n;
    const int S = p->num_sweeps;
    const int refresh = p->refresh_interval > 0 ? p->refresh_interval : std::max(S, 1);
    const int seg = xc ? xc->interval : pa ? pa->theta : refresh;
    if (xc) {
      OS_REQUIRE(!d.stream, "replica exchange supports n <= 1024");
      OS_REQUIRE(xc->K >= 2 && xc->tscale, "replica exchange needs >= 2 temperature slots");
      OS_REQUIRE(R % xc->K == 0, "num_replicas must be a multiple of the temperature slots");
      OS_REQUIRE(xc->interval >= 1, "exchange_interval must be positive");
      OS_REQUIRE(p->refresh_interval <= 0 || p->refresh_interval % xc->interval == 0,
                 "refresh_interval must be a multiple of exchange_interval");
      for (int k = 0; k < xc->K; ++k)
        OS_REQUIRE(xc->tscale[k] > 0.0f && std::isfinite(xc->tscale[k]),
                   "temperature scales must be positive");
    }
    if (pa) {
      OS_REQUIRE(pa->K >= 1 && pa->theta >= 1 && pa->tscale, "invalid population schedule");
      OS_REQUIRE(p->refresh_interval <= 0 || p->refresh_interval % pa->theta == 0,
                 "refresh_interval must be a multiple of sweeps_per_step");
      for (int k = 0; k < pa->K; ++k)
        OS_REQUIRE(pa->tscale[k] > 0.0f && std::isfinite(pa->tscale[k]),
                   "temperature scales must be positive");
    }
    // replica exchange: the slot of every replica and the replica at every
    // slot (per instance and ladder), and the per-row scales of the launch
    const int K = xc ? xc->K : 0, L = xc ? R / K : 0;
    std::vector<uint32_t> slot_of, rep_at;
    std::vector<float> tsr_h;
    if (xc) {
      slot_of.resize((size_t)I * R);
      rep_at.resize((size_t)I * R);
      for (size_t i = 0; i < (size_t)I * R; ++i) {
        slot_of[i] = (uint32_t)(i % R % K);
        rep_at[i] = (uint32_t)(i % R);
      }
      tsr_h.assign(T, xc->tscale[0]);
      d.tsr.ensure(sizeof(float) * T);
    }
    auto upload_tsr = [&]() {
      for (int s = 0; s < I; ++s)
        for (int r = 0; r < R; ++r)
          tsr_h[(size_t)s * Rp + r] = xc->tscale[slot_of[(size_t)s * R + r]];
      OS_HIP_CHECK(hipMemcpyAsync(d.tsr.ptr, tsr_h.data(), sizeof(float) * T,
                                  hipMemcpyHostToDevice, c.stream));
    };
    // the exchange rounds run on the GPU: slots, replicas at slots and the
    // slots' scales in device memory (uploaded with the first scales)
    uint32_t *d_slot = nullptr, *d_rep = nullptr;
    float *d_ts = nullptr;
    if (xc) {
      d.pt_slots.ensure(sizeof(uint32_t) * slot_of.size() * 2 + sizeof(float) * K);
      d_slot = d.pt_slots.as<uint32_t>();
      d_rep = d_slot + slot_of.size();
      d_ts = reinterpret_cast<float *>(d_rep + rep_at.size());
    }
    // stream path: pipeline depth P, P field copies (OS_STREAM_DEPTH
    // overrides the default, for tuning)
    const char *denv = std::getenv("OS_STREAM_DEPTH");
    const int P = denv ? std::min(os::Dense::kMaxDepth, std::max(2, std::atoi(denv)))
                       : os::Dense::kDepth;
    d.fields.ensure(sizeof(float) * T * np * (d.stream ? P : 1));
    d.xbits.ensure(sizeof(uint32_t) * T * npw);
    d.bestbits.ensure(sizeof(uint32_t) * T * npw);
    d.energy.ensure(sizeof(double) * T);
    d.ebest.ensure(sizeof(double) * T);
    d.tscale.ensure(sizeof(float) * std::max(S, 1));
    d.scratch.ensure(os::energy_scratch_bytes(n, (int)T));
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
    hipStream_t st = c.stream;
    auto qi = [&](int s) { return d.q64.as<double>() + s * nn; };
    auto rows = [&](int s) { return (size_t)s * Rp; };
    // FP64 energies (and optionally fields) of the states in `bits`, all
    // instances in one launch
    auto energies = [&](uint32_t *bits, double *out, float *fields) {
      os::launch_energy_fp64(st, n, npw, qi(0), Rp, bits, out, d.scratch.as<double>(), fields, np,
                             fields ? d.drift.as<unsigned int>() : nullptr, I);
    };
    OS_HIP_CHECK(hipEventRecord(t0, st));
    OS_HIP_CHECK(hipMemsetAsync(d.status.ptr, 0, sizeof(int), st));
    if (S > 0 && !xc)
      OS_HIP_CHECK(hipMemcpyAsync(d.tscale.ptr, p->tscale, sizeof(float) * S,
                                  hipMemcpyHostToDevice, st));
    if (xc) {
      upload_tsr();
      OS_HIP_CHECK(hipMemcpyAsync(d_slot, slot_of.data(), sizeof(uint32_t) * slot_of.size(),
                                  hipMemcpyHostToDevice, st));
      OS_HIP_CHECK(hipMemcpyAsync(d_rep, rep_at.data(), sizeof(uint32_t) * rep_at.size(),
                                  hipMemcpyHostToDevice, st));
      OS_HIP_CHECK(hipMemcpyAsync(d_ts, xc->tscale, sizeof(float) * K, hipMemcpyHostToDevice,
                                  st));
      c.sync();
    }
    std::vector<double> e_now(pa ? T : 0);
    std::vector<uint32_t> anc_h(pa ? (size_t)I * R : 0);
    std::vector<double> wsum(pa ? R : 0);
    if (pa) {
      d.pa_x.ensure(sizeof(uint32_t) * T * npw);
      d.pa_f.ensure(sizeof(float) * T * np);
      d.pa_e.ensure(sizeof(double) * T);
      d.pa_anc.ensure(sizeof(uint32_t) * I * R);
    }
    int round = 0;
    // initial states
    if (p->init_states) {
      d.bytes.ensure(T * n);
      OS_HIP_CHECK(hipMemsetAsync(d.bytes.ptr, 0, T * n, st));
      for (int s = 0; s < I; ++s)
        OS_HIP_CHECK(hipMemcpyAsync(d.bytes.as<uint8_t>() + rows(s) * n,
                                    p->init_states + (size_t)s * R * n, (size_t)R * n,
                                    hipMemcpyHostToDevice, st));
      os::launch_pack_bytes(st, d.bytes.as<uint8_t>(), n, (int)T, npw, d.xbits.as<uint32_t>());
    } else if (p->init_marginals) {
      d.marg.ensure(sizeof(float) * n * I);
      OS_HIP_CHECK(hipMemcpyAsync(d.marg.ptr, p->init_marginals, sizeof(float) * n * I,
                                  hipMemcpyHostToDevice, st));
      os::launch_warm_init(st, n, npw, Rp, p->seed, p->instance, d.marg.as<float>(),
                           d.xbits.as<uint32_t>(), I);
    } else {
      os::launch_sweep_init(st, n, npw, Rp, p->seed, p->instance, d.xbits.as<uint32_t>(), I);
    }
    // fields and energies of the start states (FP64), best = start
    OS_HIP_CHECK(hipMemsetAsync(d.fields.ptr, 0, sizeof(float) * T * np, st));
    OS_HIP_CHECK(hipMemsetAsync(d.drift.ptr, 0, sizeof(unsigned int), st));
    energies(d.xbits.as<uint32_t>(), d.energy.as<double>(), d.fields.as<float>());
    OS_HIP_CHECK(hipMemcpyAsync(d.ebest.ptr, d.energy.ptr, sizeof(double) * T,
                                hipMemcpyDeviceToDevice, st));
    OS_HIP_CHECK(hipMemcpyAsync(d.bestbits.ptr, d.xbits.ptr, sizeof(uint32_t) * T * npw,
                                hipMemcpyDeviceToDevice, st));
    OS_HIP_CHECK(hipMemsetAsync(d.drift.ptr, 0, sizeof(unsigned int), st));
    int launches = 0;
    // stream path: the up-to-date field copy (0 .. P-1)
    int cur = 0;
    auto fbuf = [&](int i) { return d.fields.as<float>() + (size_t)(i % P) * T * np; };
    // blocks per field pass: OS_STREAM_GROUP overrides the default (tuning)
    const char *genv = std::getenv("OS_STREAM_GROUP");
    d.group = genv ? std::min(16, std::max(1, std::atoi(genv)))
                   : (d.tuned_group ? d.tuned_group : os::Dense::kGroup);
    const int m = d.group, ring = (P + 1) * d.group;
    // decide launches: operand chunking (OS_STREAM_DECIDE_VARIANT overrides
    // the default, for tuning)
    const char *venv = std::getenv("OS_STREAM_DECIDE_VARIANT");
    const int dvar = venv ? std::atoi(venv) : os::Dense::kDecideVariant;
    // pass launches: blocks whose operands load in one round trip
    // (OS_STREAM_PASS_CHUNK overrides the default, for tuning)
    const char *cenv = std::getenv("OS_STREAM_PASS_CHUNK");
    const int chunk = cenv ? std::atoi(cenv) : os::Dense::kPassChunk;
    // pass launches: persistent work-groups per CU (OS_STREAM_PASS_WGS_PER_CU
    // overrides the default); the grid leaves room for the decide launches
    const char *wenv = std::getenv("OS_STREAM_PASS_WGS_PER_CU");
    const char *grenv = std::getenv("OS_STREAM_PASS_GRID");
    const int pgrid =
        grenv ? std::max(1, std::atoi(grenv))
              : c.num_cus * (wenv ? std::max(1, std::atoi(wenv))
                                  : (d.tuned_pass_wgs ? d.tuned_pass_wgs
                                                      : os::Dense::kPassWgsPerCu));
    // stream path, with stats: HIP events around every pass_stride-th pass
    // (at most kPassSamples of them)
    constexpr int kPassSamples = 256;
    int passes = 0, timed = 0, pass_stride = 1;
    if (d.stream) {
      const long long nb = np / os::stream_block(), total = ((long long)S * nb + m - 1) / m;
      pass_stride = (int)std::max(1LL, (total + kPassSamples - 1) / kPassSamples);
      if (stats)
        while (d.pass_ev.size() < 2 * (size_t)kPassSamples) {
          hipEvent_t e;
          OS_HIP_CHECK(hipEventCreate(&e));
          d.pass_ev.push_back(e);
        }
    }
    if (d.stream) {
      d.dblk.ensure(sizeof(uint4) * ring * T);
      if (!d.side) {
        OS_HIP_CHECK(hipStreamCreateWithFlags(&d.side, hipStreamNonBlocking));
        for (hipEvent_t *e : {&d.ev_seg, &d.ev_dec})
          OS_HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
        for (hipEvent_t &e : d.ev_pass)
          OS_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      }
    }
    for (int s0 = 0; s0 < S; s0 += seg) {
      const int ns = std::min(seg, S - s0);
      hipEvent_t a, b;
      OS_HIP_CHECK(hipEventCreate(&a));
      ev.push_back(a);
      OS_HIP_CHECK(hipEventCreate(&b));
      ev.push_back(b);
      OS_HIP_CHECK(hipEventRecord(a, st));
      if (d.stream) {
        // Block g (0 .. steps-1) of group H = g / m is decided on d.side from
        // the fields after pass H-P plus the flips of blocks m(H-P+1) .. g-1
        // (at most Pm - 1 of them); pass H (context's stream) adds the flips
        // of group H to the fields after pass H-1, in the next copy, once the
        // group is decided.  So pass H runs beside the decisions of groups
        // H+1 .. H+P-1 (with P >= 3 the two cross-stream hand-offs of a pass
        // leave the decisions' path, but the decides get slower by more than
        // they save).  Pass H lands in copy
        // cur + H + 1 (mod P), the copy the decides of group H read: it waits
        // for them.  The flip masks of block g stay in ring slot
        // g % ((P+1)m) until every reader (the decides of groups H .. H+P-1,
        // pass H) is done: block g + (P+1)m is decided after pass H+1.
        const int nb = np / os::stream_block();
        uint4 *ring_d = d.dblk.as<uint4>();
        const int steps = ns * nb;
        OS_HIP_CHECK(hipEventRecord(d.ev_seg, st));
        OS_HIP_CHECK(hipStreamWaitEvent(d.side, d.ev_seg, 0));
        for (int g = 0; g < steps; ++g) {
          const int H = g / m, sw = s0 + g / nb, blk = g % nb;
          if (g % m == 0 && H >= P)
            OS_HIP_CHECK(hipStreamWaitEvent(d.side, d.ev_pass[(H - P) % P], 0));
          const int gb = std::max(0, m * (H - P + 1));
          os::launch_stream_decide(d.side, I, n, np, Rp, g, gb, ring, sw, p->tscale[sw], p->seed,
                                   p->instance, blk == nb - 1, d.qpack.as<uint4>(),
                                   d.qdb.as<float>(), fbuf(cur + std::max(0, H - P + 1)),
                                   d.xbits.as<uint32_t>(), ring_d, d.energy.as<double>(),
                                   d.ebest.as<double>(), d.bestbits.as<uint32_t>(), dvar);
          ++launches;
          if (g % m == m - 1 || g == steps - 1) {
            OS_HIP_CHECK(hipEventRecord(d.ev_dec, d.side));
            OS_HIP_CHECK(hipStreamWaitEvent(st, d.ev_dec, 0));
            const bool timing = stats && passes % pass_stride == 0 && timed < kPassSamples;
            if (timing) OS_HIP_CHECK(hipEventRecord(d.pass_ev[2 * timed], st));
            os::launch_stream_pass(st, I, np, Rp, m * H, g - m * H + 1, ring, chunk, pgrid,
                                   d.qpack.as<uint4>(), fbuf(cur + H), fbuf(cur + H + 1),
                                   ring_d);
            if (timing) OS_HIP_CHECK(hipEventRecord(d.pass_ev[2 * timed++ + 1], st));
            ++passes;
            OS_HIP_CHECK(hipEventRecord(d.ev_pass[H % P], st));
          }
        }
        // the last pass waited for the last decide: the context's stream
        // holds every launch of the segment again
        if (steps > 0) cur = (cur + (steps - 1) / m + 1) % P;
      } else {
        os::launch_dense_sweeps(st, n, np, I, d.qpack.as<uint4>(), d.qdb.as<float>(),
                                d.tscale.as<float>() + s0, xc ? d.tsr.as<float>() : nullptr, s0,
                                ns, p->seed, p->instance, Rp,
                                d.fields.as<float>(), d.xbits.as<uint32_t>(),
                                d.bestbits.as<uint32_t>(), d.energy.as<double>(),
                                d.ebest.as<double>(), d.status.as<int>());
        ++launches;
      }
      OS_HIP_CHECK(hipEventRecord(b, st));
      // FP64 refresh of fields and running energies; records the fp32 drift
      if (!(xc || pa) || (s0 + ns) % refresh == 0 || s0 + ns >= S)
        energies(d.xbits.as<uint32_t>(), d.energy.as<double>(), fbuf(cur));
      if (pa && s0 + ns < S) {
        // resampling round k on the running energies, then every row moves
        // to its ancestor's state, fields and energy
        const int k = s0 / pa->theta;
        OS_HIP_CHECK(hipMemcpyAsync(e_now.data(), d.energy.ptr, sizeof(double) * T,
                                    hipMemcpyDeviceToHost, st));
        c.sync();
        const double dbeta = 1.0 / (double)pa->tscale[k + 1] - 1.0 / (double)pa->tscale[k];
        const uint32_t k0 = (uint32_t)p->seed, k1 = (uint32_t)(p->seed >> 32);
        for (int s = 0; s < I; ++s) {
          const uint32_t tag = ((p->instance + (uint32_t)s) << 8) | os::TAG_RESAMPLE;
          const double u = os::resample_uniform(k0, k1, tag, (uint32_t)k);
          uint32_t *anc = anc_h.data() + (size_t)s * R;
          os::resample(e_now.data() + rows(s), R, dbeta, u, anc, wsum.data());
          if (pa->ancestors)
            std::memcpy(pa->ancestors + ((size_t)s * (pa->K - 1) + k) * R, anc,
                        sizeof(uint32_t) * R);
        }
        OS_HIP_CHECK(hipMemcpyAsync(d.pa_anc.ptr, anc_h.data(), sizeof(uint32_t) * I * R,
                                    hipMemcpyHostToDevice, st));
        os::launch_dense_gather(st, I, Rp, R, npw, np, d.pa_anc.as<uint32_t>(),
                                d.xbits.as<uint32_t>(), fbuf(cur), d.energy.as<double>(),
                                d.pa_x.as<uint32_t>(), d.pa_f.as<float>(), d.pa_e.as<double>());
        OS_HIP_CHECK(hipMemcpyAsync(d.xbits.ptr, d.pa_x.ptr, sizeof(uint32_t) * T * npw,
                                    hipMemcpyDeviceToDevice, st));
        OS_HIP_CHECK(hipMemcpyAsync(fbuf(cur), d.pa_f.ptr, sizeof(float) * T * np,
                                    hipMemcpyDeviceToDevice, st));
        OS_HIP_CHECK(hipMemcpyAsync(d.energy.ptr, d.pa_e.ptr, sizeof(double) * T,
                                    hipMemcpyDeviceToDevice, st));
      }
      if (xc && s0 + ns < S) {
        // exchange round on the running energies of the launch (or of the
        // refresh); the next launch reads the new scales
        os::launch_exchange_round(st, I, L, K, R, Rp, (uint32_t)round, p->seed, p->instance,
                                  d_ts, d.energy.as<double>(), d_rep, d_slot, d.tsr.as<float>());
        ++round;
      }
    }
    // FP64 energies of the best states (the returned energies)
    energies(d.bestbits.as<uint32_t>(), d.ebest.as<double>(), nullptr);
    std::vector<double> e(T);
    OS_HIP_CHECK(hipMemcpyAsync(e.data(), d.ebest.ptr, sizeof(double) * T, hipMemcpyDeviceToHost,
                                st));
    int kstatus = 0;
    OS_HIP_CHECK(hipMemcpyAsync(&kstatus, d.status.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
    unsigned int drift_bits = 0;
    OS_HIP_CHECK(hipMemcpyAsync(&drift_bits, d.drift.ptr, sizeof(unsigned int),
                                hipMemcpyDeviceToHost, st));
    if (best_states || final_states) d.bytes.ensure(T * n);
    auto fetch = [&](uint32_t *bits, uint8_t *dst) {
      os::launch_unpack_bits(st, bits, n, (int)T, npw, d.bytes.as<uint8_t>());
      for (int s = 0; s < I; ++s)
        OS_HIP_CHECK(hipMemcpyAsync(dst + (size_t)s * R * n, d.bytes.as<uint8_t>() + rows(s) * n,
                                    (size_t)R * n, hipMemcpyDeviceToHost, st));
      c.sync();
    };
    if (best_states) fetch(d.bestbits.as<uint32_t>(), best_states);
    if (final_states) fetch(d.xbits.as<uint32_t>(), final_states);
    OS_HIP_CHECK(hipEventRecord(t1, st));
    c.sync();
    if (kstatus != 0)
      throw os::Error(OS_ERR_INTERNAL, "dense sweep kernel aborted: a work-group wait expired");
    std::vector<uint32_t> wbits;
    if (best_state) {
      wbits.resize(T * npw);
      OS_HIP_CHECK(hipMemcpy(wbits.data(), d.bestbits.ptr, sizeof(uint32_t) * T * npw,
                             hipMemcpyDeviceToHost));
    }
    for (int s = 0; s < I; ++s) {
      const double *es = e.data() + rows(s);
      const size_t bi = (size_t)(std::min_element(es, es + R) - es);
      if (best_energies) std::memcpy(best_energies + (size_t)s * R, es, sizeof(double) * R);
      if (best_energy) best_energy[s] = es[bi];
      if (best_state) {
        const uint32_t *wb = wbits.data() + (rows(s) + bi) * npw;
        for (int j = 0; j < n; ++j)
          best_state[(size_t)s * n + j] = (uint8_t)((wb[j >> 5] >> (j & 31)) & 1u);
      }
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
      stats->attempts = (uint64_t)I * (uint64_t)R * (uint64_t)S * (uint64_t)n;
      double pms = 0.0;
      for (int k = 0; k < timed; ++k) {
        float ms = 0.0f;
        OS_HIP_CHECK(hipEventElapsedTime(&ms, d.pass_ev[2 * k], d.pass_ev[2 * k + 1]));
        pms += ms;
      }
      stats->pass_avg_ms = timed > 0 ? pms / timed : 0.0;
      stats->pass_launches = passes;
    }
    if (xc && xc->final_slots)
      OS_HIP_CHECK(hipMemcpy(xc->final_slots, d_slot, sizeof(uint32_t) * I * R,
                             hipMemcpyDeviceToHost));
  });
}

}  // namespace

extern "C" {

int os_dense_anneal(os_dense *d_, const os_sweep_params *p, double *best_energies,
                    uint8_t *best_states, uint8_t *final_states, uint8_t *best_state,
                    double *best_energy, os_sweep_stats *stats) {
  return dense_run(d_, p, nullptr, nullptr, best_energies, best_states, final_states, best_state,
                   best_energy, stats);
}

int os_dense_tune(os_dense *d_, const os_sweep_params *p, int32_t *group,
                  int32_t *pass_wgs_per_cu, double *best_ms) {
  if (group) *group = 0;
  if (pass_wgs_per_cu) *pass_wgs_per_cu = 0;
  if (best_ms) *best_ms = 0.0;
  if (!d_ || !p) return dguard([&] { OS_REQUIRE(false, "null argument"); });
  if (p->num_sweeps < 1 || !p->tscale)
    return dguard([&] { OS_REQUIRE(false, "tuning needs num_sweeps >= 1 and tscale"); });
  os::Dense &d = *reinterpret_cast<os::Dense *>(d_);
  if (!d.stream) return OS_OK;  // n <= 1024: nothing to tune
  // probe runs: the caller's replicas, schedule and seed, at most 4 sweeps,
  // no outputs; each run starts from fresh states, so the probes leave the
  // caller's later runs unchanged
  os_sweep_params probe = *p;
  probe.num_sweeps = std::min(p->num_sweeps, 4);
  probe.refresh_interval = 0;
  static constexpr int kGroups[] = {3, 4, 5, 6, 8};
  static constexpr int kWgs[] = {1, 2};
  int bg = 0, bw = 0;
  double bms = 0.0;
  for (int wg : kWgs)
    for (int m : kGroups) {
      d.tuned_group = m;
      d.tuned_pass_wgs = wg;
      os_sweep_stats st{};
      const int rc = dense_run(d_, &probe, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, &st);
      if (rc != OS_OK) {
        d.tuned_group = bg;
        d.tuned_pass_wgs = bw;
        return rc;
      }
      if (bg == 0 || st.kernel_ms < bms) {
        bg = m;
        bw = wg;
        bms = st.kernel_ms;
      }
    }
  d.tuned_group = bg;
  d.tuned_pass_wgs = bw;
  if (group) *group = bg;
  if (pass_wgs_per_cu) *pass_wgs_per_cu = bw;
  if (best_ms) *best_ms = bms;
  return OS_OK;
}

int os_dense_tempering(os_dense *d_, const os_tempering_params *t, double *best_energies,
                       uint8_t *best_states, uint8_t *final_states, uint32_t *final_slots,
                       uint8_t *best_state, double *best_energy, os_sweep_stats *stats) {
  if (!t) return dguard([&] { OS_REQUIRE(false, "null argument"); });
  if (t->num_ladders < 1 || t->num_temperatures < 2 ||
      (uint64_t)t->num_ladders * t->num_temperatures > (1u << 24))
    return dguard([&] { OS_REQUIRE(false, "invalid ladder shape"); });
  os_sweep_params p{};
  p.num_replicas = t->num_ladders * t->num_temperatures;
  p.num_sweeps = t->num_sweeps;
  p.tscale = nullptr;
  p.seed = t->seed;
  p.instance = t->instance;
  p.refresh_interval = t->refresh_interval;
  Exchange xc;
  xc.K = (int)t->num_temperatures;
  xc.tscale = t->tscale;
  xc.interval = t->exchange_interval;
  xc.final_slots = final_slots;
  return dense_run(d_, &p, &xc, nullptr, best_energies, best_states, final_states, best_state,
                   best_energy, stats);
}

int os_dense_population(os_dense *d_, const os_population_params *q, double *best_energies,
                        uint8_t *best_states, uint8_t *final_states, uint32_t *ancestors,
                        uint8_t *best_state, double *best_energy, os_sweep_stats *stats) {
  if (!q) return dguard([&] { OS_REQUIRE(false, "null argument"); });
  if (q->num_steps < 1 || q->sweeps_per_step < 1 || !q->tscale ||
      (int64_t)q->num_steps * q->sweeps_per_step > (int64_t)INT32_MAX)
    return dguard([&] { OS_REQUIRE(false, "invalid population schedule"); });
  // the per-sweep scales: tscale[k] for the sweeps of step k
  std::vector<float> ts;
  try {
    ts.resize((size_t)q->num_steps * q->sweeps_per_step);
  } catch (...) {
    return dguard([&] { OS_REQUIRE(false, "population schedule too long"); });
  }
  for (int k = 0; k < q->num_steps; ++k)
    std::fill(ts.begin() + (size_t)k * q->sweeps_per_step,
              ts.begin() + (size_t)(k + 1) * q->sweeps_per_step, q->tscale[k]);
  os_sweep_params p{};
  p.num_replicas = q->num_replicas;
  p.num_sweeps = q->num_steps * q->sweeps_per_step;
  p.tscale = ts.data();
  p.seed = q->seed;
  p.instance = q->instance;
  p.refresh_interval = q->refresh_interval;
  Population pa;
  pa.K = q->num_steps;
  pa.theta = q->sweeps_per_step;
  pa.tscale = q->tscale;
  pa.ancestors = ancestors;
  return dense_run(d_, &p, nullptr, &pa, best_energies, best_states, final_states, best_state,
                   best_energy, stats);
}

}  // extern "C"
