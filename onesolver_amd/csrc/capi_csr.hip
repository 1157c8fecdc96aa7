// capi_csr.hip -- batched sparse (CSR) sweep-mode annealing entry points (os_csr_*).
//
// A batch keeps its instances resident on the GPU twice: as fp32 ELL rows
// (the sweep kernel's layout: row width padded to a multiple of 4 with the
// spin index n and weight 0) and as the caller's FP64 CSR (the energy check).
// A run anneals num_replicas replicas of every instance in one persistent
// kernel, then recomputes the FP64 energy of every best state; the three
// launches (counter reset, sweeps, energies) are captured once into a
// hipGraph and replayed, the run parameters living in device memory.  The
// launch shape (waves per work-group, resident work-groups) is picked by
// timing short runs of the candidates on the batch itself.
#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "../../include/onesolver.h"
#include "os_ctx.h"
#include "os_kernels.h"

namespace os {

struct CsrBatch {
  Ctx *ctx = nullptr;
  int I = 0;
  int max_words = 0;
  int64_t total_words = 0;  // sum of words over instances
  int64_t total_spins = 0;
  std::vector<CsrInst> inst;
  DevBuf d_inst, ell_col, ell_val, diag, diag64, row_ptr, col, val64;
  // per run shape
  int Rp = 0;
  DevBuf items, run, tscale, init, best, fin, eb, ef, e64, queue;
  int num_items = 0;
  // launch shape chosen by the tuner (0: not tuned); reg: row-prefetching kernel
  int waves = 0, groups = 0;
  bool reg = false;
  int max_width = 0;
  // captured graph and what it was captured for
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_waves = 0, g_groups = 0, g_Rp = 0;
  bool g_reg = false;
  const void *g_tscale = nullptr;

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }
  ~CsrBatch() {
    drop_graph();
    for (auto *b : {&d_inst, &ell_col, &ell_val, &diag, &diag64, &row_ptr, &col, &val64, &items,
                    &run, &tscale, &init, &best, &fin, &eb, &ef, &e64, &queue})
      b->release();
  }

  void enqueue(bool rg, int w, int g, bool energy) {
    hipStream_t st = ctx->stream;
    launch_csr_queue_reset(st, queue.as<int>());
    launch_csr_sweeps(st, rg, w, g, max_words, d_inst.as<CsrInst>(), items.as<int2>(),
                      run.as<CsrRun>(), ell_col.as<int4>(), ell_val.as<float4>(), diag.as<float>(),
                      tscale.as<float>(), init.as<uint32_t>(), best.as<uint32_t>(),
                      fin.as<uint32_t>(), eb.as<double>(), ef.as<double>(), queue.as<int>());
    if (energy)
      launch_csr_energy(st, d_inst.as<CsrInst>(), items.as<int2>(), num_items, Rp,
                        row_ptr.as<int32_t>(), col.as<int32_t>(), val64.as<double>(),
                        diag64.as<double>(), best.as<uint32_t>(), e64.as<double>());
  }
};

namespace {

// Lays out per-replica storage for Rp replicas per instance and the work
// items, largest instances (by row work) first.
void shape_for(CsrBatch &b, int Rp) {
  if (b.Rp == Rp) return;
  int64_t st = 0, rep = 0;
  for (auto &d : b.inst) {
    d.state_off = st;
    d.rep_off = rep;
    st += (int64_t)d.words * Rp;
    rep += Rp;
  }
  std::vector<int> order(b.I);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    const int64_t wa = (int64_t)b.inst[a].n * (b.inst[a].width + 8);
    const int64_t wc = (int64_t)b.inst[c].n * (b.inst[c].width + 8);
    return wa > wc;
  });
  const int groups = Rp / 64;
  std::vector<int2> items;
  items.reserve((size_t)b.I * groups);
  for (int i : order)
    for (int gidx = 0; gidx < groups; ++gidx) items.push_back(make_int2(i, gidx));
  b.num_items = (int)items.size();
  hipStream_t s = b.ctx->stream;
  b.d_inst.ensure(sizeof(CsrInst) * b.I);
  OS_HIP_CHECK(hipMemcpyAsync(b.d_inst.ptr, b.inst.data(), sizeof(CsrInst) * b.I,
                              hipMemcpyHostToDevice, s));
  b.items.ensure(sizeof(int2) * std::max<size_t>(items.size(), 1));
  OS_HIP_CHECK(hipMemcpyAsync(b.items.ptr, items.data(), sizeof(int2) * items.size(),
                              hipMemcpyHostToDevice, s));
  const size_t planes = sizeof(uint32_t) * (size_t)std::max<int64_t>(st, 1);
  b.init.ensure(planes);
  b.best.ensure(planes);
  b.fin.ensure(planes);
  const size_t reps = sizeof(double) * (size_t)std::max<int64_t>(rep, 1);
  b.eb.ensure(reps);
  b.ef.ensure(reps);
  b.e64.ensure(reps);
  b.queue.ensure(sizeof(int) * 64);
  b.run.ensure(sizeof(CsrRun));
  OS_HIP_CHECK(hipStreamSynchronize(s));
  b.Rp = Rp;
  b.drop_graph();
}

float time_launch(CsrBatch &b, bool rg, int w, int g) {
  Ctx &c = *b.ctx;
  b.enqueue(rg, w, g, false);  // warm-up
  OS_HIP_CHECK(hipEventRecord(c.ev_start, c.stream));
  b.enqueue(rg, w, g, false);
  OS_HIP_CHECK(hipEventRecord(c.ev_stop, c.stream));
  OS_HIP_CHECK(hipEventSynchronize(c.ev_stop));
  float ms = 0.f;
  OS_HIP_CHECK(hipEventElapsedTime(&ms, c.ev_start, c.ev_stop));
  return ms;
}

// Times the candidate launch shapes on a short run of the batch itself.
void tune(CsrBatch &b, const CsrRun &full) {
  Ctx &c = *b.ctx;
  CsrRun probe = full;
  probe.num_sweeps = std::min(full.num_sweeps, 2);
  OS_HIP_CHECK(hipMemcpyAsync(b.run.ptr, &probe, sizeof(CsrRun), hipMemcpyHostToDevice, c.stream));
  float best_ms = 0.f;
  for (const bool rg : {false, true}) {
    if (rg && !csr_pf_eligible(b.max_width)) continue;
    for (int w : {1, 2, 4}) {
      const int maxb = csr_max_groups_per_cu(rg, w, b.max_words);
      if (maxb < 1) continue;
      std::vector<int> gpcs{maxb};
      if (maxb > 1) gpcs.push_back(std::max(1, maxb / 2));
      for (int gpc : gpcs) {
        const int g = std::max(1, std::min(c.num_cus * gpc, (b.num_items + w - 1) / w));
        const float ms = time_launch(b, rg, w, g);
        if (b.waves == 0 || ms < best_ms) {
          best_ms = ms;
          b.reg = rg;
          b.waves = w;
          b.groups = g;
        }
      }
    }
  }
  OS_REQUIRE(b.waves > 0, "csr batch: no launch shape fits (instance too large for LDS)");
}

}  // namespace

}  // namespace os

using os::CsrBatch;
using os::Ctx;

extern "C" {

int os_csr_batch_create(os_ctx *ctx_, int32_t num_instances, const int32_t *n,
                        const int32_t *row_ptr, const int32_t *col, const double *val,
                        const double *diag, os_csr_batch **out) {
  return os::guarded([&] {
    OS_REQUIRE(ctx_ && out, "null argument");
    OS_REQUIRE(num_instances >= 1, "num_instances must be positive");
    OS_REQUIRE(n && row_ptr && diag, "null instance arrays");
    *out = nullptr;
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    auto b = std::make_unique<CsrBatch>();
    b->ctx = &c;
    b->I = num_instances;
    // validate and lay out (host)
    int64_t rows = 0, nnz = 0, spins = 0, ell = 0;
    std::vector<int32_t> ell_c;
    std::vector<float> ell_v, dg;
    for (int i = 0; i < num_instances; ++i) {
      const int ni = n[i];
      OS_REQUIRE(ni >= 1, "instance size must be positive");
      const int32_t *rp = row_ptr + rows;
      OS_REQUIRE(rp[0] == 0, "row_ptr of each instance must start at 0");
      int maxdeg = 0;
      for (int j = 0; j < ni; ++j) {
        OS_REQUIRE(rp[j + 1] >= rp[j], "row_ptr must be non-decreasing");
        maxdeg = std::max(maxdeg, rp[j + 1] - rp[j]);
      }
      const int nz = rp[ni];
      OS_REQUIRE(nz == 0 || (col && val), "null col/val");
      for (int k = 0; k < nz; ++k) {
        const int l = col[nnz + k];
        OS_REQUIRE(l >= 0 && l < ni, "column index out of range");
      }
      os::CsrInst d{};
      d.n = ni;
      d.words = (ni + 1 + 31) / 32;
      d.width = std::max(4, (maxdeg + 3) / 4 * 4);
      d.ell_off = ell;
      d.spin_off = spins;
      d.row_off = rows;
      d.csr_off = nnz;
      ell_c.resize((size_t)(ell + (int64_t)ni * d.width), ni);
      ell_v.resize(ell_c.size(), 0.0f);
      for (int j = 0; j < ni; ++j)
        for (int k = rp[j]; k < rp[j + 1]; ++k) {
          const size_t at = (size_t)(ell + (int64_t)j * d.width + (k - rp[j]));
          ell_c[at] = col[nnz + k];
          ell_v[at] = (float)val[nnz + k];
        }
      for (int j = 0; j < ni; ++j) dg.push_back((float)diag[spins + j]);
      b->max_words = std::max(b->max_words, d.words);
      b->max_width = std::max(b->max_width, d.width);
      b->total_words += d.words;
      b->inst.push_back(d);
      ell += (int64_t)ni * d.width;
      rows += ni + 1;
      nnz += nz;
      spins += ni;
    }
    b->total_spins = spins;
    OS_REQUIRE(os::csr_lds_bytes(1, b->max_words) <= 160 * 1024,
               "csr batch: instance too large for the LDS-resident state (n <= 40000)");
    hipStream_t s = c.stream;
    b->ell_col.ensure(sizeof(int32_t) * std::max<size_t>(ell_c.size(), 4));
    b->ell_val.ensure(sizeof(float) * std::max<size_t>(ell_v.size(), 4));
    b->diag.ensure(sizeof(float) * spins);
    b->diag64.ensure(sizeof(double) * spins);
    b->row_ptr.ensure(sizeof(int32_t) * rows);
    b->col.ensure(sizeof(int32_t) * std::max<int64_t>(nnz, 1));
    b->val64.ensure(sizeof(double) * std::max<int64_t>(nnz, 1));
    OS_HIP_CHECK(hipMemcpyAsync(b->ell_col.ptr, ell_c.data(), sizeof(int32_t) * ell_c.size(),
                                hipMemcpyHostToDevice, s));
    OS_HIP_CHECK(hipMemcpyAsync(b->ell_val.ptr, ell_v.data(), sizeof(float) * ell_v.size(),
                                hipMemcpyHostToDevice, s));
    OS_HIP_CHECK(hipMemcpyAsync(b->diag.ptr, dg.data(), sizeof(float) * spins,
                                hipMemcpyHostToDevice, s));
    OS_HIP_CHECK(hipMemcpyAsync(b->diag64.ptr, diag, sizeof(double) * spins,
                                hipMemcpyHostToDevice, s));
    OS_HIP_CHECK(hipMemcpyAsync(b->row_ptr.ptr, row_ptr, sizeof(int32_t) * rows,
                                hipMemcpyHostToDevice, s));
    if (nnz) {
      OS_HIP_CHECK(hipMemcpyAsync(b->col.ptr, col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
      OS_HIP_CHECK(hipMemcpyAsync(b->val64.ptr, val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    }
    OS_HIP_CHECK(hipStreamSynchronize(s));
    *out = reinterpret_cast<os_csr_batch *>(b.release());
  });
}

int os_csr_batch_destroy(os_csr_batch *b_) {
  return os::guarded([&] {
    if (!b_) return;
    CsrBatch *b = reinterpret_cast<CsrBatch *>(b_);
    b->ctx->activate();
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
  });
}

int os_csr_batch_anneal(os_csr_batch *b_, const os_csr_params *p, double *best_energies,
                        double *run_best_energies, uint8_t *best_states, uint8_t *final_states,
                        double *inst_best_energy, int32_t *inst_best_replica,
                        os_sweep_stats *stats) {
  return os::guarded([&] {
    OS_REQUIRE(b_ && p, "null argument");
    OS_REQUIRE(p->num_replicas >= 1, "num_replicas must be positive");
    OS_REQUIRE(p->num_sweeps >= 0, "num_sweeps must be non-negative");
    OS_REQUIRE(p->num_sweeps == 0 || p->tscale, "tscale is null");
    CsrBatch &b = *reinterpret_cast<CsrBatch *>(b_);
    OS_REQUIRE((uint64_t)p->instance_base + (uint64_t)b.I <= (1u << 24),
               "instance ids must stay below 2^24");
    Ctx &c = *b.ctx;
    c.activate();
    hipStream_t s = c.stream;
    const int R = (int)p->num_replicas;
    const int Rp = (R + 63) / 64 * 64;
    const int S = p->num_sweeps;
    os::shape_for(b, Rp);
    b.tscale.ensure(sizeof(float) * std::max(S, 1));
    if (S) OS_HIP_CHECK(hipMemcpyAsync(b.tscale.ptr, p->tscale, sizeof(float) * S,
                                       hipMemcpyHostToDevice, s));
    if (b.g_tscale != b.tscale.ptr) b.drop_graph();
    os::CsrRun run{};
    run.num_items = b.num_items;
    run.num_sweeps = S;
    run.Rp = Rp;
    run.init_mode = p->init_states ? 1 : 0;
    run.k0 = (uint32_t)p->seed;
    run.k1 = (uint32_t)(p->seed >> 32);
    run.instance_base = p->instance_base;
    if (p->init_states) {
      // pack the caller's states (instance-major, [R x n_i]) into planes
      std::vector<uint32_t> planes((size_t)b.total_words * Rp, 0u);
      const uint8_t *src = p->init_states;
      for (const auto &d : b.inst) {
        for (int r = 0; r < R; ++r, src += d.n)
          for (int j = 0; j < d.n; ++j)
            if (src[j]) planes[(size_t)d.state_off + (size_t)(j >> 5) * Rp + r] |= 1u << (j & 31);
      }
      OS_HIP_CHECK(hipMemcpyAsync(b.init.ptr, planes.data(), sizeof(uint32_t) * planes.size(),
                                  hipMemcpyHostToDevice, s));
      OS_HIP_CHECK(hipStreamSynchronize(s));
    }
    if (b.waves == 0) {
      if (p->tune) {
        os::tune(b, run);
      } else {
        b.reg = os::csr_pf_eligible(b.max_width);
        b.waves = 1;
        b.groups = std::max(1, std::min(c.num_cus * os::csr_max_groups_per_cu(b.reg, 1, b.max_words),
                                        b.num_items));
      }
    }
    OS_HIP_CHECK(hipMemcpyAsync(b.run.ptr, &run, sizeof(os::CsrRun), hipMemcpyHostToDevice, s));
    if (p->use_graph && (!b.gexec || b.g_reg != b.reg || b.g_waves != b.waves ||
                         b.g_groups != b.groups ||
                         b.g_Rp != Rp)) {
      b.drop_graph();
      OS_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        b.enqueue(b.reg, b.waves, b.groups, true);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      OS_HIP_CHECK(hipStreamEndCapture(s, &b.graph));
      OS_HIP_CHECK(hipGraphInstantiate(&b.gexec, b.graph, nullptr, nullptr, 0));
      b.g_reg = b.reg;
      b.g_waves = b.waves;
      b.g_groups = b.groups;
      b.g_Rp = Rp;
      b.g_tscale = b.tscale.ptr;
    }
    OS_HIP_CHECK(hipEventRecord(c.ev_start, s));
    if (p->use_graph)
      OS_HIP_CHECK(hipGraphLaunch(b.gexec, s));
    else
      b.enqueue(b.reg, b.waves, b.groups, true);
    OS_HIP_CHECK(hipEventRecord(c.ev_stop, s));
    // results
    std::vector<double> e64((size_t)b.I * Rp);
    OS_HIP_CHECK(hipMemcpyAsync(e64.data(), b.e64.ptr, sizeof(double) * e64.size(),
                                hipMemcpyDeviceToHost, s));
    std::vector<double> ebh;
    if (run_best_energies) {
      ebh.resize((size_t)b.I * Rp);
      OS_HIP_CHECK(hipMemcpyAsync(ebh.data(), b.eb.ptr, sizeof(double) * ebh.size(),
                                  hipMemcpyDeviceToHost, s));
    }
    std::vector<uint32_t> bp, fp;
    if (best_states) {
      bp.resize((size_t)b.total_words * Rp);
      OS_HIP_CHECK(hipMemcpyAsync(bp.data(), b.best.ptr, sizeof(uint32_t) * bp.size(),
                                  hipMemcpyDeviceToHost, s));
    }
    if (final_states) {
      fp.resize((size_t)b.total_words * Rp);
      OS_HIP_CHECK(hipMemcpyAsync(fp.data(), b.fin.ptr, sizeof(uint32_t) * fp.size(),
                                  hipMemcpyDeviceToHost, s));
    }
    OS_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    OS_HIP_CHECK(hipEventElapsedTime(&ms, c.ev_start, c.ev_stop));
    uint8_t *bs = best_states, *fs = final_states;
    for (int i = 0; i < b.I; ++i) {
      const auto &d = b.inst[i];
      int arg = 0;
      for (int r = 0; r < R; ++r) {
        const double e = e64[(size_t)d.rep_off + r];
        if (best_energies) best_energies[(size_t)i * R + r] = e;
        if (run_best_energies) run_best_energies[(size_t)i * R + r] = ebh[(size_t)d.rep_off + r];
        if (e < e64[(size_t)d.rep_off + arg]) arg = r;
      }
      if (inst_best_energy) inst_best_energy[i] = e64[(size_t)d.rep_off + arg];
      if (inst_best_replica) inst_best_replica[i] = arg;
      if (bs || fs)
        for (int r = 0; r < R; ++r)
          for (int j = 0; j < d.n; ++j) {
            const size_t at = (size_t)d.state_off + (size_t)(j >> 5) * Rp + r;
            if (bs) bs[(size_t)r * d.n + j] = (uint8_t)((bp[at] >> (j & 31)) & 1u);
            if (fs) fs[(size_t)r * d.n + j] = (uint8_t)((fp[at] >> (j & 31)) & 1u);
          }
      if (bs) bs += (size_t)R * d.n;
      if (fs) fs += (size_t)R * d.n;
    }
    if (stats) {
      stats->kernel_ms = ms;
      stats->total_ms = ms;
      stats->kernel_launches = 1;
      stats->max_field_drift = 0.0;
      stats->attempts = (uint64_t)b.total_spins * (uint64_t)R * (uint64_t)S;
      stats->pass_avg_ms = 0.0;
      stats->pass_launches = 0;
    }
  });
}

int os_csr_batch_launch_shape(os_csr_batch *b_, int32_t *waves_per_group, int32_t *groups,
                              int32_t *prefetch) {
  return os::guarded([&] {
    OS_REQUIRE(b_, "null argument");
    const CsrBatch &b = *reinterpret_cast<CsrBatch *>(b_);
    if (waves_per_group) *waves_per_group = b.waves;
    if (groups) *groups = b.groups;
    if (prefetch) *prefetch = b.reg ? 1 : 0;
  });
}

}  // extern "C"
