// capi_csr.hip -- batched sparse (CSR) sweep-mode annealing entry points (os_csr_*).
//
// A batch keeps its instances resident on the GPU twice: as fp32 ELL rows
// (the sweep kernel's layout: row width padded to a multiple of 4 with the
// spin index n and weight 0) and as the caller's FP64 CSR (the energy check).
// A run anneals num_replicas replicas of every instance in one persistent
// kernel, then recomputes the FP64 energy of every best state; the three
// launches (counter reset, sweeps, energies) are captured once into a
// hipGraph and replayed, the run parameters living in device memory.  The
// launch shape (waves per work-group, resident work-groups and, for the
// colour-class site order, sites per wave step) is picked by timing short
// runs of the candidates on the batch itself.  Every instance is coloured
// greedily when the batch is created (index order, smallest colour free of
// lower-index neighbours) for the colour-class site order.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>
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
  // colour classes: sites class by class (offset spin_off), class starts
  // (offset class_off)
  DevBuf corder, cstart;
  std::vector<int32_t> h_order, h_cstart;
  // ELL rows and diagonal in that visiting order
  DevBuf pell_col, pell_val, pdiag;
  // per run shape
  int Rp = 0;
  DevBuf items, run, tscale, init, best, fin, eb, ef, e64, queue;
  DevBuf marg;  // warm-start marginals (total_spins floats)
  // population annealing: best planes and energies of every replica index
  // over the steps, ancestors of a resampling round
  DevBuf pop_best, pop_eb, pop_anc;
  // replica exchange: per-replica temperature scales ([I][Rp], read when
  // CsrRun::use_tsr is set)
  DevBuf tsr;
  // replica exchange rounds on the GPU: slots, replicas at slots, scales
  DevBuf pt_slots;
  int num_items = 0;
  // launch shape chosen by the tuner (0: not tuned); reg: row-prefetching kernel
  int waves = 0, groups = 0;
  bool reg = false;
  int max_width = 0;
  // colour-class site order: its launch shape (0: not tuned), the policy
  // that gives each instance its sites per step (c_uniform: one value for
  // all, else the smallest of 4, 8, 16 with at most c_limit steps per class
  // sweep of n sites), and the work items ((instance, group of 64/sites_i
  // replicas)) built for (items_uniform, items_limit, items_c_Rp)
  int c_waves = 0, c_groups = 0, c_uniform = 0, c_limit = 0;
  DevBuf items_c;
  int num_items_c = 0, items_uniform = -1, items_limit = -1, items_c_Rp = 0;
  int c_plane_words = 0;  // LDS state words per wave: max of words_i * 64/sites_i
  std::vector<int> order_by_work;  // instances, largest first
  // the run that enqueue() launches: 0 sequential, 1 colour classes
  int mode = 0;
  // captured graph and what it was captured for
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_waves = 0, g_groups = 0, g_Rp = 0, g_mode = 0, g_uniform = 0, g_limit = 0;
  bool g_energy = false;  // the graph ends with the FP64 energies of the best states
  bool g_reg = false;
  const void *g_tscale = nullptr, *g_items_c = nullptr;

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }
  ~CsrBatch() {
    drop_graph();
    for (auto *b : {&d_inst, &ell_col, &ell_val, &diag, &diag64, &row_ptr, &col, &val64, &items,
                    &run, &tscale, &init, &best, &fin, &eb, &ef, &e64, &queue, &marg, &corder,
                    &cstart, &items_c, &pell_col, &pell_val, &pdiag, &pop_best, &pop_eb,
                    &pop_anc, &tsr, &pt_slots})
      b->release();
  }

  // mode 0: sequential sweeps with shape (rg, w, g); mode 1: colour classes
  // with w waves and g work-groups over the items in items_c
  void enqueue(bool rg, int w, int g, bool energy) {
    hipStream_t st = ctx->stream;
    launch_csr_queue_reset(st, queue.as<int>());
    if (mode == 1)
      launch_csr_col_sweeps(st, w, g, c_plane_words, d_inst.as<CsrInst>(), items_c.as<int2>(),
                            run.as<CsrRun>(), ell_col.as<int4>(), ell_val.as<float4>(),
                            diag.as<float>(), corder.as<int32_t>(), cstart.as<int32_t>(),
                            pell_col.as<int4>(), pell_val.as<float4>(), pdiag.as<float>(),
                            tscale.as<float>(), tsr.as<float>(), init.as<uint32_t>(),
                            best.as<uint32_t>(),
                            fin.as<uint32_t>(), eb.as<double>(), ef.as<double>(),
                            queue.as<int>());
    else
      launch_csr_sweeps(st, rg, w, g, max_words, d_inst.as<CsrInst>(), items.as<int2>(),
                        run.as<CsrRun>(), ell_col.as<int4>(), ell_val.as<float4>(),
                        diag.as<float>(), tscale.as<float>(), tsr.as<float>(),
                        init.as<uint32_t>(), best.as<uint32_t>(), fin.as<uint32_t>(),
                        eb.as<double>(), ef.as<double>(), queue.as<int>());
    if (energy)
      launch_csr_energy(st, d_inst.as<CsrInst>(), items.as<int2>(), num_items, Rp, max_words,
                        row_ptr.as<int32_t>(), col.as<int32_t>(), val64.as<double>(),
                        diag64.as<double>(), best.as<uint32_t>(), e64.as<double>());
  }
};

// fn(0) .. fn(count - 1) on up to 8 host threads (fn must not throw)
template <class F>
void host_parallel_for(int count, F fn) {
  const int T = std::max(1, std::min({8, count, (int)std::thread::hardware_concurrency()}));
  if (T == 1) {
    for (int i = 0; i < count; ++i) fn(i);
    return;
  }
  std::atomic<int> next{0};
  auto work = [&] {
    for (int i; (i = next++) < count;) fn(i);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(work);
  work();
  for (auto &th : pool) th.join();
}

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
  b.order_by_work = order;
  b.items_uniform = -1;
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
  b.tsr.ensure(sizeof(float) * (size_t)std::max<int64_t>(rep, 1));
  b.queue.ensure(sizeof(int) * 64);
  b.run.ensure(sizeof(CsrRun));
  OS_HIP_CHECK(hipStreamSynchronize(s));
  b.Rp = Rp;
  b.drop_graph();
}

// Sites per step of an n-site instance under (uniform, limit).
int sites_for(int n, int uniform, int limit) {
  if (uniform) return uniform;
  if ((n + 3) / 4 <= limit) return 4;
  if ((n + 7) / 8 <= limit) return 8;
  return 16;
}

// Sets every instance's sites per step for the policy (uniform, limit) and
// builds the colour kernel's work items: (instance, group of 64/sites_i
// replicas), the items of longer estimated runtime first (steps per sweep
// times row cost).
void col_items(CsrBatch &b, int uniform, int limit) {
  if (b.items_uniform == uniform && b.items_limit == limit && b.items_c_Rp == b.Rp) return;
  std::vector<int> order(b.I);
  std::iota(order.begin(), order.end(), 0);
  std::vector<int64_t> est(b.I);
  int plane = 0;
  for (int i = 0; i < b.I; ++i) {
    CsrInst &d = b.inst[i];
    d.sites = sites_for(d.n, uniform, limit);
    est[i] = ((int64_t)(d.n + d.sites - 1) / d.sites + d.ncol) * (d.width + 8);
    plane = std::max(plane, d.words * (64 / d.sites));
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return est[a] > est[c]; });
  // whole instances to the 8 XCD lists, largest total work first onto the
  // least loaded list; within a list, the longest items first
  std::vector<int> by_total(order);
  auto groups_of = [&](int i) { return b.Rp / (64 / b.inst[i].sites); };
  std::stable_sort(by_total.begin(), by_total.end(), [&](int a, int c) {
    return est[a] * groups_of(a) > est[c] * groups_of(c);
  });
  int64_t load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<int> xcd_of(b.I);
  for (int i : by_total) {
    const int x = (int)(std::min_element(load, load + 8) - load);
    xcd_of[i] = x;
    load[x] += est[i] * groups_of(i);
  }
  std::vector<int2> items;
  int meta[16];
  for (int x = 0; x < 8; ++x) {
    meta[x] = (int)items.size();
    for (int i : order) {
      if (xcd_of[i] != x) continue;
      for (int gidx = 0; gidx < groups_of(i); ++gidx) items.push_back(make_int2(i, gidx));
    }
    meta[8 + x] = (int)items.size() - meta[x];
  }
  hipStream_t s = b.ctx->stream;
  OS_HIP_CHECK(hipMemcpyAsync(b.d_inst.ptr, b.inst.data(), sizeof(CsrInst) * b.I,
                              hipMemcpyHostToDevice, s));
  b.items_c.ensure(sizeof(int2) * std::max<size_t>(items.size(), 1));
  OS_HIP_CHECK(hipMemcpyAsync(b.items_c.ptr, items.data(), sizeof(int2) * items.size(),
                              hipMemcpyHostToDevice, s));
  // list starts and lengths behind the 8 counters (queue[8..23])
  OS_HIP_CHECK(hipMemcpyAsync(b.queue.as<int>() + 8, meta, sizeof(meta), hipMemcpyHostToDevice,
                              s));
  OS_HIP_CHECK(hipStreamSynchronize(s));
  b.num_items_c = (int)items.size();
  b.c_plane_words = plane;
  b.items_uniform = uniform;
  b.items_limit = limit;
  b.items_c_Rp = b.Rp;
}

// The run parameters in device memory for the launch enqueue() makes.
void put_run(CsrBatch &b, CsrRun r) {
  r.num_items = b.mode == 1 ? b.num_items_c : b.num_items;
  OS_HIP_CHECK(hipMemcpyAsync(b.run.ptr, &r, sizeof(CsrRun), hipMemcpyHostToDevice,
                              b.ctx->stream));
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

// Times the colour kernel's candidates -- the policy for sites per step (all
// 4; at most 512, 256, 128 or 64 steps per class sweep; all 16), waves per
// work-group, resident work-groups -- on a run of up to 8 sweeps of the
// batch itself: long enough that the items of the largest instances, which
// set the length of a run, weigh as they do in a full one.
void tune_col(CsrBatch &b, const CsrRun &full) {
  Ctx &c = *b.ctx;
  CsrRun probe = full;
  probe.num_sweeps = std::min(full.num_sweeps, 8);
  float best_ms = 0.f;
  for (int limit : {1 << 30, 512, 256, 128, 64, 0}) {
    col_items(b, 0, limit);
    put_run(b, probe);
    for (int w : {1, 2, 4}) {
      const int maxb = csr_col_max_groups_per_cu(w, b.c_plane_words);
      if (maxb < 1) continue;
      std::vector<int> gpcs{maxb};
      if (maxb > 1) gpcs.push_back(std::max(1, maxb / 2));
      for (int gpc : gpcs) {
        const int g = std::max(1, std::min(c.num_cus * gpc, (b.num_items_c + w - 1) / w));
        const float ms = time_launch(b, false, w, g);
        if (b.c_waves == 0 || ms < best_ms) {
          best_ms = ms;
          b.c_waves = w;
          b.c_groups = g;
          b.c_uniform = 0;
          b.c_limit = limit;
        }
      }
    }
  }
  OS_REQUIRE(b.c_waves > 0, "csr batch: no colour-kernel shape fits");
}

// Times the candidate launch shapes on a short run of the batch itself.
void tune(CsrBatch &b, const CsrRun &full) {
  Ctx &c = *b.ctx;
  CsrRun probe = full;
  probe.num_sweeps = std::min(full.num_sweeps, 2);
  put_run(b, probe);
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

// The per-sweep scales of a run in device memory (a new buffer drops the
// captured graph).
void put_tscale(CsrBatch &b, const float *ts, int S) {
  b.tscale.ensure(sizeof(float) * std::max(S, 1));
  if (S)
    OS_HIP_CHECK(hipMemcpyAsync(b.tscale.ptr, ts, sizeof(float) * S, hipMemcpyHostToDevice,
                                b.ctx->stream));
  if (b.g_tscale != b.tscale.ptr) b.drop_graph();
}

struct Shape {
  bool rg;
  int waves, groups;
  int uniform, limit;  // colour-class policy (mode 1)
};

// The launch shape of a run in `site_order` (tuned on first use when `tune`).
Shape select_shape(CsrBatch &b, const CsrRun &run, int site_order, bool tune_it) {
  Ctx &c = *b.ctx;
  b.mode = site_order;
  if (b.mode == 0 && b.waves == 0) {
    if (tune_it) {
      tune(b, run);
    } else {
      b.reg = csr_pf_eligible(b.max_width);
      b.waves = 1;
      b.groups = std::max(1, std::min(c.num_cus * csr_max_groups_per_cu(b.reg, 1, b.max_words),
                                      b.num_items));
    }
  }
  // OS_CSR_COLOUR_SITES (4, 8 or 16) fixes the colour kernel's sites per step
  if (b.mode == 1) {
    const char *senv = std::getenv("OS_CSR_COLOUR_SITES");
    const int forced = senv ? std::atoi(senv) : 0;
    OS_REQUIRE(!senv || forced == 4 || forced == 8 || forced == 16,
               "OS_CSR_COLOUR_SITES must be 4, 8 or 16");
    if (forced && forced != b.c_uniform) {
      b.c_uniform = forced;
      b.c_limit = 0;
      b.c_waves = 1;
      col_items(b, b.c_uniform, b.c_limit);
      b.c_groups = std::max(
          1, std::min(c.num_cus * csr_col_max_groups_per_cu(1, b.c_plane_words), b.num_items_c));
    }
  }
  if (b.mode == 1 && b.c_waves == 0) {
    if (tune_it) {
      tune_col(b, run);
    } else {
      // untuned: the shape the tuner picks for the mixed config-3 batch
      // (at most 256 steps per class sweep, 2 waves per group);
      // OS_CSR_COLOUR_LIMIT overrides the step limit
      const char *lenv = std::getenv("OS_CSR_COLOUR_LIMIT");
      const int lim = lenv ? std::atoi(lenv) : 256;
      OS_REQUIRE(!lenv || lim > 0, "OS_CSR_COLOUR_LIMIT must be a positive step count");
      b.c_uniform = 0;
      b.c_limit = lim;
      b.c_waves = 2;
      col_items(b, b.c_uniform, b.c_limit);
      b.c_groups = std::max(
          1, std::min(c.num_cus * csr_col_max_groups_per_cu(2, b.c_plane_words), b.num_items_c));
    }
  }
  if (b.mode == 1) col_items(b, b.c_uniform, b.c_limit);
  Shape sh;
  sh.rg = b.mode == 0 && b.reg;
  sh.waves = b.mode == 1 ? b.c_waves : b.waves;
  sh.groups = b.mode == 1 ? b.c_groups : b.groups;
  sh.uniform = b.mode == 1 ? b.c_uniform : 0;
  sh.limit = b.mode == 1 ? b.c_limit : 0;
  return sh;
}

// Captures (reset, sweeps, FP64 energies of the best states) for `sh` unless
// the graph already holds that shape.
void ensure_graph(CsrBatch &b, const Shape &sh, bool energy) {
  hipStream_t s = b.ctx->stream;
  if (b.gexec && b.g_mode == b.mode && b.g_reg == sh.rg && b.g_waves == sh.waves &&
      b.g_groups == sh.groups && b.g_uniform == sh.uniform && b.g_limit == sh.limit &&
      b.g_items_c == b.items_c.ptr && b.g_Rp == b.Rp && b.g_energy == energy)
    return;
  b.drop_graph();
  OS_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    b.enqueue(sh.rg, sh.waves, sh.groups, energy);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  OS_HIP_CHECK(hipStreamEndCapture(s, &b.graph));
  OS_HIP_CHECK(hipGraphInstantiate(&b.gexec, b.graph, nullptr, nullptr, 0));
  b.g_mode = b.mode;
  b.g_reg = sh.rg;
  b.g_waves = sh.waves;
  b.g_groups = sh.groups;
  b.g_uniform = sh.uniform;
  b.g_limit = sh.limit;
  b.g_items_c = b.items_c.ptr;
  b.g_Rp = b.Rp;
  b.g_energy = energy;
  b.g_tscale = b.tscale.ptr;
}

// Copies a run's results to the caller: FP64 energies of the best states
// (b.e64), the states of the `best` and `fin` planes and, when `eb` is
// given, the running energies it holds; synchronises the stream.
void collect(CsrBatch &b, int R, DevBuf &best, DevBuf &fin, DevBuf *eb, double *best_energies,
             double *run_best_energies, uint8_t *best_states, uint8_t *final_states,
             double *inst_best_energy, int32_t *inst_best_replica) {
  hipStream_t s = b.ctx->stream;
  const int Rp = b.Rp;
  std::vector<double> e64((size_t)b.I * Rp);
  OS_HIP_CHECK(hipMemcpyAsync(e64.data(), b.e64.ptr, sizeof(double) * e64.size(),
                              hipMemcpyDeviceToHost, s));
  std::vector<double> ebh;
  if (run_best_energies && eb) {
    ebh.resize((size_t)b.I * Rp);
    OS_HIP_CHECK(hipMemcpyAsync(ebh.data(), eb->ptr, sizeof(double) * ebh.size(),
                                hipMemcpyDeviceToHost, s));
  }
  std::vector<uint32_t> bp, fp;
  if (best_states) {
    bp.resize((size_t)b.total_words * Rp);
    OS_HIP_CHECK(hipMemcpyAsync(bp.data(), best.ptr, sizeof(uint32_t) * bp.size(),
                                hipMemcpyDeviceToHost, s));
  }
  if (final_states) {
    fp.resize((size_t)b.total_words * Rp);
    OS_HIP_CHECK(hipMemcpyAsync(fp.data(), fin.ptr, sizeof(uint32_t) * fp.size(),
                                hipMemcpyDeviceToHost, s));
  }
  OS_HIP_CHECK(hipStreamSynchronize(s));
  for (int i = 0; i < b.I; ++i) {
    const auto &d = b.inst[i];
    int arg = 0;
    for (int r = 0; r < R; ++r) {
      const double e = e64[(size_t)d.rep_off + r];
      if (best_energies) best_energies[(size_t)i * R + r] = e;
      if (!ebh.empty()) run_best_energies[(size_t)i * R + r] = ebh[(size_t)d.rep_off + r];
      if (e < e64[(size_t)d.rep_off + arg]) arg = r;
    }
    if (inst_best_energy) inst_best_energy[i] = e64[(size_t)d.rep_off + arg];
    if (inst_best_replica) inst_best_replica[i] = arg;
  }
  if (!best_states && !final_states) return;
  // planes [words][Rp] -> bytes [R][n] per instance: word-major, so the
  // planes are read in order and each word fills a run of up to 32 bytes;
  // instances are split over a few host threads for large batches
  std::vector<size_t> out_off(b.I + 1, 0);
  for (int i = 0; i < b.I; ++i) out_off[i + 1] = out_off[i] + (size_t)R * b.inst[i].n;
  auto unpack = [&](const std::vector<uint32_t> &pl, uint8_t *dst, int i0, int i1) {
    for (int i = i0; i < i1; ++i) {
      const auto &d = b.inst[i];
      uint8_t *o = dst + out_off[i];
      for (int w = 0; w * 32 < d.n; ++w) {
        const uint32_t *src = pl.data() + (size_t)d.state_off + (size_t)w * Rp;
        const int j0 = w * 32, len = std::min(32, d.n - j0);
        for (int r = 0; r < R; ++r) {
          const uint32_t x = src[r];
          uint8_t *row = o + (size_t)r * d.n + j0;
          for (int k = 0; k < len; ++k) row[k] = (uint8_t)((x >> k) & 1u);
        }
      }
    }
  };
  auto run = [&](int i0, int i1) {
    if (best_states) unpack(bp, best_states, i0, i1);
    if (final_states) unpack(fp, final_states, i0, i1);
  };
  const size_t bytes = out_off[b.I] * ((best_states ? 1 : 0) + (final_states ? 1 : 0));
  const int T = bytes < ((size_t)8 << 20) ? 1 : std::min(8, std::max(1, b.I));
  if (T == 1) {
    run(0, b.I);
    return;
  }
  // contiguous instance ranges of about equal output size
  std::vector<int> cut(T + 1, b.I);
  cut[0] = 0;
  for (int t = 1, i = 0; t < T; ++t) {
    while (i < b.I && out_off[i] < out_off[b.I] * t / T) ++i;
    cut[t] = i;
  }
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(run, cut[t], cut[t + 1]);
  run(cut[0], cut[1]);
  for (auto &th : pool) th.join();
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
    // Pass 1 (serial, O(n) per instance): sizes, offsets and the row_ptr
    // checks, up to the first instance that fails one.
    std::vector<os::CsrInst> ds;
    const char *bad_msg = nullptr;
    for (int i = 0; i < num_instances && !bad_msg; ++i) {
      const int ni = n[i];
      if (ni < 1) {
        bad_msg = "instance size must be positive";
        break;
      }
      const int32_t *rp = row_ptr + rows;
      if (rp[0] != 0) {
        bad_msg = "row_ptr of each instance must start at 0";
        break;
      }
      int maxdeg = 0;
      for (int j = 0; j < ni && !bad_msg; ++j) {
        if (rp[j + 1] < rp[j]) bad_msg = "row_ptr must be non-decreasing";
        maxdeg = std::max(maxdeg, rp[j + 1] - rp[j]);
      }
      if (bad_msg) break;
      const int nz = rp[ni];
      if (!(nz == 0 || (col && val))) {
        bad_msg = "null col/val";
        break;
      }
      os::CsrInst d{};
      d.n = ni;
      d.words = (ni + 1 + 31) / 32;
      d.width = std::max(4, (maxdeg + 3) / 4 * 4);
      d.ell_off = ell;
      d.spin_off = spins;
      d.row_off = rows;
      d.csr_off = nnz;
      ds.push_back(d);
      ell += (int64_t)ni * d.width;
      rows += ni + 1;
      nnz += nz;
      spins += ni;
    }
    // Pass 2 (host threads, one instance each): column checks, ELL rows
    // (padded with column n), the fp32 diagonal and the greedy colouring
    // (index order, smallest colour free of lower-index neighbours).
    const int nd = (int)ds.size();
    std::vector<int32_t> ell_c((size_t)ell);
    std::vector<float> ell_v((size_t)ell, 0.0f), dg((size_t)spins);
    std::vector<std::vector<int32_t>> cst(nd), cord(nd);
    std::vector<const char *> err(nd, nullptr);
    os::host_parallel_for(nd, [&](int i) {
      os::CsrInst &d = ds[i];
      const int ni = d.n;
      const int32_t *rp = row_ptr + d.row_off;
      const int32_t *cl = col + d.csr_off;
      const double *vl = val + d.csr_off;
      for (int k = 0; k < rp[ni]; ++k)
        if (cl[k] < 0 || cl[k] >= ni) {
          err[i] = "column index out of range";
          return;
        }
      int32_t *ec = ell_c.data() + d.ell_off;
      float *ev = ell_v.data() + d.ell_off;
      std::fill(ec, ec + (size_t)ni * d.width, ni);
      int maxdeg = 0;
      for (int j = 0; j < ni; ++j) {
        maxdeg = std::max(maxdeg, rp[j + 1] - rp[j]);
        for (int k = rp[j]; k < rp[j + 1]; ++k) {
          ec[(size_t)j * d.width + (k - rp[j])] = cl[k];
          ev[(size_t)j * d.width + (k - rp[j])] = (float)vl[k];
        }
      }
      for (int j = 0; j < ni; ++j) dg[(size_t)d.spin_off + j] = (float)diag[d.spin_off + j];
      std::vector<int32_t> colour(ni);
      std::vector<uint8_t> used((size_t)maxdeg + 2, 0);
      int ncol = 0;
      for (int j = 0; j < ni; ++j) {
        for (int k = rp[j]; k < rp[j + 1]; ++k)
          if (cl[k] < j) used[colour[cl[k]]] = 1;
        int cc = 0;
        while (used[cc]) ++cc;
        colour[j] = cc;
        ncol = std::max(ncol, cc + 1);
        for (int k = rp[j]; k < rp[j + 1]; ++k)
          if (cl[k] < j) used[colour[cl[k]]] = 0;
      }
      d.ncol = ncol;
      std::vector<int32_t> &start = cst[i];
      start.assign(ncol + 1, 0);
      for (int j = 0; j < ni; ++j) ++start[colour[j] + 1];
      for (int cc = 0; cc < ncol; ++cc) start[cc + 1] += start[cc];
      std::vector<int32_t> fill(start.begin(), start.end() - 1);
      cord[i].resize(ni);
      for (int j = 0; j < ni; ++j) cord[i][fill[colour[j]]++] = j;
    });
    // the first failing instance in order, as a serial pass reports it
    for (int i = 0; i < nd; ++i) OS_REQUIRE(!err[i], err[i]);
    OS_REQUIRE(!bad_msg, bad_msg);
    // Pass 3 (serial): class offsets and the concatenated class arrays
    for (int i = 0; i < nd; ++i) {
      os::CsrInst &d = ds[i];
      d.class_off = (int64_t)b->h_cstart.size();
      b->h_cstart.insert(b->h_cstart.end(), cst[i].begin(), cst[i].end());
      b->h_order.insert(b->h_order.end(), cord[i].begin(), cord[i].end());
      b->max_words = std::max(b->max_words, d.words);
      b->max_width = std::max(b->max_width, d.width);
      b->total_words += d.words;
    }
    b->inst = ds;
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
    // the ELL rows and diagonal in each instance's colour-class order
    {
      std::vector<int32_t> pc(ell_c.size());
      std::vector<float> pv(ell_v.size()), pd(dg.size());
      os::host_parallel_for((int)b->inst.size(), [&](int i) {
        const auto &d = b->inst[i];
        const int32_t *ord = b->h_order.data() + d.spin_off;
        for (int pos = 0; pos < d.n; ++pos) {
          const int j = ord[pos];
          std::memcpy(&pc[(size_t)(d.ell_off + (int64_t)pos * d.width)],
                      &ell_c[(size_t)(d.ell_off + (int64_t)j * d.width)], sizeof(int32_t) * d.width);
          std::memcpy(&pv[(size_t)(d.ell_off + (int64_t)pos * d.width)],
                      &ell_v[(size_t)(d.ell_off + (int64_t)j * d.width)], sizeof(float) * d.width);
          pd[(size_t)(d.spin_off + pos)] = dg[(size_t)(d.spin_off + j)];
        }
      });
      b->pell_col.ensure(sizeof(int32_t) * std::max<size_t>(pc.size(), 4));
      b->pell_val.ensure(sizeof(float) * std::max<size_t>(pv.size(), 4));
      b->pdiag.ensure(sizeof(float) * std::max<size_t>(pd.size(), 1));
      OS_HIP_CHECK(hipMemcpyAsync(b->pell_col.ptr, pc.data(), sizeof(int32_t) * pc.size(),
                                  hipMemcpyHostToDevice, s));
      OS_HIP_CHECK(hipMemcpyAsync(b->pell_val.ptr, pv.data(), sizeof(float) * pv.size(),
                                  hipMemcpyHostToDevice, s));
      OS_HIP_CHECK(hipMemcpyAsync(b->pdiag.ptr, pd.data(), sizeof(float) * pd.size(),
                                  hipMemcpyHostToDevice, s));
      OS_HIP_CHECK(hipStreamSynchronize(s));
    }
    b->corder.ensure(sizeof(int32_t) * b->h_order.size());
    b->cstart.ensure(sizeof(int32_t) * b->h_cstart.size());
    OS_HIP_CHECK(hipMemcpyAsync(b->corder.ptr, b->h_order.data(),
                                sizeof(int32_t) * b->h_order.size(), hipMemcpyHostToDevice, s));
    OS_HIP_CHECK(hipMemcpyAsync(b->cstart.ptr, b->h_cstart.data(),
                                sizeof(int32_t) * b->h_cstart.size(), hipMemcpyHostToDevice, s));
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
    OS_REQUIRE(p->site_order == 0 || p->site_order == 1, "site_order must be 0 or 1");
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
    os::put_tscale(b, p->tscale, S);
    os::CsrRun run{};
    run.num_items = b.num_items;
    run.num_sweeps = S;
    run.Rp = Rp;
    OS_REQUIRE(!(p->init_states && p->init_marginals),
               "give init_states or init_marginals, not both");
    run.init_mode = p->init_states || p->init_marginals ? 1 : 0;
    run.k0 = (uint32_t)p->seed;
    run.k1 = (uint32_t)(p->seed >> 32);
    run.instance_base = p->instance_base;
    if (p->init_marginals) {
      // start planes sampled on the GPU from the marginals
      b.marg.ensure(sizeof(float) * (size_t)std::max<int64_t>(b.total_spins, 1));
      OS_HIP_CHECK(hipMemcpyAsync(b.marg.ptr, p->init_marginals,
                                  sizeof(float) * (size_t)b.total_spins, hipMemcpyHostToDevice,
                                  s));
      os::launch_csr_warm_init(s, b.d_inst.as<os::CsrInst>(), b.I, R, Rp, p->seed,
                               p->instance_base, b.marg.as<float>(), b.init.as<uint32_t>());
    }
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
    const os::Shape sh = os::select_shape(b, run, p->site_order, p->tune != 0);
    os::put_run(b, run);
    if (p->use_graph) os::ensure_graph(b, sh, true);
    OS_HIP_CHECK(hipEventRecord(c.ev_start, s));
    if (p->use_graph)
      OS_HIP_CHECK(hipGraphLaunch(b.gexec, s));
    else
      b.enqueue(sh.rg, sh.waves, sh.groups, true);
    OS_HIP_CHECK(hipEventRecord(c.ev_stop, s));
    os::collect(b, R, b.best, b.fin, run_best_energies ? &b.eb : nullptr, best_energies,
                run_best_energies, best_states, final_states, inst_best_energy,
                inst_best_replica);
    float ms = 0.f;
    OS_HIP_CHECK(hipEventElapsedTime(&ms, c.ev_start, c.ev_stop));
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

int os_csr_batch_population(os_csr_batch *b_, const os_csr_population_params *p,
                            double *best_energies, uint8_t *best_states, uint8_t *final_states,
                            uint32_t *ancestors, double *inst_best_energy,
                            int32_t *inst_best_replica, os_sweep_stats *stats) {
  return os::guarded([&] {
    OS_REQUIRE(b_ && p, "null argument");
    OS_REQUIRE(p->num_replicas >= 1, "num_replicas must be positive");
    OS_REQUIRE(p->num_steps >= 1 && p->sweeps_per_step >= 1 && p->tscale,
               "invalid population schedule");
    OS_REQUIRE((int64_t)p->num_steps * p->sweeps_per_step <= (int64_t)INT32_MAX,
               "population schedule too long");
    OS_REQUIRE(p->site_order == 0 || p->site_order == 1, "site_order must be 0 or 1");
    for (int k = 0; k < p->num_steps; ++k)
      OS_REQUIRE(p->tscale[k] > 0.0f && std::isfinite(p->tscale[k]),
                 "temperature scales must be positive");
    CsrBatch &b = *reinterpret_cast<CsrBatch *>(b_);
    OS_REQUIRE((uint64_t)p->instance_base + (uint64_t)b.I <= (1u << 24),
               "instance ids must stay below 2^24");
    Ctx &c = *b.ctx;
    c.activate();
    hipStream_t s = c.stream;
    const int R = (int)p->num_replicas;
    const int Rp = (R + 63) / 64 * 64;
    const int K = p->num_steps, theta = p->sweeps_per_step;
    os::shape_for(b, Rp);
    // the per-sweep scales: tscale[k] for the sweeps of step k
    std::vector<float> ts((size_t)K * theta);
    for (int k = 0; k < K; ++k)
      std::fill(ts.begin() + (size_t)k * theta, ts.begin() + (size_t)(k + 1) * theta,
                p->tscale[k]);
    os::put_tscale(b, ts.data(), K * theta);
    os::CsrRun run{};
    run.num_items = b.num_items;
    run.num_sweeps = theta;
    run.Rp = Rp;
    run.k0 = (uint32_t)p->seed;
    run.k1 = (uint32_t)(p->seed >> 32);
    run.instance_base = p->instance_base;
    const os::Shape sh = os::select_shape(b, run, p->site_order, p->tune != 0);
    if (p->use_graph) os::ensure_graph(b, sh, false);
    const size_t planes = sizeof(uint32_t) * (size_t)std::max<int64_t>(b.total_words * Rp, 1);
    b.pop_best.ensure(planes);
    b.pop_eb.ensure(sizeof(double) * (size_t)b.I * Rp);
    b.pop_anc.ensure(sizeof(uint32_t) * (size_t)b.I * R);
    std::vector<double> ef((size_t)b.I * Rp);
    std::vector<uint32_t> anc((size_t)b.I * R);
    std::vector<double> wsum(R);
    OS_HIP_CHECK(hipEventRecord(c.ev_start, s));
    for (int k = 0; k < K; ++k) {
      run.sweep_base = k * theta;
      run.init_mode = k > 0 ? 1 : 0;
      os::put_run(b, run);
      if (p->use_graph)
        OS_HIP_CHECK(hipGraphLaunch(b.gexec, s));
      else
        b.enqueue(sh.rg, sh.waves, sh.groups, false);
      // best of every replica index: the step's best where it is lower
      os::launch_csr_pop_best(s, b.d_inst.as<os::CsrInst>(), b.items.as<int2>(), b.num_items, Rp,
                              k == 0, b.eb.as<double>(), b.best.as<uint32_t>(),
                              b.pop_eb.as<double>(), b.pop_best.as<uint32_t>());
      if (k + 1 >= K) break;
      // resampling round k on the running energies at the step's end
      OS_HIP_CHECK(hipMemcpyAsync(ef.data(), b.ef.ptr, sizeof(double) * ef.size(),
                                  hipMemcpyDeviceToHost, s));
      OS_HIP_CHECK(hipStreamSynchronize(s));
      const double dbeta = 1.0 / (double)p->tscale[k + 1] - 1.0 / (double)p->tscale[k];
      for (int i = 0; i < b.I; ++i) {
        const uint32_t tag = ((p->instance_base + (uint32_t)i) << 8) | os::TAG_RESAMPLE;
        const double u = os::resample_uniform(run.k0, run.k1, tag, (uint32_t)k);
        uint32_t *a = anc.data() + (size_t)i * R;
        os::resample(ef.data() + b.inst[i].rep_off, R, dbeta, u, a, wsum.data());
        if (ancestors)
          std::memcpy(ancestors + ((size_t)i * (K - 1) + k) * R, a, sizeof(uint32_t) * R);
      }
      OS_HIP_CHECK(hipMemcpyAsync(b.pop_anc.ptr, anc.data(), sizeof(uint32_t) * anc.size(),
                                  hipMemcpyHostToDevice, s));
      os::launch_csr_pop_gather(s, b.d_inst.as<os::CsrInst>(), b.items.as<int2>(), b.num_items,
                                Rp, R, b.pop_anc.as<uint32_t>(), b.fin.as<uint32_t>(),
                                b.init.as<uint32_t>());
    }
    OS_HIP_CHECK(hipEventRecord(c.ev_stop, s));
    // FP64 energies of the best states of every replica index
    os::launch_csr_energy(s, b.d_inst.as<os::CsrInst>(), b.items.as<int2>(), b.num_items, Rp,
                          b.max_words, b.row_ptr.as<int32_t>(), b.col.as<int32_t>(), b.val64.as<double>(),
                          b.diag64.as<double>(), b.pop_best.as<uint32_t>(), b.e64.as<double>());
    os::collect(b, R, b.pop_best, b.fin, nullptr, best_energies, nullptr, best_states,
                final_states, inst_best_energy, inst_best_replica);
    float ms = 0.f;
    OS_HIP_CHECK(hipEventElapsedTime(&ms, c.ev_start, c.ev_stop));
    if (stats) {
      stats->kernel_ms = ms;
      stats->total_ms = ms;
      stats->kernel_launches = K;
      stats->max_field_drift = 0.0;
      stats->attempts = (uint64_t)b.total_spins * (uint64_t)R * (uint64_t)K * (uint64_t)theta;
      stats->pass_avg_ms = 0.0;
      stats->pass_launches = 0;
    }
  });
}

int os_csr_batch_tempering(os_csr_batch *b_, const os_csr_tempering_params *p,
                           double *best_energies, uint8_t *best_states, uint8_t *final_states,
                           uint32_t *final_slots, double *inst_best_energy,
                           int32_t *inst_best_replica, os_sweep_stats *stats) {
  return os::guarded([&] {
    OS_REQUIRE(b_ && p, "null argument");
    OS_REQUIRE(p->num_ladders >= 1 && p->num_temperatures >= 2 && p->tscale &&
                   (uint64_t)p->num_ladders * p->num_temperatures <= (1u << 24),
               "invalid ladder shape");
    OS_REQUIRE(p->num_sweeps >= 1, "num_sweeps must be positive");
    OS_REQUIRE(p->exchange_interval >= 1, "exchange_interval must be positive");
    OS_REQUIRE(p->site_order == 0 || p->site_order == 1, "site_order must be 0 or 1");
    const int K = (int)p->num_temperatures;
    for (int k = 0; k < K; ++k)
      OS_REQUIRE(p->tscale[k] > 0.0f && std::isfinite(p->tscale[k]),
                 "temperature scales must be positive");
    CsrBatch &b = *reinterpret_cast<CsrBatch *>(b_);
    OS_REQUIRE((uint64_t)p->instance_base + (uint64_t)b.I <= (1u << 24),
               "instance ids must stay below 2^24");
    Ctx &c = *b.ctx;
    c.activate();
    hipStream_t s = c.stream;
    const int L = (int)p->num_ladders;
    const int R = K * L;
    const int Rp = (R + 63) / 64 * 64;
    const int S = p->num_sweeps, iv = p->exchange_interval;
    os::shape_for(b, Rp);
    // per-sweep scales are not read (use_tsr); the array still covers S sweeps
    std::vector<float> ones((size_t)S, 1.0f);
    os::put_tscale(b, ones.data(), S);
    os::CsrRun run{};
    run.num_items = b.num_items;
    run.Rp = Rp;
    run.k0 = (uint32_t)p->seed;
    run.k1 = (uint32_t)(p->seed >> 32);
    run.instance_base = p->instance_base;
    run.num_sweeps = std::min(S, iv);
    const os::Shape sh = os::select_shape(b, run, p->site_order, p->tune != 0);
    if (p->use_graph) os::ensure_graph(b, sh, false);
    const size_t planes = sizeof(uint32_t) * (size_t)std::max<int64_t>(b.total_words * Rp, 1);
    b.pop_best.ensure(planes);
    b.pop_eb.ensure(sizeof(double) * (size_t)b.I * Rp);
    // slot of every replica and replica at every slot, per instance; the
    // scale of every row (padding rows: slot 0)
    std::vector<uint32_t> slot_of((size_t)b.I * R), rep_at((size_t)b.I * R);
    for (size_t i = 0; i < slot_of.size(); ++i) {
      slot_of[i] = (uint32_t)(i % R % K);
      rep_at[i] = (uint32_t)(i % R);
    }
    std::vector<float> tsr_h((size_t)b.I * Rp, p->tscale[0]);
    for (int i = 0; i < b.I; ++i)
      for (int r = 0; r < R; ++r)
        tsr_h[(size_t)b.inst[i].rep_off + r] = p->tscale[slot_of[(size_t)i * R + r]];
    // the exchange rounds run on the GPU between segments: slots, replicas
    // at slots, and the slots' scales in device memory
    b.pt_slots.ensure(sizeof(uint32_t) * slot_of.size() * 2 + sizeof(float) * K);
    uint32_t *d_slot = b.pt_slots.as<uint32_t>();
    uint32_t *d_rep = d_slot + slot_of.size();
    float *d_ts = reinterpret_cast<float *>(d_rep + rep_at.size());
    OS_HIP_CHECK(hipMemcpyAsync(d_slot, slot_of.data(), sizeof(uint32_t) * slot_of.size(),
                                hipMemcpyHostToDevice, s));
    OS_HIP_CHECK(hipMemcpyAsync(d_rep, rep_at.data(), sizeof(uint32_t) * rep_at.size(),
                                hipMemcpyHostToDevice, s));
    OS_HIP_CHECK(hipMemcpyAsync(d_ts, p->tscale, sizeof(float) * K, hipMemcpyHostToDevice, s));
    OS_HIP_CHECK(hipMemcpyAsync(b.tsr.ptr, tsr_h.data(), sizeof(float) * tsr_h.size(),
                                hipMemcpyHostToDevice, s));
    OS_HIP_CHECK(hipStreamSynchronize(s));
    uint32_t round = 0;
    int segments = 0;
    OS_HIP_CHECK(hipEventRecord(c.ev_start, s));
    for (int s0 = 0; s0 < S; s0 += iv) {
      const int ns = std::min(iv, S - s0);
      run.sweep_base = s0;
      run.num_sweeps = ns;
      run.init_mode = s0 > 0 ? 1 : 0;
      run.use_tsr = 1;
      os::put_run(b, run);
      if (p->use_graph)
        OS_HIP_CHECK(hipGraphLaunch(b.gexec, s));
      else
        b.enqueue(sh.rg, sh.waves, sh.groups, false);
      ++segments;
      os::launch_csr_pop_best(s, b.d_inst.as<os::CsrInst>(), b.items.as<int2>(), b.num_items, Rp,
                              s0 == 0, b.eb.as<double>(), b.best.as<uint32_t>(),
                              b.pop_eb.as<double>(), b.pop_best.as<uint32_t>());
      if (s0 + ns >= S) break;
      // exchange round on the running energies at the segment's end, and
      // the next segment continues from the final states
      os::launch_exchange_round(s, b.I, L, K, R, Rp, round, p->seed, p->instance_base, d_ts,
                                b.ef.as<double>(), d_rep, d_slot, b.tsr.as<float>());
      ++round;
      OS_HIP_CHECK(hipMemcpyAsync(b.init.ptr, b.fin.ptr, planes, hipMemcpyDeviceToDevice, s));
    }
    OS_HIP_CHECK(hipEventRecord(c.ev_stop, s));
    os::launch_csr_energy(s, b.d_inst.as<os::CsrInst>(), b.items.as<int2>(), b.num_items, Rp,
                          b.max_words, b.row_ptr.as<int32_t>(), b.col.as<int32_t>(), b.val64.as<double>(),
                          b.diag64.as<double>(), b.pop_best.as<uint32_t>(), b.e64.as<double>());
    os::collect(b, R, b.pop_best, b.fin, nullptr, best_energies, nullptr, best_states,
                final_states, inst_best_energy, inst_best_replica);
    if (final_slots) {
      OS_HIP_CHECK(hipMemcpyAsync(final_slots, d_slot, sizeof(uint32_t) * slot_of.size(),
                                  hipMemcpyDeviceToHost, s));
      OS_HIP_CHECK(hipStreamSynchronize(s));
    }
    float ms = 0.f;
    OS_HIP_CHECK(hipEventElapsedTime(&ms, c.ev_start, c.ev_stop));
    if (stats) {
      stats->kernel_ms = ms;
      stats->total_ms = ms;
      stats->kernel_launches = segments;
      stats->max_field_drift = 0.0;
      stats->attempts = (uint64_t)b.total_spins * (uint64_t)R * (uint64_t)S;
      stats->pass_avg_ms = 0.0;
      stats->pass_launches = 0;
    }
  });
}

}  // extern "C"

namespace os {
// The batch's device arrays and the context it lives on (for the network).
CsrView csr_view(os_csr_batch *b_, Ctx **ctx) {
  OS_REQUIRE(b_ != nullptr, "null batch");
  CsrBatch &b = *reinterpret_cast<CsrBatch *>(b_);
  // the descriptors reach the device with a batch's first run (shape_for);
  // a batch that has not run yet gets them here (n and the CSR offsets are
  // fixed at creation)
  b.ctx->activate();
  b.d_inst.ensure(sizeof(CsrInst) * std::max(b.I, 1));
  OS_HIP_CHECK(hipMemcpyAsync(b.d_inst.ptr, b.inst.data(), sizeof(CsrInst) * b.I,
                              hipMemcpyHostToDevice, b.ctx->stream));
  CsrView v;
  v.num_instances = b.I;
  v.total_spins = b.total_spins;
  for (const auto &d : b.inst) v.max_n = std::max(v.max_n, d.n);
  v.inst = b.d_inst.as<CsrInst>();
  v.row_ptr = b.row_ptr.as<int32_t>();
  v.col = b.col.as<int32_t>();
  v.val = b.val64.as<double>();
  v.diag = b.diag64.as<double>();
  *ctx = b.ctx;
  return v;
}
}  // namespace os

extern "C" {

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

int os_csr_batch_colour_shape(os_csr_batch *b_, int32_t *waves_per_group, int32_t *groups,
                              int32_t *sites_per_step) {
  return os::guarded([&] {
    OS_REQUIRE(b_, "null argument");
    const CsrBatch &b = *reinterpret_cast<CsrBatch *>(b_);
    if (waves_per_group) *waves_per_group = b.c_waves;
    if (groups) *groups = b.c_groups;
    if (sites_per_step) {
      // the common value, or 0 when the instances differ (or before a run)
      int v = b.items_uniform >= 0 ? b.inst[0].sites : 0;
      for (const auto &d : b.inst)
        if (d.sites != v) v = 0;
      *sites_per_step = v;
    }
  });
}

int os_csr_batch_colour_sites(os_csr_batch *b_, int32_t *sites_per_step) {
  return os::guarded([&] {
    OS_REQUIRE(b_ && sites_per_step, "null argument");
    const CsrBatch &b = *reinterpret_cast<CsrBatch *>(b_);
    for (int i = 0; i < b.I; ++i)
      sites_per_step[i] = b.items_uniform >= 0 ? b.inst[i].sites : 0;
  });
}

int os_csr_batch_site_order(os_csr_batch *b_, int32_t instance, int32_t *order,
                            int32_t *class_start, int32_t *num_classes) {
  return os::guarded([&] {
    OS_REQUIRE(b_, "null argument");
    const CsrBatch &b = *reinterpret_cast<CsrBatch *>(b_);
    OS_REQUIRE(instance >= 0 && instance < b.I, "instance out of range");
    const os::CsrInst &d = b.inst[instance];
    if (order)
      std::memcpy(order, b.h_order.data() + d.spin_off, sizeof(int32_t) * d.n);
    if (class_start)
      std::memcpy(class_start, b.h_cstart.data() + d.class_off, sizeof(int32_t) * (d.ncol + 1));
    if (num_classes) *num_classes = d.ncol;
  });
}

}  // extern "C"
