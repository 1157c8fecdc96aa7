// anneal_csr.hip -- batched sweep-mode annealing of sparse QUBOs (gfx950).
//
// One persistent kernel anneals a whole batch of instances of uneven sizes.
// Work items are (instance, group of 64 replicas); waves take items from a
// global counter in the order the host gives (largest first), so one launch
// drains the batch whatever its size mix.  A wave runs its 64 replicas in
// lockstep, lane = replica: the site loop is uniform, so the coupling row of
// site j (ELL layout: `width` columns padded with the spin index n, whose bit
// is always 0, and their weights) is read once per wave with scalar loads and
// shared by all 64 replicas, and each lane gathers its own neighbour bits
// from LDS, where the states are stored bit-packed and transposed
// ([word][replica], conflict-free for a wave).
//
// Per site (reference semantics of a Metropolis sweep, in the oracle's
// order): h = diag_j, then h = h + w_jk for each set neighbour in row order
// (fp32, no contraction); dE = x_j ? -h : h; flip iff dE < neg_ln_u(w) *
// tscale[t], w = Philox word of counter (j>>2, replica, t, tag), t the
// sweep's number (run->sweep_base + the launch's own sweep index); the
// FP64 running energy adds dE.  At the end of every sweep a replica whose
// running energy is below its best copies its state to the best-state planes.
#include "os_kernels.h"

namespace os {

namespace {

// This wave's LDS writes are done before its later LDS reads (other lanes
// read what a lane wrote).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Bit j of replica lane in a wave's transposed planes (the bit-field extract
// reads only the low 5 bits of its offset).
__device__ __forceinline__ uint32_t plane_bit(const uint32_t *xs, int j, int lane) {
  return __builtin_amdgcn_ubfe(xs[(j >> 5) * 64 + lane], (uint32_t)j, 1u);
}

// h + w if neighbour l's bit is set, else h: fma(1, w, h) is the rounded sum
// and fma(0, w, h) is h (up to the sign of a zero h, which no comparison or
// later sum sees), with one instruction fewer than a select.
__device__ __forceinline__ float add_if_set(const uint32_t *xs, int l, int lane, float w,
                                            float h) {
  return fmaf((float)plane_bit(xs, l, lane), w, h);
}

__device__ __forceinline__ float row_field(const uint32_t *xs, int lane, const int4 *ecol,
                                           const float4 *eval, int chunks, float h) {
  for (int c = 0; c < chunks; ++c) {
    const int4 cc = ecol[c];
    const float4 vv = eval[c];
    h = add_if_set(xs, cc.x, lane, vv.x, h);
    h = add_if_set(xs, cc.y, lane, vv.y, h);
    h = add_if_set(xs, cc.z, lane, vv.z, h);
    h = add_if_set(xs, cc.w, lane, vv.w, h);
  }
  return h;
}

}  // namespace

// Work-queue counters: queue[0] for the sequential kernels, queue[0..7] (one
// per XCD) for the colour-class kernel, whose item lists start at queue[8+x]
// and hold queue[16+x] items (written by the host, kept across resets).
__global__ void csr_queue_reset_kernel(int *queue) {
  if (threadIdx.x < 8) queue[threadIdx.x] = 0;
}

template <int kW>
__global__ __launch_bounds__(64 * kW) void csr_sweep_kernel(
    const CsrInst *__restrict__ inst, const int2 *__restrict__ items,
    const CsrRun *__restrict__ run, const int4 *__restrict__ ell_col,
    const float4 *__restrict__ ell_val, const float *__restrict__ diag,
    const float *__restrict__ tscale, const float *__restrict__ tsr,
    const uint32_t *__restrict__ init_bits, uint32_t *__restrict__ best_bits,
    uint32_t *__restrict__ final_bits, double *__restrict__ e_best,
    double *__restrict__ e_final, int *__restrict__ queue, int max_words) {
  extern __shared__ uint32_t lds[];
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  uint32_t *xs = lds + (size_t)w * max_words * 64;
  const int num_items = run->num_items, S = run->num_sweeps, Rp = run->Rp;
  const uint32_t k0 = run->k0, k1 = run->k1;
  const uint32_t base = run->instance_base;
  const int init_mode = run->init_mode;
  const int tb = run->sweep_base;
  const int use_tsr = run->use_tsr;
  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(queue, 1);
    it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
    if (it >= num_items) break;
    const int2 wi = items[it];
    const CsrInst d = inst[wi.x];
    const int n = d.n, words = d.words, chunks = d.width >> 2;
    const uint32_t r = (uint32_t)(wi.y * 64 + lane);
    const uint32_t id = (base + (uint32_t)wi.x) << 8;
    const int4 *ecol = ell_col + (d.ell_off >> 2);
    const float4 *eval = ell_val + (d.ell_off >> 2);
    const float *dg = diag + d.spin_off;
    uint32_t *bb = best_bits + d.state_off + r;
    // start state: Philox bits (spin j: counter (j>>7, r, 0, init tag), word
    // (j>>5)&3, bit j&31) or the caller's planes; bits >= n are zero
    for (int wd = 0; wd < words; ++wd) {
      uint32_t v;
      if (init_mode) {
        v = init_bits[d.state_off + (size_t)wd * Rp + r];
      } else {
        const uint4 o = philox4x32_10(make_uint4((uint32_t)(wd >> 2), r, 0u, id | TAG_SWEEP_INIT),
                                      k0, k1);
        v = pick(o, (uint32_t)wd & 3u);
      }
      const int valid = n - 32 * wd;
      if (valid < 32) v = valid <= 0 ? 0u : (v & ((1u << valid) - 1u));
      xs[wd * 64 + lane] = v;
    }
    // FP64 energy in the oracle's order: sum over set j of diag_j + off_j / 2
    double e = 0.0;
    for (int j = 0; j < n; ++j) {
      double off = 0.0;
      for (int c = 0; c < chunks; ++c) {
        const int4 cc = ecol[(size_t)j * chunks + c];
        const float4 vv = eval[(size_t)j * chunks + c];
        off = plane_bit(xs, cc.x, lane) ? off + (double)vv.x : off;
        off = plane_bit(xs, cc.y, lane) ? off + (double)vv.y : off;
        off = plane_bit(xs, cc.z, lane) ? off + (double)vv.z : off;
        off = plane_bit(xs, cc.w, lane) ? off + (double)vv.w : off;
      }
      if (plane_bit(xs, j, lane)) e += (double)dg[j] + 0.5 * off;
    }
    double eb = e;
    for (int wd = 0; wd < words; ++wd) bb[(size_t)wd * Rp] = xs[wd * 64 + lane];
    for (int t = 0; t < S; ++t) {
      const float ts = use_tsr ? tsr[d.rep_off + r] : tscale[tb + t];
      for (int j4 = 0; j4 < n; j4 += 4) {
        const uint4 o = philox4x32_10(
            make_uint4((uint32_t)(j4 >> 2), r, (uint32_t)(tb + t), id | TAG_SWEEP_THR), k0, k1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int j = j4 + q;
          if (j >= n) break;
          const float thr = neg_ln_u(pick(o, (uint32_t)q)) * ts;
          const float h = row_field(xs, lane, ecol + (size_t)j * chunks,
                                    eval + (size_t)j * chunks, chunks, dg[j]);
          uint32_t *xw = &xs[(j >> 5) * 64 + lane];
          const uint32_t word = *xw;
          const bool xj = (word >> (j & 31)) & 1u;
          const float de = xj ? -h : h;
          if (de < thr) {
            *xw = word ^ (1u << (j & 31));
            e += (double)de;
          }
        }
      }
      if (e < eb) {
        eb = e;
        for (int wd = 0; wd < words; ++wd) bb[(size_t)wd * Rp] = xs[wd * 64 + lane];
      }
    }
    uint32_t *fb = final_bits + d.state_off + r;
    for (int wd = 0; wd < words; ++wd) fb[(size_t)wd * Rp] = xs[wd * 64 + lane];
    e_best[d.rep_off + r] = eb;
    e_final[d.rep_off + r] = e;
  }
}

namespace {

// One ELL row of at most 8 entries (chunks 1 or 2; the second chunk of a
// one-chunk row is padding) and the site's diagonal, uniform across a wave.
struct SiteRow {
  int4 c0, c1;
  float4 v0, v1;
  float dj;

  // field in the oracle's order: diag, then + w for each set neighbour
  __device__ __forceinline__ float field(const uint32_t *xs, int lane, int chunks) const {
    float h = dj;
    h = add_if_set(xs, c0.x, lane, v0.x, h);
    h = add_if_set(xs, c0.y, lane, v0.y, h);
    h = add_if_set(xs, c0.z, lane, v0.z, h);
    h = add_if_set(xs, c0.w, lane, v0.w, h);
    if (chunks > 1) {
      h = add_if_set(xs, c1.x, lane, v1.x, h);
      h = add_if_set(xs, c1.y, lane, v1.y, h);
      h = add_if_set(xs, c1.z, lane, v1.z, h);
      h = add_if_set(xs, c1.w, lane, v1.w, h);
    }
    return h;
  }
};

// Row of site j, read with vector loads: the (uniform) index is passed
// through an empty asm statement so the compiler treats it as per-lane and
// does not use scalar loads, whose completion counter LDS reads share -- a
// row requested one site ahead would otherwise hold up the current site's
// LDS gathers.
__device__ __forceinline__ SiteRow load_row(const int4 *ecol, const float4 *eval, const float *dg,
                                            int j, int chunks, int n) {
  int e0 = j * chunks, jv = j;
  asm volatile("" : "+v"(e0), "+v"(jv));
  SiteRow rw;
  rw.c0 = ecol[e0];
  rw.v0 = eval[e0];
  rw.c1 = make_int4(n, n, n, n);
  rw.v1 = make_float4(0.f, 0.f, 0.f, 0.f);
  if (chunks > 1) {
    rw.c1 = ecol[e0 + 1];
    rw.v1 = eval[e0 + 1];
  }
  rw.dj = dg[jv];
  return rw;
}

}  // namespace

// Prefetching variant for batches whose ELL rows have at most 8 entries
// (1 or 2 chunks): the same kernel, but the row, weights and diagonal of site
// j+1 (uniform values, vector loads: load_row) are requested before site j is
// decided, so the per-site chain no longer waits for a row load.  Same
// arithmetic, order and results as csr_sweep_kernel.
template <int kW>
__global__ __launch_bounds__(64 * kW) void csr_sweep_pf_kernel(
    const CsrInst *__restrict__ inst, const int2 *__restrict__ items,
    const CsrRun *__restrict__ run, const int4 *__restrict__ ell_col,
    const float4 *__restrict__ ell_val, const float *__restrict__ diag,
    const float *__restrict__ tscale, const float *__restrict__ tsr,
    const uint32_t *__restrict__ init_bits, uint32_t *__restrict__ best_bits,
    uint32_t *__restrict__ final_bits, double *__restrict__ e_best,
    double *__restrict__ e_final, int *__restrict__ queue, int max_words) {
  extern __shared__ uint32_t lds[];
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  uint32_t *xs = lds + (size_t)w * max_words * 64;
  const int num_items = run->num_items, S = run->num_sweeps, Rp = run->Rp;
  const uint32_t k0 = run->k0, k1 = run->k1;
  const uint32_t base = run->instance_base;
  const int init_mode = run->init_mode;
  const int tb = run->sweep_base;
  const int use_tsr = run->use_tsr;
  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(queue, 1);
    it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
    if (it >= num_items) break;
    const int2 wi = items[it];
    const CsrInst d = inst[wi.x];
    const int n = d.n, words = d.words, chunks = d.width >> 2;
    const uint32_t r = (uint32_t)(wi.y * 64 + lane);
    const uint32_t id = (base + (uint32_t)wi.x) << 8;
    const int4 *ecol = ell_col + (d.ell_off >> 2);
    const float4 *eval = ell_val + (d.ell_off >> 2);
    const float *dg = diag + d.spin_off;
    uint32_t *bb = best_bits + d.state_off + r;
    for (int wd = 0; wd < words; ++wd) {
      uint32_t v;
      if (init_mode) {
        v = init_bits[d.state_off + (size_t)wd * Rp + r];
      } else {
        const uint4 o = philox4x32_10(make_uint4((uint32_t)(wd >> 2), r, 0u, id | TAG_SWEEP_INIT),
                                      k0, k1);
        v = pick(o, (uint32_t)wd & 3u);
      }
      const int valid = n - 32 * wd;
      if (valid < 32) v = valid <= 0 ? 0u : (v & ((1u << valid) - 1u));
      xs[wd * 64 + lane] = v;
    }
    // FP64 energy in the oracle's order: sum over set j of diag_j + off_j / 2
    double e = 0.0;
    for (int j = 0; j < n; ++j) {
      double off = 0.0;
      for (int c = 0; c < chunks; ++c) {
        const int4 cc = ecol[(size_t)j * chunks + c];
        const float4 vv = eval[(size_t)j * chunks + c];
        off = plane_bit(xs, cc.x, lane) ? off + (double)vv.x : off;
        off = plane_bit(xs, cc.y, lane) ? off + (double)vv.y : off;
        off = plane_bit(xs, cc.z, lane) ? off + (double)vv.z : off;
        off = plane_bit(xs, cc.w, lane) ? off + (double)vv.w : off;
      }
      if (plane_bit(xs, j, lane)) e += (double)dg[j] + 0.5 * off;
    }
    double eb = e;
    for (int wd = 0; wd < words; ++wd) bb[(size_t)wd * Rp] = xs[wd * 64 + lane];
    for (int t = 0; t < S; ++t) {
      const float ts = use_tsr ? tsr[d.rep_off + r] : tscale[tb + t];
      SiteRow row = load_row(ecol, eval, dg, 0, chunks, n);
      // the state word of the current sites (j >> 5), kept in a register:
      // flips update it and are written through to LDS for the gathers
      uint32_t word = 0u;
      for (int j4 = 0; j4 < n; j4 += 4) {
        const uint4 o = philox4x32_10(
            make_uint4((uint32_t)(j4 >> 2), r, (uint32_t)(tb + t), id | TAG_SWEEP_THR), k0, k1);
        uint32_t *xw = &xs[(j4 >> 5) * 64 + lane];
        if ((j4 & 31) == 0) word = *xw;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int j = j4 + q;
          if (j >= n) break;
          // the next site's row, requested before this site is decided
          const SiteRow nrow = load_row(ecol, eval, dg, min(j + 1, n - 1), chunks, n);
          const float thr = neg_ln_u(pick(o, (uint32_t)q)) * ts;
          const float h = row.field(xs, lane, chunks);
          const bool xj = (word >> (j & 31)) & 1u;
          const float de = xj ? -h : h;
          if (de < thr) {
            word ^= 1u << (j & 31);
            *xw = word;
            e += (double)de;
          }
          row = nrow;
        }
      }
      if (e < eb) {
        eb = e;
        for (int wd = 0; wd < words; ++wd) bb[(size_t)wd * Rp] = xs[wd * 64 + lane];
      }
    }
    uint32_t *fb = final_bits + d.state_off + r;
    for (int wd = 0; wd < words; ++wd) fb[(size_t)wd * Rp] = xs[wd * 64 + lane];
    e_best[d.rep_off + r] = eb;
    e_final[d.rep_off + r] = e;
  }
}

// Colour-class variant.  The sites of an instance are visited class by class
// (greedy colouring: no two sites of a class are coupled), every class in
// ascending index order; within a class the updates are independent, so a
// wave updates G sites of each of its 64/G replicas per step: lane =
// replica slot * G + site slot.  The update at position pos of the visiting
// order draws word pos&3 of Philox counter (pos>>2, r, t) -- for the order
// 0..n-1 the stream of the sequential kernels -- and each site keeps its
// field order and dE; the FP64 running energy adds the dE of a sweep per
// lane and then over the G lanes of a replica (and the start energy, per
// site as the oracle computes it, sums sites s, s+G, ... in lane s, then over
// the G lanes).  Thresholds are computed for
// windows of 4G positions (one Philox call per lane, four logarithms) into
// LDS.  Flips are atomic XORs on the LDS planes, since sites of one class
// can share a word.

// One work item of the colour-class kernel: replicas [wi.y * 64/G, +64/G) of
// instance d, G sites per step; xs: this wave's LDS planes, then the
// threshold window.
template <int G>
__device__ __forceinline__ void col_item(
    const CsrInst &d, int2 wi, uint32_t *xs, int plane_words, int lane, int S, int Rp,
    uint32_t k0, uint32_t k1, uint32_t base, int init_mode, int tb, int use_tsr,
    const float *__restrict__ tsr,
    const int4 *__restrict__ ell_col, const float4 *__restrict__ ell_val,
    const float *__restrict__ diag, const int32_t *__restrict__ corder,
    const int32_t *__restrict__ cstart, const int4 *__restrict__ pell_col,
    const float4 *__restrict__ pell_val, const float *__restrict__ pdiag,
    const float *__restrict__ tscale, const uint32_t *__restrict__ init_bits,
    uint32_t *__restrict__ best_bits, uint32_t *__restrict__ final_bits,
    double *__restrict__ e_best, double *__restrict__ e_final) {
  static_assert(G == 4 || G == 8 || G == 16, "sites per step");
  constexpr int RW = 64 / G;  // replicas per wave
  const int rr = lane / G, s = lane % G;
  // thresholds of the positions [win, win + 4G) for the RW replicas
  float *thw = reinterpret_cast<float *>(xs + plane_words);
  // bit j of this lane's replica (the bit-field extract reads only the low 5
  // bits of its offset)
  auto bit = [&](int j) { return __builtin_amdgcn_ubfe(xs[(j >> 5) * RW + rr], (uint32_t)j, 1u); };
  const int n = d.n, words = d.words, chunks = d.width >> 2, ncol = d.ncol;
  const uint32_t r = (uint32_t)(wi.y * RW + rr);
  const uint32_t id = (base + (uint32_t)wi.x) << 8;
  const int4 *ecol = ell_col + (d.ell_off >> 2);
  const float4 *eval = ell_val + (d.ell_off >> 2);
  const float *dg = diag + d.spin_off;
  const int32_t *ord = corder + d.spin_off;
  const int32_t *cs = cstart + d.class_off;
  // the rows and diagonal in visiting order (position pos: site ord[pos])
  const int4 *pcol = pell_col + (d.ell_off >> 2);
  const float4 *pval = pell_val + (d.ell_off >> 2);
  const float *pdg = pdiag + d.spin_off;
  uint32_t *bb = best_bits + d.state_off + r;
  // start state, as in csr_sweep_kernel; the G lanes of a replica share
  // its words
  for (int wd = s; wd < words; wd += G) {
    uint32_t v;
    if (init_mode) {
      v = init_bits[d.state_off + (size_t)wd * Rp + r];
    } else {
      const uint4 o = philox4x32_10(make_uint4((uint32_t)(wd >> 2), r, 0u, id | TAG_SWEEP_INIT),
                                    k0, k1);
      v = pick(o, (uint32_t)wd & 3u);
    }
    const int valid = n - 32 * wd;
    if (valid < 32) v = valid <= 0 ? 0u : (v & ((1u << valid) - 1u));
    xs[wd * RW + rr] = v;
  }
  wave_sync();
  // FP64 energy: lane s of a replica sums sites s, s + G, ... (each site as
  // the oracle does: diag + half its set couplings), then the G partial
  // sums are added
  double e = 0.0;
  for (int j = s; j < n; j += G) {
    double off = 0.0;
    for (int c = 0; c < chunks; ++c) {
      const int4 cc = ecol[(size_t)j * chunks + c];
      const float4 vv = eval[(size_t)j * chunks + c];
      off = bit(cc.x) ? off + (double)vv.x : off;
      off = bit(cc.y) ? off + (double)vv.y : off;
      off = bit(cc.z) ? off + (double)vv.z : off;
      off = bit(cc.w) ? off + (double)vv.w : off;
    }
    if (bit(j)) e += (double)dg[j] + 0.5 * off;
  }
#pragma unroll
  for (int o = 1; o < G; o <<= 1) e += __shfl_xor(e, o, 64);
  double eb = e;
  for (int wd = s; wd < words; wd += G) bb[(size_t)wd * Rp] = xs[wd * RW + rr];
  for (int t = 0; t < S; ++t) {
    const float ts = use_tsr ? tsr[d.rep_off + r] : tscale[tb + t];
    double de_sum = 0.0;
    int win = -4 * G;  // no window yet
    // steps of the sweep: positions [p0, p0 + G) of class c (ending at
    // c1); the rows of the next step are requested before this one runs
    int c = 0, p0 = cs[0], c1 = cs[1];
    // the first two chunks of a row travel with the step; rows of more
    // than 8 couplings read the rest when the step runs
    struct Step {
      int j, pos;
      float dj;
      int4 c0, c1;
      float4 v0, v1;
    };
    // loads at a uniform base plus a 32-bit byte offset
    auto at16 = [](const void *base, uint32_t off) {
      return reinterpret_cast<const char *>(base) + off;
    };
    auto fetch = [&](int q0) {
      Step st;
      const int pos = min(q0 + s, n - 1);
      const uint32_t ro = (uint32_t)pos * (uint32_t)chunks * 16u;  // row's byte offset
      st.pos = pos;
      st.j = *reinterpret_cast<const int32_t *>(at16(ord, (uint32_t)pos * 4u));
      st.dj = *reinterpret_cast<const float *>(at16(pdg, (uint32_t)pos * 4u));
      st.c0 = *reinterpret_cast<const int4 *>(at16(pcol, ro));
      st.v0 = *reinterpret_cast<const float4 *>(at16(pval, ro));
      st.c1 = make_int4(n, n, n, n);
      st.v1 = make_float4(0.f, 0.f, 0.f, 0.f);
      if (chunks > 1) {
        st.c1 = *reinterpret_cast<const int4 *>(at16(pcol, ro + 16u));
        st.v1 = *reinterpret_cast<const float4 *>(at16(pval, ro + 16u));
      }
      return st;
    };
    // one step: decides the sites of `cur` after requesting the next
    // step's rows into `nxt`; false after the sweep's last step.  Two
    // buffers in turn (no copies between steps).
    auto step = [&](const Step &cur, Step &nxt) -> bool {
      int np0 = p0 + G, nc = c, nc1 = c1;
      if (np0 >= c1) {
        nc = c + 1;
        if (nc < ncol) {
          np0 = cs[nc];
          nc1 = cs[nc + 1];
        }
      }
      const bool more = nc < ncol;
      if (more) nxt = fetch(np0);
      if (p0 + G > win + 4 * G) {
        // next window: lane (rr, s) draws positions win + 4s .. +3
        win = p0 & ~3;
        const uint4 o = philox4x32_10(make_uint4((uint32_t)((win >> 2) + s), r,
                                                 (uint32_t)(tb + t), id | TAG_SWEEP_THR),
                                      k0, k1);
#pragma unroll
        for (int q = 0; q < 4; ++q) thw[(4 * s + q) * RW + rr] = neg_ln_u(pick(o, (uint32_t)q)) * ts;
        wave_sync();
      }
      const int pos = p0 + s;
      const bool active = pos < c1;
      const float thr = thw[(pos - win) * RW + rr];
      // h + w for a set neighbour bit, h for a clear one: fma(1, w, h) is the
      // rounded sum, fma(0, w, h) is h (up to the sign of a zero h, which no
      // comparison or later sum sees)
      float h = cur.dj;
      h = fmaf((float)bit(cur.c0.x), cur.v0.x, h);
      h = fmaf((float)bit(cur.c0.y), cur.v0.y, h);
      h = fmaf((float)bit(cur.c0.z), cur.v0.z, h);
      h = fmaf((float)bit(cur.c0.w), cur.v0.w, h);
      if (chunks > 1) {
        h = fmaf((float)bit(cur.c1.x), cur.v1.x, h);
        h = fmaf((float)bit(cur.c1.y), cur.v1.y, h);
        h = fmaf((float)bit(cur.c1.z), cur.v1.z, h);
        h = fmaf((float)bit(cur.c1.w), cur.v1.w, h);
      }
      for (int cc = 2; cc < chunks; ++cc) {
        const uint32_t ro = ((uint32_t)cur.pos * (uint32_t)chunks + (uint32_t)cc) * 16u;
        const int4 ci = *reinterpret_cast<const int4 *>(at16(pcol, ro));
        const float4 vi = *reinterpret_cast<const float4 *>(at16(pval, ro));
        h = fmaf((float)bit(ci.x), vi.x, h);
        h = fmaf((float)bit(ci.y), vi.y, h);
        h = fmaf((float)bit(ci.z), vi.z, h);
        h = fmaf((float)bit(ci.w), vi.w, h);
      }
      const int j = cur.j;
      const float de = bit(j) ? -h : h;
      if (active && de < thr) {
        atomicXor(&xs[(j >> 5) * RW + rr], 1u << (j & 31));
        de_sum += (double)de;
      }
      if (!more) return false;
      // a new class reads this class's flips
      if (nc != c) wave_sync();
      p0 = np0;
      c1 = nc1;
      c = nc;
      return true;
    };
    Step sa = fetch(p0), sb = sa;
    while (step(sa, sb) && step(sb, sa)) {
    }
    wave_sync();
#pragma unroll
    for (int off = 1; off < G; off <<= 1) de_sum += __shfl_xor(de_sum, off, 64);
    e += de_sum;
    if (e < eb) {
      eb = e;
      for (int wd = s; wd < words; wd += G) bb[(size_t)wd * Rp] = xs[wd * RW + rr];
    }
  }
  uint32_t *fb = final_bits + d.state_off + r;
  for (int wd = s; wd < words; wd += G) fb[(size_t)wd * Rp] = xs[wd * RW + rr];
  if (s == 0) {
    e_best[d.rep_off + r] = eb;
    e_final[d.rep_off + r] = e;
  }
  // the next item's start state overwrites these planes
  wave_sync();
}

// Colour-class sweeps over a batch whose instances each have their own sites
// per step (CsrInst::sites: 4, 8 or 16).  plane_words: LDS words per wave for
// the state planes (the largest words * 64/sites of the batch), followed by
// 4*64 threshold words.
template <int kW>
__global__ __launch_bounds__(64 * kW) void csr_sweep_col_kernel(
    const CsrInst *__restrict__ inst, const int2 *__restrict__ items,
    const CsrRun *__restrict__ run, const int4 *__restrict__ ell_col,
    const float4 *__restrict__ ell_val, const float *__restrict__ diag,
    const int32_t *__restrict__ corder, const int32_t *__restrict__ cstart,
    const int4 *__restrict__ pell_col, const float4 *__restrict__ pell_val,
    const float *__restrict__ pdiag, const float *__restrict__ tscale,
    const float *__restrict__ tsr, const uint32_t *__restrict__ init_bits,
    uint32_t *__restrict__ best_bits, uint32_t *__restrict__ final_bits,
    double *__restrict__ e_best, double *__restrict__ e_final, int *__restrict__ queue,
    int plane_words) {
  extern __shared__ uint32_t lds[];
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  uint32_t *xs = lds + (size_t)w * (plane_words + 4 * 64);
  const int num_items = run->num_items, S = run->num_sweeps, Rp = run->Rp;
  const uint32_t k0 = run->k0, k1 = run->k1;
  const uint32_t base = run->instance_base;
  const int init_mode = run->init_mode;
  const int tb = run->sweep_base;
  const int use_tsr = run->use_tsr;
  (void)num_items;
  // XCD-aware queue: work-groups are dealt to the 8 XCDs round robin, and the
  // host gave each XCD's list whole instances, so an instance's rows stay in
  // one XCD's L2.  A wave takes items from its XCD's list, then from the
  // others in turn; it leaves once every list is exhausted (each list is
  // tried once more after it ran out, so every wave exits).
  const int x0 = blockIdx.x & 7;
  uint32_t done = 0u;  // lists this wave found exhausted (lane 0)
  for (;;) {
    int it = -1;
    if (lane == 0) {
      for (int q = 0; q < 8; ++q) {
        const int x = (x0 + q) & 7;
        if (done & (1u << x)) continue;
        const int cnt = queue[16 + x];
        const int v = cnt > 0 ? atomicAdd(&queue[x], 1) : cnt;
        if (v < cnt) {
          it = queue[8 + x] + v;
          break;
        }
        done |= 1u << x;
      }
    }
    it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
    if (it < 0) break;
    const int2 wi = items[it];
    const CsrInst d = inst[wi.x];
    const int g = __builtin_amdgcn_readfirstlane(d.sites);
#define OS_COL_ITEM(GG)                                                                         \
  col_item<GG>(d, wi, xs, plane_words, lane, S, Rp, k0, k1, base, init_mode, tb, use_tsr, tsr, \
               ell_col,                                                                        \
               ell_val, diag, corder, cstart, pell_col, pell_val, pdiag, tscale, init_bits,    \
               best_bits, final_bits, e_best, e_final)
    if (g == 4)
      OS_COL_ITEM(4);
    else if (g == 8)
      OS_COL_ITEM(8);
    else
      OS_COL_ITEM(16);
#undef OS_COL_ITEM
  }
}

// FP64 energy of states given as planes: E = sum_j x_j (Q_jj + sum_{l>j} Q_jl x_l),
// the reference's energy function (upper triangle incl. diagonal), with the
// double-precision couplings of the CSR input.  Work-group = item (64
// replicas of one instance, lane = replica): its state planes are staged in
// LDS (words x 64, dynamic), then wave w sums rows
// [n w / kEnergyWaves, n (w+1) / kEnergyWaves) in ascending order, and the
// partial sums are added in wave order, so the result is the same on every
// run.  The neighbour loop takes four entries at a time, so four bit reads
// are in flight per lane.
constexpr int kEnergyWaves = 8;

__global__ __launch_bounds__(64 * kEnergyWaves) void csr_energy_kernel(
    const CsrInst *__restrict__ inst, const int2 *__restrict__ items, int num_items, int Rp,
    const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ diag64,
    const uint32_t *__restrict__ bits, double *__restrict__ out) {
  __shared__ double part[kEnergyWaves][64];
  extern __shared__ uint32_t eplanes[];  // [words][64]
  const int it = blockIdx.x;  // < num_items: the grid has one work-group per item
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int2 wi = items[it];
  const CsrInst d = inst[wi.x];
  const uint32_t r = (uint32_t)(wi.y * 64 + lane);
  {
    const uint32_t *x = bits + d.state_off + (uint32_t)wi.y * 64;
    for (int i = threadIdx.x; i < d.words * 64; i += 64 * kEnergyWaves)
      eplanes[i] = x[(size_t)(i >> 6) * Rp + (i & 63)];
  }
  __syncthreads();
  const int32_t *rp = row_ptr + d.row_off;
  const int32_t *cl = col + d.csr_off;
  const double *vl = val + d.csr_off;
  const double *dg = diag64 + d.spin_off;
  const int j0 = (int)((int64_t)d.n * w / kEnergyWaves);
  const int j1 = (int)((int64_t)d.n * (w + 1) / kEnergyWaves);
  auto bit = [&](int l) { return (eplanes[(l >> 5) * 64 + lane] >> (l & 31)) & 1u; };
  double e = 0.0;
  for (int j = j0; j < j1; ++j) {
    // s gains +0.0 for a neighbour that is not set or not above j, and e
    // gains +0.0 for an unset spin: the same sums as skipping them
    double s = dg[j];
    int k = rp[j];
    const int k1 = rp[j + 1];
    for (; k + 4 <= k1; k += 4) {
      int l[4];
      uint32_t b[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) l[q] = cl[k + q];
#pragma unroll
      for (int q = 0; q < 4; ++q) b[q] = l[q] > j ? bit(l[q]) : 0u;
#pragma unroll
      for (int q = 0; q < 4; ++q) s += b[q] ? vl[k + q] : 0.0;
    }
    for (; k < k1; ++k) {
      const int l = cl[k];
      s += (l > j && bit(l)) ? vl[k] : 0.0;
    }
    e += bit(j) ? s : 0.0;
  }
  part[w][lane] = e;
  __syncthreads();
  if (w == 0) {
    double t = part[0][lane];
#pragma unroll
    for (int q = 1; q < kEnergyWaves; ++q) t += part[q][lane];
    out[d.rep_off + r] = t;
  }
}

// Population annealing: thread = (item, lane) = replica r of the item's
// instance; only replicas below R exist for the caller, the rest keep their
// planes.
__global__ __launch_bounds__(64) void csr_pop_best_kernel(
    const CsrInst *__restrict__ inst, const int2 *__restrict__ items, int num_items, int Rp,
    int first, const double *__restrict__ eb, const uint32_t *__restrict__ best,
    double *__restrict__ pop_eb, uint32_t *__restrict__ pop_best) {
  const int it = blockIdx.x;
  if (it >= num_items) return;
  const int2 wi = items[it];
  const CsrInst d = inst[wi.x];
  const uint32_t r = (uint32_t)(wi.y * 64 + threadIdx.x);
  const double e = eb[d.rep_off + r];
  if (!first && !(e < pop_eb[d.rep_off + r])) return;
  pop_eb[d.rep_off + r] = e;
  for (int w = 0; w < d.words; ++w)
    pop_best[d.state_off + (size_t)w * Rp + r] = best[d.state_off + (size_t)w * Rp + r];
}

__global__ __launch_bounds__(64) void csr_pop_gather_kernel(
    const CsrInst *__restrict__ inst, const int2 *__restrict__ items, int num_items, int Rp,
    int R, const uint32_t *__restrict__ anc, const uint32_t *__restrict__ fin,
    uint32_t *__restrict__ init) {
  const int it = blockIdx.x;
  if (it >= num_items) return;
  const int2 wi = items[it];
  const CsrInst d = inst[wi.x];
  const int r = wi.y * 64 + (int)threadIdx.x;
  const int src = r < R ? (int)anc[(size_t)wi.x * R + r] : r;
  for (int w = 0; w < d.words; ++w)
    init[d.state_off + (size_t)w * Rp + r] = fin[d.state_off + (size_t)w * Rp + src];
}

void launch_csr_pop_best(hipStream_t st, const CsrInst *inst, const int2 *items, int num_items,
                         int Rp, bool first, const double *eb, const uint32_t *best,
                         double *pop_eb, uint32_t *pop_best) {
  csr_pop_best_kernel<<<num_items, 64, 0, st>>>(inst, items, num_items, Rp, first ? 1 : 0, eb,
                                                best, pop_eb, pop_best);
  OS_HIP_CHECK(hipGetLastError());
}

void launch_csr_pop_gather(hipStream_t st, const CsrInst *inst, const int2 *items, int num_items,
                           int Rp, int R, const uint32_t *anc, const uint32_t *fin,
                           uint32_t *init) {
  csr_pop_gather_kernel<<<num_items, 64, 0, st>>>(inst, items, num_items, Rp, R, anc, fin, init);
  OS_HIP_CHECK(hipGetLastError());
}

int csr_lds_bytes(int waves, int max_words) { return waves * max_words * 64 * 4; }

void launch_csr_queue_reset(hipStream_t st, int *queue) {
  csr_queue_reset_kernel<<<1, 64, 0, st>>>(queue);
  OS_HIP_CHECK(hipGetLastError());
}

bool csr_pf_eligible(int max_width) { return max_width <= 8; }

void launch_csr_sweeps(hipStream_t st, bool pf, int waves, int groups, int max_words,
                       const CsrInst *inst, const int2 *items, const CsrRun *run,
                       const int4 *ell_col, const float4 *ell_val, const float *diag,
                       const float *tscale, const float *tsr, const uint32_t *init_bits,
                       uint32_t *best_bits, uint32_t *final_bits, double *e_best,
                       double *e_final, int *queue) {
  const size_t lds = (size_t)csr_lds_bytes(waves, max_words);
  switch (waves) {
#define OS_CSR_CASE(W)                                                                       \
  case W:                                                                                    \
    if (pf) {                                                                                \
      if (lds > 65536)                                                                       \
        OS_HIP_CHECK(hipFuncSetAttribute((const void *)csr_sweep_pf_kernel<W>,               \
                                         hipFuncAttributeMaxDynamicSharedMemorySize,         \
                                         (int)lds));                                         \
      csr_sweep_pf_kernel<W><<<groups, 64 * W, lds, st>>>(inst, items, run, ell_col, ell_val, \
                                                          diag, tscale, tsr, init_bits,       \
                                                          best_bits,                          \
                                                          final_bits, e_best, e_final, queue, \
                                                          max_words);                         \
      break;                                                                                 \
    }                                                                                        \
    if (lds > 65536)                                                                         \
      OS_HIP_CHECK(hipFuncSetAttribute((const void *)csr_sweep_kernel<W>,                    \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
    csr_sweep_kernel<W><<<groups, 64 * W, lds, st>>>(inst, items, run, ell_col, ell_val, diag,  \
                                                     tscale, tsr, init_bits, best_bits,         \
                                                     final_bits,                                \
                                                     e_best, e_final, queue, max_words);        \
    break;
    OS_CSR_CASE(1)
    OS_CSR_CASE(2)
    OS_CSR_CASE(4)
#undef OS_CSR_CASE
    default:
      throw Error(-3, "csr sweep kernel: unsupported waves per group");
  }
  OS_HIP_CHECK(hipGetLastError());
}

int csr_max_groups_per_cu(bool pf, int waves, int max_words) {
  const size_t lds = (size_t)csr_lds_bytes(waves, max_words);
  int blocks = 0;
  hipError_t e = hipSuccess;
  if (pf) {
    switch (waves) {
      case 1:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, csr_sweep_pf_kernel<1>, 64, lds);
        break;
      case 2:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, csr_sweep_pf_kernel<2>, 128, lds);
        break;
      case 4:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, csr_sweep_pf_kernel<4>, 256, lds);
        break;
      default:
        throw Error(-3, "csr sweep kernel: unsupported waves per group");
    }
    OS_HIP_CHECK(e);
    return blocks;
  }
  switch (waves) {
    case 1:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, csr_sweep_kernel<1>, 64, lds);
      break;
    case 2:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, csr_sweep_kernel<2>, 128, lds);
      break;
    case 4:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, csr_sweep_kernel<4>, 256, lds);
      break;
    default:
      throw Error(-3, "csr sweep kernel: unsupported waves per group");
  }
  OS_HIP_CHECK(e);
  return blocks;
}

int csr_col_lds_bytes(int waves, int plane_words) { return waves * (plane_words + 4 * 64) * 4; }

namespace {

const void *col_kernel_for(int waves) {
  switch (waves) {
    case 1: return (const void *)csr_sweep_col_kernel<1>;
    case 2: return (const void *)csr_sweep_col_kernel<2>;
    case 4: return (const void *)csr_sweep_col_kernel<4>;
    default: throw Error(-3, "csr colour sweep kernel: unsupported waves per group");
  }
}

}  // namespace

int csr_col_max_groups_per_cu(int waves, int plane_words) {
  const size_t lds = (size_t)csr_col_lds_bytes(waves, plane_words);
  int blocks = 0;
  OS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, col_kernel_for(waves),
                                                           64 * waves, lds));
  return blocks;
}

void launch_csr_col_sweeps(hipStream_t st, int waves, int groups, int plane_words,
                           const CsrInst *inst, const int2 *items, const CsrRun *run,
                           const int4 *ell_col, const float4 *ell_val, const float *diag,
                           const int32_t *corder, const int32_t *cstart, const int4 *pell_col,
                           const float4 *pell_val, const float *pdiag, const float *tscale,
                           const float *tsr, const uint32_t *init_bits, uint32_t *best_bits,
                           uint32_t *final_bits, double *e_best, double *e_final, int *queue) {
  const size_t lds = (size_t)csr_col_lds_bytes(waves, plane_words);
  const void *k = col_kernel_for(waves);
  if (lds > 65536)
    OS_HIP_CHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
#define OS_COL_CASE(W)                                                                          \
  case W:                                                                                       \
    csr_sweep_col_kernel<W><<<groups, 64 * W, lds, st>>>(                                       \
        inst, items, run, ell_col, ell_val, diag, corder, cstart, pell_col, pell_val, pdiag,    \
        tscale, tsr, init_bits, best_bits, final_bits, e_best, e_final, queue, plane_words);   \
    break;
  switch (waves) {
    OS_COL_CASE(1)
    OS_COL_CASE(2)
    OS_COL_CASE(4)
    default:
      throw Error(-3, "csr colour sweep kernel: unsupported waves per group");
  }
#undef OS_COL_CASE
  OS_HIP_CHECK(hipGetLastError());
}

// Warm start of a batch from per-instance marginals p (instance i at
// p + spin_off): x_j = (u < p_j), u = (w >> 8) 2^-24, w word j&3 of Philox
// counter (j>>2, r, 0, ((instance_base + i) << 8) | TAG_WARM), the dense
// path's rule.  Work-group = instance; replicas r >= R and bit n (the ELL
// padding target) stay zero.
__global__ __launch_bounds__(256) void csr_warm_init_kernel(const CsrInst *__restrict__ inst,
                                                            int R, int Rp, uint32_t k0,
                                                            uint32_t k1, uint32_t instance_base,
                                                            const float *__restrict__ p,
                                                            uint32_t *__restrict__ planes) {
  const CsrInst d = inst[blockIdx.x];
  const uint32_t tag = ((instance_base + blockIdx.x) << 8) | TAG_WARM;
  const float *pi = p + d.spin_off;
  uint32_t *pl = planes + d.state_off;
  for (int idx = threadIdx.x; idx < d.words * Rp; idx += blockDim.x) {
    const int w = idx / Rp, r = idx % Rp;
    uint32_t v = 0;
    if (r < R) {
      for (int q = 0; q < 8; ++q) {
        const int j0 = w * 32 + 4 * q;
        if (j0 >= d.n) break;
        const uint4 o =
            philox4x32_10(make_uint4((uint32_t)(j0 >> 2), (uint32_t)r, 0u, tag), k0, k1);
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int j = j0 + jj;
          if (j < d.n && (float)(pick(o, jj) >> 8) * 0x1p-24f < pi[j]) v |= 1u << (4 * q + jj);
        }
      }
    }
    pl[(size_t)w * Rp + r] = v;
  }
}

void launch_csr_warm_init(hipStream_t st, const CsrInst *inst, int num_instances, int R, int Rp,
                          uint64_t seed, uint32_t instance_base, const float *p,
                          uint32_t *planes) {
  if (num_instances == 0) return;
  csr_warm_init_kernel<<<num_instances, 256, 0, st>>>(inst, R, Rp, (uint32_t)seed,
                                                      (uint32_t)(seed >> 32), instance_base, p,
                                                      planes);
  OS_HIP_CHECK(hipGetLastError());
}

void launch_csr_energy(hipStream_t st, const CsrInst *inst, const int2 *items, int num_items,
                       int Rp, int max_words, const int32_t *row_ptr, const int32_t *col,
                       const double *val, const double *diag64, const uint32_t *bits,
                       double *out) {
  if (num_items == 0) return;
  const size_t lds = (size_t)max_words * 64 * sizeof(uint32_t);
  csr_energy_kernel<<<num_items, 64 * kEnergyWaves, lds, st>>>(inst, items, num_items, Rp,
                                                                row_ptr, col, val, diag64, bits,
                                                                out);
  OS_HIP_CHECK(hipGetLastError());
}

}  // namespace os
