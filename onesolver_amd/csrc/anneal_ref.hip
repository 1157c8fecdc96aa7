// anneal_ref.hip -- reference-semantics simulated annealing on the GPU.
//
// Restates the chain of sa::anneal (include/simulated_annealing/annealing.hpp:85-126)
// with O(N) local fields instead of the O(N^2) energy recompute of :104:
//   - one wavefront per replica; its 64 lanes share the replica's fields
//     h_j = Q_jj + sum_{l!=j} Q_jl x_l (FP64, in LDS when they fit);
//   - attempt t: Philox draw -> site s = floor(w0*N/2^32), u = w1*2^-32;
//     dE = (1-2x_s) h_s; accept iff dE < 0 || exp(-dE/beta) > u
//     (annealing.hpp:106-107 with E_cur - E_new = -dE);
//   - an accepted flip updates all fields with row s of Q (coalesced,
//     Q is symmetric), E += dE;
//   - best state copied whenever E < E_best (annealing.hpp:115-121).
#include "os_kernels.h"

namespace os {

template <bool kLdsFields>
__global__ __launch_bounds__(64) void ref_anneal_kernel(int n, const double *__restrict__ q,
                                                        const double *__restrict__ beta,
                                                        uint64_t total_attempts, int spb,
                                                        uint32_t k0, uint32_t k1,
                                                        double *__restrict__ gfields,
                                                        uint32_t *__restrict__ best_bits, int W,
                                                        double *__restrict__ best_run) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t replica = blockIdx.x;
  const int lane = threadIdx.x;
  double *h = kLdsFields ? reinterpret_cast<double *>(smem) : gfields + (size_t)replica * n;
  uint8_t *x = smem + (kLdsFields ? sizeof(double) * (size_t)n : 0);
  uint8_t *best = x + n;

  for (int j = lane; j < n; j += 64) {
    const uint4 o = philox4x32_10(make_uint4((uint32_t)(j >> 7), replica, 0u, TAG_REF_INIT), k0, k1);
    const uint8_t b = (uint8_t)((pick(o, (j >> 5) & 3) >> (j & 31)) & 1u);
    x[j] = b;
    best[j] = b;
    h[j] = q[(size_t)j * n + j];
  }
  __syncthreads();
  // h_j += sum_{l != j, x_l = 1} Q_lj (row l of the symmetric Q, coalesced)
  for (int l = 0; l < n; ++l) {
    if (x[l]) {
      const double *row = q + (size_t)l * n;
      for (int j = lane; j < n; j += 64)
        if (j != l) h[j] += row[j];
    }
  }
  __syncthreads();
  double part = 0.0;
  for (int j = lane; j < n; j += 64)
    if (x[j]) part += 0.5 * (h[j] + q[(size_t)j * n + j]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  double e_cur = part;
  double e_best = part;

  for (uint64_t t = 0; t < total_attempts; ++t) {
    const double b = beta[t / (uint64_t)spb];
    const uint4 o =
        philox4x32_10(make_uint4((uint32_t)t, replica, (uint32_t)(t >> 32), TAG_REF_STEP), k0, k1);
    const int s = (int)__umulhi(o.x, (uint32_t)n);
    const double u = (double)o.y * 0x1p-32;
    const uint8_t xs = x[s];
    const double hs = h[s];
    const double de = xs ? -hs : hs;
    const bool accept = (de < 0.0) || (exp(-de / b) > u);
    if (accept) {
      const double d = xs ? -1.0 : 1.0;
      const double *row = q + (size_t)s * n;
      for (int j = lane; j < n; j += 64)
        if (j != s) h[j] = fma(row[j], d, h[j]);
      if (lane == 0) x[s] = (uint8_t)(1 - xs);
      e_cur += de;
    }
    __syncthreads();
    if (e_cur < e_best) {
      e_best = e_cur;
      for (int j = lane; j < n; j += 64) best[j] = x[j];
    }
  }
  __syncthreads();
  uint32_t *dst = best_bits + (size_t)replica * W;
  for (int w = lane; w < W; w += 64) {
    uint32_t v = 0;
    for (int bb = 0; bb < 32; ++bb) {
      const int j = w * 32 + bb;
      if (j < n && best[j]) v |= 1u << bb;
    }
    dst[w] = v;
  }
  if (lane == 0) best_run[replica] = e_best;
}

// n <= 64: the same chain with the replica in registers, kLook attempts per
// step by speculation.  Lane j holds h_j; the state and best state are
// wave-uniform 64-bit masks; Q is staged in LDS.  The draws of an attempt
// do not depend on the state, so lane 2^m - 1 + p (m < kLook, p < 2^m)
// evaluates attempt t + m as if attempts t .. t+m-1 had been accepted where
// bit i of p is set: it replays their field updates on h at its own site, in
// the chain's order and with the chain's fma, and takes the acceptance test.
// A ballot of the 63 answers, walked from m = 0 along the decisions made,
// gives the chain's kLook decisions; they are then applied in order.  Every
// FP64 operation that reaches the state, energy or best state is the one of
// ref_anneal_kernel, so the chains are the same.
constexpr int kRefSmallN = 64;
constexpr int kLook = 6;

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__global__ __launch_bounds__(64) void ref_anneal_small_kernel(int n, const double *__restrict__ q,
                                                              const double *__restrict__ beta,
                                                              uint64_t total_attempts, int spb,
                                                              uint32_t k0, uint32_t k1,
                                                              uint32_t *__restrict__ best_bits,
                                                              int W,
                                                              double *__restrict__ best_run) {
  extern __shared__ __align__(16) double qs[];  // n x n
  const uint32_t replica = blockIdx.x;
  const int lane = threadIdx.x;
  for (int i = lane; i < n * n; i += 64) qs[i] = q[i];
  const bool live = lane < n;
  uint32_t b0 = 0;
  {
    const uint4 o = philox4x32_10(make_uint4(0u, replica, 0u, TAG_REF_INIT), k0, k1);
    if (live) b0 = (pick(o, (lane >> 5) & 3) >> (lane & 31)) & 1u;
  }
  uint64_t x = __ballot(b0 != 0u);
  uint64_t best = x;
  __syncthreads();
  const double qjj = live ? qs[lane * n + lane] : 0.0;
  double h = qjj;
  for (int l = 0; l < n; ++l)
    if ((x >> l) & 1u) {
      const double v = live ? qs[l * n + lane] : 0.0;
      if (live && lane != l) h += v;
    }
  double part = 0.0;
  if (live && ((x >> lane) & 1u)) part += 0.5 * (h + qjj);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  double e_cur = part;
  double e_best = part;

  // this lane's candidate: attempt offset m_l, assumed decisions p_l (lane
  // 63 evaluates nothing); its draw is computed one step ahead
  const int m_l = 31 - __clz(lane + 1);
  const uint32_t p_l = (uint32_t)(lane + 1 - (1 << m_l));
  const int m_c = min(m_l, kLook - 1);
  auto draw = [&](uint64_t t) {
    const uint64_t tm = t + (uint64_t)m_c;
    return philox4x32_10(make_uint4((uint32_t)tm, replica, (uint32_t)(tm >> 32), TAG_REF_STEP),
                         k0, k1);
  };
  uint4 o = draw(0);
  int bi = 0, bc = 0;  // beta index of attempt t and its remainder: t = bi * spb + bc
  for (uint64_t t = 0; t < total_attempts; t += kLook) {
    const int kk = (int)min((uint64_t)kLook, total_attempts - t);
    const int s_l = (int)__umulhi(o.x, (uint32_t)n);
    const double u_l = (double)o.y * 0x1p-32;
    o = draw(t + kLook);
    // sites and temperatures of the kLook attempts (uniform)
    int sm[kLook];
    double bm[kLook];
    int ci = bi, cc = bc, bi_next = bi, bc_next = bc;
#pragma unroll
    for (int m = 0; m < kLook; ++m) {
      sm[m] = __builtin_amdgcn_readlane(s_l, (1 << m) - 1);
      bm[m] = m < kk ? beta[ci] : 1.0;
      if (++cc == spb) {
        cc = 0;
        ++ci;
      }
      if (m == kk - 1) {
        bi_next = ci;
        bc_next = cc;
      }
    }
    // rows of Q at the kLook sites, for the updates applied below
    double qrow[kLook];
#pragma unroll
    for (int m = 0; m < kLook; ++m) qrow[m] = live ? qs[sm[m] * n + lane] : 0.0;
    // this lane's field h_{s_l} at time t (from lane s_l) and temperature
    const uint64_t hb = (uint64_t)__double_as_longlong(h);
    const uint32_t hlo = (uint32_t)__builtin_amdgcn_ds_bpermute(s_l << 2, (int)(uint32_t)hb);
    const uint32_t hhi =
        (uint32_t)__builtin_amdgcn_ds_bpermute(s_l << 2, (int)(uint32_t)(hb >> 32));
    double h_l = __longlong_as_double((long long)(((uint64_t)hhi << 32) | hlo));
    double b_l = bm[0];
#pragma unroll
    for (int m = 1; m < kLook; ++m) b_l = m_c == m ? bm[m] : b_l;
    // replay the assumed decisions of attempts t .. t + m_l - 1 (selects,
    // not branches: the fma is the chain's whether or not it is kept)
    uint64_t fm = 0;  // spins flipped on the assumed path
#pragma unroll
    for (int i = 0; i < kLook - 1; ++i) {
      const int si = sm[i];
      const double qv = qs[si * n + s_l];
      const bool on = i < m_l && ((p_l >> i) & 1u);
      const bool xi = ((x ^ fm) >> si) & 1u;
      const double hn = fma(qv, xi ? -1.0 : 1.0, h_l);
      h_l = on && si != s_l ? hn : h_l;
      fm ^= on ? 1ull << si : 0ull;
    }
    const bool xl = ((x ^ fm) >> s_l) & 1u;
    const double de = xl ? -h_l : h_l;
    const bool acc = m_l < kk && ((de < 0.0) || (exp(-de / b_l) > u_l));
    const uint64_t A = __ballot(acc);
    // the chain's decisions, applied in order
    uint32_t p = 0;
#pragma unroll
    for (int m = 0; m < kLook; ++m) {
      if (m < kk) {
        const int L = (1 << m) - 1 + (int)p;
        if ((A >> L) & 1u) {
          const int s = sm[m];
          const double d = ((x >> s) & 1u) ? -1.0 : 1.0;
          const double hn = fma(qrow[m], d, h);
          h = live && lane != s ? hn : h;
          x ^= 1ull << s;
          e_cur += readlane_f64(de, L);
          p |= 1u << m;
        }
        if (e_cur < e_best) {
          e_best = e_cur;
          best = x;
        }
      }
    }
    bi = bi_next;
    bc = bc_next;
  }
  if (lane == 0) {
    uint32_t *dst = best_bits + (size_t)replica * W;
    dst[0] = (uint32_t)best;
    if (W > 1) dst[1] = (uint32_t)(best >> 32);
    best_run[replica] = e_best;
  }
}

int ref_lds_limit() { return 4096; }

size_t ref_lds_bytes(int n) {
  return (n <= ref_lds_limit() ? sizeof(double) * (size_t)n : 0) + 2 * (size_t)n;
}

void launch_ref_anneal(hipStream_t st, int n, const double *q, const double *beta, int num_iter,
                       int spb, uint32_t num_tries, uint64_t seed, double *gfields,
                       uint32_t *best_bits, double *best_run) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const uint64_t total = (uint64_t)num_iter * (uint64_t)spb;
  const int W = (n + 31) / 32;
  const size_t lds = (ref_lds_bytes(n) + 15) & ~(size_t)15;
  if (n <= kRefSmallN) {
    ref_anneal_small_kernel<<<num_tries, 64, sizeof(double) * (size_t)n * n, st>>>(
        n, q, beta, total, spb, k0, k1, best_bits, W, best_run);
  } else if (n <= ref_lds_limit()) {
    ref_anneal_kernel<true><<<num_tries, 64, lds, st>>>(n, q, beta, total, spb, k0, k1, nullptr,
                                                         best_bits, W, best_run);
  } else {
    ref_anneal_kernel<false><<<num_tries, 64, lds, st>>>(n, q, beta, total, spb, k0, k1, gfields,
                                                          best_bits, W, best_run);
  }
  OS_HIP_CHECK(hipGetLastError());
}

}  // namespace os
