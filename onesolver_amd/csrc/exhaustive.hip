// exhaustive.hip -- the exact QUBO minimum by enumeration of all 2^n states.
//
// Restates exhaustive::solve (include/exhaustive/exhaustive.hpp:29-167):
// E(x) = sum_{i <= j} q[i][j] x_i x_j over the upper triangle of q (diagonal:
// the linear terms), bit i of the state index is variable i, energies are
// FP64 sums in the reference's order (i ascending, then j >= i ascending,
// exhaustive.hpp:113-128), and of equal energies the lowest state index wins
// (each work-item keeps its first minimum, exhaustive.hpp:129-132, and the
// host takes the first minimal work-item, :158-160).
//
// n <= 20: one thread per state computes that literal energy; a minimum over
// (energy, index) pairs follows.
//
// 20 < n <= 40: thread t owns the states whose top n-L bits equal t.  It walks
// them in Gray-code order in segments of 2^S states, starting each segment
// from its literal energy.  The low kLow bits flip in an unrolled order with
// their local fields h_j = q_jj + sum_{i != j} q_ij x_i in registers, so a
// state costs kLow FMAs; a higher bit B flips once per 2^kLow states, with
// its field summed from the LDS copy of q.
//   pass 1: every thread's minimum of its walk energies, and their minimum
//           E_w;
//   pass 2: only the threads whose minimum is within tol = 1e-9 (|E_w| + 1)
//           of E_w walk again; every state whose walk energy is within tol of
//           E_w gets its literal energy, and the (literal energy, index)
//           minimum of those is the answer.
// A walk energy differs from the literal one by the rounding of at most 2^S
// steps, far below tol, so the literal minimum is always among the
// candidates and the result equals the reference's.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "../../include/onesolver.h"
#include "os_kernels.h"

namespace os {

namespace {

constexpr int kMaxN = 40;
constexpr int kLiteralMaxN = 20;
constexpr int kLow = 5;      // bits flipped in the unrolled order
constexpr int kSegBits = 12;  // literal restart every 2^kSegBits states
constexpr int kMaxTopBits = 20;
constexpr int kThreads = 256;

__device__ __forceinline__ bool better(double e, uint64_t x, double be, uint64_t bx) {
  return e < be || (e == be && x < bx);
}

// q (n x n, row stride n, upper triangle used) into LDS as a symmetric
// matrix qs[i][j] = q[min][max] with row stride kMaxN.
__device__ void load_q(int n, const double *__restrict__ q, double *qs) {
  for (int k = threadIdx.x; k < n * n; k += blockDim.x) {
    const int i = k / n, j = k % n;
    const int a = min(i, j), b = max(i, j);
    qs[i * kMaxN + j] = q[(size_t)a * n + b];
  }
}

// The reference's per-state energy and summation order (exhaustive.hpp:113-128).
__device__ double literal_energy(int n, const double *qs, uint64_t x) {
  double e = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!((x >> i) & 1u)) continue;
    e += qs[i * kMaxN + i];
    for (int j = i + 1; j < n; ++j)
      if ((x >> j) & 1u) e += qs[i * kMaxN + j];
  }
  return e;
}

// Field of bit b in state x: q_bb + sum_{i != b} q_bi x_i.
__device__ __forceinline__ double field(int n, const double *qs, uint64_t x, int b) {
  double h = qs[b * kMaxN + b];
  for (int i = 0; i < n; ++i)
    if (i != b && ((x >> i) & 1u)) h += qs[b * kMaxN + i];
  return h;
}

// Minimum over the block of (e, x) pairs, lexicographic; thread 0 writes it.
__device__ void block_min(double e, uint64_t x, double *out_e, uint64_t *out_x) {
  __shared__ double se[kThreads];
  __shared__ uint64_t sx[kThreads];
  se[threadIdx.x] = e;
  sx[threadIdx.x] = x;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && better(se[threadIdx.x + s], sx[threadIdx.x + s], se[threadIdx.x],
                                       sx[threadIdx.x])) {
      se[threadIdx.x] = se[threadIdx.x + s];
      sx[threadIdx.x] = sx[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out_e[blockIdx.x] = se[0];
    out_x[blockIdx.x] = sx[0];
  }
}

}  // namespace

// One thread per state, n <= 30.  An energy that is NaN or not below DBL_MAX
// never wins (the reference's strict comparison against its DBL_MAX start,
// exhaustive.hpp:108-132); state 0 (energy 0) always competes.
__global__ __launch_bounds__(kThreads) void exh_literal_kernel(int n, const double *__restrict__ q,
                                                               double *__restrict__ out_e,
                                                               uint64_t *__restrict__ out_x) {
  __shared__ double qs[kMaxN * kMaxN];
  load_q(n, q, qs);
  __syncthreads();
  const uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  double e = DBL_MAX;
  uint64_t bx = 0;
  if (x < (1ull << n)) {
    const double v = literal_energy(n, qs, x);
    if (v < DBL_MAX) {
      e = v;
      bx = x;
    }
  }
  block_min(e, bx, out_e, out_x);
}

// Gray-code walks, 20 < n <= 40; L = n - P bits per thread, 2^P threads.
// PASS 1 writes every thread's minimum walk energy to tmin and per-block
// minima of them (out_x unused); PASS 2 walks the threads with tmin <= thr
// and writes per-block (literal energy, index) minima of the candidates
// e_walk <= thr.
template <int PASS>
__global__ __launch_bounds__(kThreads) void exh_gray_kernel(int n, int L, const double *__restrict__ q,
                                                            double thr, double *__restrict__ tmin,
                                                            double *__restrict__ out_e,
                                                            uint64_t *__restrict__ out_x) {
  __shared__ double qs[kMaxN * kMaxN];
  const int S = min(L, kSegBits);
  const uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  bool active = t < (1ull << (n - L));
  if (PASS == 2) {
    if (active) active = tmin[t] <= thr;
    if (!__syncthreads_or(active)) {  // no candidate in this block
      if (threadIdx.x == 0) {
        out_e[blockIdx.x] = DBL_MAX;
        out_x[blockIdx.x] = 0;
      }
      return;
    }
  }
  load_q(n, q, qs);
  __syncthreads();
  // the off-diagonal part of the low kLow x kLow block of q in registers:
  // ql[i][j] = q_ij for i < j
  double ql[kLow][kLow];
#pragma unroll
  for (int i = 0; i < kLow; ++i)
#pragma unroll
    for (int j = i + 1; j < kLow; ++j) ql[i][j] = qs[i * kMaxN + j];
  double emin = DBL_MAX;
  uint64_t xmin = 0;
  double eg[8];  // PASS 1: the energies of 8 consecutive states
  const int segs = active ? 1 << (L - S) : 0;
  const int blocks = 1 << (S - kLow);
  for (int seg = 0; seg < segs; ++seg) {
    // segment start: low S bits zero
    uint64_t xhi = (t << L) | ((uint64_t)seg << S);
    double e = literal_energy(n, qs, xhi);
    double h[kLow], s[kLow];
#pragma unroll
    for (int j = 0; j < kLow; ++j) {
      h[j] = field(n, qs, xhi, j);
      s[j] = 1.0;  // 1 - 2 x_j
    }
    for (int blk = 0; blk < blocks; ++blk) {
      // low kLow bits at the block start: bit kLow-1 of the Gray code of 2^kLow * blk
      const uint32_t low0 = (blk & 1) ? 1u << (kLow - 1) : 0u;
#pragma unroll
      for (int pos = 0; pos < (1 << kLow); ++pos) {
        if (pos > 0) {
          const int b = __builtin_ctz(pos);
          e = fma(s[b], h[b], e);
#pragma unroll
          for (int j = 0; j < kLow; ++j)
            if (j != b) h[j] = fma(s[b], j < b ? ql[j][b] : ql[b][j], h[j]);
          s[b] = -s[b];
        }
        if (PASS == 1) {
          // minimum of each 8 states as a tree, then into emin; the barrier
          // keeps the compiler from holding a whole block of energies
          // (register pressure, occupancy)
          eg[pos & 7] = e;
          if ((pos & 7) == 7) {
            const double m = fmin(fmin(fmin(eg[0], eg[1]), fmin(eg[2], eg[3])),
                                  fmin(fmin(eg[4], eg[5]), fmin(eg[6], eg[7])));
            emin = fmin(emin, m);
            __builtin_amdgcn_sched_barrier(0);
          }
        } else if (e <= thr) {
          const uint64_t x = xhi | (low0 ^ (uint32_t)(pos ^ (pos >> 1)));
          const double v = literal_energy(n, qs, x);
          if (better(v, x, emin, xmin)) {
            emin = v;
            xmin = x;
          }
        }
      }
      if (blk + 1 < blocks) {
        // the low bits end at low0 ^ 2^(kLow-1); flip bit B
        const int B = kLow + __builtin_ctz((unsigned)(blk + 1));
        const uint64_t x = xhi | (low0 ^ (1u << (kLow - 1)));
        const double sb = ((x >> B) & 1u) ? -1.0 : 1.0;
        e = fma(sb, field(n, qs, x, B), e);
#pragma unroll
        for (int j = 0; j < kLow; ++j) h[j] = fma(sb, qs[j * kMaxN + B], h[j]);
        xhi ^= 1ull << B;
      }
    }
  }
  if (PASS == 1 && t < (1ull << (n - L))) tmin[t] = emin;
  block_min(emin, xmin, out_e, out_x);
}

int exhaustive_max_n() { return kMaxN; }

// Threads and per-thread bits of the Gray walk for n (20 < n <= 40).
static void gray_shape(int n, int *P, int *L) {
  *P = std::min(n - kLow, kMaxTopBits);
  *L = n - *P;
}

void exhaustive_solve(hipStream_t st, int n, const double *q_dev, bool finite, double *tmin,
                      double *part_e, uint64_t *part_x, size_t part_cap, double *best_e,
                      uint64_t *best_x) {
  auto reduce = [&](int nb) {
    std::vector<double> e(nb);
    std::vector<uint64_t> x(nb);
    OS_HIP_CHECK(hipMemcpyAsync(e.data(), part_e, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    OS_HIP_CHECK(hipMemcpyAsync(x.data(), part_x, sizeof(uint64_t) * nb, hipMemcpyDeviceToHost, st));
    OS_HIP_CHECK(hipStreamSynchronize(st));
    double be = DBL_MAX;
    uint64_t bx = 0;
    for (int i = 0; i < nb; ++i)
      if (e[i] < be || (e[i] == be && x[i] < bx)) {
        be = e[i];
        bx = x[i];
      }
    *best_e = be;
    *best_x = bx;
  };
  if (n <= kLiteralMaxN || (!finite && n <= 30)) {
    const int nb = (int)(((1ull << n) + kThreads - 1) / kThreads);
    OS_REQUIRE((size_t)nb <= part_cap, "partials buffer too small");
    exh_literal_kernel<<<nb, kThreads, 0, st>>>(n, q_dev, part_e, part_x);
    OS_HIP_CHECK(hipGetLastError());
    reduce(nb);
    return;
  }
  if (!finite)
    throw Error(OS_ERR_UNSUPPORTED, "exhaustive search with non-finite coefficients supports n <= 30");
  OS_REQUIRE(n <= kMaxN, "exhaustive search supports n <= 40");
  int P, L;
  gray_shape(n, &P, &L);
  const int nb = (int)(((1ull << P) + kThreads - 1) / kThreads);
  OS_REQUIRE((size_t)nb <= part_cap, "partials buffer too small");
  exh_gray_kernel<1><<<nb, kThreads, 0, st>>>(n, L, q_dev, 0.0, tmin, part_e, part_x);
  OS_HIP_CHECK(hipGetLastError());
  reduce(nb);
  const double ew = *best_e;
  const double thr = ew + 1e-9 * (std::fabs(ew) + 1.0);
  exh_gray_kernel<2><<<nb, kThreads, 0, st>>>(n, L, q_dev, thr, tmin, part_e, part_x);
  OS_HIP_CHECK(hipGetLastError());
  reduce(nb);
}

size_t exhaustive_thread_slots(int n) {
  if (n <= kLiteralMaxN) return 0;
  int P, L;
  gray_shape(std::min(n, kMaxN), &P, &L);
  return (size_t)1 << P;
}

size_t exhaustive_partials(int n) {
  if (n <= 30) return (size_t)(((1ull << std::min(n, 30)) + kThreads - 1) / kThreads);
  return (size_t)1 << (kMaxTopBits - 8);
}

}  // namespace os
