// anneal_dense.hip -- sweep-mode simulated annealing of dense QUBOs with
// delayed field updates on MFMA (gfx950).
//
// Replicas run Metropolis sweeps in a fixed site order j = 0..N-1.  Local
// fields h_rj = Q_jj + sum_{l!=j} Q_jl x_rl live in MFMA accumulators (fp32).
// Spins are processed in blocks of k = 16:
//   1. the wave that owns block b's columns brings that 16-column tile up to
//      date, transposes it through LDS (lane = replica) and decides the 16
//      spins in order; flips inside the block correct the later spins' fields
//      with the fp32 16x16 diagonal sub-block of Q;
//   2. the flip deltas D (replicas x 16, values -1/0/+1, exact in bf16) are
//      written to LDS and every wave applies F += [D D] * [Qhi; Qlo] to its
//      columns with v_mfma_f32_16x16x32_bf16 (Q = Qhi + Qlo, both bf16; the
//      K=32 operand pairs D with the hi and the lo half).
// Acceptance: flip iff dE < neg_ln_u(w) * tscale[t] with w the Philox word of
// counter (j>>2, replica, sweep, tag) -- i.e. u < exp(-dE / tscale^-1).
// Helper waves compute the next block's thresholds while the owner decides.
//
// Work-group: 8 waves x 64 lanes, 64 replicas; wave w owns columns
// [16*TPW*w, 16*TPW*(w+1)).  Fields per wave: 4 replica tiles x TPW column
// tiles x 4 fp32 = 16*TPW VGPRs per lane.
#include "os_kernels.h"

namespace os {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kWaves = 8;
constexpr int kRT = 64;  // replicas per work-group

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// thresholds of spins [16*bn + 4q, +4) of replica r for sweep (sweep0+sw)
__device__ __forceinline__ void block_thresholds(int item, int bn, int sw, int n, uint32_t rbase,
                                                 int sweep0, uint32_t k0, uint32_t k1,
                                                 uint32_t tag, float ts, float (*thr)[17]) {
  const int r = item >> 2, q = item & 3;
  const int j0 = bn * 16 + 4 * q;
  const uint4 o = philox4x32_10(make_uint4((uint32_t)(j0 >> 2), rbase + (uint32_t)r,
                                           (uint32_t)(sweep0 + sw), tag),
                                k0, k1);
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const int j = j0 + jj;
    thr[r][4 * q + jj] = j < n ? neg_ln_u(pick(o, jj)) * ts : -INFINITY;
  }
}

}  // namespace

template <int TPW>
__global__ __launch_bounds__(512, 1) void dense_sweep_kernel(
    int n, const uint4 *__restrict__ qpack, const float *__restrict__ qdb,
    const float *__restrict__ tscale, int sweep0, int nsweeps, uint32_t k0, uint32_t k1,
    uint32_t tag, float *__restrict__ fields, uint32_t *__restrict__ xbits,
    uint32_t *__restrict__ bestbits, double *__restrict__ energy, double *__restrict__ ebest) {
  constexpr int NP = kWaves * 16 * TPW;
  constexpr int NB = NP / 16;
  constexpr int NPW = NP / 32;
  __shared__ uint4 dbuf[2][kRT][2];
  __shared__ float fblk[kRT][17];
  __shared__ float thr[2][kRT][17];
  __shared__ uint32_t xs[kRT][NPW + 1];
  __shared__ double es[kRT];
  __shared__ double ebs[kRT];

  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int lane = tid & 63;
  const int lrow = lane & 15;
  const int kg = lane >> 4;
  const uint32_t rbase = blockIdx.x * kRT;
  const int col0 = w * 16 * TPW;

  for (int i = tid; i < kRT * NPW; i += 512)
    xs[i / NPW][i % NPW] = xbits[(size_t)(rbase + i / NPW) * NPW + i % NPW];
  if (tid < kRT) {
    es[tid] = energy[rbase + tid];
    ebs[tid] = ebest[rbase + tid];
  }
  f32x4 C[4][TPW];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        C[m][tt][i] =
            fields[(size_t)(rbase + 16 * m + 4 * kg + i) * NP + col0 + 16 * tt + lrow];
  if (tid < 256 && nsweeps > 0)
    block_thresholds(tid, 0, 0, n, rbase, sweep0, k0, k1, tag, tscale[0], thr[0]);
  __syncthreads();

  const int G = nsweeps * NB;
  for (int g = 0; g < G; ++g) {
    const int b = g % NB;
    const int owner = b / TPW;
    const int tb = b % TPW;
    const int bp = (b + NB - 1) % NB;
    bf16x8 A[4];
    if (g > 0) {
#pragma unroll
      for (int m = 0; m < 4; ++m) A[m] = as_bf16x8(dbuf[(g - 1) & 1][16 * m + lrow][kg & 1]);
    }
    const uint4 *qb = qpack + ((size_t)bp * NP + col0 + lrow) * 4 + kg;
    if (w == owner) {
      // bring tile tb up to date and transpose it (static register indices:
      // tb is wave-uniform, so this is a uniform branch per tile)
#pragma unroll
      for (int tt = 0; tt < TPW; ++tt) {
        if (tt != tb) continue;
        if (g > 0) {
          const bf16x8 B = as_bf16x8(qb[(size_t)16 * tt * 4]);
#pragma unroll
          for (int m = 0; m < 4; ++m)
            C[m][tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, C[m][tt], 0, 0, 0);
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int i = 0; i < 4; ++i) fblk[16 * m + 4 * kg + i][lrow] = C[m][tt][i];
      }
      wave_sync();
      // ---- sequential decisions of block b, lane = replica ----
      const float(*tb_thr)[17] = thr[g & 1];
      float f[16], c[16];
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) {
        f[jj] = fblk[lane][jj];
        c[jj] = 0.0f;
      }
      const uint32_t word = xs[lane][b >> 1];
      const int sh = (b & 1) * 16;
      uint32_t xb = (word >> sh) & 0xFFFFu;
      uint32_t dpos = 0, dneg = 0;
      float desum = 0.0f;
      const float *qd = qdb + (size_t)b * 256;
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) {
        const float h = f[jj] + c[jj];
        const bool xj = (xb >> jj) & 1u;
        const float de = xj ? -h : h;
        const bool acc = de < tb_thr[lane][jj];
        const float d = acc ? (xj ? -1.0f : 1.0f) : 0.0f;
        xb ^= acc ? (1u << jj) : 0u;
        desum += acc ? de : 0.0f;
        dpos |= (acc && !xj) ? (1u << jj) : 0u;
        dneg |= (acc && xj) ? (1u << jj) : 0u;
#pragma unroll
        for (int mm = jj + 1; mm < 16; ++mm) c[mm] = fmaf(qd[mm * 16 + jj], d, c[mm]);
      }
      xs[lane][b >> 1] = (word & ~(0xFFFFu << sh)) | (xb << sh);
      uint32_t pk[8];
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        const uint32_t lo = ((dpos >> (2 * p)) & 1u) ? 0x3F80u : (((dneg >> (2 * p)) & 1u) ? 0xBF80u : 0u);
        const uint32_t hi =
            ((dpos >> (2 * p + 1)) & 1u) ? 0x3F80u : (((dneg >> (2 * p + 1)) & 1u) ? 0xBF80u : 0u);
        pk[p] = lo | (hi << 16);
      }
      dbuf[g & 1][lane][0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
      dbuf[g & 1][lane][1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
      const double e_now = es[lane] + (double)desum;
      es[lane] = e_now;
      if (b == NB - 1 && e_now < ebs[lane]) {  // end of a sweep: best-state tracking
        ebs[lane] = e_now;
        uint32_t *dst = bestbits + (size_t)(rbase + lane) * NPW;
        for (int q = 0; q < NPW; ++q) dst[q] = xs[lane][q];
      }
      if (g > 0) {
#pragma unroll
        for (int tt = 0; tt < TPW; ++tt) {
          if (tt == tb) continue;
          const bf16x8 B = as_bf16x8(qb[(size_t)16 * tt * 4]);
#pragma unroll
          for (int m = 0; m < 4; ++m)
            C[m][tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, C[m][tt], 0, 0, 0);
        }
      }
    } else {
      if (g > 0) {
#pragma unroll
        for (int tt = 0; tt < TPW; ++tt) {
          const bf16x8 B = as_bf16x8(qb[(size_t)16 * tt * 4]);
#pragma unroll
          for (int m = 0; m < 4; ++m)
            C[m][tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, C[m][tt], 0, 0, 0);
        }
      }
      const int hw = (w - owner - 1 + kWaves) % kWaves;  // 0..6
      if (hw < 4 && g + 1 < G) {
        const int g1 = g + 1;
        block_thresholds(hw * 64 + lane, g1 % NB, g1 / NB, n, rbase, sweep0, k0, k1, tag,
                         tscale[g1 / NB], thr[g1 & 1]);
      }
    }
    __syncthreads();
  }
  if (G > 0) {
    const int bp = NB - 1;
    bf16x8 A[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) A[m] = as_bf16x8(dbuf[(G - 1) & 1][16 * m + lrow][kg & 1]);
    const uint4 *qb = qpack + ((size_t)bp * NP + col0 + lrow) * 4 + kg;
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) {
      const bf16x8 B = as_bf16x8(qb[(size_t)16 * tt * 4]);
#pragma unroll
      for (int m = 0; m < 4; ++m)
        C[m][tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, C[m][tt], 0, 0, 0);
    }
  }
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        fields[(size_t)(rbase + 16 * m + 4 * kg + i) * NP + col0 + 16 * tt + lrow] = C[m][tt][i];
  __syncthreads();
  for (int i = tid; i < kRT * NPW; i += 512)
    xbits[(size_t)(rbase + i / NPW) * NPW + i % NPW] = xs[i / NPW][i % NPW];
  if (tid < kRT) {
    energy[rbase + tid] = es[tid];
    ebest[rbase + tid] = ebs[tid];
  }
}

// Random initial bits: spin j of replica r from counter (j>>7, r, 0, tag_init).
__global__ __launch_bounds__(256) void sweep_init_kernel(int n, int npw, int R, uint32_t k0,
                                                         uint32_t k1, uint32_t tag,
                                                         uint32_t *__restrict__ xbits) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)R * npw) return;
  const uint32_t r = (uint32_t)(idx / npw);
  const int wd = (int)(idx % npw);
  const int j0 = wd * 32;
  const uint4 o = philox4x32_10(make_uint4((uint32_t)(j0 >> 7), r, 0u, tag), k0, k1);
  uint32_t v = pick(o, (j0 >> 5) & 3);
  const int valid = n - j0;
  if (valid < 32) v = valid <= 0 ? 0u : (v & ((1u << valid) - 1u));
  xbits[idx] = v;
}

// Warm start: x_j = (u < p_j), u = (w >> 8) * 2^-24 from counter (j>>2, r, 0, tag_warm).
__global__ __launch_bounds__(256) void warm_init_kernel(int n, int npw, int R, uint32_t k0,
                                                        uint32_t k1, uint32_t tag,
                                                        const float *__restrict__ p,
                                                        uint32_t *__restrict__ xbits) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)R * npw) return;
  const uint32_t r = (uint32_t)(idx / npw);
  const int wd = (int)(idx % npw);
  uint32_t v = 0;
  for (int q = 0; q < 8; ++q) {
    const int j0 = wd * 32 + 4 * q;
    if (j0 >= n) break;
    const uint4 o = philox4x32_10(make_uint4((uint32_t)(j0 >> 2), r, 0u, tag), k0, k1);
    for (int jj = 0; jj < 4; ++jj) {
      const int j = j0 + jj;
      if (j < n && (float)(pick(o, jj) >> 8) * 0x1p-24f < p[j]) v |= 1u << (4 * q + jj);
    }
  }
  xbits[idx] = v;
}

int dense_max_n() { return kWaves * 16 * 8; }

int dense_padded_n(int n) {
  const int per = kWaves * 16;
  return ((n + per - 1) / per) * per;
}

void launch_dense_sweeps(hipStream_t st, int n, int np, const uint4 *qpack, const float *qdb,
                         const float *tscale, int sweep0, int nsweeps, uint64_t seed,
                         uint32_t instance, int R_pad, float *fields, uint32_t *xbits,
                         uint32_t *bestbits, double *energy, double *ebest) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const uint32_t tag = (instance << 8) | TAG_SWEEP_THR;
  const dim3 grid(R_pad / kRT), block(512);
  switch (np / (kWaves * 16)) {
#define OS_DENSE_CASE(T)                                                                    \
  case T:                                                                                   \
    dense_sweep_kernel<T><<<grid, block, 0, st>>>(n, qpack, qdb, tscale, sweep0, nsweeps, k0, \
                                                  k1, tag, fields, xbits, bestbits, energy, \
                                                  ebest);                                   \
    break;
    OS_DENSE_CASE(1)
    OS_DENSE_CASE(2)
    OS_DENSE_CASE(3)
    OS_DENSE_CASE(4)
    OS_DENSE_CASE(5)
    OS_DENSE_CASE(6)
    OS_DENSE_CASE(7)
    OS_DENSE_CASE(8)
#undef OS_DENSE_CASE
    default:
      throw Error(-3, "dense sweep kernel: unsupported padded size");
  }
  OS_HIP_CHECK(hipGetLastError());
}

void launch_sweep_init(hipStream_t st, int n, int npw, int R, uint64_t seed, uint32_t instance,
                       uint32_t *xbits) {
  const size_t total = (size_t)R * npw;
  sweep_init_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(
      n, npw, R, (uint32_t)seed, (uint32_t)(seed >> 32), (instance << 8) | TAG_SWEEP_INIT, xbits);
  OS_HIP_CHECK(hipGetLastError());
}

void launch_warm_init(hipStream_t st, int n, int npw, int R, uint64_t seed, uint32_t instance,
                      const float *p, uint32_t *xbits) {
  const size_t total = (size_t)R * npw;
  warm_init_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(
      n, npw, R, (uint32_t)seed, (uint32_t)(seed >> 32), (instance << 8) | TAG_WARM, p, xbits);
  OS_HIP_CHECK(hipGetLastError());
}

}  // namespace os
