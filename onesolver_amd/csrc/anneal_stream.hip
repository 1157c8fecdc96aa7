// anneal_stream.hip -- sweep-mode annealing of large dense QUBOs (n > 1024).
//
// When R x N fp32 fields no longer fit the register file (N=8192, R=4096:
// 128 MiB), they live in HBM and every block of k = 64 spins costs one pass:
// launch g reads each 64x64 field tile, applies the previous block's flip
// deltas (F += D_{g-1} * Qhi + D_{g-1} * Qlo over the 64 rows of block
// g-1, bf16 v_mfma_f32_32x32x16_bf16, fp32 accumulation) and writes it back;
// the work-groups holding block b = g mod NB then decide its 64 spins for
// their 64 replicas (lane = replica, in order, in-block corrections with the
// fp32 64x64 diagonal sub-block of Q) and publish D_g as +/- bit masks.  So
// each launch moves the field array through HBM once (read + write) and the
// decisions ride along in a few work-groups.  Thresholds are the sweep-mode
// ones of the register-resident kernel: flip iff dE < neg_ln_u(w) *
// tscale[t], w = Philox word of counter (j>>2, replica, sweep, tag).
#include "os_kernels.h"

namespace os {

typedef __bf16 sbf16x8 __attribute__((ext_vector_type(8)));
typedef float sf32x16 __attribute__((ext_vector_type(16)));

namespace {

constexpr int kSB = 64;        // spins per block
constexpr int kST = kSB + 4;   // row stride of the transposed field tile (floats)

__device__ __forceinline__ sf32x16 smfma(uint4 a, uint4 b, sf32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(sbf16x8, a),
                                                 __builtin_bit_cast(sbf16x8, b), c, 0, 0, 0);
}

// 8 spins (bits of `plus` / `minus` at offset sh) as 8 bf16: +1, -1 or 0
__device__ __forceinline__ uint4 spins8(uint32_t plus, uint32_t minus, int sh) {
  uint32_t v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t p0 = (plus >> (sh + 2 * q)) & 1u, p1 = (plus >> (sh + 2 * q + 1)) & 1u;
    const uint32_t m0 = (minus >> (sh + 2 * q)) & 1u, m1 = (minus >> (sh + 2 * q + 1)) & 1u;
    v[q] = (p0 ? 0x3F80u : 0u) | (m0 ? 0xBF80u : 0u) | (p1 ? 0x3F800000u : 0u) |
           (m1 ? 0xBF800000u : 0u);
  }
  return make_uint4(v[0], v[1], v[2], v[3]);
}

}  // namespace

// Grid (np/64 column blocks, Rp/64 replica groups, I instances), 256 threads.
// bprev >= 0: apply D (dprev) of block bprev to the tile; b >= 0: the tiles of
// column block b decide block b of sweep `sweep` into dcur.
__global__ __launch_bounds__(256) void stream_block_kernel(
    int n, int np, int Rp, int b, int bprev, int sweep, float ts, uint32_t k0, uint32_t k1,
    uint32_t inst_base, int last_of_sweep, const uint4 *__restrict__ qpack,
    const float *__restrict__ qdb, float *__restrict__ F, uint32_t *__restrict__ xbits,
    const uint4 *__restrict__ dprev, uint4 *__restrict__ dcur, double *__restrict__ energy,
    double *__restrict__ ebest, uint32_t *__restrict__ bestbits) {
  __shared__ __align__(16) float fl[kSB][kST];   // decided block: column-major fields
  __shared__ __align__(16) float qd[kSB][kSB];   // its diagonal sub-block
  __shared__ float thr[kSB][kSB + 1];            // thresholds [replica][spin]
  const int cx = blockIdx.x, ry = blockIdx.y, s = blockIdx.z;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int lc = lane & 31, kh = lane >> 5;
  const int mt = w >> 1, nt = w & 1;
  const int NB = np / kSB, npw = np / 32;
  const size_t row0 = (size_t)s * Rp + (size_t)ry * 64;  // first state row of the group
  const int col = cx * kSB + 32 * nt + lc;
  const bool decide = b >= 0 && cx == b;
  float *Fp = F + (row0 + 32 * mt + 4 * kh) * np + col;
  sf32x16 C;
#pragma unroll
  for (int i = 0; i < 16; ++i) C[i] = Fp[(size_t)(8 * (i >> 2) + (i & 3)) * np];
  if (bprev >= 0) {
    const uint4 dd = dprev[row0 + 32 * mt + lc];  // (plus lo, plus hi, minus lo, minus hi)
    const uint4 *qb = qpack + (((size_t)s * NB + bprev) * np + col) * 16;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const uint32_t pw = ks < 2 ? dd.x : dd.y, mw = ks < 2 ? dd.z : dd.w;
      const uint4 a = spins8(pw, mw, 16 * (ks & 1) + 8 * kh);
#pragma unroll
      for (int p = 0; p < 2; ++p) C = smfma(a, qb[p * 8 + 2 * ks + kh], C);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) Fp[(size_t)(8 * (i >> 2) + (i & 3)) * np] = C[i];
  }
  if (!decide) return;
  // ---- decisions of block b for the 64 replicas of this group ----
  const uint32_t tag = ((inst_base + (uint32_t)s) << 8) | TAG_SWEEP_THR;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    *reinterpret_cast<float4 *>(&fl[32 * nt + lc][32 * mt + 8 * q + 4 * kh]) =
        make_float4(C[4 * q], C[4 * q + 1], C[4 * q + 2], C[4 * q + 3]);
  const float *qsrc = qdb + ((size_t)s * NB + b) * kSB * kSB;
  for (int k = tid; k < kSB * kSB; k += 256) qd[k >> 6][k & 63] = qsrc[k];
  // thresholds: 64 replicas x 16 Philox calls (4 spins each), 4 per thread
  for (int it = tid; it < 64 * 16; it += 256) {
    const int r = it >> 4, q4 = it & 15;
    const int j0 = b * kSB + 4 * q4;
    const uint4 o = philox4x32_10(make_uint4((uint32_t)(j0 >> 2), (uint32_t)(ry * 64 + r),
                                             (uint32_t)sweep, tag), k0, k1);
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const float x = neg_ln_u(pick(o, jj)) * ts;
      thr[r][4 * q4 + jj] = j0 + jj < n ? x : -INFINITY;
    }
  }
  __syncthreads();
  if (w != 0) return;
  const size_t rr = row0 + lane;
  uint32_t *xw = xbits + rr * npw + 2 * b;
  const uint32_t w0 = xw[0], w1 = xw[1];
  uint32_t x0 = w0, x1 = w1;
  float c[kSB];
#pragma unroll
  for (int j = 0; j < kSB; ++j) c[j] = 0.0f;
  float desum = 0.0f;
#pragma unroll
  for (int j = 0; j < kSB; ++j) {
    const float h = fl[j][lane] + c[j];
    uint32_t &xwd = j < 32 ? x0 : x1;
    const int bit = j & 31;
    const bool xj = (xwd >> bit) & 1u;
    const float de = xj ? -h : h;
    const bool acc = de < thr[lane][j];
    const float d = acc ? (xj ? -1.0f : 1.0f) : 0.0f;
    xwd ^= acc ? (1u << bit) : 0u;
    desum += acc ? de : 0.0f;
    const float4 *qrow = reinterpret_cast<const float4 *>(qd[j]);
#pragma unroll
    for (int q = (j + 1) / 4; q < kSB / 4; ++q) {
      const float4 v = qrow[q];
      if (4 * q > j) c[4 * q] = fmaf(v.x, d, c[4 * q]);
      if (4 * q + 1 > j) c[4 * q + 1] = fmaf(v.y, d, c[4 * q + 1]);
      if (4 * q + 2 > j) c[4 * q + 2] = fmaf(v.z, d, c[4 * q + 2]);
      if (4 * q + 3 > j) c[4 * q + 3] = fmaf(v.w, d, c[4 * q + 3]);
    }
  }
  xw[0] = x0;
  xw[1] = x1;
  const uint32_t f0 = x0 ^ w0, f1 = x1 ^ w1;
  dcur[rr] = make_uint4(f0 & ~w0, f1 & ~w1, f0 & w0, f1 & w1);
  const double e_now = energy[rr] + (double)desum;
  energy[rr] = e_now;
  if (last_of_sweep && e_now < ebest[rr]) {
    ebest[rr] = e_now;
    const uint32_t *src = xbits + rr * npw;
    uint32_t *dst = bestbits + rr * npw;
    for (int q = 0; q < npw; ++q) dst[q] = src[q];
  }
}

// qpack[s][b][j][0..63] = hi(Q[64b+k][j]), [64..127] = lo(...), diagonal zeroed;
// qdb[s][b][k][c] = fp32 Q[64b+k][64b+c] (diagonal sub-blocks, diagonal zeroed)
__global__ __launch_bounds__(256) void pack_q64_kernel(int n, int np, const double *__restrict__ q,
                                                       uint16_t *__restrict__ qpack,
                                                       float *__restrict__ qdb) {
  const int s = blockIdx.y;
  const int NB = np / kSB;
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;  // (b, j, k)
  if (t >= (size_t)NB * np * kSB) return;
  const int k = (int)(t & 63);
  const int j = (int)((t >> 6) % np);
  const int b = (int)((t >> 6) / np);
  const int i = kSB * b + k;
  const double *qs = q + (size_t)s * n * n;
  float qf = 0.0f;
  if (i < n && j < n && i != j) qf = (float)qs[(size_t)i * n + j];
  uint16_t hi = 0, lo = 0;
  if (qf != 0.0f) {
    uint32_t u = __float_as_uint(qf);
    u += 0x7fffu + ((u >> 16) & 1u);
    hi = (uint16_t)(u >> 16);
    uint32_t r = __float_as_uint(qf - __uint_as_float((uint32_t)hi << 16));
    r += 0x7fffu + ((r >> 16) & 1u);
    lo = (uint16_t)(r >> 16);
  }
  uint16_t *dst = qpack + (size_t)s * NB * np * 128 + ((size_t)b * np + j) * 128;
  dst[k] = hi;
  dst[64 + k] = lo;
  if (j / kSB == b) qdb[((size_t)s * NB + b) * kSB * kSB + k * kSB + (j % kSB)] = qf;
}

int stream_block() { return kSB; }

void launch_pack_q64(hipStream_t st, int I, int n, int np, const double *q, uint16_t *qpack,
                     float *qdb) {
  const size_t work = (size_t)(np / kSB) * np * kSB;
  pack_q64_kernel<<<dim3((unsigned)((work + 255) / 256), I), 256, 0, st>>>(n, np, q, qpack, qdb);
  OS_HIP_CHECK(hipGetLastError());
}

void launch_stream_block(hipStream_t st, int I, int n, int np, int Rp, int b, int bprev,
                         int sweep, float ts, uint64_t seed, uint32_t inst_base,
                         bool last_of_sweep, const uint4 *qpack, const float *qdb, float *F,
                         uint32_t *xbits, const uint4 *dprev, uint4 *dcur, double *energy,
                         double *ebest, uint32_t *bestbits) {
  const dim3 grid(np / kSB, Rp / 64, I);
  stream_block_kernel<<<grid, 256, 0, st>>>(n, np, Rp, b, bprev, sweep, ts, (uint32_t)seed,
                                            (uint32_t)(seed >> 32), inst_base,
                                            last_of_sweep ? 1 : 0, qpack, qdb, F, xbits, dprev,
                                            dcur, energy, ebest, bestbits);
  OS_HIP_CHECK(hipGetLastError());
}

}  // namespace os
