// energy.hip -- FP64 QUBO energies and field refresh on the GPU.
//
// For states X (R x N, bit-packed rows of W = ceil(N/32) words) and the
// symmetric FP64 matrix Q, Y = X Q is accumulated in FP64 over LDS tiles
// (128 states x 128 columns x 16 k; 8 x 8 per thread, so every FMA pair
// costs one LDS read), and
//   E_r   = 1/2 * sum_j x_rj (Y_rj + Q_jj)           (= sum_{i<=j} Q_ij x_ri x_rj,
//                                                      annealing.hpp:31-40)
//   h_rj  = Q_jj + Y_rj - Q_jj x_rj                  (local field of spin j)
// Energies are reduced in a fixed order (16-lane shuffle tree, then column
// blocks in ascending order), so results are bit-reproducible run to run.
#include "os_kernels.h"

namespace os {

namespace {
constexpr int kTR = 128;  // states per tile
constexpr int kTN = 128;  // columns per tile
constexpr int kTK = 16;   // k per LDS stage
constexpr int kPT = 8;    // rows and columns per thread (16 x 16 threads)
}  // namespace

__global__ __launch_bounds__(256) void pack_bytes_kernel(const uint8_t *__restrict__ bytes, int n,
                                                          int count, int W,
                                                          uint32_t *__restrict__ bits) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)count * W;
  if (idx >= total) return;
  const size_t r = idx / W;
  const int w = (int)(idx % W);
  uint32_t v = 0;
  for (int b = 0; b < 32; ++b) {
    const int j = w * 32 + b;
    if (j < n && bytes[r * n + j]) v |= 1u << b;
  }
  bits[idx] = v;
}

__global__ __launch_bounds__(256) void unpack_bits_kernel(const uint32_t *__restrict__ bits, int n,
                                                           int count, int W,
                                                           uint8_t *__restrict__ bytes) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)count * n;
  if (idx >= total) return;
  const size_t r = idx / n;
  const int j = (int)(idx % n);
  bytes[idx] = (uint8_t)((bits[r * W + (j >> 5)] >> (j & 31)) & 1u);
}

// kFields: also write h_rj as fp32 into fields[r*ldf + j] and raise
// drift_max (fp32 bits, non-negative) to max |h_rj - old fields value|.
template <bool kFields>
__global__ __launch_bounds__(256) void xq_tile_kernel(int n, int W, const double *__restrict__ q,
                                                       int R, const uint32_t *__restrict__ bits,
                                                       double *__restrict__ partial, int ncb,
                                                       float *__restrict__ fields, int ldf,
                                                       unsigned int *__restrict__ drift_max) {
  __shared__ double qs[kTK][kTN];
  __shared__ double xs[kTK][kTR];
  __shared__ double qdiag[kTN];
  const int tid = threadIdx.x;
  const int tx = tid & 15;
  const int ty = tid >> 4;
  const int r0 = blockIdx.y * kTR;
  const int c0 = blockIdx.x * kTN;
  // instance blockIdx.z: its matrix and its R rows of states / outputs
  q += (size_t)blockIdx.z * n * n;
  bits += (size_t)blockIdx.z * R * W;
  partial += (size_t)blockIdx.z * R * ncb;
  if (kFields) fields += (size_t)blockIdx.z * R * ldf;
  if (tid < kTN) {
    const int c = c0 + tid;
    qdiag[tid] = c < n ? q[(size_t)c * n + c] : 0.0;
  }
  // acc[i][j]: state r0 + ty + 16i, column c0 + tx + 16j
  double acc[kPT][kPT];
#pragma unroll
  for (int i = 0; i < kPT; ++i)
#pragma unroll
    for (int j = 0; j < kPT; ++j) acc[i][j] = 0.0;

  for (int k0 = 0; k0 < n; k0 += kTK) {
    for (int e = tid; e < kTK * kTN; e += 256) {
      const int kk = e / kTN, cc = e % kTN;
      const int gk = k0 + kk, gc = c0 + cc;
      qs[kk][cc] = (gk < n && gc < n) ? q[(size_t)gk * n + gc] : 0.0;
    }
    {
      // the 16 bits k0 .. k0+15 of a state are in one word (k0 is a multiple
      // of 16): two threads per state, one load and 8 bits each
      static_assert(kTK == 16 && 2 * kTR == 256, "x staging: two threads per state row");
      const int rr = tid >> 1, k8 = 8 * (tid & 1);
      const int gr = r0 + rr;
      const uint32_t wd = gr < R ? bits[(size_t)gr * W + (k0 >> 5)] >> (k0 & 31) : 0u;
#pragma unroll
      for (int kk = 0; kk < 8; ++kk)
        xs[k8 + kk][rr] = k0 + k8 + kk < n ? (double)((wd >> (k8 + kk)) & 1u) : 0.0;
    }
    __syncthreads();
#pragma unroll 2
    for (int kk = 0; kk < kTK; ++kk) {
      double a[kPT], b[kPT];
#pragma unroll
      for (int i = 0; i < kPT; ++i) a[i] = xs[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < kPT; ++j) b[j] = qs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < kPT; ++i)
#pragma unroll
        for (int j = 0; j < kPT; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }

  float dmax = 0.0f;
#pragma unroll
  for (int i = 0; i < kPT; ++i) {
    const int r = r0 + ty + 16 * i;
    double s = 0.0;
    if (r < R) {
#pragma unroll
      for (int j = 0; j < kPT; ++j) {
        const int c = c0 + tx + 16 * j;
        if (c < n) {
          const uint32_t xb = (bits[(size_t)r * W + (c >> 5)] >> (c & 31)) & 1u;
          const double qcc = qdiag[tx + 16 * j];
          if (xb) s += acc[i][j] + qcc;
          if (kFields) {
            const double h = qcc + acc[i][j] - (xb ? qcc : 0.0);
            float *fp = fields + (size_t)r * ldf + c;
            const float old = *fp;
            dmax = fmaxf(dmax, (float)fabs(h - (double)old));
            *fp = (float)h;
          }
        }
      }
    }
    // reduce s over the 16 lanes sharing this row (fixed shuffle tree)
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) s += __shfl_xor(s, off, 16);
    if (tx == 0 && r < R) partial[(size_t)r * ncb + blockIdx.x] = 0.5 * s;
  }
  if (kFields) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) dmax = fmaxf(dmax, __shfl_xor(dmax, off, 64));
    if ((tid & 63) == 0) atomicMax(drift_max, __float_as_uint(dmax));
  }
}

__global__ __launch_bounds__(256) void reduce_partials_kernel(const double *__restrict__ partial,
                                                               int R, int ncb,
                                                               double *__restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double s = 0.0;
  for (int b = 0; b < ncb; ++b) s += partial[(size_t)r * ncb + b];
  out[r] = s;
}

void launch_pack_bytes(hipStream_t st, const uint8_t *bytes, int n, int count, int W,
                       uint32_t *bits) {
  const size_t total = (size_t)count * W;
  if (total == 0) return;
  pack_bytes_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(bytes, n, count, W, bits);
  OS_HIP_CHECK(hipGetLastError());
}

void launch_unpack_bits(hipStream_t st, const uint32_t *bits, int n, int count, int W,
                        uint8_t *bytes) {
  const size_t total = (size_t)count * n;
  if (total == 0) return;
  unpack_bits_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(bits, n, count, W, bytes);
  OS_HIP_CHECK(hipGetLastError());
}

size_t energy_scratch_bytes(int n, int count) {
  const int ncb = (n + kTN - 1) / kTN;
  return sizeof(double) * (size_t)count * ncb;
}

// n <= 64, energies only: one thread per state, in the reference's order
// (sa::energy, annealing.hpp:31-40): one FP64 accumulator over the upper
// triangle row by row, q[i][j] * x_i * x_j for every j >= i (zero terms
// included, as the reference's loop has them).  The work-group stages Q in
// LDS first; every thread then reads the same addresses (broadcasts).
__global__ __launch_bounds__(256) void energy_small_kernel(int n, int W, const double *__restrict__ q,
                                                          int count,
                                                          const uint32_t *__restrict__ bits,
                                                          double *__restrict__ out) {
  __shared__ double ql[64 * 64];
  const int s = blockIdx.y;
  const double *qs = q + (size_t)s * n * n;
  for (int k = threadIdx.x; k < n * n; k += 256) ql[k] = qs[k];
  __syncthreads();
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= count) return;
  const uint32_t *x = bits + ((size_t)s * count + r) * W;
  const uint64_t m = (uint64_t)x[0] | (W > 1 ? (uint64_t)x[1] << 32 : 0ull);
  double e = 0.0;
  for (int i = 0; i < n; ++i) {
    const double xi = (double)((m >> i) & 1ull);
    const double *row = ql + i * n;
#pragma unroll 8
    for (int j = i; j < n; ++j) e += row[j] * xi * (double)((m >> j) & 1ull);
  }
  out[(size_t)s * count + r] = e;
}

void launch_energy_fp64(hipStream_t st, int n, int W, const double *q, int count,
                        const uint32_t *bits, double *out, double *scratch, float *fields, int ldf,
                        unsigned int *drift_max, int num_instances) {
  if (count <= 0 || num_instances <= 0) return;
  if (!fields && n <= 64) {
    energy_small_kernel<<<dim3((count + 255) / 256, num_instances), 256, 0, st>>>(n, W, q, count,
                                                                                bits, out);
    OS_HIP_CHECK(hipGetLastError());
    return;
  }
  const int ncb = (n + kTN - 1) / kTN;
  dim3 grid(ncb, (count + kTR - 1) / kTR, num_instances);
  if (fields)
    xq_tile_kernel<true><<<grid, 256, 0, st>>>(n, W, q, count, bits, scratch, ncb, fields, ldf,
                                                drift_max);
  else
    xq_tile_kernel<false><<<grid, 256, 0, st>>>(n, W, q, count, bits, scratch, ncb, nullptr, 0,
                                                 nullptr);
  OS_HIP_CHECK(hipGetLastError());
  const int rows = count * num_instances;
  reduce_partials_kernel<<<(rows + 255) / 256, 256, 0, st>>>(scratch, rows, ncb, out);
  OS_HIP_CHECK(hipGetLastError());
}

}  // namespace os
