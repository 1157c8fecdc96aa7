// gen.hip -- dense instances generated and packed on the GPU.
//
// Also the uniform instances of the dense configurations (Q_ij i.i.d.
// U[-1,1) fp32 on and above the diagonal), restating instances.dense_uniform.
//
// SK instances of the dense benchmark distribution are produced in device
// memory (no host generation or PCIe transfer), so that batches of thousands
// of N=1024 instances for training and evaluation cost only HBM bandwidth.
// The generator is the Philox restatement of the host generator
// (onesolver_amd/instances.py, sk_qubo): J_ij, i<j, from the Box-Muller pair
// of counter (i*n + j, 0, TAG_GEN_GAUSS) under key (seed, instance),
// J_ij = fp32(z / sqrt(n)); Q_ij = 4 J_ij, Q_ii = fp32(-2 sum_{j!=i} J_ij).
#include "os_kernels.h"

namespace os {

namespace {

constexpr uint32_t TAG_GEN_UNIFORM = 0xC1;
constexpr uint32_t TAG_GEN_GAUSS = 0xC2;

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

}  // namespace

__global__ __launch_bounds__(256) void gen_sk_offdiag_kernel(int n, uint32_t k0,
                                                             uint32_t inst_base,
                                                             double *__restrict__ q) {
  const int s = blockIdx.y;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)n * n) return;
  const int i = (int)(idx / n), j = (int)(idx % n);
  double *qs = q + (size_t)s * n * n;
  if (i == j) return;
  const uint64_t a = (uint64_t)min(i, j), b = (uint64_t)max(i, j);
  const uint64_t e = a * (uint64_t)n + b;
  const uint4 o = philox4x32_10(make_uint4((uint32_t)e, (uint32_t)(e >> 32), 0u, TAG_GEN_GAUSS),
                                k0, inst_base + (uint32_t)s);
  const double u1 = ((double)(o.x >> 8) + 1.0) * 0x1p-24;
  const double u2 = (double)(o.y >> 8) * 0x1p-24;
  const double z = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
  const float v = (float)(z / sqrt((double)n));
  qs[idx] = (double)(4.0f * v);
}

// Dense uniform instances (instances.py, dense_uniform): Q_ij, i <= j, =
// fp32((w >> 8) * 2^-23 - 1) from counter (i*n + j, 0, 0, TAG_GEN_UNIFORM),
// mirrored below the diagonal.
__global__ __launch_bounds__(256) void gen_uniform_kernel(int n, uint32_t k0, uint32_t inst_base,
                                                          double *__restrict__ q) {
  const int s = blockIdx.y;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)n * n) return;
  const int i = (int)(idx / n), j = (int)(idx % n);
  const uint64_t a = (uint64_t)min(i, j), b = (uint64_t)max(i, j);
  const uint64_t e = a * (uint64_t)n + b;
  const uint4 o = philox4x32_10(
      make_uint4((uint32_t)e, (uint32_t)(e >> 32), 0u, TAG_GEN_UNIFORM), k0, inst_base + (uint32_t)s);
  q[(size_t)s * n * n + idx] = (double)((float)(o.x >> 8) * 0x1p-23f - 1.0f);
}

__global__ __launch_bounds__(256) void gen_sk_diag_kernel(int I, int n, double *__restrict__ q) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)I * n) return;
  const int s = (int)(t / n), i = (int)(t % n);
  double *row = q + ((size_t)s * n + i) * n;
  double sum = 0.0;
  for (int j = 0; j < n; ++j)
    if (j != i) sum += row[j] * 0.25;  // J_ij (exact)
  row[i] = (double)(float)(-2.0 * sum);
}

// qpack[s][b][j][k] = hi(Q[16b+k][j]), [16+k] = lo(...), diagonal zeroed;
// qdb[s][b][k][c] = fp32 Q[16b+k][16b+c] (diagonal sub-blocks).
__global__ __launch_bounds__(256) void pack_q_kernel(int n, int np, const double *__restrict__ q,
                                                     uint16_t *__restrict__ qpack,
                                                     float *__restrict__ qdb) {
  const int s = blockIdx.y;
  const int nb = np / 16;
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;  // (b, j, k)
  if (t >= (size_t)nb * np * 16) return;
  const int k = (int)(t & 15);
  const int j = (int)((t >> 4) % np);
  const int b = (int)((t >> 4) / np);
  const int i = 16 * b + k;
  const double *qs = q + (size_t)s * n * n;
  float qf = 0.0f;
  if (i < n && j < n) qf = (float)qs[(size_t)i * n + j];
  uint16_t hi = 0, lo = 0;
  if (i != j && i < n && j < n) {
    hi = bf16_rne(qf);
    lo = bf16_rne(qf - __uint_as_float((uint32_t)hi << 16));
  }
  uint16_t *dst = qpack + (size_t)s * nb * np * 32 + ((size_t)b * np + j) * 32;
  dst[k] = hi;
  dst[16 + k] = lo;
  if (j / 16 == b) qdb[(size_t)s * nb * 256 + (size_t)b * 256 + k * 16 + (j % 16)] = (i != j) ? qf : 0.0f;
}

void launch_gen_sk(hipStream_t st, int I, int n, uint64_t seed, uint32_t inst_base, double *q) {
  const size_t nn = (size_t)n * n;
  gen_sk_offdiag_kernel<<<dim3((unsigned)((nn + 255) / 256), I), 256, 0, st>>>(
      n, (uint32_t)seed, inst_base, q);
  OS_HIP_CHECK(hipGetLastError());
  gen_sk_diag_kernel<<<(unsigned)(((size_t)I * n + 255) / 256), 256, 0, st>>>(I, n, q);
  OS_HIP_CHECK(hipGetLastError());
}

void launch_gen_uniform(hipStream_t st, int I, int n, uint64_t seed, uint32_t inst_base,
                        double *q) {
  const size_t nn = (size_t)n * n;
  gen_uniform_kernel<<<dim3((unsigned)((nn + 255) / 256), I), 256, 0, st>>>(n, (uint32_t)seed,
                                                                            inst_base, q);
  OS_HIP_CHECK(hipGetLastError());
}

void launch_pack_q(hipStream_t st, int I, int n, int np, const double *q, uint16_t *qpack,
                   float *qdb) {
  const size_t work = (size_t)(np / 16) * np * 16;
  pack_q_kernel<<<dim3((unsigned)((work + 255) / 256), I), 256, 0, st>>>(n, np, q, qpack, qdb);
  OS_HIP_CHECK(hipGetLastError());
}

}  // namespace os
