// mpnn.hip -- training and inference kernels of the warm-start network.
//
// A 3-layer message-passing network over the coupling graph of a dense
// instance predicts per-spin marginals of good annealed states.  Per layer
//   M = A r(H)                (neighbour aggregation, bf16 MFMA, fp32 acc)
//   Z = H r(Ws) + M r(Wm) + b,  H' = relu(Z)          (node update, fp32)
// with A = offdiag(Q) / scale (bf16), r = bf16 rounding of the weights and
// of the aggregation input; head: logit = H r(wo) + bo, p = sigmoid(logit).
// Features of spin i: (Q_ii, sum_j Q_ij) / scale, [i == 0], 1; padded to 64
// columns so every layer has 64 inputs (the zero columns of layer 1 stay
// zero in value and gradient).  Rows are padded to a multiple of 64 with
// zero features and zero couplings; they take no part in the loss, and
// their gradients are zero.
//
// Activations of a minibatch of B instances live in [B][Np][64] fp32 arrays
// (row = spin) and, as MFMA B operands, in [B][64][Np] bf16 (feature-major).
// Weight gradients are reduced deterministically: fixed 64-row chunks, then
// the chunk partials in order.
#include "os_kernels.h"

namespace os {

typedef __bf16 mp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mp_f32x16 __attribute__((ext_vector_type(16)));

namespace {

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_round(float f) {
  return __uint_as_float((uint32_t)to_bf16(f) << 16);
}

}  // namespace

// ---- features --------------------------------------------------------------

// rowstat[b][i] = (sum_{j != i} Q_ij^2, sum_{j != i} Q_ij) of instance q[b]
__global__ __launch_bounds__(256) void mp_rowstat_kernel(int B, int n, const double *__restrict__ q,
                                                         double *__restrict__ rowstat) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)B * n) return;
  const int b = (int)(t / n), i = (int)(t % n);
  const double *row = q + ((size_t)b * n + i) * n;
  double s2 = 0.0, s1 = 0.0;
  for (int j = 0; j < n; ++j) {
    if (j == i) continue;
    const double v = row[j];
    s2 += v * v;
    s1 += v;
  }
  rowstat[2 * t] = s2;
  rowstat[2 * t + 1] = s1;
}

__global__ void mp_scale_kernel(int B, int n, const double *__restrict__ rowstat,
                                double *__restrict__ scale) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += rowstat[2 * ((size_t)b * n + i)];
  s = sqrt(s / n);
  scale[b] = s > 0.0 ? s : 1.0;
}

// A[b][i][j] (bf16) and the features X0[b][i][:], XT0[b][:][i] (bf16)
__global__ __launch_bounds__(256) void mp_fill_kernel(int n, int Np, const double *__restrict__ q,
                                                      const double *__restrict__ rowstat,
                                                      const double *__restrict__ scale,
                                                      uint16_t *__restrict__ A,
                                                      float *__restrict__ X0,
                                                      uint16_t *__restrict__ XT0) {
  const int b = blockIdx.y;
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)Np * Np) return;
  const int i = (int)(t / Np), j = (int)(t % Np);
  const double sc = scale[b];
  const double *qb = q + (size_t)b * n * n;
  float a = 0.0f;
  if (i < n && j < n && i != j) a = (float)(qb[(size_t)i * n + j] / sc);
  A[(size_t)b * Np * Np + t] = to_bf16(a);
  if (j < 64) {
    float f = 0.0f;
    if (i < n) {
      if (j == 0) f = (float)(qb[(size_t)i * n + i] / sc);
      else if (j == 1) f = (float)(rowstat[2 * ((size_t)b * n + i) + 1] / sc);
      else if (j == 2) f = i == 0 ? 1.0f : 0.0f;
      else if (j == 3) f = 1.0f;
    }
    X0[((size_t)b * Np + i) * 64 + j] = f;
    XT0[((size_t)b * 64 + j) * Np + i] = to_bf16(f);
  }
}

// ---- sparse instances (CSR batch) -------------------------------------------
// The network on a batch of CSR instances, as on each instance's dense
// matrix: one row per spin, instance i at rows spin_off.. (instance-major,
// padded to a multiple of 64 with zero rows).  The scale of instance i is
// the dense one over scale_rows[i] rows (its size in the dense layout the
// network was trained on; n when null): rows without couplings add nothing
// but the row count.  A_jl = bf16((float)(Q_jl / scale)) as in mp_fill_kernel.

// scale[i] = sqrt(sum_j (sum_{l != j} Q_jl^2) / rows_i), the row sums first
__global__ void mp_csr_scale_kernel(int I, const CsrInst *__restrict__ inst,
                                    const int32_t *__restrict__ row_ptr,
                                    const double *__restrict__ val,
                                    const int32_t *__restrict__ scale_rows,
                                    double *__restrict__ scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  const CsrInst d = inst[i];
  const int32_t *rp = row_ptr + d.row_off;
  const double *vl = val + d.csr_off;
  double s = 0.0;
  for (int j = 0; j < d.n; ++j) {
    double s2 = 0.0;
    for (int k = rp[j]; k < rp[j + 1]; ++k) s2 += vl[k] * vl[k];
    s += s2;
  }
  const int rows = scale_rows ? scale_rows[i] : d.n;
  s = sqrt(s / rows);
  scale[i] = s > 0.0 ? s : 1.0;
}

// features of spin j: (Q_jj, sum_l Q_jl) / scale, 0, 1 (grid: rows / 256, I)
__global__ __launch_bounds__(256) void mp_csr_features_kernel(
    const CsrInst *__restrict__ inst, const int32_t *__restrict__ row_ptr,
    const double *__restrict__ val, const double *__restrict__ diag,
    const double *__restrict__ scale, float *__restrict__ X0) {
  const CsrInst d = inst[blockIdx.y];
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= d.n) return;
  const int32_t *rp = row_ptr + d.row_off;
  const double *vl = val + d.csr_off;
  const double sc = scale[blockIdx.y];
  double s1 = 0.0;
  for (int k = rp[j]; k < rp[j + 1]; ++k) s1 += vl[k];
  float *x = X0 + (size_t)(d.spin_off + j) * 64;
  x[0] = (float)(diag[d.spin_off + j] / sc);
  x[1] = (float)(s1 / sc);
  x[2] = 0.0f;
  x[3] = 1.0f;
  for (int f = 4; f < 64; ++f) x[f] = 0.0f;
}

// M[row j] = sum_k A_{j col_k} bf16(Hin[col_k]) in fp32 (one wave per spin,
// lane = feature; grid: rows / 4, I)
__global__ __launch_bounds__(256) void mp_csr_agg_kernel(
    const CsrInst *__restrict__ inst, const int32_t *__restrict__ row_ptr,
    const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ scale, const float *__restrict__ Hin, float *__restrict__ M) {
  const CsrInst d = inst[blockIdx.y];
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int f = threadIdx.x & 63;
  if (j >= d.n) return;
  const int32_t *rp = row_ptr + d.row_off;
  const int32_t *cl = col + d.csr_off;
  const double *vl = val + d.csr_off;
  const double sc = scale[blockIdx.y];
  const float *h = Hin + (size_t)d.spin_off * 64 + f;
  float acc = 0.0f;
  for (int k = rp[j]; k < rp[j + 1]; ++k) {
    const float a = bf16_round((float)(vl[k] / sc));
    acc = fmaf(a, bf16_round(h[(size_t)cl[k] * 64]), acc);
  }
  M[(size_t)(d.spin_off + j) * 64 + f] = acc;
}

void mp_launch_csr_features(hipStream_t st, int I, int max_n, const CsrInst *inst,
                            const int32_t *row_ptr, const double *val, const double *diag,
                            const int32_t *scale_rows, double *scale, float *X0) {
  if (I == 0 || max_n == 0) return;
  mp_csr_scale_kernel<<<(I + 63) / 64, 64, 0, st>>>(I, inst, row_ptr, val, scale_rows, scale);
  OS_HIP_CHECK(hipGetLastError());
  mp_csr_features_kernel<<<dim3((max_n + 255) / 256, I), 256, 0, st>>>(inst, row_ptr, val, diag,
                                                                       scale, X0);
  OS_HIP_CHECK(hipGetLastError());
}

void mp_launch_csr_agg(hipStream_t st, int I, int max_n, const CsrInst *inst,
                       const int32_t *row_ptr, const int32_t *col, const double *val,
                       const double *scale, const float *Hin, float *M) {
  if (I == 0 || max_n == 0) return;
  mp_csr_agg_kernel<<<dim3((max_n + 3) / 4, I), 256, 0, st>>>(inst, row_ptr, col, val, scale,
                                                              Hin, M);
  OS_HIP_CHECK(hipGetLastError());
}

// ---- neighbour aggregation: Y[b] (+)= A[b] . X[b], X given feature-major bf16
template <bool kAcc>
__global__ __launch_bounds__(256) void mp_agg_kernel(int Np, const uint16_t *__restrict__ A,
                                                     const uint16_t *__restrict__ XT,
                                                     float *__restrict__ Y) {
  const int b = blockIdx.y;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lc = lane & 31, kh = lane >> 5;
  const int r0 = blockIdx.x * 64 + 32 * (w >> 1);
  const int f0 = 32 * (w & 1);
  const uint4 *arow = reinterpret_cast<const uint4 *>(A + ((size_t)b * Np + r0 + lc) * Np) + kh;
  const uint4 *xrow = reinterpret_cast<const uint4 *>(XT + ((size_t)b * 64 + f0 + lc) * Np) + kh;
  mp_f32x16 acc;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
  const int steps = Np / 16;
  for (int s = 0; s < steps; ++s) {
    const uint4 a = arow[2 * s];
    const uint4 x = xrow[2 * s];
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mp_bf16x8, a),
                                                  __builtin_bit_cast(mp_bf16x8, x), acc, 0, 0,
                                                  0);
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = r0 + 8 * (i >> 2) + 4 * kh + (i & 3);
    float *y = Y + ((size_t)b * Np + row) * 64 + f0 + lc;
    if (kAcc)
      *y += acc[i];
    else
      *y = acc[i];
  }
}

// ---- node update -----------------------------------------------------------
constexpr int kRowsPerGroup = 16;

// Z = Hin Ws + M Wm + b, H = relu(Z), HT = bf16(H) feature-major (optional)
__global__ __launch_bounds__(256) void mp_node_fwd_kernel(
    int Np, const float *__restrict__ Hin, const float *__restrict__ M,
    const float *__restrict__ wl, float *__restrict__ Z, float *__restrict__ H,
    uint16_t *__restrict__ HT) {
  __shared__ float ws[64][64], wm[64][64], bias[64];
  __shared__ float xs[kRowsPerGroup][64], ms[kRowsPerGroup][64];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  for (int k = tid; k < 4096; k += 256) {
    ws[k >> 6][k & 63] = wl[k];
    wm[k >> 6][k & 63] = wl[4096 + k];
  }
  if (tid < 64) bias[tid] = wl[8192 + tid];
  const size_t row0 = (size_t)b * Np + (size_t)blockIdx.x * kRowsPerGroup;
  for (int k = tid; k < kRowsPerGroup * 64; k += 256) {
    xs[k >> 6][k & 63] = Hin[row0 * 64 + k];
    ms[k >> 6][k & 63] = M[row0 * 64 + k];
  }
  __syncthreads();
  const int o = tid & 63;
  for (int r = tid >> 6; r < kRowsPerGroup; r += 4) {
    float z = bias[o];
    for (int a = 0; a < 64; ++a) z = fmaf(xs[r][a], ws[a][o], z);
    for (int a = 0; a < 64; ++a) z = fmaf(ms[r][a], wm[a][o], z);
    const size_t row = row0 + r;
    Z[row * 64 + o] = z;
    const float h = fmaxf(z, 0.0f);
    H[row * 64 + o] = h;
    if (HT) {
      const int i = (int)(row - (size_t)b * Np);
      HT[((size_t)b * 64 + o) * Np + i] = to_bf16(h);
    }
  }
}

// dZ = dH * (Z > 0); dHprev = dZ Ws^T; dMsgT = bf16(dZ Wm^T) feature-major
__global__ __launch_bounds__(256) void mp_node_bwd_kernel(
    int Np, const float *__restrict__ dH, const float *__restrict__ Z,
    const float *__restrict__ wl, float *__restrict__ dZ, float *__restrict__ dHprev,
    uint16_t *__restrict__ dMsgT) {
  __shared__ float ws[64][65], wm[64][65];
  __shared__ float dz[kRowsPerGroup][64];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  for (int k = tid; k < 4096; k += 256) {
    ws[k >> 6][k & 63] = wl[k];
    wm[k >> 6][k & 63] = wl[4096 + k];
  }
  const size_t row0 = (size_t)b * Np + (size_t)blockIdx.x * kRowsPerGroup;
  for (int k = tid; k < kRowsPerGroup * 64; k += 256) {
    const size_t at = row0 * 64 + k;
    const float g = Z[at] > 0.0f ? dH[at] : 0.0f;
    dz[k >> 6][k & 63] = g;
    dZ[at] = g;
  }
  __syncthreads();
  if (!dHprev) return;
  const int a = tid & 63;
  for (int r = tid >> 6; r < kRowsPerGroup; r += 4) {
    float s1 = 0.0f, s2 = 0.0f;
    for (int o = 0; o < 64; ++o) {
      s1 = fmaf(dz[r][o], ws[a][o], s1);
      s2 = fmaf(dz[r][o], wm[a][o], s2);
    }
    const size_t row = row0 + r;
    dHprev[row * 64 + a] = s1;
    const int i = (int)(row - (size_t)b * Np);
    dMsgT[((size_t)b * 64 + a) * Np + i] = to_bf16(s2);
  }
}

// ---- head ------------------------------------------------------------------
// logit = bo + H3 wo; p = sigmoid; with labels: BCE terms, dlogit and dH3.
// Canonical labels: y_i = x_i xor (1 - x_0) (spin 0 is pinned to 1).
__global__ __launch_bounds__(256) void mp_head_kernel(
    int B, int n, int Np, const float *__restrict__ H3, const float *__restrict__ wo,
    const uint8_t *__restrict__ labels, float inv_count, float *__restrict__ p,
    double *__restrict__ lossterm, float *__restrict__ dlogit, float *__restrict__ dH3) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)B * Np) return;
  const int b = (int)(t / Np), i = (int)(t % Np);
  const float *h = H3 + t * 64;
  float l = wo[64];
  for (int a = 0; a < 64; ++a) l = fmaf(h[a], wo[a], l);
  const float pr = 1.0f / (1.0f + expf(-l));
  if (p) p[t] = pr;
  if (!labels) return;
  float g = 0.0f;
  double term = 0.0;
  if (i < n) {
    const uint8_t *x = labels + (size_t)b * n;
    const float y = (float)(x[i] ^ (uint8_t)(1 - x[0]));
    const double ld = (double)l;
    term = (ld > 0.0 ? ld : 0.0) - ld * (double)y + log1p(exp(-fabs(ld)));
    g = (pr - y) * inv_count;
  }
  lossterm[t] = term;
  dlogit[t] = g;
  for (int a = 0; a < 64; ++a) dH3[t * 64 + a] = g * wo[a];
}

// ---- weight gradients --------------------------------------------------------
// partial[c][a][o] = sum over rows [64c, 64c+64) of X[row][a] * G[row][o]
// (a = 64: the constant 1, i.e. the bias gradient); W = width of G (64 or 1)
__global__ __launch_bounds__(256) void mp_wgrad_partial_kernel(int W, const float *__restrict__ X,
                                                               const float *__restrict__ G,
                                                               float *__restrict__ partial) {
  __shared__ float xs[64][65];
  __shared__ float gs[64][64];
  const int c = blockIdx.x;
  const int tid = threadIdx.x;
  const size_t row0 = (size_t)c * 64;
  for (int k = tid; k < 64 * 64; k += 256) xs[k >> 6][k & 63] = X[row0 * 64 + k];
  for (int k = tid; k < 64 * W; k += 256) gs[k / W][k % W] = G[row0 * W + k];
  __syncthreads();
  float *out = partial + (size_t)c * 65 * W;
  for (int idx = tid; idx < 65 * W; idx += 256) {
    const int a = idx / W, o = idx % W;
    float s = 0.0f;
    if (a < 64)
      for (int r = 0; r < 64; ++r) s = fmaf(xs[r][a], gs[r][o], s);
    else
      for (int r = 0; r < 64; ++r) s += gs[r][o];
    out[idx] = s;
  }
}

// grad_w[a][o] = sum_c partial[c][a][o] (a < 64), grad_b[o] = sum_c partial[c][64][o]
__global__ __launch_bounds__(256) void mp_wgrad_sum_kernel(int W, int chunks,
                                                           const float *__restrict__ partial,
                                                           float *__restrict__ grad_w,
                                                           float *__restrict__ grad_b) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= 65 * W) return;
  float s = 0.0f;
  for (int c = 0; c < chunks; ++c) s += partial[(size_t)c * 65 * W + idx];
  if (idx < 64 * W)
    grad_w[idx] = s;
  else if (grad_b)
    grad_b[idx - 64 * W] = s;
}

// sum of the loss terms in a fixed order (one work-group)
__global__ __launch_bounds__(256) void mp_loss_sum_kernel(size_t count,
                                                          const double *__restrict__ terms,
                                                          double *__restrict__ out) {
  __shared__ double part[256];
  double s = 0.0;
  for (size_t k = threadIdx.x; k < count; k += 256) s += terms[k];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

// Adam on the fp32 master weights; the compute copy rounds weight matrices
// and the head weights to bf16 (biases stay fp32).  mask[k] = 0: padding.
__global__ __launch_bounds__(256) void mp_adam_kernel(int P, const float *__restrict__ grad,
                                                      const uint8_t *__restrict__ kind,
                                                      float *__restrict__ w,
                                                      float *__restrict__ m,
                                                      float *__restrict__ v,
                                                      float *__restrict__ wc, float lr,
                                                      float b1, float b2, float eps, float c1,
                                                      float c2) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= P) return;
  const uint8_t kd = kind[k];
  if (kd == 0) return;
  const float g = grad[k];
  const float mk = b1 * m[k] + (1.0f - b1) * g;
  const float vk = b2 * v[k] + (1.0f - b2) * g * g;
  m[k] = mk;
  v[k] = vk;
  const float wk = w[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
  w[k] = wk;
  wc[k] = kd == 1 ? bf16_round(wk) : wk;
}

__global__ __launch_bounds__(256) void mp_compute_copy_kernel(int P,
                                                              const uint8_t *__restrict__ kind,
                                                              const float *__restrict__ w,
                                                              float *__restrict__ wc) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= P) return;
  const uint8_t kd = kind[k];
  wc[k] = kd == 1 ? bf16_round(w[k]) : (kd == 2 ? w[k] : 0.0f);
}

// ---- launchers -----------------------------------------------------------------

void mp_launch_features(hipStream_t st, int B, int n, int Np, const double *q, double *rowstat,
                        double *scale, uint16_t *A, float *X0, uint16_t *XT0) {
  mp_rowstat_kernel<<<(unsigned)(((size_t)B * n + 255) / 256), 256, 0, st>>>(B, n, q, rowstat);
  OS_HIP_CHECK(hipGetLastError());
  mp_scale_kernel<<<(B + 63) / 64, 64, 0, st>>>(B, n, rowstat, scale);
  OS_HIP_CHECK(hipGetLastError());
  mp_fill_kernel<<<dim3((unsigned)(((size_t)Np * Np + 255) / 256), B), 256, 0, st>>>(
      n, Np, q, rowstat, scale, A, X0, XT0);
  OS_HIP_CHECK(hipGetLastError());
}

void mp_launch_agg(hipStream_t st, int B, int Np, const uint16_t *A, const uint16_t *XT, float *Y,
                   bool accumulate) {
  const dim3 grid(Np / 64, B);
  if (accumulate)
    mp_agg_kernel<true><<<grid, 256, 0, st>>>(Np, A, XT, Y);
  else
    mp_agg_kernel<false><<<grid, 256, 0, st>>>(Np, A, XT, Y);
  OS_HIP_CHECK(hipGetLastError());
}

void mp_launch_node_fwd(hipStream_t st, int B, int Np, const float *Hin, const float *M,
                        const float *wl, float *Z, float *H, uint16_t *HT) {
  mp_node_fwd_kernel<<<dim3(Np / kRowsPerGroup, B), 256, 0, st>>>(Np, Hin, M, wl, Z, H, HT);
  OS_HIP_CHECK(hipGetLastError());
}

void mp_launch_node_bwd(hipStream_t st, int B, int Np, const float *dH, const float *Z,
                        const float *wl, float *dZ, float *dHprev, uint16_t *dMsgT) {
  mp_node_bwd_kernel<<<dim3(Np / kRowsPerGroup, B), 256, 0, st>>>(Np, dH, Z, wl, dZ, dHprev,
                                                                  dMsgT);
  OS_HIP_CHECK(hipGetLastError());
}

void mp_launch_head(hipStream_t st, int B, int n, int Np, const float *H3, const float *wo,
                    const uint8_t *labels, float inv_count, float *p, double *lossterm,
                    float *dlogit, float *dH3) {
  mp_head_kernel<<<(unsigned)(((size_t)B * Np + 255) / 256), 256, 0, st>>>(
      B, n, Np, H3, wo, labels, inv_count, p, lossterm, dlogit, dH3);
  OS_HIP_CHECK(hipGetLastError());
}

void mp_launch_wgrad(hipStream_t st, int rows, int W, const float *X, const float *G,
                     float *partial, float *grad_w, float *grad_b) {
  const int chunks = rows / 64;
  mp_wgrad_partial_kernel<<<chunks, 256, 0, st>>>(W, X, G, partial);
  OS_HIP_CHECK(hipGetLastError());
  mp_wgrad_sum_kernel<<<(65 * W + 255) / 256, 256, 0, st>>>(W, chunks, partial, grad_w, grad_b);
  OS_HIP_CHECK(hipGetLastError());
}

void mp_launch_loss_sum(hipStream_t st, size_t count, const double *terms, double *out) {
  mp_loss_sum_kernel<<<1, 256, 0, st>>>(count, terms, out);
  OS_HIP_CHECK(hipGetLastError());
}

void mp_launch_adam(hipStream_t st, int P, const float *grad, const uint8_t *kind, float *w,
                    float *m, float *v, float *wc, float lr, float b1, float b2, float eps,
                    int step) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  mp_adam_kernel<<<(P + 255) / 256, 256, 0, st>>>(P, grad, kind, w, m, v, wc, lr, b1, b2, eps,
                                                  c1, c2);
  OS_HIP_CHECK(hipGetLastError());
}

void mp_launch_compute_copy(hipStream_t st, int P, const uint8_t *kind, const float *w,
                            float *wc) {
  mp_compute_copy_kernel<<<(P + 255) / 256, 256, 0, st>>>(P, kind, w, wc);
  OS_HIP_CHECK(hipGetLastError());
}

}  // namespace os
