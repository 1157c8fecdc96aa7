// os_common.h -- shared device/host definitions of the MI355X annealing library.
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11) with
// key = (seed_lo, seed_hi) and a counter that names what the draw is for, so
// every draw is a pure function of (seed, instance, replica, step, site) and
// results do not depend on launch shape or scheduling.  This replaces the
// per-work-item oneMKL engine of include/simulated_annealing/annealing.hpp:87
// (engine(1234, i)) and the distributions of device_rng.hpp:26-30.
//
// Counter layouts (word 3 = (instance << 8) | tag):
//   reference-mode init bit j of replica i : (j>>7, i, 0, TAG_REF_INIT)  word (j>>5)&3, bit j&31
//   reference-mode attempt t of replica i  : (t_lo, i, t_hi, TAG_REF_STEP)
//        site = floor(w0 * N / 2^32), u = w1 * 2^-32
//   sweep-mode init bit j of replica r     : (j>>7, r, 0, TAG_SWEEP_INIT)
//   sweep-mode threshold of spin j, sweep t: (j>>2, r, t, TAG_SWEEP_THR) word j&3
//   warm-start draw of spin j, replica r   : (j>>2, r, 0, TAG_WARM)      word j&3
//   replica exchange, slots (k, k+1) of ladder l in round t:
//                                            (k, l, t, TAG_EXCHANGE)     word 0
//   population annealing, resampling round k: (k, 0, 0, TAG_RESAMPLE)   words 0, 1
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <stdexcept>
#include <string>

namespace os {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

enum : uint32_t {
  TAG_REF_INIT = 0x01,
  TAG_REF_STEP = 0x02,
  TAG_SWEEP_INIT = 0x10,
  TAG_SWEEP_THR = 0x11,
  TAG_WARM = 0x12,
  TAG_EXCHANGE = 0x13,
  TAG_RESAMPLE = 0x14,
};

__host__ __device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += kPhiloxW0;
      k1 += kPhiloxW1;
    }
    const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
    const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
    c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1,
                   (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0);
  }
  return c;
}

__host__ __device__ __forceinline__ uint32_t pick(const uint4 &v, uint32_t i) {
  return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}

// -ln(u) of the open-interval 24-bit uniform u = (2*(w>>9)+1) * 2^-24.
// u = m * 2^e with m in [sqrt(1/2), sqrt(2)); ln(1+f), f = m - 1, is
// f + f^2 * P(f) with P a degree-6 least-squares fit (relative error
// <= 2.2e-7 over the interval).  Each step is exact or one correctly rounded
// fp32 operation, so host and device results agree bit for bit.
__host__ __device__ __forceinline__ float neg_ln_u(uint32_t w) {
  const float u = (float)((w >> 9) * 2u + 1u) * 0x1p-24f;
  int e;
  float m = frexpf(u, &e);
  if (m < 0x1.6a09e6p-1f) {
    m = m + m;
    e = e - 1;
  }
  const float f = m - 1.0f;
  float p = -0x1.9190e0p-4f;
  p = fmaf(p, f, 0x1.4953eep-3f);
  p = fmaf(p, f, -0x1.621a72p-3f);
  p = fmaf(p, f, 0x1.97f830p-3f);
  p = fmaf(p, f, -0x1.ff5ac2p-3f);
  p = fmaf(p, f, 0x1.5557dep-2f);
  p = fmaf(p, f, -0x1.000080p-1f);
  const float f2 = f * f;
  float l = fmaf(f2, p, f);
  l = fmaf((float)e, 0x1.62e430p-1f, l);
  return -l;
}

// Population annealing, host side.  The uniform of resampling round k:
// words 0 and 1 of Philox counter (k, 0, 0, tag), 53 bits.
inline double resample_uniform(uint32_t k0, uint32_t k1, uint32_t tag, uint32_t k) {
  const uint4 o = philox4x32_10(make_uint4(k, 0u, 0u, tag), k0, k1);
  return (double)(((uint64_t)(o.x >> 5) << 26) | (uint64_t)(o.y >> 6)) * 0x1p-53;
}

// Systematic resampling of R replicas with energies E from scale 1/beta to
// 1/(beta + dbeta): anc[j] = first i with W_i > (j + u) / R * W_{R-1}, W the
// running sums of exp(-dbeta (E_i - min E)).  W: R doubles of scratch.
inline void resample(const double *E, int R, double dbeta, double u, uint32_t *anc, double *W) {
  double emin = E[0];
  for (int i = 1; i < R; ++i) emin = E[i] < emin ? E[i] : emin;
  double acc = 0.0;
  for (int i = 0; i < R; ++i) {
    acc += std::exp(-dbeta * (E[i] - emin));
    W[i] = acc;
  }
  int i = 0;
  for (int j = 0; j < R; ++j) {
    const double target = ((double)j + u) / (double)R * W[R - 1];
    while (i + 1 < R && W[i] <= target) ++i;
    anc[j] = (uint32_t)i;
  }
}

// Library error with a return code; caught at the C-ABI boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

}  // namespace os

#define OS_HIP_CHECK(expr)                                                          \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess)                                                           \
      throw ::os::Error(-2, std::string(#expr) + ": " + hipGetErrorString(_e));     \
  } while (0)

#define OS_REQUIRE(cond, msg)                                                       \
  do {                                                                              \
    if (!(cond)) throw ::os::Error(-1, (msg));                                      \
  } while (0)
