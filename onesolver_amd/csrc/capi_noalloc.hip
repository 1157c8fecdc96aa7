// capi_noalloc.hip -- the entry point that allocates no device memory.
//
// os_anneal_reference_noalloc runs the kernels of os_anneal_reference (the
// reference-mode chains, the FP64 energies of the best states, the byte
// unpacking) with every array the caller does not pass -- bit-packed best
// states, running best energies, FP64 energies, energy partials and, above
// the LDS limit, the FP64 fields -- in `scratch`, a fixed device array of
// kScratchBytes per GPU.  The address of each GPU's copy of the array is
// looked up once per device when a context is created on it
// (noalloc_setup, called by the Ctx constructor with the device current) and
// kept in g_scratch; host code uses only those addresses, never the array's
// name.  Every call on a GPU shares that GPU's array: calls that overlap on
// one GPU (from any context on it) race.  Nothing in the array has to start
// at zero: each kernel writes every element it later reads.
#include <mutex>
#include <string>

#include "../../include/onesolver.h"
#include "os_ctx.h"
#include "os_kernels.h"

namespace os {

constexpr size_t kScratchBytes = (size_t)OS_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
constexpr size_t kAlign = 256;

std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// Replica with the lowest energy, first index on ties
// (include/simulated_annealing/annealing.hpp:134-136: std::min_element).
__global__ __launch_bounds__(256) void argmin_kernel(const double *__restrict__ e, uint32_t count,
                                                     int32_t *__restrict__ best) {
  __shared__ double se[256];
  __shared__ uint32_t si[256];
  double be = 0.0;
  uint32_t bi = 0xFFFFFFFFu;
  for (uint32_t i = threadIdx.x; i < count; i += 256) {
    const double v = e[i];
    if (bi == 0xFFFFFFFFu || v < be) {
      be = v;
      bi = i;
    }
  }
  se[threadIdx.x] = be;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      const double v = se[threadIdx.x + off];
      const uint32_t i = si[threadIdx.x + off];
      const uint32_t mi = si[threadIdx.x];
      if (i != 0xFFFFFFFFu &&
          (mi == 0xFFFFFFFFu || v < se[threadIdx.x] || (v == se[threadIdx.x] && i < mi))) {
        se[threadIdx.x] = v;
        si[threadIdx.x] = i;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *best = (int32_t)si[0];
}

}  // namespace

void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  g_scratch[device] = e == hipSuccess ? p : nullptr;
  g_scratch_done[device] = true;
}

// Bytes of scratch one call needs (the layout of os_anneal_reference_noalloc).
static size_t noalloc_need(int n, uint32_t num_tries, size_t *off_run, size_t *off_e,
                           size_t *off_part, size_t *off_fields) {
  const size_t W = (size_t)(n + 31) / 32;
  size_t at = align_up(sizeof(uint32_t) * num_tries * W);  // best bits
  *off_run = at;
  at += align_up(sizeof(double) * num_tries);
  *off_e = at;
  at += align_up(sizeof(double) * num_tries);
  *off_part = at;
  at += align_up(energy_scratch_bytes(n, (int)num_tries));
  *off_fields = at;
  if (n > ref_lds_limit()) at += align_up(sizeof(double) * (size_t)n * num_tries);
  return at;
}

}  // namespace os

using os::Ctx;
using os::guarded;

extern "C" {

size_t os_noalloc_scratch_bytes(void) { return os::kScratchBytes; }

int os_anneal_reference_noalloc(os_ctx *ctx_, int n, const double *d_q_sym, const double *d_beta,
                                int num_iter, int sweeps_per_beta, uint32_t num_tries,
                                uint64_t seed, uint8_t *d_best_states, double *d_best_energies,
                                int32_t *d_best_index) {
  return guarded([&] {
    OS_REQUIRE(ctx_ != nullptr, "context is null");
    OS_REQUIRE(n >= 1, "n must be positive");
    OS_REQUIRE(d_q_sym != nullptr, "q_sym is null");
    OS_REQUIRE(num_iter >= 0, "num_iter must be non-negative");
    OS_REQUIRE(num_iter == 0 || d_beta != nullptr, "beta is null");
    OS_REQUIRE(sweeps_per_beta >= 1, "sweeps_per_beta must be positive");
    OS_REQUIRE(num_tries >= 1, "num_tries must be positive");
    if (n > 32768)
      throw os::Error(OS_ERR_UNSUPPORTED, "reference-mode annealing supports n <= 32768");
    size_t off_run, off_e, off_part, off_fields;
    const size_t need = os::noalloc_need(n, num_tries, &off_run, &off_e, &off_part, &off_fields);
    if (need > os::kScratchBytes)
      throw os::Error(OS_ERR_UNSUPPORTED,
                      "os_anneal_reference_noalloc: n = " + std::to_string(n) + " with " +
                          std::to_string(num_tries) + " tries needs " + std::to_string(need) +
                          " bytes of scratch; the fixed scratch holds " +
                          std::to_string(os::kScratchBytes));
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    unsigned char *base = nullptr;
    {
      std::lock_guard<std::mutex> lock(os::g_scratch_mu);
      if (c.device < 0 || c.device >= os::kMaxDevices || !os::g_scratch_done[c.device])
        throw os::Error(OS_ERR_HIP, "no scratch address for this device");
      if (!os::g_scratch[c.device])
        throw os::Error(OS_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                        hipGetErrorString(os::g_scratch_err[c.device]));
      base = static_cast<unsigned char *>(os::g_scratch[c.device]);
    }
    const int W = (n + 31) / 32;
    uint32_t *bits = reinterpret_cast<uint32_t *>(base);
    double *run = reinterpret_cast<double *>(base + off_run);
    double *energies = reinterpret_cast<double *>(base + off_e);
    double *partials = reinterpret_cast<double *>(base + off_part);
    double *fields =
        n > os::ref_lds_limit() ? reinterpret_cast<double *>(base + off_fields) : nullptr;
    os::launch_ref_anneal(c.stream, n, d_q_sym, d_beta, num_iter, sweeps_per_beta, num_tries,
                          seed, fields, bits, run);
    os::launch_energy_fp64(c.stream, n, W, d_q_sym, (int)num_tries, bits, energies, partials,
                           nullptr, 0, nullptr);
    if (d_best_energies)
      OS_HIP_CHECK(hipMemcpyAsync(d_best_energies, energies, sizeof(double) * num_tries,
                                  hipMemcpyDeviceToDevice, c.stream));
    if (d_best_states) os::launch_unpack_bits(c.stream, bits, n, (int)num_tries, W, d_best_states);
    if (d_best_index) {
      os::argmin_kernel<<<1, 256, 0, c.stream>>>(energies, num_tries, d_best_index);
      OS_HIP_CHECK(hipGetLastError());
    }
    c.sync();
  });
}

}  // extern "C"
