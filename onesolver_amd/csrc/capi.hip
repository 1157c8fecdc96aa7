// capi.hip -- extern "C" entry points of libonesolver.so (declared in include/onesolver.h).
//
// Each entry point validates its arguments, runs on the context's stream and
// converts every failure into a negative return code plus a thread-local
// message, mirroring the reference's exception types (std::invalid_argument
// for bad input, sycl::exception for device failures).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/onesolver.h"
#include "os_ctx.h"
#include "os_kernels.h"

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace os {

void set_last_error(const char *msg) { g_last_error = msg; }

void DevBuf::ensure(size_t bytes) {
  if (bytes <= cap) return;
  release();
  if (bytes == 0) return;
  OS_HIP_CHECK(hipMalloc(&ptr, bytes));
  cap = bytes;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  cap = 0;
}

Ctx::Ctx(int dev) : device(dev) {
  int count = 0;
  OS_HIP_CHECK(hipGetDeviceCount(&count));
  OS_REQUIRE(dev >= 0 && dev < count, "device index out of range");
  OS_HIP_CHECK(hipSetDevice(dev));
  hipDeviceProp_t prop;
  OS_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  name = prop.name;
  arch = prop.gcnArchName;
  num_cus = prop.multiProcessorCount;
  // the marketing name can be empty (e.g. without access to the board's
  // product data): name the device by its architecture and CU count then
  if (name.empty()) name = arch + " (" + std::to_string(num_cus) + " CUs)";
  OS_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  OS_HIP_CHECK(hipEventCreate(&ev_start));
  OS_HIP_CHECK(hipEventCreate(&ev_stop));
  noalloc_setup(dev);
}

Ctx::~Ctx() {
  (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto *b : {&b_q, &b_beta, &b_bits, &b_run, &b_energy, &b_scratch, &b_fields, &b_bytes})
    b->release();
  for (DevBuf &b : ingest) b.release();
  for (int k = 0; k < 2; ++k) {
    if (pinned[k]) (void)hipHostFree(pinned[k]);
    if (pinned_ev[k]) (void)hipEventDestroy(pinned_ev[k]);
  }
  if (stage) (void)hipHostFree(stage);
  for (hipEvent_t e : ingest_ev)
    if (e) (void)hipEventDestroy(e);
  if (ev_start) (void)hipEventDestroy(ev_start);
  if (ev_stop) (void)hipEventDestroy(ev_stop);
  if (stream) (void)hipStreamDestroy(stream);
}

void Ctx::activate() const { OS_HIP_CHECK(hipSetDevice(device)); }

void *Ctx::staging(size_t bytes) {
  OS_REQUIRE(bytes <= kStageMax, "staging request above its limit");
  if (!stage) {
    OS_HIP_CHECK(hipHostMalloc(&stage, kStageMax, hipHostMallocDefault));
    stage_bytes = kStageMax;
  }
  return stage;
}

void Ctx::sync() const { OS_HIP_CHECK(hipStreamSynchronize(stream)); }

}  // namespace os

using os::Ctx;
using os::guarded;

extern "C" {

int os_version(void) { return 10000; }

const char *os_last_error(void) { return g_last_error.c_str(); }

int os_device_count(int *count) {
  return guarded([&] {
    OS_REQUIRE(count != nullptr, "count is null");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e == hipErrorNoDevice) c = 0;
    else OS_HIP_CHECK(e);
    *count = c;
  });
}

int os_ctx_create(int device, os_ctx **out) {
  return guarded([&] {
    OS_REQUIRE(out != nullptr, "out is null");
    *out = reinterpret_cast<os_ctx *>(new Ctx(device));
  });
}

int os_ctx_destroy(os_ctx *ctx) {
  return guarded([&] { delete reinterpret_cast<Ctx *>(ctx); });
}

int os_ctx_device_name(os_ctx *ctx, char *buf, size_t len) {
  return guarded([&] {
    OS_REQUIRE(ctx && buf && len > 0, "invalid argument");
    const std::string &nm = reinterpret_cast<Ctx *>(ctx)->name;
    const size_t k = std::min(len - 1, nm.size());
    std::memcpy(buf, nm.data(), k);
    buf[k] = 0;
  });
}

int os_energy_fp64(os_ctx *ctx_, int n, const double *q_sym, uint32_t count,
                   const uint8_t *states, double *energies) {
  return guarded([&] {
    OS_REQUIRE(ctx_ != nullptr, "context is null");
    OS_REQUIRE(n >= 1, "n must be positive");
    OS_REQUIRE(q_sym && energies && (states || count == 0), "null array");
    if (count == 0) return;
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    const int W = (n + 31) / 32;
    const size_t qbytes = sizeof(double) * (size_t)n * n;
    c.b_q.ensure(qbytes);
    c.b_bytes.ensure((size_t)count * n);
    c.b_bits.ensure(sizeof(uint32_t) * (size_t)count * W);
    c.b_energy.ensure(sizeof(double) * count);
    c.b_scratch.ensure(os::energy_scratch_bytes(n, count));
    OS_HIP_CHECK(hipMemcpyAsync(c.b_q.ptr, q_sym, qbytes, hipMemcpyHostToDevice, c.stream));
    OS_HIP_CHECK(hipMemcpyAsync(c.b_bytes.ptr, states, (size_t)count * n, hipMemcpyHostToDevice,
                                c.stream));
    os::launch_pack_bytes(c.stream, c.b_bytes.as<uint8_t>(), n, count, W, c.b_bits.as<uint32_t>());
    os::launch_energy_fp64(c.stream, n, W, c.b_q.as<double>(), count, c.b_bits.as<uint32_t>(),
                           c.b_energy.as<double>(), c.b_scratch.as<double>(), nullptr, 0, nullptr);
    OS_HIP_CHECK(hipMemcpyAsync(energies, c.b_energy.ptr, sizeof(double) * count,
                                hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

int os_exhaustive(os_ctx *ctx_, int n, const double *q_upper, uint8_t *best_state,
                  double *best_energy) {
  return guarded([&] {
    OS_REQUIRE(ctx_ != nullptr, "context is null");
    OS_REQUIRE(n >= 1, "n must be positive");
    OS_REQUIRE(q_upper && best_state && best_energy, "null array");
    if (n > os::exhaustive_max_n())
      throw os::Error(OS_ERR_UNSUPPORTED,
                      "exhaustive search supports n <= " + std::to_string(os::exhaustive_max_n()));
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    bool finite = true;
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j) finite = finite && std::isfinite(q_upper[(size_t)i * n + j]);
    const size_t qbytes = sizeof(double) * (size_t)n * n;
    const size_t parts = os::exhaustive_partials(n);
    c.b_q.ensure(qbytes);
    c.b_energy.ensure(sizeof(double) * parts);
    c.b_scratch.ensure(sizeof(uint64_t) * parts);
    c.b_fields.ensure(sizeof(double) * std::max<size_t>(1, os::exhaustive_thread_slots(n)));
    OS_HIP_CHECK(hipMemcpyAsync(c.b_q.ptr, q_upper, qbytes, hipMemcpyHostToDevice, c.stream));
    double e = 0.0;
    uint64_t x = 0;
    os::exhaustive_solve(c.stream, n, c.b_q.as<double>(), finite, c.b_fields.as<double>(),
                         c.b_energy.as<double>(), c.b_scratch.as<uint64_t>(), parts, &e, &x);
    for (int i = 0; i < n; ++i) best_state[i] = (uint8_t)((x >> i) & 1u);
    *best_energy = e;
  });
}

int os_anneal_reference(os_ctx *ctx_, int n, const double *q_sym, const double *beta,
                        int num_iter, int sweeps_per_beta, uint32_t num_tries, uint64_t seed,
                        uint8_t *best_states, double *best_energies, uint8_t *best_state,
                        double *best_energy) {
  return guarded([&] {
    OS_REQUIRE(ctx_ != nullptr, "context is null");
    OS_REQUIRE(n >= 1, "n must be positive");
    OS_REQUIRE(q_sym != nullptr, "q_sym is null");
    OS_REQUIRE(num_iter >= 0, "num_iter must be non-negative");
    OS_REQUIRE(num_iter == 0 || beta != nullptr, "beta is null");
    OS_REQUIRE(sweeps_per_beta >= 1, "sweeps_per_beta must be positive");
    OS_REQUIRE(num_tries >= 1, "num_tries must be positive");
    if (n > 32768)
      throw os::Error(OS_ERR_UNSUPPORTED, "reference-mode annealing supports n <= 32768");
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    const int W = (n + 31) / 32;
    const size_t qbytes = sizeof(double) * (size_t)n * n;
    c.b_q.ensure(qbytes);
    c.b_beta.ensure(sizeof(double) * std::max(num_iter, 1));
    c.b_bits.ensure(sizeof(uint32_t) * (size_t)num_tries * W);
    c.b_run.ensure(sizeof(double) * num_tries);
    c.b_energy.ensure(sizeof(double) * num_tries);
    c.b_scratch.ensure(os::energy_scratch_bytes(n, num_tries));
    if (n > os::ref_lds_limit()) c.b_fields.ensure(sizeof(double) * (size_t)num_tries * n);
    // Small runs move Q and beta up in one copy and every replica's state
    // words and energies down in one copy, through the context's pinned
    // staging buffer: beta sits behind Q in b_q, the energies behind the
    // words in b_bits.  No pageable copies, one synchronisation per call.
    const size_t bbytes = sizeof(double) * (size_t)num_iter, ebytes = sizeof(double) * num_tries,
                 wbytes = sizeof(uint32_t) * (size_t)num_tries * W,
                 wpad = (wbytes + 255) & ~(size_t)255;
    const bool staged = qbytes + bbytes + wpad + ebytes <= Ctx::kStageMax;
    char *h = staged ? static_cast<char *>(c.staging(qbytes + bbytes + wpad + ebytes)) : nullptr;
    const double *beta_dev = c.b_beta.as<double>();
    double *e_dev = c.b_energy.as<double>();
    if (staged) {
      c.b_q.ensure(qbytes + bbytes);
      c.b_bits.ensure(wpad + ebytes);
      beta_dev = reinterpret_cast<const double *>(c.b_q.as<char>() + qbytes);
      e_dev = reinterpret_cast<double *>(c.b_bits.as<char>() + wpad);
      std::memcpy(h, q_sym, qbytes);
      if (num_iter > 0) std::memcpy(h + qbytes, beta, bbytes);
      OS_HIP_CHECK(hipMemcpyAsync(c.b_q.ptr, h, qbytes + bbytes, hipMemcpyHostToDevice,
                                  c.stream));
    } else {
      OS_HIP_CHECK(hipMemcpyAsync(c.b_q.ptr, q_sym, qbytes, hipMemcpyHostToDevice, c.stream));
      if (num_iter > 0)
        OS_HIP_CHECK(hipMemcpyAsync(c.b_beta.ptr, beta, bbytes, hipMemcpyHostToDevice,
                                    c.stream));
    }
    os::launch_ref_anneal(c.stream, n, c.b_q.as<double>(), beta_dev, num_iter, sweeps_per_beta,
                          num_tries, seed,
                          n > os::ref_lds_limit() ? c.b_fields.as<double>() : nullptr,
                          c.b_bits.as<uint32_t>(), c.b_run.as<double>());
    os::launch_energy_fp64(c.stream, n, W, c.b_q.as<double>(), (int)num_tries,
                           c.b_bits.as<uint32_t>(), e_dev, c.b_scratch.as<double>(), nullptr, 0,
                           nullptr);
    std::vector<double> ev(staged ? 0 : num_tries);
    double *e = staged ? reinterpret_cast<double *>(h + qbytes + bbytes + wpad) : ev.data();
    const uint32_t *words = staged ? reinterpret_cast<const uint32_t *>(h + qbytes + bbytes)
                                   : nullptr;
    if (staged)
      OS_HIP_CHECK(hipMemcpyAsync(h + qbytes + bbytes, c.b_bits.ptr, wpad + ebytes,
                                  hipMemcpyDeviceToHost, c.stream));
    else
      OS_HIP_CHECK(hipMemcpyAsync(e, e_dev, ebytes, hipMemcpyDeviceToHost, c.stream));
    if (best_states) {
      c.b_bytes.ensure((size_t)num_tries * n);
      os::launch_unpack_bits(c.stream, c.b_bits.as<uint32_t>(), n, num_tries, W,
                             c.b_bytes.as<uint8_t>());
      OS_HIP_CHECK(hipMemcpyAsync(best_states, c.b_bytes.ptr, (size_t)num_tries * n,
                                  hipMemcpyDeviceToHost, c.stream));
    }
    c.sync();
    // argmin over replicas, first index on ties (annealing.hpp:134-139)
    const size_t best_idx = (size_t)(std::min_element(e, e + num_tries) - e);
    if (best_energies) std::memcpy(best_energies, e, ebytes);
    if (best_energy) *best_energy = e[best_idx];
    if (best_state) {
      std::vector<uint32_t> w(W);
      if (staged)
        std::memcpy(w.data(), words + best_idx * W, sizeof(uint32_t) * W);
      else
        OS_HIP_CHECK(hipMemcpy(w.data(), c.b_bits.as<uint32_t>() + best_idx * W,
                               sizeof(uint32_t) * W, hipMemcpyDeviceToHost));
      for (int j = 0; j < n; ++j) best_state[j] = (uint8_t)((w[j >> 5] >> (j & 31)) & 1u);
    }
  });
}

}  // extern "C"
