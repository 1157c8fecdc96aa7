// os_ctx.h -- the context behind the C ABI's os_ctx handle.
#pragma once

#include <string>

#include "os_common.h"

namespace os {

// Grow-only device buffer.
struct DevBuf {
  void *ptr = nullptr;
  size_t cap = 0;
  void ensure(size_t bytes);
  void release();
  template <typename T>
  T *as() const {
    return static_cast<T *>(ptr);
  }
};

struct Ctx {
  int device = 0;
  std::string name;
  std::string arch;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr;
  hipEvent_t ev_stop = nullptr;
  // scratch buffers of the one-shot entry points
  DevBuf b_q, b_beta, b_bits, b_run, b_energy, b_scratch, b_fields, b_bytes;
  // .qubo ingest work arrays (qubo_text.hip), kept between parses, and two
  // pinned staging buffers for host text (hipHostMalloc, kPinnedBytes each)
  DevBuf ingest[16];
  void *pinned[2] = {nullptr, nullptr};
  hipEvent_t pinned_ev[2] = {nullptr, nullptr};
  // pinned staging of the one-shot entry points' small transfers (up to
  // kStageMax bytes; idle whenever no call is running)
  static constexpr size_t kStageMax = 1u << 20;
  void *stage = nullptr;
  size_t stage_bytes = 0;
  void *staging(size_t bytes);
  hipEvent_t ingest_ev[4] = {nullptr, nullptr, nullptr, nullptr};  // parse phase timing

  explicit Ctx(int dev);
  ~Ctx();
  Ctx(const Ctx &) = delete;
  Ctx &operator=(const Ctx &) = delete;
  void activate() const;
  void sync() const;
};

// Thread-local message behind os_last_error().
void set_last_error(const char *msg);

// Runs `f`, mapping exceptions to the C ABI's return codes.
template <typename F>
int guarded(F &&f) {
  try {
    f();
    set_last_error("");
    return 0;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("out of host memory");
    return -4;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -4;
  }
}

}  // namespace os
