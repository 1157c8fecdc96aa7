/*! file philox.hpp
 *
 * Host Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11) and the stream tags
 * of the annealing library.  The GPU kernels draw from the same counters, so
 * host and device runs of one seed follow the same chains.  Counter word 3 is
 * (instance << 8) | tag.
 */
#ifndef ONESOLVER_PHILOX_HPP_
#define ONESOLVER_PHILOX_HPP_

#include <cstdint>

namespace philox {

enum : uint32_t {
  TAG_REF_INIT = 0x01,
  TAG_REF_STEP = 0x02,
  TAG_SWEEP_INIT = 0x10,
  TAG_SWEEP_THR = 0x11,
  TAG_WARM = 0x12,
};

inline void draw(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                 uint32_t out[4]) {
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

}  // namespace philox

#endif
