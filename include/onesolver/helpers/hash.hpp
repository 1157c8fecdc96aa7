/*! file hash.hpp
 *
 * Hash functor for std::pair keys of QuadraticCoef (same role as oneSolver's
 * include/helpers/hash.hpp:19-35).  Packs both 32-bit halves into one 64-bit
 * word and mixes it (splitmix64 finaliser), so (i, j) and (j, i) do not
 * collide.
 */
#ifndef ONESOLVER_HASH_HPP_
#define ONESOLVER_HASH_HPP_

#include <cstddef>
#include <cstdint>
#include <functional>
#include <utility>

namespace helpers {

struct hash_pair {
  template <class T1, class T2>
  size_t operator()(const std::pair<T1, T2> &p) const {
    uint64_t a = (uint64_t)std::hash<T1>{}(p.first);
    uint64_t b = (uint64_t)std::hash<T2>{}(p.second);
    uint64_t z = (a << 32) ^ (b & 0xffffffffull) ^ (b >> 32);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return (size_t)(z ^ (z >> 31));
  }
};

}  // namespace helpers

#endif
