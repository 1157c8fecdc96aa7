/*! file ulong_to_vec.hpp
 *
 * State index -> vector of N bits, bit i = variable i (oneSolver
 * include/helpers/ulong_to_vec.hpp:23-32; throws for more than 32 bits).
 */
#ifndef ONESOLVER_ULONG_TO_VEC_HPP_
#define ONESOLVER_ULONG_TO_VEC_HPP_

#include <stdexcept>
#include <vector>

namespace helpers {

inline std::vector<char> ulong_to_vec(unsigned long val, unsigned n_bits) {
  if (n_bits > 32) throw std::invalid_argument("state can be up to 32 bit");
  std::vector<char> state(n_bits);
  for (unsigned i = 0; i < n_bits; ++i) state[i] = (char)((val >> i) & 1ul);
  return state;
}

}  // namespace helpers

#endif
