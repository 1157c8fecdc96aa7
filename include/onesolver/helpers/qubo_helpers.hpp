/*! file qubo_helpers.hpp
 *
 * Model -> dense symmetric N x N matrix (oneSolver
 * include/helpers/qubo_helpers.hpp:26-44): entry (i, i) is the linear
 * coefficient of i, and each stored coupling (i, j), i != j, is added to both
 * (i, j) and (j, i).  Keys outside [0, N) and quadratic keys with i == j are
 * never looked up by the reference loop and are ignored here too.  Walks the
 * maps instead of doing N^2 lookups; the result is the same matrix.
 */
#ifndef ONESOLVER_QUBO_HELPERS_HPP_
#define ONESOLVER_QUBO_HELPERS_HPP_

#include <vector>

#include "../model/qubo.hpp"

namespace helpers {

template <class CoeffType>
std::vector<CoeffType> flatten_qubo(qubo::QUBOModel<int, CoeffType> instance) {
  const long n = (long)instance.get_nodes();
  std::vector<CoeffType> result((size_t)(n * n), CoeffType(0));
  for (const auto &kv : instance.linear_terms()) {
    if (kv.first >= 0 && kv.first < n) result[(size_t)kv.first * (n + 1)] = kv.second;
  }
  for (const auto &kv : instance.quadratic_terms()) {
    const long i = kv.first.first, j = kv.first.second;
    if (i == j || i < 0 || j < 0 || i >= n || j >= n) continue;
    result[(size_t)(i + j * n)] += kv.second;
    result[(size_t)(j + i * n)] += kv.second;
  }
  return result;
}

}  // namespace helpers

#endif
