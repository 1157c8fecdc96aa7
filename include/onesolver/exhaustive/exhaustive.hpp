/*! file exhaustive.hpp
 *
 * exhaustive::solve of the MI355X library: same semantics as oneSolver's
 * include/exhaustive/exhaustive.hpp:29-167, with the sycl::queue replaced by a
 * devices::Device.
 *
 * Semantics kept from the reference:
 *   - the coefficients form an upper-triangular buffer: get_variable(i) on the
 *     diagonal, get_connection({i, j}) for j > i (:44-61; a connection stored
 *     under the key {j, i} with j > i is not read);
 *   - E(x) is summed over i ascending, then j >= i ascending, in double
 *     (:113-128), bit i of the state index being variable i;
 *   - of equal minima the lowest state index wins (each work-item keeps its
 *     first minimum, :129-132, and the first minimal work-item is taken,
 *     :158-160);
 *   - the state is converted by helpers::ulong_to_vec, which throws
 *     std::invalid_argument for more than 32 variables.
 * Difference: the reference counts states with an int shift, valid up to 30
 * variables (:67); the enumeration here is 64-bit, so 31 and 32 work, and the
 * 32-variable limit is checked before enumerating.
 * "gpu" runs os_exhaustive; "host" and "cpu" run the literal per-state loop
 * over contiguous index ranges on the device's threads, as the reference's
 * work-items do.  Device failures print the reference's message to stderr and
 * throw std::runtime_error.
 */
#ifndef ONESOLVER_EXHAUSTIVE_HPP_
#define ONESOLVER_EXHAUSTIVE_HPP_

#include <cfloat>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <thread>
#include <utility>
#include <vector>

#include "../../onesolver.h"
#include "../helpers/devices.hpp"
#include "../helpers/ulong_to_vec.hpp"
#include "../model/qubo.hpp"
#include "../model/solution.hpp"

namespace exhaustive {

namespace detail {

// Lowest energy (reference order) and its first index over states [s0, s1).
inline void literal_range(int n, const std::vector<double> &q, uint64_t s0, uint64_t s1,
                          double *e_best, uint64_t *s_best) {
  double eb = DBL_MAX;
  uint64_t sb = 0;
  for (uint64_t s = s0; s < s1; ++s) {
    double e = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j)
        if (((s >> i) & 1u) && ((s >> j) & 1u)) e += q[(size_t)i * n + j];
    if (e < eb) {
      eb = e;
      sb = s;
    }
  }
  *e_best = eb;
  *s_best = sb;
}

}  // namespace detail

/*! The exact minimum of `qubos` over all 2^N states. */
template <class NodeType, class CoefType>
qubo::Solution solve(devices::Device &device, qubo::QUBOModel<NodeType, CoefType> &qubos) {
  const int n = (int)qubos.get_nodes();
  if (n > 32) throw std::invalid_argument("state can be up to 32 bit");
  std::vector<double> q((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j)
      q[(size_t)i * n + j] =
          i == j ? qubos.get_variable(i) : qubos.get_connection(std::make_pair(i, j));
  uint64_t best_state = 0;
  double best_energy = 0.0;
  if (n == 0) {
    // one state, the empty one, of energy 0
  } else if (device.is_gpu()) {
    std::vector<uint8_t> st((size_t)n);
    const int rc = os_exhaustive(device.ctx(), n, q.data(), st.data(), &best_energy);
    if (rc != OS_OK) {
      if (rc == OS_ERR_INVALID) throw std::invalid_argument(os_last_error());
      std::cerr << "Execution of kernel failed. See below exception info for details";
      throw std::runtime_error(os_last_error());
    }
    for (int i = 0; i < n; ++i) best_state |= (uint64_t)(st[i] & 1u) << i;
  } else {
    const uint64_t n_states = 1ull << n;
    const uint64_t units = std::min<uint64_t>(device.threads(), n_states);
    std::vector<double> energies(units);
    std::vector<uint64_t> states(units);
    std::vector<std::thread> pool;
    uint64_t start = 0;
    for (uint64_t u = 0; u < units; ++u) {
      // contiguous ranges, the first n_states % units one state longer
      const uint64_t len = n_states / units + (u < n_states % units ? 1 : 0);
      pool.emplace_back(detail::literal_range, n, std::cref(q), start, start + len,
                        &energies[u], &states[u]);
      start += len;
    }
    for (auto &t : pool) t.join();
    size_t k = 0;
    for (size_t u = 1; u < units; ++u)
      if (energies[u] < energies[k]) k = u;
    best_energy = energies[k];
    best_state = states[k];
  }
  auto result = helpers::ulong_to_vec((unsigned long)best_state, (unsigned)n);
  return qubo::Solution(result.begin(), result.end(), best_energy);
}

}  // namespace exhaustive

#endif
