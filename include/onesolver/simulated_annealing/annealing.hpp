/*! file annealing.hpp
 *
 * sa::energy and sa::anneal of the MI355X annealing library: same signatures
 * and semantics as oneSolver's include/simulated_annealing/annealing.hpp:31-140,
 * with the sycl::queue replaced by a devices::Device.
 *
 * Semantics kept from the reference:
 *   - num_tries independent replicas, random {0,1} start (:90-95);
 *   - num_iter * sweeps_per_beta single-site attempts per replica, each at a
 *     uniformly random site, with beta[iter] (:97-101);
 *   - accept iff E_new < E_cur || exp((E_cur - E_new)/beta) > u (:106-113);
 *   - best state tracked after every attempt (:115-121);
 *   - the returned Solution is the lowest-energy replica, first index on
 *     ties (:134-139).
 * Differences: ΔE comes from O(N) local fields instead of an O(N^2)
 * recompute; random numbers are Philox4x32-10 draws with a counter per
 * (replica, attempt) keyed by `seed` (the reference's engine is seeded 1234,
 * the default here); the returned energy is an FP64 recompute of the
 * returned state.  Kernel failures print the reference's message to stderr
 * and throw std::runtime_error; invalid arguments throw std::invalid_argument.
 */
#ifndef ONESOLVER_ANNEALING_HPP_
#define ONESOLVER_ANNEALING_HPP_

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <thread>
#include <vector>

#include "../../onesolver.h"
#include "../helpers/devices.hpp"
#include "../helpers/qubo_helpers.hpp"
#include "../model/qubo.hpp"
#include "../model/solution.hpp"
#include "philox.hpp"

namespace sa {

/*! E(x) = sum_{i<=j} F[i*N+j] x_i x_j over the flattened (symmetric) QUBO,
 *  i ascending then j ascending, in double (annealing.hpp:31-40). */
template <typename QuboArray, typename StateArray>
double energy(const QuboArray flat_qubo, StateArray state, int N) {
  double result = 0.0;
  for (int i = 0; i < N; i++)
    for (int j = i; j < N; j++) result += flat_qubo[i * N + j] * state[i] * state[j];
  return result;
}

namespace detail {

inline void throw_os(int rc) {
  if (rc == OS_ERR_INVALID) throw std::invalid_argument(os_last_error());
  std::cerr << "Execution of kernel failed. See below exception info for details";
  throw std::runtime_error(os_last_error());
}

// One replica of the reference chain on the host (local-field form).
inline void cpu_replica(int n, const double *q, const double *beta, uint64_t total, int spb,
                        uint32_t replica, uint64_t seed, char *best, double *best_e) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  std::vector<char> x((size_t)n);
  std::vector<double> h((size_t)n);
  uint32_t w[4];
  for (int j = 0; j < n; ++j) {
    if ((j & 127) == 0) philox::draw((uint32_t)(j >> 7), replica, 0u, philox::TAG_REF_INIT, k0, k1, w);
    x[j] = (char)((w[(j >> 5) & 3] >> (j & 31)) & 1u);
    h[j] = q[(size_t)j * n + j];
  }
  for (int l = 0; l < n; ++l)
    if (x[l])
      for (int j = 0; j < n; ++j)
        if (j != l) h[j] += q[(size_t)l * n + j];
  double e = 0.0;
  for (int j = 0; j < n; ++j)
    if (x[j]) e += 0.5 * (h[j] + q[(size_t)j * n + j]);
  double e_best = e;
  std::copy(x.begin(), x.end(), best);
  for (uint64_t t = 0; t < total; ++t) {
    const double b = beta[t / (uint64_t)spb];
    philox::draw((uint32_t)t, replica, (uint32_t)(t >> 32), philox::TAG_REF_STEP, k0, k1, w);
    const int s = (int)(((uint64_t)w[0] * (uint64_t)(uint32_t)n) >> 32);
    const double u = (double)w[1] * 0x1p-32;
    const double de = x[s] ? -h[s] : h[s];
    if ((de < 0.0) || (std::exp(-de / b) > u)) {
      const double d = x[s] ? -1.0 : 1.0;
      const double *row = q + (size_t)s * n;
      for (int j = 0; j < n; ++j)
        if (j != s) h[j] = std::fma(row[j], d, h[j]);
      x[s] = (char)(1 - x[s]);
      e += de;
    }
    if (e < e_best) {
      e_best = e;
      std::copy(x.begin(), x.end(), best);
    }
  }
  *best_e = e_best;
}

}  // namespace detail

/*! Solve `instance` by simulated annealing on `device`. */
template <typename T>
qubo::Solution anneal(qubo::QUBOModel<int, T> instance, devices::Device &device,
                      std::vector<double> &h_beta_schedule, int num_iter, unsigned int num_tries,
                      int sweeps_per_beta = 1, uint64_t seed = 1234) {
  const int N = (int)instance.get_nodes();
  if (N < 1) throw std::invalid_argument("empty model");
  if (num_iter < 0 || (size_t)num_iter > h_beta_schedule.size())
    throw std::invalid_argument("beta schedule shorter than num_iter");
  if (num_tries < 1) throw std::invalid_argument("num_tries must be positive");
  if (sweeps_per_beta < 1) throw std::invalid_argument("sweeps_per_beta must be positive");
  std::vector<T> flat_t = helpers::flatten_qubo(instance);
  std::vector<double> flat(flat_t.begin(), flat_t.end());
  std::vector<char> state((size_t)N);
  double best_energy = 0.0;
  if (device.is_gpu()) {
    std::vector<uint8_t> st((size_t)N);
    const int rc = os_anneal_reference(device.ctx(), N, flat.data(), h_beta_schedule.data(),
                                       num_iter, sweeps_per_beta, num_tries, seed, nullptr,
                                       nullptr, st.data(), &best_energy);
    if (rc != OS_OK) detail::throw_os(rc);
    std::copy(st.begin(), st.end(), state.begin());
  } else {
    std::vector<char> states((size_t)num_tries * N);
    std::vector<double> energies(num_tries);
    const uint64_t total = (uint64_t)num_iter * (uint64_t)sweeps_per_beta;
    const unsigned nt = std::min<unsigned>(device.threads(), num_tries);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t) {
      pool.emplace_back([&, t] {
        for (unsigned r = t; r < num_tries; r += nt)
          detail::cpu_replica(N, flat.data(), h_beta_schedule.data(), total, sweeps_per_beta, r,
                              seed, states.data() + (size_t)r * N, &energies[r]);
      });
    }
    for (auto &th : pool) th.join();
    for (unsigned r = 0; r < num_tries; ++r)
      energies[r] = energy(flat.data(), states.data() + (size_t)r * N, N);
    const size_t bi = (size_t)(std::min_element(energies.begin(), energies.end()) - energies.begin());
    std::copy(states.begin() + bi * N, states.begin() + (bi + 1) * N, state.begin());
    best_energy = energies[bi];
  }
  return qubo::Solution(state.begin(), state.end(), best_energy);
}

}  // namespace sa

#endif
