/*! file schedules.hpp
 *
 * Beta schedules of the one-solver-anneal CLI (oneSolver
 * app/one-solver-anneal.cpp:23-39):
 *   linear     beta_i = beta_min + beta_max * i / (n-1)   (ends at beta_min+beta_max;
 *              n == 1 gives NaN)
 *   geometric  beta_0 = beta_min, beta_i = beta_{i-1} * alpha,
 *              alpha = (beta_max/beta_min)^(1/(n-1))  (cumulative product)
 * `schedule` must hold num_iter elements; num_iter == 0 is rejected (the
 * reference writes schedule[0] out of bounds).
 */
#ifndef ONESOLVER_SCHEDULES_HPP_
#define ONESOLVER_SCHEDULES_HPP_

#include <cmath>
#include <stdexcept>
#include <vector>

inline void construct_linear_beta_schedule(std::vector<double> &schedule, double beta_min,
                                           double beta_max, unsigned num_iter) {
  if (schedule.size() < num_iter) throw std::invalid_argument("schedule is too short");
  const double denom = static_cast<double>(num_iter - 1u);
  for (unsigned i = 0; i < num_iter; i++) schedule[i] = beta_min + beta_max * i / denom;
}

inline void construct_geometric_beta_schedule(std::vector<double> &schedule, double beta_min,
                                              double beta_max, unsigned num_iter) {
  if (num_iter == 0) throw std::invalid_argument("num_iter must be at least 1");
  if (schedule.size() < num_iter) throw std::invalid_argument("schedule is too short");
  schedule[0] = beta_min;
  const double alpha = std::pow(beta_max / beta_min, 1.0 / static_cast<double>(num_iter - 1u));
  for (unsigned i = 1; i < num_iter; i++) schedule[i] = schedule[i - 1] * alpha;
}

#endif
