/*! file sweep.hpp
 *
 * sa::anneal_sweeps: the MI355X library's sweep-mode annealers behind the
 * argument list of oneSolver's sa::anneal (include/simulated_annealing/
 * annealing.hpp:55-58: model by value, device, beta schedule, num_iter,
 * num_tries, sweeps_per_beta).  The reference has no sweep mode; this is the
 * opt-in fast path, and sa::anneal keeps the reference's semantics.
 *
 * Semantics: each replica makes num_iter * sweeps_per_beta sweeps; a sweep
 * visits the sites in the order 0..N-1 and site j flips with probability
 * min(1, exp(-dE / T)), where T = beta[t / sweeps_per_beta] when beta is a
 * temperature (the reference's use, annealing.hpp:107, and the default) or
 * T = 1 / beta[...] with BetaKind::inverse.  The best state of each replica
 * is tracked at sweep ends; the returned Solution is the lowest FP64 energy
 * over the replicas, first replica on ties (annealing.hpp:134-139).
 *
 * Paths (SweepOptions::path):
 *   - dense (os_dense_*): bf16 hi+lo MFMA field updates, fields in registers
 *     for N <= 1024 and in HBM up to N = 16384;
 *   - sparse (os_csr_batch_*): CSR gathers, for any N;
 *   - automatic: dense when N <= 16384 and at least 1/8 of the couplings are
 *     set, or when a warm start is given; sparse otherwise.
 * Warm start: SweepOptions::warm_params holds the parameters of a trained
 * warm-start network (os_mpnn_num_params() floats, the packed order of
 * include/onesolver.h); its per-spin marginals seed the replicas on the dense
 * path.
 *
 * Errors: invalid arguments throw std::invalid_argument (host or cpu device
 * included: sweep mode runs on the gpu device); device failures print the
 * reference's kernel-failure message and throw std::runtime_error.
 */
#ifndef ONESOLVER_SWEEP_HPP_
#define ONESOLVER_SWEEP_HPP_

#include <algorithm>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../onesolver.h"
#include "../helpers/devices.hpp"
#include "../helpers/qubo_helpers.hpp"
#include "../model/qubo.hpp"
#include "../model/solution.hpp"
#include "annealing.hpp"

namespace sa {

enum class BetaKind { temperature, inverse };
enum class SweepPath { automatic, dense, sparse };

struct SweepOptions {
  BetaKind beta_kind = BetaKind::temperature;
  SweepPath path = SweepPath::automatic;
  uint64_t seed = 1234;
  int refresh_interval = 0;        // sweeps between FP64 field refreshes (dense); 0: at the end
  std::vector<float> warm_params;  // trained warm-start network; empty: random start states
};

/*! What a sweep-mode run did (optional output of anneal_sweeps). */
struct SweepReport {
  std::string path;              // "dense" or "sparse"
  double kernel_ms = 0.0;        // summed sweep-kernel time
  double max_field_drift = 0.0;  // dense: max |h_fp32 - h_fp64| over the refreshes
  uint64_t attempts = 0;         // replicas x sweeps x N
};

/*! Parameters of a warm-start network from a file of os_mpnn_num_params()
 *  little-endian fp32 values. */
inline std::vector<float> load_warm_params(const std::string &file) {
  std::ifstream f(file, std::ios::binary);
  if (!f) throw std::invalid_argument("can not open warm-start file: " + file);
  const int count = os_mpnn_num_params();
  std::vector<float> p((size_t)count);
  f.read(reinterpret_cast<char *>(p.data()), (std::streamsize)(sizeof(float) * p.size()));
  if (f.gcount() != (std::streamsize)(sizeof(float) * p.size()) || f.peek() != EOF)
    throw std::invalid_argument("warm-start file " + file + " does not hold " +
                                std::to_string(count) + " fp32 parameters");
  return p;
}

namespace detail {

inline void check_os(int rc) {
  if (rc != OS_OK) throw_os(rc);
}

// Symmetric off-diagonal couplings of the flattened matrix (both (i, j) and
// (j, i), duplicates summed as helpers::flatten_qubo adds them) as CSR rows
// sorted by column, and the diagonal.
template <typename T>
void model_csr(const qubo::QUBOModel<int, T> &m, int n, std::vector<int32_t> &rp,
               std::vector<int32_t> &col, std::vector<double> &val, std::vector<double> &diag) {
  diag.assign((size_t)n, 0.0);
  for (const auto &kv : m.linear_terms())
    if (kv.first >= 0 && kv.first < n) diag[(size_t)kv.first] = (double)kv.second;
  std::vector<std::pair<int64_t, double>> e;
  e.reserve(2 * m.quadratic_terms().size());
  for (const auto &kv : m.quadratic_terms()) {
    const int64_t i = kv.first.first, j = kv.first.second;
    if (i == j || i < 0 || j < 0 || i >= n || j >= n) continue;
    e.push_back({i * n + j, (double)kv.second});
    e.push_back({j * n + i, (double)kv.second});
  }
  std::sort(e.begin(), e.end(), [](const std::pair<int64_t, double> &a,
                                   const std::pair<int64_t, double> &b) { return a.first < b.first; });
  rp.assign((size_t)n + 1, 0);
  col.clear();
  val.clear();
  for (size_t k = 0; k < e.size(); ++k) {
    if (!col.empty() && k > 0 && e[k].first == e[k - 1].first) {
      val.back() += e[k].second;  // (i, j) and (j, i) both stored
      continue;
    }
    col.push_back((int32_t)(e[k].first % n));
    val.push_back(e[k].second);
    ++rp[(size_t)(e[k].first / n) + 1];
  }
  for (int i = 0; i < n; ++i) rp[(size_t)i + 1] += rp[(size_t)i];
}

}  // namespace detail

/*! Sweep-mode annealing of `instance` on a gpu `device`. */
template <typename T>
qubo::Solution anneal_sweeps(qubo::QUBOModel<int, T> instance, devices::Device &device,
                             std::vector<double> &h_beta_schedule, int num_iter,
                             unsigned int num_tries, int sweeps_per_beta = 1,
                             const SweepOptions &opt = SweepOptions(),
                             SweepReport *report = nullptr) {
  const int N = (int)instance.get_nodes();
  if (N < 1) throw std::invalid_argument("empty model");
  if (!device.is_gpu()) throw std::invalid_argument("sweep mode runs on the gpu device");
  if (num_iter < 0 || (size_t)num_iter > h_beta_schedule.size())
    throw std::invalid_argument("beta schedule shorter than num_iter");
  if (num_tries < 1) throw std::invalid_argument("num_tries must be positive");
  if (sweeps_per_beta < 1) throw std::invalid_argument("sweeps_per_beta must be positive");
  const size_t sweeps = (size_t)num_iter * (size_t)sweeps_per_beta;
  if (sweeps > (size_t)INT32_MAX) throw std::invalid_argument("too many sweeps");
  std::vector<float> ts(sweeps);
  for (size_t t = 0; t < sweeps; ++t) {
    const double b = h_beta_schedule[t / (size_t)sweeps_per_beta];
    ts[t] = (float)(opt.beta_kind == BetaKind::temperature ? b : 1.0 / b);
  }
  const bool warm = !opt.warm_params.empty();
  if (warm && opt.path == SweepPath::sparse)
    throw std::invalid_argument("the warm start runs on the dense path");
  if (warm && opt.warm_params.size() != (size_t)os_mpnn_num_params())
    throw std::invalid_argument("warm_params must hold os_mpnn_num_params() values");
  size_t couplings = 0;
  for (const auto &kv : instance.quadratic_terms())
    if (kv.first.first != kv.first.second) ++couplings;
  bool dense = opt.path == SweepPath::dense;
  if (opt.path == SweepPath::automatic)
    dense = warm || (N <= 16384 && 16.0 * (double)couplings >= (double)N * (double)(N - 1));
  if (report) report->path = dense ? "dense" : "sparse";

  os_ctx *ctx = device.ctx();
  std::vector<uint8_t> state((size_t)N);
  double best_energy = 0.0;
  os_sweep_stats st{};
  if (dense) {
    const std::vector<T> flat_t = helpers::flatten_qubo(instance);
    const std::vector<double> flat(flat_t.begin(), flat_t.end());
    os_dense *d = nullptr;
    detail::check_os(os_dense_create(ctx, N, flat.data(), &d));
    std::vector<float> marg;
    int rc = OS_OK;
    if (warm) {
      os_mpnn *net = nullptr;
      rc = os_mpnn_create(ctx, 0, &net);
      if (rc == OS_OK) rc = os_mpnn_set_params(net, opt.warm_params.data());
      if (rc == OS_OK) {
        marg.resize((size_t)N);
        rc = os_mpnn_predict(net, d, 0, 1, marg.data());
      }
      if (net) os_mpnn_destroy(net);
    }
    if (rc == OS_OK) {
      os_sweep_params p{};
      p.num_replicas = num_tries;
      p.num_sweeps = (int32_t)sweeps;
      p.tscale = ts.data();
      p.seed = opt.seed;
      p.refresh_interval = opt.refresh_interval;
      p.init_marginals = warm ? marg.data() : nullptr;
      rc = os_dense_anneal(d, &p, nullptr, nullptr, nullptr, state.data(), &best_energy, &st);
    }
    os_dense_destroy(d);
    detail::check_os(rc);
  } else {
    std::vector<int32_t> rp, col;
    std::vector<double> val, diag;
    detail::model_csr(instance, N, rp, col, val, diag);
    const int32_t n32 = N;
    os_csr_batch *b = nullptr;
    detail::check_os(
        os_csr_batch_create(ctx, 1, &n32, rp.data(), col.data(), val.data(), diag.data(), &b));
    os_csr_params p{};
    p.num_replicas = num_tries;
    p.num_sweeps = (int32_t)sweeps;
    p.tscale = ts.data();
    p.seed = opt.seed;
    p.use_graph = 1;
    p.tune = 0;
    std::vector<uint8_t> states((size_t)num_tries * N);
    int32_t best_r = 0;
    const int rc = os_csr_batch_anneal(b, &p, nullptr, nullptr, states.data(), nullptr,
                                       &best_energy, &best_r, &st);
    os_csr_batch_destroy(b);
    detail::check_os(rc);
    std::copy(states.begin() + (size_t)best_r * N, states.begin() + (size_t)(best_r + 1) * N,
              state.begin());
  }
  if (report) {
    report->kernel_ms = st.kernel_ms;
    report->max_field_drift = st.max_field_drift;
    report->attempts = st.attempts;
  }
  std::vector<char> x(state.begin(), state.end());
  return qubo::Solution(x.begin(), x.end(), best_energy);
}

}  // namespace sa

#endif
