// Schedules (app/one-solver-anneal.cpp:23-39) and sa::energy
// (annealing.hpp:31-40); host-device sa::anneal invariants.
#include <cmath>
#include <sstream>

#include "harness.hpp"
#include "onesolver/simulated_annealing/annealing.hpp"
#include "onesolver/simulated_annealing/schedules.hpp"

TEST_CASE(schedules_test, linear_ends_at_min_plus_max) {
  std::vector<double> s(5);
  construct_linear_beta_schedule(s, 0.1, 1.0, 5);
  REQUIRE(s[0] == 0.1);
  REQUIRE(s[2] == 0.1 + 1.0 * 2 / 4.0);
  REQUIRE(s[4] == 0.1 + 1.0);
  std::vector<double> one(1);
  construct_linear_beta_schedule(one, 0.1, 1.0, 1);
  REQUIRE(std::isnan(one[0]));
}

TEST_CASE(schedules_test, geometric_is_a_cumulative_product) {
  std::vector<double> s(1000);
  construct_geometric_beta_schedule(s, 0.1, 10.0, 1000);
  const double alpha = std::pow(10.0 / 0.1, 1.0 / 999.0);
  double b = 0.1;
  for (int i = 0; i < 1000; ++i) {
    REQUIRE(s[i] == b);
    b *= alpha;
  }
  REQUIRE(std::fabs(s[999] - 10.0) < 1e-9);
  std::vector<double> z;
  CHECK_THROW(construct_geometric_beta_schedule(z, 0.1, 1.0, 0), std::invalid_argument);
}

TEST_CASE(annealing_test, energy_matches_hand_computation) {
  // test1.qubo of benchmarks/exhaustive_search/examples: ground -12 at 1,0,1,1... checked by value
  std::stringstream st("p qubo 0 4 4 6\n0 0 -5\n0 1 2\n0 2 4\n0 3 0\n1 1 -3\n1 2 1\n1 3 0\n"
                       "2 2 -8\n2 3 5\n3 3 -6\n");
  st.unsetf(std::ios::skipws);
  auto m = qubo::QUBOModel<int, double>::load(st);
  auto flat = helpers::flatten_qubo(m);
  const char x[] = {1, 0, 0, 1};
  REQUIRE(sa::energy(flat.data(), x, 4) == -5.0 - 6.0 + 0.0);
  const char y[] = {0, 1, 1, 0};
  REQUIRE(sa::energy(flat.data(), y, 4) == -3.0 - 8.0 + 1.0);
}

TEST_CASE(annealing_test, host_device_anneal_returns_a_consistent_solution) {
  std::stringstream st("p qubo 0 4 4 6\n0 0 -5\n0 1 2\n0 2 4\n0 3 0\n1 1 -3\n1 2 1\n1 3 0\n"
                       "2 2 -8\n2 3 5\n3 3 -6\n");
  st.unsetf(std::ios::skipws);
  auto m = qubo::QUBOModel<int, double>::load(st);
  devices::Device dev(devices::host_selector{});
  std::vector<double> beta(200);
  construct_geometric_beta_schedule(beta, 0.1, 1.0, 200);
  auto sol = sa::anneal(m, dev, beta, 200, 16);
  auto flat = helpers::flatten_qubo(m);
  REQUIRE(sol.state.size() == 4);
  REQUIRE(sol.energy == sa::energy(flat.data(), sol.state.data(), 4));
  REQUIRE(sol.energy == -12.0);  // exhaustive ground energy of this instance
  devices::Device cpu(devices::cpu_selector{});
  auto sol2 = sa::anneal(m, cpu, beta, 200, 16);
  REQUIRE(sol2.state == sol.state);  // same streams, any thread count
}
