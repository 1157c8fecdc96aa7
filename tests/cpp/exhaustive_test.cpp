// exhaustive::solve (include/exhaustive/exhaustive.hpp:29-167) on the host
// devices: the reference's known answers (tests/exhaustive_test.cpp:12-156,
// tolerance 1e-13), the lowest index among equal minima, the N=3 rule of the
// reader and the 32-variable limit of ulong_to_vec.
#include <unistd.h>

#include <cmath>
#include <fstream>
#include <sstream>

#include "harness.hpp"
#include "onesolver/exhaustive/exhaustive.hpp"

namespace {

qubo::Solution solve_model(qubo::QUBOModel<int, double> &m, const char *device = "cpu") {
  devices::Device d(*devices::construct_device_selector(device));
  return exhaustive::solve(d, m);
}

qubo::Solution solve_text(const std::string &text, const char *device = "cpu") {
  std::stringstream st(text);
  st.unsetf(std::ios::skipws);
  auto m = qubo::QUBOModel<int, double>::load(st);
  return solve_model(m, device);
}

// tests/golden/data/kat, found from this executable (onesolver_amd/bin)
std::string kat_dir() {
  char buf[4096];
  const ssize_t k = readlink("/proc/self/exe", buf, sizeof buf - 1);
  std::string p(buf, k > 0 ? (size_t)k : 0);
  p = p.substr(0, p.rfind('/'));
  return p + "/../../tests/golden/data/kat/";
}

}  // namespace

TEST_CASE(exhaustive_test, known_answers) {
  // the five files of tests/exhaustive_test.cpp:12-142 and their energies
  const double want[] = {-12.0, -1.2, -22.0, -14.0, -32.0};
  for (int k = 0; k < 5; ++k) {
    std::ifstream f(kat_dir() + "kat_" + std::to_string(k) + ".qubo");
    REQUIRE(f.good());
    std::stringstream text;
    text << f.rdbuf();
    for (const char *dev : {"host", "cpu"}) {
      auto s = solve_text(text.str(), dev);
      REQUIRE(std::abs(s.energy - want[k]) < 1e-13);
    }
  }
}

TEST_CASE(exhaustive_test, ground_state_and_ties) {
  // -x0 - x1 + 2 x0 x1: states 1 and 2 both -1; the lower index (x0 = 1) wins
  auto s = solve_text("p qubo 0 2 2 1\n0 0 -1\n1 1 -1\n0 1 2\n");
  REQUIRE(s.energy == -1.0);
  REQUIRE(s.state.size() == 2 && s.state[0] == 1 && s.state[1] == 0);
  // test1.qubo: ground state 1,1,0,1 of energy -12
  auto t = solve_text("p qubo 0 4 4 6\n0 0 -5\n0 1 2\n0 2 4\n0 3 0\n1 1 -3\n1 2 1\n1 3 0\n"
                      "2 2 -8\n2 3 5\n3 3 -6\n");
  REQUIRE(t.energy == -12.0);
  REQUIRE(t.state[0] == 1 && t.state[1] == 1 && t.state[2] == 0 && t.state[3] == 1);
}

TEST_CASE(exhaustive_test, connection_keys_below_the_diagonal_are_not_read) {
  // the reference's buffer holds get_connection({i, j}) for j > i only
  qubo::LinearCoef<int, double> lin;
  qubo::QuadraticCoef<int, double> quad;
  qubo::QUBOModel<int, double> m(lin, quad);
  m.add_variable(0, 1.0);
  m.add_variable(1, 1.0);
  m.add_connection(std::make_pair(1, 0), -5.0);
  m.set_nodes(2);
  auto s = solve_model(m);
  REQUIRE(s.energy == 0.0);
}

TEST_CASE(exhaustive_test, more_than_32_variables_are_rejected) {
  CHECK_THROW(solve_text("p qubo 0 33 1 0\n32 32 1\n"), std::invalid_argument);
}
