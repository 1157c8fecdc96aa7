// Flatten case of oneSolver tests/qubo_helpers_test.cpp:12-30.
#include "harness.hpp"
#include "onesolver/helpers/qubo_helpers.hpp"
#include "onesolver/helpers/ulong_to_vec.hpp"

TEST_CASE(qubo_helpers_test, qubo_flattening) {
  qubo::LinearCoef<int, double> linear_c{{0, 0.5}, {1, -2.0}, {2, 1.0}, {4, -1.5}};
  qubo::QuadraticCoef<int, double> quadratic_c{
      {std::pair<int, int>(0, 1), 1.0}, {std::pair<int, int>(0, 3), 7.2},
      {std::pair<int, int>(1, 4), -1.0}, {std::pair<int, int>(2, 3), 2.0},
      {std::pair<int, int>(2, 4), -1.5}, {std::pair<int, int>(3, 4), -3.5}};
  qubo::QUBOModel<int, double> instance(linear_c, quadratic_c);
  instance.set_nodes(5);
  auto flat_qubo = helpers::flatten_qubo(instance);
  std::vector<double> expected_result({0.5, 1.0,  0.0, 7.2, 0.0,  1.0,  -2.0, 0.0, 0.0,
                                       -1.0, 0.0, 0.0, 1.0, 2.0, -1.5, 7.2,  0.0,  2.0,
                                       0.0, -3.5, 0.0, -1.0, -1.5, -3.5, -1.5});
  REQUIRE(flat_qubo == expected_result);
}

TEST_CASE(qubo_helpers_test, ulong_to_vec_lsb_is_variable_zero) {
  auto v = helpers::ulong_to_vec(0b1101ul, 5);
  REQUIRE(v == std::vector<char>({1, 0, 1, 1, 0}));
  CHECK_THROW(helpers::ulong_to_vec(0, 33), std::invalid_argument);
}
