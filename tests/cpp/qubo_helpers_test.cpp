// Flatten case of oneSolver tests/qubo_helpers_test.cpp:12-30.
#include "harness.hpp"
#include "onesolver/helpers/qubo_helpers.hpp"
#include "onesolver/helpers/ulong_to_vec.hpp"
#include "onesolver/simulated_annealing/sweep.hpp"

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

// The sparse path's rows (sa::detail::model_csr) hold exactly the off-diagonal
// entries of helpers::flatten_qubo, (i, j) and (j, i) keys summed, keys out of
// range and i == j quadratic keys ignored, columns ascending.
TEST_CASE(qubo_helpers_test, sweep_csr_equals_flattened_matrix) {
  qubo::LinearCoef<int, double> linear_c{{0, 0.5}, {2, 1.0}, {4, -1.5}, {9, 3.0}};
  qubo::QuadraticCoef<int, double> quadratic_c{
      {std::pair<int, int>(0, 1), 1.0},  {std::pair<int, int>(1, 0), 0.25},
      {std::pair<int, int>(0, 3), 7.2},  {std::pair<int, int>(4, 1), -1.0},
      {std::pair<int, int>(2, 2), 9.0},  {std::pair<int, int>(3, 4), -3.5},
      {std::pair<int, int>(2, 7), 11.0}};
  qubo::QUBOModel<int, double> instance(linear_c, quadratic_c);
  instance.set_nodes(5);
  const auto flat = helpers::flatten_qubo(instance);
  std::vector<int32_t> rp, col;
  std::vector<double> val, diag;
  sa::detail::model_csr(instance, 5, rp, col, val, diag);
  REQUIRE(rp.size() == 6u && rp[0] == 0 && (size_t)rp[5] == col.size());
  std::vector<double> dense(25, 0.0);
  for (int i = 0; i < 5; ++i) {
    REQUIRE(diag[i] == flat[i * 6]);
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      if (k > rp[i]) REQUIRE(col[k] > col[k - 1]);
      REQUIRE(col[k] != i);
      dense[i * 5 + col[k]] = val[k];
    }
  }
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j)
      if (i != j) REQUIRE(dense[i * 5 + j] == flat[i * 5 + j]);
}

TEST_CASE(qubo_helpers_test, sweep_mode_needs_the_gpu_device) {
  qubo::LinearCoef<int, double> linear_c{{0, 1.0}};
  qubo::QuadraticCoef<int, double> quadratic_c{{std::pair<int, int>(0, 1), -2.0}};
  qubo::QUBOModel<int, double> instance(linear_c, quadratic_c);
  instance.set_nodes(2);
  devices::Device host(*devices::construct_device_selector("host"));
  std::vector<double> beta{0.5, 1.0};
  CHECK_THROW(sa::anneal_sweeps(instance, host, beta, 2, 8u), std::invalid_argument);
}
