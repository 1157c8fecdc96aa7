// Model API cases of oneSolver tests/qubo_test.cpp:14-55.
#include "harness.hpp"
#include "onesolver/model/qubo.hpp"

TEST_CASE(qubo_test, qubo_linear_variable) {
  qubo::LinearCoef<int, double> linear_c{{1, -0.4}, {2, 1.1}, {3, -1.0}, {4, -1.2}};
  qubo::QuadraticCoef<int, double> quadratic_c{
      {std::make_pair(1, 2), 1.0}, {std::make_pair(1, 3), 7.2}, {std::make_pair(1, 4), 1.0},
      {std::make_pair(2, 3), 2.0}, {std::make_pair(2, 4), 1.9}, {std::make_pair(3, 4), 3.0}};
  qubo::QUBOModel<int, double> qubos(linear_c, quadratic_c);
  qubos.add_variable(1, 20);
  REQUIRE(qubos.get_variable(1) == 20);
}

TEST_CASE(qubo_test, qubo_quadratic_variable) {
  qubo::LinearCoef<int, double> linear_c{};
  qubo::QuadraticCoef<int, double> quadratic_c{};
  qubo::QUBOModel<int, double> qubos(linear_c, quadratic_c);
  qubos.add_connection(std::make_pair(1, 2), 20);
  REQUIRE(qubos.get_connection(std::make_pair(1, 2)) == 20);
}

TEST_CASE(qubo_test, qubo_print) {
  qubo::LinearCoef<int, int> linear_c{};
  qubo::QuadraticCoef<int, int> quadratic_c{};
  qubo::QUBOModel<int, int> qubos(linear_c, quadratic_c);
  qubos.add_variable(1, 10);
  qubos.add_connection(std::make_pair(1, 2), 20);
  REQUIRE(qubos.str() == "QUBO model 1--1:10 1--2:20");
}

TEST_CASE(qubo_test, absent_coefficients_are_zero) {
  qubo::LinearCoef<int, double> l{};
  qubo::QuadraticCoef<int, double> q{};
  qubo::QUBOModel<int, double> m(l, q);
  REQUIRE(m.get_variable(7) == 0.0);
  REQUIRE(m.get_connection(std::make_pair(0, 1)) == 0.0);
  m.add_connection(std::make_pair(0, 1), 3.0);
  m.add_connection(std::make_pair(0, 1), 4.0);  // overwrite (insert_model)
  REQUIRE(m.get_connection(std::make_pair(0, 1)) == 4.0);
  REQUIRE(m.get_connection(std::make_pair(1, 0)) == 0.0);  // keys are ordered
}
