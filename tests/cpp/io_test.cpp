// Reader and writer cases of oneSolver tests/io_test.cpp:12-118 plus the
// grammar edge cases of include/model/qubo.hpp:392-417.
#include <sstream>

#include "harness.hpp"
#include "onesolver/model/qubo.hpp"
#include "onesolver/model/solution.hpp"

namespace {
const std::string qubo_file_contents(
    "c qubo Target MaxNodes NumNodes NumLinks\n"
    "p qubo 0 100 2 2\n"
    "c comment\n"
    "0 0 -0.5\n"
    "0 1 2.0\n"
    "1 2 4\n"
    "2 2 -0.7\n");

const std::string grammatically_incorrect_file_contents[]{
    std::string("p qubo 0 1 100 3 492\n0 0 -0.5\n0 1 2.0\n"),
    std::string("p qubo 0 100 3 492\n0 0 1 -0.5\n0 1 2.0\n"),
    std::string("p qubo 0 100 3 492\n0 0 -0.5\n0 2.0\n"),
    std::string("p qubo 0 100 492\n0 0 -0.5\n0 1 2.0\n"),
    std::string("p qubo 0 100 3 492\n0 0 -0.5\nunexpected string\n0 1 2.0\n")};

qubo::QUBOModel<int, double> load(const std::string &text) {
  std::stringstream stream(text);
  stream.unsetf(std::ios::skipws);
  return qubo::QUBOModel<int, double>::load(stream);
}
}  // namespace

TEST_CASE(io_test, number_of_variable_agrees_with_given_in_input_stream) {
  REQUIRE(load(qubo_file_contents).get_nodes() == 3);
}

TEST_CASE(io_test, all_linear_coefficients_are_present_in_the_model) {
  auto p = load(qubo_file_contents);
  REQUIRE(p.get_variable(0) == -0.5);
  REQUIRE(p.get_variable(2) == -0.7);
}

TEST_CASE(io_test, all_quadratic_coefficients_are_present_in_the_model) {
  auto p = load(qubo_file_contents);
  REQUIRE(p.get_connection(std::pair<int, int>(0, 1)) == 2.0);
  REQUIRE(p.get_connection(std::pair<int, int>(1, 2)) == 4.0);
}

TEST_CASE(io_test, read_qubo_throws_if_interaction_is_placed_in_reversed_order) {
  CHECK_THROW(load("p qubo 0 100 2 2\n0 0 -0.5\n1 0 2.0\n1 2 4\n2 2 -0.7\n"),
              std::invalid_argument);
}

TEST_CASE(io_test, read_qubo_throws_if_stream_contents_dont_adhere_to_grammar) {
  for (const auto &sample : grammatically_incorrect_file_contents)
    CHECK_THROW(load(sample), std::invalid_argument);
}

TEST_CASE(io_test, number_of_quadratic_terms_is_equal_to_the_declared_one) {
  CHECK_THROW(load("p qubo 0 100 3 3\n0 0 -0.5\n0 1 2.0\n1 2 4\n2 2 -0.7\n"),
              std::invalid_argument);
}

TEST_CASE(io_test, saved_solution_has_correct_structure) {
  char state[] = {0, 1, 1, 0, 1};
  qubo::Solution solution(state, state + 5, -12.5);
  std::stringstream stream("");
  solution.save(stream);
  REQUIRE(stream.str() == "0,1,2,3,4,energy\n0,1,1,0,1,-12.5\n");
}

// --- grammar edge cases (qubo.hpp:392-417), beyond the reference's tests ---

TEST_CASE(io_test, header_needs_exactly_one_space_in_p_qubo) {
  CHECK_THROW(load("p  qubo 0 4 1 0\n0 0 1\n"), std::invalid_argument);
  CHECK_THROW(load("p\tqubo 0 4 1 0\n0 0 1\n"), std::invalid_argument);
  REQUIRE(load("p qubo   0\t4 1 0\n0 0 1\n").get_nodes() == 1);
}

TEST_CASE(io_test, index_must_be_followed_by_a_space) {
  CHECK_THROW(load("p qubo 0 4 1 0\n0\t0 1\n"), std::invalid_argument);
  CHECK_THROW(load("p qubo 0 4 1 0\n0 0\t1\n"), std::invalid_argument);
  REQUIRE(load("p qubo 0 4 1 0\n  0   0 \t 1.5  \n").get_variable(0) == 1.5);
}

TEST_CASE(io_test, blank_lines_are_rejected) {
  CHECK_THROW(load("p qubo 0 4 1 0\n\n0 0 1\n"), std::invalid_argument);
  CHECK_THROW(load("p qubo 0 4 1 0\n0 0 1\n\n"), std::invalid_argument);
  CHECK_THROW(load("\np qubo 0 4 1 0\n0 0 1\n"), std::invalid_argument);
  REQUIRE(load("p qubo 0 4 1 0\n0 0 1").get_nodes() == 1);  // no trailing newline
  REQUIRE(load("p qubo 0 4 1 0\r\n0 0 1\r\n").get_nodes() == 1);
}

TEST_CASE(io_test, duplicate_keys_keep_the_first_value_and_count_once) {
  auto p = load("p qubo 0 4 1 1\n0 0 1\n0 0 2\n0 1 3\n0 1 4\n");
  REQUIRE(p.get_variable(0) == 1.0);
  REQUIRE(p.get_connection(std::pair<int, int>(0, 1)) == 3.0);
}

TEST_CASE(io_test, header_counts_are_checked) {
  CHECK_THROW(load("0 0 1\n"), std::invalid_argument);                       // no header
  CHECK_THROW(load("p qubo 0 1 2 0\n0 0 1\n1 1 1\n"), std::invalid_argument);  // linear > max
  CHECK_THROW(load("p qubo 0 4 2 0\n0 0 1\n"), std::invalid_argument);       // linear count
  CHECK_THROW(load(""), std::invalid_argument);                              // empty
  CHECK_THROW(load("p qubo 0 4 0 0\n"), std::invalid_argument);              // no coefficients
}

TEST_CASE(io_test, comments_anywhere_and_numbers_in_all_forms) {
  auto p = load("c a\nc\np qubo 0 9 3 1\nc mid\t comment\n0 0 .5\n1 1 -2.\n2 2 1e1\n0 2 -1.25E-1\n");
  REQUIRE(p.get_variable(0) == 0.5);
  REQUIRE(p.get_variable(1) == -2.0);
  REQUIRE(p.get_variable(2) == 10.0);
  REQUIRE(p.get_connection(std::pair<int, int>(0, 2)) == -0.125);
  REQUIRE(p.get_nodes() == 3);
  auto q = load("p qubo 0 4 1 0\n0 0 -inf\n");
  REQUIRE(std::isinf(q.get_variable(0)) && q.get_variable(0) < 0);
  CHECK_THROW(load("p qubo 0 4 1 0\n0 0 -\n"), std::invalid_argument);
  CHECK_THROW(load("p qubo 0 4 1 0\n-1 0 1\n"), std::invalid_argument);
}

TEST_CASE(io_test, n_is_max_index_plus_one) {
  REQUIRE(load("p qubo 0 100 1 1\n7 7 1\n2 40 3\n").get_nodes() == 41);
}
