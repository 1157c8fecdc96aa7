/*! file solution.hpp
 *
 * Solution of a QUBO model and its two-line CSV form (oneSolver
 * include/model/solution.hpp:19-68):
 *   "0,1,...,N-1,energy\n" then "x0,...,xN-1,E\n", E in default ostream
 *   formatting (6 significant digits).
 */
#ifndef ONESOLVER_SOLUTION_HPP_
#define ONESOLVER_SOLUTION_HPP_

#include <ostream>
#include <vector>

namespace qubo {

class Solution {
 public:
  /*! i-th element is the value of variable i. */
  std::vector<char> state;
  /*! Energy of `state`. */
  double energy;

  template <typename InputIt>
  Solution(InputIt begin, InputIt end, double energy) : state(begin, end), energy(energy) {}

  void save(std::ostream &stream) {
    for (size_t i = 0; i < state.size(); ++i) stream << i << ',';
    stream << "energy" << std::endl;
    for (char bit : state) stream << int(bit) << ',';
    stream << energy << std::endl;
  }
};

}  // namespace qubo

#endif
