/*! file insert.hpp
 *
 * Insert-or-overwrite into a coefficient map (oneSolver
 * include/helpers/insert.hpp:27-38): add_variable / add_connection replace an
 * existing value, unlike the parser, which keeps the first one.
 */
#ifndef ONESOLVER_INSERT_HPP_
#define ONESOLVER_INSERT_HPP_

#include <unordered_map>

namespace helpers {

template <class NodeType, class CoefType, class Hash>
void insert_model(std::unordered_map<NodeType, CoefType, Hash> &um, const NodeType &key,
                  const CoefType &val) {
  um[key] = val;
}

}  // namespace helpers

#endif
