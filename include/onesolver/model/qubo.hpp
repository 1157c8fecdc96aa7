/*! file qubo.hpp
 *
 * QUBO model container and `.qubo` reader of the MI355X annealing library.
 * Same API as oneSolver's include/model/qubo.hpp (QUBOModel :53-240,
 * QUBOBuilder :282-379, parse_qubo :392-417, load :425-430), without Boost:
 * the grammar is parsed by a hand-written recursive-descent reader that
 * accepts and rejects exactly what the reference's Spirit grammar does:
 *
 *   file        := comment* header? (comment | coefficient)*
 *   comment     := 'c' print* (eol | eoi)
 *   header      := "p qubo" uint uint uint uint (eol | eoi)
 *   coefficient := index index real (eol | eoi)
 *   index       := uint ' '            (the space must follow the digits)
 *
 * Blanks (space, tab) are skipped before every token; eol is \n, \r\n or \r.
 * Every violation, and every semantic check of QUBOBuilder::build_qubo, throws
 * std::invalid_argument.
 */
#ifndef ONESOLVER_QUBO_MODEL_HPP_
#define ONESOLVER_QUBO_MODEL_HPP_

#include <cctype>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <iterator>
#include <limits>
#include <ostream>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>

#include "../helpers/hash.hpp"
#include "../helpers/insert.hpp"

namespace qubo {

/** Linear coefficients (node -> value), include/model/qubo.hpp:35-36. */
template <class NodeType, class CoefType>
using LinearCoef = std::unordered_map<NodeType, CoefType>;

/** Quadratic coefficients ((i, j) -> value), include/model/qubo.hpp:44-46. */
template <class NodeType, class CoefType>
using QuadraticCoef =
    std::unordered_map<std::pair<NodeType, NodeType>, CoefType, helpers::hash_pair>;

/** QUBO model: linear and quadratic maps plus the number of nodes. */
template <class NodeType, class CoefType>
class QUBOModel {
 protected:
  LinearCoef<NodeType, CoefType> linear;
  QuadraticCoef<NodeType, CoefType> quadratic;
  unsigned long num_nodes = 0;

 public:
  QUBOModel(LinearCoef<NodeType, CoefType> &c_linear,
            QuadraticCoef<NodeType, CoefType> &c_quadratic)
      : linear(c_linear), quadratic(c_quadratic) {}

  /** "QUBO model" followed by " i--i:v " per linear term, then "i--j:v" per
   *  quadratic term (std::to_string formatting), as qubo.hpp:182-202. */
  std::string str() const {
    std::string s("QUBO model");
    for (const auto &kv : linear) {
      s += " " + std::to_string(kv.first) + "--" + std::to_string(kv.first) + ":" +
           std::to_string(kv.second) + " ";
    }
    for (const auto &kv : quadratic) {
      s += std::to_string(kv.first.first) + "--" + std::to_string(kv.first.second) + ":" +
           std::to_string(kv.second);
    }
    return s;
  }

  /** Insert or overwrite a linear coefficient (qubo.hpp:204-208). */
  void add_variable(const NodeType &vi, const CoefType &hi) {
    helpers::insert_model(linear, vi, hi);
  }

  /** Insert or overwrite a quadratic coefficient (qubo.hpp:210-214). */
  void add_connection(const std::pair<NodeType, NodeType> &connection, const CoefType &Ji) {
    helpers::insert_model(quadratic, connection, Ji);
  }

  /** Value of a linear coefficient, 0 if absent (qubo.hpp:229-240). */
  const CoefType get_variable(const NodeType &vi) const {
    auto it = linear.find(vi);
    return it == linear.end() ? CoefType(0) : it->second;
  }

  /** Value of a quadratic coefficient, 0 if absent (qubo.hpp:216-227). */
  const CoefType get_connection(const std::pair<NodeType, NodeType> &connection) const {
    auto it = quadratic.find(connection);
    return it == quadratic.end() ? CoefType(0) : it->second;
  }

  void set_nodes(int number) { num_nodes = (unsigned long)number; }
  unsigned long get_nodes() const { return num_nodes; }

  const LinearCoef<NodeType, CoefType> &linear_terms() const { return linear; }
  const QuadraticCoef<NodeType, CoefType> &quadratic_terms() const { return quadratic; }

  /** Parse a `.qubo` stream.  Like the reference (one-solver-anneal.cpp:136),
   *  callers unset std::ios::skipws first: characters are read through
   *  std::istream_iterator<char>, which honours that flag. */
  static QUBOModel<int, double> load(std::istream &stream);
};

template <class NodeType, class CoefType>
std::ostream &operator<<(std::ostream &os, const QUBOModel<NodeType, CoefType> &qubos) {
  os << qubos.str();
  return os;
}

/*! Builder fed by the parser (qubo.hpp:282-379). */
struct QUBOBuilder {
 private:
  LinearCoef<int, double> linear_c{};
  QuadraticCoef<int, double> quadratic_c{};
  int num_quadratic = -1;
  int num_linear = -1;
  int max_nodes = -1;
  std::set<int> used_nodes;

 public:
  /** Add (i, j, q_ij): i == j linear, i < j quadratic, first value of a
   *  repeated key wins; i > j throws. */
  void add_element(int i, int j, double coef) {
    if (i == j) {
      linear_c.insert({i, coef});
    } else if (i < j) {
      quadratic_c.insert({{i, j}, coef});
    } else {
      throw std::invalid_argument(
          "Incorrect file, encountered coefficient from lower triangle of QUBO matrix.");
    }
    used_nodes.insert(i);
    used_nodes.insert(j);
  }
  void set_num_quadratic(int value) { num_quadratic = value; }
  void set_num_linear(int value) { num_linear = value; }
  void set_max_nodes(int value) { max_nodes = value; }

  QUBOModel<int, double> build_qubo() {
    if (linear_c.empty() && quadratic_c.empty())
      throw std::invalid_argument("An empty input, no coefficients defined.");
    if (num_quadratic == -1 || num_linear == -1 || max_nodes == -1)
      throw std::invalid_argument("No header line or header line misformatted.");
    if (num_linear > max_nodes)
      throw std::invalid_argument("Number of linear terms is greater than num nodes.");
    if (quadratic_c.size() != (size_t)num_quadratic)
      throw std::invalid_argument("Number of quadratic terms is not equal to the declared one.");
    if (linear_c.size() != (size_t)num_linear)
      throw std::invalid_argument("Number of linear terms is not equal to the declared one.");
    QUBOModel<int, double> q(linear_c, quadratic_c);
    q.set_nodes(*used_nodes.rbegin() + 1);
    return q;
  }
};

namespace detail {

// Builder: QUBOBuilder, or any type with its four set/add member functions.
template <typename It, typename Builder = QUBOBuilder>
class QuboReader {
 public:
  QuboReader(It first, It last, Builder &b) : p_(first), end_(last), b_(b) {}

  // true iff the whole input matched the grammar
  bool run() {
    while (comment()) {
    }
    header();
    while (comment() || coefficient()) {
    }
    skip_blanks();
    return p_ == end_;
  }

 private:
  It p_, end_;
  Builder &b_;

  static bool blank(char c) { return c == ' ' || c == '\t'; }
  void skip_blanks() {
    while (p_ != end_ && blank(*p_)) ++p_;
  }
  bool eol_or_eoi() {
    skip_blanks();
    if (p_ == end_) return true;
    if (*p_ == '\n') {
      ++p_;
      return true;
    }
    if (*p_ == '\r') {
      ++p_;
      if (p_ != end_ && *p_ == '\n') ++p_;
      return true;
    }
    return false;
  }
  // unsigned decimal, no sign, fails on overflow of 32 bits
  bool uint_(It &q, unsigned &out) {
    if (q == end_ || !std::isdigit((unsigned char)*q)) return false;
    uint64_t v = 0;
    while (q != end_ && std::isdigit((unsigned char)*q)) {
      v = v * 10 + (uint64_t)(*q - '0');
      if (v > std::numeric_limits<unsigned>::max()) return false;
      ++q;
    }
    out = (unsigned)v;
    return true;
  }
  bool match_ci(It &q, const char *word) {
    It t = q;
    for (; *word; ++word, ++t) {
      if (t == end_ || std::tolower((unsigned char)*t) != *word) return false;
    }
    q = t;
    return true;
  }
  // real number: sign? (nan("(...)")? | inf(inity)? | digits [. digits] [exp])
  bool real_(It &q, double &out) {
    It t = q;
    std::string tok;
    if (t != end_ && (*t == '+' || *t == '-')) tok += *t++;
    if (t != end_ && (*t == 'n' || *t == 'N')) {
      if (!match_ci(t, "nan")) return false;
      if (t != end_ && *t == '(') {
        It u = t;
        ++u;
        while (u != end_ && *u != ')') ++u;
        if (u != end_) t = ++u;
      }
      out = std::numeric_limits<double>::quiet_NaN();
      if (!tok.empty() && tok[0] == '-') out = -out;
      q = t;
      return true;
    }
    if (t != end_ && (*t == 'i' || *t == 'I')) {
      if (!match_ci(t, "inf")) return false;
      match_ci(t, "inity");
      out = std::numeric_limits<double>::infinity();
      if (!tok.empty() && tok[0] == '-') out = -out;
      q = t;
      return true;
    }
    bool digits = false;
    while (t != end_ && std::isdigit((unsigned char)*t)) {
      tok += *t++;
      digits = true;
    }
    if (t != end_ && *t == '.') {
      It u = t;
      ++u;
      bool frac = false;
      std::string f(".");
      while (u != end_ && std::isdigit((unsigned char)*u)) {
        f += *u++;
        frac = true;
      }
      if (digits || frac) {
        tok += f;
        t = u;
        digits = true;
      }
    }
    if (!digits) return false;
    if (t != end_ && (*t == 'e' || *t == 'E')) {
      It u = t;
      ++u;
      std::string e("e");
      if (u != end_ && (*u == '+' || *u == '-')) e += *u++;
      bool ed = false;
      while (u != end_ && std::isdigit((unsigned char)*u)) {
        e += *u++;
        ed = true;
      }
      if (ed) {
        tok += e;
        t = u;
      }
    }
    out = std::strtod(tok.c_str(), nullptr);
    q = t;
    return true;
  }
  bool comment() {
    It save = p_;
    skip_blanks();
    if (p_ == end_ || *p_ != 'c') {
      p_ = save;
      return false;
    }
    ++p_;
    for (;;) {
      skip_blanks();
      if (p_ != end_ && (unsigned char)*p_ < 0x80 && std::isprint((unsigned char)*p_))
        ++p_;
      else
        break;
    }
    if (!eol_or_eoi()) {
      p_ = save;
      return false;
    }
    return true;
  }
  void header() {
    It save = p_;
    skip_blanks();
    It q = p_;
    const char *lit = "p qubo";
    for (; *lit; ++lit, ++q) {
      if (q == end_ || *q != *lit) {
        p_ = save;
        return;
      }
    }
    p_ = q;
    unsigned v[4];
    for (int k = 0; k < 4; ++k) {
      skip_blanks();
      if (!uint_(p_, v[k])) {
        p_ = save;
        return;
      }
      if (k == 1) b_.set_max_nodes((int)v[1]);
      if (k == 2) b_.set_num_linear((int)v[2]);
      if (k == 3) b_.set_num_quadratic((int)v[3]);
    }
    if (!eol_or_eoi()) p_ = save;
  }
  bool index_(unsigned &out) {
    skip_blanks();
    It q = p_;
    if (!uint_(q, out)) return false;
    if (q == end_ || *q != ' ') return false;
    p_ = ++q;
    return true;
  }
  bool coefficient() {
    It save = p_;
    unsigned i, j;
    double v;
    if (!index_(i) || !index_(j)) {
      p_ = save;
      return false;
    }
    skip_blanks();
    if (!real_(p_, v)) {
      p_ = save;
      return false;
    }
    b_.add_element((int)i, (int)j, v);
    if (!eol_or_eoi()) {
      p_ = save;
      return false;
    }
    return true;
  }
};

}  // namespace detail

/*! Parse `.qubo` text in [first, last) (qubo.hpp:392-417). */
template <typename Iterator>
QUBOModel<int, double> parse_qubo(Iterator first, Iterator last) {
  QUBOBuilder builder;
  detail::QuboReader<Iterator> reader(first, last, builder);
  if (!reader.run()) throw std::invalid_argument("Parsing failed. Incorrect file format.");
  return builder.build_qubo();
}

template <class NodeType, class CoefType>
QUBOModel<int, double> QUBOModel<NodeType, CoefType>::load(std::istream &stream) {
  std::string text{std::istream_iterator<char>(stream), std::istream_iterator<char>()};
  return parse_qubo(text.cbegin(), text.cend());
}

}  // namespace qubo

#endif
