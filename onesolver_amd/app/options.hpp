/*! file options.hpp
 *
 * Command-line parsing shared by the CLIs, in the style of
 * Boost.Program_options as the reference's apps use it: "--name value" or
 * "--name=value", unambiguous prefixes accepted, each option at most once, no
 * positional arguments.  Errors throw std::runtime_error with Boost's
 * messages.
 */
#ifndef ONESOLVER_APP_OPTIONS_HPP_
#define ONESOLVER_APP_OPTIONS_HPP_

#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace cli {

struct Options {
  std::map<std::string, std::string> values;
  bool help = false;
};

inline std::string resolve(const std::vector<std::string> &names, const std::string &given) {
  for (const auto &n : names)
    if (n == given) return n;
  std::vector<std::string> hits;
  for (const auto &n : names)
    if (n.compare(0, given.size(), given) == 0) hits.push_back(n);
  if (hits.empty()) throw std::runtime_error("unrecognised option '--" + given + "'");
  if (hits.size() > 1) {
    std::string m = "option '--" + given + "' is ambiguous and matches ";
    for (size_t i = 0; i < hits.size(); ++i)
      m += (i ? (i + 1 == hits.size() ? ", and " : ", ") : "") + ("'--" + hits[i] + "'");
    throw std::runtime_error(m);
  }
  return hits[0];
}

inline Options parse(int argc, char *argv[], const std::vector<std::string> &names) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    std::string tok = argv[i];
    if (tok.size() < 3 || tok.compare(0, 2, "--") != 0) {
      if (tok.size() >= 2 && tok[0] == '-') throw std::runtime_error("unrecognised option '" + tok + "'");
      throw std::runtime_error("too many positional options have been specified on the command line");
    }
    std::string body = tok.substr(2), val;
    bool has_val = false;
    const size_t eq = body.find('=');
    if (eq != std::string::npos) {
      val = body.substr(eq + 1);
      body = body.substr(0, eq);
      has_val = true;
    }
    const std::string name = resolve(names, body);
    if (name == "help") {
      if (has_val) throw std::runtime_error("option '--help' does not take any arguments");
      o.help = true;
      continue;
    }
    if (!has_val) {
      if (i + 1 >= argc)
        throw std::runtime_error("the required argument for option '--" + name + "' is missing");
      val = argv[++i];
    }
    if (o.values.count(name))
      throw std::runtime_error("option '--" + name + "' cannot be specified more than once");
    o.values[name] = val;
  }
  return o;
}

template <typename T>
T convert(const Options &o, const std::string &name, T def) {
  auto it = o.values.find(name);
  if (it == o.values.end()) return def;
  std::istringstream is(it->second);
  T v;
  if (!(is >> v) || !is.eof() || (std::is_unsigned<T>::value && it->second.find('-') != std::string::npos))
    throw std::runtime_error("the argument ('" + it->second + "') for option '--" + name +
                             "' is invalid");
  return v;
}

template <>
inline std::string convert<std::string>(const Options &o, const std::string &name,
                                        std::string def) {
  auto it = o.values.find(name);
  return it == o.values.end() ? def : it->second;
}

}  // namespace cli

#endif
