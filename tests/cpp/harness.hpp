// harness.hpp -- minimal test harness for the C++ unit tests (no Boost/GTest
// in this toolchain).  Suites and cases register themselves; the runner
// accepts --run_test=<suite> like Boost.Test and prints one line per case.
#pragma once

#include <cmath>
#include <functional>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace harness {

struct Case {
  std::string suite, name;
  std::function<void()> fn;
};

inline std::vector<Case> &registry() {
  static std::vector<Case> r;
  return r;
}

struct Registrar {
  Registrar(const char *suite, const char *name, std::function<void()> fn) {
    registry().push_back({suite, name, std::move(fn)});
  }
};

struct Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline std::string where(const char *file, int line, const char *what) {
  std::ostringstream s;
  s << file << ":" << line << ": " << what;
  return s.str();
}

inline int run(int argc, char **argv) {
  std::string only;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a.rfind("--run_test=", 0) == 0) only = a.substr(11);
  }
  int failed = 0, ran = 0;
  for (auto &c : registry()) {
    if (!only.empty() && c.suite != only) continue;
    ++ran;
    try {
      c.fn();
      std::cout << "PASS " << c.suite << "/" << c.name << "\n";
    } catch (const std::exception &e) {
      ++failed;
      std::cout << "FAIL " << c.suite << "/" << c.name << ": " << e.what() << "\n";
    }
  }
  std::cout << ran - failed << "/" << ran << " passed\n";
  if (ran == 0) {
    std::cout << "no test matched\n";
    return 2;
  }
  return failed ? 1 : 0;
}

}  // namespace harness

#define HCAT2(a, b) a##b
#define HCAT(a, b) HCAT2(a, b)
#define TEST_CASE(suite, name)                                                     \
  static void HCAT(suite##_, name)();                                              \
  static harness::Registrar HCAT(reg_##suite##_, name)(#suite, #name,              \
                                                       HCAT(suite##_, name));      \
  static void HCAT(suite##_, name)()
#define REQUIRE(cond)                                                              \
  do {                                                                             \
    if (!(cond)) throw harness::Failure(harness::where(__FILE__, __LINE__, #cond)); \
  } while (0)
#define CHECK_THROW(expr, Exc)                                                     \
  do {                                                                             \
    bool _thrown = false;                                                          \
    try {                                                                          \
      (void)(expr);                                                                \
    } catch (const Exc &) {                                                        \
      _thrown = true;                                                              \
    }                                                                              \
    if (!_thrown)                                                                  \
      throw harness::Failure(harness::where(__FILE__, __LINE__, "no " #Exc " from " #expr)); \
  } while (0)
