// main-test: C++ unit tests of the host-side model, parser, writer, helpers
// and schedules (the reference's Boost.Test module tests/*.cpp, ported to a
// local harness).  Run: main-test [--run_test=<suite>]
#include "harness.hpp"

int main(int argc, char **argv) { return harness::run(argc, argv); }
