/*! file one-solver-exhaustive.cpp
 *
 * Drop-in for oneSolver's one-solver-exhaustive
 * (app/one-solver-exhaustive.cpp:23-105): same flags, default device, messages,
 * exit codes and output file.  `--device-type gpu` runs the search on the
 * MI355X through libonesolver.so (os_exhaustive); `host` uses the calling
 * thread and `cpu` all host cores.  Options parse as in one-solver-anneal.
 * If no device of the requested type can be initialised the reference goes on
 * with a null queue; this program exits 1 after the same message.
 */
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "options.hpp"
#include "onesolver/exhaustive/exhaustive.hpp"
#include "onesolver/helpers/devices.hpp"
#include "onesolver/model/qubo.hpp"
#include "onesolver/model/solution.hpp"

namespace {

const char *kHelp =
    "Allowed options:\n"
    "  --help                    produce help message\n"
    "  --input arg               input file\n"
    "  --output arg              output file\n"
    "  --device-type arg (=host) device type to use (cpu, gpu or host)\n";

const std::vector<std::string> kNames = {"help", "input", "output", "device-type"};

}  // namespace

int main(int argc, char *argv[]) {
  std::string input_file, output_file, device_type;
  try {
    cli::Options opt = cli::parse(argc, argv, kNames);
    if (opt.help) {
      std::cout << kHelp << std::endl;
      return 0;
    }
    device_type = cli::convert<std::string>(opt, "device-type", "host");
    if (!opt.values.count("input")) {
      std::cerr << "No input file provided." << std::endl;
      return -1;
    }
    input_file = opt.values["input"];
    if (!opt.values.count("output")) {
      std::cerr << "No output file provided." << std::endl;
      return -1;
    }
    output_file = opt.values["output"];
    if (device_type != "cpu" && device_type != "gpu" && device_type != "host") {
      std::cerr << "Unknown device type: " << device_type << std::endl;
      return -1;
    }
    std::cout << "Reading input from: " << input_file << std::endl;
    std::cout << "Output will be saved to: " << output_file << std::endl;

    std::ifstream qubo_file(input_file);
    if (!qubo_file) {
      std::cerr << "can not open input file: " << input_file << std::endl;
      return -1;
    }
    qubo_file.unsetf(std::ios::skipws);
    auto instance = qubo::QUBOModel<int, double>::load(qubo_file);

    std::unique_ptr<devices::Device> device;
    try {
      device.reset(new devices::Device(*devices::construct_device_selector(device_type)));
    } catch (std::runtime_error &) {
      std::cerr << "No devices of given type could be initialized." << std::endl;
      return 1;
    }
    std::cout << "Using device: " << device->name() << std::endl;

    auto solution = exhaustive::solve(*device, instance);

    std::ofstream results_file(output_file);
    solution.save(results_file);
    results_file.close();
  } catch (std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  } catch (...) {
    std::cerr << "Exception of unknown type!\n";
  }
  return 0;
}
