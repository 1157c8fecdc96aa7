/*! file one-solver-anneal.cpp
 *
 * Drop-in for oneSolver's one-solver-anneal (app/one-solver-anneal.cpp:41-178):
 * same flags, defaults, validation messages, stdout log lines, exit codes and
 * output file.  `--device-type gpu` runs on the MI355X through
 * libonesolver.so; `host` uses the calling thread and `cpu` all host cores.
 *
 * Options are parsed like Boost.Program_options' command-line style:
 * "--name value" or "--name=value", unambiguous prefixes accepted, each
 * option at most once, no positional arguments.  Parse errors print
 * "error: <message>" and exit 1; validation failures exit with -1 (255).
 *
 * Extensions (optional, reference-compatible defaults): --seed (the
 * reference's fixed engine seed, 1234) and --sweeps-per-beta (the library
 * parameter of sa::anneal that the reference CLI leaves at 1).  The fast
 * path is opt-in: --mode sweep runs sa::anneal_sweeps (sweep.hpp) on the gpu
 * device -- num-iter beta values, each for sweeps-per-beta sweeps of N
 * attempts in site order -- with --beta-kind (temperature, the reference's
 * use of beta, or inverse), --path (auto, dense or sparse),
 * --refresh-interval (dense: sweeps between FP64 field refreshes) and
 * --warm-start (a file of trained warm-start network parameters).  In the
 * default --mode reference the output is the reference's, line for line.
 */
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "options.hpp"
#include "onesolver/helpers/devices.hpp"
#include "onesolver/model/qubo.hpp"
#include "onesolver/model/solution.hpp"
#include "onesolver/simulated_annealing/annealing.hpp"
#include "onesolver/simulated_annealing/schedules.hpp"
#include "onesolver/simulated_annealing/sweep.hpp"

namespace {

const char *kHelp =
    "Allowed options:\n"
    "  --help                                produce help message\n"
    "  --input arg                           input file\n"
    "  --output arg                          output file\n"
    "  --num-iter arg (=100)                 number of iterations of the algorithm\n"
    "  --num-tries arg (=100)                number of trajectories to try\n"
    "  --schedule-type arg (=geometric)      type of beta schedule tu use, either \n"
    "                                        linear or geometric\n"
    "  --beta-min arg (=0.10000000000000001) minimum value of beta in the annealing \n"
    "                                        schedule (default 0.1)\n"
    "  --beta-max arg (=1)                   maximum value of beta in the annealing \n"
    "                                        schedule (default 1.0)\n"
    "  --device-type arg (=host)             device type to use (cpu, gpu or host)\n"
    "\n"
    "Extensions:\n"
    "  --seed arg (=1234)                    seed of the Philox4x32-10 streams\n"
    "  --sweeps-per-beta arg (=1)            attempts (sweep mode: sweeps) per beta\n"
    "                                        value\n"
    "  --mode arg (=reference)               reference (random-site attempts, as\n"
    "                                        the reference) or sweep (gpu only)\n"
    "  --beta-kind arg (=temperature)        sweep mode: beta as a temperature (as\n"
    "                                        the reference) or inverse\n"
    "  --path arg (=auto)                    sweep mode: auto, dense or sparse\n"
    "  --refresh-interval arg (=0)           sweep mode, dense: sweeps between FP64\n"
    "                                        field refreshes (0: at the end)\n"
    "  --warm-start arg                      sweep mode: file of warm-start network\n"
    "                                        parameters (fp32)\n";

const std::vector<std::string> kNames = {
    "help",        "input",           "output",   "num-iter",  "num-tries", "schedule-type",
    "beta-min",    "beta-max",        "device-type", "seed",   "sweeps-per-beta", "mode",
    "beta-kind",   "path",            "refresh-interval", "warm-start"};

}  // namespace

int main(int argc, char *argv[]) {
  std::string input_file, output_file, schedule_type, device_type;
  unsigned num_iter, num_tries;
  double beta_min, beta_max;
  try {
    cli::Options opt = cli::parse(argc, argv, kNames);
    if (opt.help) {
      std::cout << kHelp << std::endl;
      return 0;
    }
    num_iter = cli::convert<unsigned>(opt, "num-iter", 100u);
    num_tries = cli::convert<unsigned>(opt, "num-tries", 100u);
    schedule_type = cli::convert<std::string>(opt, "schedule-type", "geometric");
    beta_min = cli::convert<double>(opt, "beta-min", 0.1);
    beta_max = cli::convert<double>(opt, "beta-max", 1.0);
    device_type = cli::convert<std::string>(opt, "device-type", "host");
    const uint64_t seed = cli::convert<uint64_t>(opt, "seed", 1234u);
    const unsigned sweeps_per_beta = cli::convert<unsigned>(opt, "sweeps-per-beta", 1u);
    const std::string mode = cli::convert<std::string>(opt, "mode", "reference");
    const std::string beta_kind = cli::convert<std::string>(opt, "beta-kind", "temperature");
    const std::string path = cli::convert<std::string>(opt, "path", "auto");
    const int refresh_interval = cli::convert<int>(opt, "refresh-interval", 0);
    const std::string warm_file = cli::convert<std::string>(opt, "warm-start", "");

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
    if (schedule_type != "linear" && schedule_type != "geometric") {
      std::cerr << "Unknown beta schedule: " << schedule_type << std::endl;
      return -1;
    }
    if (beta_max < 0 || beta_min < 0) {
      std::cerr << "Invalid schedule, both ends of beta range need to be positive" << std::endl;
      return -1;
    }
    if (beta_min >= beta_max) {
      std::cerr << "Invalid schedule, initial beta is not lesser than final beta" << std::endl;
      return -1;
    }
    if (sweeps_per_beta < 1) {
      std::cerr << "Invalid sweeps per beta, need at least 1" << std::endl;
      return -1;
    }
    if (mode != "reference" && mode != "sweep") {
      std::cerr << "Unknown mode: " << mode << std::endl;
      return -1;
    }
    if (beta_kind != "temperature" && beta_kind != "inverse") {
      std::cerr << "Unknown beta kind: " << beta_kind << std::endl;
      return -1;
    }
    if (path != "auto" && path != "dense" && path != "sparse") {
      std::cerr << "Unknown path: " << path << std::endl;
      return -1;
    }
    if (refresh_interval < 0) {
      std::cerr << "Invalid refresh interval, need 0 or more" << std::endl;
      return -1;
    }
    const bool sweep_only = opt.values.count("beta-kind") || opt.values.count("path") ||
                            opt.values.count("refresh-interval") || !warm_file.empty();
    if (mode == "reference" && sweep_only) {
      std::cerr << "--beta-kind, --path, --refresh-interval and --warm-start need --mode sweep"
                << std::endl;
      return -1;
    }
    if (mode == "sweep" && device_type != "gpu") {
      std::cerr << "Sweep mode runs on the gpu device" << std::endl;
      return -1;
    }

    std::cout << "Reading input from: " << input_file << std::endl;
    std::cout << "Output will be saved to: " << output_file << std::endl;
    std::cout << "Schedule type: " << schedule_type << std::endl;
    std::cout << "Beta range: [" << beta_min << ", " << beta_max << "]" << std::endl;
    std::cout << "Number of iterations: " << num_iter << std::endl;
    std::cout << "Number of tries: " << num_tries << std::endl;

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

    std::vector<double> beta_schedule(num_iter);
    if (num_iter > 0) {
      if (schedule_type == "linear")
        construct_linear_beta_schedule(beta_schedule, beta_min, beta_max, num_iter);
      else
        construct_geometric_beta_schedule(beta_schedule, beta_min, beta_max, num_iter);
    }

    auto run = [&]() {
      if (mode != "sweep")
        return sa::anneal(instance, *device, beta_schedule, (int)num_iter, num_tries,
                          (int)sweeps_per_beta, seed);
      sa::SweepOptions so;
      so.beta_kind =
          beta_kind == "temperature" ? sa::BetaKind::temperature : sa::BetaKind::inverse;
      so.path = path == "dense"    ? sa::SweepPath::dense
                : path == "sparse" ? sa::SweepPath::sparse
                                   : sa::SweepPath::automatic;
      so.seed = seed;
      so.refresh_interval = refresh_interval;
      if (!warm_file.empty()) so.warm_params = sa::load_warm_params(warm_file);
      sa::SweepReport rep;
      auto sol = sa::anneal_sweeps(instance, *device, beta_schedule, (int)num_iter, num_tries,
                                   (int)sweeps_per_beta, so, &rep);
      std::cout << "Mode: sweep (" << rep.path << " path, " << rep.attempts << " attempts, "
                << rep.kernel_ms << " ms in the sweep kernels)" << std::endl;
      return sol;
    };
    auto solution = run();

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
