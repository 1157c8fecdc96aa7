/*! file one-solver-batch.cpp
 *
 * Many `.qubo` instances in one run: a directory (every `*.qubo` and binary
 * `*.osqb` file, in name order) or a manifest (one path per line, relative to
 * the manifest's directory; blank lines and lines starting with '#'
 * skipped).  Each instance gets its own Solution CSV in --output, named after
 * its file (stem + ".csv"), in the format of oneSolver's qubo::Solution::save
 * (include/model/solution.hpp:58-67).  No reference counterpart: the
 * reference's one-solver-anneal (app/one-solver-anneal.cpp:41-178) takes one
 * file per run.
 *
 * Files are parsed on the GPU (os_qubo_parse / os_qubo_read_binary: the
 * reference's grammar, rejections and messages).  All instances are annealed
 * together in sweep mode:
 *   - dense (os_dense_batch_*): every instance has the same n <= 16384 and at
 *     least a quarter of its couplings set; or --path dense;
 *   - sparse (os_csr_batch_*): otherwise; or --path sparse.  Sweeps visit
 *     the sites class by class of a greedy colouring (--site-order colour,
 *     the default: several independent sites per wave step) or in index
 *     order (--site-order sequential).
 * Instance k draws its random numbers as stream k.  The beta schedule is the
 * reference's (linear or geometric over --num-iter values, one sweep per
 * value); --beta-kind temperature (default) uses beta as the reference does,
 * as a temperature: a flip is accepted with probability
 * min(1, exp(-dE / beta)); --beta-kind inverse uses exp(-beta dE).
 * Each instance's answer is its lowest FP64 energy over --num-tries replicas
 * (first replica on ties).
 *
 * Exit codes: 0 success, -1 (255) invalid options, 1 errors (file, parse,
 * device), as one-solver-anneal.
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <dirent.h>
#include <fstream>
#include <iostream>
#include <iterator>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>

#include "onesolver.h"
#include "onesolver/model/solution.hpp"
#include "onesolver/simulated_annealing/schedules.hpp"
#include "options.hpp"

namespace {

const char *kHelp =
    "Allowed options:\n"
    "  --help                                produce help message\n"
    "  --input arg                           directory of .qubo/.osqb files, or a manifest\n"
    "  --output arg                          directory for the per-instance CSV files\n"
    "  --num-iter arg (=100)                 sweeps (one per beta value)\n"
    "  --num-tries arg (=100)                replicas per instance\n"
    "  --schedule-type arg (=geometric)      linear or geometric\n"
    "  --beta-min arg (=0.10000000000000001) minimum value of beta\n"
    "  --beta-max arg (=1)                   maximum value of beta\n"
    "  --beta-kind arg (=temperature)        temperature (as the reference) or inverse\n"
    "  --seed arg (=1234)                    seed of the Philox4x32-10 streams\n"
    "  --path arg (=auto)                    auto, dense or sparse\n"
    "  --site-order arg (=colour)            sparse path: colour or sequential\n";

const std::vector<std::string> kNames = {"help",      "input",       "output",   "num-iter",
                                         "num-tries", "schedule-type", "beta-min", "beta-max",
                                         "beta-kind", "seed",        "path",     "site-order"};

bool is_dir(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

bool ends_with(const std::string &s, const std::string &suf) {
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

std::string stem(const std::string &path) {
  std::string b = path.substr(path.find_last_of('/') + 1);
  const size_t dot = b.find_last_of('.');
  return dot == std::string::npos || dot == 0 ? b : b.substr(0, dot);
}

std::vector<std::string> list_inputs(const std::string &input) {
  std::vector<std::string> files;
  if (is_dir(input)) {
    DIR *d = opendir(input.c_str());
    if (!d) throw std::runtime_error("can not open input directory: " + input);
    while (dirent *e = readdir(d)) {
      const std::string name = e->d_name;
      if (ends_with(name, ".qubo") || ends_with(name, ".osqb")) files.push_back(input + "/" + name);
    }
    closedir(d);
    std::sort(files.begin(), files.end());
    return files;
  }
  std::ifstream m(input);
  if (!m) throw std::runtime_error("can not open input file: " + input);
  const size_t slash = input.find_last_of('/');
  const std::string dir = slash == std::string::npos ? "" : input.substr(0, slash + 1);
  std::string line;
  while (std::getline(m, line)) {
    while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t'))
      line.pop_back();
    const size_t b = line.find_first_not_of(" \t");
    if (b == std::string::npos || line[b] == '#') continue;
    line = line.substr(b);
    files.push_back(line[0] == '/' ? line : dir + line);
  }
  return files;
}

struct Instance {
  std::string file;
  int32_t n = 0;
  int64_t nl = 0, nq = 0;
  std::vector<int32_t> ti, tj;
  std::vector<double> tv;
};

void check(int rc, const std::string &what) {
  if (rc != OS_OK) throw std::runtime_error(what + ": " + os_last_error());
}

Instance load(os_ctx *ctx, const std::string &file) {
  std::ifstream f(file, std::ios::binary);
  if (!f) throw std::runtime_error("can not open input file: " + file);
  std::string data{std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>()};
  os_qubo *q = nullptr;
  const bool binary = data.size() >= 4 && data.compare(0, 4, "OSQB") == 0;
  check(binary ? os_qubo_read_binary(ctx, data.data(), data.size(), &q)
               : os_qubo_parse(ctx, data.data(), data.size(), &q),
        file);
  Instance in;
  in.file = file;
  os_qubo_info(q, &in.n, nullptr, &in.nl, &in.nq);
  const size_t T = (size_t)(in.nl + in.nq);
  in.ti.resize(T);
  in.tj.resize(T);
  in.tv.resize(T);
  const int rc = os_qubo_terms(q, in.ti.data(), in.tj.data(), in.tv.data());
  os_qubo_destroy(q);
  check(rc, file);
  return in;
}

// The reference's flattened matrix (include/helpers/qubo_helpers.hpp:26-44).
void flatten(const Instance &in, double *q) {
  const size_t n = (size_t)in.n;
  std::fill(q, q + n * n, 0.0);
  for (size_t k = 0; k < in.ti.size(); ++k) {
    const size_t i = (size_t)in.ti[k], j = (size_t)in.tj[k];
    if (i == j) {
      q[i * n + i] = in.tv[k];
    } else {
      q[i * n + j] += in.tv[k];
      q[j * n + i] += in.tv[k];
    }
  }
}

// fn(0) .. fn(count - 1) on up to `threads` host threads (fn must not throw).
template <class F>
void parallel_for(size_t count, size_t threads, F fn) {
  std::atomic<size_t> next{0};
  auto work = [&] {
    for (size_t k; (k = next++) < count;) fn(k);
  };
  std::vector<std::thread> pool;
  for (size_t t = 1; t < std::min(threads, count); ++t) pool.emplace_back(work);
  work();
  for (auto &th : pool) th.join();
}

size_t host_threads() {
  const unsigned hw = std::thread::hardware_concurrency();
  return std::max<size_t>(1, std::min<size_t>(8, hw));
}

// Symmetric off-diagonal rows sorted by column, and the diagonal.  The
// entries are bucketed by row (counting sort), then each row is sorted by
// column, stably: O(nnz) for rows of a few entries.
void to_csr(const Instance &in, std::vector<int32_t> &rp, std::vector<int32_t> &col,
            std::vector<double> &val, std::vector<double> &diag) {
  const int32_t n = in.n;
  std::vector<int32_t> start(n + 1, 0);
  std::vector<double> d(n, 0.0);
  for (size_t k = 0; k < in.ti.size(); ++k)
    if (in.ti[k] != in.tj[k]) {
      ++start[in.ti[k] + 1];
      ++start[in.tj[k] + 1];
    }
  for (int32_t i = 0; i < n; ++i) start[i + 1] += start[i];
  const size_t base = col.size();
  col.resize(base + start[n]);
  val.resize(base + start[n]);
  std::vector<int32_t> fill(start.begin(), start.end() - 1);
  for (size_t k = 0; k < in.ti.size(); ++k) {
    const int32_t i = in.ti[k], j = in.tj[k];
    if (i == j) {
      d[i] = in.tv[k];
    } else {
      const size_t a = base + fill[i]++, b = base + fill[j]++;
      col[a] = j;
      val[a] = in.tv[k];
      col[b] = i;
      val[b] = in.tv[k];
    }
  }
  std::vector<std::pair<int32_t, double>> row;
  for (int32_t i = 0; i < n; ++i) {
    rp.push_back(start[i]);
    int32_t *c = col.data() + base + start[i];
    double *v = val.data() + base + start[i];
    const int32_t len = start[i + 1] - start[i];
    if (len <= 16) {
      for (int32_t x = 1; x < len; ++x)
        for (int32_t y = x; y > 0 && c[y - 1] > c[y]; --y) {
          std::swap(c[y - 1], c[y]);
          std::swap(v[y - 1], v[y]);
        }
    } else {
      row.clear();
      for (int32_t x = 0; x < len; ++x) row.push_back({c[x], v[x]});
      std::stable_sort(row.begin(), row.end(),
                       [](const std::pair<int32_t, double> &p, const std::pair<int32_t, double> &q) {
                         return p.first < q.first;
                       });
      for (int32_t x = 0; x < len; ++x) {
        c[x] = row[x].first;
        v[x] = row[x].second;
      }
    }
  }
  rp.push_back(start[n]);
  diag.insert(diag.end(), d.begin(), d.end());
}

}  // namespace

int main(int argc, char *argv[]) {
  try {
    cli::Options opt = cli::parse(argc, argv, kNames);
    if (opt.help) {
      std::cout << kHelp << std::endl;
      return 0;
    }
    const unsigned num_iter = cli::convert<unsigned>(opt, "num-iter", 100u);
    const unsigned num_tries = cli::convert<unsigned>(opt, "num-tries", 100u);
    const std::string schedule_type = cli::convert<std::string>(opt, "schedule-type", "geometric");
    const double beta_min = cli::convert<double>(opt, "beta-min", 0.1);
    const double beta_max = cli::convert<double>(opt, "beta-max", 1.0);
    const std::string beta_kind = cli::convert<std::string>(opt, "beta-kind", "temperature");
    const uint64_t seed = cli::convert<uint64_t>(opt, "seed", 1234u);
    const std::string path = cli::convert<std::string>(opt, "path", "auto");
    const std::string site_order = cli::convert<std::string>(opt, "site-order", "colour");
    if (!opt.values.count("input")) {
      std::cerr << "No input file provided." << std::endl;
      return -1;
    }
    if (!opt.values.count("output")) {
      std::cerr << "No output directory provided." << std::endl;
      return -1;
    }
    const std::string input = opt.values["input"], output = opt.values["output"];
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
    if (beta_kind != "temperature" && beta_kind != "inverse") {
      std::cerr << "Unknown beta kind: " << beta_kind << std::endl;
      return -1;
    }
    if (site_order != "colour" && site_order != "sequential") {
      std::cerr << "Unknown site order: " << site_order << std::endl;
      return -1;
    }
    if (path != "auto" && path != "dense" && path != "sparse") {
      std::cerr << "Unknown path: " << path << std::endl;
      return -1;
    }
    if (num_tries < 1) {
      std::cerr << "Invalid number of tries, need at least 1" << std::endl;
      return -1;
    }
    if (!is_dir(output)) {
      std::cerr << "can not open output directory: " << output << std::endl;
      return -1;
    }

    const std::vector<std::string> files = list_inputs(input);
    if (files.empty()) throw std::runtime_error("no .qubo or .osqb files in " + input);
    std::cout << "Instances: " << files.size() << std::endl;

    using clock = std::chrono::steady_clock;
    auto ms_since = [](clock::time_point t) {
      return std::chrono::duration<double, std::milli>(clock::now() - t).count();
    };
    os_ctx *ctx = nullptr;
    if (os_ctx_create(0, &ctx) != OS_OK) {
      std::cerr << "No devices of given type could be initialized." << std::endl;
      return 1;
    }
    auto t0 = clock::now();
    // Files are parsed by up to kParseThreads host threads, each on a
    // context (stream) of its own, so that one file's launches and
    // read-backs overlap another's.  A failure stops further files from
    // being taken; every file before it was taken already, so the error
    // reported is the first in file order, as in a serial run.
    std::vector<Instance> insts(files.size());
    {
      constexpr size_t kParseThreads = 4;
      std::vector<std::string> err(files.size());
      std::atomic<size_t> next{0};
      std::atomic<bool> failed{false};
      auto work = [&](os_ctx *c) {
        for (size_t k; !failed && (k = next++) < files.size();) {
          try {
            insts[k] = load(c, files[k]);
          } catch (std::exception &e) {
            err[k] = e.what();
            failed = true;
          }
        }
      };
      std::vector<std::thread> pool;
      for (size_t t = 1; t < std::min(kParseThreads, files.size()); ++t)
        pool.emplace_back([&] {
          os_ctx *c = nullptr;
          if (os_ctx_create(0, &c) != OS_OK) return;  // the other threads take its files
          work(c);
          os_ctx_destroy(c);
        });
      work(ctx);
      for (auto &th : pool) th.join();
      for (const auto &e : err)
        if (!e.empty()) throw std::runtime_error(e);
    }
    const double parse_ms = ms_since(t0);
    t0 = clock::now();

    std::vector<double> beta(num_iter);
    if (num_iter > 0) {
      if (schedule_type == "linear")
        construct_linear_beta_schedule(beta, beta_min, beta_max, num_iter);
      else
        construct_geometric_beta_schedule(beta, beta_min, beta_max, num_iter);
    }
    std::vector<float> ts(num_iter);
    for (unsigned t = 0; t < num_iter; ++t)
      ts[t] = (float)(beta_kind == "temperature" ? beta[t] : 1.0 / beta[t]);

    const int32_t I = (int32_t)insts.size(), R = (int32_t)num_tries;
    bool dense = path == "dense";
    if (path == "auto") {
      dense = true;
      for (const auto &in : insts)
        dense = dense && in.n == insts[0].n && in.n <= 16384 &&
                (double)in.nq * 4.0 >= 0.5 * (double)in.n * (in.n - 1);
    }
    std::cout << "Path: " << (dense ? "dense" : "sparse")
              << (dense ? "" : site_order == "colour" ? " (colour-class order)" : " (sequential order)")
              << std::endl;
    std::vector<double> best_e(I);
    std::vector<std::vector<uint8_t>> best_x(I);
    if (dense) {
      const int32_t n = insts[0].n;
      for (const auto &in : insts)
        if (in.n != n) throw std::runtime_error("the dense path needs instances of one size");
      std::vector<double> q((size_t)I * n * n);
      for (int32_t k = 0; k < I; ++k) flatten(insts[k], q.data() + (size_t)k * n * n);
      os_dense *d = nullptr;
      check(os_dense_batch_create(ctx, I, n, q.data(), &d), "dense batch");
      os_sweep_params p{};
      p.num_replicas = (uint32_t)R;
      p.num_sweeps = (int32_t)num_iter;
      p.tscale = ts.data();
      p.seed = seed;
      p.instance = 0;
      std::vector<double> be((size_t)I * R);
      std::vector<uint8_t> bs((size_t)I * R * n);
      const int rc = os_dense_anneal(d, &p, be.data(), bs.data(), nullptr, nullptr, nullptr,
                                     nullptr);
      os_dense_destroy(d);
      check(rc, "dense anneal");
      for (int32_t k = 0; k < I; ++k) {
        int32_t r = 0;
        for (int32_t s = 1; s < R; ++s)
          if (be[(size_t)k * R + s] < be[(size_t)k * R + r]) r = s;
        best_e[k] = be[(size_t)k * R + r];
        const uint8_t *x = bs.data() + ((size_t)k * R + r) * n;
        best_x[k].assign(x, x + n);
      }
    } else {
      // each instance's rows on a host thread, then the batch arrays in order
      std::vector<std::vector<int32_t>> rps(I), cols(I);
      std::vector<std::vector<double>> vals(I), diags(I);
      std::vector<std::string> cerr_k(I);
      parallel_for((size_t)I, host_threads(), [&](size_t k) {
        try {
          to_csr(insts[k], rps[k], cols[k], vals[k], diags[k]);
        } catch (std::exception &e) {
          cerr_k[k] = e.what();
        }
      });
      for (const auto &e : cerr_k)
        if (!e.empty()) throw std::runtime_error(e);
      std::vector<int32_t> ns, rp, col;
      std::vector<double> val, diag;
      size_t nrp = 0, nnz = 0, nsp = 0;
      for (int32_t k = 0; k < I; ++k) {
        nrp += rps[k].size();
        nnz += cols[k].size();
        nsp += diags[k].size();
      }
      rp.reserve(nrp);
      col.reserve(nnz);
      val.reserve(nnz);
      diag.reserve(nsp);
      for (int32_t k = 0; k < I; ++k) {
        ns.push_back(insts[k].n);
        rp.insert(rp.end(), rps[k].begin(), rps[k].end());
        col.insert(col.end(), cols[k].begin(), cols[k].end());
        val.insert(val.end(), vals[k].begin(), vals[k].end());
        diag.insert(diag.end(), diags[k].begin(), diags[k].end());
      }
      os_csr_batch *b = nullptr;
      check(os_csr_batch_create(ctx, I, ns.data(), rp.data(), col.data(), val.data(), diag.data(),
                                &b),
            "sparse batch");
      os_csr_params p{};
      p.num_replicas = (uint32_t)R;
      p.num_sweeps = (int32_t)num_iter;
      p.tscale = ts.data();
      p.seed = seed;
      p.instance_base = 0;
      p.use_graph = 1;
      p.tune = 0;
      p.site_order = site_order == "colour" ? 1 : 0;
      size_t tot = 0;
      for (int32_t n : ns) tot += (size_t)n * R;
      std::vector<double> be((size_t)I * R);
      std::unique_ptr<uint8_t[]> bs(new uint8_t[tot]);  // filled by the call
      std::vector<double> ib(I);
      std::vector<int32_t> ir(I);
      const int rc = os_csr_batch_anneal(b, &p, be.data(), nullptr, bs.get(), nullptr, ib.data(),
                                         ir.data(), nullptr);
      os_csr_batch_destroy(b);
      check(rc, "sparse anneal");
      size_t off = 0;
      for (int32_t k = 0; k < I; ++k) {
        best_e[k] = ib[k];
        const uint8_t *x = bs.get() + off + (size_t)ir[k] * ns[k];
        best_x[k].assign(x, x + ns[k]);
        off += (size_t)ns[k] * R;
      }
    }
    os_ctx_destroy(ctx);
    const double anneal_ms = ms_since(t0);
    t0 = clock::now();

    // the CSVs on host threads; the log lines, and an error for the first
    // file in order that could not be written, as a serial loop gives them
    std::vector<std::string> werr(I);
    parallel_for((size_t)I, host_threads(), [&](size_t k) {
      const std::string out = output + "/" + stem(insts[k].file) + ".csv";
      std::ofstream f(out);
      if (!f) {
        werr[k] = "can not open output file: " + out;
        return;
      }
      try {
        std::vector<char> x(best_x[k].begin(), best_x[k].end());
        qubo::Solution(x.begin(), x.end(), best_e[k]).save(f);
      } catch (std::exception &e) {
        werr[k] = e.what();
      }
    });
    for (int32_t k = 0; k < I; ++k) {
      if (!werr[k].empty()) {
        std::cout.flush();
        throw std::runtime_error(werr[k]);
      }
      std::cout << insts[k].file << ": n=" << insts[k].n << " energy=" << best_e[k] << '\n';
    }
    std::cout.flush();
    std::cout << "Read and parse: " << parse_ms << " ms" << std::endl;
    std::cout << "Anneal: " << anneal_ms << " ms" << std::endl;
    std::cout << "Write: " << ms_since(t0) << " ms" << std::endl;
  } catch (std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
