/*! file devices.hpp
 *
 * Device selection (oneSolver include/helpers/devices.hpp:27-42) and the
 * device handle that replaces the sycl::queue of the reference:
 *   "gpu"  -> an MI355X context of libonesolver.so (os_ctx);
 *   "cpu"  -> the host's cores (one thread per core);
 *   "host" -> the calling thread only (the SYCL host device).
 * Any other string throws std::invalid_argument("Unknown device type: ...").
 */
#ifndef ONESOLVER_DEVICES_HPP_
#define ONESOLVER_DEVICES_HPP_

#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>

#include "../../onesolver.h"

namespace devices {

struct device_selector {
  virtual ~device_selector() = default;
  virtual std::string type() const = 0;
};
struct host_selector : device_selector {
  std::string type() const override { return "host"; }
};
struct cpu_selector : device_selector {
  std::string type() const override { return "cpu"; }
};
struct gpu_selector : device_selector {
  std::string type() const override { return "gpu"; }
};

using device_selector_ptr = std::unique_ptr<device_selector>;

inline device_selector_ptr construct_device_selector(std::string device_type) {
  if (device_type == "cpu") return device_selector_ptr(new cpu_selector());
  if (device_type == "gpu") return device_selector_ptr(new gpu_selector());
  if (device_type == "host") return device_selector_ptr(new host_selector());
  std::ostringstream error_stream;
  error_stream << "Unknown device type: " << device_type;
  throw std::invalid_argument(error_stream.str());
}

/*! A device to run on.  Construction fails with std::runtime_error when no
 *  device of the selected type can be initialised. */
class Device {
 public:
  explicit Device(const device_selector &sel, int gpu_index = 0) : type_(sel.type()) {
    if (type_ == "gpu") {
      int count = 0;
      if (os_device_count(&count) != OS_OK || count <= gpu_index)
        throw std::runtime_error("no GPU device available");
      if (os_ctx_create(gpu_index, &ctx_) != OS_OK)
        throw std::runtime_error(std::string("GPU context creation failed: ") + os_last_error());
      char buf[256];
      os_ctx_device_name(ctx_, buf, sizeof buf);
      name_ = buf;
    } else {
      name_ = cpu_model_name();
      threads_ = type_ == "host" ? 1u : std::max(1u, std::thread::hardware_concurrency());
    }
  }
  ~Device() {
    if (ctx_) os_ctx_destroy(ctx_);
  }
  Device(const Device &) = delete;
  Device &operator=(const Device &) = delete;

  bool is_gpu() const { return ctx_ != nullptr; }
  os_ctx *ctx() const { return ctx_; }
  const std::string &name() const { return name_; }
  const std::string &type() const { return type_; }
  unsigned threads() const { return threads_; }

 private:
  static std::string cpu_model_name() {
    std::ifstream f("/proc/cpuinfo");
    std::string line;
    while (std::getline(f, line)) {
      if (line.rfind("model name", 0) == 0) {
        auto p = line.find(':');
        if (p != std::string::npos) return line.substr(p + 2);
      }
    }
    return "host CPU";
  }
  std::string type_;
  std::string name_;
  os_ctx *ctx_ = nullptr;
  unsigned threads_ = 1;
};

}  // namespace devices

#endif
