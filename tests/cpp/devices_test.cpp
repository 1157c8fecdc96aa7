// Selector mapping of oneSolver tests/devices_test.cpp:14-30 (no device needed).
#include "harness.hpp"
#include "onesolver/helpers/devices.hpp"

TEST_CASE(devices_test, specifying_host_as_device_type_constructs_host_selector) {
  auto p = devices::construct_device_selector("host");
  REQUIRE(dynamic_cast<devices::host_selector *>(p.get()) != nullptr);
}
TEST_CASE(devices_test, specifying_cpu_as_device_type_constructs_cpu_selector) {
  auto p = devices::construct_device_selector("cpu");
  REQUIRE(dynamic_cast<devices::cpu_selector *>(p.get()) != nullptr);
}
TEST_CASE(devices_test, specifying_gpu_as_device_type_constructs_gpu_selector) {
  auto p = devices::construct_device_selector("gpu");
  REQUIRE(dynamic_cast<devices::gpu_selector *>(p.get()) != nullptr);
}
TEST_CASE(devices_test, unknown_device_type_throws) {
  CHECK_THROW(devices::construct_device_selector("tpu"), std::invalid_argument);
}
