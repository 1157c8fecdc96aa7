// Prints rocRAND's philox4x32_10 output words for (seed, offset) pairs read
// from stdin, as an independent known-answer source for the Philox tests.
// Built host-only against the ROCm headers: g++ -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

int main() {
  unsigned long long seed, offset;
  while (std::scanf("%llu %llu", &seed, &offset) == 2) {
    rocrand_device::philox4x32_10_engine e(seed, 0, offset);
    for (int i = 0; i < 4; ++i) std::printf("%u ", (unsigned)e());
    std::printf("\n");
  }
  return 0;
}
