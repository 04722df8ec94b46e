// Writes Philox4x32-10 vectors produced by rocRAND's host-callable engine
// (an implementation independent of this project) as JSON on stdout.
// Built and run by make_philox_fixture.sh; the output is committed as
// philox_rocrand.json and checked against the oracle and the HIP kernels.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

int main() {
  const unsigned long long seeds[] = {0ULL, 1ULL, 0xdeadbeefcafef00dULL, 0xffffffffffffffffULL, 12345ULL};
  const unsigned long long ctr01[] = {0ULL, 1ULL, 7ULL, 0x100000000ULL, 0x3ffffffffffffffeULL};
  const unsigned long long ctr23[] = {0ULL, 3ULL, 0x200000001ULL, 0xffffffffffffffffULL};
  std::printf("[\n");
  bool first = true;
  for (unsigned long long s : seeds)
    for (unsigned long long a : ctr01)
      for (unsigned long long b : ctr23) {
        rocrand_device::philox4x32_10_engine e(s, b, 4ULL * a);
        unsigned int o[4];
        for (int i = 0; i < 4; ++i) o[i] = e();
        std::printf("%s {\"key\": [%u, %u], \"ctr\": [%u, %u, %u, %u], \"out\": [%u, %u, %u, %u]}",
                    first ? "" : ",\n", (unsigned)s, (unsigned)(s >> 32), (unsigned)a, (unsigned)(a >> 32),
                    (unsigned)b, (unsigned)(b >> 32), o[0], o[1], o[2], o[3]);
        first = false;
      }
  std::printf("\n]\n");
  return 0;
}
