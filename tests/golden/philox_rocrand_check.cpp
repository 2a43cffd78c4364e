// Prints Philox4x32-10 outputs computed with rocRAND's open-source engine
// (rocrand_philox4x32_10.h) for the counters/keys given on stdin, one
// "c0 c1 c2 c3 k0 k1" hex line each.  Used by tests/test_oracle_pinning.py to
// cross-check the oracle's and the engine's Philox against a library.
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

struct Probe : rocrand_device::philox4x32_10_engine {
  uint4 run(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};

int main() {
  unsigned c0, c1, c2, c3, k0, k1;
  Probe p;
  while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
    uint4 r = p.run(uint4{c0, c1, c2, c3}, uint2{k0, k1});
    std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
  }
  return 0;
}
