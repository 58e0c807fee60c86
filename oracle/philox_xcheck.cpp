// Host-side driver around rocRAND's open-source Philox4x32-10 engine
// (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h).  Test infrastructure:
// reads "c0 c1 c2 c3 k0 k1" hex lines on stdin and prints the ten-round output,
// so tests can compare the oracle's and the HIP kernel's Philox with it.
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

struct Probe : rocrand_device::philox4x32_10_engine {
  uint4 rounds(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
  unsigned c0, c1, c2, c3, k0, k1;
  Probe p;
  while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
    uint4 c = {c0, c1, c2, c3};
    uint2 k = {k0, k1};
    uint4 r = p.rounds(c, k);
    std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
  }
  return 0;
}
