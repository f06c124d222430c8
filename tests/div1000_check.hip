// Host check of div1000 (gvv_math.hpp): bit-identical to the IEEE binary32 division x / 1000
// for every finite non-negative binary32 x (the function is odd-symmetric above 2^-120, and
// below that it is the division itself); negative x sampled.  Runs on the host only.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../gvv-differentiable-cuda-renderer_amd/csrc/gvv_math.hpp"

static float f_of(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

int main() {
  const int T = 8;
  std::vector<uint64_t> bad(T, 0);
  std::vector<std::thread> th;
  const uint64_t n = 1ull << 31;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t] {
      for (uint64_t k = n * t / T; k < n * (t + 1) / T; ++k) {
        const float x = f_of((uint32_t)k);
        if (!std::isfinite(x)) continue;
        volatile float ref = x / 1000.f;
        if (u_of(gvv::div1000(x)) != u_of(ref)) ++bad[t];
      }
    });
  for (auto& x : th) x.join();
  uint64_t total = 0;
  for (auto b : bad) total += b;
  uint64_t neg_bad = 0;
  for (uint64_t k = 0; k < (1ull << 31); k += 4099) {
    const float x = f_of((uint32_t)k | 0x80000000u);
    if (!std::isfinite(x)) continue;
    volatile float ref = x / 1000.f;
    if (u_of(gvv::div1000(x)) != u_of(ref)) ++neg_bad;
  }
  std::printf("mismatches %llu negative_sample_mismatches %llu\n", (unsigned long long)total, (unsigned long long)neg_bad);
  return total == 0 && neg_bad == 0 ? 0 : 1;
}
