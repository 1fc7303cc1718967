// Host build of div_by() (device_math.hpp): the reciprocal quotient with one
// correction against the correctly rounded x / h, for spacings dpa's grids use
// (powers of two, 2*pi/(n-1), decimal fractions).  Every bit pattern of
// |x| in [2^-93, 2^-87] (around the switch to plain division) and in
// [h/2, 2h], both signs, plus 4e6 random patterns of [0, 2(n-1)h] per spacing.
// Prints one line per spacing: "h mismatches" and exits non-zero on any.
#include <hip/hip_runtime.h>
#undef __device__
#define __device__ __attribute__((host)) __attribute__((device))
#include "../../dpa_amd/csrc/device_math.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

using namespace dpah;

static uint32_t bits(float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; }
static float from(uint32_t u) { float x; std::memcpy(&x, &u, 4); return x; }

static uint64_t check(float x, float h, float r) {
  return (bits(div_by(x, h, r)) != bits(x / h)) + (bits(div_by(-x, h, r)) != bits(-x / h));
}

int main() {
  const float hs[] = {1.0f, 0.5f, 2.0f, 0.75f, 0.1f, 0.01f, 3.0f,
                      (float)(2.0 * M_PI / 511.0), (float)(2.0 * M_PI / 63.0), 1.0f / 3.0f};
  const int ns[] = {256, 256, 128, 40, 1000, 512, 100, 512, 64, 300};
  std::mt19937_64 rng(12345);
  uint64_t total = 0;
  for (int k = 0; k < (int)(sizeof(hs) / sizeof(hs[0])); ++k) {
    const float h = hs[k], r = 1.0f / h;
    uint64_t bad = 0;
    for (uint32_t u = bits(0x1p-93f); u <= bits(0x1p-87f); ++u) bad += check(from(u), h, r);
    for (uint32_t u = bits(0.5f * h); u <= bits(2.0f * h); ++u) bad += check(from(u), h, r);
    std::uniform_int_distribution<uint32_t> pick(0, bits(2.0f * (float)(ns[k] - 1) * h));
    for (int i = 0; i < 4000000; ++i) bad += check(from(pick(rng)), h, r);
    std::printf("%.9g %llu\n", h, (unsigned long long)bad);
    total += bad;
  }
  // power-of-two spacings: d * (1/h) is d / h exactly (quotient() takes it
  // without any check); every pattern of [h/4, 4h] and the denormal range
  for (float h : {1.0f, 0.5f, 2.0f, 0.25f, 4.0f, 0x1p-20f, 0x1p20f}) {
    const float r = 1.0f / h;
    uint64_t bad = 0;
    auto exact = [&](float x) {
      return (uint64_t)(bits(x * r) != bits(x / h)) + (bits(-x * r) != bits(-x / h));
    };
    for (uint32_t u = bits(0.25f * h); u <= bits(4.0f * h); ++u) bad += exact(from(u));
    for (uint32_t u = 0; u < 0x00800000u; u += 7) bad += exact(from(u));
    std::printf("%.9g %llu exact\n", h, (unsigned long long)bad);
    total += bad;
  }
  return total != 0;
}
