// Host build of the device sampling code: a sample through a warm cell cache
// (kept across steps, with its face-neighbour partial reloads) must equal, bit
// for bit, a sample through an empty cache (full eight-corner load), for steady
// and time-variant fields, over random walks that cross cells in every
// direction.  Prints "samples S shifts F diffs D" and exits non-zero on a diff.
#include <hip/hip_runtime.h>
#undef __device__
#define __device__ __attribute__((host)) __attribute__((device))
#include "../../dpa_amd/csrc/device_math.hpp"

#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

using namespace dpah;

template <bool kUnsteady>
static long walk(const FieldView& f, std::mt19937& rng, long& samples, long& shifts) {
  std::uniform_real_distribution<float> u(-1.0f, 1.0f);
  std::uniform_real_distribution<float> p0(0.0f, (float)(f.n0 - 1)), p1(0.0f, (float)(f.n1 - 1)),
      p2(0.0f, (float)(f.n2 - 1)), pt(0.0f, (float)(f.nt - 1));
  long diffs = 0;
  for (int k = 0; k < 4000; ++k) {
    float x0 = p0(rng), x1 = p1(rng), x2 = p2(rng), t = kUnsteady ? pt(rng) : 0.0f;
    CellCache warm;
    cache_clear(warm);
    const float reach = (k & 1) ? 1.2f : 0.3f;   // cell-crossing and in-cell walks
    for (int s = 0; s < 300; ++s) {
      const int32_t c0 = warm.c0, c1 = warm.c1, c2 = warm.c2, ct = warm.ct;
      V3 a{}, b{};
      CellCache cold;
      cache_clear(cold);
      const bool ok = sample<kUnsteady>(f, x0, x1, x2, t, a, warm);
      const bool ok2 = sample<kUnsteady>(f, x0, x1, x2, t, b, cold);
      if (ok != ok2) { ++diffs; break; }
      if (!ok) break;
      ++samples;
      const int moved = (warm.c0 != c0) + (warm.c1 != c1) + (warm.c2 != c2);
      if (c0 >= 0 && moved == 1 && (!kUnsteady || warm.ct == ct)) ++shifts;
      if (std::memcmp(&a, &b, sizeof a) != 0) ++diffs;
      x0 += reach * u(rng); x1 += reach * u(rng); x2 += reach * u(rng);
      if (kUnsteady) t += 0.05f * reach * u(rng);
    }
  }
  return diffs;
}

int main() {
  const int n0 = 19, n1 = 23, n2 = 29, nt = 5;
  std::vector<float> data(3ull * n0 * n1 * n2 * nt);
  std::mt19937 rng(20240607);
  std::uniform_real_distribution<float> u(-1.0f, 1.0f);
  for (auto& x : data) x = u(rng);
  FieldView f{};
  f.data = data.data();
  f.slice = (int64_t)n0 * n1 * n2;
  f.n0 = n0; f.n1 = n1; f.n2 = n2; f.nt = 1;
  f.o0 = -1.5f; f.o1 = 0.25f; f.o2 = 2.0f; f.to = 0.0f;
  f.h0 = 0.5f; f.h1 = 1.0f; f.h2 = 0.75f; f.th = 1.0f;
  long samples = 0, shifts = 0;
  // positions in cell units: scale the walk into physical coordinates
  FieldView unit = f;
  unit.o0 = unit.o1 = unit.o2 = 0.0f; unit.h0 = unit.h1 = unit.h2 = 1.0f;
  long diffs = walk<false>(unit, rng, samples, shifts);
  diffs += walk<false>(f, rng, samples, shifts);
  FieldView g = unit;
  g.nt = nt;
  diffs += walk<true>(g, rng, samples, shifts);
  std::printf("samples %ld shifts %ld diffs %ld\n", samples, shifts, diffs);
  return diffs != 0;
}
