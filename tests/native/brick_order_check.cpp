// Host build of lattice_brick_index (the queue order of brick-ordered lattice
// launches): for lattices whose sides are and are not multiples of the brick,
// the map from queue position to C-order lattice index must be a bijection on
// [0, n), visit the bricks in C order of their brick coordinates, give each
// brick one contiguous run of positions, and visit the points inside a brick
// in C order.  The grouped walk (lattice_walk_index, groups of G^3 points
// above the bricks) must be a bijection too, visit the groups in C order, one
// contiguous run each, and inside a group visit its bricks in C order with
// the points of a brick in C order.  Prints "lattices L points P errors E";
// exits non-zero on an error.
#include <hip/hip_runtime.h>
#undef __device__
#define __device__ __attribute__((host)) __attribute__((device))
#include "../../dpa_amd/csrc/advect_kernels.hpp"

#include <cstdio>
#include <vector>

static long check(uint32_t d0, uint32_t d1, uint32_t d2, uint32_t B) {
  const uint64_t n = (uint64_t)d0 * d1 * d2;
  std::vector<uint8_t> seen(n, 0);
  long errors = 0;
  uint64_t prev_brick = 0, prev_inner = 0;
  bool first = true;
  for (uint64_t q = 0; q < n; ++q) {
    const uint64_t idx = dpah::lattice_brick_index(d0, d1, d2, B, (uint32_t)q);
    if (idx >= n || seen[idx]) { ++errors; continue; }
    seen[idx] = 1;
    const uint64_t i2 = idx % d2, i1 = idx / d2 % d1, i0 = idx / d2 / d1;
    const uint64_t nb1 = (d1 + B - 1) / B, nb2 = (d2 + B - 1) / B;
    const uint64_t brick = (i0 / B * nb1 + i1 / B) * nb2 + i2 / B;
    const uint64_t inner = ((i0 % B) * B + i1 % B) * B + i2 % B;
    if (!first) {
      // bricks in C order, each one contiguous run; points in C order inside
      if (brick < prev_brick || (brick == prev_brick && inner <= prev_inner)) ++errors;
    }
    first = false;
    prev_brick = brick;
    prev_inner = inner;
  }
  return errors;
}

static long check_grouped(uint32_t d0, uint32_t d1, uint32_t d2, uint32_t B, uint32_t G) {
  const uint64_t n = (uint64_t)d0 * d1 * d2;
  std::vector<uint8_t> seen(n, 0);
  long errors = 0;
  uint64_t prev_group = 0, prev_brick = 0, prev_inner = 0;
  bool first = true;
  const uint64_t ng1 = (d1 + G - 1) / G, ng2 = (d2 + G - 1) / G;
  const uint64_t nbg = G / B;   // bricks per group edge (G is a multiple of B)
  for (uint64_t q = 0; q < n; ++q) {
    const uint64_t idx = dpah::lattice_walk_index(d0, d1, d2, B, G, (uint32_t)q);
    if (idx >= n || seen[idx]) { ++errors; continue; }
    seen[idx] = 1;
    const uint64_t i2 = idx % d2, i1 = idx / d2 % d1, i0 = idx / d2 / d1;
    const uint64_t group = (i0 / G * ng1 + i1 / G) * ng2 + i2 / G;
    const uint64_t j0 = i0 % G, j1 = i1 % G, j2 = i2 % G;   // inside the group
    const uint64_t brick = (j0 / B * nbg + j1 / B) * nbg + j2 / B;
    const uint64_t inner = ((j0 % B) * B + j1 % B) * B + j2 % B;
    if (!first) {
      if (group < prev_group) ++errors;
      if (group == prev_group &&
          (brick < prev_brick || (brick == prev_brick && inner <= prev_inner)))
        ++errors;
    }
    first = false;
    prev_group = group;
    prev_brick = brick;
    prev_inner = inner;
  }
  return errors;
}

int main() {
  const uint32_t shapes[][3] = {{4, 4, 4},   {1, 1, 1},  {5, 7, 9},    {3, 1, 8},
                                {8, 12, 4},  {17, 33, 65}, {4, 5, 4},  {9, 4, 13},
                                {128, 512, 512}};
  long errors = 0, points = 0, lattices = 0;
  for (const auto& s : shapes) {
    for (uint32_t B : {4u, 3u, 1u}) {
      errors += check(s[0], s[1], s[2], B);
      points += (long)s[0] * s[1] * s[2];
      ++lattices;
    }
  }
  const uint32_t grouped[][5] = {{17, 33, 65, 4, 8},  {128, 512, 512, 4, 32},
                                 {128, 512, 512, 2, 16}, {5, 7, 9, 2, 4},
                                 {9, 4, 13, 4, 16},   {64, 64, 64, 4, 64},
                                 {1, 1, 1, 4, 8},     {33, 17, 40, 1, 3}};
  for (const auto& g : grouped) {
    errors += check_grouped(g[0], g[1], g[2], g[3], g[4]);
    points += (long)g[0] * g[1] * g[2];
    ++lattices;
  }
  std::printf("lattices %ld points %ld errors %ld\n", lattices, points, errors);
  return errors != 0;
}
