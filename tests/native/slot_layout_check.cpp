// Host build of the record-slot layout (slot_float, slot_offset,
// slot_region_floats, tile_slot_float in advect_kernels.hpp), for slot shifts
// 0 and 3 and strides that are and are not multiples of a block:
//   - shift 0 is the one-slot-run layout 3 ((r / 64 * stride + j) * 64 + r % 64);
//   - every slot of every row of a region of `rows` rows maps to a distinct
//     float triple inside slot_region_floats, and no two slots overlap;
//   - a row's slot j sits slot_offset(j) past its slot 0;
//   - shift 3: every block starts on a 16-byte boundary (whole-block stores);
//   - the compaction's chunk offsets (tile_slot_float from block b0) land on
//     slot_float.
// Prints "layouts L slots S errors E"; exits non-zero on an error.
#include <hip/hip_runtime.h>
#undef __device__
#define __device__ __attribute__((host)) __attribute__((device))
#include "../../dpa_amd/csrc/advect_kernels.hpp"

#include <cstdio>
#include <vector>

using namespace dpah;

static long check(uint64_t rows, uint64_t stride, uint32_t g, long& slots) {
  long errors = 0;
  const uint64_t region = slot_region_floats(rows, stride, g);
  std::vector<uint8_t> used(region, 0);
  for (uint64_t r = 0; r < tiled_rows(rows); ++r) {
    const uint64_t base = slot_float(r, stride, 0, g);
    const uint64_t tile = slot_float(r / kTileRows * kTileRows, stride, 0, g);
    for (uint64_t j = 0; j < stride; ++j) {
      ++slots;
      const uint64_t f = slot_float(r, stride, j, g);
      if (g == 0 && f != 3 * ((r / kTileRows * stride + j) * kTileRows + r % kTileRows)) ++errors;
      if (f != base + slot_offset(j, g)) ++errors;
      if (f + 3 > region) { ++errors; continue; }
      for (int c = 0; c < 3; ++c) {
        if (used[f + c]) ++errors;
        used[f + c] = 1;
      }
      if (g == kStagedShift && (j & (kSlotGroup - 1)) == 0 && f % 4 != 0) ++errors;
      for (uint32_t b0 : {0u, (uint32_t)(j >> g)}) {   // chunk starts at or before j's block
        const uint64_t from = tile + (uint64_t)b0 * (kSlotPitch << g);
        if (from + tile_slot_float((uint32_t)j, b0, (uint32_t)(r % kTileRows), g) != f) ++errors;
      }
    }
  }
  return errors;
}

int main() {
  long errors = 0, slots = 0, layouts = 0;
  for (uint32_t g : {0u, kStagedShift})
    for (uint64_t rows : {1ull, 63ull, 64ull, 65ull, 200ull})
      for (uint64_t stride : {1ull, 2ull, 7ull, 8ull, 9ull, 63ull, 64ull, 258ull}) {
        errors += check(rows, stride, g, slots);
        ++layouts;
      }
  std::printf("layouts %ld slots %ld errors %ld\n", layouts, slots, errors);
  return errors != 0;
}
