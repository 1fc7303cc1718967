// Host build of the device sampling code: a sample through a warm cell cache
// (kept across steps) must equal, bit for bit, a sample through an empty cache
// (full eight-corner load from the C-order nodes) that divides by the
// spacings, for steady and time-variant fields, over random walks that cross
// cells in every direction, with every layout a reload can read: C-order
// nodes, cell records (C order and bricks), node bricks.
// The warm sample runs each division mode a launch can take: plain division,
// the reciprocal with one exact-residual correction (kDivFast, on power-of-two
// and on a non-power-of-two grid), and on grids whose spacings are all powers
// of two the reciprocal multiply with fmod as one fmaf (kDivExact); so every
// mode is pinned against the correctly rounded divisions and fmodf, also at
// boundary probes (signed zeros, denormals, faces, the last cell).  Each
// axis's floor and weight are also pinned directly to floorf and fmodf, and
// kDivFast's fast path with its exact branch on spacings (0.1, 1/3, 2 pi / 511)
// where rounded quotients land on the next integer at ordinary offsets.
// Prints "samples S shifts F roundups R diffs D" and exits non-zero on a diff.
#include <hip/hip_runtime.h>
#undef __device__
#define __device__ __attribute__((host)) __attribute__((device))
#include "../../dpa_amd/csrc/device_math.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

using namespace dpah;

template <bool kUnsteady, int kDiv>
static long walk(const FieldView& f, std::mt19937& rng, long& samples, long& shifts) {
  // the cold reference divides on every axis and reads the nodes through
  // 64-bit pointer offsets (no cell layout, no buffer offsets)
  FieldView plain = f;
  plain.cells = nullptr;
  plain.nbrick = nullptr;
  plain.idx32 = 0;
  plain.node_bytes32 = plain.cell_bytes32 = 0;
  std::uniform_real_distribution<float> u(-1.0f, 1.0f);
  std::uniform_real_distribution<float> p0(0.0f, (float)(f.n0 - 1)), p1(0.0f, (float)(f.n1 - 1)),
      p2(0.0f, (float)(f.n2 - 1)), pt(0.0f, (float)(f.nt - 1));
  long diffs = 0;
  for (int k = 0; k < 4000; ++k) {
    float x0 = f.o0 + f.h0 * p0(rng), x1 = f.o1 + f.h1 * p1(rng), x2 = f.o2 + f.h2 * p2(rng);
    float t = kUnsteady ? f.to + f.th * pt(rng) : 0.0f;
    CellCache warm;
    cache_clear(warm);
    const float reach = (k & 1) ? 1.2f : 0.3f;   // cell-crossing and in-cell walks
    for (int s = 0; s < 300; ++s) {
      const float c0 = warm.s0, c1 = warm.s1, c2 = warm.s2, ct = warm.st;
      V3 a{}, b{};
      CellCache cold;
      cache_clear(cold);
      const bool ok = sample<kUnsteady, kDiv>(f, x0, x1, x2, t, a, warm);
      const bool ok2 = sample<kUnsteady, kDivide>(plain, x0, x1, x2, t, b, cold);
      if (ok != ok2) { ++diffs; break; }
      if (!ok) break;
      ++samples;
      const int moved = (warm.s0 != c0) + (warm.s1 != c1) + (warm.s2 != c2);
      if (c0 >= 0 && moved == 1 && (!kUnsteady || warm.st == ct)) ++shifts;
      if (std::memcmp(&a, &b, sizeof a) != 0) ++diffs;
      x0 += reach * f.h0 * u(rng); x1 += reach * f.h1 * u(rng); x2 += reach * f.h2 * u(rng);
      if (kUnsteady) t += 0.05f * reach * f.th * u(rng);
    }
  }
  return diffs;
}

// Positions at the edges of the power-of-two path, one axis at a time with the
// others mid-grid: zeros of both signs, denormals, tiny negatives, exact cell
// faces, the last valid cell and just past it.  Returns the number of samples
// (inside flag or value bits) that differ from the dividing path.
template <int kDiv>
static long probes(const FieldView& e, long& samples) {
  FieldView plain = e;   // dividing, pointer reads of the nodes
  plain.cells = nullptr;
  plain.nbrick = nullptr;
  plain.idx32 = 0;
  plain.node_bytes32 = plain.cell_bytes32 = 0;
  const float tiny[] = {0.0f, -0.0f, 1e-45f, -1e-45f, 1e-38f, -1e-38f, 1e-30f, -1e-30f,
                        -1e-7f, 1e-7f};
  long diffs = 0;
  for (int ax = 0; ax < 3; ++ax) {
    const float o = ax == 0 ? e.o0 : ax == 1 ? e.o1 : e.o2;
    const float h = ax == 0 ? e.h0 : ax == 1 ? e.h1 : e.h2;
    const int32_t n = ax == 0 ? e.n0 : ax == 1 ? e.n1 : e.n2;
    std::vector<float> xs;
    for (float d : tiny) xs.push_back(o + d);
    for (int k = 0; k < n; ++k) {
      const float face = o + h * (float)k;
      xs.push_back(face);
      xs.push_back(std::nextafter(face, -1e30f));
      xs.push_back(std::nextafter(face, 1e30f));
    }
    for (float x : xs) {
      float p[3] = {e.o0 + e.h0 * (e.n0 - 1) * 0.5f, e.o1 + e.h1 * (e.n1 - 1) * 0.5f,
                    e.o2 + e.h2 * (e.n2 - 1) * 0.5f};
      p[ax] = x;
      V3 a{}, b{};
      CellCache c1, c2;
      cache_clear(c1);
      cache_clear(c2);
      const bool ok = sample<false, kDiv>(e, p[0], p[1], p[2], 0.0f, a, c1);
      const bool ok2 = sample<false, kDivide>(plain, p[0], p[1], p[2], 0.0f, b, c2);
      ++samples;
      if (ok != ok2 || (ok && std::memcmp(&a, &b, sizeof a) != 0)) ++diffs;
    }
  }
  return diffs;
}

// One axis's floor and weight against libm, as regular_grid.hpp:61-62 computes
// them: s = floorf(d / h) and w = fmodf(d, h) / h, bit for bit (signed zeros
// included), in division mode kDiv, for boundary offsets d (signed zeros,
// denormals, tiny values, every face and its neighbours) and random ones in
// [-h, n h).  Weights are compared where the cell is inside, as only there are
// they used.
template <int kDiv>
static long axis_probes(float o, float h, int32_t n, std::mt19937& rng, long& samples) {
  const float rh = 1.0f / h;
  const float tiny[] = {0.0f, -0.0f, 1e-45f, -1e-45f, 1e-38f, -1e-38f, 1e-30f, -1e-30f,
                        -1e-7f, 1e-7f};
  std::vector<float> xs;
  for (float d : tiny) xs.push_back(o + d);
  for (int k = 0; k < n; ++k) {
    const float face = o + h * (float)k;
    xs.push_back(face);
    xs.push_back(std::nextafter(face, -1e30f));
    xs.push_back(std::nextafter(face, 1e30f));
  }
  std::uniform_real_distribution<float> u(-h, h * (float)n);
  for (int k = 0; k < 20000; ++k) xs.push_back(o + u(rng));
  long diffs = 0;
  for (float x : xs) {
    for (float d : {x - o, -0.0f, 0.0f}) {
      const float s = axis_floor<kDiv>(d, h, rh);
      const float ref_s = floorf(d / h);
      ++samples;
      if (std::memcmp(&s, &ref_s, 4) != 0) ++diffs;
      if (!axis_inside(ref_s, n)) continue;
      const float w = axis_weight<kDiv>(d, s, h, rh);
      const float ref_w = fmodf(d, h) / h;
      if (std::memcmp(&w, &ref_w, 4) != 0) ++diffs;
    }
  }
  return diffs;
}

// kDivFast's location as sample() runs it (locate_fast: the fast path, the
// exact branch for tiny offsets and inside cells with a rounded-up quotient), against floorf(d / h) and fmodf(d, h) / h.  Offsets are
// the boundary set of axis_probes plus, for every k < n, the floats nearest
// below k h whose rounded quotient is k (the floor of the rounded quotient is
// one above the exact floor there: on spacings such as 0.1, 1/3 or 2 pi / 511
// this happens at ordinary positions like d = 0.5).  `roundups` counts those
// offsets, so a run shows that it met them.
static long split_probes(float h, int32_t n, std::mt19937& rng, long& samples, long& roundups) {
  // the same offset on all three axes (and time) of a grid with spacing h
  FieldView g{};
  g.n0 = g.n1 = g.n2 = g.nt = n;
  g.h0 = g.h1 = g.h2 = g.th = h;
  g.r0 = g.r1 = g.r2 = g.rt = 1.0f / h;
  const float tiny[] = {0.0f, -0.0f, 1e-45f, -1e-45f, 1e-38f, -1e-38f, 1e-30f, -1e-30f,
                        -1e-7f, 1e-7f};
  std::vector<float> xs(std::begin(tiny), std::end(tiny));
  for (int k = 0; k < n; ++k) {
    const double exact = (double)k * (double)h;   // exact: a float times a small integer
    float d = (float)exact;
    for (int j = 0; j < 4; ++j) d = std::nextafter(d, -1e30f);
    for (int j = 0; j < 9; ++j) {
      xs.push_back(d);
      if ((double)d < exact && floorf(d / h) >= (float)k) ++roundups;
      d = std::nextafter(d, 1e30f);
    }
  }
  std::uniform_real_distribution<float> u(-h, h * (float)n);
  for (int k = 0; k < 20000; ++k) xs.push_back(u(rng));
  long diffs = 0;
  for (float d : xs) {
    Axis a[4];
    const uint32_t in = locate_fast<true>(g, d, d, d, d, a[0], a[1], a[2], a[3]);
    const float ref_s = floorf(d / h);
    ++samples;
    if (in != axis_inside(ref_s, n)) ++diffs;
    for (const Axis& x : a) {
      if (std::memcmp(&x.s, &ref_s, 4) != 0) ++diffs;
      if (!axis_inside(ref_s, n)) continue;
      const float ref_w = fmodf(d, h) / h;
      if (std::memcmp(&x.w, &ref_w, 4) != 0) ++diffs;
    }
  }
  return diffs;
}

// The cell layout of a steady field as the library's k_build_cells writes it:
// record cell_record(c0, c1, c2) (C order, or bricks of 2^k cells per axis)
// holds the eight corners in order 4a + 2b + c, 24 floats padded to 32.
static std::vector<float4> build_cells(FieldView& f, const std::vector<float>& nodes, uint32_t k) {
  auto bricks = [&](int32_t n) { return (int32_t)(((uint32_t)n - 1 + (1u << k) - 1) >> k); };
  f.cell_bits = k;
  f.cb1 = k ? bricks(f.n1) : 0;
  f.cb2 = k ? bricks(f.n2) : 0;
  f.cell_slice = k ? ((int64_t)bricks(f.n0) * f.cb1 * f.cb2) << (3 * k)
                   : (int64_t)(f.n0 - 1) * (f.n1 - 1) * (f.n2 - 1);
  std::vector<float4> cells(8 * (size_t)f.cell_slice, make_float4(0, 0, 0, 0));
  for (int32_t c0 = 0; c0 + 1 < f.n0; ++c0)
    for (int32_t c1 = 0; c1 + 1 < f.n1; ++c1)
      for (int32_t c2 = 0; c2 + 1 < f.n2; ++c2) {
        float* rec = (float*)&cells[8 * (size_t)cell_record(f, 0, c0, c1, c2)];
        for (int q = 0; q < 8; ++q) {
          const size_t node = ((size_t)(c0 + (q >> 2)) * f.n1 + (c1 + ((q >> 1) & 1))) * f.n2 +
                              (c2 + (q & 1));
          for (int c = 0; c < 3; ++c) rec[3 * q + c] = nodes[3 * node + c];
        }
      }
  return cells;
}

// The node bricks of a field as the library's k_build_nbrick writes them:
// bricks of 4^3 nodes (768 bytes) in C order over the bricks, nodes in C order
// inside a brick, one slice after another; padding nodes left at zero.
static std::vector<float> build_nbrick(FieldView& f, const std::vector<float>& nodes) {
  const uint32_t nb0 = ((uint32_t)f.n0 + 3) / 4, nb1 = ((uint32_t)f.n1 + 3) / 4,
                 nb2 = ((uint32_t)f.n2 + 3) / 4;
  f.nb1 = nb1;
  f.nb2 = nb2;
  f.nbrick_slice = nb0 * nb1 * nb2 * 768u;
  std::vector<float> out((size_t)f.nbrick_slice / 4 * (size_t)f.nt, 0.0f);
  for (int32_t t = 0; t < f.nt; ++t)
    for (int32_t i0 = 0; i0 < f.n0; ++i0)
      for (int32_t i1 = 0; i1 < f.n1; ++i1)
        for (int32_t i2 = 0; i2 < f.n2; ++i2) {
          const size_t brick = ((size_t)(i0 / 4) * nb1 + i1 / 4) * nb2 + i2 / 4;
          const size_t at = (size_t)t * f.nbrick_slice / 4 +
                            3 * (brick * 64 + (size_t)(i0 % 4) * 16 + (i1 % 4) * 4 + i2 % 4);
          const size_t node = (((size_t)t * f.n0 + i0) * f.n1 + i1) * f.n2 + i2;
          for (int c = 0; c < 3; ++c) out[at + c] = nodes[3 * node + c];
        }
  return out;
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
  f.r0 = 1.0f / f.h0; f.r1 = 1.0f / f.h1; f.r2 = 1.0f / f.h2; f.rt = 1.0f / f.th;
  // positions in cell units: scale the walk into physical coordinates
  FieldView unit = f;
  unit.o0 = unit.o1 = unit.o2 = 0.0f; unit.h0 = unit.h1 = unit.h2 = 1.0f;
  unit.r0 = unit.r1 = unit.r2 = 1.0f;
  long diffs = 0;
  for (float h : {0.5f, 1.0f, 0.75f, 2.0f, 0.25f}) {   // spacings the walks below use
    diffs += axis_probes<kDivide>(0.0f, h, 23, rng, samples);
    diffs += axis_probes<kDivide>(-1.5f, h, 23, rng, samples);
    diffs += axis_probes<kDivFast>(0.0f, h, 23, rng, samples);
    diffs += axis_probes<kDivFast>(2.0f, h, 23, rng, samples);
  }
  for (float h : {0.5f, 1.0f, 2.0f, 0.25f}) {          // powers of two
    diffs += axis_probes<kDivExact>(0.0f, h, 23, rng, samples);
    diffs += axis_probes<kDivExact>(0.25f, h, 23, rng, samples);
  }
  diffs += walk<false, kDivide>(unit, rng, samples, shifts);
  diffs += walk<false, kDivide>(f, rng, samples, shifts);
  diffs += walk<false, kDivFast>(f, rng, samples, shifts);   // 0.75 is not a power of two
  FieldView g = unit;
  g.nt = nt;
  diffs += walk<true, kDivide>(g, rng, samples, shifts);
  diffs += walk<true, kDivFast>(g, rng, samples, shifts);
  // power-of-two spacings (0.5, 2, 0.25, time 0.5)
  FieldView e = f;
  e.h0 = 0.5f; e.h1 = 2.0f; e.h2 = 0.25f; e.th = 0.5f;
  e.r0 = 2.0f; e.r1 = 0.5f; e.r2 = 4.0f; e.rt = 2.0f;
  diffs += walk<false, kDivExact>(unit, rng, samples, shifts);
  diffs += walk<false, kDivExact>(e, rng, samples, shifts);
  diffs += walk<false, kDivFast>(e, rng, samples, shifts);
  FieldView et = e;
  et.nt = nt;
  diffs += walk<true, kDivExact>(et, rng, samples, shifts);
  diffs += probes<kDivExact>(e, samples);
  diffs += probes<kDivExact>(unit, samples);
  diffs += probes<kDivFast>(e, samples);
  diffs += probes<kDivFast>(f, samples);
  // zero origins, so that the tiny probes reach the divisions unabsorbed (a
  // tiny negative d whose quotient underflows to -0 stays inside and keeps d
  // as its fmod)
  FieldView e0 = e, f0 = f;
  e0.o0 = e0.o1 = e0.o2 = 0.0f;
  f0.o0 = f0.o1 = f0.o2 = 0.0f;
  diffs += probes<kDivExact>(e0, samples);
  diffs += probes<kDivFast>(e0, samples);
  diffs += probes<kDivFast>(f0, samples);
  diffs += probes<kDivFast>(unit, samples);
  // spacings whose rounded quotients land on the next integer at ordinary
  // offsets (the verified-reciprocal spacings of the tests' fields); the
  // kDivFast sample on such a grid against the dividing one, at its faces
  // and their neighbours
  long roundups = 0;
  for (float h : {0.1f, 1.0f / 3.0f, (float)(2.0 * M_PI / 511.0), (float)(2.0 * M_PI / 63.0),
                  0.75f, 3.0f})
    diffs += split_probes(h, 200, rng, samples, roundups);
  FieldView r = f0;
  r.h0 = 0.1f; r.h1 = 1.0f / 3.0f; r.h2 = (float)(2.0 * M_PI / 511.0);
  r.r0 = 1.0f / r.h0; r.r1 = 1.0f / r.h1; r.r2 = 1.0f / r.h2;
  diffs += probes<kDivFast>(r, samples);
  diffs += walk<false, kDivFast>(r, rng, samples, shifts);
  // 32-bit reloads (24-bit offset arithmetic into a buffer resource or from a
  // uniform base, and 32-bit float / record offsets) against the 64-bit
  // pointer path: the node array of steady and time-variant fields, and cell
  // layouts in C order and in bricks of 2 and 8 cells
  {
    FieldView b = f;
    b.node_bytes32 = (uint32_t)(data.size() * sizeof(float));
    diffs += walk<false, kDivFast>(b, rng, samples, shifts);
    diffs += walk<false, kDivide>(b, rng, samples, shifts);
    FieldView bt = g;
    bt.node_bytes32 = (uint32_t)(data.size() * sizeof(float));
    diffs += walk<true, kDivide>(bt, rng, samples, shifts);
    diffs += walk<true, kDivFast>(bt, rng, samples, shifts);
    for (uint32_t k : {0u, 1u, 3u}) {
      FieldView c = f;
      const std::vector<float4> cells = build_cells(c, data, k);
      c.cells = cells.data();
      c.cell_bytes32 = (uint32_t)(cells.size() * sizeof(float4));
      diffs += walk<false, kDivFast>(c, rng, samples, shifts);
      diffs += probes<kDivFast>(c, samples);
      // the same layout through 32-bit record indices (fields past 2^32 bytes)
      c.cell_bytes32 = 0;
      c.idx32 = 1;
      diffs += walk<false, kDivFast>(c, rng, samples, shifts);
    }
    // node bricks (4^3 nodes per brick; the grid's sides are not multiples of
    // 4, so cells at the brick faces, the padded bricks and the last cell are
    // all walked), steady and time-variant
    {
      FieldView nb = f;
      const std::vector<float> bricks = build_nbrick(nb, data);
      nb.nbrick = (const char*)bricks.data();
      diffs += walk<false, kDivFast>(nb, rng, samples, shifts);
      diffs += walk<false, kDivide>(nb, rng, samples, shifts);
      diffs += probes<kDivFast>(nb, samples);
      FieldView nbt = g;
      const std::vector<float> tbricks = build_nbrick(nbt, data);
      nbt.nbrick = (const char*)tbricks.data();
      diffs += walk<true, kDivide>(nbt, rng, samples, shifts);
      diffs += walk<true, kDivFast>(nbt, rng, samples, shifts);
    }
    // nodes through 32-bit float offsets, steady and time-variant
    FieldView i32 = f;
    i32.idx32 = 1;
    diffs += walk<false, kDivFast>(i32, rng, samples, shifts);
    FieldView i32t = g;
    i32t.idx32 = 1;
    diffs += walk<true, kDivFast>(i32t, rng, samples, shifts);
  }
  std::printf("samples %ld shifts %ld roundups %ld diffs %ld\n", samples, shifts, roundups, diffs);
  if (roundups < 100) {
    std::printf("too few offsets whose rounded quotient is one above the floor\n");
    return 1;
  }
  return diffs != 0;
}
