// device_math.hpp -- per-step arithmetic of dpa's advection stage for gfx950.
//
// Every operation is one IEEE binary32 rounding, in the order dpa's CPU build
// performs it; the library is compiled with -ffp-contract=off and correctly
// rounded division/sqrt, so no multiply-add is fused behind our back.  The
// only fused operation is the explicit fmaf in exact_fmod(), whose result is
// exact (see there).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dpah {

// A regular float3 grid resident in HBM (regular_grid.hpp:148-151):
// node (i0,i1,i2) of time slice it lives at data + 3*(it*slice + (i0*n1+i1)*n2+i2).
// Axis 0 is the slowest; nt == 1 for steady fields.
struct FieldView {
  const float* data;
  int64_t slice;     // n0*n1*n2
  int32_t n0, n1, n2, nt;
  float o0, o1, o2, to;
  float h0, h1, h2, th;
  // Optional cell layout: record ((it*(n0-1) + c0)*(n1-1) + c1)*(n2-1) + c2 holds
  // the eight corners of cell (c0, c1, c2) of slice it as 24 floats in corner
  // order 4a+2b+c, padded to 128 B, so a reload reads one aligned line instead
  // of four rows.  The values are copies of the nodes; null when absent.
  const float4* cells;
  int64_t cell_slice;  // records per time slice
  // Record order of the cell layout: 0 = C order over (c0, c1, c2); k > 0 =
  // bricks of 2^k cells per axis stored one after another (bricks in C order,
  // cells in C order inside a brick), so that the cells a particle crosses
  // next, and those of its neighbours, share DRAM pages and address
  // translations.  cb1, cb2: bricks along axes 1 and 2 (the cell grid is
  // padded to whole bricks; padding records are never read).
  uint32_t cell_bits;
  int32_t cb1, cb2;
  // Reciprocals of the spacings and, per axis (bit 0..2, bit 3 = time), whether
  // div_by() with that reciprocal was verified at field creation to equal the
  // correctly rounded quotient for every float in the axis's range.
  float r0, r1, r2, rt;
  uint32_t fastdiv;
  // Per axis (same bits): the spacing is a power of two, so its reciprocal is
  // exact and d * r is the same real number as d / h -- the same rounding, for
  // every d, with no check needed.
  uint32_t exactdiv;
  // 1: every node's float offset (all slices) and every cell record's float4
  // index fit 32 bits, so a reload forms its offset in 32-bit arithmetic and
  // adds it to the base once
  uint32_t idx32;
};

enum Status : uint8_t { kActive = 0, kCompleted = 1, kExited = 2, kZero = 3 };
enum Integrator : int { kDpaRk4 = 0, kRk4 = 1 };

struct Axis {
  float w;   // fmod(d, h) / h
  float s;   // floor(d / h): the cell's start index, an integral float
};

// fmodf(d, h) for d >= 0 (or -0) and h > 0, given s = floorf(fl(d / h)).
// fl(d/h) is monotone in the exact quotient, so s is the exact floor k or k+1.
// When s == k the exact remainder d - k*h is representable (it is fmod's
// result), so one fmaf returns it without rounding; when s == k+1 the fmaf
// result is negative (rounding keeps the sign of a non-zero value) and the
// fmaf with k is exact.  s = +-0 with d < 0 (a tiny negative d whose quotient
// underflowed to -0) keeps d, as fmodf does.  Both candidates are computed and
// one selected: no branch.  fmodf's result carries d's sign, also when it is
// zero: for d = -0, s = -0, and -s * h + d would be +0 + -0 = +0, so the
// product is formed as -|s| * h (= -0 * h = -0, and -0 + -0 = -0; for every
// s >= 0 it is -s * h): fmodf(-0, h) = -0, and the weight -0 / h = -0 matters
// wherever it multiplies a corner value of -0.  The |.| and the negation are
// operand modifiers of the fma: no extra instruction.
__device__ __forceinline__ float exact_fmod(float d, float h, float s) {
  const float r = __builtin_fmaf(-__builtin_fabsf(s), h, d);
  const float k = s - 1.0f;
  const float r1 = __builtin_fmaf(-k, h, d);
  return (r < 0.0f && s != 0.0f) ? r1 : r;
}

// d / h from r = fl(1/h): q0 = fl(d*r), the residual d - q0*h is exact in one
// fmaf, and one more fmaf corrects q0 -- three instructions where a correctly
// rounded division takes about twelve.  Below |d| = 2^-90 the residual could
// underflow and lose exactness, so zeros, tiny values and NaN divide as usual.
// It is used only on axes where the field's creation checked it, bit for bit,
// against d / h for every float d in [-B, B], B = 2 (n-1) h
// (verify_division_kernel); outside that range the quotient only decides
// "outside the grid", and both forms agree on that (same sign, magnitude
// beyond n-1).
__device__ __forceinline__ bool div_tiny(float d) { return !(__builtin_fabsf(d) >= 0x1p-90f); }

__device__ __forceinline__ float div_by_fast(float d, float h, float r) {
  const float q0 = d * r;
  const float e = __builtin_fmaf(-q0, h, d);
  return __builtin_fmaf(e, r, q0);
}

__device__ __forceinline__ float div_by(float d, float h, float r) {
  if (div_tiny(d)) return d / h;
  return div_by_fast(d, h, r);
}

// div_by on three axes with one test for the rare operands: the three fast
// quotients, then one branch (taken by no lane in practice) that divides the
// axes whose operand is tiny, zero or NaN -- the same bits as three div_by
// calls, with one divergent branch instead of three.
__device__ __forceinline__ void div_by3(float d0, float h0, float r0, float d1, float h1, float r1,
                                        float d2, float h2, float r2, float& q0, float& q1,
                                        float& q2) {
  q0 = div_by_fast(d0, h0, r0);
  q1 = div_by_fast(d1, h1, r1);
  q2 = div_by_fast(d2, h2, r2);
  const bool t0 = div_tiny(d0), t1 = div_tiny(d1), t2 = div_tiny(d2);
  if (t0 || t1 || t2) {
    if (t0) q0 = d0 / h0;
    if (t1) q1 = d1 / h1;
    if (t2) q2 = d2 / h2;
  }
}

// Counts the floats x in [-B, B] (bit patterns 0..hi of |x|, both signs) where
// div_by(x, h, r) and x / h differ in any bit.  Grid-stride over 2^32-bounded
// indices; one atomic per mismatch.
static __global__ __launch_bounds__(256) void verify_division_kernel(float h, float r, uint32_t hi,
                                                                     unsigned long long* bad) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= hi; i += stride) {
    const float x = __uint_as_float((uint32_t)i);
    const float qp = x / h, qn = (-x) / h;
    const bool ok = __float_as_uint(div_by(x, h, r)) == __float_as_uint(qp) &&
                    __float_as_uint(div_by(-x, h, r)) == __float_as_uint(qn);
    if (!ok) atomicAdd(bad, 1ull);
  }
}

// How a launch divides by the grid spacings, chosen per launch from the
// field's per-axis checks (FieldView::exactdiv / fastdiv) over every axis it
// locates on (time too, for a time-variant field):
//   kDivExact  every spacing is a power of two: its reciprocal is exact, so
//              d * r is the same real number as d / h and rounds the same;
//   kDivFast   every axis is a power of two or passed verify_division_kernel:
//              div_by() (on a power-of-two axis its correction is zero);
//   kDivide    otherwise (or with the tuning's fastdiv off): correctly rounded
//              d / h.
// The mode is a template parameter, so the step loop reads no per-axis flags.
enum DivMode : int { kDivide = 0, kDivExact = 1, kDivFast = 2 };

__host__ __device__ inline int div_mode(uint32_t exactdiv, uint32_t fastdiv, bool unsteady) {
  const uint32_t axes = unsteady ? 15u : 7u;
  if ((exactdiv & axes) == axes) return kDivExact;
  if (((exactdiv | fastdiv) & axes) == axes) return kDivFast;
  return kDivide;
}

template <int kDiv>
__device__ __forceinline__ float div_spacing(float d, float h, float rh) {
  if constexpr (kDiv == kDivExact) return d * rh;
  else if constexpr (kDiv == kDivFast) return div_by(d, h, rh);
  else return d / h;
}

// One axis of regular_grid::contains (regular_grid.hpp:41-50) and of the
// weight / start index of regular_grid::interpolate (52-75), split so that a
// sample computes every axis's floor, tests them together and computes the
// weights only inside.
//
// kDivExact: s = floor(d * rh) is the exact floor of the quotient (or 0 / -0
// for a quotient that underflows, where |d| < h), and fmaf(-s, h, d) is
// d - s*h exactly: fmod's result, with -0 and tiny negative d returned
// unchanged as fmodf returns them.  The other modes take fmod from
// exact_fmod(), which also handles a rounded quotient one above the floor.
template <int kDiv>
__device__ __forceinline__ float axis_floor(float d, float h, float rh) {
  return floorf(div_spacing<kDiv>(d, h, rh));
}

// 1 when 0 <= s <= n - 2, else 0 (also for NaN); an integer so that the axes
// combine with & and no branch.
__device__ __forceinline__ uint32_t axis_inside(float s, int32_t n) {
  return (s >= 0.0f && s < (float)(n - 1)) ? 1u : 0u;
}

template <int kDiv>
__device__ __forceinline__ float axis_weight(float d, float s, float h, float rh) {
  if constexpr (kDiv == kDivExact) {
    const float r = __builtin_fmaf(-__builtin_fabsf(s), h, d);   // -0 for d = -0, as exact_fmod
    return r * rh;
  } else {
    const float r = exact_fmod(d, h, s);
    return div_spacing<kDiv>(r, h, rh);   // r in [0, h) or -0: inside the verified range
  }
}

struct V3 { float a, b, c; };

// (1 - w) * lo + w * hi per component, two products then one sum.
__device__ __forceinline__ V3 lerp3(const V3& lo, const V3& hi, float w, float om) {
  V3 r;
  { const float x = om * lo.a; const float y = w * hi.a; r.a = x + y; }
  { const float x = om * lo.b; const float y = w * hi.b; r.b = x + y; }
  { const float x = om * lo.c; const float y = w * hi.c; r.c = x + y; }
  return r;
}

__device__ __forceinline__ V3 node(const FieldView& f, const float* base, int64_t off) {
  const float* p = base + 3 * off;
  return V3{p[0], p[1], p[2]};
}

// Record index of cell (c0, c1, c2) of time slice it in the cell layout.
__host__ __device__ inline int64_t cell_record(const FieldView& f, int32_t it, int32_t c0,
                                               int32_t c1, int32_t c2) {
  if (f.cell_bits) {   // field-uniform
    const uint32_t k = f.cell_bits, m = (1u << k) - 1u;
    const int64_t brick = ((int64_t)(c0 >> k) * f.cb1 + (c1 >> k)) * f.cb2 + (c2 >> k);
    const int64_t inner = ((int64_t)(c0 & m) << (2 * k)) | ((c1 & m) << k) | (c2 & m);
    return (int64_t)it * f.cell_slice + (brick << (3 * k)) + inner;
  }
  return (((int64_t)it * (f.n0 - 1) + c0) * (f.n1 - 1) + c1) * (int64_t)(f.n2 - 1) + c2;
}

// cell_record for fields with idx32 (the record index and 8 times it fit 32 bits).
__device__ __forceinline__ uint32_t cell_record32(const FieldView& f, uint32_t it, uint32_t c0,
                                                  uint32_t c1, uint32_t c2) {
  if (f.cell_bits) {   // field-uniform
    const uint32_t k = f.cell_bits, m = (1u << k) - 1u;
    const uint32_t brick = ((c0 >> k) * (uint32_t)f.cb1 + (c1 >> k)) * (uint32_t)f.cb2 + (c2 >> k);
    const uint32_t inner = ((c0 & m) << (2 * k)) | ((c1 & m) << k) | (c2 & m);
    return it * (uint32_t)f.cell_slice + (brick << (3 * k)) + inner;
  }
  return ((it * (uint32_t)(f.n0 - 1) + c0) * (uint32_t)(f.n1 - 1) + c1) * (uint32_t)(f.n2 - 1) +
         c2;
}

// Eight corners from one cell record: six 16-B loads of one 128-B line.
__device__ __forceinline__ void load_cell(const float4* rec, V3& L0, V3& L1, V3& L2, V3& L3,
                                          V3& L4, V3& L5, V3& L6, V3& L7) {
  const float4 a = rec[0], b = rec[1], c = rec[2], d = rec[3], e = rec[4], g = rec[5];
  L0 = V3{a.x, a.y, a.z}; L1 = V3{a.w, b.x, b.y}; L2 = V3{b.z, b.w, c.x}; L3 = V3{c.y, c.z, c.w};
  L4 = V3{d.x, d.y, d.z}; L5 = V3{d.w, e.x, e.y}; L6 = V3{e.z, e.w, g.x}; L7 = V3{g.y, g.z, g.w};
}

// The eight corners of the cell a particle last sampled, kept in registers.
// A particle moves a fraction of a cell per step, so most steps sample the
// same cell again and need no memory access at all; only lanes that cross
// into a new cell reload.  The values are the grid's, so results are
// unchanged.  The cell is kept as the floats the sample computes (integral
// values; -0 and +0 name the same cell and compare equal), so the test needs
// no conversion.  s0 < 0 marks an empty cache; clearing sets every coordinate,
// so the cell test never compares an indeterminate value.
struct CellCache {
  float s0, s1, s2, st;
  V3 L0, L1, L2, L3, L4, L5, L6, L7;   // corners of slice ct (the only slice if steady)
  V3 M0, M1, M2, M3, M4, M5, M6, M7;   // corners of slice ct + 1 (time-variant fields)
};

__device__ __forceinline__ void cache_clear(CellCache& cc) {
  cc.s0 = -1.0f; cc.s1 = -1.0f; cc.s2 = -1.0f; cc.st = -1.0f;
}

// Trilinear sample of one time slice, corner order and lerp cascade of
// regular_grid.hpp:59-74: list index 4a+2b+c, axis 2 first, then 1, then 0.
__device__ __forceinline__ V3 trilinear(const FieldView& f, const float* slice,
                                        const Axis& a0, const Axis& a1, const Axis& a2,
                                        CellCache& cc) {
  if (a0.s != cc.s0 || a1.s != cc.s1 || a2.s != cc.s2) {
    const int32_t c0 = (int32_t)a0.s, c1 = (int32_t)a1.s, c2 = (int32_t)a2.s;
    if (f.cells) {   // field-uniform branch
      const float4* rec = f.idx32 ? f.cells + 8u * cell_record32(f, 0u, c0, c1, c2)
                                  : f.cells + 8 * cell_record(f, 0, c0, c1, c2);
      load_cell(rec, cc.L0, cc.L1, cc.L2, cc.L3, cc.L4, cc.L5, cc.L6, cc.L7);
    } else {
      // one base address per reload; the four rows sit at launch-uniform
      // distances and each row's two nodes 12 bytes apart
      const float *p0, *p1, *p2, *p3;
      if (f.idx32) {   // field-uniform branch: the four row offsets in 32 bits
        const uint32_t u2 = 3u * (uint32_t)f.n2, u12 = u2 * (uint32_t)f.n1;
        const uint32_t o0 = 3u * (((uint32_t)c0 * (uint32_t)f.n1 + (uint32_t)c1) *
                                  (uint32_t)f.n2 + (uint32_t)c2);
        p0 = slice + o0;
        p1 = slice + (o0 + u2);
        p2 = slice + (o0 + u12);
        p3 = slice + (o0 + u12 + u2);
      } else {
        const int64_t n2 = f.n2, n12 = (int64_t)f.n1 * f.n2;
        p0 = slice + 3 * ((int64_t)c0 * n12 + (int64_t)c1 * n2 + c2);
        p1 = p0 + 3 * n2;
        p2 = p0 + 3 * n12;
        p3 = p2 + 3 * n2;
      }
      cc.L0 = V3{p0[0], p0[1], p0[2]}; cc.L1 = V3{p0[3], p0[4], p0[5]};
      cc.L2 = V3{p1[0], p1[1], p1[2]}; cc.L3 = V3{p1[3], p1[4], p1[5]};
      cc.L4 = V3{p2[0], p2[1], p2[2]}; cc.L5 = V3{p2[3], p2[4], p2[5]};
      cc.L6 = V3{p3[0], p3[1], p3[2]}; cc.L7 = V3{p3[3], p3[4], p3[5]};
    }
    cc.s0 = a0.s; cc.s1 = a1.s; cc.s2 = a2.s;
  }
  const float w2 = a2.w, m2 = 1.0f - w2;
  const V3 A0 = lerp3(cc.L0, cc.L1, w2, m2), A1 = lerp3(cc.L2, cc.L3, w2, m2);
  const V3 A2 = lerp3(cc.L4, cc.L5, w2, m2), A3 = lerp3(cc.L6, cc.L7, w2, m2);
  const float w1 = a1.w, m1 = 1.0f - w1;
  const V3 B0 = lerp3(A0, A1, w1, m1), B1 = lerp3(A2, A3, w1, m1);
  const float w0 = a0.w, m0 = 1.0f - w0;
  return lerp3(B0, B1, w0, m0);
}

// Four-dimensional sample of a time-variant field (regular_grid<vector3,4>,
// regular_fields.hpp:27): the 16 corners in permute_for order have time as the
// innermost axis, so the cascade lerps time first, then axes 2, 1, 0.  The
// cache holds the corners of both bracketing slices.
__device__ __forceinline__ V3 quadlinear(const FieldView& f, const Axis& a0, const Axis& a1,
                                         const Axis& a2, const Axis& at, CellCache& cc) {
  if (a0.s != cc.s0 || a1.s != cc.s1 || a2.s != cc.s2 || at.s != cc.st) {
    const int32_t c0 = (int32_t)a0.s, c1 = (int32_t)a1.s, c2 = (int32_t)a2.s;
    const int32_t ct = (int32_t)at.s;
    if (f.cells) {   // field-uniform branch
      const float4* r0 = f.idx32 ? f.cells + 8u * cell_record32(f, ct, c0, c1, c2)
                                 : f.cells + 8 * cell_record(f, ct, c0, c1, c2);
      const float4* r1 = r0 + 8 * f.cell_slice;
      load_cell(r0, cc.L0, cc.L1, cc.L2, cc.L3, cc.L4, cc.L5, cc.L6, cc.L7);
      load_cell(r1, cc.M0, cc.M1, cc.M2, cc.M3, cc.M4, cc.M5, cc.M6, cc.M7);
    } else {
      int64_t n2, n12, b;
      const float* s0;
      if (f.idx32) {   // field-uniform branch
        const uint32_t u2 = (uint32_t)f.n2, u12 = (uint32_t)f.n1 * u2;
        n2 = u2;
        n12 = u12;
        b = 0;
        s0 = f.data + 3u * ((uint32_t)ct * (uint32_t)f.slice + (uint32_t)c0 * u12 +
                            (uint32_t)c1 * u2 + (uint32_t)c2);
      } else {
        n2 = f.n2;
        n12 = (int64_t)f.n1 * f.n2;
        b = (int64_t)c0 * n12 + (int64_t)c1 * n2 + c2;
        s0 = f.data + 3 * (int64_t)ct * f.slice;
      }
      const float* s1 = s0 + 3 * f.slice;
      cc.L0 = node(f, s0, b);            cc.M0 = node(f, s1, b);
      cc.L1 = node(f, s0, b + 1);        cc.M1 = node(f, s1, b + 1);
      cc.L2 = node(f, s0, b + n2);       cc.M2 = node(f, s1, b + n2);
      cc.L3 = node(f, s0, b + n2 + 1);   cc.M3 = node(f, s1, b + n2 + 1);
      cc.L4 = node(f, s0, b + n12);      cc.M4 = node(f, s1, b + n12);
      cc.L5 = node(f, s0, b + n12 + 1);  cc.M5 = node(f, s1, b + n12 + 1);
      cc.L6 = node(f, s0, b + n12 + n2); cc.M6 = node(f, s1, b + n12 + n2);
      cc.L7 = node(f, s0, b + n12 + n2 + 1);
      cc.M7 = node(f, s1, b + n12 + n2 + 1);
    }
    cc.s0 = a0.s; cc.s1 = a1.s; cc.s2 = a2.s; cc.st = at.s;
  }
  const float wt = at.w, mt = 1.0f - wt;
  const V3 T0 = lerp3(cc.L0, cc.M0, wt, mt), T1 = lerp3(cc.L1, cc.M1, wt, mt);
  const V3 T2 = lerp3(cc.L2, cc.M2, wt, mt), T3 = lerp3(cc.L3, cc.M3, wt, mt);
  const V3 T4 = lerp3(cc.L4, cc.M4, wt, mt), T5 = lerp3(cc.L5, cc.M5, wt, mt);
  const V3 T6 = lerp3(cc.L6, cc.M6, wt, mt), T7 = lerp3(cc.L7, cc.M7, wt, mt);
  const float w2 = a2.w, m2 = 1.0f - w2;
  const V3 A0 = lerp3(T0, T1, w2, m2), A1 = lerp3(T2, T3, w2, m2);
  const V3 A2 = lerp3(T4, T5, w2, m2), A3 = lerp3(T6, T7, w2, m2);
  const float w1 = a1.w, m1 = 1.0f - w1;
  const V3 B0 = lerp3(A0, A1, w1, m1), B1 = lerp3(A2, A3, w1, m1);
  const float w0 = a0.w, m0 = 1.0f - w0;
  return lerp3(B0, B1, w0, m0);
}

// contains + interpolate at (p, t).  Returns false when p (or t, for a
// time-variant field) lies outside; v is the stored (unswizzled) vector.
// kDiv: the launch's DivMode.
template <bool kUnsteady, int kDiv = kDivide>
__device__ __forceinline__ bool sample(const FieldView& f, float p0, float p1, float p2, float t,
                                       V3& v, CellCache& cc) {
  const float d0 = p0 - f.o0, d1 = p1 - f.o1, d2 = p2 - f.o2;
  Axis a0, a1, a2;
  if constexpr (kDiv == kDivFast) {   // axis_floor on three axes, one rare-operand branch
    float q0, q1, q2;
    div_by3(d0, f.h0, f.r0, d1, f.h1, f.r1, d2, f.h2, f.r2, q0, q1, q2);
    a0.s = floorf(q0); a1.s = floorf(q1); a2.s = floorf(q2);
  } else {
    a0.s = axis_floor<kDiv>(d0, f.h0, f.r0);
    a1.s = axis_floor<kDiv>(d1, f.h1, f.r1);
    a2.s = axis_floor<kDiv>(d2, f.h2, f.r2);
  }
  // every axis tested together: one branch instead of one per axis
  uint32_t in = axis_inside(a0.s, f.n0) & axis_inside(a1.s, f.n1) & axis_inside(a2.s, f.n2);
  if constexpr (kUnsteady) {
    const float dt = t - f.to;
    Axis at;
    at.s = axis_floor<kDiv>(dt, f.th, f.rt);
    in = in & axis_inside(at.s, f.nt);
    if (!in) return false;
    a0.w = axis_weight<kDiv>(d0, a0.s, f.h0, f.r0);
    a1.w = axis_weight<kDiv>(d1, a1.s, f.h1, f.r1);
    a2.w = axis_weight<kDiv>(d2, a2.s, f.h2, f.r2);
    at.w = axis_weight<kDiv>(dt, at.s, f.th, f.rt);
    v = quadlinear(f, a0, a1, a2, at, cc);
  } else {
    (void)t;
    if (!in) return false;
    if constexpr (kDiv == kDivFast) {   // axis_weight on three axes, one rare-operand branch
      div_by3(exact_fmod(d0, f.h0, a0.s), f.h0, f.r0, exact_fmod(d1, f.h1, a1.s), f.h1, f.r1,
              exact_fmod(d2, f.h2, a2.s), f.h2, f.r2, a0.w, a1.w, a2.w);
    } else {
      a0.w = axis_weight<kDiv>(d0, a0.s, f.h0, f.r0);
      a1.w = axis_weight<kDiv>(d1, a1.s, f.h1, f.r1);
      a2.w = axis_weight<kDiv>(d2, a2.s, f.h2, f.r2);
    }
    v = trilinear(f, f.data, a0, a1, a2, cc);
  }
  return true;
}

// Eigen isZero() at float dummy_precision (particle_advector.cpp:363).
__device__ __forceinline__ bool is_zero(const V3& v) {
  return fabsf(v.a) <= 1e-5f && fabsf(v.b) <= 1e-5f && fabsf(v.c) <= 1e-5f;
}

// Step coefficients a_j = fl(b_j * dt) with b = (fl(1/6), fl(1/3), fl(1/3), fl(1/6)).
struct StepCoef {
  float a16, a13, half_dt, zero_dt, dt;
};
__host__ __device__ inline StepCoef make_coef(float dt) {
  const float b16 = 1.0f / 6.0f, b13 = 1.0f / 3.0f, half = 0.5f, zero = 0.0f;
  StepCoef c;
  c.a16 = b16 * dt;
  c.a13 = b13 * dt;
  c.half_dt = half * dt;
  c.zero_dt = zero * dt;
  c.dt = dt;
  return c;
}

// x + a16 k + a13 k + a13 k + a16 k, left to right (odeint runge_kutta4 final
// stage over vector_space_algebra, with dpa's single sample in every stage).
__device__ __forceinline__ float rk_sum_dpa(float x, float k, const StepCoef& c) {
  const float t1 = c.a16 * k; x = x + t1;
  const float t2 = c.a13 * k; x = x + t2;
  const float t3 = c.a13 * k; x = x + t3;
  const float t4 = c.a16 * k; x = x + t4;
  return x;
}

__device__ __forceinline__ void step_dpa(float& p0, float& p1, float& p2, const V3& v,
                                         const StepCoef& c) {
  // dxdt = (v2, v1, v0): axis k moves by stored component 2-k.
  p0 = rk_sum_dpa(p0, v.c, c);
  p1 = rk_sum_dpa(p1, v.b, c);
  p2 = rk_sum_dpa(p2, v.a, c);
}

// Classic four-sample RK4 in odeint's explicit_generic_rk order (not dpa's
// mode; an extension).  A stage point outside the grid returns false and leaves
// the position unchanged.
template <bool kUnsteady, int kDiv = kDivide>
__device__ __forceinline__ bool step_rk4(const FieldView& f, float& p0, float& p1, float& p2,
                                         const V3& v, float t, const StepCoef& c, CellCache& cc) {
  const float k10 = v.c, k11 = v.b, k12 = v.a;
  float x0, x1, x2;
  { const float s = c.half_dt * k10; x0 = p0 + s; }
  { const float s = c.half_dt * k11; x1 = p1 + s; }
  { const float s = c.half_dt * k12; x2 = p2 + s; }
  V3 s2;
  if (!sample<kUnsteady, kDiv>(f, x0, x1, x2, t + c.half_dt, s2, cc)) return false;
  const float k20 = s2.c, k21 = s2.b, k22 = s2.a;
  { float y = p0; const float a = c.zero_dt * k10; y = y + a; const float b = c.half_dt * k20; x0 = y + b; }
  { float y = p1; const float a = c.zero_dt * k11; y = y + a; const float b = c.half_dt * k21; x1 = y + b; }
  { float y = p2; const float a = c.zero_dt * k12; y = y + a; const float b = c.half_dt * k22; x2 = y + b; }
  V3 s3;
  if (!sample<kUnsteady, kDiv>(f, x0, x1, x2, t + c.half_dt, s3, cc)) return false;
  const float k30 = s3.c, k31 = s3.b, k32 = s3.a;
  const float one_dt = 1.0f * c.dt;
  { float y = p0; const float a = c.zero_dt * k10; y = y + a; const float b = c.zero_dt * k20; y = y + b; const float d = one_dt * k30; x0 = y + d; }
  { float y = p1; const float a = c.zero_dt * k11; y = y + a; const float b = c.zero_dt * k21; y = y + b; const float d = one_dt * k31; x1 = y + d; }
  { float y = p2; const float a = c.zero_dt * k12; y = y + a; const float b = c.zero_dt * k22; y = y + b; const float d = one_dt * k32; x2 = y + d; }
  V3 s4;
  if (!sample<kUnsteady, kDiv>(f, x0, x1, x2, t + one_dt, s4, cc)) return false;
  const float k40 = s4.c, k41 = s4.b, k42 = s4.a;
  auto fin = [&](float x, float q1, float q2, float q3, float q4) {
    const float a = c.a16 * q1; x = x + a;
    const float b = c.a13 * q2; x = x + b;
    const float d = c.a13 * q3; x = x + d;
    const float e = c.a16 * q4; x = x + e;
    return x;
  };
  p0 = fin(p0, k10, k20, k30, k40);
  p1 = fin(p1, k11, k21, k31, k41);
  p2 = fin(p2, k12, k22, k32, k42);
  return true;
}

}  // namespace dpah
