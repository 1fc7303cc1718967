// device_math.hpp -- per-step arithmetic of dpa's advection stage for gfx950.
//
// Every operation is one IEEE binary32 rounding, in the order dpa's CPU build
// performs it; the library is compiled with -ffp-contract=off and correctly
// rounded division/sqrt, so no multiply-add is fused behind our back.  The
// explicit fmaf calls compute exact quantities -- fmod's remainder
// (exact_fmod, fast_axis) and the residual of a reciprocal quotient before
// its one correction (div_by_fast) -- so each result is still the correctly
// rounded one the reference computes (see there).
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
  // Byte sizes of the node array (all slices) and of the cell layout when a
  // reload can address them with 32-bit byte offsets formed by 24-bit
  // multiplies (the cell layout as one buffer resource): the size is below
  // 2^32 and every factor of the offset's products below 2^24 (node_byte24,
  // cell_byte24); 0 otherwise.
  uint32_t node_bytes32, cell_bytes32;
  // Optional node bricks: the nodes again, in bricks of 4^3 nodes (768 B)
  // stored brick after brick in C order over the bricks, nodes in C order
  // inside a brick (the grid padded to whole bricks; padding nodes are never
  // read).  A cell's eight corners then lie in two or three lines instead of
  // the four to eight of four C-order rows, and a wave's neighbouring
  // particles share them: read by launches that sample nodes (lattice
  // seeds).  nb1, nb2: bricks along axes 1 and 2; nbrick_slice: bytes per
  // time slice.  Present only when all slices fit 2^32 bytes and the bricks
  // of a slice number below 2^24 (nbrick_byte24); null otherwise.
  const char* nbrick;
  uint32_t nb1, nb2, nbrick_slice;
};

// Word 3 of a raw buffer resource on gfx9-family GPUs: 32-bit data format,
// no swizzle; loads beyond the resource's byte size return zero.
constexpr int kBufferWord3 = 0x00020000;

// a * b for a, b < 2^24 and a product below 2^32: one full-rate 24-bit
// multiply (v_mul_u32_u24 / v_mad_u32_u24) where a 32-bit multiply is a
// quarter-rate instruction.  The masks tell the compiler the operand ranges.
__device__ __forceinline__ uint32_t mul24(uint32_t a, uint32_t b) {
  return (a & 0xFFFFFFu) * (b & 0xFFFFFFu);
}

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

// kDivFast locates one axis with no test on the way: the quotient, its floor,
// the remainder as one fmaf and the weight, all from the verified reciprocal.
// Where that is not the correctly rounded result, the caller takes
// exact_axis() in one branch for all axes; the branch is needed
//   * when |d| < 2^-90 (zero, tiny or NaN): the fast quotient may differ from
//     d / h (a tiny negative d whose quotient rounds to -0 lies in cell -0,
//     inside; the fast floor may say -1);
//   * when the cell is inside and some remainder m = d - |s| h is negative.
// Otherwise d is not tiny, so the fast quotient is d / h's (the reciprocal
// form was verified for every such d in the axis's range; beyond it both
// floors lie outside the grid) and s is the floor of the rounded quotient, as
// regular_grid.hpp computes it: the inside test on the fast s is the
// reference's.  s is the floor k of the exact quotient or k + 1 (exact_fmod);
// k + 1 makes the exact d - (k+1) h negative and non-zero -- a multiple of the
// coarser of d's and h's ulps -- so its rounding m stays negative, and m >= 0
// means s == k: m is fmod's exact result.  m is then +0 (a node plane: the
// fast quotient of 0 is +0, as 0 / h) or at least the spacing's ulp when
// s >= 1, or d itself when s == 0; fast division needs at least 2^-90, so
// verified spacings are at least 2^-60 (field creation enforces it).  Lanes
// outside the grid use no weight.
__device__ __forceinline__ void fast_axis(float d, float h, float r, float& s, float& w,
                                          float& m) {
  s = floorf(div_by_fast(d, h, r));
  m = __builtin_fmaf(-__builtin_fabsf(s), h, d);
  w = div_by_fast(m, h, r);
}

// 1 or 0, integers so that the axes combine with | and & and no branch
__device__ __forceinline__ uint32_t fast_axis_tiny(float d) {
  return __builtin_fabsf(d) >= 0x1p-90f ? 0u : 1u;
}
__device__ __forceinline__ uint32_t fast_axis_roundup(float m) { return m < 0.0f ? 1u : 0u; }

__device__ __forceinline__ void exact_axis(float d, float h, float r, float& s, float& w) {
  s = floorf(div_by(d, h, r));
  w = div_by(exact_fmod(d, h, s), h, r);
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

// Byte offset of node (c0, c1, c2) of slice 0, for fields with node_bytes32
// (the plane's bytes 12 n1 n2 below 2^24; the offset below 2^32): three
// 24-bit multiply-adds, the 12 bytes of a node folded into the factors.
__device__ __forceinline__ uint32_t node_byte24(const FieldView& f, uint32_t c0, uint32_t c1,
                                                uint32_t c2) {
  return mul24(c0, 12u * (uint32_t)f.n1 * (uint32_t)f.n2) + mul24(c1, 12u * (uint32_t)f.n2) +
         mul24(c2, 12u);
}

// Byte offset of cell record (c0, c1, c2) of slice 0, for fields with
// cell_bytes32 (the brick counts, or the cell counts in C order, and their
// plane products below 2^24).
__device__ __forceinline__ uint32_t cell_byte24(const FieldView& f, uint32_t c0, uint32_t c1,
                                                uint32_t c2) {
  uint32_t rec;
  if (f.cell_bits) {   // field-uniform
    const uint32_t k = f.cell_bits, m = (1u << k) - 1u;
    const uint32_t brick = mul24(c0 >> k, (uint32_t)f.cb1 * (uint32_t)f.cb2) +
                           mul24(c1 >> k, (uint32_t)f.cb2) + (c2 >> k);
    rec = (brick << (3 * k)) + (((c0 & m) << (2 * k)) | ((c1 & m) << k) | (c2 & m));
  } else {
    rec = mul24(c0, (uint32_t)(f.n1 - 1) * (uint32_t)(f.n2 - 1)) + mul24(c1, (uint32_t)(f.n2 - 1)) +
          c2;
  }
  return rec << 7;   // 128 bytes per record
}

// Loads through a raw buffer resource over [base, base + bytes): voffset in
// VGPRs, a uniform soffset, constant parts in the instruction's offset field.
// The host build of this header (the tests' CPU checks) reads the same bytes
// through a pointer.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __amdgpu_buffer_rsrc_t BufferRes;
__device__ __forceinline__ BufferRes buffer_res(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           kBufferWord3);
}
__device__ __forceinline__ float4 buf_load4(BufferRes r, uint32_t voff, uint32_t soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
  return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                     __uint_as_float(v[3]));
}
#else
typedef const char* BufferRes;
__device__ __forceinline__ BufferRes buffer_res(const void* base, uint32_t) {
  return (const char*)base;
}
__device__ __forceinline__ float4 buf_load4(BufferRes r, uint32_t voff, uint32_t soff) {
  const float* p = (const float*)(r + voff + soff);
  return make_float4(p[0], p[1], p[2], p[3]);
}
#endif

// Byte offset of node (c0, c1, c2) in a slice of the node bricks.
__device__ __forceinline__ uint32_t nbrick_byte24(const FieldView& f, uint32_t c0,
                                                           uint32_t c1, uint32_t c2) {
  const uint32_t brick = mul24(c0 >> 2, f.nb1 * f.nb2) + mul24(c1 >> 2, f.nb2) + (c2 >> 2);
  const uint32_t local = ((c0 & 3u) << 4) | ((c1 & 3u) << 2) | (c2 & 3u);
  return mul24(brick, 768u) + local * 12u;
}

// One node (12 bytes) at a 32-bit byte offset from a uniform base.
__device__ __forceinline__ V3 load_node(const char* base, uint32_t off) {
  const float* p = (const float*)(base + off);
  return V3{p[0], p[1], p[2]};
}

// The eight corners of cell (c0, c1, c2) from the node bricks at byte offset
// `at` of the slice holding them: per axis the step to the next node stays
// inside the brick, or goes to the next brick's first node on that axis.
__device__ __forceinline__ void load_nbrick_cell(const FieldView& f, const char* b, uint32_t at,
                                                 uint32_t c0, uint32_t c1, uint32_t c2,
                                                 V3& L0, V3& L1, V3& L2, V3& L3, V3& L4,
                                                 V3& L5, V3& L6, V3& L7) {
  const uint32_t d2 = (c2 & 3u) != 3u ? 12u : 768u - 36u;
  const uint32_t d1 = (c1 & 3u) != 3u ? 48u : f.nb2 * 768u - 144u;
  const uint32_t d0 = (c0 & 3u) != 3u ? 192u : f.nb1 * f.nb2 * 768u - 576u;
  L0 = load_node(b, at);
  L1 = load_node(b, at + d2);
  L2 = load_node(b, at + d1);
  L3 = load_node(b, at + (d1 + d2));
  L4 = load_node(b, at + d0);
  L5 = load_node(b, at + (d0 + d2));
  L6 = load_node(b, at + (d0 + d1));
  L7 = load_node(b, at + (d0 + d1 + d2));
}

// One node row of a cell (two nodes along axis 2, 24 bytes) at a 32-bit byte
// offset from a uniform base: the loads take the base in scalar registers and
// the offset in one vector register.
__device__ __forceinline__ void load_row(const char* base, uint32_t off, V3& a, V3& b) {
  const float* p = (const float*)(base + off);
  a = V3{p[0], p[1], p[2]};
  b = V3{p[3], p[4], p[5]};
}

// A cell record (six 16-byte pieces of one 128-byte line) from a buffer.
__device__ __forceinline__ void load_cell_buf(BufferRes r, uint32_t voff, uint32_t soff, V3& L0,
                                              V3& L1, V3& L2, V3& L3, V3& L4, V3& L5, V3& L6,
                                              V3& L7) {
  const float4 a = buf_load4(r, voff, soff), b = buf_load4(r, voff + 16u, soff);
  const float4 c = buf_load4(r, voff + 32u, soff), d = buf_load4(r, voff + 48u, soff);
  const float4 e = buf_load4(r, voff + 64u, soff), g = buf_load4(r, voff + 80u, soff);
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
    if (f.cells && f.cell_bytes32) {   // field-uniform branches
      load_cell_buf(buffer_res(f.cells, f.cell_bytes32), cell_byte24(f, c0, c1, c2), 0u, cc.L0,
                    cc.L1, cc.L2, cc.L3, cc.L4, cc.L5, cc.L6, cc.L7);
    } else if (f.cells) {
      const float4* rec = f.idx32 ? f.cells + 8u * cell_record32(f, 0u, c0, c1, c2)
                                  : f.cells + 8 * cell_record(f, 0, c0, c1, c2);
      load_cell(rec, cc.L0, cc.L1, cc.L2, cc.L3, cc.L4, cc.L5, cc.L6, cc.L7);
    } else if (f.nbrick) {   // slice 0, like every caller
      const uint32_t u0 = (uint32_t)c0, u1 = (uint32_t)c1, u2 = (uint32_t)c2;
      load_nbrick_cell(f, f.nbrick, nbrick_byte24(f, u0, u1, u2), u0, u1, u2, cc.L0, cc.L1,
                       cc.L2, cc.L3, cc.L4, cc.L5, cc.L6, cc.L7);
    } else if (f.node_bytes32) {   // (slice is f.data: every caller samples slice 0)
      // one 32-bit byte offset per reload from the array's uniform base (a
      // buffer resource here would add scalar registers that lattice
      // instances spill); the four rows at uniform distances
      const char* b = (const char*)f.data;
      const uint32_t v = node_byte24(f, c0, c1, c2);
      const uint32_t r1 = 12u * (uint32_t)f.n2, r2 = 12u * (uint32_t)f.n1 * (uint32_t)f.n2;
      load_row(b, v, cc.L0, cc.L1);
      load_row(b, v + r1, cc.L2, cc.L3);
      load_row(b, v + r2, cc.L4, cc.L5);
      load_row(b, v + (r2 + r1), cc.L6, cc.L7);
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
    if (f.cells && f.cell_bytes32) {   // field-uniform branches
      const BufferRes r = buffer_res(f.cells, f.cell_bytes32);
      const uint32_t sb = (uint32_t)f.cell_slice << 7;   // bytes per slice
      const uint32_t v = (uint32_t)ct * sb + cell_byte24(f, c0, c1, c2);
      load_cell_buf(r, v, 0u, cc.L0, cc.L1, cc.L2, cc.L3, cc.L4, cc.L5, cc.L6, cc.L7);
      load_cell_buf(r, v, sb, cc.M0, cc.M1, cc.M2, cc.M3, cc.M4, cc.M5, cc.M6, cc.M7);
    } else if (f.cells) {
      const float4* r0 = f.idx32 ? f.cells + 8u * cell_record32(f, ct, c0, c1, c2)
                                 : f.cells + 8 * cell_record(f, ct, c0, c1, c2);
      const float4* r1 = r0 + 8 * f.cell_slice;
      load_cell(r0, cc.L0, cc.L1, cc.L2, cc.L3, cc.L4, cc.L5, cc.L6, cc.L7);
      load_cell(r1, cc.M0, cc.M1, cc.M2, cc.M3, cc.M4, cc.M5, cc.M6, cc.M7);
    } else if (f.nbrick) {
      const uint32_t u0 = (uint32_t)c0, u1 = (uint32_t)c1, u2 = (uint32_t)c2;
      const uint32_t v = (uint32_t)ct * f.nbrick_slice + nbrick_byte24(f, u0, u1, u2);
      load_nbrick_cell(f, f.nbrick, v, u0, u1, u2, cc.L0, cc.L1, cc.L2, cc.L3, cc.L4, cc.L5,
                       cc.L6, cc.L7);
      load_nbrick_cell(f, f.nbrick, v + f.nbrick_slice, u0, u1, u2, cc.M0, cc.M1, cc.M2, cc.M3,
                       cc.M4, cc.M5, cc.M6, cc.M7);
    } else if (f.node_bytes32) {
      const char* b = (const char*)f.data;
      const uint32_t sb = 12u * (uint32_t)f.slice;   // bytes per slice
      const uint32_t v = (uint32_t)ct * sb + node_byte24(f, c0, c1, c2);
      const uint32_t r1 = 12u * (uint32_t)f.n2, r2 = 12u * (uint32_t)f.n1 * (uint32_t)f.n2;
      load_row(b, v, cc.L0, cc.L1);                   load_row(b, v + sb, cc.M0, cc.M1);
      load_row(b, v + r1, cc.L2, cc.L3);              load_row(b, v + (sb + r1), cc.M2, cc.M3);
      load_row(b, v + r2, cc.L4, cc.L5);              load_row(b, v + (sb + r2), cc.M4, cc.M5);
      load_row(b, v + (r2 + r1), cc.L6, cc.L7);       load_row(b, v + (sb + r2 + r1), cc.M6, cc.M7);
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

// The kDivFast location of a sample: every axis (and time, for a time-variant
// field) on the fast path, then one branch for the lanes that need the exact
// forms -- a tiny, zero or non-finite offset on any axis, or a cell inside the
// grid with a quotient rounded up to the next integer on some axis
// (fast_axis).  Returns 1 when the cell is inside.  dt is t - to (unused for
// steady fields).
template <bool kUnsteady>
__device__ __forceinline__ uint32_t locate_fast(const FieldView& f, float d0, float d1, float d2,
                                                float dt, Axis& a0, Axis& a1, Axis& a2,
                                                Axis& at) {
  float m0, m1, m2;
  fast_axis(d0, f.h0, f.r0, a0.s, a0.w, m0);
  fast_axis(d1, f.h1, f.r1, a1.s, a1.w, m1);
  fast_axis(d2, f.h2, f.r2, a2.s, a2.w, m2);
  uint32_t in = axis_inside(a0.s, f.n0) & axis_inside(a1.s, f.n1) & axis_inside(a2.s, f.n2);
  uint32_t tiny = fast_axis_tiny(d0) | fast_axis_tiny(d1) | fast_axis_tiny(d2);
  uint32_t up = fast_axis_roundup(m0) | fast_axis_roundup(m1) | fast_axis_roundup(m2);
  if constexpr (kUnsteady) {
    float mt;
    fast_axis(dt, f.th, f.rt, at.s, at.w, mt);
    in = in & axis_inside(at.s, f.nt);
    tiny = tiny | fast_axis_tiny(dt);
    up = up | fast_axis_roundup(mt);
  }
  if (tiny | (in & up)) {
    exact_axis(d0, f.h0, f.r0, a0.s, a0.w);
    exact_axis(d1, f.h1, f.r1, a1.s, a1.w);
    exact_axis(d2, f.h2, f.r2, a2.s, a2.w);
    in = axis_inside(a0.s, f.n0) & axis_inside(a1.s, f.n1) & axis_inside(a2.s, f.n2);
    if constexpr (kUnsteady) {
      exact_axis(dt, f.th, f.rt, at.s, at.w);
      in = in & axis_inside(at.s, f.nt);
    }
  }
  return in;
}

// contains + interpolate at (p, t).  Returns false when p (or t, for a
// time-variant field) lies outside; v is the stored (unswizzled) vector.
// kDiv: the launch's DivMode.
template <bool kUnsteady, int kDiv = kDivide>
__device__ __forceinline__ bool sample(const FieldView& f, float p0, float p1, float p2, float t,
                                       V3& v, CellCache& cc) {
  const float d0 = p0 - f.o0, d1 = p1 - f.o1, d2 = p2 - f.o2;
  Axis a0, a1, a2;
  if constexpr (kDiv == kDivFast) {
    Axis at;
    const uint32_t in = locate_fast<kUnsteady>(f, d0, d1, d2, t - f.to, a0, a1, a2, at);
    if constexpr (kUnsteady) {
      if (!in) return false;
      v = quadlinear(f, a0, a1, a2, at, cc);
    } else {
      (void)t;
      if (!in) return false;
      v = trilinear(f, f.data, a0, a1, a2, cc);
    }
    return true;
  } else {
    a0.s = axis_floor<kDiv>(d0, f.h0, f.r0);
    a1.s = axis_floor<kDiv>(d1, f.h1, f.r1);
    a2.s = axis_floor<kDiv>(d2, f.h2, f.r2);
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
      a0.w = axis_weight<kDiv>(d0, a0.s, f.h0, f.r0);
      a1.w = axis_weight<kDiv>(d1, a1.s, f.h1, f.r1);
      a2.w = axis_weight<kDiv>(d2, a2.s, f.h2, f.r2);
      v = trilinear(f, f.data, a0, a1, a2, cc);
    }
    return true;
  }
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
