// ftle_kernels.hpp -- K_ftle: the flow-map gradient, right Cauchy-Green tensor
// and largest eigenvalue of ftle_estimator::estimate (ftle_estimator.cpp:7-65,
// regular_grid.hpp:87-108), plus the auxiliary-neighbour variant.
//
// J and C = J^T J are float32.  The largest eigenvalue uses the closed-form
// trigonometric solution of the characteristic cubic in float32; when the two
// largest eigenvalues nearly coincide (r = det(B)/2 -> -1, where acos has an
// unbounded derivative) it is recomputed in float64 from the same float32 C.
#pragma once

#include "advect_kernels.hpp"

namespace dpah {

template <class T>
__device__ __forceinline__ T max_eig_sym3(T a00, T a01, T a02, T a11, T a12, T a22, bool& ill) {
  const T p1 = a01 * a01 + a02 * a02 + a12 * a12;
  const T q = (a00 + a11 + a22) / T(3);
  const T b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
  const T p2 = b00 * b00 + b11 * b11 + b22 * b22 + T(2) * p1;
  ill = false;
  if (!(p2 > T(0))) return q;                 // C = q I
  const T p = sqrt(p2 / T(6));
  const T ip = T(1) / p;
  const T c00 = b00 * ip, c11 = b11 * ip, c22 = b22 * ip;
  const T c01 = a01 * ip, c02 = a02 * ip, c12 = a12 * ip;
  const T det = c00 * (c11 * c22 - c12 * c12) - c01 * (c01 * c22 - c12 * c02) +
                c02 * (c01 * c12 - c11 * c02);
  T r = det / T(2);
  r = r < T(-1) ? T(-1) : (r > T(1) ? T(1) : r);
  ill = r < T(-0.99);
  const T phi = acos(r) / T(3);
  return q + T(2) * p * cos(phi);
}

// FTLE from a Jacobian given column by column: J[:, d] = (ja_d, jb_d, jc_d).
// ln(sqrt(lmax)) / T in float32, as ftle_estimator.cpp:57 does.
__device__ __forceinline__ float ftle_from_jacobian(const float J[3][3], float T_abs) {
  // C = J^T J:  C(i, j) = sum_k J(k, i) J(k, j)
  float C[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) {
      const float x = J[0][i] * J[0][j], y = J[1][i] * J[1][j], z = J[2][i] * J[2][j];
      C[i][j] = x + (y + z);
    }
  bool ill;
  float l = max_eig_sym3<float>(C[0][0], C[0][1], C[0][2], C[1][1], C[1][2], C[2][2], ill);
  if (ill) {
    bool ill64;
    l = (float)max_eig_sym3<double>(C[0][0], C[0][1], C[0][2], C[1][1], C[1][2], C[2][2], ill64);
  }
  return logf(sqrtf(l)) / T_abs;
}

// ---------------------------------------------------------- auxiliary ----

struct FtleAuxArgs {
  // six per point, [points][6][3]: -d0, +d0, -d1, +d1, -d2, +d2; or, with
  // `shared`, the particles of aux_shared_particle over a d0 x d1 x d2 chunk
  const float* finals;
  uint64_t points;
  float twod0, twod1, twod2;      // fl(2 * delta_d)
  float T_abs;
  float* out;            // [points]
  uint32_t shared;
  uint64_t d0, d1, d2;
};

// Per point: J[:, d] = (final(+delta_d) - final(-delta_d)) / (2 delta_d), then
// C = J^T J, lambda_max and ln(sqrt(lambda_max)) / T (ftle_from_jacobian).
static __global__ __launch_bounds__(256) void k_ftle_aux_reduce(FtleAuxArgs a) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.points) return;
  const float* lo[3];
  const float* hi[3];
  if (a.shared) {
    // point p's own +delta particles; its -delta particles are its lower
    // neighbours' +delta ones, or the chunk's face particles on a low face
    const uint64_t i2 = p % a.d2, r = p / a.d2, i1 = r % a.d1, i0 = r / a.d1;
    const uint64_t f0 = 3 * a.points, f1 = f0 + a.d1 * a.d2, f2 = f1 + a.d0 * a.d2;
    for (int d = 0; d < 3; ++d) hi[d] = a.finals + 3 * (3 * p + d);
    lo[0] = a.finals + 3 * (i0 > 0 ? 3 * (p - a.d1 * a.d2) + 0 : f0 + i1 * a.d2 + i2);
    lo[1] = a.finals + 3 * (i1 > 0 ? 3 * (p - a.d2) + 1 : f1 + i0 * a.d2 + i2);
    lo[2] = a.finals + 3 * (i2 > 0 ? 3 * (p - 1) + 2 : f2 + i0 * a.d1 + i1);
  } else {
    for (int d = 0; d < 3; ++d) {
      lo[d] = a.finals + 18 * p + 6 * d;
      hi[d] = lo[d] + 3;
    }
  }
  float J[3][3];
  const float den[3] = {a.twod0, a.twod1, a.twod2};
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      const float diff = hi[d][e] - lo[d][e];
      J[e][d] = diff / den[d];
    }
  a.out[p] = ftle_from_jacobian(J, a.T_abs);
}

// ------------------------------------------------------- dpa lattice ----

struct FtleScatterArgs {
  const float* original;     // [count][3], or null: lattice seeds from `lat`
  SeedLattice lat;
  const float* finals;       // [count][3]
  uint64_t count;
  float o0, o1, o2;          // flow-map offset (the field's offset)
  float s0, s1, s2;          // strided spacing fl(h * stride)
  uint64_t m0, m1, m2;       // strided shape
  unsigned long long* owner; // [m]: pid + 1 of the particle owning the cell, 0 = none
  float* flow;               // [m][3], zero where no particle lands
  unsigned long long* misses;  // particles whose cell lies outside the map
  float T_abs;               // |fl(dt * iterations)|
  float* out;                // [m]
};

// flow_map.cell(original_position) (ftle_estimator.cpp:41; regular_grid.hpp:28-34):
// floor((p - o) / s) per axis.  Returns false when it falls outside the map.
__device__ __forceinline__ bool flow_cell(const FtleScatterArgs& a, uint64_t pid, uint64_t& lin) {
  float p0, p1, p2;
  if (a.original) {
    p0 = a.original[3 * pid]; p1 = a.original[3 * pid + 1]; p2 = a.original[3 * pid + 2];
  } else {
    lattice_point(a.lat, pid, p0, p1, p2);
  }
  const float c0 = floorf((p0 - a.o0) / a.s0);
  const float c1 = floorf((p1 - a.o1) / a.s1);
  const float c2 = floorf((p2 - a.o2) / a.s2);
  if (!(c0 >= 0.0f) || !(c1 >= 0.0f) || !(c2 >= 0.0f)) return false;
  if (c0 >= (float)a.m0 || c1 >= (float)a.m1 || c2 >= (float)a.m2) return false;
  lin = ((uint64_t)c0 * a.m1 + (uint64_t)c1) * a.m2 + (uint64_t)c2;
  return true;
}

static __global__ __launch_bounds__(256) void k_ftle_clear(FtleScatterArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t m = a.m0 * a.m1 * a.m2;
  if (i < m) {
    a.owner[i] = 0ull;
    a.flow[3 * i] = 0.0f; a.flow[3 * i + 1] = 0.0f; a.flow[3 * i + 2] = 0.0f;
  }
  if (i == 0) *a.misses = 0ull;
}

// Two passes make the scatter deterministic: on a collision the highest
// particle index wins (dpa's TBB scatter leaves that order unspecified).
static __global__ __launch_bounds__(256) void k_ftle_claim(FtleScatterArgs a) {
  const uint64_t pid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pid >= a.count) return;
  uint64_t lin;
  if (flow_cell(a, pid, lin)) atomicMax(a.owner + lin, (unsigned long long)(pid + 1));
  else atomicAdd(a.misses, 1ull);
}

static __global__ __launch_bounds__(256) void k_ftle_scatter(FtleScatterArgs a) {
  const uint64_t pid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pid >= a.count) return;
  uint64_t lin;
  if (flow_cell(a, pid, lin) && a.owner[lin] == pid + 1) {
    a.flow[3 * lin] = a.finals[3 * pid];
    a.flow[3 * lin + 1] = a.finals[3 * pid + 1];
    a.flow[3 * lin + 2] = a.finals[3 * pid + 2];
  }
}

// regular_grid::gradient (regular_grid.hpp:95-105): clamped neighbours, the
// divisor stays fl(2 * s) at the borders; then C, lmax and the FTLE.
static __global__ __launch_bounds__(256) void k_ftle_gradient(FtleScatterArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t m = a.m0 * a.m1 * a.m2;
  if (i >= m) return;
  const uint64_t i2 = i % a.m2, r = i / a.m2, i1 = r % a.m1, i0 = r / a.m1;
  const uint64_t idx[3] = {i0, i1, i2};
  const uint64_t dim[3] = {a.m0, a.m1, a.m2};
  const uint64_t step[3] = {a.m1 * a.m2, a.m2, 1};
  const float two_s[3] = {2.0f * a.s0, 2.0f * a.s1, 2.0f * a.s2};
  float J[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const uint64_t prev = idx[d] > 0 ? i - step[d] : i;
    const uint64_t next = idx[d] < dim[d] - 1 ? i + step[d] : i;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      const float diff = a.flow[3 * next + e] - a.flow[3 * prev + e];
      J[e][d] = diff / two_s[d];
    }
  }
  a.out[i] = ftle_from_jacobian(J, a.T_abs);
}

}  // namespace dpah
