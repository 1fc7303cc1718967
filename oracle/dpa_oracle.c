/*
 * dpa_oracle.c -- CPU restatement of dpa's particle-advection path.
 *
 * TEST INFRASTRUCTURE ONLY (see dpa_oracle.h).  Only tests/, the smoke check
 * and bench.py's cpu_baseline leg load this library, as the checker.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp, no -march).
 *
 * What pins each part (tests/test_oracle.py unless noted):
 *   contains / interpolate   multilinear fields reproduced to rounding, corner
 *                            order from an index-encoding field, face / last
 *                            cell / outside / negative special cases, fmodf
 *                            identity d = k*h + r.
 *   dpo_step_dpa             uniform field straight line, rotation field radius
 *                            growth sqrt(1+(w dt)^2) per step (dpa mode) vs
 *                            O((w dt)^6) drift (rk4 mode), four-term sum vs
 *                            x + dt*v sensitivity, swizzle moves axis 0 from
 *                            stored component 2.
 *   advect loop / rounds     zero field -> [seed, T]; outside seeds -> [seed, T];
 *                            brute-force exit counts in a uniform field; vertex
 *                            count sum(k+2); tail-first round order; ppr splits.
 *   prune / indices / colors closed forms (concat [b+j, b+j+1]); straight line
 *                            colours 0; right-angle zig-zag pi/2 with the
 *                            vertex-1 quirk; dpa's remove semantics on planted
 *                            sentinels.
 *   time-variant sampling    (tests/test_oracle_unsteady.py) fields multilinear
 *                            in (x0, x1, x2, t); a slice-index field; the 4-D
 *                            sample equals the 3-D sample of the nodes blended
 *                            in time first, bit for bit; time-axis contains;
 *                            exit iteration by brute force; a linear time ramp
 *                            against its closed-form sum / integral.  Each
 *                            planted slip (DPO_PLANT) fails its pins.
 *   lattice seeds            count and C order closed form.
 *   aux seeds                the moved coordinate against the exactly rounded
 *                            rational origin + round(spacing * i +- delta) for
 *                            random spacings, offsets and indices; with delta =
 *                            spacing / 2, point k's +delta particle equals point
 *                            k + 1's -delta particle bit for bit (and differs
 *                            from the two-rounding form at some k).
 *   ftle                     diagonal linear field closed form, general linear
 *                            field vs numpy float64 SVD, eigen special cases vs
 *                            numpy eigvalsh.
 * Parity unpinned (third-party arithmetic absent from the dpa tree, restated
 * from its published source): Boost.Odeint's runge_kutta4 summation order,
 * Eigen's isZero threshold 1e-5 and its length-3 reduction order
 * x0 + (x1 + x2) used by dot / squaredNorm.  The float SelfAdjointEigenSolver
 * of ftle_estimator.cpp:52 is replaced by a float64 solver: its bit pattern is
 * not reproducible without Eigen.
 */
#include "dpa_oracle.h"

/* DPO_PLANT: test builds only (tests/test_oracle_unsteady.py compiles this
 * file with -DDPO_PLANT=k into a temporary directory).  Each value plants one
 * known mistake in the time-variant (4-D) path, so that the pins of that path
 * can be shown to fail on it.  The default build (oracle/Makefile) defines
 * nothing and is 0: no plant.
 *   1  time lerped last (time the outermost corner bit, not the innermost)
 *   2  time slice off by one (the next interval, clamped to the last)
 *   3  time weight swapped (1 - w instead of w)
 *   4  time of iteration i taken as (i + 1) * dt
 *   5  contains admits the last time node (s <= nt - 1) */
#ifndef DPO_PLANT
#define DPO_PLANT 0
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- grid ---- */

static int grid_dims(const dpo_grid* g) { return g->nt > 1 ? 4 : 3; }

static void grid_coord(const dpo_grid* g, int axis, const float p[3], float t,
                       float* o, float* h, int64_t* n, float* x) {
  if (axis < 3) { *o = g->offset[axis]; *h = g->spacing[axis]; *n = g->n[axis]; *x = p[axis]; }
  else          { *o = g->t_offset;     *h = g->t_spacing;     *n = g->nt;      *x = t; }
}

/* regular_grid.hpp:41-50: subscript = floor((p - o) / h);
 * outside iff int64(subscript) < 0 or size_t(subscript) >= shape - 1. */
int dpo_contains(const dpo_grid* g, const float p[3], float t) {
  const int dims = grid_dims(g);
  for (int i = 0; i < dims; ++i) {
    float o, h, x; int64_t n;
    grid_coord(g, i, p, t, &o, &h, &n, &x);
    const float d = x - o;
    const float s = floorf(d / h);
    if (!(s >= 0.0f)) return 0;             /* negative, or NaN */
#if DPO_PLANT == 5
    if (i == 3) n += 1;
#endif
    if (s >= (float)(n - 1)) return 0;      /* also +inf */
  }
  return 1;
}

static const float* grid_node(const dpo_grid* g, int64_t i0, int64_t i1, int64_t i2, int64_t it) {
  const int64_t slice = g->n[0] * g->n[1] * g->n[2];
  return g->data + 3 * (it * slice + (i0 * g->n[1] + i1) * g->n[2] + i2);
}

/* regular_grid.hpp:52-75.  Corner list order is permute_for's
 * (permute_for.hpp:12-34): dimension 0 outermost, so list index
 * = sum_i bit_i * 2^(dims-1-i); then the cascade lerps the last dimension
 * first: L[j] = (1 - w_i) * L[2j] + w_i * L[2j+1], j < 2^i, i = dims-1 .. 0. */
void dpo_interpolate(const dpo_grid* g, const float p[3], float t, float out[3]) {
  const int dims = grid_dims(g);
  float   w[4];
  int64_t c[4] = {0, 0, 0, 0};
  for (int i = 0; i < dims; ++i) {
    float o, h, x; int64_t n;
    grid_coord(g, i, p, t, &o, &h, &n, &x);
    const float d = x - o;
    w[i] = fmodf(d, h) / h;
    c[i] = (int64_t)floorf(d / h);
  }
  /* axis[j]: the grid axis of corner bit j (bit dims-1-j of the list index) */
  int axis[4] = {0, 1, 2, 3};
#if DPO_PLANT == 1
  if (dims == 4) { axis[0] = 3; axis[1] = 0; axis[2] = 1; axis[3] = 2; }
#elif DPO_PLANT == 2
  if (dims == 4 && c[3] + 2 < g->nt) c[3] += 1;
#elif DPO_PLANT == 3
  if (dims == 4) w[3] = 1.0f - w[3];
#endif
  float L[16][3];
  const int corners = 1 << dims;
  for (int k = 0; k < corners; ++k) {
    int64_t idx[4] = {c[0], c[1], c[2], c[3]};
    for (int j = 0; j < dims; ++j)
      idx[axis[j]] += (k >> (dims - 1 - j)) & 1;
    const float* v = grid_node(g, idx[0], idx[1], idx[2], dims == 4 ? idx[3] : 0);
    L[k][0] = v[0]; L[k][1] = v[1]; L[k][2] = v[2];
  }
  for (int i = dims - 1; i >= 0; --i) {
    const float wi = w[axis[i]];
    const float om = 1.0f - wi;
    for (int j = 0; j < (1 << i); ++j)
      for (int e = 0; e < 3; ++e) {
        const float a = om * L[2 * j][e];
        const float b = wi * L[2 * j + 1][e];
        L[j][e] = a + b;
      }
  }
  out[0] = L[0][0]; out[1] = L[0][1]; out[2] = L[0][2];
}

/* Eigen DenseBase::isZero(): every |v_i| <= 1 * dummy_precision<float>() = 1e-5. */
int dpo_is_zero(const float v[3]) {
  return fabsf(v[0]) <= 1e-5f && fabsf(v[1]) <= 1e-5f && fabsf(v[2]) <= 1e-5f;
}

/* odeint runge_kutta4 (explicit_generic_rk, rk4 coefficients, Value = float)
 * last stage: x_out = 1*x + (b1 dt) k1 + (b2 dt) k2 + (b3 dt) k3 + (b4 dt) k4,
 * evaluated left to right per component by Eigen.  In dpa all k_j are the
 * same swizzled sample k = (v2, v1, v0) (particle_advector.cpp:369). */
void dpo_step_dpa(float x[3], const float v[3], float dt) {
  const float b16 = 1.0f / 6.0f, b13 = 1.0f / 3.0f;
  const float a1 = b16 * dt, a2 = b13 * dt, a3 = b13 * dt, a4 = b16 * dt;
  const float k[3] = {v[2], v[1], v[0]};
  for (int e = 0; e < 3; ++e) {
    float y = x[e];
    const float t1 = a1 * k[e]; y = y + t1;
    const float t2 = a2 * k[e]; y = y + t2;
    const float t3 = a3 * k[e]; y = y + t3;
    const float t4 = a4 * k[e]; y = y + t4;
    x[e] = y;
  }
}

/* Swizzled field sample (particle_advector.cpp:369): dxdt = (v2, v1, v0). */
static int sample_swizzled(const dpo_grid* g, const float x[3], float t, float k[3]) {
  if (!dpo_contains(g, x, t)) return 0;
  float v[3];
  dpo_interpolate(g, x, t, v);
  k[0] = v[2]; k[1] = v[1]; k[2] = v[0];
  return 1;
}

/* Classic four-sample RK4 in odeint's generic_rk order (rk4 coefficients
 * a = {1/2}, {0, 1/2}, {0, 0, 1}; b = {1/6, 1/3, 1/3, 1/6}; c = {0, 1/2, 1/2, 1}).
 * This mode is not dpa's: dpa always samples once.  A stage point outside the
 * grid retires the particle before the step (returns 0, x unchanged). */
static int step_rk4(const dpo_grid* g, float x[3], const float v[3], float t, float dt) {
  const float half = 1.0f / 2.0f, zero = 0.0f, one = 1.0f;
  const float b16 = 1.0f / 6.0f, b13 = 1.0f / 3.0f;
  float k1[3] = {v[2], v[1], v[0]}, k2[3], k3[3], k4[3], xt[3];
  const float hdt = half * dt, zdt = zero * dt, odt = one * dt;
  for (int e = 0; e < 3; ++e) { const float s = hdt * k1[e]; xt[e] = x[e] + s; }
  if (!sample_swizzled(g, xt, t + hdt, k2)) return 0;
  for (int e = 0; e < 3; ++e) {
    float y = x[e];
    const float s1 = zdt * k1[e]; y = y + s1;
    const float s2 = hdt * k2[e]; y = y + s2;
    xt[e] = y;
  }
  if (!sample_swizzled(g, xt, t + hdt, k3)) return 0;
  for (int e = 0; e < 3; ++e) {
    float y = x[e];
    const float s1 = zdt * k1[e]; y = y + s1;
    const float s2 = zdt * k2[e]; y = y + s2;
    const float s3 = odt * k3[e]; y = y + s3;
    xt[e] = y;
  }
  if (!sample_swizzled(g, xt, t + odt, k4)) return 0;
  const float a1 = b16 * dt, a2 = b13 * dt, a3 = b13 * dt, a4 = b16 * dt;
  for (int e = 0; e < 3; ++e) {
    float y = x[e];
    const float s1 = a1 * k1[e]; y = y + s1;
    const float s2 = a2 * k2[e]; y = y + s2;
    const float s3 = a3 * k3[e]; y = y + s3;
    const float s4 = a4 * k4[e]; y = y + s4;
    x[e] = y;
  }
  return 1;
}

/* particle_advector.cpp:325-396 for one particle. */
uint64_t dpo_advect_particle(const dpo_grid* g, float p[3], uint64_t* remaining,
                             float dt, float t0, int integrator, float* rec,
                             uint8_t* status) {
  uint64_t it = 0;
  uint8_t st = DPO_ACTIVE;
  if (rec) { rec[0] = p[0]; rec[1] = p[1]; rec[2] = p[2]; }
  for (; *remaining > 0; ++it, --*remaining) {
#if DPO_PLANT == 4
    const float ft = (float)(it + 1) * dt;
#else
    const float ft = (float)it * dt;     /* iteration_index * step_size_ (375) */
#endif
    const float t  = t0 + ft;
    if (!dpo_contains(g, p, t)) { st = DPO_EXITED; break; }
    float v[3];
    dpo_interpolate(g, p, t, v);
    if (dpo_is_zero(v)) { st = DPO_ZERO; break; }
    if (integrator == DPO_INTEGRATOR_RK4) {
      if (!step_rk4(g, p, v, t, dt)) { st = DPO_EXITED; break; }
    } else {
      dpo_step_dpa(p, v, dt);
    }
    if (rec) { float* r = rec + 3 * (it + 1); r[0] = p[0]; r[1] = p[1]; r[2] = p[2]; }
  }
  if (rec) { float* r = rec + 3 * (it + 1); r[0] = -2.0f; r[1] = -2.0f; r[2] = -2.0f; }
  if (*remaining == 0) st = DPO_COMPLETED;
  *status = st;
  return it;
}

/* --------------------------------------------------- post-processing ---- */

static int is_vec(const float* v, float s) { return v[0] == s && v[1] == s && v[2] == s; }

/* std::remove of every vertex equal to invalid_value (-1,-1,-1)
 * (particle_advector.cpp:402-408, basic_types.hpp:51-55). */
uint64_t dpo_prune(float* v, uint64_t count) {
  uint64_t out = 0;
  for (uint64_t i = 0; i < count; ++i) {
    if (is_vec(v + 3 * i, -1.0f)) continue;
    if (out != i) memmove(v + 3 * out, v + 3 * i, 3 * sizeof(float));
    ++out;
  }
  return out;
}

/* index_generator.cpp:10-39: 2V placeholders (all ones); (i, i+1) at 2i when
 * neither vertex is the terminal value; then remove the placeholders. */
uint64_t dpo_indices(const float* v, uint64_t count, void* out, int use_64_bit) {
  if (count == 0) return 0;
  uint64_t n = 0;
  if (use_64_bit) {
    uint64_t* o = (uint64_t*)out;
    for (uint64_t i = 0; i + 1 < count; ++i)
      if (!is_vec(v + 3 * i, -2.0f) && !is_vec(v + 3 * (i + 1), -2.0f)) {
        const uint64_t a = i, b = i + 1;
        if (a != UINT64_MAX) o[n++] = a;
        if (b != UINT64_MAX) o[n++] = b;
      }
  } else {
    uint32_t* o = (uint32_t*)out;
    for (uint64_t i = 0; i + 1 < count; ++i)
      if (!is_vec(v + 3 * i, -2.0f) && !is_vec(v + 3 * (i + 1), -2.0f)) {
        const uint32_t a = (uint32_t)i, b = (uint32_t)(i + 1);
        if (a != UINT32_MAX) o[n++] = a;
        if (b != UINT32_MAX) o[n++] = b;
      }
  }
  return n;
}

/* Eigen fixed-size-3 reductions: redux_novec_unroller splits [0,3) into
 * [0,1) and [1,3), giving x0 + (x1 + x2). */
static float sum3(float a, float b, float c) { const float bc = b + c; return a + bc; }
static float sqnorm3(const float v[3]) {
  const float a = v[0] * v[0], b = v[1] * v[1], c = v[2] * v[2];
  return sum3(a, b, c);
}
static float dot3(const float x[3], const float y[3]) {
  const float a = x[0] * y[0], b = x[1] * y[1], c = x[2] * y[2];
  return sum3(a, b, c);
}
/* Eigen MatrixBase::normalized(): v / sqrt(squaredNorm()) if that is > 0. */
static void normalized3(const float v[3], float o[3]) {
  const float z = sqnorm3(v);
  if (z > 0.0f) {
    const float s = sqrtf(z);
    o[0] = v[0] / s; o[1] = v[1] / s; o[2] = v[2] / s;
  } else {
    o[0] = v[0]; o[1] = v[1]; o[2] = v[2];
  }
}
static void cross3(const float a[3], const float b[3], float o[3]) {
  float p, q;
  p = a[1] * b[2]; q = a[2] * b[1]; o[0] = p - q;
  p = a[2] * b[0]; q = a[0] * b[2]; o[1] = p - q;
  p = a[0] * b[1]; q = a[1] * b[0]; o[2] = p - q;
}

/* color_generator.cpp:37-58 (note "vertex_index > 1" at line 50). */
void dpo_colors(const float* v, uint64_t count, float* out) {
  for (uint64_t i = 0; i < count; ++i) {
    out[i] = 0.0f;
    if (is_vec(v + 3 * i, -2.0f)) continue;
    float prev[3] = {0, 0, 0}, next[3] = {0, 0, 0}, d[3];
    if (i > 1 && !is_vec(v + 3 * (i - 1), -2.0f)) {
      for (int e = 0; e < 3; ++e) d[e] = v[3 * i + e] - v[3 * (i - 1) + e];
      normalized3(d, prev);
    }
    if (i + 1 < count && !is_vec(v + 3 * (i + 1), -2.0f)) {
      for (int e = 0; e < 3; ++e) d[e] = v[3 * (i + 1) + e] - v[3 * i + e];
      normalized3(d, next);
    }
    float c[3];
    cross3(prev, next, c);
    out[i] = atan2f(sqrtf(sqnorm3(c)), dot3(prev, next));
  }
}

/* ------------------------------------------------------------ rounds ---- */

typedef struct dpo_round {
  float*   vertices;
  float*   colors;
  void*    indices;
  uint64_t vertex_count;
  uint64_t index_count;
} dpo_round;

struct dpo_rounds {
  dpo_round* rounds;
  uint64_t   count;
  float*     final_pos;
  uint64_t*  final_rem;
  uint8_t*   final_status;
  uint64_t*  steps;
  uint64_t   total_steps;
};

static int clamp_threads(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

dpo_rounds* dpo_run_rounds(const dpo_grid* g, const float* seeds,
                           const uint64_t* remaining, uint64_t n,
                           uint64_t ppr, float dt, float t0, int integrator,
                           int record, int use_64_bit, int nthreads) {
  if (ppr == 0 && n > 0) return NULL;   /* dpa would loop forever */
  const int nt = clamp_threads(nthreads);
  dpo_rounds* R = (dpo_rounds*)calloc(1, sizeof(dpo_rounds));
  R->final_pos    = (float*)malloc(3 * (n ? n : 1) * sizeof(float));
  R->final_rem    = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
  R->final_status = (uint8_t*)malloc(n ? n : 1);
  R->steps        = (uint64_t*)calloc(n ? n : 1, sizeof(uint64_t));
  memcpy(R->final_pos, seeds, 3 * n * sizeof(float));
  memcpy(R->final_rem, remaining, n * sizeof(uint64_t));

  uint64_t active = n;
  uint64_t cap = 0;
  do {   /* pipeline.cpp:100-146: runs once even with no particle */
    const uint64_t count = ppr < active ? ppr : active;
    const uint64_t first = active - count;          /* tail of the vector (327) */
    uint64_t max_rem = 0;
    for (uint64_t i = 0; i < count; ++i)
      if (R->final_rem[first + i] > max_rem) max_rem = R->final_rem[first + i];
    const uint64_t stride = max_rem + 2;
    float* slots = NULL;
    if (record) {
      const uint64_t nslot = count * stride;
      slots = (float*)malloc(3 * (nslot > 0 ? nslot : 1) * sizeof(float));
      for (uint64_t i = 0; i < 3 * count * stride; ++i) slots[i] = -1.0f;
    }
    int64_t sc = (int64_t)count;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t i = 0; i < sc; ++i) {
      const uint64_t id = first + (uint64_t)i;
      R->steps[id] = dpo_advect_particle(g, R->final_pos + 3 * id, R->final_rem + id,
                                         dt, t0, integrator,
                                         slots ? slots + 3 * (uint64_t)i * stride : NULL,
                                         R->final_status + id);
    }
    active = first;
    if (record) {
      if (R->count == cap) {
        cap = cap ? 2 * cap : 4;
        R->rounds = (dpo_round*)realloc(R->rounds, cap * sizeof(dpo_round));
      }
      dpo_round* rd = &R->rounds[R->count++];
      rd->vertex_count = dpo_prune(slots, count * stride);
      rd->vertices = slots;
      rd->colors = (float*)malloc((rd->vertex_count ? rd->vertex_count : 1) * sizeof(float));
      const uint64_t ncap = 2 * rd->vertex_count;
      rd->indices = malloc((ncap > 0 ? ncap : 1) * (use_64_bit ? 8 : 4));
      rd->index_count = dpo_indices(slots, rd->vertex_count, rd->indices, use_64_bit);
      dpo_colors(slots, rd->vertex_count, rd->colors);
    }
  } while (active > 0);
  for (uint64_t i = 0; i < n; ++i) R->total_steps += R->steps[i];
  return R;
}

uint64_t dpo_rounds_count(const dpo_rounds* r) { return r->count; }
uint64_t dpo_rounds_vertex_count(const dpo_rounds* r, uint64_t i) { return r->rounds[i].vertex_count; }
uint64_t dpo_rounds_index_count(const dpo_rounds* r, uint64_t i) { return r->rounds[i].index_count; }
const float* dpo_rounds_vertices(const dpo_rounds* r, uint64_t i) { return r->rounds[i].vertices; }
const float* dpo_rounds_colors(const dpo_rounds* r, uint64_t i) { return r->rounds[i].colors; }
const void* dpo_rounds_indices(const dpo_rounds* r, uint64_t i) { return r->rounds[i].indices; }
const float* dpo_rounds_final_positions(const dpo_rounds* r) { return r->final_pos; }
const uint64_t* dpo_rounds_final_remaining(const dpo_rounds* r) { return r->final_rem; }
const uint8_t* dpo_rounds_final_status(const dpo_rounds* r) { return r->final_status; }
const uint64_t* dpo_rounds_steps(const dpo_rounds* r) { return r->steps; }
uint64_t dpo_rounds_total_steps(const dpo_rounds* r) { return r->total_steps; }

void dpo_rounds_free(dpo_rounds* r) {
  if (!r) return;
  for (uint64_t i = 0; i < r->count; ++i) {
    free(r->rounds[i].vertices); free(r->rounds[i].colors); free(r->rounds[i].indices);
  }
  free(r->rounds); free(r->final_pos); free(r->final_rem); free(r->final_status);
  free(r->steps); free(r);
}

uint64_t dpo_advect_endpoints(const dpo_grid* g, float* pos, uint64_t* remaining,
                              uint8_t* status, uint64_t n, float dt, float t0,
                              int integrator, int nthreads) {
  const int nt = clamp_threads(nthreads);
  uint64_t total = 0;
  int64_t sn = (int64_t)n;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(+ : total)
  for (int64_t i = 0; i < sn; ++i)
    total += dpo_advect_particle(g, pos + 3 * i, remaining + i, dt, t0, integrator,
                                 NULL, status + i);
  return total;
}

/* ----------------------------------------- dpa-structured CPU variant ---- */
/* The same arithmetic as dpo_interpolate, with the structure of dpa's
 * regular_grid::interpolate (regular_grid.hpp:52-75): a heap-allocated corner
 * list per call (the std::vector reserved at 65-66) filled by permute_for's
 * recursive walk through a callback (permute_for.hpp:12-34, dimension 0
 * outermost, one call per level and corner).  Used only as the CPU baseline
 * that costs what dpa's loop costs; its results equal the oracle's bit for
 * bit (tests/test_oracle.py). */

typedef struct {
  const dpo_grid* g;
  float* list;
  int count;
  int dims;
} corner_list;

typedef void (*corner_fn)(void* ctx, const int64_t idx[4]);

static void push_corner(void* c, const int64_t idx[4]) {
  corner_list* cl = (corner_list*)c;
  const float* v = grid_node(cl->g, idx[0], idx[1], idx[2], cl->dims == 4 ? idx[3] : 0);
  float* o = cl->list + 3 * cl->count++;
  o[0] = v[0]; o[1] = v[1]; o[2] = v[2];
}

static void permute_walk(int depth, int dims, int64_t idx[4], const int64_t start[4],
                         const int64_t end[4], corner_fn fn, void* ctx) {
  if (depth < dims) {
    for (int64_t i = start[depth]; i < end[depth]; ++i) {
      idx[depth] = i;
      permute_walk(depth + 1, dims, idx, start, end, fn, ctx);
    }
  } else {
    fn(ctx, idx);
  }
}

static void interpolate_dpa_structured(const dpo_grid* g, const float p[3], float t, float out[3]) {
  const int dims = grid_dims(g);
  float w[4];
  int64_t start[4] = {0, 0, 0, 0}, end[4] = {0, 0, 0, 0}, idx[4] = {0, 0, 0, 0};
  for (int i = 0; i < dims; ++i) {
    float o, h, x; int64_t n;
    grid_coord(g, i, p, t, &o, &h, &n, &x);
    const float d = x - o;
    w[i] = fmodf(d, h) / h;
    start[i] = (int64_t)floorf(d / h);
    end[i] = start[i] + 2;
  }
  corner_list cl = {g, (float*)malloc(sizeof(float) * 3 * (1u << dims)), 0, dims};
  permute_walk(0, dims, idx, start, end, push_corner, &cl);
  float* L = cl.list;
  for (int i = dims - 1; i >= 0; --i) {
    const float wi = w[i];
    const float om = 1.0f - wi;
    for (int j = 0; j < (1 << i); ++j)
      for (int e = 0; e < 3; ++e) {
        const float a = om * L[3 * (2 * j) + e];
        const float b = wi * L[3 * (2 * j + 1) + e];
        L[3 * j + e] = a + b;
      }
  }
  out[0] = L[0]; out[1] = L[1]; out[2] = L[2];
  free(L);
}

/* dpo_advect_particle's loop (dpa mode) with the dpa-structured interpolation. */
static uint64_t advect_particle_dpa_structured(const dpo_grid* g, float p[3], uint64_t* remaining,
                                               float dt, float t0, uint8_t* status) {
  uint64_t it = 0;
  uint8_t st = DPO_ACTIVE;
  for (; *remaining > 0; ++it, --*remaining) {
    const float ft = (float)it * dt;
    const float t  = t0 + ft;
    if (!dpo_contains(g, p, t)) { st = DPO_EXITED; break; }
    float v[3];
    interpolate_dpa_structured(g, p, t, v);
    if (dpo_is_zero(v)) { st = DPO_ZERO; break; }
    dpo_step_dpa(p, v, dt);
  }
  if (*remaining == 0) st = DPO_COMPLETED;
  *status = st;
  return it;
}

uint64_t dpo_advect_endpoints_dpa_structured(const dpo_grid* g, float* pos, uint64_t* remaining,
                                             uint8_t* status, uint64_t n, float dt, float t0,
                                             int nthreads) {
  const int nt = clamp_threads(nthreads);
  uint64_t total = 0;
  int64_t sn = (int64_t)n;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(+ : total)
  for (int64_t i = 0; i < sn; ++i)
    total += advect_particle_dpa_structured(g, pos + 3 * i, remaining + i, dt, t0, status + i);
  return total;
}

/* ------------------------------------------------------------- seeds ---- */

uint64_t dpo_lattice_seeds(const float offset[3], const float size[3],
                           const float stride[3], uint64_t per_dim[3], float* out) {
  uint64_t d[3];
  for (int i = 0; i < 3; ++i) {
    const float q = size[i] / stride[i];
    d[i] = q > 0.0f ? (uint64_t)q : 0;    /* (size / stride).cast<size_t>() (30) */
    if (per_dim) per_dim[i] = d[i];
  }
  const uint64_t count = d[0] * d[1] * d[2];
  if (out)
    for (uint64_t index = 0; index < count; ++index) {
      uint64_t sub[3], rest = index;
      for (int i = 2; i >= 0; --i) { sub[i] = rest % d[i]; rest /= d[i]; }
      for (int i = 0; i < 3; ++i) {
        const float m = stride[i] * (float)sub[i];
        out[3 * index + i] = offset[i] + m;
      }
    }
  return count;
}

void dpo_aux_seeds(const float origin[3], const float spacing[3], const uint64_t shape[3],
                   const float delta[3], float* out) {
  const uint64_t count = shape[0] * shape[1] * shape[2];
  for (uint64_t index = 0; index < count; ++index) {
    uint64_t sub[3], rest = index;
    for (int i = 2; i >= 0; --i) { sub[i] = rest % shape[i]; rest /= shape[i]; }
    float base[3];
    for (int i = 0; i < 3; ++i) {
      const float m = spacing[i] * (float)sub[i];
      base[i] = origin[i] + m;
    }
    for (int a = 0; a < 6; ++a) {
      float* p = out + 3 * (6 * index + (uint64_t)a);
      const int ax = a / 2;
      for (int i = 0; i < 3; ++i) p[i] = base[i];
      const float m = fmaf(spacing[ax], (float)sub[ax], (a & 1) ? delta[ax] : -delta[ax]);
      p[ax] = origin[ax] + m;
    }
  }
}

/* -------------------------------------------------------------- ftle ---- */

double dpo_sym3_max_eigenvalue(const double m[9]) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = 0.5 * (m[3 * i + j] + m[3 * j + i]);
  for (int sweep = 0; sweep < 64; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off == 0.0 || off <= 1e-40 * dia) break;
    static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int r = 0; r < 3; ++r) {
      const int p = P[r], q = Q[r];
      if (a[p][q] == 0.0) continue;
      const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
      const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      for (int k = 0; k < 3; ++k) {        /* A <- A G (columns p, q) */
        const double akp = a[k][p], akq = a[k][q];
        a[k][p] = c * akp - s * akq;
        a[k][q] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; ++k) {        /* A <- G^T A (rows p, q) */
        const double apk = a[p][k], aqk = a[q][k];
        a[p][k] = c * apk - s * aqk;
        a[q][k] = s * apk + c * aqk;
      }
      a[p][q] = a[q][p] = 0.0;
    }
  }
  double mx = a[0][0];
  if (a[1][1] > mx) mx = a[1][1];
  if (a[2][2] > mx) mx = a[2][2];
  return mx;
}

static double ftle_of_jacobian(const double J[3][3], double T) {
  double C[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C[3 * i + j] = J[0][i] * J[0][j] + J[1][i] * J[1][j] + J[2][i] * J[2][j];
  const double l = dpo_sym3_max_eigenvalue(C);
  return log(sqrt(l)) / T;
}

void dpo_ftle_dpa(const int64_t field_n[3], const float offset[3],
                  const float spacing[3], const float stride[3],
                  uint64_t iterations, float dt, const float* orig,
                  const float* fin, uint64_t count, float* out,
                  uint64_t out_shape[3], uint64_t* collisions) {
  uint64_t m[3];
  float ss[3];
  for (int i = 0; i < 3; ++i) {
    m[i]  = (uint64_t)((float)field_n[i] / stride[i]);   /* ftle_estimator.cpp:16-21 */
    ss[i] = spacing[i] * stride[i];                       /* (22) */
    if (out_shape) out_shape[i] = m[i];
  }
  const uint64_t total = m[0] * m[1] * m[2];
  if (!out) return;
  float*    flow  = (float*)calloc(3 * (total ? total : 1), sizeof(float));
  int64_t*  owner = (int64_t*)malloc((total ? total : 1) * sizeof(int64_t));
  for (uint64_t i = 0; i < total; ++i) owner[i] = -1;
  uint64_t coll = 0;
  for (uint64_t p = 0; p < count; ++p) {                 /* scatter (39-43) */
    int64_t idx[3]; int ok = 1;
    for (int i = 0; i < 3; ++i) {
      const float d = orig[3 * p + i] - offset[i];
      const float s = floorf(d / ss[i]);                  /* cell_index (regular_grid.hpp:28-34) */
      if (!(s >= 0.0f) || s >= (float)m[i]) { ok = 0; break; }
      idx[i] = (int64_t)s;
    }
    if (!ok) continue;
    const uint64_t lin = ((uint64_t)idx[0] * m[1] + (uint64_t)idx[1]) * m[2] + (uint64_t)idx[2];
    if (owner[lin] >= 0) ++coll;
    owner[lin] = (int64_t)p;
    flow[3 * lin] = fin[3 * p]; flow[3 * lin + 1] = fin[3 * p + 1]; flow[3 * lin + 2] = fin[3 * p + 2];
  }
  if (collisions) *collisions = coll;
  const float  Tf = dt * (float)iterations;               /* step_size * iterations (57) */
  const double T  = fabs((double)Tf);
  for (uint64_t i0 = 0; i0 < m[0]; ++i0)
    for (uint64_t i1 = 0; i1 < m[1]; ++i1)
      for (uint64_t i2 = 0; i2 < m[2]; ++i2) {
        const uint64_t id[3] = {i0, i1, i2};
        double J[3][3];
        for (int d = 0; d < 3; ++d) {                     /* gradient (regular_grid.hpp:95-105) */
          uint64_t pv[3] = {i0, i1, i2}, nx[3] = {i0, i1, i2};
          if (id[d] > 0) pv[d] -= 1;
          if (id[d] < m[d] - 1) nx[d] += 1;
          const float* a = flow + 3 * ((pv[0] * m[1] + pv[1]) * m[2] + pv[2]);
          const float* b = flow + 3 * ((nx[0] * m[1] + nx[1]) * m[2] + nx[2]);
          const double two_h = 2.0 * (double)ss[d];
          for (int e = 0; e < 3; ++e) J[e][d] = ((double)b[e] - (double)a[e]) / two_h;
        }
        out[(i0 * m[1] + i1) * m[2] + i2] = (float)ftle_of_jacobian(J, T);
      }
  free(flow); free(owner);
}

void dpo_ftle_aux(const float* fin, uint64_t points, const float delta[3],
                  double T, float* out) {
  for (uint64_t p = 0; p < points; ++p) {
    const float* f = fin + 18 * p;
    double J[3][3];
    for (int d = 0; d < 3; ++d) {
      const float* lo = f + 6 * d;
      const float* hi = f + 6 * d + 3;
      const double den = 2.0 * (double)delta[d];
      for (int e = 0; e < 3; ++e) J[e][d] = ((double)hi[e] - (double)lo[e]) / den;
    }
    out[p] = (float)ftle_of_jacobian(J, fabs(T));
  }
}
