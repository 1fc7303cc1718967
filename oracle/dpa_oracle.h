/*
 * dpa_oracle.h -- CPU restatement of dpa's particle-advection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py load it to
 * compute expected results.  The product library (libdpa_hip.so) never links,
 * loads or calls it, and nothing here is ever the thing that is measured as the
 * GPU path.
 *
 * Every function restates the reference behaviour cited next to it
 * (paths relative to the dpa source tree).  Arithmetic is IEEE binary32 with
 * one rounding per operation and no contraction (built with -O2
 * -ffp-contract=off, no -march, no fast-math), which is what dpa's Release
 * build (CMakeLists.txt:7-10) does on x86-64.
 *
 * Pins (what checks each function beyond retyping it) are listed in the
 * header comment of oracle/dpa_oracle.c.
 */
#ifndef DPA_ORACLE_H
#define DPA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Regular grid of float3 samples, C order, axis 0 slowest
 * (regular_grid.hpp:148-151, regular_grid_loader.cpp:65-88).
 * nt == 1: steady field, data = float[n0][n1][n2][3].
 * nt  > 1: time-variant field (regular_fields.hpp:27, regular_grid<vector3,4>),
 *          stored as nt slices: data = float[nt][n0][n1][n2][3]; time is the
 *          grid's fourth axis with origin t_offset and spacing t_spacing. */
typedef struct dpo_grid {
  const float* data;
  int64_t      n[3];
  int64_t      nt;
  float        offset[3];
  float        spacing[3];
  float        t_offset;
  float        t_spacing;
} dpo_grid;

enum { DPO_INTEGRATOR_DPA_RK4 = 0, DPO_INTEGRATOR_RK4 = 1 };
enum { DPO_ACTIVE = 0, DPO_COMPLETED = 1, DPO_EXITED = 2, DPO_ZERO = 3 };

/* regular_grid::contains (regular_grid.hpp:41-50); t used only when nt > 1. */
int  dpo_contains   (const dpo_grid* g, const float p[3], float t);
/* regular_grid::interpolate (regular_grid.hpp:52-75 with permute_for.hpp:12-34). */
void dpo_interpolate(const dpo_grid* g, const float p[3], float t, float out[3]);
/* Eigen isZero() with float dummy_precision 1e-5 (particle_advector.cpp:363). */
int  dpo_is_zero    (const float v[3]);
/* One dpa runge_kutta_4 step with a single field sample
 * (particle_advector.cpp:369-375, odeint runge_kutta4 over vector_space_algebra). */
void dpo_step_dpa   (float x[3], const float v[3], float dt);

/* Per-particle loop body (particle_advector.cpp:325-396).  rec (nullable) gets
 * the seed at rec[0], one vertex per step and the terminal vertex.  Returns the
 * number of steps taken; *status gets DPO_COMPLETED / EXITED / ZERO. */
uint64_t dpo_advect_particle(const dpo_grid* g, float p[3], uint64_t* remaining,
                             float dt, float t0, int integrator, float* rec,
                             uint8_t* status);

/* The single-rank round loop of pipeline.cpp:98-147 with load_balancer none:
 * compute_round_state (particle_advector.cpp:261-310), allocate (311-316),
 * advect (317-401), prune (402-408); then index_generator (index_generator.cpp:
 * 10-39) and color_generator::generate_from_angular_velocities
 * (color_generator.cpp:37-58) over every round.  nthreads <= 0 uses OpenMP's
 * default; the parallel loop is schedule(dynamic,1), as tbb::parallel_for with
 * grain 1. */
typedef struct dpo_rounds dpo_rounds;
dpo_rounds*     dpo_run_rounds(const dpo_grid* g, const float* seeds,
                               const uint64_t* remaining, uint64_t n,
                               uint64_t particles_per_round, float dt, float t0,
                               int integrator, int record, int use_64_bit,
                               int nthreads);
uint64_t        dpo_rounds_count        (const dpo_rounds* r);
uint64_t        dpo_rounds_vertex_count (const dpo_rounds* r, uint64_t round);
uint64_t        dpo_rounds_index_count  (const dpo_rounds* r, uint64_t round);
const float*    dpo_rounds_vertices     (const dpo_rounds* r, uint64_t round);
const float*    dpo_rounds_colors       (const dpo_rounds* r, uint64_t round);
const void*     dpo_rounds_indices      (const dpo_rounds* r, uint64_t round);
/* Final particle states indexed by seed id (the inactive set, particle_advector.hpp:87). */
const float*    dpo_rounds_final_positions(const dpo_rounds* r);
const uint64_t* dpo_rounds_final_remaining(const dpo_rounds* r);
const uint8_t*  dpo_rounds_final_status   (const dpo_rounds* r);
const uint64_t* dpo_rounds_steps          (const dpo_rounds* r);
uint64_t        dpo_rounds_total_steps    (const dpo_rounds* r);
void            dpo_rounds_free           (dpo_rounds* r);

/* Endpoint-only advection of n independent particles (record off), in place.
 * Returns the total number of steps.  Used for FTLE and for the CPU baseline. */
uint64_t dpo_advect_endpoints(const dpo_grid* g, float* pos, uint64_t* remaining,
                              uint8_t* status, uint64_t n, float dt, float t0,
                              int integrator, int nthreads);
/* The same in dpa's structure (heap corner list, recursive corner walk per
 * interpolation), dpa mode only: a CPU baseline; equal results. */
uint64_t dpo_advect_endpoints_dpa_structured(const dpo_grid* g, float* pos, uint64_t* remaining,
                                             uint8_t* status, uint64_t n, float dt, float t0,
                                             int nthreads);

/* Literal post-processing stages on one round's vertex array. */
uint64_t dpo_prune(float* vertices, uint64_t count);                     /* particle_advector.cpp:402-408 */
uint64_t dpo_indices(const float* vertices, uint64_t count, void* out, int use_64_bit); /* index_generator.cpp:10-39 */
void     dpo_colors (const float* vertices, uint64_t count, float* out); /* color_generator.cpp:37-58 */

/* Lattice seeds (uniform_seed_generator.cpp:16-45, indexing.hpp:10-20).
 * per_dim[i] = size_t(size[i] / stride[i]); returns the count; out (nullable)
 * gets count x 3 floats in C order of the lattice. */
uint64_t dpo_lattice_seeds(const float offset[3], const float size[3],
                           const float stride[3], uint64_t per_dim[3], float* out);

/* Seeds of the auxiliary-neighbour FTLE (the FTLE benchmark's +-delta mode,
 * no dpa counterpart): per lattice point i = (i0, i1, i2) of a d0 x d1 x d2
 * lattice in C order, six particles -delta_0, +delta_0, -delta_1, +delta_1,
 * -delta_2, +delta_2; each is the lattice point origin + fl(spacing * i)
 * (uniform_seed_generator.cpp:35-36) with the one moved coordinate
 * origin + fl(spacing * i +- delta): the offset position rounded once (fmaf),
 * so that with delta = spacing / 2 the +delta particle of point k and the
 * -delta particle of point k + 1 are the same float.  out: points x 6 x 3. */
void dpo_aux_seeds(const float origin[3], const float spacing[3], const uint64_t shape[3],
                   const float delta[3], float* out);

/* ftle_estimator::estimate (ftle_estimator.cpp:7-65) with
 * regular_grid::gradient (regular_grid.hpp:87-108), restated in float64 from
 * the float32 flow map: scatter final positions to the cell of their original
 * positions, central differences with clamped borders and the 2h divisor,
 * C = J^T J, largest eigenvalue, ln(sqrt(lmax)) / |dt * iterations|.
 * shape/out are the strided lattice; on a scatter collision the highest
 * particle index wins and *collisions counts them. */
void dpo_ftle_dpa(const int64_t field_n[3], const float offset[3],
                  const float spacing[3], const float stride[3],
                  uint64_t iterations, float dt, const float* original_positions,
                  const float* final_positions, uint64_t count, float* out,
                  uint64_t out_shape[3], uint64_t* collisions);

/* Auxiliary-neighbour FTLE: finals holds, per lattice point, the final
 * positions of its six neighbours at -/+ delta_i along axes 0,1,2 (in that
 * order).  J[:,d] = (phi(+d) - phi(-d)) / (2 delta_d), float64 thereafter. */
void dpo_ftle_aux(const float* finals, uint64_t points, const float delta[3],
                  double T, float* out);

/* Largest eigenvalue of a symmetric 3x3 matrix (row major, float64, cyclic
 * Jacobi to convergence). */
double dpo_sym3_max_eigenvalue(const double m[9]);

#ifdef __cplusplus
}
#endif
#endif
