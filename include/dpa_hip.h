/*
 * dpa_hip.h -- C ABI of the MI355X particle-advection library (libdpa_hip.so).
 *
 * Drop-in boundary for one dpa rank's advection stage.  dpa has no C API: the
 * entry points below replace the C++ stage objects a single rank of dpa's
 * pipeline calls (source/pipeline.cpp:91-171), with plain pointers and sizes.
 * Each declaration cites the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns DPAH_OK (0) or a DPAH_ERR_* code; no C++
 *     exception crosses the ABI.  dpah_last_error() gives the calling
 *     thread's last message.
 *   - Positions are float32 (x0, x1, x2) triples in dpa's component order:
 *     component i indexes grid axis i (regular_grid.hpp:32, 45, 61-62).
 *   - Fields are AoS float32 [n0][n1][n2][3], axis 0 slowest, exactly the
 *     layout dpa's loader reads (regular_grid_loader.cpp:65-88).  The vector
 *     applied to axis i is stored component 2-i (particle_advector.cpp:369).
 *   - Functions named *_device take device pointers (from dpah_device_alloc);
 *     all others take host pointers.
 *   - A context serialises its own calls; use one context per thread for
 *     concurrency.
 */
#ifndef DPA_HIP_H
#define DPA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPAH_OK              0
#define DPAH_ERR_INVALID     1  /* bad argument                               */
#define DPAH_ERR_HIP         2  /* HIP runtime error                          */
#define DPAH_ERR_NOMEM       3  /* allocation failed or budget exceeded       */
#define DPAH_ERR_ADMISSION   4  /* request does not fit the memory budget     */
#define DPAH_ERR_CANCELLED   5  /* request was cancelled                      */
#define DPAH_ERR_UNSUPPORTED 6  /* configuration not supported                */
#define DPAH_ERR_TIMEOUT     7  /* wait timed out                             */
#define DPAH_ERR_PENDING     8  /* result not ready yet                       */

/* particle_advector_integrator values (particle_advector.cpp:29-36).
 * DPA_RK4 is dpa's "runge_kutta_4": one field sample per step, odeint's
 * four-term sum.  RK4 samples four times (classic RK4; not a dpa mode). */
#define DPAH_INTEGRATOR_DPA_RK4 0
#define DPAH_INTEGRATOR_RK4     1

/* Particle status after advection (where dpa moves the particle to
 * output.inactive_particles, particle_advector.cpp:353, 365, 395). */
#define DPAH_STATUS_ACTIVE    0
#define DPAH_STATUS_COMPLETED 1  /* remaining_iterations reached 0        */
#define DPAH_STATUS_EXITED    2  /* failed regular_grid::contains         */
#define DPAH_STATUS_ZERO      3  /* interpolated vector isZero()          */

typedef struct dpah_context dpah_context;
typedef struct dpah_field   dpah_field;
typedef struct dpah_result  dpah_result;

typedef struct dpah_options {
  uint64_t    memory_budget_bytes;  /* 0: 90% of free device memory at create */
  uint64_t    slot_budget_bytes;    /* record-mode staging per chunk; 0: 4 GiB */
  int         kernel_timing;        /* 1: time advect kernels with HIP events */
  const char* autotune_cache_path;  /* nullable */
} dpah_options;

/* Last error message of the calling thread ("" if none). */
const char* dpah_last_error(void);
int dpah_version(void);

/* Context on one GPU (replaces the per-rank process set up at pipeline.cpp:24-47). */
int dpah_create(int device, const dpah_options* options, dpah_context** out);
int dpah_destroy(dpah_context* ctx);
int dpah_synchronize(dpah_context* ctx);

/* Device memory helpers for the *_device entry points. */
int dpah_device_alloc(dpah_context* ctx, uint64_t bytes, void** out);
int dpah_device_free(dpah_context* ctx, void* ptr);
int dpah_copy_to_device(dpah_context* ctx, void* dst, const void* src, uint64_t bytes);
int dpah_copy_to_host(dpah_context* ctx, void* dst, const void* src, uint64_t bytes);

/* regular_grid_loader::load_vector_field (regular_grid_loader.cpp:65-88) and the
 * regular_vector_field_3d it returns (regular_fields.hpp:14): copies the field
 * to HBM, where it stays resident until released.  offset/spacing are the
 * grid's physical origin and node spacing (offset = ghosted_offset * spacing,
 * 86; on one rank 0). */
int dpah_field_create(dpah_context* ctx, const float* xyzv, const uint64_t shape[3],
                      const float offset[3], const float spacing[3], dpah_field** out);
/* Time-variant field regular_time_variant_vector_field_3d (regular_fields.hpp:27):
 * time_steps slices of [n0][n1][n2][3], slice t at time_offset + t*time_spacing. */
int dpah_field_create_unsteady(dpah_context* ctx, const float* xyzv, const uint64_t shape[3],
                               uint64_t time_steps, const float offset[3],
                               const float spacing[3], float time_offset, float time_spacing,
                               dpah_field** out);
int dpah_field_release(dpah_field* field);

/* Constructor arguments of particle_advector (particle_advector.hpp:124-131)
 * that the single-GPU path uses, plus index_generator's width flag
 * (index_generator.hpp:11; dpa's rule is pipeline.cpp:155). */
typedef struct dpah_advect_params {
  uint64_t particles_per_round;  /* particle_advector_particles_per_round */
  float    step_size;            /* particle_advector_step_size           */
  float    start_time;           /* time-variant fields: t at iteration 0 */
  int      integrator;           /* DPAH_INTEGRATOR_*                     */
  int      record;               /* particle_advector_record              */
  int      use_64_bit_indices;   /* index_generator::generate use_64_bit  */
} dpah_advect_params;

/* One rank's advection stage with load_balancer "none": the round loop of
 * pipeline.cpp:98-147 (compute_round_state, allocate_integral_curves, advect,
 * prune_integral_curves; particle_advector.cpp:261-408) followed by
 * index_generator::generate and color_generator::generate_from_angular_velocities
 * (pipeline.cpp:153-159).  seeds: count x 3 floats; remaining: count values or
 * NULL to give every particle uniform_remaining (seed_generation_iterations).
 * Rounds take the tail of the seed array first, as dpa does (327). */
int dpah_advect(dpah_context* ctx, const dpah_field* field, const float* seeds,
                const uint64_t* remaining, uint64_t uniform_remaining, uint64_t count,
                const dpah_advect_params* params, dpah_result** out);

/* One seed set of a batch: the arguments of one dpah_advect call. */
typedef struct dpah_job {
  const dpah_field*  field;
  const float*       seeds;              /* host, count x 3 */
  const uint64_t*    remaining;          /* host, count, or NULL */
  uint64_t           uniform_remaining;
  uint64_t           count;
  dpah_advect_params params;
} dpah_job;

/* Many independent advection stages at once (no counterpart in dpa, whose
 * ranks balance load with diffusion between neighbours,
 * particle_advector.cpp:65-260).  Jobs sharing field, step size, start time,
 * integrator and record flag run in ONE persistent launch whose on-device work
 * queue holds every round of every job, longest rows first.  results[i] gets a
 * result equal, bit for bit, to dpah_advect on jobs[i] alone. */
int dpah_advect_batch(dpah_context* ctx, const dpah_job* jobs, uint64_t job_count,
                      dpah_result** results);

/* The per-round integral_curve (integral_curves.hpp:11-18) and the inactive
 * particle set (particle_advector.hpp:85-89) of a finished advection.
 * index_count counts index entries (two per segment), like indices.size(). */
int dpah_result_round_count(const dpah_result* r, uint64_t* rounds);
int dpah_result_round_sizes(const dpah_result* r, uint64_t round, uint64_t* vertex_count,
                            uint64_t* index_count, int* index_bytes);
int dpah_result_copy_round(const dpah_result* r, uint64_t round, float* vertices, float* colors,
                           void* indices);
int dpah_result_final_state(const dpah_result* r, float* positions, uint64_t* remaining,
                            uint8_t* status);
int dpah_result_total_steps(const dpah_result* r, uint64_t* steps);
int dpah_result_free(dpah_result* r);

/* Endpoint advection (record off) on device-resident particles, the form the
 * FTLE and batched paths use: the particle loop of particle_advector.cpp:325-396
 * over count particles.  seeds/positions_out are device count x 3 floats;
 * remaining may be NULL (uniform_remaining for all).  remaining_out / status_out
 * may be NULL.  If total_steps is non-NULL the call waits for the kernel and
 * stores the number of particle-steps taken; otherwise it returns after the
 * launch (stream-ordered on the context's stream). */
int dpah_advect_endpoints_device(dpah_context* ctx, const dpah_field* field, const float* seeds,
                                 const uint64_t* remaining, uint64_t uniform_remaining,
                                 uint64_t count, float step_size, float start_time,
                                 int integrator, float* positions_out, uint64_t* remaining_out,
                                 uint8_t* status_out, uint64_t* total_steps);

/* Allocation-free form of dpah_advect_endpoints_device.  No counterpart in dpa
 * (whose CPU loop allocates freely): it is for callers that cannot allocate
 * device memory -- code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  Same arguments and results, bit for bit; it calls no
 * hipMalloc / hipFree and launches untimed (kernel_timing records nothing for
 * it).  Its scratch (step counter, queue heads, queue-ordering keys, order
 * table and histogram) is one fixed device array of DPAH_NOALLOC_SCRATCH_BYTES
 * per GPU, held from dpah_create until the process ends, so count is limited
 * to DPAH_NOALLOC_MAX_PARTICLES (larger: DPAH_ERR_INVALID).  Every call on a
 * GPU shares that GPU's array: calls from one context are stream-ordered, but
 * calls that overlap on one GPU from two contexts race. */
#define DPAH_NOALLOC_MAX_PARTICLES (1ull << 21)
#define DPAH_NOALLOC_SCRATCH_BYTES (18ull << 20)
int dpah_advect_endpoints_device_noalloc(dpah_context* ctx, const dpah_field* field,
                                         const float* seeds, const uint64_t* remaining,
                                         uint64_t uniform_remaining, uint64_t count,
                                         float step_size, float start_time, int integrator,
                                         float* positions_out, uint64_t* remaining_out,
                                         uint8_t* status_out, uint64_t* total_steps);

/* ftle_estimator::estimate (ftle_estimator.cpp:7-65) on device data: the final
 * positions are scattered onto the strided lattice of the original positions
 * (shape size_t(n_i / stride_i), spacing h_i * stride_i, cell of
 * floor((p - offset) / spacing)), differentiated by regular_grid::gradient
 * (regular_grid.hpp:87-108: clamped neighbours, divisor 2 * spacing), and
 * ln(sqrt(lmax(J^T J))) / |step_size * iterations| is written per lattice cell.
 * field_shape/offset/spacing describe the advected field; stride is in cells.
 * out_shape always receives the lattice shape; ftle_out (device, may be NULL
 * to query the shape) receives out_shape[0]*out_shape[1]*out_shape[2] floats.
 * Two particles landing in one cell: the higher particle index wins (dpa leaves
 * the order to TBB). */
int dpah_ftle_estimate_device(dpah_context* ctx, const uint64_t field_shape[3],
                              const float offset[3], const float spacing[3],
                              const float stride[3], uint64_t iterations, float step_size,
                              const float* original_positions, const float* final_positions,
                              uint64_t count, float* ftle_out, uint64_t out_shape[3]);

/* One rank's FTLE stage as dpa's pipeline runs it (pipeline.cpp:65-72, 91-147,
 * 165-167): lattice seeds of uniform_seed_generator::generate
 * (uniform_seed_generator.cpp:16-45) with stride in cells over the whole field,
 * endpoint advection for `iterations` steps, then ftle_estimator::estimate.
 * ftle_out is host memory (NULL: only out_shape is filled). */
int dpah_ftle_lattice(dpah_context* ctx, const dpah_field* field, const float stride[3],
                      uint64_t iterations, float step_size, int integrator, float* ftle_out,
                      uint64_t out_shape[3], uint64_t* total_steps);

/* FTLE with auxiliary neighbours (no counterpart in dpa; replaces the scatter
 * and one-sided borders of ftle_estimator.cpp:39-45 by central differences
 * that hold at the border): for every lattice point p = origin + spacing * i
 * (C order over lattice_shape, each coordinate origin + fl(spacing * i))
 * particles start at p -/+ delta_d along each axis d, that coordinate being
 * origin_d + fl(spacing_d * i_d -/+ delta_d) (the offset rounded once);
 * J[:, d] = (phi(p + delta_d) - phi(p - delta_d)) / (2 delta_d);
 * ftle_out (device, one float per point) = ln(sqrt(lmax(J^T J))) /
 * |step_size * iterations|.  When delta is half the spacing on every axis,
 * point i's +delta particle and point i + e_d's -delta particle are the same
 * float and are advected once (three particles per point plus the low faces'
 * -delta particles); *total_steps counts the particle-steps advected, a
 * shared particle once (a lattice processed in chunks -- above 2^25 points,
 * or finals above a quarter of the memory budget -- repeats one slab of
 * particles per chunk boundary and counts it again). */
int dpah_ftle_aux_device(dpah_context* ctx, const dpah_field* field, const float origin[3],
                         const float lattice_spacing[3], const uint64_t lattice_shape[3],
                         const float delta[3], uint64_t iterations, float step_size,
                         float start_time, int integrator, float* ftle_out,
                         uint64_t* total_steps);

/* ------------------------------------------------------------- serving ----
 * No counterpart in dpa (a batch pipeline): a request loop on one GPU that
 * keeps fields resident, admits advection and FTLE queries by memory and
 * batches them continuously.  One engine thread owns the device.  In
 * DPAH_SERVE_BATCHED mode admitted requests are launched as soon as one of
 * three streams is free; the requests waiting at that moment that share
 * (field, integrator, record) go into ONE launch, which advances each particle
 * by up to steps_per_tick steps (0: to retirement) and puts unfinished
 * requests back in line; results are compacted on the launch's stream.
 * With steps_per_tick 0 an FTLE request runs as its own persistent launch of
 * the auxiliary lattice (the kernel dpah_ftle_aux_device runs) on one of the
 * streams.  Two comparators run each request alone in arrival order:
 * DPAH_SERVE_PER_REQUEST allocates (hipMalloc), uploads, launches one thread
 * per particle, compacts and frees (hipFree) per request; DPAH_SERVE_SERIAL
 * goes through the single-request path (dpah_advect / dpah_ftle_aux_device,
 * caching allocator, persistent kernel).  results_on_host applies to every
 * mode alike.  A request's result is the same in every mode and mix, bit for
 * bit. */
#define DPAH_REQUEST_ADVECT   0   /* dpah_advect arguments                    */
#define DPAH_REQUEST_FTLE_AUX 1   /* dpah_ftle_aux_device arguments           */

#define DPAH_SERVE_BATCHED     0
#define DPAH_SERVE_PER_REQUEST 1
#define DPAH_SERVE_SERIAL      2

#define DPAH_TICKET_PENDING   0
#define DPAH_TICKET_RUNNING   1
#define DPAH_TICKET_DONE      2
#define DPAH_TICKET_FAILED    3
#define DPAH_TICKET_CANCELLED 4

typedef struct dpah_server dpah_server;

typedef struct dpah_server_options {
  uint64_t memory_budget_bytes;  /* device bytes for admitted requests; 0: half the context's */
  uint32_t max_active;           /* requests admitted at once; 0: 256                        */
  uint32_t steps_per_tick;       /* per particle per launch; 0: run to retirement            */
  int      mode;                 /* DPAH_SERVE_*                                              */
  int      results_on_host;      /* 1: each result is copied to host memory before its ticket
                                    is ready; 0: results stay in device memory (every mode)   */
} dpah_server_options;

typedef struct dpah_request {
  int                kind;               /* DPAH_REQUEST_*                              */
  const dpah_field*  field;
  /* DPAH_REQUEST_ADVECT (host arrays, copied at submit) */
  const float*       seeds;
  const uint64_t*    remaining;
  uint64_t           uniform_remaining;
  uint64_t           count;
  dpah_advect_params params;             /* FTLE uses step_size, start_time, integrator */
  /* DPAH_REQUEST_FTLE_AUX */
  float              origin[3];
  float              lattice_spacing[3];
  uint64_t           lattice_shape[3];
  float              delta[3];
  uint64_t           iterations;
} dpah_request;

typedef struct dpah_request_info {
  int      state;                /* DPAH_TICKET_*                         */
  double   queue_seconds;        /* submit -> admitted                    */
  double   latency_seconds;      /* submit -> result ready (on device)    */
  uint64_t particle_steps;
  uint64_t footprint_bytes;
  double   launch_seconds;       /* submit -> first kernel launch enqueued */
  double   device_seconds;       /* submit -> last advection kernel seen done */
} dpah_request_info;

typedef struct dpah_server_stats {
  uint64_t submitted, completed, rejected, cancelled, failed, pending, ticks;
  uint64_t particle_steps;
  uint64_t bytes_in_use;
  double   busy_seconds;
} dpah_server_stats;

int dpah_server_create(dpah_context* ctx, const dpah_server_options* options, dpah_server** out);
/* Queues a request; DPAH_ERR_ADMISSION if it can never fit the budget. */
int dpah_server_submit(dpah_server* server, const dpah_request* request, uint64_t* ticket);
int dpah_server_poll(dpah_server* server, uint64_t ticket, int* state);
/* Waits (timeout_ms < 0: forever) and hands over the result (free it with
 * dpah_result_free); DPAH_ERR_TIMEOUT / DPAH_ERR_CANCELLED / the failure code. */
int dpah_server_wait(dpah_server* server, uint64_t ticket, int64_t timeout_ms, dpah_result** out);
int dpah_server_request_info(dpah_server* server, uint64_t ticket, dpah_request_info* info);
int dpah_server_cancel(dpah_server* server, uint64_t ticket);
/* Drops a finished ticket's bookkeeping (and its result if not taken). */
int dpah_server_forget(dpah_server* server, uint64_t ticket);
int dpah_server_get_stats(dpah_server* server, dpah_server_stats* stats);
/* Cancels what is still queued or running, stops the engine, frees. */
int dpah_server_destroy(dpah_server* server);
/* FTLE volume of a DPAH_REQUEST_FTLE_AUX result (out may be NULL: count only). */
int dpah_result_ftle(const dpah_result* r, float* out, uint64_t* count);

/* A repeated endpoint pass captured into a hipGraph: the arguments of
 * dpah_advect_endpoints_device (device pointers, fixed for the pass's life),
 * captured once (counter reset, queue ordering, persistent advect kernel) and
 * replayed by dpah_pass_launch on the context's stream with no host launch
 * work and no allocation.  Each replay recomputes every output from the seeds.
 * dpah_pass_total_steps waits and returns the last replay's particle-steps.
 * The field and the buffers must outlive the pass.  No counterpart in dpa (one
 * CPU pass per run). */
typedef struct dpah_pass dpah_pass;
int dpah_pass_create_endpoints(dpah_context* ctx, const dpah_field* field, const float* seeds,
                               const uint64_t* remaining, uint64_t uniform_remaining,
                               uint64_t count, float step_size, float start_time, int integrator,
                               float* positions_out, uint64_t* remaining_out, uint8_t* status_out,
                               dpah_pass** out);
int dpah_pass_launch(dpah_pass* pass);
int dpah_pass_total_steps(dpah_pass* pass, uint64_t* steps);
int dpah_pass_destroy(dpah_pass* pass);

/* A repeated batch captured into a hipGraph: the jobs of dpah_advect_batch
 * (host arrays, read at creation only) planned, uploaded and captured once --
 * per group of jobs sharing field, step size, start time, integrator and
 * record flag: the counter reset, the persistent launch over the group's
 * segment table and, for recorded groups, every round's vertex / index counts
 * and scan.  dpah_batch_pass_launch replays them on the context's stream with
 * no host launch work and no allocation; each replay recomputes every output
 * from the uploaded seeds.  dpah_batch_pass_results waits for the last
 * replay, reads the totals and final states back and writes every round's
 * vertices, colours and indices (sizes the replay decides, so this part is
 * not captured), returning one result per job exactly as dpah_advect_batch
 * does; the caller frees them with dpah_result_free.  The jobs' fields must
 * outlive the pass (the graph holds their device addresses).  No counterpart
 * in dpa (one CPU pass per run). */
typedef struct dpah_batch_pass dpah_batch_pass;
int dpah_batch_pass_create(dpah_context* ctx, const dpah_job* jobs, uint64_t job_count,
                           dpah_batch_pass** out);
int dpah_batch_pass_launch(dpah_batch_pass* pass);
int dpah_batch_pass_results(dpah_batch_pass* pass, dpah_result** results);
int dpah_batch_pass_destroy(dpah_batch_pass* pass);

/* Kernel timing recorded with HIP events on the context's stream (enabled by
 * dpah_options.kernel_timing): launches and summed milliseconds of the named
 * kernel ("advect_persistent", "advect_simple", "compact_write", ...). */
int dpah_kernel_timing(dpah_context* ctx, const char* kernel, uint64_t* launches, double* total_ms);
int dpah_kernel_timing_reset(dpah_context* ctx);

/* Roofline instrumentation (no counterpart in dpa).  While on, endpoint
 * launches of DPAH_INTEGRATOR_DPA_RK4 on steady fields run a kernel instance
 * that also counts the register cell cache's reloads -- one per step that
 * crosses into a new cell (and one per particle's first sample): a 128-byte
 * cell record for launches that read the cell layout, the cell's eight nodes
 * (96 bytes) for launches that read the nodes.  These are the bytes the
 * launch must bring past the L2.  Results are unchanged.  dpah_cell_loads
 * returns the counts accumulated since counting was last turned on. */
int dpah_set_count_cell_loads(dpah_context* ctx, int on);
int dpah_cell_loads(dpah_context* ctx, uint64_t* record_loads, uint64_t* node_loads);

/* Launch shape of the persistent advect kernel and the order of its work
 * queue.  No counterpart in dpa, whose CPU loop has no launch shape;
 * dpah_autotune_endpoints picks these by timing a sample of a workload. */
typedef struct dpah_tuning {
  int block;           /* threads per workgroup: 64, 128, 192 or 256          */
  int blocks_per_cu;   /* persistent grid = CUs x blocks_per_cu, 1..32; 0: by
                          launch kind (4 endpoint seed arrays; lattices of
                          steady fields 3 on C-order nodes, 5 auxiliary and 4
                          of points on node bricks; 4 lattices of time-variant
                          fields; 6 recorded launches)                         */
  int order;           /* 1: endpoint launches of seed arrays and lattices take
                          their queue in brick Morton order, split across the
                          XCDs (auxiliary FTLE lattices are already compact;
                          recorded launches write rows in seed order)          */
  int parts;           /* queue parts, 1..8 (the XCD count by default)         */
  int refill;          /* idle lanes (1..64) that make a wave draw new particles:
                          1 keeps every lane busy, 64 keeps a wave's particles
                          together until all have retired; 0: 8 for endpoint
                          launches of seed arrays, 64 for lattices and for
                          recorded launches                                    */
  int cells;           /* 0: node layout only.  1: fields created afterwards also
                          keep a cell layout (each cell's 8 corners in one 128-B
                          record, 10.7x the node bytes) when it fits a tenth of
                          the memory budget; launches of seed arrays read it,
                          lattice-seeded launches (FTLE) read the nodes.
                          2: every launch reads the cells when present.        */
  int fastdiv;         /* 1: divisions by a grid spacing use its reciprocal with
                          one correction on axes where field creation verified
                          it bit for bit against the correctly rounded quotient
                          over the axis's range; 0: always divide              */
  int cell_brick;      /* record order of the cell layout of fields created
                          afterwards: 0 C order over the cells; 2..5 bricks of
                          2^k cells per axis stored one after another         */
  int order_brick;     /* queue ordering of seed arrays: bricks of at least
                          2^order_brick cells per axis, 1..5; 0: 4 cells      */
  int lattice_brick;   /* whole-lattice launches (FTLE) walk their points in
                          bricks of this many points per axis: 2, 4 or 8;
                          0: lattices of points 4 (2 on node bricks),
                          auxiliary lattices 2 (4 with shared particles)      */
  int slot_stage;      /* recorded launches: 0 by launch kind (a single
                          request's rounds one thread per row; a batch
                          persistent, its slots in 8-slot blocks staged in LDS
                          from 2^15 rows), 1 persistent with one-slot runs,
                          2 persistent with staged blocks, 3 one thread per
                          row with one-slot runs                              */
  int lattice_group;   /* whole-lattice launches walk groups of this many
                          bricks per axis one after another, bricks inside a
                          group: 2, 4, 8 or 16; 1: no groups; 0: lattices of
                          points 2 (4 on node bricks), none for auxiliary
                          ones                                                 */
  int node_brick;      /* 1: steady fields created afterwards also keep their
                          nodes in bricks of 4^3 nodes (the node bytes again)
                          when all slices fit 2^32 bytes and a tenth of the
                          memory budget; launches that sample nodes (lattice
                          seeds, fields without a cell layout) read them.
                          2: time-variant fields too.  0: C-order nodes only */
} dpah_tuning;
int dpah_get_tuning(dpah_context* ctx, dpah_tuning* out);
/* Sets the context's tuning and drops autotuned entries. */
int dpah_set_tuning(dpah_context* ctx, const dpah_tuning* tuning);

/* Times candidate tunings (workgroups of 128 or 256 threads x 12/16/20/24/32
 * waves per CU x refill 4/8/16/64 x queue order off or on with 4- or 8-cell
 * sort bricks x cell/node layout) on a sample of an endpoint workload --
 * seeds: count x 3 device floats, as dpah_advect_endpoints_device takes them --
 * and keeps the fastest for launches with the same key (field shape,
 * integrator, seed array, endpoints).  With dpah_options.autotune_cache_path
 * set, the choice is written there and read back by later contexts on the same
 * kind of device.  No counterpart in dpa (its CPU loop has no launch shape). */
int dpah_autotune_endpoints(dpah_context* ctx, const dpah_field* field, const float* seeds,
                            uint64_t count, uint64_t uniform_remaining, float step_size,
                            float start_time, int integrator, dpah_tuning* best,
                            double* best_ms);

/* Times candidate launch shapes of recorded (or endpoint) batch launches --
 * workgroups per CU 3..8 x refill 16 / 64 x slot layout (one-slot runs or
 * 8-slot blocks staged in LDS) -- on each group of the jobs of
 * dpah_advect_batch (the groups that function forms), each candidate timed on
 * the group's own persistent launch, and keeps the fastest per launch key
 * (field shape, integrator, seed array, record flag).  best / best_ms: the
 * group with the most particles.  No counterpart in dpa. */
int dpah_autotune_batch(dpah_context* ctx, const dpah_job* jobs, uint64_t job_count,
                        dpah_tuning* best, double* best_ms);

/* Times candidate launch shapes of a whole-lattice launch -- workgroups per CU
 * 2..8 x refill 16 / 32 / 64 x queue walk (C order, or bricks of 2, 4 or 8
 * points per axis) -- and keeps the fastest for its launch key.  delta non-NULL:
 * the six auxiliary particles per point of dpah_ftle_aux_device (origin,
 * spacing, shape as there); NULL: the lattice points themselves, as
 * dpah_ftle_lattice seeds them (origin = field offset, spacing = field spacing
 * x stride, shape = points per axis).  No counterpart in dpa. */
int dpah_autotune_lattice(dpah_context* ctx, const dpah_field* field, const float origin[3],
                          const float spacing[3], const uint64_t shape[3], const float* delta,
                          uint64_t iterations, float step_size, float start_time, int integrator,
                          dpah_tuning* best, double* best_ms);

/* The check behind dpah_tuning.fastdiv, exposed for diagnostics: the number of
 * floats x in [-bound, bound] whose reciprocal quotient (x*r corrected once by
 * the exact residual) differs in any bit from the correctly rounded x / h.
 * Replaces the divisions of regular_grid.hpp:45-46, 63-64 on verified axes. */
int dpah_verify_division(dpah_context* ctx, float h, float r, float bound, uint64_t* mismatches);

#ifdef __cplusplus
}
#endif
#endif
