/*
 * stellar.h -- C ABI of libstellar_hip, the MI355X (gfx950) Barnes-Hut step.
 *
 * The library replaces the launch-wrapper layer of pentilm/StellarSolver
 * (Particle_cuda.cuh:5-13, Particle_cuda.cu:8-58): nine void functions over
 * caller-owned raw device pointers, all on the legacy default stream, with no
 * error reporting.  Here a context owns the device buffers and one HIP
 * stream; every entry point returns a status code and never aborts or waits
 * on stdin (the reference's gpuErrchk blocks on getchar(),
 * BarnesHutParticleSystem.cpp:14-22).
 *
 * Threading: calls on one context must be serialised by the caller (as the
 * reference's single host thread does).  Different contexts are independent.
 *
 * Stage order per step (BarnesHutParticleSystem.cpp:308-315):
 *   stellar_bbox -> stellar_build -> stellar_com -> stellar_sort
 *   -> stellar_forces -> stellar_update      (stellar_step runs all six)
 * A stage called out of order returns STELLAR_ERR_STATE.
 *
 * Body arrays crossing the ABI are host pointers in body-id order.
 */
#ifndef STELLAR_H
#define STELLAR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define STELLAR_OK 0
#define STELLAR_ERR_ARG (-1)      /* bad argument or size                    */
#define STELLAR_ERR_OOM (-2)      /* device or host allocation failed        */
#define STELLAR_ERR_POOL (-3)     /* more cells than node_capacity            */
#define STELLAR_ERR_STACK (-4)    /* walk stack overflow                      */
#define STELLAR_ERR_DEPTH (-5)    /* too many bodies share one 32-level cell  */
#define STELLAR_ERR_DEVICE (-6)   /* HIP runtime error                        */
#define STELLAR_ERR_STATE (-7)    /* stage called out of order                */
#define STELLAR_ERR_BUCKET (-8)   /* >4 bodies share a deepest-level cell     */
#define STELLAR_ERR_SIZE (-9)     /* input above an allocation-free maximum    */

/* Deepest cell level: bodies that share all 64 slot codes (coincident or
 * closer than FP32 midpoints can separate) are kept, up to four, in one
 * level-64 cell by rank instead of splitting forever. */
#define STELLAR_MAX_LEVEL 64

/* Force-path precision.  FP32: every term in FP32.  MIXED: an accepted
 * cell's monopole + quadrupole term is evaluated as a dimensionless bracket
 * in packed FP16 (two cells per v_pk_fma_f16, moments Q/(m ws^2) stored in
 * FP16 by the centre-of-mass pass) and added to Kahan-compensated FP32 sums;
 * bodies (the near field), the acceptance vote and the update stay FP32.
 * MIXED needs STELLAR_FLAG_QUADRUPOLE and group 64 (else STELLAR_ERR_ARG).
 * Moments are summed in FP64 in every mode. */
#define STELLAR_PRECISION_FP32 0  /* FP32 walk and update                    */
#define STELLAR_PRECISION_MIXED 1 /* packed-FP16 far field, FP32 near field  */

typedef struct stellar_params {
  /* Opening test: a cell of x-width w is accepted when every body of the
   * walk group has w*w/theta <= dx*dx+dy*dy+eps2 (kernels.cu:376,395,406;
   * the reference hard-codes theta=0.5, kernels.cu:9). */
  float theta;
  /* Softening squared (reference: eps2=0.025, kernels.cu:8). */
  float eps2;
  /* dt and velocity dampening d of v+=dt*a; x+=d*dt*v
   * (SimulationParameters.h:23,25; kernels.cu:522-537). */
  float timestep;
  float dampening;
  /* Kept for the interface; like the reference the step does not use it
   * (kernels.cu:340 receives g and ignores it). */
  float gravity;
  /* Walk-group size: 64 (one wavefront, the native choice) or 32 (the
   * reference's warp; two groups per wavefront). */
  int32_t group;
  int32_t precision; /* STELLAR_PRECISION_* */
  int32_t flags;     /* STELLAR_FLAG_* */
} stellar_params;

/* Record per-stage HIP event timings on every step (small overhead). */
#define STELLAR_FLAG_STAGE_TIMING 1
/* Accepted cells add their quadrupole term (not in the reference), the
 * second-order term of the softened potential sum m / sqrt(|d + s|^2 + eps2)
 * expanded about the cell's centre of mass:
 * a += -Q.d/r^5 + (5/2)(d.Q.d - eps2 tr Q) d/r^7 with r^2 = |d|^2 + eps2,
 * Q = sum m (3 s s^T - |s|^2 I) over the cell's bodies (in-plane), moments
 * accumulated bottom-up in FP64. */
#define STELLAR_FLAG_QUADRUPOLE 2
/* Opening test w/d <= theta on the unsoftened distance d (the textbook test
 * of the reference's older, commented-out kernel, kernels.cu:488-489),
 * evaluated as w^2/theta^2 + eps2 <= d^2 + eps2 in FP32.  The reference's
 * own test puts eps2 on one side only, so every cell narrower than about
 * theta*eps passes, including the cell that holds the body itself. */
#define STELLAR_FLAG_TEXTBOOK_OPENING 4
/* Opening test with a per-cell radius instead of the cell width: a cell is
 * accepted when every body of the walk group has bmax/theta <= d (unsoftened,
 * as bmax^2/theta^2 + eps2 <= d^2 + eps2 in FP32),
 * where bmax bounds the distance from the cell's centre of mass to its
 * bodies (bottom-up in FP64: bmax = max over children of |c_k - C| +
 * bmax_k).  Not in the reference; with it theta is a true angle bound. */
#define STELLAR_FLAG_BMAX_OPENING 16
/* stellar_run replays a captured HIP graph of two whole steps (captured on
 * first use for each arrangement of the context's double buffers; an odd
 * last step runs directly).  Stage times are not recorded in graph runs. */
#define STELLAR_FLAG_GRAPH 8
/* Record HIP events around the forces stage only (two per step instead of
 * eight): stellar_stage_times then reports the forces time and zeros.
 * Ignored when STELLAR_FLAG_STAGE_TIMING is set. */
#define STELLAR_FLAG_FORCES_TIMING 32
/* Treat a body alone in a root slot as the reference does (kernels.cu:
 * 373-387, 397-398): it is pushed like a cell and its child slots are read.
 * Those are empty for every body but body 0 (of each instance), so such a
 * body exerts no force; body 0's slots are the root's own (child[0..3]), so
 * the root's children are visited again with the next level's threshold and
 * the accepted ones counted twice.  Without the flag the lone body is
 * accumulated like any other body.  For the level-threshold walks only (not
 * with STELLAR_FLAG_QUADRUPOLE or STELLAR_FLAG_BMAX_OPENING). */
#define STELLAR_FLAG_REFERENCE_ROOT 64
/* Store bodies and number cells along a Hilbert curve of the same tree (not
 * in the reference): the children at levels 1..16 are visited along their
 * parent's Hilbert curve, deeper children in slot order (an instance's order
 * is the same alone and in an ensemble).  The tree -- its cells, each
 * child's slot, counts, moments -- is the same; the walk groups, which are
 * consecutive bodies of the storage order, become more compact (fewer
 * interactions per body).  Exported trees stay in the reference's numbering
 * and sort order (stellar_export_tree renumbers them). */
#define STELLAR_FLAG_HILBERT_ORDER 128

typedef struct stellar_ctx stellar_ctx;

/* Defaults matching the reference: theta 0.5, eps2 0.025, dt 0.001,
 * dampening 1.0, gravity 1.0 (kernels.cu:8-9, main.cpp:94-97), group 64. */
void stellar_default_params(stellar_params* p);

/* Create a context for n bodies on the current HIP device.  node_capacity
 * is the maximum number of cells; <= 0 selects 2n+12000, the reference's
 * pool size (BarnesHutParticleSystem.cpp:33).  Replaces the 19 cudaMallocs
 * of BarnesHutParticleSystem.cpp:54-83. */
int stellar_create(const stellar_params* p, int32_t n, int64_t node_capacity,
                   stellar_ctx** out);
/* Create a context for an ensemble of ninst independent simulations of
 * counts[i] >= 1 bodies each (no counterpart in the reference, which runs one
 * simulation per process).  Instance i owns body ids and positions
 * [offsets[i], offsets[i+1]) with offsets the running sum of counts; every
 * stage advances all instances in one launch (per-instance boxes, roots and
 * node pools; the sort keeps each instance in its own range).  Body arrays
 * crossing the ABI are the instances' arrays concatenated.  node_capacity
 * <= 0 selects sum(2*counts[i] + 12000).  ninst <= 65536. */
int stellar_batch_create(const stellar_params* p, int32_t ninst,
                         const int32_t* counts, int64_t node_capacity,
                         stellar_ctx** out);
int32_t stellar_num_instances(const stellar_ctx* ctx);
/* ninst + 1 offsets. */
int stellar_instance_offsets(const stellar_ctx* ctx, int32_t* offsets);
int stellar_destroy(stellar_ctx* ctx);
int stellar_set_params(stellar_ctx* ctx, const stellar_params* p);
int stellar_get_params(const stellar_ctx* ctx, stellar_params* p);
int32_t stellar_num_bodies(const stellar_ctx* ctx);

/* Host -> device state in body-id order (BarnesHutParticleSystem.cpp:365-371).
 * Accelerations are zeroed. */
int stellar_upload(stellar_ctx* ctx, const float* mass, const float* x,
                   const float* y, const float* vx, const float* vy);
/* Device -> host, body-id order.  Any pointer may be NULL. */
int stellar_download(stellar_ctx* ctx, float* mass, float* x, float* y,
                     float* vx, float* vy, float* ax, float* ay);
/* Interleaved (x0,y0,x1,y1,...) body positions, body-id order: the first 2n
 * floats of the reference's output buffer (kernels.cu:541-553). */
int stellar_download_positions(stellar_ctx* ctx, float* out_xy);

/* Headless rendering of the current positions: the reference draws every
 * body as a one-pixel point of colour (1, 1, 1) and alpha 0.1, blended
 * additively over black (glBlendFunc(GL_SRC_ALPHA, GL_ONE)) through
 * glm::ortho(-25, 25, -25, 25) into a 1000 x 1000 window
 * (StellarSolverVisualizer.cpp:13-19, 208-253).  view = {left, right,
 * bottom, top}; the body at (x, y) falls in column floor((x - left) * sx)
 * and row height - 1 - floor((y - bottom) * sy), sx = width / (right - left),
 * sy = height / (top - bottom), all FP32 as written (row 0 at the top);
 * bodies outside the image are clipped.  counts (host, width x height, may be
 * NULL): bodies per pixel; grey (host, width x height bytes, may be NULL):
 * min(255, round(25.5 count)), the blend's white at alpha 0.1 per body,
 * clamped and quantised once.  Synchronous; reads the state, changes nothing.
 * STELLAR_ERR_ARG for an empty view or more than 2^28 pixels. */
int stellar_render(stellar_ctx* ctx, const float* view, int32_t width, int32_t height,
                   uint32_t* counts, uint8_t* grey);

/* ---- the six stages (kernels.cu:62-537) -------------------------------- */
int stellar_bbox(stellar_ctx* ctx);    /* compute_bounding_box_kernel      */
int stellar_build(stellar_ctx* ctx);   /* build_tree_kernel                */
int stellar_com(stellar_ctx* ctx);     /* centre_of_mass_kernel            */
int stellar_sort(stellar_ctx* ctx);    /* sort_kernel                      */
int stellar_forces(stellar_ctx* ctx);  /* compute_forces_kernel            */
int stellar_update(stellar_ctx* ctx);  /* update_kernel (params dt, d)     */

/* One whole step: the six stages plus the output buffer
 * (BarnesHutParticleSystem::update, BarnesHutParticleSystem.cpp:301-347).
 * Synchronous, like the reference (it waits on its stop event, :318). */
int stellar_step(stellar_ctx* ctx);
/* `steps` steps back to back with a single synchronisation at the end. */
int stellar_run(stellar_ctx* ctx, int32_t steps);
/* Wait for queued work and report any sticky device error. */
int stellar_sync(stellar_ctx* ctx);

/* ---- allocation-free bounding box --------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the bounding box of stellar_bbox -- {left, right,
 * bottom, top} of n bodies with the origin included (kernels.cu:62-114) --
 * over caller-owned device arrays, with no context.  x, y: device pointers
 * to n floats; bbox4: device pointer to 4 floats (the result); stream: the
 * hipStream_t to launch on (NULL: the null stream).  Asynchronous: bbox4
 * holds the result once the stream's work before the next synchronisation
 * completes.  The per-tile partial results live in a fixed device array of
 * the library (STELLAR_BBOX_NOALLOC_MAX_TILES x 16 bytes on each GPU it runs
 * on, held until the process ends); nothing is allocated or freed, and
 * n above STELLAR_BBOX_NOALLOC_MAX_BODIES returns STELLAR_ERR_SIZE.  All
 * calls on one GPU share that array: calls that overlap in time on one GPU
 * (two streams, two threads) race and must be serialised by the caller. */
#define STELLAR_BBOX_NOALLOC_MAX_TILES 4096
#define STELLAR_BBOX_NOALLOC_MAX_BODIES (STELLAR_BBOX_NOALLOC_MAX_TILES * 4096)
int stellar_bbox_noalloc(const float* x, const float* y, int32_t n, float* bbox4,
                         void* stream);

/* ---- inspection --------------------------------------------------------- */
/* Bounding box {left, right, bottom, top} of the last stellar_bbox
 * (instance 0 / instance inst). */
int stellar_get_bbox(stellar_ctx* ctx, float* bbox4);
int stellar_get_bbox_instance(stellar_ctx* ctx, int32_t inst, float* bbox4);
/* Number of cells of the last stellar_build (all instances). */
int stellar_num_cells(stellar_ctx* ctx, int64_t* ncells);

/* Canonical tree of the last build: cells renumbered in DFS pre-order over
 * slots 0..3 (the order sort_kernel, kernels.cu:292-336, lays bodies out).
 * Child codes: -1 empty, 0..n-1 body id, n+c cell c.  root4 = the four root
 * slots (the reference's child[0..3]).  Arrays sized by capacity (>= ncells;
 * sorted: n).  com_* need stellar_com; any pointer may be NULL. */
int stellar_export_tree(stellar_ctx* ctx, int32_t* root4, int32_t* child,
                        int32_t* count, int32_t* start, int32_t* level,
                        int32_t* sorted, float* com_x, float* com_y,
                        float* com_m, int64_t capacity, int64_t* ncells);
/* The same for instance inst of an ensemble, with body ids and positions
 * local to the instance (0 .. counts[inst]-1). */
int stellar_export_tree_instance(stellar_ctx* ctx, int32_t inst, int32_t* root4,
                                 int32_t* child, int32_t* count, int32_t* start,
                                 int32_t* level, int32_t* sorted, float* com_x,
                                 float* com_y, float* com_m, int64_t capacity,
                                 int64_t* ncells);

/* Per-body interaction counts and accelerations of the last stellar_forces,
 * body-id order.  Either pointer may be NULL. */
int stellar_export_forces(stellar_ctx* ctx, float* ax, float* ay,
                          int32_t* interactions);

/* Sum of interactions of the last walk (exact, 64-bit), and how many of
 * them added a cell's quadrupole term. */
int stellar_last_interactions(stellar_ctx* ctx, int64_t* total);
int stellar_last_quad_interactions(stellar_ctx* ctx, int64_t* total);

/* Per-stage device milliseconds of the last step (mean over the steps of
 * the last stellar_run) when STELLAR_FLAG_STAGE_TIMING is set: [bbox,
 * build, com, sort, forces, update, output]; with STELLAR_FLAG_FORCES_TIMING
 * only the forces entry.  A stage that launched no kernel reports 0: sort
 * (the build already wrote the bodies in tree order; the stage swaps
 * buffers), output (filled when read) and update when the walk applied it
 * (group-64 walks that pop one node at a time). */
int stellar_stage_times(stellar_ctx* ctx, float* ms7);

/* ---- launch-shape autotuner (no counterpart in the reference) ---------- */
typedef struct stellar_tuning {
  int32_t group;         /* walk-group size: 64, 32 or 16                   */
  int32_t nodes_per_pop; /* group 64: tree nodes fetched per round trip      */
  int32_t block_threads; /* group 64: threads per walk workgroup             */
  float walk_ms;         /* mean walk time of the choice when measured       */
  int32_t schedule;      /* group 64: STELLAR_WALK_STATIC or _QUEUE          */
} stellar_tuning;
/* Time the tree walk for every candidate launch shape and schedule (the
 * static launch and the persistent work queue, see stellar_set_walk_schedule)
 * on the current state
 * (reps timed walks each, after one untimed) and keep the fastest for this
 * context.  group_mask selects the walk-group sizes tried: 1 = 64, 2 = 32,
 * 4 = 16; 0 means 64 only, which keeps the set of interactions unchanged
 * (the shapes then differ only in speed and in summation order).  The
 * choice is remembered per (log2 bodies per instance, ensemble size bucket,
 * flags, group_mask) for the rest of the process, and a context in a known
 * bucket takes it without measuring.  Leaves the state sorted (stage
 * stellar_sort done).  out may be NULL. */
int stellar_autotune(stellar_ctx* ctx, int32_t group_mask, int32_t reps,
                     stellar_tuning* out);
/* Testing hook, no counterpart in the reference: caps the keys a bucket of
 * the key sort sorts in one block pass (1 .. 2048, default 2048), so that tests
 * can send every bucket through the multi-pass path that larger buckets take;
 * cap = -1 instead sorts the keys with the device-wide radix sort that
 * contexts of more than 2^24 bodies use (and cap = 1 .. 2048 returns to the
 * sample sort). */
int stellar_debug_sort_cap(stellar_ctx* ctx, int32_t cap);

/* Set the group-64 walk launch shape by hand: nodes_per_pop 1, 2 or 4
 * (4 is capped to 2 when a node's second block is read), block_threads 64,
 * 128, 256, 512 or 1024. */
int stellar_set_walk_shape(stellar_ctx* ctx, int32_t nodes_per_pop,
                           int32_t block_threads);

/* Scheduling of the group-64 walk (no counterpart in the reference, whose
 * force kernel is one grid-stride launch, kernels.cu:352-435 with the fixed
 * shape of Particle_cuda.cu:5-6):
 *   STELLAR_WALK_STATIC      one workgroup per WAVES groups, each XCD one
 *                            contiguous range of workgroups (default);
 *   STELLAR_WALK_QUEUE       persistent: the resident workgroups take units
 *                            of one group per wavefront from per-XCD work
 *                            queues, then from the other XCDs' queues;
 *   STELLAR_WALK_QUEUE_WAVE  the same with every wavefront claiming its own
 *                            groups.
 * The set of interactions is the same in every schedule.  Queues apply to
 * workgroups of 256 or 512 threads; other shapes run the static launch. */
#define STELLAR_WALK_STATIC 0
#define STELLAR_WALK_QUEUE 1
#define STELLAR_WALK_QUEUE_WAVE 2
int stellar_set_walk_schedule(stellar_ctx* ctx, int32_t schedule);

/* Message for the last non-OK status on this context. */
const char* stellar_last_error(const stellar_ctx* ctx);
const char* stellar_status_string(int status);

/* ---- FP64 diagnostics on the device ------------------------------------ */
/* Direct-sum accelerations in FP64 for the listed body ids (all if targets
 * is NULL, then ntargets must be n), softening eps2 as given; sources are
 * the bodies of the target's instance. */
int stellar_direct_fp64(stellar_ctx* ctx, int32_t ntargets,
                        const int32_t* targets, double eps2, double* ax,
                        double* ay);
/* Kinetic and Plummer-softened potential energy, linear momentum (summed
 * over the instances of an ensemble, each with its own bodies only). */
int stellar_energy_fp64(stellar_ctx* ctx, double eps2, double* kinetic,
                        double* potential, double* px, double* py);
/* The same per instance: arrays of ninst values, any pointer may be NULL. */
int stellar_energy_fp64_instances(stellar_ctx* ctx, double eps2, double* kinetic,
                                  double* potential, double* px, double* py);

/* Conserved quantities of one instance, FP64.  No counterpart in the
 * reference (it reports none); its update kernels.cu:522-537 fixes what
 * "velocity" means here. */
typedef struct stellar_diagnostics {
  double kinetic;    /* sum m |v|^2 / 2                                    */
  double potential;  /* -sum_{i<j} m_i m_j / sqrt(r_ij^2 + eps2)            */
  double px, py;     /* sum m v                                             */
  double lz;         /* sum m (x vy - y vx), about the origin               */
} stellar_diagnostics;
/* out: ninst records.  half_dt != 0 synchronises the velocities with the
 * positions first, v + half_dt * a, using the accelerations of the last
 * stellar_forces call: the kick-drift update stores v half a step behind x,
 * so half_dt = dt/2 with forces evaluated at the stored positions gives the
 * time-centred energy.  Then stellar_forces must be the last stage run
 * (STELLAR_ERR_STATE otherwise). */
int stellar_diagnostics_fp64(stellar_ctx* ctx, double eps2, double half_dt,
                             stellar_diagnostics* out);

/* ---- initial conditions (host) ------------------------------------------ */
#define STELLAR_IC_DISK 0          /* central mass + uniform-radius disk     */
#define STELLAR_IC_PLUMMER 1       /* 2D-projected Plummer sphere            */
#define STELLAR_IC_COLLIDING 2     /* two disks on a collision course        */
#define STELLAR_IC_REF_DISK 10     /* BarnesHutParticleSystem.cpp:439-483    */
#define STELLAR_IC_REF_PLUMMER 11  /* BarnesHutParticleSystem.cpp:390-428    */
#define STELLAR_IC_REF_COLLIDING 12 /* BarnesHutParticleSystem.cpp:494-560   */

typedef struct stellar_ic_params {
  int32_t model;       /* STELLAR_IC_* */
  int32_t pad;
  uint64_t seed;
  double gravity;      /* G of the velocity set-up                           */
  double central_mass; /* disk: mass of the central body                     */
  double rmin, rmax;   /* disk: radii ~ U[rmin, rmax]                        */
  double jitter;       /* disk: N(0, jitter) added to each velocity component */
  double plummer_a;    /* plummer: scale radius a                            */
  double plummer_rcut; /* plummer: truncation radius in units of a           */
  double offset_x;     /* colliding: disks centred at (+-offset_x, 0)        */
  double bulk_vx, bulk_vy; /* colliding: the +x disk moves with (bulk_vx,
                              bulk_vy), the other with the negative          */
} stellar_ic_params;

/* Fill the parameters of a named workload: 0 disk N=4096 (M=1000,
 * r~U[2,100]); 1 Plummer a=1; 2 colliding disks (M=1e4, r~U[2,200],
 * +-300, -+5/+-1); 4 deep Plummer a=0.01.  Seeds 0,1,2,4. */
int stellar_ic_defaults(int32_t workload, stellar_ic_params* p);
/* Generate n bodies (host arrays, body-id order). */
int stellar_generate_ic(const stellar_ic_params* p, int32_t n, float* mass,
                        float* x, float* y, float* vx, float* vy);

#ifdef __cplusplus
}
#endif
#endif
