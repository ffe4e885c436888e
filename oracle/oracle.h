/*
 * oracle.h -- CPU restatement of the reference Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libstellar_hip) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Reference files
 * are pentilm/StellarSolver: kernels.cu, BarnesHutParticleSystem.cpp,
 * debug.cpp.
 *
 * Tree formats
 * ------------
 * "reference" arrays follow BarnesHutParticleSystem.cpp:33,47-50: a node pool
 * of m entries, bodies at [0,n), cells at [n, n+ncells), child[4*m] with the
 * four root slots aliasing child[0..3] (kernels.cu:142,159,300-301).
 *
 * "canonical" arrays renumber the cells in DFS pre-order over slots 0..3
 * (cell c = 0..C-1).  Child codes: -1 empty, 0..n-1 body id, n+c cell c.
 * root[4] holds the four root slots in the same code.  start[c]/count[c] are
 * the offset of the cell's first body in sorted[] and its body count.  This
 * form is unique for a given set of FP32 positions and bounding box.
 */
#ifndef STELLAR_ORACLE_H
#define STELLAR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_OK 0
#define OR_ERR_ARG (-1)
#define OR_ERR_POOL (-3)   /* more cells than the pool holds */
#define OR_ERR_DEPTH (-5)  /* a split would go below the deepest level */
#define OR_ERR_BUCKET (-8) /* more than 4 bodies share a deepest-level cell */

/* Deepest cell level.  Cells live at levels 1..OR_MAX_LEVEL; a cell at
 * OR_MAX_LEVEL keeps its bodies in slots 0..k-1 by (key, id) rank. */
#define OR_MAX_LEVEL 64

/* bbox[4] = {left, right, bottom, top} */
int or_bbox(int32_t n, const float* x, const float* y, float* bbox);

/* Per-body path keys: k1 = slot codes of levels 1..32, k2 = levels 33..64,
 * two bits per level, most significant level first. */
void or_path_keys(int32_t n, const float* x, const float* y, const float* bbox,
                  uint64_t* k1, uint64_t* k2);

/* Sequential restatement of build_tree_kernel.  Arrays sized for a pool of
 * m nodes; cell sums (x,y,mass as sum m*x, sum m*y, sum m) in FP32. */
int or_build_insert(int32_t n, int32_t m, const float* x, const float* y,
                    const float* mass, const float* bbox, int32_t* child,
                    int32_t* count, int32_t* start, float* cx, float* cy,
                    float* cm, int32_t* ncells);

/* centre_of_mass_kernel: cx/=cm, cy/=cm over cells [n, n+ncells). */
void or_com_ref(int32_t n, int32_t ncells, float* cx, float* cy,
                const float* cm);

/* sort_kernel, sequential: fills start[] of cells and sorted[0..n). */
void or_sort_ref(int32_t n, int32_t ncells, const int32_t* child,
                 const int32_t* count, int32_t* start, int32_t* sorted);

/* Renumber reference arrays into canonical form.  Outputs sized ncells. */
int or_canonicalize(int32_t n, int32_t ncells, const int32_t* child,
                    const int32_t* count, const int32_t* start,
                    const float* cx, const float* cy, const float* cm,
                    int32_t* root, int32_t* cchild, int32_t* ccount,
                    int32_t* cstart, int32_t* clevel, float* ccx, float* ccy,
                    float* ccm);

/* Path-key construction of the same tree, directly in canonical form.
 * Capacity = max cells; returns the cell count in *ncells. */
int or_build_keys(int32_t n, const float* x, const float* y, const float* bbox,
                  int64_t capacity, int32_t* sorted, int32_t* root,
                  int32_t* child, int32_t* count, int32_t* start,
                  int32_t* level, int64_t* ncells);

/* Bottom-up FP64 centre of mass in slot order over a canonical tree. */
void or_com_exact(int32_t n, int64_t ncells, const int32_t* child,
                  const float* x, const float* y, const float* mass,
                  float* ccx, float* ccy, float* ccm);

/* Traceless quadrupole moments of every cell about its centre of mass
 * (Qxx, Qxy, Qyy of sum m (3 d d^T - |d|^2 I) for points in the plane),
 * accumulated bottom-up in FP64 in slot order with the parallel-axis shift
 * from each child's FP64 centre of mass; written as FP32, 3 per cell. */
void or_quad_exact(int32_t n, int64_t ncells, const int32_t* child,
                   const float* x, const float* y, const float* mass,
                   float* cq);

/* compute_forces_kernel restated: group-of-G vote walk over a canonical tree.
 * Positions/mass are by body id; ax/ay/inter are written by body id. */
typedef struct or_walk_params {
  float theta;
  float eps2;
  int32_t group;     /* walk-group size (32 = reference warp) */
  int32_t threads;   /* OpenMP threads, <=0 = all */
  int32_t opening;   /* 0: w^2/theta <= r (reference);
                        1: w^2/theta^2 + eps2 <= r (textbook, unsoftened d) */
  int32_t mixed;     /* 1: accepted cells' terms in FP16 (see or_quad_mixed),
                        needs cqh and clevel */
  const float* cq;   /* NULL: monopole; else 3 quadrupole terms per cell
                        (Qxx, Qxy, Qyy), applied with the softened form
                        -Q.d/r^5 + (5/2)(d.Q.d - eps2 tr Q) d/r^7 */
  const float* cthr; /* NULL: per-level thresholds; else one per cell
                        (or_bmax_exact), compared with r like dp */
  const float* cqh;  /* mixed: Qh per cell (3 FP16 values, or_quad_mixed) */
  const int32_t* clevel; /* mixed: level of each cell */
  int32_t ref_root;  /* 1: a body alone in a root slot is pushed like a cell
                        (kernels.cu:373-387): body 0's child slots are the
                        root's (the root's children are visited again one
                        level down), any other body's are empty */
  int32_t pad2;
  const int32_t* order; /* NULL: walk groups are G consecutive bodies of
                           sorted[]; else of this body order (a permutation of
                           0..n-1, e.g. or_hilbert_order's) */
} or_walk_params;

int or_walk(int32_t n, const float* x, const float* y, const float* mass,
            const float* bbox, const int32_t* sorted, const int32_t* root,
            const int32_t* child, int64_t ncells, const float* ccx,
            const float* ccy, const float* ccm, const or_walk_params* p,
            float* ax, float* ay, int32_t* inter);

/* Hilbert order of the bodies of a canonical tree (not in the reference):
 * the depth-first order that visits the children of a node at level L < l1
 * (the root is level 0) along a Hilbert curve -- base order SW, NW, NE, SE;
 * the first child's frame transposed, the last child's transposed and
 * turned by 180 degrees -- and the children of deeper nodes in slot order.
 * Slot code c = west | south << 1.  order[k] = the k-th body id. */
int or_hilbert_order(int32_t n, const int32_t* root, const int32_t* child, int32_t l1,
                     int32_t* order);

/* Mixed precision (not in the reference).  An accepted cell's term
 * m d/r^3 - Q.d/r^5 + (5/2)(d.Q.d - eps2 tr Q) d/r^7 is k B with k = m rs^2
 * (FP32) and
 *   B = e + t (-Qh.e + (5/2)(e.Qh.e - tr(Qh) eu) e),  e = d rs, t = rs^2 ws^2,
 * eu = eps2 rs^2 (FP32 product, converted toward zero like e and t),
 * tr(Qh) = Qxx + Qyy in FP16,
 * rs = 1/sqrt(|d|^2 + eps2), ws^2 = ldexpf(S*S, -2 L) for a cell at level L
 * (S the larger extent of the box, FP32), Qh = Q / (m ws^2) rounded to FP32
 * and then to FP16 (nearest even, clamped to +-65504).  B is evaluated in
 * FP16 arithmetic (each operation rounded to nearest even; e, t converted
 * toward zero), two accepted cells at a time in visiting order (the last
 * with zeros), and fmaf(k_a, B_a, k_b * B_b) enters an FP32 Kahan sum per
 * body; bodies stay FP32 monopoles.  or_quad_mixed writes Qh (3 per cell, as
 * floats holding FP16 values). */
/* x rounded to FP16 (to_zero = 0: nearest even; 1: toward zero). */
double or_f16_round(double x, int32_t to_zero);
void or_quad_mixed(int32_t n, int64_t ncells, const int32_t* child,
                   const int32_t* clevel, const float* x, const float* y,
                   const float* mass, const float* bbox, float* qh);

/* Opening thresholds from the bottom-up bound bmax (the largest distance
 * from a cell's FP64 centre of mass to any of its bodies, bounded through
 * the children: bmax = max_k |c_k - C| + bmax_k, bodies 0):
 * cthr[c] = (float)(bmax^2 / theta^2 + eps2), |c_k - C| = sqrt(fma(dx,dx,dy*dy)).
 * A cell is accepted when every body of the group has cthr <= r, i.e.
 * bmax/theta <= d on the unsoftened distance. */
void or_bmax_exact(int32_t n, int64_t ncells, const int32_t* child,
                   const float* x, const float* y, const float* mass,
                   float theta, float eps2, float* cthr);

/* Recursive (stack-free) statement of the same acceptance rule; FP64
 * accumulation.  Used to pin or_walk's traversal. */
int or_walk_recursive(int32_t n, const float* x, const float* y,
                      const float* mass, const float* bbox,
                      const int32_t* sorted, const int32_t* root,
                      const int32_t* child, int64_t ncells, const float* ccx,
                      const float* ccy, const float* ccm,
                      const or_walk_params* p, double* ax, double* ay,
                      int32_t* inter);

/* FP64 O(N^2) direct sum for the listed targets (all if targets==NULL). */
void or_direct_fp64(int32_t n, const float* x, const float* y,
                    const float* mass, double eps2, int32_t ntargets,
                    const int32_t* targets, double* ax, double* ay,
                    int32_t threads);

/* update_kernel: v += dt*a; x += (d*dt)*v (FMA-contracted). */
void or_update(int32_t n, float* x, float* y, float* vx, float* vy,
               const float* ax, const float* ay, float dt, float d);

/* debug.cpp invariants restated; bit i set = check i failed.
 * bit0 COUNT1, bit1 CHILD1, bit2 CHILD2, bit3 CHILD3, bit4 CHILD4,
 * bit5 SORTED1, bit6 SORTED2, bit7 BBOX1. */
int32_t or_invariants(int32_t n, int32_t m, const int32_t* count,
                      const int32_t* sorted, const int32_t* child,
                      const float* bbox);

/* FP64 diagnostics: kinetic, potential (Plummer-softened), momentum. */
void or_energy(int32_t n, const float* x, const float* y, const float* vx,
               const float* vy, const float* mass, double eps2,
               int32_t threads, double* kinetic, double* potential,
               double* px, double* py);
/* out[5] = kinetic, potential, px, py, Lz (about the origin); with ax, ay
 * non-NULL the velocities are v + half_dt * a. */
void or_diagnostics(int32_t n, const float* x, const float* y, const float* vx,
                    const float* vy, const float* mass, const float* ax,
                    const float* ay, double half_dt, double eps2,
                    int32_t threads, double* out);

/* One whole reference step on the CPU (bbox, insertion build, FP32 CoM,
 * sort, group-of-G walk, update), state by body id.  Work arrays are
 * allocated inside.  Returns OR_OK or an error. */
typedef struct or_step_params {
  float theta, eps2, dt, dampening;
  int32_t group;
  int32_t threads;
  int32_t m; /* pool size; <=0 means 2n+12000 */
} or_step_params;

int or_step(int32_t n, float* x, float* y, float* vx, float* vy,
            const float* mass, float* ax, float* ay, const or_step_params* p,
            double* stage_seconds /* [6] or NULL */);

int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
