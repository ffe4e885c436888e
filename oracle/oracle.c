/*
 * oracle.c -- CPU restatement of the reference Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libstellar_hip, loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Never part of the
 * product path.
 *
 * Built with -ffp-contract=off so that every FP32 expression is evaluated as
 * written (explicit fmaf where the reference's nvcc build contracts).
 *
 * Parity pins (see tests/test_oracle.py):
 *   or_bbox            closed form min/max with the origin; reference debug.cpp
 *                      BBOX check via oracle/_ref.
 *   or_build_insert    hand-computed known answers; agreement with the
 *                      independent path-key construction (or_build_keys);
 *                      the reference's own debug.cpp checks via oracle/_ref.
 *   or_com_ref/_exact  closed form sum(m x)/sum(m) over each cell's body range.
 *   or_sort_ref        permutation + equality with key argsort.
 *   or_walk            theta->0 equals the direct sum; 2-body closed form;
 *                      symmetric ring; recursive restatement with exact
 *                      interaction counts.
 *   or_direct_fp64     2-body and ring closed forms.
 *   or_update          closed form.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static int nthreads(int32_t t) { return t > 0 ? t : or_num_threads(); }

/* ------------------------------------------------------------------------ */
/* Box subdivision: kernels.cu:144-157 (root), :165-178 (descent),
 * :206-211 (old body in a split), :229-242 (new body in a split).
 * mid = 0.5*(l+r): the FP32 sum l+r promoted to double by the double literal
 * 0.5; the comparison x < mid happens in double; the new bound is the double
 * mid rounded back to FP32.  Slot code: +1 if x < mid (west), +2 if y < mid
 * (south); ties go to the east/north side. */
static inline int subdivide(float px, float py, float* l, float* r, float* b,
                            float* t) {
  double mx = 0.5 * (double)(*l + *r);
  double my = 0.5 * (double)(*t + *b);
  int c = 0;
  if ((double)px < mx) {
    c += 1;
    *r = (float)mx;
  } else {
    *l = (float)mx;
  }
  if ((double)py < my) {
    c += 2;
    *t = (float)my;
  } else {
    *b = (float)my;
  }
  return c;
}

static inline int slot_of(float px, float py, float l, float r, float b,
                          float t) {
  double mx = 0.5 * (double)(l + r);
  double my = 0.5 * (double)(t + b);
  return ((double)px < mx ? 1 : 0) + ((double)py < my ? 2 : 0);
}

/* ------------------------------------------------------------------------ */
/* compute_bounding_box_kernel (kernels.cu:62-114) after reset_arrays_kernel
 * seeds left=right=bottom=top=0 (kernels.cu:51-58): the box always holds the
 * origin.  A negative zero never lowers a bound below +0 (x < 0 is false). */
int or_bbox(int32_t n, const float* x, const float* y, float* bbox) {
  float l = 0.f, r = 0.f, b = 0.f, t = 0.f;
  for (int32_t i = 0; i < n; ++i) {
    if (x[i] < l) l = x[i];
    if (x[i] > r) r = x[i];
    if (y[i] < b) b = y[i];
    if (y[i] > t) t = y[i];
  }
  bbox[0] = l;
  bbox[1] = r;
  bbox[2] = b;
  bbox[3] = t;
  return OR_OK;
}

/* Path key of one body: the slot codes along its own descent, which depend
 * only on the body and the box (kernels.cu:144-178). */
void or_path_keys(int32_t n, const float* x, const float* y, const float* bbox,
                  uint64_t* k1, uint64_t* k2) {
  for (int32_t i = 0; i < n; ++i) {
    float l = bbox[0], r = bbox[1], b = bbox[2], t = bbox[3];
    uint64_t a = 0, c = 0;
    for (int lev = 1; lev <= 32; ++lev)
      a = (a << 2) | (uint64_t)subdivide(x[i], y[i], &l, &r, &b, &t);
    for (int lev = 33; lev <= 64; ++lev)
      c = (c << 2) | (uint64_t)subdivide(x[i], y[i], &l, &r, &b, &t);
    k1[i] = a;
    k2[i] = c;
  }
}

/* ------------------------------------------------------------------------ */
/* build_tree_kernel (kernels.cu:117-271) with bodies inserted one at a time
 * in id order, which is one legal interleaving of the concurrent kernel.
 * Cell sums follow the kernel: every cell passed on the descent gets
 * m*x, m*y, m, 1 added (:180-183); a split chain adds the old body then the
 * new one to each new cell (:221-224, :247-250).  Pool: cells take ids
 * n, n+1, ... (:199) and must stay below m.  The reference splits forever
 * when two bodies share every slot code; here a cell at OR_MAX_LEVEL is an
 * error instead (the path-key build stores such bodies as a bucket). */
int or_build_insert(int32_t n, int32_t m, const float* x, const float* y,
                    const float* mass, const float* bbox, int32_t* child,
                    int32_t* count, int32_t* start, float* cx, float* cy,
                    float* cm, int32_t* ncells) {
  if (n < 0 || m < n) return OR_ERR_ARG;
  /* reset_arrays_kernel (kernels.cu:25-59) */
  for (int64_t i = 0; i < 4 * (int64_t)m; ++i) child[i] = -1;
  for (int32_t i = 0; i < m; ++i) {
    count[i] = i < n ? 1 : 0;
    start[i] = -1;
    if (i >= n) cx[i] = cy[i] = cm[i] = 0.f;
  }
  int32_t index = n;
  for (int32_t i = 0; i < n; ++i) {
    const float px = x[i], py = y[i], pm = mass[i];
    float l = bbox[0], r = bbox[1], b = bbox[2], t = bbox[3];
    int32_t temp = 0;
    int lev = 1; /* level of the box (l,r,b,t) */
    int32_t path = subdivide(px, py, &l, &r, &b, &t);
    int32_t ci = child[4 * temp + path];
    while (ci >= n) {
      temp = ci;
      path = subdivide(px, py, &l, &r, &b, &t);
      ++lev;
      cx[temp] += pm * px;
      cy[temp] += pm * py;
      cm[temp] += pm;
      count[temp] += 1;
      ci = child[4 * temp + path];
    }
    if (ci == -1) {
      child[4 * temp + path] = i;
      continue;
    }
    /* split: ci is a body sharing this slot */
    const int32_t locked = 4 * temp + path;
    int32_t patch = INT32_MAX;
    while (ci >= 0 && ci < n) {
      if (lev >= OR_MAX_LEVEL) return OR_ERR_DEPTH;
      if (index >= m) return OR_ERR_POOL;
      const int32_t cell = index++;
      if (cell < patch) patch = cell;
      if (patch != cell) child[4 * temp + path] = cell;
      const int oc = slot_of(x[ci], y[ci], l, r, b, t);
      cx[cell] += mass[ci] * x[ci];
      cy[cell] += mass[ci] * y[ci];
      cm[cell] += mass[ci];
      count[cell] += count[ci];
      child[4 * cell + oc] = ci;
      start[cell] = -1;
      temp = cell;
      path = subdivide(px, py, &l, &r, &b, &t);
      ++lev;
      cx[cell] += pm * px;
      cy[cell] += pm * py;
      cm[cell] += pm;
      count[cell] += count[i];
      ci = child[4 * temp + path];
    }
    child[4 * temp + path] = i;
    child[locked] = patch;
  }
  *ncells = index - n;
  return OR_OK;
}

/* centre_of_mass_kernel (kernels.cu:275-288) */
void or_com_ref(int32_t n, int32_t ncells, float* cx, float* cy,
                const float* cm) {
  for (int32_t c = n; c < n + ncells; ++c) {
    cx[c] /= cm[c];
    cy[c] /= cm[c];
  }
}

/* sort_kernel (kernels.cu:292-336), run in cell-id order: a parent cell
 * always has a smaller id than its children in a sequential insertion, so
 * start[parent] is known when a child is reached (the kernel spin-waits for
 * it instead). */
void or_sort_ref(int32_t n, int32_t ncells, const int32_t* child,
                 const int32_t* count, int32_t* start, int32_t* sorted) {
  int32_t s = 0;
  for (int k = 0; k < 4; ++k) {
    const int32_t node = child[k];
    if (node >= n) {
      start[node] = s;
      s += count[node];
    } else if (node >= 0) {
      sorted[s++] = node;
    }
  }
  for (int32_t cell = n; cell < n + ncells; ++cell) {
    s = start[cell];
    for (int k = 0; k < 4; ++k) {
      const int32_t node = child[4 * cell + k];
      if (node >= n) {
        start[node] = s;
        s += count[node];
      } else if (node >= 0) {
        sorted[s++] = node;
      }
    }
  }
}

/* DFS pre-order renumbering over slots 0..3, root slots first. */
int or_canonicalize(int32_t n, int32_t ncells, const int32_t* child,
                    const int32_t* count, const int32_t* start,
                    const float* cx, const float* cy, const float* cm,
                    int32_t* root, int32_t* cchild, int32_t* ccount,
                    int32_t* cstart, int32_t* clevel, float* ccx, float* ccy,
                    float* ccm) {
  int32_t* newid = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncells + 1));
  int32_t* stk = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncells + 8));
  int32_t* stl = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncells + 8));
  if (!newid || !stk || !stl) {
    free(newid);
    free(stk);
    free(stl);
    return OR_ERR_ARG;
  }
  int32_t sp = 0, next = 0;
  for (int k = 3; k >= 0; --k)
    if (child[k] >= n) {
      stk[sp] = child[k];
      stl[sp++] = 1;
    }
  while (sp > 0) {
    const int32_t node = stk[--sp];
    const int32_t lev = stl[sp];
    const int32_t c = next++;
    newid[node - n] = c;
    clevel[c] = lev;
    ccount[c] = count[node];
    cstart[c] = start[node];
    if (ccx) {
      ccx[c] = cx[node];
      ccy[c] = cy[node];
      ccm[c] = cm[node];
    }
    for (int k = 3; k >= 0; --k) {
      const int32_t ch = child[4 * node + k];
      if (ch >= n) {
        stk[sp] = ch;
        stl[sp++] = lev + 1;
      }
    }
  }
  int rc = next == ncells ? OR_OK : OR_ERR_ARG;
  for (int k = 0; k < 4; ++k) {
    const int32_t ch = child[k];
    root[k] = ch >= n ? n + newid[ch - n] : ch;
  }
  for (int32_t old = 0; old < ncells && rc == OR_OK; ++old) {
    const int32_t c = newid[old];
    for (int k = 0; k < 4; ++k) {
      const int32_t ch = child[4 * (n + old) + k];
      cchild[4 * c + k] = ch >= n ? n + newid[ch - n] : ch;
    }
  }
  free(newid);
  free(stk);
  free(stl);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Path-key construction: a cell at level L exists exactly when two or more
 * bodies share their first L slot codes (the PR-quadtree that the insertion
 * kernel grows).  Bodies are ordered by (k1, k2, id); every cell is the
 * contiguous run of bodies sharing its prefix, split into four sub-runs by
 * the next code.  Cells are numbered in DFS pre-order as they are created. */
typedef struct keyctx {
  int32_t n;
  const uint64_t* k1;
  const uint64_t* k2;
  const int32_t* sorted;
  int32_t* child;
  int32_t* count;
  int32_t* start;
  int32_t* level;
  int64_t capacity;
  int64_t next;
  int err;
} keyctx;

static inline int digit_at(const keyctx* K, int32_t s, int lev) {
  const int32_t b = K->sorted[s];
  if (lev <= 32) return (int)((K->k1[b] >> (2 * (32 - lev))) & 3u);
  return (int)((K->k2[b] >> (2 * (64 - lev))) & 3u);
}

/* Node code for the run [a,e] of bodies sharing their first lev codes. */
static int32_t key_assign(keyctx* K, int32_t a, int32_t e, int lev) {
  if (a == e) return K->sorted[a];
  if (K->next >= K->capacity) {
    K->err = OR_ERR_POOL;
    return -1;
  }
  const int64_t c = K->next++;
  K->level[c] = lev;
  K->start[c] = a;
  K->count[c] = e - a + 1;
  int32_t* ch = K->child + 4 * c;
  ch[0] = ch[1] = ch[2] = ch[3] = -1;
  if (lev == OR_MAX_LEVEL) {
    if (e - a + 1 > 4) {
      K->err = OR_ERR_BUCKET;
      return -1;
    }
    for (int32_t s = a; s <= e; ++s) ch[s - a] = K->sorted[s];
    return K->n + (int32_t)c;
  }
  int32_t u = a;
  while (u <= e && !K->err) {
    const int d = digit_at(K, u, lev + 1);
    int32_t v = u;
    while (v + 1 <= e && digit_at(K, v + 1, lev + 1) == d) ++v;
    ch[d] = key_assign(K, u, v, lev + 1);
    u = v + 1;
  }
  return K->n + (int32_t)c;
}

static const uint64_t* g_k1;
static const uint64_t* g_k2;
static int cmp_key(const void* pa, const void* pb) {
  const int32_t a = *(const int32_t*)pa, b = *(const int32_t*)pb;
  if (g_k1[a] != g_k1[b]) return g_k1[a] < g_k1[b] ? -1 : 1;
  if (g_k2[a] != g_k2[b]) return g_k2[a] < g_k2[b] ? -1 : 1;
  return a < b ? -1 : (a > b);
}

int or_build_keys(int32_t n, const float* x, const float* y, const float* bbox,
                  int64_t capacity, int32_t* sorted, int32_t* root,
                  int32_t* child, int32_t* count, int32_t* start,
                  int32_t* level, int64_t* ncells) {
  if (n < 0) return OR_ERR_ARG;
  uint64_t* k1 = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n + 1));
  uint64_t* k2 = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n + 1));
  if (!k1 || !k2) {
    free(k1);
    free(k2);
    return OR_ERR_ARG;
  }
  or_path_keys(n, x, y, bbox, k1, k2);
  for (int32_t i = 0; i < n; ++i) sorted[i] = i;
  g_k1 = k1;
  g_k2 = k2;
  qsort(sorted, (size_t)n, sizeof(int32_t), cmp_key);
  keyctx K = {n, k1, k2, sorted, child, count, start, level, capacity, 0, 0};
  root[0] = root[1] = root[2] = root[3] = -1;
  int32_t u = 0;
  while (u < n && !K.err) {
    const int d = digit_at(&K, u, 1);
    int32_t v = u;
    while (v + 1 < n && digit_at(&K, v + 1, 1) == d) ++v;
    root[d] = key_assign(&K, u, v, 1);
    u = v + 1;
  }
  *ncells = K.next;
  free(k1);
  free(k2);
  return K.err;
}

/* Bottom-up FP64 sums in slot order; cells are visited from the highest
 * pre-order id down so every child cell is complete before its parent. */
void or_com_exact(int32_t n, int64_t ncells, const int32_t* child,
                  const float* x, const float* y, const float* mass,
                  float* ccx, float* ccy, float* ccm) {
  double* S = (double*)malloc(sizeof(double) * 3 * (size_t)(ncells + 1));
  for (int64_t c = ncells - 1; c >= 0; --c) {
    double sm = 0.0, sx = 0.0, sy = 0.0;
    for (int k = 0; k < 4; ++k) {
      const int32_t ch = child[4 * c + k];
      if (ch < 0) continue;
      if (ch < n) {
        const double m = (double)mass[ch];
        sm += m;
        sx += m * (double)x[ch];
        sy += m * (double)y[ch];
      } else {
        const double* q = S + 3 * (size_t)(ch - n);
        sm += q[0];
        sx += q[1];
        sy += q[2];
      }
    }
    S[3 * c + 0] = sm;
    S[3 * c + 1] = sx;
    S[3 * c + 2] = sy;
    /* centre of mass as the sums times the reciprocal mass */
    const double inv = sm != 0.0 ? 1.0 / sm : 0.0;
    ccm[c] = (float)sm;
    ccx[c] = (float)(sx * inv);
    ccy[c] = (float)(sy * inv);
  }
  free(S);
}

/* ------------------------------------------------------------------------ */
/* compute_forces_kernel (kernels.cu:340-435).  Bodies are taken in groups of
 * G consecutive entries of sorted[] (the reference's warps of 32, aligned by
 * the grid stride).  Each group walks one shared stack:
 *   radius = 0.5*(right-left)   (x extent only, :349)
 *   root cells pushed with depth = radius^2/theta (:370-388)
 *   pop: dp = 0.25*depth (:394-395); for each child slot 0..3 (:397):
 *     r = dx*dx + dy*dy + eps2 (:403-405)
 *     body, or every group member has dp <= r (:406): accumulate
 *        f = m*rsqrt(r)^3, acc += f*(dx,dy)  (:407-411)
 *     else push (child, dp) (:413-419)
 * A body sitting alone in a root slot is accumulated here; the reference
 * pushes it and then walks its (empty) child slots, so it drops out.
 *
 * Extensions (not in the reference, off by default):
 *   opening = 1: depth = radius^2/theta^2 and the test is dp + eps2 <= r,
 *     i.e. the textbook w/d <= theta of the reference's older, commented-out
 *     kernel (kernels.cu:488-489) on the unsoftened distance d.  (With eps2
 *     in both sides' r only, as the reference has it, every cell narrower
 *     than about theta*eps passes -- including the cell that holds the body
 *     itself, whose centre of mass then stands in for the body's neighbours.)
 *   cthr != NULL: a cell's own threshold (or_bmax_exact) replaces dp.
 *   cq != NULL: accepted cells add their quadrupole term, the second-order
 *     term of the softened potential's expansion about the cell's centre of
 *     mass: a += -Q.d/r^5 + (5/2)(d.Q.d - eps2 tr Q) d/r^7, r^2 = |d|^2 +
 *     eps2 (the eps2 tr Q part is what softening adds to the unsoftened
 *     form). */
#define OR_STACK 256

static inline float depth0_of(const float* bbox, float theta, int opening) {
  const float radius = (float)(0.5 * (double)(bbox[1] - bbox[0]));
  const float rr = radius * radius;
  const float t = opening ? theta * theta : theta;
  return rr / t;
}

static inline float rsqrt_ref(float r) { return (float)(1.0 / sqrt((double)r)); }

/* r = dx*dx + dy*dy + eps2 (kernels.cu:403-405).  The reference leaves the
 * rounding to nvcc's FMA contraction; this build fixes one contraction,
 * dx*dx + (dy*dy + eps2) as two fused multiply-adds, on both sides. */
static inline float dist2(float dx, float dy, float eps2) {
  return fmaf(dx, dx, fmaf(dy, dy, eps2));
}

/* FP32 accumulation of one accepted node for one body.  With q (a cell's
 * quadrupole, tr Q = q[0] + q[2]) the terms
 *   m d/r^3 - Q.d/r^5 + (5/2)(d.Q.d - eps2 tr Q) d/r^7
 * -- the expansion of the softened sum m_i / sqrt(|d + s_i|^2 + eps2) to
 * second order in the offsets s_i -- are evaluated as
 *   rs^5 [ (m r + (2.5 (d.Q.d) + T) rs^2) d - Q.d ],
 * r = |d|^2 + eps2, rs = 1/sqrt(r), T = (q[0] + q[2]) kq with kq = -2.5 eps2
 * (FP32): the form the device walk uses. */
static inline void acc_node(float* fx, float* fy, float dx, float dy, float r,
                            float cmv, const float* q, float kq) {
  const float rs = rsqrt_ref(r);
  if (!q) {
    const float f = cmv * rs * rs * rs;
    *fx = fmaf(f, dx, *fx);
    *fy = fmaf(f, dy, *fy);
    return;
  }
  const float u = rs * rs;
  const float rs5 = (u * u) * rs;
  const float px = fmaf(q[0], dx, q[1] * dy);
  const float py = fmaf(q[1], dx, q[2] * dy);
  const float s = fmaf(dx, px, dy * py);
  const float tq = (q[0] + q[2]) * kq; /* -2.5 eps2 tr Q */
  const float c = fmaf(fmaf(s, 2.5f, tq), u, cmv * r);
  *fx = fmaf(rs5, fmaf(c, dx, -px), *fx);
  *fy = fmaf(rs5, fmaf(c, dy, -py), *fy);
}

/* ---- FP16 arithmetic (mixed precision), emulated in double -------------- */
/* x rounded to an FP16 value: nearest even (to_zero = 0) or toward zero,
 * with subnormals (quantum 2^-24); nearest-even overflow gives +-inf, toward
 * zero saturates at +-65504. */
static double f16_round(double x, int to_zero) {
  if (x == 0.0 || x != x) return x;
  const double ax = fabs(x);
  int e;
  (void)frexp(ax, &e); /* ax = m 2^e, m in [0.5, 1): leading bit 2^(e-1) */
  const int lead = e - 1 < -14 ? -14 : e - 1;
  const double q = ldexp(1.0, lead - 10);
  double r = ax / q;
  r = to_zero ? trunc(r) : nearbyint(r);
  double v = r * q;
  if (v > 65504.0) v = to_zero ? 65504.0 : INFINITY;
  return x < 0 ? -v : v;
}
static inline double h16(double x) { return f16_round(x, 0); }
double or_f16_round(double x, int32_t to_zero) { return f16_round(x, to_zero); }
static inline double hmul(double a, double b) { return h16(a * b); }
static inline double hfma(double a, double b, double c) { return h16(fma(a, b, c)); }

/* Per-body state of the mixed far field. */
typedef struct mixacc {
  float sx, sy, cx, cy;     /* Kahan sums */
  float pex, pey, pk, pt;   /* the held cell */
  float peu;
} mixacc;

static inline void kahan(float* sum, float* comp, float v) {
  const float yv = v - *comp;
  const float t = *sum + yv;
  *comp = (t - *sum) - yv;
  *sum = t;
}

/* Two accepted cells (a: held, b: current) for one body. */
static void mix_pair(mixacc* M, const float* qa, float ex, float ey, float k, float t,
                     float eu, const float* qb) {
  const double EXa = f16_round(M->pex, 1), EXb = f16_round(ex, 1);
  const double EYa = f16_round(M->pey, 1), EYb = f16_round(ey, 1);
  const double Ta = f16_round(M->pt, 1), Tb = f16_round(t, 1);
  const double EUa = f16_round(M->peu, 1), EUb = f16_round(eu, 1);
  double BX[2], BY[2];
  const double EX[2] = {EXa, EXb}, EY[2] = {EYa, EYb}, T[2] = {Ta, Tb}, EU[2] = {EUa, EUb};
  const float* Q[2] = {qa, qb};
  for (int h = 0; h < 2; ++h) {
    const double qxx = Q[h][0], qxy = Q[h][1], qyy = Q[h][2];
    const double qex = hfma(qxx, EX[h], hmul(qxy, EY[h]));
    const double qey = hfma(qxy, EX[h], hmul(qyy, EY[h]));
    const double s = hfma(EX[h], qex, hmul(EY[h], qey));
    /* softening: e.Qh.e - tr(Qh) eps2 rs^2 */
    const double trh = h16(qxx + qyy);
    const double cc = hmul(hfma(-trh, EU[h], s), 2.5);
    BX[h] = hfma(T[h], hfma(cc, EX[h], -qex), EX[h]);
    BY[h] = hfma(T[h], hfma(cc, EY[h], -qey), EY[h]);
  }
  kahan(&M->sx, &M->cx, fmaf(M->pk, (float)BX[0], k * (float)BX[1]));
  kahan(&M->sy, &M->cy, fmaf(M->pk, (float)BY[0], k * (float)BY[1]));
}

/* Hilbert frames: base order of the four quadrants (x: 1 east, y: 1 north)
 * SW, NW, NE, SE; frame bit 0 transposes, bit 1 turns by 180 degrees (they
 * commute); the child at curve position h gets frame ^ kHilbertChildFrame[h]. */
static const int kHilbertChildFrame[4] = {1, 0, 0, 3};

/* Slot of the child at curve position h in frame f. */
static int hilbert_slot(int f, int h) {
  static const int bx[4] = {0, 0, 1, 1}, by[4] = {0, 1, 1, 0};
  int x = bx[h], y = by[h];
  if (f & 1) {
    const int t = x;
    x = y;
    y = t;
  }
  if (f & 2) {
    x = 1 - x;
    y = 1 - y;
  }
  return (1 - x) | ((1 - y) << 1);
}

int or_hilbert_order(int32_t n, const int32_t* root, const int32_t* child, int32_t l1,
                     int32_t* order) {
  /* explicit stack of (node code, level, frame); node -1 is the root */
  enum { kStack = 4 * (OR_MAX_LEVEL + 2) };
  int32_t sn[kStack], sl[kStack], sf[kStack];
  int sp = 0, k = 0;
  sn[sp] = -1;
  sl[sp] = 0;
  sf[sp++] = 0;
  while (sp > 0) {
    --sp;
    const int32_t node = sn[sp], lev = sl[sp], f = sf[sp];
    if (node >= 0 && node < n) {
      if (k >= n) return OR_ERR_ARG;
      order[k++] = node;
      continue;
    }
    const int32_t* kids = node < 0 ? root : child + 4 * (int64_t)(node - n);
    /* children pushed in reverse visiting order */
    for (int h = 3; h >= 0; --h) {
      const int hil = lev + 1 <= l1;
      const int slot = hil ? hilbert_slot(f, h) : h;
      const int32_t ch = kids[slot];
      if (ch < 0) continue;
      if (sp >= kStack) return OR_ERR_DEPTH;
      sn[sp] = ch;
      sl[sp] = lev + 1;
      sf[sp++] = hil ? (f ^ kHilbertChildFrame[h]) : f;
    }
  }
  return k == n ? OR_OK : OR_ERR_ARG;
}

int or_walk(int32_t n, const float* x, const float* y, const float* mass,
            const float* bbox, const int32_t* sorted, const int32_t* root,
            const int32_t* child, int64_t ncells, const float* ccx,
            const float* ccy, const float* ccm, const or_walk_params* p,
            float* ax, float* ay, int32_t* inter) {
  (void)ncells;
  const int32_t G = p->group;
  if (G <= 0 || G > 1024) return OR_ERR_ARG;
  const float eps2 = p->eps2;
  const float depth0 = depth0_of(bbox, p->theta, p->opening);
  const float* cq = p->cq;
  const int32_t* grp = p->order ? p->order : sorted; /* body order of the walk groups */
  const int32_t ngroups = (n + G - 1) / G;
  const int mixed = p->mixed && p->cqh && p->clevel;
  float s2 = 0.f;
  if (mixed) {
    const float w = bbox[1] - bbox[0], h = bbox[3] - bbox[2];
    const float sz = w > h ? w : h;
    s2 = sz * sz;
  }
  static const float qzero[3] = {0.f, 0.f, 0.f};
  int err = OR_OK;
#pragma omp parallel num_threads(nthreads(p->threads))
  {
    float* px = (float*)malloc(sizeof(float) * (size_t)G * 5);
    float* py = px + G;
    float* fx = py + G;
    float* fy = fx + G;
    float* rr = fy + G;
    mixacc* mx = (mixacc*)malloc(sizeof(mixacc) * (size_t)G);
    const float* pq = qzero; /* the held cell's Qh (the whole group holds one) */
    int pend = 0;
    int32_t stk[OR_STACK];
    float dep[OR_STACK];
#pragma omp for schedule(dynamic, 8)
    for (int32_t g = 0; g < ngroups; ++g) {
      const int32_t s0 = g * G;
      const int32_t gs = (n - s0) < G ? (n - s0) : G;
      int32_t cnt = 0;
      for (int32_t j = 0; j < gs; ++j) {
        const int32_t b = grp[s0 + j];
        px[j] = x[b];
        py[j] = y[b];
        fx[j] = 0.f;
        fy[j] = 0.f;
        memset(&mx[j], 0, sizeof(mixacc));
      }
      pend = 0;
      pq = qzero;
      int sp = 0;
      for (int k = 0; k < 4; ++k) {
        const int32_t ch = root[k];
        if (ch < 0) continue;
        if (ch < n && p->ref_root) {
          /* pushed like a cell; -1 marks the root's own slots (body 0) */
          if (ch == 0) {
            stk[sp] = -1;
            dep[sp++] = depth0;
          }
          continue;
        }
        if (ch < n) {
          for (int32_t j = 0; j < gs; ++j) {
            const float dx = x[ch] - px[j], dy = y[ch] - py[j];
            const float r = dist2(dx, dy, eps2);
            acc_node(&fx[j], &fy[j], dx, dy, r, mass[ch], NULL, 0.f);
          }
          ++cnt;
        } else {
          stk[sp] = ch;
          dep[sp++] = depth0;
        }
      }
      while (sp > 0) {
        --sp;
        const int32_t node = stk[sp];
        const float dp = 0.25f * dep[sp];
        const int32_t* kids = node < 0 ? root : child + 4 * (int64_t)(node - n);
        for (int k = 0; k < 4; ++k) {
          const int32_t ch = kids[k];
          if (ch < 0) continue;
          float cxv, cyv, cmv;
          const float* q = NULL;
          if (ch < n) {
            cxv = x[ch];
            cyv = y[ch];
            cmv = mass[ch];
          } else {
            cxv = ccx[ch - n];
            cyv = ccy[ch - n];
            cmv = ccm[ch - n];
            if (cq) q = cq + 3 * (int64_t)(ch - n);
          }
          const float th = (p->cthr && ch >= n) ? p->cthr[ch - n] : (p->opening ? dp + eps2 : dp);
          int all = 1;
          for (int32_t j = 0; j < gs; ++j) {
            const float dx = cxv - px[j], dy = cyv - py[j];
            rr[j] = dist2(dx, dy, eps2);
            if (!(th <= rr[j])) all = 0;
          }
          if (mixed && ch >= n && all) {
            const float* qc = p->cqh + 3 * (int64_t)(ch - n);
            const float ws2 = ldexpf(s2, -2 * p->clevel[ch - n]);
            for (int32_t j = 0; j < gs; ++j) {
              const float dx = cxv - px[j], dy = cyv - py[j];
              const float rs = rsqrt_ref(rr[j]);
              const float u = rs * rs;
              const float ex = dx * rs, ey = dy * rs, k = cmv * u, t = u * ws2;
              const float eu = eps2 * u;
              if (pend) {
                mix_pair(&mx[j], pq, ex, ey, k, t, eu, qc);
              } else {
                mx[j].pex = ex;
                mx[j].pey = ey;
                mx[j].pk = k;
                mx[j].pt = t;
                mx[j].peu = eu;
              }
            }
            if (pend) {
              pend = 0;
            } else {
              pend = 1;
              pq = qc;
            }
            ++cnt;
          } else if (ch < n || all) {
            for (int32_t j = 0; j < gs; ++j) {
              const float dx = cxv - px[j], dy = cyv - py[j];
              acc_node(&fx[j], &fy[j], dx, dy, rr[j], cmv, q, -2.5f * eps2);
            }
            ++cnt;
          } else {
            if (sp >= OR_STACK) {
#pragma omp atomic write
              err = OR_ERR_DEPTH;
              sp = 0;
              break;
            }
            stk[sp] = ch;
            dep[sp++] = dp;
          }
        }
      }
      if (mixed) {
        for (int32_t j = 0; j < gs; ++j) {
          if (pend) mix_pair(&mx[j], pq, 0.f, 0.f, 0.f, 0.f, 0.f, qzero);
          fx[j] = fx[j] + mx[j].sx;
          fy[j] = fy[j] + mx[j].sy;
        }
      }
      for (int32_t j = 0; j < gs; ++j) {
        const int32_t b = grp[s0 + j];
        ax[b] = fx[j];
        ay[b] = fy[j];
        if (inter) inter[b] = cnt;
      }
    }
    free(px);
    free(mx);
  }
  return err;
}

/* Same acceptance rule stated recursively, accumulated in FP64. */
typedef struct recctx {
  int32_t n, gs, opening;
  const int32_t* root;
  const float *x, *y, *mass, *ccx, *ccy, *ccm, *cq, *cthr;
  const int32_t* child;
  float eps2;
  const float* px;
  const float* py;
  double* fx;
  double* fy;
  int32_t cnt;
} recctx;

static void rec_acc(recctx* R, float cxv, float cyv, float cmv, const float* q) {
  for (int32_t j = 0; j < R->gs; ++j) {
    const float dx = cxv - R->px[j], dy = cyv - R->py[j];
    const float r = dist2(dx, dy, R->eps2);
    const double rs = 1.0 / sqrt((double)r);
    const double X = dx, Y = dy;
    double gx = (double)cmv * rs * rs * rs * X, gy = (double)cmv * rs * rs * rs * Y;
    if (q) {
      const double rs2 = rs * rs, rs5 = rs2 * rs2 * rs, rs7 = rs5 * rs2;
      const double qdx = (double)q[0] * X + (double)q[1] * Y;
      const double qdy = (double)q[1] * X + (double)q[2] * Y;
      const double dqd = X * qdx + Y * qdy - ((double)q[0] + (double)q[2]) * (double)R->eps2;
      gx += -rs5 * qdx + 2.5 * dqd * rs7 * X;
      gy += -rs5 * qdy + 2.5 * dqd * rs7 * Y;
    }
    R->fx[j] += gx;
    R->fy[j] += gy;
  }
  R->cnt++;
}

/* node < 0: the root's own slots (a lone body 0 in the reference's handling) */
static void rec_open(recctx* R, int32_t node, float depth) {
  const float dp = 0.25f * depth;
  for (int k = 0; k < 4; ++k) {
    const int32_t ch = node < 0 ? R->root[k] : R->child[4 * (int64_t)(node - R->n) + k];
    if (ch < 0) continue;
    if (ch < R->n) {
      rec_acc(R, R->x[ch], R->y[ch], R->mass[ch], NULL);
      continue;
    }
    const int32_t c = ch - R->n;
    const float th = R->cthr ? R->cthr[c] : (R->opening ? dp + R->eps2 : dp);
    int all = 1;
    for (int32_t j = 0; j < R->gs && all; ++j) {
      const float dx = R->ccx[c] - R->px[j], dy = R->ccy[c] - R->py[j];
      const float r = dist2(dx, dy, R->eps2);
      if (!(th <= r)) all = 0;
    }
    if (all)
      rec_acc(R, R->ccx[c], R->ccy[c], R->ccm[c], R->cq ? R->cq + 3 * (int64_t)c : NULL);
    else
      rec_open(R, ch, dp);
  }
}

int or_walk_recursive(int32_t n, const float* x, const float* y,
                      const float* mass, const float* bbox,
                      const int32_t* sorted, const int32_t* root,
                      const int32_t* child, int64_t ncells, const float* ccx,
                      const float* ccy, const float* ccm,
                      const or_walk_params* p, double* ax, double* ay,
                      int32_t* inter) {
  (void)ncells;
  const int32_t G = p->group;
  if (G <= 0 || G > 1024) return OR_ERR_ARG;
  const float depth0 = depth0_of(bbox, p->theta, p->opening);
  const int32_t ngroups = (n + G - 1) / G;
  const int32_t* grp = p->order ? p->order : sorted; /* body order of the walk groups */
#pragma omp parallel num_threads(nthreads(p->threads))
  {
    float* px = (float*)malloc(sizeof(float) * (size_t)G * 2);
    double* f = (double*)malloc(sizeof(double) * (size_t)G * 2);
#pragma omp for schedule(dynamic, 8)
    for (int32_t g = 0; g < ngroups; ++g) {
      const int32_t s0 = g * G;
      const int32_t gs = (n - s0) < G ? (n - s0) : G;
      for (int32_t j = 0; j < gs; ++j) {
        px[j] = x[grp[s0 + j]];
        px[G + j] = y[grp[s0 + j]];
        f[j] = f[G + j] = 0.0;
      }
      recctx R = {n, gs, p->opening, root, x, y, mass, ccx, ccy, ccm, p->cq, p->cthr, child,
                  p->eps2, px, px + G, f, f + G, 0};
      for (int k = 0; k < 4; ++k) {
        const int32_t ch = root[k];
        if (ch < 0) continue;
        if (ch < n && p->ref_root) {
          if (ch == 0) rec_open(&R, -1, depth0);
        } else if (ch < n) {
          rec_acc(&R, x[ch], y[ch], mass[ch], NULL);
        } else {
          rec_open(&R, ch, depth0);
        }
      }
      for (int32_t j = 0; j < gs; ++j) {
        const int32_t b = grp[s0 + j];
        ax[b] = f[j];
        ay[b] = f[G + j];
        if (inter) inter[b] = R.cnt;
      }
    }
    free(px);
    free(f);
  }
  return OR_OK;
}

/* FP64 moments per cell, bottom-up in slot order: S[6c..6c+5] = sum m,
 * sum m x, sum m y, Qxx, Qxy, Qyy (traceless, about the FP64 centre of
 * mass).  The caller frees S. */
static double* quad_moments(int32_t n, int64_t ncells, const int32_t* child,
                            const float* x, const float* y, const float* mass) {
  double* S = (double*)malloc(sizeof(double) * 6 * (size_t)(ncells + 1));
  for (int64_t c = ncells - 1; c >= 0; --c) {
    double cm[4], cx[4], cy[4];
    const double* qq[4];
    double sm = 0.0, sx = 0.0, sy = 0.0;
    for (int k = 0; k < 4; ++k) {
      const int32_t ch = child[4 * c + k];
      qq[k] = NULL;
      cm[k] = 0.0;
      cx[k] = cy[k] = 0.0;
      if (ch < 0) continue;
      if (ch < n) {
        cm[k] = (double)mass[ch];
        cx[k] = (double)x[ch];
        cy[k] = (double)y[ch];
        sm += cm[k];
        sx += cm[k] * cx[k];
        sy += cm[k] * cy[k];
      } else {
        const double* q = S + 6 * (size_t)(ch - n);
        cm[k] = q[0];
        const double ci = q[0] != 0.0 ? 1.0 / q[0] : 0.0;
        cx[k] = q[1] * ci;
        cy[k] = q[2] * ci;
        qq[k] = q + 3;
        sm += q[0];
        sx += q[1];
        sy += q[2];
      }
    }
    const double inv = sm != 0.0 ? 1.0 / sm : 0.0;
    const double ox = sx * inv, oy = sy * inv;
    double qxx = 0.0, qxy = 0.0, qyy = 0.0;
    for (int k = 0; k < 4; ++k) {
      if (child[4 * c + k] < 0) continue;
      const double dx = cx[k] - ox, dy = cy[k] - oy;
      const double t1 = dx * dx, t2 = dy * dy;
      double axx = cm[k] * (2.0 * t1 - t2);
      double axy = cm[k] * (3.0 * (dx * dy));
      double ayy = cm[k] * (2.0 * t2 - t1);
      if (qq[k]) {
        axx = qq[k][0] + axx;
        axy = qq[k][1] + axy;
        ayy = qq[k][2] + ayy;
      }
      qxx += axx;
      qxy += axy;
      qyy += ayy;
    }
    double* o = S + 6 * (size_t)c;
    o[0] = sm;
    o[1] = sx;
    o[2] = sy;
    o[3] = qxx;
    o[4] = qxy;
    o[5] = qyy;
  }
  return S;
}

/* Quadrupole moments, bottom-up in FP64 (see oracle.h). */
void or_quad_exact(int32_t n, int64_t ncells, const int32_t* child,
                   const float* x, const float* y, const float* mass,
                   float* cq) {
  double* S = quad_moments(n, ncells, child, x, y, mass);
  for (int64_t c = 0; c < ncells; ++c) {
    cq[3 * c + 0] = (float)S[6 * c + 3];
    cq[3 * c + 1] = (float)S[6 * c + 4];
    cq[3 * c + 2] = (float)S[6 * c + 5];
  }
  free(S);
}

void or_quad_mixed(int32_t n, int64_t ncells, const int32_t* child,
                   const int32_t* clevel, const float* x, const float* y,
                   const float* mass, const float* bbox, float* qh) {
  double* S = quad_moments(n, ncells, child, x, y, mass);
  const float w = bbox[1] - bbox[0], h = bbox[3] - bbox[2];
  const float sz = w > h ? w : h;
  const float s2 = sz * sz;
  for (int64_t c = 0; c < ncells; ++c) {
    const float ws2 = ldexpf(s2, -2 * clevel[c]);
    const double den = S[6 * c] * (double)ws2;
    for (int k = 0; k < 3; ++k) {
      float v = den != 0.0 ? (float)(S[6 * c + 3 + k] / den) : 0.f;
      v = v > 65504.f ? 65504.f : (v < -65504.f ? -65504.f : v);
      qh[3 * c + k] = (float)h16((double)v);
    }
  }
  free(S);
}

/* Opening thresholds from the bmax bound (not in the reference). */
void or_bmax_exact(int32_t n, int64_t ncells, const int32_t* child,
                   const float* x, const float* y, const float* mass,
                   float theta, float eps2, float* cthr) {
  double* S = (double*)malloc(sizeof(double) * 4 * (size_t)(ncells + 1));
  const double t2 = (double)theta * (double)theta;
  for (int64_t c = ncells - 1; c >= 0; --c) {
    double cx[4], cy[4], cb[4];
    double sm = 0.0, sx = 0.0, sy = 0.0;
    for (int k = 0; k < 4; ++k) {
      const int32_t ch = child[4 * c + k];
      cx[k] = cy[k] = cb[k] = 0.0;
      if (ch < 0) continue;
      if (ch < n) {
        const double m = (double)mass[ch];
        cx[k] = (double)x[ch];
        cy[k] = (double)y[ch];
        sm += m;
        sx += m * cx[k];
        sy += m * cy[k];
      } else {
        const double* q = S + 4 * (size_t)(ch - n);
        const double ci = q[0] != 0.0 ? 1.0 / q[0] : 0.0;
        cx[k] = q[1] * ci;
        cy[k] = q[2] * ci;
        cb[k] = q[3];
        sm += q[0];
        sx += q[1];
        sy += q[2];
      }
    }
    const double inv = sm != 0.0 ? 1.0 / sm : 0.0;
    const double ox = sx * inv, oy = sy * inv;
    double b = 0.0;
    for (int k = 0; k < 4; ++k) {
      if (child[4 * c + k] < 0) continue;
      const double dx = cx[k] - ox, dy = cy[k] - oy;
      const double d = sqrt(fma(dx, dx, dy * dy)) + cb[k];
      b = d > b ? d : b;
    }
    double* o = S + 4 * (size_t)c;
    o[0] = sm;
    o[1] = sx;
    o[2] = sy;
    o[3] = b;
    cthr[c] = (float)((b * b) / t2 + (double)eps2);
  }
  free(S);
}

/* ------------------------------------------------------------------------ */
void or_direct_fp64(int32_t n, const float* x, const float* y,
                    const float* mass, double eps2, int32_t ntargets,
                    const int32_t* targets, double* ax, double* ay,
                    int32_t threads) {
  const int32_t nt = targets ? ntargets : n;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads(threads))
  for (int32_t q = 0; q < nt; ++q) {
    const int32_t i = targets ? targets[q] : q;
    const double xi = x[i], yi = y[i];
    double sx = 0.0, sy = 0.0;
    for (int32_t j = 0; j < n; ++j) {
      const double dx = (double)x[j] - xi, dy = (double)y[j] - yi;
      const double r2 = dx * dx + dy * dy + eps2;
      const double inv = 1.0 / (r2 * sqrt(r2));
      sx += (double)mass[j] * dx * inv;
      sy += (double)mass[j] * dy * inv;
    }
    ax[q] = sx;
    ay[q] = sy;
  }
}

/* update_kernel (kernels.cu:522-537): v += dt*a, then x += d*dt*v; nvcc
 * contracts both to fused multiply-adds with d*dt rounded first. */
void or_update(int32_t n, float* x, float* y, float* vx, float* vy,
               const float* ax, const float* ay, float dt, float d) {
  const float ddt = d * dt;
  for (int32_t i = 0; i < n; ++i) {
    vx[i] = fmaf(dt, ax[i], vx[i]);
    vy[i] = fmaf(dt, ay[i], vy[i]);
    x[i] = fmaf(ddt, vx[i], x[i]);
    y[i] = fmaf(ddt, vy[i], y[i]);
  }
}

/* debug.cpp:6-178 DEBUG_RUN_TESTS, as a failure bitmask. */
int32_t or_invariants(int32_t n, int32_t m, const int32_t* count,
                      const int32_t* sorted, const int32_t* child,
                      const float* bbox) {
  int32_t fail = 0;
  for (int32_t i = n; i < m; ++i) { /* COUNT 1, :23-41 */
    int64_t c = count[i];
    for (int j = 0; j < 4; ++j) {
      const int32_t ch = child[4 * (int64_t)i + j];
      if (ch != -1 && ch >= 0 && ch < m) c -= count[ch];
    }
    if (c != 0) fail |= 1;
  }
  int64_t sum = 0;
  for (int64_t j = 0; j < 4 * (int64_t)m; ++j) /* CHILD 1, :50-61 */
    if (child[j] >= 0 && child[j] < n) sum += child[j];
  if (sum != (int64_t)n * (int64_t)(n - 1) / 2) fail |= 2;
  for (int64_t j = 0; j < 4 * (int64_t)m; ++j) { /* CHILD 2/3, :65-91 */
    if (child[j] >= m) fail |= 4;
    if (child[j] < -1) fail |= 8;
  }
  int32_t mn = 2 * n; /* CHILD 4, :95-112 */
  for (int k = 0; k < 4; ++k)
    if (child[k] < mn) mn = child[k];
  for (int64_t j = 4; j < 4 * (int64_t)(mn - 1); ++j)
    if (child[j] != -1) fail |= 16;
  sum = 0; /* SORTED 1, :121-130 */
  for (int32_t i = 0; i < m; ++i) sum += sorted[i];
  if (sum != (int64_t)n * (int64_t)(n - 1) / 2) fail |= 32;
  char* seen = (char*)calloc((size_t)n + 1, 1); /* SORTED 2, :134-153 */
  for (int32_t i = 0; i < n; ++i)
    if (sorted[i] >= 0 && sorted[i] < n) seen[sorted[i]] = 1;
  for (int32_t i = 0; i < n; ++i)
    if (!seen[i]) fail |= 64;
  free(seen);
  if (bbox[0] > bbox[1] || bbox[2] > bbox[3]) fail |= 128; /* BBOX 1 */
  return fail;
}

void or_energy(int32_t n, const float* x, const float* y, const float* vx,
               const float* vy, const float* mass, double eps2,
               int32_t threads, double* kinetic, double* potential,
               double* px, double* py) {
  double out[5];
  or_diagnostics(n, x, y, vx, vy, mass, NULL, NULL, 0.0, eps2, threads, out);
  *kinetic = out[0];
  *potential = out[1];
  *px = out[2];
  *py = out[3];
}

/* Velocities are taken as v + half_dt * a (FP64) when ax is given: with the
 * kick-drift update of kernels.cu:522-537 the stored v lags the positions by
 * half a step, and a = the accelerations at the stored positions brings it
 * level.  Angular momentum is about the origin: Lz = sum m (x vy - y vx). */
void or_diagnostics(int32_t n, const float* x, const float* y, const float* vx,
                    const float* vy, const float* mass, const float* ax,
                    const float* ay, double half_dt, double eps2,
                    int32_t threads, double* out) {
  double k = 0.0, u = 0.0, mx = 0.0, my = 0.0, lz = 0.0;
  for (int32_t i = 0; i < n; ++i) {
    const double m = mass[i];
    const double vu = ax ? (double)vx[i] + half_dt * (double)ax[i] : (double)vx[i];
    const double vv = ax ? (double)vy[i] + half_dt * (double)ay[i] : (double)vy[i];
    k += 0.5 * m * (vu * vu + vv * vv);
    mx += m * vu;
    my += m * vv;
    lz += m * ((double)x[i] * vv - (double)y[i] * vu);
  }
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : u) \
    num_threads(nthreads(threads))
  for (int32_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int32_t j = i + 1; j < n; ++j) {
      const double dx = (double)x[j] - x[i], dy = (double)y[j] - y[i];
      s += (double)mass[j] / sqrt(dx * dx + dy * dy + eps2);
    }
    u -= (double)mass[i] * s;
  }
  out[0] = k;
  out[1] = u;
  out[2] = mx;
  out[3] = my;
  out[4] = lz;
}

/* ------------------------------------------------------------------------ */
static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* BarnesHutParticleSystem::update (BarnesHutParticleSystem.cpp:301-316):
 * reset, bbox, build, CoM, sort, forces, update, in that order. */
int or_step(int32_t n, float* x, float* y, float* vx, float* vy,
            const float* mass, float* ax, float* ay, const or_step_params* p,
            double* stage_seconds) {
  const int32_t m = p->m > 0 ? p->m : 2 * n + 12000;
  int32_t* child = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)m);
  int32_t* count = (int32_t*)malloc(sizeof(int32_t) * (size_t)m);
  int32_t* start = (int32_t*)malloc(sizeof(int32_t) * (size_t)m);
  int32_t* sorted = (int32_t*)calloc((size_t)m, sizeof(int32_t));
  float* cx = (float*)malloc(sizeof(float) * (size_t)m);
  float* cy = (float*)malloc(sizeof(float) * (size_t)m);
  float* cm = (float*)malloc(sizeof(float) * (size_t)m);
  int rc = OR_OK;
  double t[7];
  float bbox[4];
  int32_t ncells = 0;
  if (!child || !count || !start || !sorted || !cx || !cy || !cm) {
    rc = OR_ERR_ARG;
    goto out;
  }
  t[0] = now_s();
  or_bbox(n, x, y, bbox);
  t[1] = now_s();
  rc = or_build_insert(n, m, x, y, mass, bbox, child, count, start, cx, cy, cm,
                       &ncells);
  if (rc) goto out;
  t[2] = now_s();
  or_com_ref(n, ncells, cx, cy, cm);
  t[3] = now_s();
  or_sort_ref(n, ncells, child, count, start, sorted);
  t[4] = now_s();
  {
    /* the walk reads the pool layout: cells at n.. with CoM in cx/cy/cm */
    or_walk_params wp = {p->theta, p->eps2, p->group, p->threads, 0, 0, NULL, NULL, NULL, NULL, 0, 0, NULL};
    rc = or_walk(n, x, y, mass, bbox, sorted, child, child + 4 * (size_t)n,
                 ncells, cx + n, cy + n, cm + n, &wp, ax, ay, NULL);
    if (rc) goto out;
  }
  t[5] = now_s();
  or_update(n, x, y, vx, vy, ax, ay, p->dt, p->dampening);
  t[6] = now_s();
  if (stage_seconds)
    for (int k = 0; k < 6; ++k) stage_seconds[k] = t[k + 1] - t[k];
out:
  free(child);
  free(count);
  free(start);
  free(sorted);
  free(cx);
  free(cy);
  free(cm);
  return rc;
}
