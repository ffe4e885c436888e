// Tree-walk force kernel for gfx950 (compute_forces_kernel, kernels.cu:340-435).
//
// A walk group is G consecutive bodies in tree order (the reference uses warps
// of 32).  The group shares one traversal: for each child of a visited node it
// evaluates r = dx*dx + dy*dy + eps2 per body and accepts the child when it is
// a body or when every body of the group passes dp <= r (the reference's
// __all vote); otherwise the child is opened.  Accepted children add
// f = m*rsqrt(r)^3 * (dx, dy) to every body of the group.
//
// gfx950 mapping (G = 64, walk64_kernel, one wavefront per group):
//   * node indices are wavefront-uniform, so a node's four child records (one
//     64-byte block) arrive through scalar loads, and each iteration pops up
//     to KPOP nodes so that one memory round trip serves several nodes;
//   * the vote is one 64-lane ballot; lanes past the last body repeat the
//     last body, so their votes change nothing and need no mask;
//   * the stack (in LDS, one region per wavefront, written by lane 0) holds
//     node index and level in one word; the opening threshold of a node's
//     children (they share a level) comes from a per-lane table of level
//     thresholds, one v_readlane per visited node, except with per-cell
//     thresholds (bmax) or mixed precision, where each cell child's
//     threshold comes with the node's second block;
//   * a child's code decides empty, body or cell with one compare; a body is
//     accepted without a vote;
//   * the eight wavefronts of a workgroup take consecutive groups and each
//     XCD one contiguous range of workgroups (scalar-cache and L2 reuse);
//   * bodies are stored in tree order: the group's positions and outputs are
//     contiguous 256-byte rows.
// G = 32 and 16 (walk_kernel) run 2 or 4 independent groups per wavefront
// with per-group ballot masks, reproducing the reference's warp-of-32 vote.
//
// Which children are accepted depends only on the child and the group, never
// on the visiting order, so every variant makes the same decisions as the CPU
// restatement (bit-exact interaction counts); the order of summation differs.
// One deliberate difference from the reference, unless
// STELLAR_FLAG_REFERENCE_ROOT is set: a body alone in a root slot is
// accumulated; the reference pushes it like a cell and then reads its child
// slots (kernels.cu:373-387, 397-398), so that body exerts no force -- and
// body 0, whose child slots are the root's, makes the group revisit the
// root's children one level down.
//
// QUAD adds each accepted cell's quadrupole term (a second 64-byte block of
// {Qxx, Qxy, Qyy, w} per node); bodies stay monopoles.  The term is the
// second-order term of the softened potential sum m / sqrt(|d + s|^2 + eps2)
// (s a body's offset from the cell's centre of mass), which carries, beside
// the unsoftened form's two terms, -(5/2) tr(Q) eps2 d / r^7: w holds
// T = -2.5 eps2 (Qxx + Qyy) for it.  With per-cell opening thresholds (bmax)
// or mixed precision w holds the threshold instead; bmax walks compute T from
// Q, and the mixed bracket carries the term in FP16 (MixedAcc).
#include <map>
#include <mutex>

#include "engine.h"

namespace stellar {

namespace {

// Walk groups are G consecutive tree positions of one instance: group g is
// [gtab[g], gtab[g+1]) (64, 32 or 16 positions, fewer at an instance's end).
struct WalkArgs {
  const float* x;
  const float* y;
  const ChildRec* blk;
  const float4* qblk;
  int32_t n;
  float eps2;
  float theta_div;  // theta (reference opening) or theta^2 (textbook)
  DevScalars* dev;
  float* ax;
  float* ay;
  int32_t* inter;
  const int32_t* gtab;
  int64_t ngroups;
  const int32_t* inst_of;
  const float4* ibbox;
  int unsoft;  // level thresholds compare with the unsoftened distance
  // STELLAR_FLAG_REFERENCE_ROOT: a body alone in a root slot is handled as
  // the reference handles it (see walk64_group); needs the body ids and the
  // instances' first ids
  int ref_root;
  const int32_t* id;
  const int32_t* ioff;
  // update != 0 (G = 64 walks that pop one node at a time): each group then
  // applies the step's update to its own bodies (update_kernel's arithmetic,
  // v += dt a; x += ddt v with ddt = d dt) right after its forces.  Only the
  // group reads its bodies' x, y (every other group sees them through the
  // tree's copies), so the result equals the separate update launch.
  int update;
  float dt;
  float ddt;
  float* xw;
  float* yw;
  float* vx;
  float* vy;
};

// The fused update's arguments, copied to LDS (one copy per wavefront) at the
// kernel's start and read back after the walk: held as kernel arguments they
// would occupy scalar registers through the walk's loop, which needs all of
// them (spills otherwise).
struct WalkUpdate {
  float* x;
  float* y;
  float* vx;
  float* vy;
  float dt;
  float ddt;
  int on;
  int pad;
};

__device__ __forceinline__ void stage_update(const WalkArgs& A, WalkUpdate* u) {
  if ((threadIdx.x & 63) == 0) *u = WalkUpdate{A.xw, A.yw, A.vx, A.vy, A.dt, A.ddt, A.update, 0};
}

// r = dx*dx + dy*dy + eps2 as dx*dx + (dy*dy + eps2), two fused
// multiply-adds (the CPU restatement evaluates the same form).
__device__ __forceinline__ float dist2(float dx, float dy, float eps2) {
  return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, eps2));
}

// The softening term T = (Qxx + Qyy) kq, kq = -2.5 eps2 (FP32; the
// centre-of-mass pass stores the same product).
__device__ __forceinline__ float quad_soft_term(float4 q, float kq) { return (q.x + q.z) * kq; }

struct Acc {
  float fx = 0.f, fy = 0.f;
  int32_t cnt = 0, qcnt = 0;
  __device__ __forceinline__ void mono(float cm, float r, float dx, float dy) {
    const float rs = __builtin_amdgcn_rsqf(r);
    const float f = cm * rs * rs * rs;
    fx = __builtin_fmaf(f, dx, fx);
    fy = __builtin_fmaf(f, dy, fy);
    ++cnt;
  }
  // Monopole plus quadrupole of an accepted cell (Q the traceless in-plane
  // moments, tr Q = Qxx + Qyy = sum m |s|^2):
  //   a = m d/r^3 - Q.d/r^5 + (5/2)(d.Q.d - tr Q eps2) d/r^7
  //     = rs^5 [ (m r^2 + (2.5 (d.Q.d) + T) rs^2) d - Q.d ],
  //   rs = 1/r, r^2 = |d|^2 + eps2, T = -2.5 eps2 tr Q
  // (20 issue slots with the rsqrt instead of 27 for the terms apart).
  // c25 = 2.5 held in a VGPR (see walk64_group), so that the multiply-add
  // 2.5 s + T takes T from a scalar register without a copy
  __device__ __forceinline__ void quad(float cm, float r, float dx, float dy,
                                       float4 q, float tq, float c25) {
    const float rs = __builtin_amdgcn_rsqf(r);
    const float u = rs * rs;
    const float rs5 = (u * u) * rs;
    const float px = __builtin_fmaf(q.x, dx, q.y * dy);  // Q.d
    const float py = __builtin_fmaf(q.y, dx, q.z * dy);
    const float s = __builtin_fmaf(dx, px, dy * py);      // d.Q.d
    const float c = __builtin_fmaf(__builtin_fmaf(s, c25, tq), u, cm * r);
    fx = __builtin_fmaf(rs5, __builtin_fmaf(c, dx, -px), fx);
    fy = __builtin_fmaf(rs5, __builtin_fmaf(c, dy, -py), fy);
    ++cnt;
    ++qcnt;
  }
  // Either of the two above, chosen by a wavefront-uniform flag, with the
  // accumulating multiply-adds shared after the branch: both sides produce
  // (a, bx, by) and add a*bx, a*by, so the results stay in place instead of
  // being copied out of one side (the same arithmetic as mono / quad).
  __device__ __forceinline__ void term(bool with_quad, float cm, float r,
                                       float dx, float dy, float4 q, float tq, float c25) {
    const float rs = __builtin_amdgcn_rsqf(r);
    float a, bx, by;
    if (with_quad) {
      const float u = rs * rs;
      a = (u * u) * rs;
      const float px = __builtin_fmaf(q.x, dx, q.y * dy);
      const float py = __builtin_fmaf(q.y, dx, q.z * dy);
      const float s = __builtin_fmaf(dx, px, dy * py);
      const float c = __builtin_fmaf(__builtin_fmaf(s, c25, tq), u, cm * r);
      bx = __builtin_fmaf(c, dx, -px);
      by = __builtin_fmaf(c, dy, -py);
      ++qcnt;
    } else {
      a = cm * rs * rs * rs;
      bx = dx;
      by = dy;
    }
    fx = __builtin_fmaf(a, bx, fx);
    fy = __builtin_fmaf(a, by, fy);
    ++cnt;
  }
};

// ------------------------------------------------ mixed-precision far field --
// STELLAR_PRECISION_MIXED: an accepted cell's term m d/r^3 - Q.d/r^5 +
// (5/2)(d.Q.d - eps2 tr Q) d/r^7 is written as k B with k = m rs^2 (FP32)
// and the dimensionless bracket
//   B = e + t (-Qh.e + (5/2)(e.Qh.e - tr(Qh) eu) e),  e = d rs, t = rs^2 ws^2,
// eu = eps2 rs^2 (the softening term, tr(Qh) = Qxx + Qyy in FP16),
// where rs = 1/sqrt(|d|^2 + eps2), ws = S / 2^L is the cell's size at level L
// (S the larger extent of the instance's box) and Qh = Q / (m ws^2) is stored
// by the centre-of-mass pass as three FP16 values.  |e| <= 1 and t, |Qh| are
// O(1), so B is evaluated in packed FP16 (v_pk_fma_f16), two accepted cells
// per instruction: a cell is held until the next one comes, and the last is
// paired with zeros.  The conversions to FP16 round toward zero (one
// v_cvt_pkrtz_f16_f32 per pair).  k B enters an FP32 sum with Kahan
// compensation; bodies (the near field) stay FP32 monopoles as in the plain
// walk, and the acceptance vote is the plain walk's, in FP32.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

struct MixedAcc {
  float sx = 0.f, sy = 0.f, cx = 0.f, cy = 0.f;  // compensated far-field sums
  bool pend = false;
  float pex = 0.f, pey = 0.f, pk = 0.f, pt = 0.f, peu = 0.f;  // the held cell (per lane)
  uint32_t pq0 = 0u, pq1 = 0u;                     // its Qh halves (uniform)

  __device__ __forceinline__ static void kahan(float& sum, float& comp, float v) {
    const float y = v - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
  }
  // q0 = Qxx | Qxy << 16, q1 = Qyy (FP16 bits)
  __device__ __forceinline__ void pair(float ex, float ey, float k, float t, float eu,
                                       uint32_t q0, uint32_t q1) {
    const half2_t EX = __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(pex, ex));
    const half2_t EY = __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(pey, ey));
    const half2_t T = __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(pt, t));
    const half2_t EU = __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(peu, eu));
    const half2_t QXX = __builtin_bit_cast(half2_t, (pq0 & 0xffffu) | (q0 << 16));
    const half2_t QXY = __builtin_bit_cast(half2_t, (pq0 >> 16) | (q0 & 0xffff0000u));
    const half2_t QYY = __builtin_bit_cast(half2_t, (pq1 & 0xffffu) | (q1 << 16));
    const half2_t QEX = __builtin_elementwise_fma(QXX, EX, QXY * EY);
    const half2_t QEY = __builtin_elementwise_fma(QXY, EX, QYY * EY);
    const half2_t S = __builtin_elementwise_fma(EX, QEX, EY * QEY);
    const half2_t TRH = QXX + QYY;
    const half2_t C = __builtin_elementwise_fma(-TRH, EU, S) * (half2_t)(_Float16)2.5f;
    const half2_t BX = __builtin_elementwise_fma(T, __builtin_elementwise_fma(C, EX, -QEX), EX);
    const half2_t BY = __builtin_elementwise_fma(T, __builtin_elementwise_fma(C, EY, -QEY), EY);
    kahan(sx, cx, __builtin_fmaf(pk, (float)BX.x, k * (float)BX.y));
    kahan(sy, cy, __builtin_fmaf(pk, (float)BY.x, k * (float)BY.y));
  }
  // One accepted cell (r = |d|^2 + eps2, ws2 = ws^2 of its level, q0/q1 its Qh).
  __device__ __forceinline__ void cell(float cm, float r, float dx, float dy, float ws2,
                                       float eps2, uint32_t q0, uint32_t q1) {
    const float rs = __builtin_amdgcn_rsqf(r);
    const float u = rs * rs;
    const float ex = dx * rs, ey = dy * rs, k = cm * u, t = u * ws2, eu = eps2 * u;
    if (pend) {
      pair(ex, ey, k, t, eu, q0, q1);
      pend = false;
    } else {
      pex = ex;
      pey = ey;
      pk = k;
      pt = t;
      peu = eu;
      pq0 = q0;
      pq1 = q1;
      pend = true;
    }
  }
  __device__ __forceinline__ void flush() {
    if (pend) pair(0.f, 0.f, 0.f, 0.f, 0.f, 0u, 0u);
    pend = false;
  }
};

// ---------------------------------------------------------------- G = 64 --
// Stack word: block index in the low 26 bits, (level - 1) in the top 6.
constexpr int kIdxBits = 26;
constexpr uint32_t kIdxMask = (1u << kIdxBits) - 1u;
// Words per wavefront.  A one-at-a-time walk never holds more than
// 1 + 3*kMaxLevel; popping KPOP at once can hold more, so pushes past the
// end are dropped and the walk stops with kErrStack.
constexpr int kStack64 = 1024;
// SGPRs per wavefront at which a SIMD still holds eight walk wavefronts on
// MI355X (measured: 81 SGPRs let only six in, whatever the compiler's
// occupancy estimate says).
constexpr int kWalkSgprs = 80;

// One walk group on one wavefront: the traversal of group g, its forces and
// interaction counts.  The tree arrays are separate const __restrict__
// arguments: that is what lets the compiler read them through the scalar
// data cache.
template <bool QUAD, bool THR, int KPOP, int STACK = kStack64, bool MIXED = false>
__device__ __forceinline__ void walk64_group(const ChildRec* __restrict__ blk,
                                             const float4* __restrict__ qblk,
                                             const WalkArgs& A, int64_t g,
                                             uint32_t* st, const WalkUpdate* up) {
  static_assert(!MIXED || (QUAD && KPOP == 1), "the mixed far field pairs cells in visiting order");
  const float eps2 = A.eps2;
  const int lane = threadIdx.x & 63;
  const int32_t s0 = __builtin_amdgcn_readfirstlane(A.gtab[g]);
  const int32_t cnt = __builtin_amdgcn_readfirstlane(A.gtab[g + 1]) - s0;
  const int32_t inst = __builtin_amdgcn_readfirstlane(A.inst_of[s0]);
  const int32_t s = s0 + lane;
  const bool valid = lane < cnt;
  const int32_t sl = valid ? s : s0 + cnt - 1;  // padding lanes repeat the last body
  const float px = A.x[sl];
  const float py = A.y[sl];

  // With per-cell thresholds (bmax) or mixed precision, a node's second block
  // carries each cell child's threshold; otherwise
  // thr(level) = threshold for the children of a node at `level`:
  // 0.25*depth(level), depth(1) = depth0, depth(l+1) = 0.25*depth(l)
  // (kernels.cu:394-395), one level per lane (read with v_readlane once per
  // visited node: every child of a node has the same level).
  constexpr bool kQ = QUAD || THR;  // the node's second block is needed
  constexpr bool kTable = !THR && !MIXED;  // thresholds from the per-lane table
  const float kq = -2.5f * eps2;            // quadrupole softening factor
  // 2.5 as an opaque vector value: a literal would make the compiler use
  // v_fmamk, which cannot read T from a scalar register (one extra copy per
  // accepted cell)
  float c25 = 2.5f;
  asm volatile("" : "+v"(c25));
  float thr = 0.f;
  if (kTable) {
    thr = walk_root_depth(A.ibbox[inst], A.theta_div);
    for (int i = 0; i <= lane; ++i) thr = 0.25f * thr;
    // textbook test on the unsoftened distance: dp + eps2 <= d^2 + eps2
    if (A.unsoft) thr = thr + eps2;
  }

  int sp = 0;
  Acc acc;
  MixedAcc mix;
  // mixed: the square of the larger box extent (a cell at level L has size
  // S / 2^L)
  float s2 = 0.f;
  if (MIXED) {
    const float4 bb = A.ibbox[inst];
    const float w = bb.y - bb.x, h = bb.w - bb.z;
    const float sz = w > h ? w : h;
    s2 = sz * sz;
  }

  const ChildRec* root = blk + 4 * (int64_t)inst;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const ChildRec c = root[k];
    if (c.code == 0) continue;
    if (c.code < 0) {
      if (!kQ && A.ref_root) {
        // the reference pushes a root slot's lone body like a cell
        // (kernels.cu:373-387) and then reads that body's child slots
        // (:397-398): empty for every body but body 0, whose slots are the
        // root's own (child[0..3]), so the root's children are visited once
        // more with the next level's threshold; any other lone body exerts no
        // force
        const int32_t b = __builtin_amdgcn_readfirstlane(A.id[-1 - c.code]);
        if (b == __builtin_amdgcn_readfirstlane(A.ioff[inst])) {
          if (lane == 0) st[sp] = (uint32_t)inst;  // the root block, level field 0
          ++sp;
        }
        continue;
      }
      const float dx = c.x - px, dy = c.y - py;
      acc.mono(c.m, dist2(dx, dy, eps2), dx, dy);
    } else {
      if (lane == 0) st[sp] = (uint32_t)c.code;  // level 1
      ++sp;
    }
  }

  while (sp > 0) {
    const int take = sp < KPOP ? sp : KPOP;
    sp -= take;
    uint32_t node[KPOP];
    float dp[KPOP];
    uint32_t lvl[KPOP];
#pragma unroll
    for (int j = 0; j < KPOP; ++j) {
      // entries beyond `take` repeat the first one: a harmless extra load
      const uint32_t w =
          __builtin_amdgcn_readfirstlane(st[sp + (j < take ? j : 0)]);
      node[j] = w & kIdxMask;
      lvl[j] = w >> kIdxBits;  // level - 1
      dp[j] = !kTable ? 0.f
                      : __int_as_float(__builtin_amdgcn_readlane(__float_as_int(thr), (int)lvl[j]));
    }
    int4 q[KPOP][4];
    float4 qm[kQ ? KPOP : 1][4];
#pragma unroll
    for (int j = 0; j < KPOP; ++j) {
      const int4* Bq = reinterpret_cast<const int4*>(blk + 4 * (int64_t)node[j]);
#pragma unroll
      for (int k = 0; k < 4; ++k) q[j][k] = Bq[k];
      if (kQ) {
#pragma unroll
        for (int k = 0; k < 4; ++k) qm[kQ ? j : 0][k] = qblk[4 * (int64_t)node[j] + k];
      }
    }
    if (kQ) {
      // keep the second blocks' loads here, batched with the first blocks,
      // instead of letting them sink to each use (one wait per child)
#pragma unroll
      for (int j = 0; j < KPOP; ++j)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float4 v = qm[kQ ? j : 0][k];
          asm volatile("" ::"s"(v.x), "s"(v.y), "s"(v.z), "s"(v.w));
        }
    }
#pragma unroll
    for (int jj = KPOP - 1; jj >= 0; --jj) {
      if (jj >= take) continue;
      const uint32_t child_lvl = (lvl[jj] + 1u) << kIdxBits;
      // mixed: ws^2 of the children (level lvl + 2)
      const float ws2 = MIXED ? __builtin_ldexpf(s2, -2 * (int)(lvl[jj] + 2u)) : 0.f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int32_t code = q[jj][k].w;
        const float dx = __int_as_float(q[jj][k].x) - px;
        const float dy = __int_as_float(q[jj][k].y) - py;
        const float cmv = __int_as_float(q[jj][k].z);
        // a body is always accepted (a scalar branch on its code, no vote;
        // its slot in the second block is not written); code 0 is empty
        if (code <= 0) {
          if (code < 0) acc.mono(cmv, dist2(dx, dy, eps2), dx, dy);
          continue;
        }
        const float r = dist2(dx, dy, eps2);
        // a cell: the group's vote against its threshold
        const int32_t ti = kTable ? __float_as_int(dp[jj]) : __float_as_int(qm[kQ ? jj : 0][k].w);
        const float t = __int_as_float(__builtin_amdgcn_readfirstlane(ti));
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(!(t <= r)) == 0, 1)) {
          if (MIXED) {
            ++acc.cnt;
            ++acc.qcnt;
            const float4 q = qm[kQ ? jj : 0][k];
            mix.cell(cmv, r, dx, dy, ws2, eps2, __float_as_uint(q.x), __float_as_uint(q.y));
          } else {
            const float4 qc = qm[kQ ? jj : 0][k];
            acc.term(QUAD, cmv, r, dx, dy, qc, THR ? quad_soft_term(qc, kq) : qc.w, c25);
          }
        } else {
          if (lane == 0 && sp < STACK) st[sp] = (uint32_t)code | child_lvl;
          ++sp;
        }
      }
    }
    if (sp > STACK) {
      if (lane == 0) atomicOr(&A.dev->err, kErrStack);
      break;
    }
  }
  if (MIXED) {
    mix.flush();
    acc.fx = acc.fx + mix.sx;
    acc.fy = acc.fy + mix.sy;
  }
  if (valid) {
    A.ax[s] = acc.fx;
    A.ay[s] = acc.fy;
    A.inter[s] = acc.cnt;
  }
  // the fused update (incomplete forces, sp > 0 after a stack overflow,
  // leave the state as it was); only walks that pop one node at a time are
  // given it
  if (KPOP == 1 && sp == 0 && __builtin_amdgcn_readfirstlane(up->on) && valid) {
    const WalkUpdate U = *up;
    const float u = __builtin_fmaf(U.dt, acc.fx, U.vx[s]);
    const float v = __builtin_fmaf(U.dt, acc.fy, U.vy[s]);
    U.vx[s] = u;
    U.vy[s] = v;
    U.x[s] = __builtin_fmaf(U.ddt, u, px);
    U.y[s] = __builtin_fmaf(U.ddt, v, py);
  }
  if (lane == 0) {
    atomicAdd(&A.dev->interactions, (unsigned long long)acc.cnt * (unsigned long long)cnt);
    if (QUAD)
      atomicAdd(&A.dev->quad_interactions,
                (unsigned long long)acc.qcnt * (unsigned long long)cnt);
  }
}

// One wavefront per walk group, the WAVES wavefronts of a workgroup on
// consecutive groups: they start together and visit mostly the same nodes
// at about the same time, which the CU's scalar cache serves once.
// Workgroups are mapped so that each XCD walks one contiguous range of
// groups (xcd_block), which its L2 serves.
template <bool QUAD, bool THR, int KPOP, int WAVES, bool MIXED = false>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_num_sgpr(kWalkSgprs))) void walk64_kernel(
    const ChildRec* __restrict__ blk, const float4* __restrict__ qblk,
    WalkArgs A) {
  __shared__ uint32_t stk[WAVES][kStack64];
  __shared__ WalkUpdate upd[WAVES];
  const int wave = threadIdx.x >> 6;
  const int64_t g = (int64_t)xcd_block(blockIdx.x, gridDim.x) * WAVES + wave;
  if (g >= A.ngroups) return;
  if (A.dev->err & kBuildErrors) return;  // the tree of this step is incomplete
  stage_update(A, &upd[wave]);
  walk64_group<QUAD, THR, KPOP, kStack64, MIXED>(blk, qblk, A, g, stk[wave], &upd[wave]);
}

// Persistent form: exactly the resident workgroups, each taking work from
// per-XCD queues.  XCD x owns a contiguous range of units (as the one-shot
// launch gives it a contiguous range of workgroups); a workgroup first drains
// its own XCD's queue, then the other XCDs' in turn, so an XCD that finishes
// early takes the last units of the slower ones.
//   CLAIM_WAVE = false: a unit is WAVES consecutive groups, claimed by the
//     workgroup (its wavefronts start each unit together and share node
//     visits through the scalar cache);
//   CLAIM_WAVE = true: every wavefront claims one group at a time, so no
//     wavefront waits for the others of its workgroup.
// The queues (dev->wq) are zeroed before each walk; claims past a range's
// end are harmless, and every workgroup stops after one failed claim on each
// of the eight queues.  Which workgroup sits on which XCD follows the
// round-robin dispatch (workgroup b on XCD b % 8).
// The claim word takes the last stack entry of the last wavefront, so that the
// workgroup's LDS stays at 4 KiB per wavefront (four 512-thread workgroups fit
// a CU).
constexpr int kStackQ = kStack64 - 1;
template <bool QUAD, bool THR, int KPOP, int WAVES, bool CLAIM_WAVE, bool MIXED = false>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_num_sgpr(kWalkSgprs))) void walk64_queue_kernel(
    const ChildRec* __restrict__ blk, const float4* __restrict__ qblk, WalkArgs A) {
  __shared__ uint32_t stk[WAVES * kStack64];
  __shared__ WalkUpdate upd[WAVES];
  uint32_t& claim = stk[WAVES * kStack64 - 1];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  if (A.dev->err & kBuildErrors) return;  // the tree of this step is incomplete
  stage_update(A, &upd[wave]);
  const int64_t units = CLAIM_WAVE ? A.ngroups : (A.ngroups + WAVES - 1) / WAVES;
  const int x0 = (int)(blockIdx.x % kXcds);
  int k = 0;
  while (k < kXcds) {
    const int x = (x0 + k) % kXcds;
    const int64_t lo = units * x / kXcds, hi = units * (x + 1) / kXcds;
    int64_t u;
    if (CLAIM_WAVE) {
      uint32_t v = 0;
      if (lane == 0) v = atomicAdd(&A.dev->wq[x], 1u);
      u = lo + (int64_t)__builtin_amdgcn_readfirstlane(v);
    } else {
      if (threadIdx.x == 0) claim = atomicAdd(&A.dev->wq[x], 1u);
      __syncthreads();
      u = lo + (int64_t)claim;
      __syncthreads();  // every wavefront has read the claim before the next one
    }
    if (u >= hi) {
      ++k;
      continue;
    }
    const int64_t g = CLAIM_WAVE ? u : u * WAVES + wave;
    if (g < A.ngroups)
      walk64_group<QUAD, THR, KPOP, kStackQ, MIXED>(blk, qblk, A, g, stk + wave * kStack64,
                                                    &upd[wave]);
  }
}

// ----------------------------------------------------------- G = 32, 16 --
template <int G, bool QUAD, bool THR>
__global__ __launch_bounds__(kWalkBlock) void walk_kernel(
    const ChildRec* __restrict__ blk, const float4* __restrict__ qblk,
    WalkArgs A) {
  constexpr int kGroupsPerWave = 64 / G;
  constexpr int kWaves = kWalkBlock / 64;
  __shared__ int32_t stk_node[kWaves * kGroupsPerWave][kWalkStack];
  __shared__ float stk_dep[kWaves * kGroupsPerWave][kWalkStack];

  const float eps2 = A.eps2;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int sub = lane / G;             // group within the wavefront
  const int glane = lane & (G - 1);     // lane within the group
  const int slot = wave * kGroupsPerWave + sub;
  const int64_t g0 =
      ((int64_t)xcd_block(blockIdx.x, gridDim.x) * kWaves + wave) * kGroupsPerWave;
  if (g0 >= A.ngroups) return;  // the whole wavefront
  if (A.dev->err & kBuildErrors) return;  // the tree of this step is incomplete
  const int64_t g = g0 + sub;
  const bool gvalid = g < A.ngroups;
  const int32_t s0 = gvalid ? A.gtab[g] : 0;
  const int32_t cnt = gvalid ? A.gtab[g + 1] - s0 : 0;
  const int32_t s = s0 + glane;
  const bool valid = glane < cnt;
  const float px = valid ? A.x[s] : 0.f;
  const float py = valid ? A.y[s] : 0.f;
  const int32_t inst = gvalid ? A.inst_of[s0] : 0;
  const float depth0 = walk_root_depth(A.ibbox[inst], A.theta_div);
  const uint64_t gmask = ((G >= 64) ? ~0ull : ((1ull << (G & 63)) - 1ull)) << (sub * G);

  Acc acc;
  int sp = 0;
  bool overflow = false;

  // root block: level-1 cells are always opened (kernels.cu:370-388)
  const ChildRec* root = blk + 4 * (int64_t)inst;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!gvalid) break;
    const ChildRec c = root[k];
    if (c.code == 0) continue;
    if (c.code < 0) {
      if (!THR && A.ref_root) {
        // the reference's handling of a lone root body (see walk64_group):
        // body 0 revisits the root's children, any other exerts no force
        if (A.id[-1 - c.code] == A.ioff[inst]) {
          if (glane == 0) {
            stk_node[slot][sp] = inst;
            stk_dep[slot][sp] = depth0;
          }
          ++sp;
        }
        continue;
      }
      const float dx = c.x - px, dy = c.y - py;
      acc.mono(c.m, dist2(dx, dy, eps2), dx, dy);
    } else {
      if (glane == 0) {
        stk_node[slot][sp] = c.code;
        stk_dep[slot][sp] = depth0;
      }
      ++sp;
    }
  }

  while (sp > 0) {
    --sp;
    const int32_t node = stk_node[slot][sp];
    const float dp = 0.25f * stk_dep[slot][sp];
    const int4* B = reinterpret_cast<const int4*>(blk + 4 * (int64_t)node);
    int4 q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = B[k];
    float4 qm[4];
    if (QUAD || THR) {
#pragma unroll
      for (int k = 0; k < 4; ++k) qm[k] = qblk[4 * (int64_t)node + k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int32_t code = q[k].w;
      if (code == 0) continue;
      const float dx = __int_as_float(q[k].x) - px;
      const float dy = __int_as_float(q[k].y) - py;
      const float cmv = __int_as_float(q[k].z);
      const float r = dist2(dx, dy, eps2);
      if (code < 0) {
        acc.mono(cmv, r, dx, dy);
        continue;
      }
      const float th = THR ? qm[k].w : (A.unsoft ? dp + eps2 : dp);
      const uint64_t bad = __ballot(valid && !(th <= r)) & gmask;
      if (bad == 0) {
        if (QUAD)
          acc.quad(cmv, r, dx, dy, qm[k], THR ? quad_soft_term(qm[k], -2.5f * eps2) : qm[k].w,
                   2.5f);
        else
          acc.mono(cmv, r, dx, dy);
      } else {
        if (sp >= kWalkStack) {
          overflow = true;
          break;
        }
        if (glane == 0) {
          stk_node[slot][sp] = code;
          stk_dep[slot][sp] = dp;
        }
        ++sp;
      }
    }
    if (overflow) break;
  }
  if (overflow && glane == 0) atomicOr(&A.dev->err, kErrStack);
  if (valid) {
    A.ax[s] = acc.fx;
    A.ay[s] = acc.fy;
    A.inter[s] = acc.cnt;
  }
  if (glane == 0 && gvalid) {
    atomicAdd(&A.dev->interactions, (unsigned long long)acc.cnt * (unsigned long long)cnt);
    if (QUAD)
      atomicAdd(&A.dev->quad_interactions,
                (unsigned long long)acc.qcnt * (unsigned long long)cnt);
  }
}

}  // namespace

// Workgroups of a kernel that are resident at once on the whole device.
int resident_blocks(const void* kernel, int threads) {
  static std::mutex mu;
  static std::map<const void*, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(kernel);
  if (it != cache.end()) return it->second;
  int dev = 0, per_cu = 0, cus = 0;
  (void)hipGetDevice(&dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    cus = kXcds;
  const int nb = per_cu * cus;
  cache[kernel] = nb;
  return nb;
}

namespace {

template <bool QUAD, bool THR, int KPOP, int WAVES, bool CLAIM_WAVE, bool MIXED>
void launch64_queue(stellar_ctx* c, int64_t ngroups, const WalkArgs& A) {
  auto k = walk64_queue_kernel<QUAD, THR, KPOP, WAVES, CLAIM_WAVE, MIXED>;
  const int64_t units = CLAIM_WAVE ? ngroups : (ngroups + WAVES - 1) / WAVES;
  int64_t grid = resident_blocks(reinterpret_cast<const void*>(k), 64 * WAVES);
  if (grid > units) grid = units;
  k<<<(unsigned)grid, 64 * WAVES, 0, c->stream>>>(A.blk, A.qblk, A);
}

template <bool QUAD, bool THR, int KPOP, bool MIXED = false>
void launch64(stellar_ctx* c, int64_t ngroups, const WalkArgs& A) {
  // default shape (measured on MI355X, N = 2^20): 8 wavefronts per workgroup
  const int w = c->walk_waves > 0 ? c->walk_waves : 8;
  if (c->walk_queue != 0 && (w == 4 || w == 8)) {
    const bool per_wave = c->walk_queue == 2;
    if (w == 8) {
      if (per_wave)
        launch64_queue<QUAD, THR, KPOP, 8, true, MIXED>(c, ngroups, A);
      else
        launch64_queue<QUAD, THR, KPOP, 8, false, MIXED>(c, ngroups, A);
    } else {
      if (per_wave)
        launch64_queue<QUAD, THR, KPOP, 4, true, MIXED>(c, ngroups, A);
      else
        launch64_queue<QUAD, THR, KPOP, 4, false, MIXED>(c, ngroups, A);
    }
    return;
  }
  const unsigned grid = (unsigned)((ngroups + w - 1) / w);
  switch (w) {
    case 1:
      walk64_kernel<QUAD, THR, KPOP, 1, MIXED><<<grid, 64, 0, c->stream>>>(A.blk, A.qblk, A);
      break;
    case 2:
      walk64_kernel<QUAD, THR, KPOP, 2, MIXED><<<grid, 128, 0, c->stream>>>(A.blk, A.qblk, A);
      break;
    case 8:
      walk64_kernel<QUAD, THR, KPOP, 8, MIXED><<<grid, 512, 0, c->stream>>>(A.blk, A.qblk, A);
      break;
    case 16:
      walk64_kernel<QUAD, THR, KPOP, 16, MIXED><<<grid, 1024, 0, c->stream>>>(A.blk, A.qblk, A);
      break;
    default:
      walk64_kernel<QUAD, THR, KPOP, 4, MIXED><<<grid, 256, 0, c->stream>>>(A.blk, A.qblk, A);
      break;
  }
}

// Nodes per pop: default 2 when a node is one 64-byte block, 1 when it also
// brings its second block (quadrupoles or per-cell thresholds: two nodes'
// four blocks exceed the scalar registers and the loads serialise); at most
// 2 in that case.
template <bool QUAD, bool THR>
void launch64_kpop(stellar_ctx* c, int64_t ngroups, const WalkArgs& A) {
  const int kp = c->walk_kpop > 0 ? c->walk_kpop : ((QUAD || THR) ? 1 : 2);
  if (kp == 1)
    launch64<QUAD, THR, 1>(c, ngroups, A);
  else if (kp == 2 || QUAD || THR)
    launch64<QUAD, THR, 2>(c, ngroups, A);
  else
    launch64<QUAD, THR, 4>(c, ngroups, A);
}

template <int G, bool QUAD, bool THR>
void launch_shape(stellar_ctx* c, int64_t ngroups, const WalkArgs& A) {
  if (G == 64) {
    launch64_kpop<QUAD, THR>(c, ngroups, A);
    return;
  }
  constexpr int64_t per_wave = 64 / G;
  const int64_t waves = (ngroups + per_wave - 1) / per_wave;
  const unsigned grid = (unsigned)((waves + (kWalkBlock / 64) - 1) / (kWalkBlock / 64));
  walk_kernel<G, QUAD, THR><<<grid, kWalkBlock, 0, c->stream>>>(A.blk, A.qblk, A);
}

template <int G>
void launch_g(stellar_ctx* c, int64_t ngroups, const WalkArgs& A) {
  const bool quad = (c->p.flags & STELLAR_FLAG_QUADRUPOLE) != 0;
  const bool thr = (c->p.flags & STELLAR_FLAG_BMAX_OPENING) != 0;
  if (G == 64 && c->p.precision == STELLAR_PRECISION_MIXED) {
    // quadrupoles are required (params_ok); one node per pop keeps the
    // pairing of accepted cells in visiting order
    if (thr)
      launch64<true, true, 1, true>(c, ngroups, A);
    else
      launch64<true, false, 1, true>(c, ngroups, A);
    return;
  }
  if (quad && thr)
    launch_shape<G, true, true>(c, ngroups, A);
  else if (quad)
    launch_shape<G, true, false>(c, ngroups, A);
  else if (thr)
    launch_shape<G, false, true>(c, ngroups, A);
  else
    launch_shape<G, false, false>(c, ngroups, A);
}

}  // namespace

bool launch_forces(stellar_ctx* c, bool zero_counter, bool with_update) {
  const int32_t n = c->n;
  if (n == 0) return false;
  Buffers& B = c->b;
  const int gi = c->p.group == 32 ? 1 : c->p.group == 16 ? 2 : 0;
  const int64_t ngroups = c->ngroups[gi];
  const float theta = c->p.theta;
  const float tdiv = (c->p.flags & STELLAR_FLAG_TEXTBOOK_OPENING) ? theta * theta : theta;
  // the update is fused into G = 64 walks that pop one node at a time: their
  // stack cannot overflow (at most 1 + 3 per level), so no group's forces
  // can be incomplete once another group has updated its bodies
  const bool quad_or_thr =
      (c->p.flags & (STELLAR_FLAG_QUADRUPOLE | STELLAR_FLAG_BMAX_OPENING)) != 0;
  const int kpop = c->walk_kpop > 0 ? c->walk_kpop : (quad_or_thr ? 1 : 2);
  const bool fuse = with_update && gi == 0 &&
                    (c->p.precision == STELLAR_PRECISION_MIXED || kpop == 1);
  WalkArgs A{B.x,       B.y,       B.blk,       B.qblk,   n,
             c->p.eps2, tdiv,      B.dev,       B.ax,     B.ay,
             B.inter,   B.gtab[gi], ngroups,    B.inst_of, B.ibbox,
             (c->p.flags & STELLAR_FLAG_TEXTBOOK_OPENING) ? 1 : 0,
             (c->p.flags & STELLAR_FLAG_REFERENCE_ROOT) ? 1 : 0, B.id, B.ioff,
             fuse ? 1 : 0, c->p.timestep, c->p.dampening * c->p.timestep,
             B.x, B.y, B.vx, B.vy};
  if (zero_counter) launch_zero_counter(c);
  switch (c->p.group) {
    case 32:
      launch_g<32>(c, ngroups, A);
      break;
    case 16:
      launch_g<16>(c, ngroups, A);
      break;
    default:
      launch_g<64>(c, ngroups, A);
      break;
  }
  return fuse;
}

}  // namespace stellar
