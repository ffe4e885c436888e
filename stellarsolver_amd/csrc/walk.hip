// Tree-walk force kernel for gfx950 (compute_forces_kernel, kernels.cu:340-435).
//
// A walk group is G consecutive bodies in tree order (the reference uses warps
// of 32).  The group shares one traversal: for each child of a popped node it
// evaluates r = dx*dx + dy*dy + eps2 per body and accepts the child when it is
// a body or when every body of the group passes dp <= r (the reference's
// __all vote); otherwise the child is pushed.  Accepted children add
// f = m*rsqrt(r)^3 * (dx, dy) to every body of the group.
//
// gfx950 mapping (G = 64, one wavefront per group):
//   * the node index is wavefront-uniform, so a node's four child records
//     (one 64-byte block) arrive through scalar loads;
//   * the vote is one 64-lane ballot;
//   * the traversal stack lives in LDS, one region per wavefront, written by
//     lane 0 and read back as a uniform value;
//   * bodies are stored in tree order, so the group's positions and the
//     outputs are contiguous 256-byte rows.
// G = 32 and 16 run 2 or 4 independent groups per wavefront (ballot masks
// per group), which reproduces the reference's warp-of-32 vote exactly.
//
// Traversal order matches the reference (children in slot order, accepted
// ones accumulated at once, opened ones pushed and popped LIFO), so the set
// of interactions and the summation order are those of the CPU restatement.
// One deliberate difference: a body alone in a root slot is accumulated; the
// reference pushes it like a cell and then finds no children (kernels.cu:
// 373-387, 397-398), so that body exerts no force.
//
// QUAD adds each accepted cell's quadrupole term (a second 64-byte block of
// Qxx, Qxy, Qyy per node); bodies stay monopoles.
#include "engine.h"

namespace stellar {

namespace {

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
};

template <int G, bool QUAD>
__global__ __launch_bounds__(kWalkBlock) void walk_kernel(WalkArgs A) {
  constexpr int kGroupsPerWave = 64 / G;
  constexpr int kWaves = kWalkBlock / 64;
  __shared__ int32_t stk_node[kWaves * kGroupsPerWave][kWalkStack];
  __shared__ float stk_dep[kWaves * kGroupsPerWave][kWalkStack];

  const ChildRec* __restrict__ blk = A.blk;
  const float4* __restrict__ qblk = A.qblk;
  const int32_t n = A.n;
  const float eps2 = A.eps2;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int sub = lane / G;             // group within the wavefront
  const int glane = lane & (G - 1);     // lane within the group
  const int slot = wave * kGroupsPerWave + sub;
  const int64_t s0 = ((int64_t)blockIdx.x * kWaves + wave) * 64 + sub * G;
  if (((int64_t)blockIdx.x * kWaves + wave) * 64 >= n) return;
  if (A.dev->err & kBuildErrors) return;  // the tree of this step is incomplete
  const int64_t s = s0 + glane;
  const bool valid = s < n;
  const float px = valid ? A.x[s] : 0.f;
  const float py = valid ? A.y[s] : 0.f;

  // depth = radius^2/theta with radius = half the x extent (kernels.cu:349,376)
  const float bl = A.dev->bbox[0], br = A.dev->bbox[1];
  const float radius = (float)(0.5 * (double)(br - bl));
  const float depth0 = (float)((double)(radius * radius) / (double)A.theta_div);

  const uint64_t gmask =
      G == 64 ? ~0ull : (((1ull << G) - 1ull) << (sub * G));

  float fx = 0.f, fy = 0.f;
  int32_t cnt = 0;
  int sp = 0;
  bool overflow = false;

  auto accumulate = [&](float cm, float r, float dx, float dy) {
    const float rs = __builtin_amdgcn_rsqf(r);
    const float f = cm * rs * rs * rs;
    fx = __builtin_fmaf(f, dx, fx);
    fy = __builtin_fmaf(f, dy, fy);
    ++cnt;
  };
  auto accumulate_q = [&](float cm, float r, float dx, float dy, float4 q) {
    const float rs = __builtin_amdgcn_rsqf(r);
    const float f = cm * rs * rs * rs;
    const float rs2 = rs * rs;
    const float rs5 = rs2 * rs2 * rs;
    const float rs7 = rs5 * rs2;
    const float qdx = q.x * dx + q.y * dy;
    const float qdy = q.y * dx + q.z * dy;
    const float dqd = dx * qdx + dy * qdy;
    const float g = f + 2.5f * dqd * rs7;
    fx = __builtin_fmaf(-rs5, qdx, __builtin_fmaf(g, dx, fx));
    fy = __builtin_fmaf(-rs5, qdy, __builtin_fmaf(g, dy, fy));
    ++cnt;
  };

  // root block: level-1 cells are always opened (kernels.cu:370-388)
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const ChildRec c = blk[k];
    if (c.code == 0) continue;
    if (c.code < 0) {
      const float dx = c.x - px, dy = c.y - py;
      const float r = dx * dx + dy * dy + eps2;
      accumulate(c.m, r, dx, dy);
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
    int32_t node = stk_node[slot][sp];
    float depth = stk_dep[slot][sp];
    if (G == 64) {
      node = __builtin_amdgcn_readfirstlane(node);
      depth = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(depth)));
    }
    const float dp = 0.25f * depth;
    // the whole 64-byte block first (scalar loads for G = 64), then the four
    // children from registers
    const int4* B = reinterpret_cast<const int4*>(blk + 4 * (int64_t)node);
    int4 q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = B[k];
    float4 qm[4];
    if (QUAD) {
#pragma unroll
      for (int k = 0; k < 4; ++k) qm[k] = qblk[4 * (int64_t)node + k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const ChildRec c{__int_as_float(q[k].x), __int_as_float(q[k].y),
                       __int_as_float(q[k].z), q[k].w};
      if (c.code == 0) continue;
      const float dx = c.x - px, dy = c.y - py;
      const float r = dx * dx + dy * dy + eps2;
      if (c.code < 0) {
        accumulate(c.m, r, dx, dy);
        continue;
      }
      const uint64_t bad = __ballot(valid && !(dp <= r)) & gmask;
      if (bad == 0) {
        if (QUAD)
          accumulate_q(c.m, r, dx, dy, qm[k]);
        else
          accumulate(c.m, r, dx, dy);
      } else {
        if (sp >= kWalkStack) {
          overflow = true;
          break;
        }
        if (glane == 0) {
          stk_node[slot][sp] = c.code;
          stk_dep[slot][sp] = dp;
        }
        ++sp;
      }
    }
    if (overflow) break;
  }
  if (overflow && glane == 0) atomicOr(&A.dev->err, kErrStack);
  if (valid) {
    A.ax[s] = fx;
    A.ay[s] = fy;
    A.inter[s] = cnt;
  }
  if (glane == 0) {
    const int64_t members = (n - s0) < G ? (n - s0) : G;
    atomicAdd(&A.dev->interactions,
              (unsigned long long)cnt * (unsigned long long)members);
  }
}

// G = 64 with several nodes in flight: each iteration pops up to KPOP stack
// entries and issues all their block loads before using any, so one memory
// round trip serves KPOP nodes.  The set of interactions does not depend on
// the order nodes are visited (a node is visited iff every ancestor was
// opened, and each decision depends only on the node and the group), so this
// walk makes exactly the decisions of walk_kernel<64>; only the order in
// which accepted terms are summed differs.
template <bool QUAD, int KPOP>
__global__ __launch_bounds__(kWalkBlock) void walk64_kernel(WalkArgs A) {
  constexpr int kWaves = kWalkBlock / 64;
  __shared__ int32_t stk_node[kWaves][kWalkStack];
  __shared__ float stk_dep[kWaves][kWalkStack];

  const ChildRec* __restrict__ blk = A.blk;
  const float4* __restrict__ qblk = A.qblk;
  const int32_t n = A.n;
  const float eps2 = A.eps2;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t s0 = ((int64_t)blockIdx.x * kWaves + wave) * 64;
  if (s0 >= n) return;
  if (A.dev->err & kBuildErrors) return;  // the tree of this step is incomplete
  const int64_t s = s0 + lane;
  const bool valid = s < n;
  const float px = valid ? A.x[s] : 0.f;
  const float py = valid ? A.y[s] : 0.f;
  const float bl = A.dev->bbox[0], br = A.dev->bbox[1];
  const float radius = (float)(0.5 * (double)(br - bl));
  const float depth0 = (float)((double)(radius * radius) / (double)A.theta_div);
  int32_t* snode = stk_node[wave];
  float* sdep = stk_dep[wave];

  float fx = 0.f, fy = 0.f;
  int32_t cnt = 0;
  int sp = 0;
  bool overflow = false;

  auto accumulate = [&](float cm, float r, float dx, float dy) {
    const float rs = __builtin_amdgcn_rsqf(r);
    const float f = cm * rs * rs * rs;
    fx = __builtin_fmaf(f, dx, fx);
    fy = __builtin_fmaf(f, dy, fy);
    ++cnt;
  };
  auto accumulate_q = [&](float cm, float r, float dx, float dy, float4 q) {
    const float rs = __builtin_amdgcn_rsqf(r);
    const float f = cm * rs * rs * rs;
    const float rs2 = rs * rs;
    const float rs5 = rs2 * rs2 * rs;
    const float rs7 = rs5 * rs2;
    const float qdx = q.x * dx + q.y * dy;
    const float qdy = q.y * dx + q.z * dy;
    const float dqd = dx * qdx + dy * qdy;
    const float g = f + 2.5f * dqd * rs7;
    fx = __builtin_fmaf(-rs5, qdx, __builtin_fmaf(g, dx, fx));
    fy = __builtin_fmaf(-rs5, qdy, __builtin_fmaf(g, dy, fy));
    ++cnt;
  };

#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const ChildRec c = blk[k];
    if (c.code == 0) continue;
    if (c.code < 0) {
      const float dx = c.x - px, dy = c.y - py;
      const float r = dx * dx + dy * dy + eps2;
      accumulate(c.m, r, dx, dy);
    } else {
      if (lane == 0) {
        snode[sp] = c.code;
        sdep[sp] = depth0;
      }
      ++sp;
    }
  }

  while (sp > 0) {
    const int take = sp < KPOP ? sp : KPOP;
    sp -= take;
    int32_t node[KPOP];
    float dp[KPOP];
#pragma unroll
    for (int j = 0; j < KPOP; ++j) {
      // entries beyond `take` repeat the first one: a harmless extra load
      const int e = sp + (j < take ? j : 0);
      node[j] = __builtin_amdgcn_readfirstlane(snode[e]);
      dp[j] = 0.25f * __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(sdep[e])));
    }
    int4 q[KPOP][4];
    float4 qm[QUAD ? KPOP : 1][4];
#pragma unroll
    for (int j = 0; j < KPOP; ++j) {
      const int4* Bq = reinterpret_cast<const int4*>(blk + 4 * (int64_t)node[j]);
#pragma unroll
      for (int k = 0; k < 4; ++k) q[j][k] = Bq[k];
      if (QUAD) {
#pragma unroll
        for (int k = 0; k < 4; ++k) qm[QUAD ? j : 0][k] = qblk[4 * (int64_t)node[j] + k];
      }
    }
#pragma unroll
    for (int jj = KPOP - 1; jj >= 0; --jj) {
      if (jj >= take) continue;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int32_t code = q[jj][k].w;
        if (code == 0) continue;
        const float cxv = __int_as_float(q[jj][k].x), cyv = __int_as_float(q[jj][k].y);
        const float cmv = __int_as_float(q[jj][k].z);
        const float dx = cxv - px, dy = cyv - py;
        const float r = dx * dx + dy * dy + eps2;
        if (code < 0) {
          accumulate(cmv, r, dx, dy);
          continue;
        }
        const uint64_t bad = __ballot(valid && !(dp[jj] <= r));
        if (bad == 0) {
          if (QUAD)
            accumulate_q(cmv, r, dx, dy, qm[QUAD ? jj : 0][k]);
          else
            accumulate(cmv, r, dx, dy);
        } else {
          if (sp >= kWalkStack) {
            overflow = true;
            break;
          }
          if (lane == 0) {
            snode[sp] = code;
            sdep[sp] = dp[jj];
          }
          ++sp;
        }
      }
      if (overflow) break;
    }
    if (overflow) break;
  }
  if (overflow && lane == 0) atomicOr(&A.dev->err, kErrStack);
  if (valid) {
    A.ax[s] = fx;
    A.ay[s] = fy;
    A.inter[s] = cnt;
  }
  if (lane == 0) {
    const int64_t members = (n - s0) < 64 ? (n - s0) : 64;
    atomicAdd(&A.dev->interactions,
              (unsigned long long)cnt * (unsigned long long)members);
  }
}

template <int G>
void launch_g(stellar_ctx* c, unsigned grid, const WalkArgs& A) {
  const bool quad = (c->p.flags & STELLAR_FLAG_QUADRUPOLE) != 0;
  if (G == 64) {
    if (quad)
      walk64_kernel<true, 2><<<grid, kWalkBlock, 0, c->stream>>>(A);
    else
      walk64_kernel<false, 4><<<grid, kWalkBlock, 0, c->stream>>>(A);
    return;
  }
  if (quad)
    walk_kernel<G, true><<<grid, kWalkBlock, 0, c->stream>>>(A);
  else
    walk_kernel<G, false><<<grid, kWalkBlock, 0, c->stream>>>(A);
}

}  // namespace

void launch_forces(stellar_ctx* c, bool zero_counter) {
  const int32_t n = c->n;
  if (n == 0) return;
  Buffers& B = c->b;
  const int64_t waves = (n + 63) / 64;
  const unsigned grid = (unsigned)((waves + (kWalkBlock / 64) - 1) / (kWalkBlock / 64));
  const float theta = c->p.theta;
  const float tdiv = (c->p.flags & STELLAR_FLAG_TEXTBOOK_OPENING) ? theta * theta : theta;
  WalkArgs A{B.x, B.y, B.blk, B.qblk, n, c->p.eps2, tdiv, B.dev, B.ax, B.ay, B.inter};
  if (zero_counter) launch_zero_counter(c);
  switch (c->p.group) {
    case 32:
      launch_g<32>(c, grid, A);
      break;
    case 16:
      launch_g<16>(c, grid, A);
      break;
    default:
      launch_g<64>(c, grid, A);
      break;
  }
}

}  // namespace stellar
