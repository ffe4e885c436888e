// Bottom-up centre of mass (and optional quadrupole moments) for gfx950.
//
// The reference accumulates sum(m*x), sum(m*y), sum(m) into every cell on a
// body's insertion path with FP32 atomics (kernels.cu:180-183, 221-250) and
// divides afterwards (centre_of_mass_kernel, kernels.cu:275-288); the order
// of the atomics, and so the rounding, changes from run to run, and every
// level-1 cell takes about N/4 atomics.
//
// Here each cell is summed once from its four child records, in slot order,
// in FP64, and the FP32 centre of mass is the FP64 quotient rounded once, so
// the result does not depend on scheduling.
//
// Scheduling without device-wide fences: on gfx950 a device-scope release
// writes back the issuing XCD's whole L2, so a per-cell fence costs
// milliseconds over a million cells.  Cells are numbered in DFS pre-order, so
// every subtree is a contiguous id range [c, cend[c]).  Pass 1 gives each
// workgroup a range of kRange cell ids and finishes every cell whose subtree
// lies inside that range, with LDS arrival counters and workgroup-scope
// ordering only.  The cells left over are ancestors of a range boundary
// (at most one root-to-boundary path per range); pass 2 finishes them in a
// single workgroup, again with workgroup-scope ordering only.  Results of
// pass 1 reach pass 2 through the kernel boundary.
#include "engine.h"

namespace stellar {

namespace {

constexpr int kRange = 4096;      // cells per pass-1 workgroup
constexpr int kComThreads = 1024;
constexpr int kSpineThreads = 1024;
// A range's leftover cells are ancestors of the first cell after the range,
// so there are at most kMaxLevel of them: each range owns kMaxLevel slots.
constexpr int kSpineSlots = kMaxLevel;

struct ComArgs {
  ChildRec* blk;
  float4* qblk;
  const int32_t* cparent;
  const uint8_t* cslot;
  const uint8_t* ckc;
  const int32_t* cend;
  int32_t* arrive;
  double* csum;  // kSumStride per block
  float4* ccom;
  int32_t* spine;   // kSpineSlots per range
  int32_t* scount;  // leftover cells per range
  uint8_t* ready;
  DevScalars* dev;
  int quad;
};

// Moments of block b from its four child records, in slot order.  Writes the
// FP64 sums, the FP32 centre of mass and the parent's child record.
__device__ void cell_moments(const ComArgs& A, int32_t b) {
  const ChildRec* rec = A.blk + 4 * (int64_t)b;
  double cm[4], cx[4], cy[4];
  const double* qq[4];
  double sm = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const ChildRec r = rec[k];
    qq[k] = nullptr;
    cm[k] = cx[k] = cy[k] = 0.0;
    if (r.code == 0) continue;
    if (r.code < 0) {
      cm[k] = (double)r.m;
      cx[k] = (double)r.x;
      cy[k] = (double)r.y;
      sm += cm[k];
      sx += cm[k] * cx[k];
      sy += cm[k] * cy[k];
    } else {
      const double* q = A.csum + kSumStride * (int64_t)r.code;
      cm[k] = q[0];
      sm += q[0];
      sx += q[1];
      sy += q[2];
      if (A.quad) {
        cx[k] = q[0] != 0.0 ? q[1] / q[0] : 0.0;
        cy[k] = q[0] != 0.0 ? q[2] / q[0] : 0.0;
        qq[k] = q + 3;
      }
    }
  }
  double* out = A.csum + kSumStride * (int64_t)b;
  out[0] = sm;
  out[1] = sx;
  out[2] = sy;
  const float fm = (float)sm;
  const float fx = sm != 0.0 ? (float)(sx / sm) : 0.f;
  const float fy = sm != 0.0 ? (float)(sy / sm) : 0.f;
  A.ccom[b] = make_float4(fx, fy, fm, 0.f);
  const int32_t p = A.cparent[b];
  const int s = A.cslot[b];
  ChildRec* slot = A.blk + 4 * (int64_t)p + s;
  slot->x = fx;
  slot->y = fy;
  slot->m = fm;
  if (A.quad) {
    // traceless in-plane quadrupole about the FP64 centre of mass, with the
    // parallel-axis shift of each child (same operation order as the CPU
    // restatement, so the FP32 results agree bit for bit)
    const double ox = sm != 0.0 ? sx / sm : 0.0, oy = sm != 0.0 ? sy / sm : 0.0;
    double qxx = 0.0, qxy = 0.0, qyy = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (rec[k].code == 0) continue;
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
    out[3] = qxx;
    out[4] = qxy;
    out[5] = qyy;
    A.qblk[4 * (int64_t)p + s] = make_float4((float)qxx, (float)qxy, (float)qyy, 0.f);
  }
}

__global__ __launch_bounds__(kComThreads) void com_local_kernel(ComArgs A) {
  __shared__ int32_t larr[kRange];
  __shared__ int32_t lspine;
  const int32_t ncells = A.dev->ncells;
  const int32_t a = blockIdx.x * kRange;  // cell ids [a, hi)
  if (a >= ncells) return;
  if (A.dev->err & kBuildErrors) return;  // no complete tree to sum
  const int32_t hi = (a + kRange < ncells) ? a + kRange : ncells;
  for (int i = threadIdx.x; i < kRange; i += blockDim.x) larr[i] = 0;
  if (threadIdx.x == 0) lspine = 0;
  __syncthreads();
  int32_t* myspine = A.spine + (int64_t)blockIdx.x * kSpineSlots;
  for (int32_t c = a + threadIdx.x; c < hi; c += blockDim.x) {
    const int32_t b = c + 1;
    if (A.cend[b] > hi) {
      const int32_t k = atomicAdd(&lspine, 1);
      if (k < kSpineSlots) myspine[k] = b;
    }
    if (A.ckc[b] != 0) continue;
    int32_t cur = b;
    while (true) {
      cell_moments(A, cur);
      const int32_t p = A.cparent[cur];
      if (p == 0 || p >= cur) break;
      const int32_t pc = p - 1;
      if (pc < a || A.cend[p] > hi) {
        // parent finished in pass 2; the kernel boundary publishes our sums
        atomicAdd(&A.arrive[p], 1);
        break;
      }
      __threadfence_block();
      const int32_t old = atomicAdd(&larr[pc - a], 1);
      if (old != (int32_t)A.ckc[p] - 1) break;
      __threadfence_block();
      cur = p;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) A.scount[blockIdx.x] = lspine < kSpineSlots ? lspine : kSpineSlots;
}

__global__ __launch_bounds__(kSpineThreads) void com_spine_kernel(ComArgs A) {
  if (A.dev->err & kBuildErrors) return;
  const int32_t nranges = (A.dev->ncells + kRange - 1) / kRange;
  const int32_t nslots = nranges * kSpineSlots;
  // a spine cell is ready at the start iff all its cell children were
  // finished in pass 1; decide that before any pass-2 arrival happens
  for (int32_t i = threadIdx.x; i < nslots; i += blockDim.x) {
    const int32_t r = i / kSpineSlots, k = i % kSpineSlots;
    uint8_t rd = 0;
    if (k < A.scount[r]) {
      const int32_t b = A.spine[i];
      rd = A.arrive[b] == (int32_t)A.ckc[b] ? 1 : 0;
    }
    A.ready[i] = rd;
  }
  __syncthreads();
  for (int32_t i = threadIdx.x; i < nslots; i += blockDim.x) {
    if (!A.ready[i]) continue;
    int32_t cur = A.spine[i];
    while (true) {
      cell_moments(A, cur);
      const int32_t p = A.cparent[cur];
      if (p == 0 || p >= cur) break;
      __threadfence_block();
      const int32_t old = atomicAdd(&A.arrive[p], 1);
      if (old != (int32_t)A.ckc[p] - 1) break;
      __threadfence_block();
      cur = p;
    }
  }
}

}  // namespace

void launch_com(stellar_ctx* c) {
  if (c->n == 0 || c->cap == 0) return;
  Buffers& B = c->b;
  ComArgs A{B.blk,    B.qblk,  B.cparent, B.cslot,  B.ckc,
            B.cend,   B.arrive, B.csum,   B.ccom,   B.spine,
            B.scount, B.ready, B.dev,
            (c->p.flags & STELLAR_FLAG_QUADRUPOLE) ? 1 : 0};
  const unsigned grid = (unsigned)((c->cap + kRange - 1) / kRange);
  com_local_kernel<<<grid, kComThreads, 0, c->stream>>>(A);
  com_spine_kernel<<<1, kSpineThreads, 0, c->stream>>>(A);
}

}  // namespace stellar
