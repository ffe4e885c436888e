// Bottom-up centre of mass for gfx950.
//
// The reference accumulates sum(m*x), sum(m*y), sum(m) into every cell on a
// body's insertion path with FP32 atomics (kernels.cu:180-183, 221-250) and
// divides afterwards (centre_of_mass_kernel, kernels.cu:275-288); the order
// of the atomics, and so the rounding, changes from run to run, and every
// level-1 cell takes about N/4 atomics.
//
// Here each cell is summed once from its four child records, in slot order,
// in FP64, and the FP32 centre of mass is the FP64 quotient rounded once.
// Work starts at cells whose children are all bodies; a per-cell arrival
// counter lets the last-arriving child carry the sum up to its parent, so no
// thread ever waits and the result is independent of scheduling.
#include "engine.h"

namespace stellar {

namespace {

__global__ __launch_bounds__(256) void com_kernel(
    ChildRec* __restrict__ blk, const int32_t* __restrict__ cparent,
    const uint8_t* __restrict__ cslot, const uint8_t* __restrict__ ckc,
    int32_t* __restrict__ arrive, double* __restrict__ csum,
    float4* __restrict__ ccom, const DevScalars* __restrict__ dev) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t ncells = dev->ncells;
  if (c >= ncells) return;
  int32_t cur = c + 1;  // block index of cell c
  if (ckc[cur] != 0) return;
  while (true) {
    const ChildRec* rec = blk + 4 * (int64_t)cur;
    double sm = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const ChildRec r = rec[k];
      if (r.code == 0) continue;
      if (r.code < 0) {
        const double m = (double)r.m;
        sm += m;
        sx += m * (double)r.x;
        sy += m * (double)r.y;
      } else {
        const double* q = csum + 3 * (int64_t)r.code;
        sm += q[0];
        sx += q[1];
        sy += q[2];
      }
    }
    double* out = csum + 3 * (int64_t)cur;
    out[0] = sm;
    out[1] = sx;
    out[2] = sy;
    const float cm = (float)sm;
    const float cx = sm != 0.0 ? (float)(sx / sm) : 0.f;
    const float cy = sm != 0.0 ? (float)(sy / sm) : 0.f;
    ccom[cur] = make_float4(cx, cy, cm, 0.f);
    const int32_t p = cparent[cur];
    ChildRec* slot = blk + 4 * (int64_t)p + cslot[cur];
    slot->x = cx;
    slot->y = cy;
    slot->m = cm;
    if (p == 0) break;  // the root needs no centre of mass
    __threadfence();
    const int32_t old = atomicAdd(&arrive[p], 1);
    if (old != (int32_t)ckc[p] - 1) break;
    __threadfence();
    cur = p;
  }
}

}  // namespace

void launch_com(stellar_ctx* c) {
  if (c->n == 0 || c->cap == 0) return;
  Buffers& B = c->b;
  const int64_t cells = c->cap;
  const unsigned grid = (unsigned)((cells + 255) / 256);
  com_kernel<<<grid, 256, 0, c->stream>>>(B.blk, B.cparent, B.cslot, B.ckc,
                                          B.arrive, B.csum, B.ccom, B.dev);
}

}  // namespace stellar
