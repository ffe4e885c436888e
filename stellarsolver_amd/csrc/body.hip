// Per-body stages for gfx950: tree-order sort of the body state, the
// leapfrog/dampening update, the output buffer, and FP64 diagnostics.
#include "engine.h"

namespace stellar {

namespace {

inline unsigned grid_for(int64_t items, int threads) {
  return (unsigned)((items + threads - 1) / threads);
}


// update_kernel (kernels.cu:522-537): v += dt*a; x += d*dt*v, both as fused
// multiply-adds with d*dt rounded first, as nvcc contracts the reference.
// Skipped when the step failed (build errors, walk stack overflow): the
// accelerations are then incomplete, and the state stays as it was.
__global__ __launch_bounds__(256) void update_kernel(
    int32_t n, float* __restrict__ x, float* __restrict__ y,
    float* __restrict__ vx, float* __restrict__ vy,
    const float* __restrict__ ax, const float* __restrict__ ay, float dt,
    float d, const DevScalars* __restrict__ dev) {
  if (__builtin_amdgcn_readfirstlane(dev->err) & (kBuildErrors | kErrStack)) return;
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t n4 = n >> 2;
  const float ddt = d * dt;
  if (i < n4) {
    float4* x4 = reinterpret_cast<float4*>(x);
    float4* y4 = reinterpret_cast<float4*>(y);
    float4* vx4 = reinterpret_cast<float4*>(vx);
    float4* vy4 = reinterpret_cast<float4*>(vy);
    const float4 a = reinterpret_cast<const float4*>(ax)[i];
    const float4 b = reinterpret_cast<const float4*>(ay)[i];
    float4 u = vx4[i], v = vy4[i], p = x4[i], q = y4[i];
    u.x = __builtin_fmaf(dt, a.x, u.x);
    u.y = __builtin_fmaf(dt, a.y, u.y);
    u.z = __builtin_fmaf(dt, a.z, u.z);
    u.w = __builtin_fmaf(dt, a.w, u.w);
    v.x = __builtin_fmaf(dt, b.x, v.x);
    v.y = __builtin_fmaf(dt, b.y, v.y);
    v.z = __builtin_fmaf(dt, b.z, v.z);
    v.w = __builtin_fmaf(dt, b.w, v.w);
    p.x = __builtin_fmaf(ddt, u.x, p.x);
    p.y = __builtin_fmaf(ddt, u.y, p.y);
    p.z = __builtin_fmaf(ddt, u.z, p.z);
    p.w = __builtin_fmaf(ddt, u.w, p.w);
    q.x = __builtin_fmaf(ddt, v.x, q.x);
    q.y = __builtin_fmaf(ddt, v.y, q.y);
    q.z = __builtin_fmaf(ddt, v.z, q.z);
    q.w = __builtin_fmaf(ddt, v.w, q.w);
    vx4[i] = u;
    vy4[i] = v;
    x4[i] = p;
    y4[i] = q;
  }
  const int32_t t = 4 * n4 + i;
  if (i < 4 && t < n) {
    vx[t] = __builtin_fmaf(dt, ax[t], vx[t]);
    vy[t] = __builtin_fmaf(dt, ay[t], vy[t]);
    x[t] = __builtin_fmaf(ddt, vx[t], x[t]);
    y[t] = __builtin_fmaf(ddt, vy[t], y[t]);
  }
}

// copy_kernel (kernels.cu:541-553): interleaved positions in body-id order,
// followed by the cell centres of mass.
__global__ __launch_bounds__(256) void output_kernel(
    int32_t n, int64_t cap, int32_t nroot, const float* __restrict__ x,
    const float* __restrict__ y, const int32_t* __restrict__ id,
    const ChildRec* __restrict__ blk, const uint64_t* __restrict__ cmeta,
    const DevScalars* __restrict__ dev,
    float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int32_t b = id[i];
    reinterpret_cast<float2*>(out)[b] = make_float2(x[i], y[i]);
    return;
  }
  const int64_t c = i - n;
  if (c < dev->ncells && c < cap) {
    // a cell's centre of mass is its record in the parent's block
    const int64_t b = c + nroot;
    const uint64_t meta = cmeta[b];
    const ChildRec q = blk[4 * (int64_t)meta_parent(meta) + meta_slot(meta)];
    reinterpret_cast<float2*>(out)[n + c] = make_float2(q.x, q.y);
  }
}

// Headless form of the reference's point drawing (StellarSolverVisualizer.cpp:
// 208-253): every body adds one to the count of the pixel it falls in,
// column floor((x - left) * sx), row height-1-floor((y - bottom) * sy) (row 0
// at the top, as an image is stored), FP32 as written; bodies outside the
// view are clipped.  Bodies are in tree order, so neighbouring lanes mostly
// share a pixel: each run of equal pixels within a wavefront adds its length
// with one atomic from its first lane.
__global__ __launch_bounds__(256) void render_count_kernel(
    int32_t n, const float* __restrict__ x, const float* __restrict__ y, float left,
    float bottom, float sx, float sy, int32_t width, int32_t height,
    uint32_t* __restrict__ counts) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int32_t p = -1;
  if (i < n) {
    const float fx = (x[i] - left) * sx;
    const float fy = (y[i] - bottom) * sy;
    // also false for NaN
    if (fx >= 0.f && fx < (float)width && fy >= 0.f && fy < (float)height) {
      const int32_t col = min((int32_t)fx, width - 1);
      const int32_t row = height - 1 - min((int32_t)fy, height - 1);
      p = row * width + col;
    }
  }
  const int32_t prev = __shfl_up(p, 1);
  const uint64_t starts = __builtin_amdgcn_ballot_w64(lane == 0 || prev != p);
  if (p >= 0 && (lane == 0 || prev != p)) {
    // the run ends at the next start (or the wavefront's end)
    const uint64_t later = lane == 63 ? 0ull : starts >> (lane + 1);
    const int len = later ? __builtin_ctzll(later) + 1 : 64 - lane;
    atomicAdd(&counts[p], (uint32_t)len);
  }
}

__global__ __launch_bounds__(256) void fill_u32_kernel(int64_t len, uint32_t* __restrict__ a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) a[i] = 0u;
}

// The reference's blend, white at alpha 0.1 added per body over black and
// clamped (glBlendFunc(GL_SRC_ALPHA, GL_ONE), colour (1, 1, 1, 0.1)),
// quantised once: min(255, round(25.5 count)) = min(255, (51 count + 1) / 2).
__global__ __launch_bounds__(256) void render_grey_kernel(int64_t len,
                                                          const uint32_t* __restrict__ counts,
                                                          uint8_t* __restrict__ grey) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) {
    const uint32_t c = counts[i];
    grey[i] = c >= 10u ? (uint8_t)255 : (uint8_t)((51u * c + 1u) / 2u);
  }
}

__global__ void zero_counter_kernel(DevScalars* dev) {
  dev->interactions = 0;
  dev->quad_interactions = 0;
  for (int x = 0; x < kXcds; ++x) dev->wq[x] = 0u;
}

// Positions scattered to body-id order (for diagnostics by id).
__global__ __launch_bounds__(256) void by_id_kernel(
    int32_t n, const float* __restrict__ x, const float* __restrict__ y,
    const int32_t* __restrict__ id, float* __restrict__ xi,
    float* __restrict__ yi) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  xi[id[s]] = x[s];
  yi[id[s]] = y[s];
}

constexpr int kTile = 256;

// FP64 direct sum for a list of targets (body ids) over the sources of each
// target's instance (sources staged through LDS in tiles over the union of
// the block's instance ranges).
__global__ __launch_bounds__(kTile) void direct_fp64_kernel(
    int32_t nt, const int32_t* __restrict__ targets,
    const int32_t* __restrict__ inst_of, const int32_t* __restrict__ ioff,
    const float* __restrict__ xid, const float* __restrict__ yid,
    const float* __restrict__ x, const float* __restrict__ y,
    const float* __restrict__ m, double eps2, double* __restrict__ ax,
    double* __restrict__ ay) {
  __shared__ float sx[kTile], sy[kTile], sm[kTile];
  __shared__ int32_t range[2];
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = q < nt;
  const int32_t b = active ? (targets ? targets[q] : q) : 0;
  const int32_t inst = inst_of[b];
  const int32_t lo = ioff[inst], hi = ioff[inst + 1];
  if (threadIdx.x == 0) {
    range[0] = INT32_MAX;
    range[1] = 0;
  }
  __syncthreads();
  if (active) {
    atomicMin(&range[0], lo);
    atomicMax(&range[1], hi);
  }
  __syncthreads();
  const int32_t blo = range[0], bhi = range[1];
  const double xi = active ? (double)xid[b] : 0.0;
  const double yi = active ? (double)yid[b] : 0.0;
  double fx = 0.0, fy = 0.0;
  for (int32_t base = blo; base < bhi; base += kTile) {
    const int32_t j = base + threadIdx.x;
    __syncthreads();
    if (j < bhi) {
      sx[threadIdx.x] = x[j];
      sy[threadIdx.x] = y[j];
      sm[threadIdx.x] = m[j];
    } else {
      sx[threadIdx.x] = 0.f;
      sy[threadIdx.x] = 0.f;
      sm[threadIdx.x] = 0.f;
    }
    __syncthreads();
    const int32_t k0 = lo > base ? lo - base : 0;
    const int32_t k1 = (hi - base) < kTile ? (hi - base) : kTile;
    for (int32_t k = k0; k < k1; ++k) {
      const double dx = (double)sx[k] - xi, dy = (double)sy[k] - yi;
      const double r2 = dx * dx + dy * dy + eps2;
      const double inv = 1.0 / (r2 * sqrt(r2));
      fx += (double)sm[k] * dx * inv;
      fy += (double)sm[k] * dy * inv;
    }
  }
  if (active) {
    ax[q] = fx;
    ay[q] = fy;
  }
}

// Plummer-softened potential, kinetic energy, momentum and angular momentum
// about the origin of kTile targets (slice blockIdx.y of tile blockIdx.x,
// positions of one instance) summed over the instance's sources:
// part[kEnergyVals*(t*kSlices + slice) + {0..4}] = pot, kin, px, py, lz with
// pot = -1/2 sum_i m_i sum_{j != i} m_j / sqrt(r^2 + eps2).  With ax
// non-null the velocities are v + half_dt * a (FP64): the kick-drift update
// leaves v half a step behind the positions, and the accelerations of the
// stored positions bring it level.
constexpr int kSlices = kEnergySlices;
constexpr int kVals = kEnergyVals;
static_assert(kSlices * kTile == kBboxTile, "slices cover a tile");
__global__ __launch_bounds__(kTile) void energy_fp64_kernel(
    const int32_t* __restrict__ tstart, const int32_t* __restrict__ inst_of,
    const int32_t* __restrict__ ioff, const float* __restrict__ x,
    const float* __restrict__ y, const float* __restrict__ vx,
    const float* __restrict__ vy, const float* __restrict__ m,
    const float* __restrict__ ax, const float* __restrict__ ay, double half_dt,
    double eps2, double* __restrict__ part) {
  __shared__ float sx[kTile], sy[kTile], sm[kTile];
  __shared__ double red[kVals][kTile / 64];
  const int32_t t0 = tstart[blockIdx.x], t1 = tstart[blockIdx.x + 1];
  const int32_t inst = inst_of[t0];
  const int32_t lo = ioff[inst], hi = ioff[inst + 1];
  double vals[kVals] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const int32_t tb = t0 + (int32_t)blockIdx.y * kTile;
  if (tb >= t1) {  // a short tile: this slice has no targets
    if (threadIdx.x < kVals)
      part[kVals * ((int64_t)blockIdx.x * kSlices + blockIdx.y) + threadIdx.x] = 0.0;
    return;
  }
  {
    const int32_t i = tb + threadIdx.x;
    const bool active = i < t1;
    const double xi = active ? (double)x[i] : 0.0;
    const double yi = active ? (double)y[i] : 0.0;
    double pot = 0.0;
    for (int32_t base = lo; base < hi; base += kTile) {
      const int32_t j = base + threadIdx.x;
      __syncthreads();
      sx[threadIdx.x] = j < hi ? x[j] : 0.f;
      sy[threadIdx.x] = j < hi ? y[j] : 0.f;
      sm[threadIdx.x] = j < hi ? m[j] : 0.f;
      __syncthreads();
      const int32_t cnt = (hi - base) < kTile ? (hi - base) : kTile;
      for (int32_t k = 0; k < cnt; ++k) {
        if (base + k == i) continue;
        const double dx = (double)sx[k] - xi, dy = (double)sy[k] - yi;
        pot += (double)sm[k] / sqrt(dx * dx + dy * dy + eps2);
      }
    }
    if (active) {
      const double mi = m[i];
      double u = vx[i], v = vy[i];
      if (ax) {
        u += half_dt * (double)ax[i];
        v += half_dt * (double)ay[i];
      }
      vals[0] += -0.5 * mi * pot;
      vals[1] += 0.5 * mi * (u * u + v * v);
      vals[2] += mi * u;
      vals[3] += mi * v;
      vals[4] += mi * (xi * v - yi * u);
    }
  }
  for (int k = 0; k < kVals; ++k)
    for (int off = 32; off > 0; off >>= 1) vals[k] += __shfl_xor(vals[k], off);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < kVals; ++k) red[k][w] = vals[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < kVals; ++k) {
      double s = 0.0;
      for (int q = 0; q < kTile / 64; ++q) s += red[k][q];
      part[kVals * ((int64_t)blockIdx.x * kSlices + blockIdx.y) + k] = s;
    }
  }
}

}  // namespace

// The sort stage: the build's link kernel already wrote the body state in
// tree order into the spare buffers (kernels.cu:292-336 sort the bodies by
// tree order; here the order is the build's), so the stage swaps them in.
void launch_sort(stellar_ctx* c) {
  Buffers& B = c->b;
  std::swap(B.x, B.x2);
  std::swap(B.y, B.y2);
  std::swap(B.vx, B.vx2);
  std::swap(B.vy, B.vy2);
  std::swap(B.m, B.m2);
  std::swap(B.id, B.id2);
}

void launch_update(stellar_ctx* c) {
  const int32_t n = c->n;
  if (n == 0) return;
  Buffers& B = c->b;
  const int32_t threads = (n >> 2) > 4 ? (n >> 2) : 4;
  update_kernel<<<grid_for(threads, 256), 256, 0, c->stream>>>(
      n, B.x, B.y, B.vx, B.vy, B.ax, B.ay, c->p.timestep, c->p.dampening, B.dev);
}

void launch_output(stellar_ctx* c) {
  const int32_t n = c->n;
  if (n == 0) return;
  Buffers& B = c->b;
  output_kernel<<<grid_for(n + c->cap, 256), 256, 0, c->stream>>>(
      n, c->cap, c->ninst, B.x, B.y, B.id, B.blk, B.cmeta, B.dev, B.out);
}

void launch_zero_counter(stellar_ctx* c) {
  zero_counter_kernel<<<1, 1, 0, c->stream>>>(c->b.dev);
}

void launch_render(stellar_ctx* c, const float view[4], int32_t width, int32_t height,
                   uint32_t* dcounts, uint8_t* dgrey) {
  Buffers& B = c->b;
  const int64_t px = (int64_t)width * height;
  fill_u32_kernel<<<grid_for(px, 256), 256, 0, c->stream>>>(px, dcounts);
  const float sx = (float)width / (view[1] - view[0]);
  const float sy = (float)height / (view[3] - view[2]);
  if (c->n > 0)
    render_count_kernel<<<grid_for(c->n, 256), 256, 0, c->stream>>>(
        c->n, B.x, B.y, view[0], view[2], sx, sy, width, height, dcounts);
  if (dgrey) render_grey_kernel<<<grid_for(px, 256), 256, 0, c->stream>>>(px, dcounts, dgrey);
}

void launch_direct_fp64(stellar_ctx* c, int32_t nt, const int32_t* dtargets,
                        double eps2, double* dax, double* day) {
  Buffers& B = c->b;
  const int32_t n = c->n;
  // positions by id, staged in the sort's spare velocity buffers (the spare
  // position buffers hold the build's tree-order copy until the sort stage)
  by_id_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(n, B.x, B.y, B.id,
                                                         B.vx2, B.vy2);
  direct_fp64_kernel<<<grid_for(nt, kTile), kTile, 0, c->stream>>>(
      nt, dtargets, B.inst_of, B.ioff, B.vx2, B.vy2, B.x, B.y, B.m, eps2, dax, day);
}

int energy_parts(stellar_ctx* c) { return c->ntiles * kSlices; }

void launch_energy_fp64(stellar_ctx* c, double eps2, double half_dt, bool sync,
                        double* dpart) {
  Buffers& B = c->b;
  energy_fp64_kernel<<<dim3((unsigned)c->ntiles, kSlices), kTile, 0, c->stream>>>(
      B.tstart, B.inst_of, B.ioff, B.x, B.y, B.vx, B.vy, B.m, sync ? B.ax : nullptr,
      sync ? B.ay : nullptr, half_dt, eps2, dpart);
}

}  // namespace stellar
