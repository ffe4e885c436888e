// Bounding box and tree build for gfx950.
//
// The reference grows its quadtree by concurrent top-down insertion with
// per-slot locks and float atomics (kernels.cu:117-271), which makes cell ids
// depend on scheduling.  The tree it builds is nevertheless unique for a
// given set of FP32 positions: a cell at level L exists exactly when two or
// more bodies share their first L slot codes, and a body's slot codes depend
// only on its own position and the bounding box (the FP32 midpoints of
// kernels.cu:144-178, recomputed along its own path).  This file builds that
// same tree without locks:
//
//   1. keys   : each body's 32 first slot codes -> one 64-bit key
//   2. sort   : LSD radix sort of (key, storage index)
//   3. ties   : runs of equal keys are ordered by levels 33..64 and body id
//   4. lambda : common-prefix depth of neighbours -> cells opened per pair
//   5. scan   : exclusive scan -> every cell gets its DFS pre-order id
//   6. link   : each cell finds its body range and its four children
//
// The result is directly the canonical form (cells in DFS pre-order, bodies
// in sort_kernel order), so it is compared bit for bit with the reference
// algorithm restated on the CPU.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "engine.h"

namespace stellar {

namespace {

constexpr int kBboxThreads = 256;

// min/max that never replace +0 by -0 and ignore NaN (the reference's
// fminf/fmaxf over a box seeded with zeros, kernels.cu:51-58,62-114).
__device__ __forceinline__ float mn(float a, float v) { return v < a ? v : a; }
__device__ __forceinline__ float mx(float a, float v) { return v > a ? v : a; }

// Pass 1: per-block min/max into part[blockIdx]; pass 2 (one block) folds
// the partials.  Two launches instead of a last-block ticket: the ticket
// pattern needs a device-scope release per block, which on gfx950 writes
// back the issuing XCD's L2.
__global__ __launch_bounds__(kBboxThreads) void bbox_kernel(
    const float* __restrict__ x, const float* __restrict__ y, int32_t n,
    float4* __restrict__ part) {
  float l = 0.f, r = 0.f, b = 0.f, t = 0.f;
  const int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t stride = gridDim.x * blockDim.x;
  const int32_t n4 = n >> 2;
  const float4* x4 = reinterpret_cast<const float4*>(x);
  const float4* y4 = reinterpret_cast<const float4*>(y);
  for (int32_t i = tid; i < n4; i += stride) {
    const float4 a = x4[i], c = y4[i];
    l = mn(mn(mn(mn(l, a.x), a.y), a.z), a.w);
    r = mx(mx(mx(mx(r, a.x), a.y), a.z), a.w);
    b = mn(mn(mn(mn(b, c.x), c.y), c.z), c.w);
    t = mx(mx(mx(mx(t, c.x), c.y), c.z), c.w);
  }
  for (int32_t i = 4 * n4 + tid; i < n; i += stride) {
    l = mn(l, x[i]);
    r = mx(r, x[i]);
    b = mn(b, y[i]);
    t = mx(t, y[i]);
  }
  // wavefront reduction
  for (int off = 32; off > 0; off >>= 1) {
    l = mn(l, __shfl_xor(l, off));
    r = mx(r, __shfl_xor(r, off));
    b = mn(b, __shfl_xor(b, off));
    t = mx(t, __shfl_xor(t, off));
  }
  __shared__ float red[4][kBboxThreads / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = l;
    red[1][w] = r;
    red[2][w] = b;
    red[3][w] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBboxThreads / 64; ++k) {
      l = mn(l, red[0][k]);
      r = mx(r, red[1][k]);
      b = mn(b, red[2][k]);
      t = mx(t, red[3][k]);
    }
    part[blockIdx.x] = make_float4(l, r, b, t);
  }
}

__global__ __launch_bounds__(kBboxThreads) void bbox_final_kernel(
    const float4* __restrict__ part, int32_t nparts,
    DevScalars* __restrict__ dev) {
  float l = 0.f, r = 0.f, b = 0.f, t = 0.f;
  for (int32_t i = threadIdx.x; i < nparts; i += blockDim.x) {
    const float4 q = part[i];
    l = mn(l, q.x);
    r = mx(r, q.y);
    b = mn(b, q.z);
    t = mx(t, q.w);
  }
  for (int off = 32; off > 0; off >>= 1) {
    l = mn(l, __shfl_xor(l, off));
    r = mx(r, __shfl_xor(r, off));
    b = mn(b, __shfl_xor(b, off));
    t = mx(t, __shfl_xor(t, off));
  }
  __shared__ float red[4][kBboxThreads / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = l;
    red[1][w] = r;
    red[2][w] = b;
    red[3][w] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBboxThreads / 64; ++k) {
      l = mn(l, red[0][k]);
      r = mx(r, red[1][k]);
      b = mn(b, red[2][k]);
      t = mx(t, red[3][k]);
    }
    dev->bbox[0] = l;
    dev->bbox[1] = r;
    dev->bbox[2] = b;
    dev->bbox[3] = t;
  }
}

__global__ void init_scalars_kernel(DevScalars* dev) {
  dev->err = 0;
  dev->ncells = 0;
  dev->bbox[0] = dev->bbox[1] = dev->bbox[2] = dev->bbox[3] = 0.f;
  dev->interactions = 0;
}

// One level of box subdivision along one axis (kernels.cu:144-157): the
// FP32 sum lo+hi is halved in double, p < mid is decided in double and the
// new bound is mid rounded to FP32.  When |lo+hi| >= 2^-125 the halving is
// exact in FP32, so the FP32 path gives the identical decision and bound.
__device__ __forceinline__ uint32_t split1(float p, float& lo, float& hi) {
  const float s = lo + hi;
  bool lt;
  float mid;
  if (fabsf(s) >= 0x1p-125f) {
    mid = s * 0.5f;
    lt = p < mid;
  } else {
    const double md = 0.5 * (double)s;
    lt = (double)p < md;
    mid = (float)md;
  }
  if (lt)
    hi = mid;
  else
    lo = mid;
  return lt ? 1u : 0u;
}

// Slot code of the next level: +1 west (x < mid), +2 south (y < mid).
__device__ __forceinline__ uint32_t split2(float px, float py, float& l,
                                           float& r, float& b, float& t) {
  const uint32_t cx = split1(px, l, r);
  const uint32_t cy = split1(py, b, t);
  return cx | (cy << 1);
}

__global__ __launch_bounds__(256) void keys_kernel(
    const float* __restrict__ x, const float* __restrict__ y, int32_t n,
    const DevScalars* __restrict__ dev, uint64_t* __restrict__ key,
    uint32_t* __restrict__ perm) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float l = dev->bbox[0], r = dev->bbox[1], b = dev->bbox[2], t = dev->bbox[3];
  const float px = x[i], py = y[i];
  uint64_t k = 0;
#pragma unroll 4
  for (int lev = 0; lev < 32; ++lev) k = (k << 2) | split2(px, py, l, r, b, t);
  key[i] = k;
  perm[i] = (uint32_t)i;
}

// Levels 33..64 of one body (its path is replayed from the root).
__device__ uint64_t deep_key(float px, float py, const float* bbox) {
  float l = bbox[0], r = bbox[1], b = bbox[2], t = bbox[3];
  for (int lev = 0; lev < 32; ++lev) (void)split2(px, py, l, r, b, t);
  uint64_t k = 0;
  for (int lev = 0; lev < 32; ++lev) k = (k << 2) | split2(px, py, l, r, b, t);
  return k;
}

// Runs of equal 32-level keys (bodies closer than a level-32 cell): order
// each run by (levels 33..64, body id) with an in-place insertion sort.
// Such runs are rare and short; longer than kMaxTieRun is an error.
__global__ __launch_bounds__(256) void ties_kernel(
    const float* __restrict__ x, const float* __restrict__ y,
    const int32_t* __restrict__ id, int32_t n, const uint64_t* __restrict__ key,
    uint32_t* __restrict__ perm, uint64_t* __restrict__ key2,
    DevScalars* __restrict__ dev) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s + 1 >= n) return;
  const uint64_t k = key[s];
  if (key[s + 1] != k || (s > 0 && key[s - 1] == k)) return;
  int32_t e = s + 1;
  while (e + 1 < n && key[e + 1] == k && e - s + 1 < kMaxTieRun) ++e;
  if (e + 1 < n && key[e + 1] == k) {
    atomicOr(&dev->err, kErrTieRun);
    return;
  }
  float bbox[4] = {dev->bbox[0], dev->bbox[1], dev->bbox[2], dev->bbox[3]};
  for (int32_t j = s; j <= e; ++j) {
    const uint32_t p = perm[j];
    key2[j] = deep_key(x[p], y[p], bbox);
  }
  for (int32_t j = s + 1; j <= e; ++j) {
    const uint64_t k2 = key2[j];
    const uint32_t p = perm[j];
    const int32_t b = id[p];
    int32_t q = j - 1;
    while (q >= s &&
           (key2[q] > k2 || (key2[q] == k2 && id[perm[q]] > b))) {
      key2[q + 1] = key2[q];
      perm[q + 1] = perm[q];
      --q;
    }
    key2[q + 1] = k2;
    perm[q + 1] = p;
  }
}

// Number of leading slot codes shared by tree positions s and t (s < t).
// Equal 32-level keys only occur inside a tie run, where key2 is valid.
__device__ __forceinline__ int common_levels(const uint64_t* __restrict__ key,
                                             const uint64_t* __restrict__ key2,
                                             int32_t s, int32_t t) {
  const uint64_t a = key[s] ^ key[t];
  if (a) return __clzll(a) >> 1;
  const uint64_t c = key2[s] ^ key2[t];
  if (c) return 32 + (__clzll(c) >> 1);
  return kMaxLevel;
}

__device__ __forceinline__ int digit_at(const uint64_t* __restrict__ key,
                                        const uint64_t* __restrict__ key2,
                                        int32_t s, int lev) {
  return lev <= 32 ? (int)((key[s] >> (2 * (32 - lev))) & 3u)
                   : (int)((key2[s] >> (2 * (64 - lev))) & 3u);
}

// Cells opened by pair (s, s+1): levels lambda(s-1)+1 .. lambda(s).
__global__ __launch_bounds__(256) void lambda_kernel(
    int32_t n, const uint64_t* __restrict__ key,
    const uint64_t* __restrict__ key2, int32_t* __restrict__ ncell) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int v = 0;
  if (s + 1 < n) {
    const int lam = common_levels(key, key2, s, s + 1);
    const int prev = s > 0 ? common_levels(key, key2, s - 1, s) : 0;
    v = lam > prev ? lam - prev : 0;
  }
  ncell[s] = v;
}

struct LinkArgs {
  int32_t n;
  int64_t cap;
  const uint64_t* key;
  const uint64_t* key2;
  const int32_t* coff;
  const int32_t* ncell;
  const uint32_t* perm;
  const float* x;
  const float* y;
  const float* m;
  ChildRec* blk;
  int32_t* cstart;
  int32_t* ccount;
  int32_t* cparent;
  uint8_t* clevel;
  uint8_t* cslot;
  int32_t* cend;
  DevScalars* dev;
};

// Largest j in [s, n-1] sharing at least L levels with s.
__device__ int32_t range_end(const LinkArgs& A, int32_t s, int L) {
  int32_t lo = s, step = 1;
  while (lo + step <= A.n - 1 && common_levels(A.key, A.key2, s, lo + step) >= L) {
    lo += step;
    step <<= 1;
  }
  int32_t hi = lo + step < A.n ? lo + step : A.n;
  while (hi - lo > 1) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if (common_levels(A.key, A.key2, s, mid) >= L)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Largest v in [u, e] whose code at level lev is <= d (codes ascend there).
__device__ int32_t digit_end(const LinkArgs& A, int32_t u, int32_t e, int lev,
                             int d) {
  int32_t lo = u, step = 1;
  while (lo + step <= e && digit_at(A.key, A.key2, lo + step, lev) <= d) {
    lo += step;
    step <<= 1;
  }
  int32_t hi = lo + step <= e ? lo + step : e + 1;
  while (hi - lo > 1) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if (digit_at(A.key, A.key2, mid, lev) <= d)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int lam_before(const LinkArgs& A, int32_t u) {
  return u > 0 ? common_levels(A.key, A.key2, u - 1, u) : 0;
}

// Fill block `blkid` for the node holding tree positions [a, e] whose
// bodies share `lev` levels (lev = 0: the root).
__device__ void link_node(const LinkArgs& A, int32_t blkid, int32_t a,
                          int32_t e, int lev) {
  ChildRec rec[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) rec[k] = ChildRec{0.f, 0.f, 0.f, 0};
  if (lev == kMaxLevel) {
    // deepest level: keep up to four bodies by rank
    if (e - a + 1 > 4) {
      atomicOr(&A.dev->err, kErrBucket);
      e = a + 3;
    }
    for (int32_t u = a; u <= e; ++u) {
      const uint32_t p = A.perm[u];
      rec[u - a] = ChildRec{A.x[p], A.y[p], A.m[p], -1 - u};
    }
  } else {
    int32_t u = a;
    while (u <= e) {
      const int d = digit_at(A.key, A.key2, u, lev + 1);
      const int32_t v = digit_end(A, u, e, lev + 1, d);
      if (u == v) {
        const uint32_t p = A.perm[u];
        rec[d] = ChildRec{A.x[p], A.y[p], A.m[p], -1 - u};
      } else {
        const int64_t c2 = (int64_t)A.coff[u] + (lev + 1 - lam_before(A, u) - 1);
        // a child cell always follows its parent in pre-order; anything else
        // means the keys were left inconsistent by an earlier error
        if (c2 < A.cap && c2 + 1 > blkid) {
          const int32_t b2 = (int32_t)(c2 + 1);
          rec[d] = ChildRec{0.f, 0.f, 0.f, b2};
          A.cparent[b2] = blkid;
          A.cslot[b2] = (uint8_t)d;
        }
      }
      u = v + 1;
    }
  }
  ChildRec* dst = A.blk + 4 * (int64_t)blkid;
#pragma unroll
  for (int k = 0; k < 4; ++k) dst[k] = rec[k];
  A.cend[blkid] = A.coff[e];  // pairs s..e-1 open the subtree's cells
  A.cstart[blkid] = a;
  A.ccount[blkid] = e - a + 1;
  A.clevel[blkid] = (uint8_t)lev;
}

// One thread per adjacent pair s: creates the cells that pair opens, at
// levels lambda(s-1)+1 .. lambda(s), each with id coff[s] + offset.  The
// last thread (s = n-1, which opens none) links the root block.
__global__ __launch_bounds__(256) void link_kernel(LinkArgs A) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= A.n) return;
  if (s == A.n - 1) {
    const int64_t total = (int64_t)A.coff[s] + A.ncell[s];
    if (total > A.cap) atomicOr(&A.dev->err, kErrPool);
    A.dev->ncells = (int32_t)(total < A.cap ? total : A.cap);
    link_node(A, 0, 0, A.n - 1, 0);
    return;
  }
  const int nc = A.ncell[s];
  if (nc == 0) return;
  const int lam = common_levels(A.key, A.key2, s, s + 1);
  const int first = lam - nc + 1;
  for (int L = lam; L >= first; --L) {
    const int64_t c = (int64_t)A.coff[s] + (L - first);
    if (c >= A.cap) continue;
    const int32_t e = range_end(A, s, L);
    link_node(A, (int32_t)(c + 1), s, e, L);
  }
}

inline unsigned grid_for(int64_t items, int threads) {
  return (unsigned)((items + threads - 1) / threads);
}

}  // namespace

size_t build_scratch_bytes(int32_t n) {
  size_t a = 0, b = 0;
  uint64_t* k = nullptr;
  uint32_t* v = nullptr;
  int32_t* q = nullptr;
  rocprim::double_buffer<uint64_t> kb(k, k);
  rocprim::double_buffer<uint32_t> vb(v, v);
  (void)rocprim::radix_sort_pairs(nullptr, a, kb, vb, (size_t)n, 0, 64);
  (void)rocprim::exclusive_scan(nullptr, b, q, q, 0, (size_t)n,
                                rocprim::plus<int32_t>());
  return (a > b ? a : b) + 256;
}

void launch_init_scalars(stellar_ctx* c) {
  init_scalars_kernel<<<1, 1, 0, c->stream>>>(c->b.dev);
}

void launch_bbox(stellar_ctx* c) {
  const int32_t n = c->n;
  int blocks = (int)grid_for((n + 3) / 4, kBboxThreads * 4);
  if (blocks < 1) blocks = 1;
  if (blocks > kBboxMaxBlocks) blocks = kBboxMaxBlocks;
  bbox_kernel<<<blocks, kBboxThreads, 0, c->stream>>>(c->b.x, c->b.y, n,
                                                       c->b.bbox_part);
  bbox_final_kernel<<<1, kBboxThreads, 0, c->stream>>>(c->b.bbox_part, blocks,
                                                        c->b.dev);
}

void launch_build(stellar_ctx* c) {
  Buffers& B = c->b;
  const int32_t n = c->n;
  if (n == 0) return;
  keys_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(B.x, B.y, n, B.dev,
                                                        B.key, B.perm);
  rocprim::double_buffer<uint64_t> kb(B.key, B.key_alt);
  rocprim::double_buffer<uint32_t> vb(B.perm, B.perm_alt);
  size_t bytes = B.tmp_bytes;
  (void)rocprim::radix_sort_pairs(B.tmp, bytes, kb, vb, (size_t)n, 0, 64,
                                  c->stream);
  // keep the sorted data in the primary buffers
  if (kb.current() != B.key) std::swap(B.key, B.key_alt);
  if (vb.current() != B.perm) std::swap(B.perm, B.perm_alt);
  ties_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(B.x, B.y, B.id, n, B.key,
                                                        B.perm, B.key2, B.dev);
  lambda_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(n, B.key, B.key2,
                                                          B.ncell);
  bytes = B.tmp_bytes;
  (void)rocprim::exclusive_scan(B.tmp, bytes, B.ncell, B.coff, 0, (size_t)n,
                                rocprim::plus<int32_t>(), c->stream);
  LinkArgs A{n,        c->cap,    B.key,     B.key2,    B.coff,
             B.ncell,  B.perm,    B.x,       B.y,       B.m,
             B.blk,    B.cstart,  B.ccount,  B.cparent, B.clevel,
             B.cslot,  B.cend,    B.dev};
  link_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(A);
}

}  // namespace stellar
