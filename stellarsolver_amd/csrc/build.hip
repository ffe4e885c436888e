// Bounding box and tree build for gfx950.
//
// The reference grows its quadtree by concurrent top-down insertion with
// per-slot locks and float atomics (kernels.cu:117-271), which makes cell ids
// depend on scheduling.  The tree it builds is nevertheless unique for a
// given set of FP32 positions: a cell at level L exists exactly when two or
// more bodies share their first L slot codes, and a body's slot codes depend
// only on its own position and the bounding box (the FP32 midpoints of
// kernels.cu:144-178, recomputed along its own path).  This file builds that
// same tree without locks, one launch per step of
//
//   1. keys   : each body's first slot codes -> one 64-bit key (instance
//               bits on top); between steps the same kernel counts the keys
//               into the sample sort's buckets (sort.hip);
//   2. sort   : sample sort of (key, storage index) seeded with the previous
//               sort's quantiles (sort.hip); rocPRIM's radix sort above 2^24
//               bodies;
//   3. ties   : runs of equal keys are ordered by the deeper levels and body
//               id; the kernel also reads off the next sort's splitters;
//   4. lambda : common-prefix depth of neighbours -> cells opened per pair,
//               their count per 256 positions, and the first level of a
//               minimum tree over the depths; clears the tree blocks;
//   5. mintree: the rest of the minimum tree and the scan of the per-block
//               counts (every cell's DFS pre-order id is then a block total
//               plus a block-local prefix);
//   6. link   : each position creates the cells it opens, links them and its
//               body into their parents (a previous-smaller search over the
//               minimum tree), writes the cells' metadata and the bodies'
//               x, y, m in tree order.
//
// The result is directly the canonical form (cells in DFS pre-order, bodies
// in sort_kernel order), so it is compared bit for bit with the reference
// algorithm restated on the CPU.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <mutex>

#include "engine.h"

namespace stellar {

namespace {

constexpr int kBboxThreads = 512;

// Sort configuration for captured steps: rocPRIM's merge-sort path at every
// size (its onesweep path, chosen above 2^20 items, clears its counters
// with memset nodes, which captured graphs must not contain).
using MergeSortConfig = rocprim::radix_sort_config<rocprim::default_config,
                                                   rocprim::default_config,
                                                   rocprim::default_config, (size_t)1 << 40>;

// min/max that never replace +0 by -0 and ignore NaN (the reference's
// fminf/fmaxf over a box seeded with zeros, kernels.cu:51-58,62-114).
__device__ __forceinline__ float mn(float a, float v) { return v < a ? v : a; }
__device__ __forceinline__ float mx(float a, float v) { return v > a ? v : a; }

// Pass 1: min/max of one tile (kBboxTile positions of one instance) into
// part[tile]; pass 2: one wavefront per instance folds its tiles.  Two
// launches instead of a last-block ticket: the ticket pattern needs a
// device-scope release per block, which on gfx950 writes back the issuing
// XCD's L2.
// Tile bounds come from tstart, or (tstart == nullptr) are the uniform tiles
// [k*kBboxTile, (k+1)*kBboxTile) of n positions.
__global__ __launch_bounds__(kBboxThreads) void bbox_kernel(
    const float* __restrict__ x, const float* __restrict__ y,
    const int32_t* __restrict__ tstart, int32_t n, float4* __restrict__ part) {
  float l = 0.f, r = 0.f, b = 0.f, t = 0.f;
  const int32_t lo = tstart ? tstart[blockIdx.x] : (int32_t)blockIdx.x * kBboxTile;
  const int32_t hi = tstart ? tstart[blockIdx.x + 1] : (n - lo < kBboxTile ? n : lo + kBboxTile);
  for (int32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const float px = x[i], py = y[i];
    l = mn(l, px);
    r = mx(r, px);
    b = mn(b, py);
    t = mx(t, py);
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
    const float4* __restrict__ part, const int32_t* __restrict__ itile,
    int32_t ninst, float4* __restrict__ ibbox) {
  const int32_t inst = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (inst >= ninst) return;  // whole wavefronts
  float l = 0.f, r = 0.f, b = 0.f, t = 0.f;
  for (int32_t i = itile[inst] + lane; i < itile[inst + 1]; i += 64) {
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
  if (lane == 0) ibbox[inst] = make_float4(l, r, b, t);
}

__global__ void init_scalars_kernel(DevScalars* dev) {
  dev->err = 0;
  dev->ncells = 0;
  dev->bbox[0] = dev->bbox[1] = dev->bbox[2] = dev->bbox[3] = 0.f;
  dev->interactions = 0;
  dev->quad_interactions = 0;
}

// Sort keys.  With one instance a key holds slot codes 1..32.  With several,
// its top ibits bits hold the instance index, so that one sort orders
// (instance, path) and every instance keeps its own position range, and the
// codes below hold levels 1..l1 = 32 - ibits/2.  Levels past l1 are computed
// only inside runs of equal keys (ties_kernel): l1+1..l1+32 in key2, the
// last ibits/2 in key3 (left-aligned).
struct KeyView {
  const uint64_t* key;
  const uint64_t* key2;
  const uint32_t* key3;
  int l1;
  int ibits;
  // key holds Hilbert curve positions instead of slot codes for levels
  // 1..kHilbertLevels (<= l1); deeper levels are slot codes either way
  int hilbert;
};

template <bool HILBERT>
__global__ __launch_bounds__(256) void keys_kernel(
    const float* __restrict__ x, const float* __restrict__ y, int32_t n,
    const int32_t* __restrict__ inst_of, const float4* __restrict__ ibbox,
    int l1, int ibits, uint64_t* __restrict__ key, uint32_t* __restrict__ perm) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t inst = inst_of[i];
  key[i] = path_key<HILBERT>(x[i], y[i], ibbox[inst], inst, l1, ibits);
  perm[i] = (uint32_t)i;
}

// Levels l1+1..l1+32 (k2) and l1+33..64 (k3, left-aligned) of one body
// (its path is replayed from the root).
__device__ void deep_key(float px, float py, float4 bb, int l1, uint64_t& k2,
                         uint32_t& k3) {
  float l = bb.x, r = bb.y, b = bb.z, t = bb.w;
  for (int lev = 0; lev < l1; ++lev) (void)split2(px, py, l, r, b, t);
  k2 = 0;
  for (int lev = 0; lev < 32; ++lev) k2 = (k2 << 2) | split2(px, py, l, r, b, t);
  k3 = 0;
  const int rest = kMaxLevel - l1 - 32;
  for (int lev = 0; lev < rest; ++lev) k3 |= split2(px, py, l, r, b, t) << (30 - 2 * lev);
}

// Runs of equal keys (bodies of one instance closer than a level-l1 cell):
// order each run by (deeper levels, body id) with an in-place insertion
// sort.  Such runs are rare and short; longer than kMaxTieRun is an error.
// With the sample sort (spl != nullptr) the first nb - 1 threads also read
// off the next sort's splitters: the sorted keys at the bucket quantiles.
__global__ __launch_bounds__(256) void ties_kernel(
    const float* __restrict__ x, const float* __restrict__ y,
    const int32_t* __restrict__ id, int32_t n, const uint64_t* __restrict__ key,
    const int32_t* __restrict__ inst_of, const float4* __restrict__ ibbox, int l1,
    uint32_t* __restrict__ perm, uint64_t* __restrict__ key2,
    uint32_t* __restrict__ key3, DevScalars* __restrict__ dev, int32_t nb,
    uint64_t* __restrict__ spl) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (spl && s < nb - 1) spl[s] = key[(int64_t)(s + 1) * n / nb - 1];
  if (s + 1 >= n) return;
  const uint64_t k = key[s];  // the instance is part of the key
  if (key[s + 1] != k || (s > 0 && key[s - 1] == k)) return;
  int32_t e = s + 1;
  while (e + 1 < n && key[e + 1] == k && e - s + 1 < kMaxTieRun) ++e;
  if (e + 1 < n && key[e + 1] == k) {
    atomicOr(&dev->err, kErrTieRun);
    return;
  }
  const float4 bb = ibbox[inst_of[s]];
  for (int32_t j = s; j <= e; ++j) {
    const uint32_t p = perm[j];
    deep_key(x[p], y[p], bb, l1, key2[j], key3[j]);
  }
  for (int32_t j = s + 1; j <= e; ++j) {
    const uint64_t k2 = key2[j];
    const uint32_t k3 = key3[j];
    const uint32_t p = perm[j];
    const int32_t b = id[p];
    int32_t q = j - 1;
    while (q >= s && (key2[q] > k2 ||
                      (key2[q] == k2 && (key3[q] > k3 || (key3[q] == k3 && id[perm[q]] > b))))) {
      key2[q + 1] = key2[q];
      key3[q + 1] = key3[q];
      perm[q + 1] = perm[q];
      --q;
    }
    key2[q + 1] = k2;
    key3[q + 1] = k3;
    perm[q + 1] = p;
  }
}

// Number of leading slot codes shared by tree positions s and t (s < t) of
// one instance.  Equal keys only occur inside a tie run, where key2 and key3
// are valid.
__device__ __forceinline__ int common_levels(const KeyView& K, int32_t s, int32_t t) {
  const uint64_t a = K.key[s] ^ K.key[t];
  if (a) return (__clzll(a) - K.ibits) >> 1;
  const uint64_t c = K.key2[s] ^ K.key2[t];
  if (c) return K.l1 + (__clzll(c) >> 1);
  const uint32_t d = K.key3[s] ^ K.key3[t];
  if (d) return K.l1 + 32 + (__clz(d) >> 1);
  return kMaxLevel;
}

// Slot codes of one tree position at consecutive levels.  With Hilbert keys
// a level <= kHilbertLevels holds the curve position h, read in the frame that the levels
// above give it: the XOR of their frame changes {1, 0, 0, 3}, whose bit 0 is
// the parity of the digits above with both bits equal (h = 0 or 3) and bit 1
// the parity of those equal to 3 -- two popcounts of the key's prefix for the
// first level asked for, then one frame change per level.
struct SlotWalk {
  int lev;         // the level slot_code answers for
  uint32_t frame;  // Hilbert frame of that level
  uint64_t k;
};

__device__ __forceinline__ SlotWalk slot_walk(const KeyView& K, int32_t s, int lev) {
  SlotWalk w{lev, 0u, K.key[s]};
  if (K.hilbert && lev <= kHilbertLevels && lev > 1) {
    constexpr uint64_t kEven = 0x5555555555555555ull;
    // digits of levels 1..lev-1 (the instance bits above them masked off)
    const uint64_t p = w.k >> (2 * (K.l1 - lev + 1));
    const uint64_t m = (((uint64_t)1 << (2 * (lev - 1))) - 1) & kEven;
    const uint64_t lo = p & m, hi = (p >> 1) & m;
    w.frame = (uint32_t)(__popcll(~(lo ^ hi) & m) & 1) |
              (uint32_t)((__popcll(lo & hi) & 1) << 1);
  }
  return w;
}

__device__ __forceinline__ int slot_code(const KeyView& K, int32_t s, const SlotWalk& w) {
  const int lev = w.lev;
  if (lev <= K.l1) {
    const uint32_t h = (uint32_t)(w.k >> (2 * (K.l1 - lev))) & 3u;
    return (int)(K.hilbert && lev <= kHilbertLevels ? hilbert_slot(w.frame, h) : h);
  }
  if (lev <= K.l1 + 32) return (int)((K.key2[s] >> (2 * (K.l1 + 32 - lev))) & 3u);
  return (int)((K.key3[s] >> (30 - 2 * (lev - K.l1 - 33))) & 3u);
}

__device__ __forceinline__ void slot_next(const KeyView& K, SlotWalk& w) {
  if (K.hilbert && w.lev <= kHilbertLevels)
    w.frame = hilbert_child_frame(w.frame, (uint32_t)(w.k >> (2 * (K.l1 - w.lev))) & 3u);
  ++w.lev;
}

// Minimum tree over lam: level 0 is lam itself (one byte per tree position,
// lam[n-1] = 0), level k >= 1 holds the minimum of 64 entries of level k-1.
// Every level is padded to whole 64-byte groups so that a group is read with
// four 16-byte loads; entries past a level's length are never used.
constexpr int kMinTreeLevels = 6;  // up to 64^6 positions
struct MinTree {
  const uint8_t* lv[kMinTreeLevels];
  int32_t len[kMinTreeLevels];
  int top;  // highest level, len[top] <= 64
};

MinTree mintree_layout(int32_t n, const uint8_t* lam, const uint8_t* lmin,
                       int64_t* bytes) {
  MinTree t{};
  t.lv[0] = lam;
  t.len[0] = n;
  t.top = 0;
  int64_t off = 0;
  while (t.len[t.top] > 64 && t.top + 1 < kMinTreeLevels) {
    const int32_t m = (t.len[t.top] + 63) / 64;
    ++t.top;
    t.lv[t.top] = lmin ? lmin + off : nullptr;
    t.len[t.top] = m;
    off += ((int64_t)m + 63) / 64 * 64;
  }
  if (bytes) *bytes = off + 64;
  return t;
}

// Cells opened by pair (s, s+1): levels lam(s-1)+1 .. lam(s).  Writes lam,
// the exclusive prefix of the cell counts within each block of kScanBlock
// positions (clocal) and each block's total (bsum; the minimum-tree kernel
// scans the totals, so coff(s) = bsum[s / kScanBlock] + clocal[s] is the id
// of the first cell pair s opens), and, per wavefront of 64 positions, level
// 1 of the minimum tree.  Also clears tree blocks [0, nclear) with coalesced
// 16-byte stores (nclear = n: a tree has fewer cells than bodies unless many
// bodies pair up closely; the minimum-tree kernel clears any blocks beyond).
constexpr int kScanBlock = 256;

__global__ __launch_bounds__(kScanBlock) void lambda_kernel(
    int32_t n, KeyView K, const int32_t* __restrict__ inst_of, int32_t* __restrict__ clocal,
    int32_t* __restrict__ bsum, uint8_t* __restrict__ lam, uint8_t* __restrict__ lmin1,
    float4* __restrict__ blk, int64_t nclear) {
  __shared__ int32_t wsum[kScanBlock / 64];
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  {
    const int64_t r0 = 4 * (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t q = r0 + (int64_t)r * blockDim.x;
      if (q < 4 * nclear) blk[q] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  int v = 0, lm = 0;
  if (s + 1 < n) {
    // a pair that straddles two instances shares no level (lam = 0, like the
    // end of the last instance)
    const int32_t is = inst_of[s];
    lm = inst_of[s + 1] == is ? common_levels(K, s, s + 1) : 0;
    const int prev = s > 0 && inst_of[s - 1] == is ? common_levels(K, s - 1, s) : 0;
    v = lm > prev ? lm - prev : 0;
  }
  if (s < n) lam[s] = (uint8_t)lm;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lmin1) {
    int w = s < n ? lm : 255;
    for (int off = 32; off > 0; off >>= 1) w = min(w, __shfl_xor(w, off));
    if (lane == 0 && s < n) lmin1[s >> 6] = (uint8_t)w;
  }
  int incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int o = __shfl_up(incl, off);
    if (lane >= off) incl += o;
  }
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  int before = incl - v, tot = 0;
  for (int w = 0; w < kScanBlock / 64; ++w) {
    if (w < wv) before += wsum[w];
    tot += wsum[w];
  }
  if (s < n) clocal[s] = before;
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Levels 2..top of the minimum tree (one workgroup; each entry reads its 64
// children as four 16-byte loads), and the exclusive scan of the lambda
// kernel's per-block cell counts (in place, nblk of them).
__global__ __launch_bounds__(1024) void mintree_kernel(MinTree t, int32_t* __restrict__ bsum,
                                                       int32_t nblk, float4* __restrict__ blk,
                                                       int64_t cleared, int64_t max_blocks,
                                                       int32_t nroot) {
  {
    __shared__ int32_t part[1024 / 64];
    const int per = (nblk + (int)blockDim.x - 1) / (int)blockDim.x;
    const int b0 = (int)threadIdx.x * per;
    const int b1 = b0 + per < nblk ? b0 + per : nblk;
    int mine = 0;
    for (int j = b0; j < b1; ++j) mine += bsum[j];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int incl = mine;
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(incl, off);
      if (lane >= off) incl += o;
    }
    if (lane == 63) part[wv] = incl;
    __syncthreads();
    int run = incl - mine;
    for (int w = 0; w < wv; ++w) run += part[w];
    for (int j = b0; j < b1; ++j) {
      const int c = bsum[j];
      bsum[j] = run;
      run += c;
    }
    // blocks beyond those the lambda kernel cleared, up to the last cell's
    // (the last thread's running total is the cell count)
    __shared__ int32_t total;
    if (threadIdx.x == blockDim.x - 1) total = run;
    __syncthreads();
    int64_t used = (int64_t)total + nroot;
    if (used > max_blocks) used = max_blocks;
    for (int64_t q = 4 * cleared + threadIdx.x; q < 4 * used; q += blockDim.x)
      blk[q] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  for (int k = 2; k <= t.top; ++k) {
    const uint8_t* src = t.lv[k - 1];
    uint8_t* dst = const_cast<uint8_t*>(t.lv[k]);
    for (int32_t e = threadIdx.x; e < t.len[k]; e += blockDim.x) {
      const int32_t j0 = 64 * e;
      const int cnt = t.len[k - 1] - j0 < 64 ? t.len[k - 1] - j0 : 64;
      const uint4* q = reinterpret_cast<const uint4*>(src + j0);
      const uint4 a[4] = {q[0], q[1], q[2], q[3]};
      int w = 255;
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const uint32_t word = u % 4 == 0 ? a[u / 4].x
                              : u % 4 == 1 ? a[u / 4].y
                              : u % 4 == 2 ? a[u / 4].z
                                           : a[u / 4].w;
#pragma unroll
        for (int b = 0; b < 4; ++b)
          if (4 * u + b < cnt) w = min(w, (int)((word >> (8 * b)) & 255u));
      }
      dst[e] = (uint8_t)w;
    }
    __syncthreads();
  }
}

// Bit 7 of byte i set when byte i of w is below v (bytes <= 255, v <= 65):
// (b | 0x80) - v keeps bit 7 exactly when (b & 0x7f) >= v, and never borrows
// from the next byte.  Entries are <= 64 or 255, so (b & 0x7f) < v <=> b < v.
__device__ __forceinline__ uint64_t below_mask(uint64_t w, uint32_t v) {
  const uint64_t hi = 0x8080808080808080ull;
  return ~((w | hi) - 0x0101010101010101ull * v) & hi;
}

// Largest j < cnt with g[j] < v among the 64 entries at g (64-byte aligned),
// or -1.
__device__ __forceinline__ int last_below(const uint8_t* g, int cnt, uint32_t v) {
  const uint4* q = reinterpret_cast<const uint4*>(g);
  const uint4 a0 = q[0], a1 = q[1], a2 = q[2], a3 = q[3];
  const uint64_t w[8] = {
      (uint64_t)a0.x | ((uint64_t)a0.y << 32), (uint64_t)a0.z | ((uint64_t)a0.w << 32),
      (uint64_t)a1.x | ((uint64_t)a1.y << 32), (uint64_t)a1.z | ((uint64_t)a1.w << 32),
      (uint64_t)a2.x | ((uint64_t)a2.y << 32), (uint64_t)a2.z | ((uint64_t)a2.w << 32),
      (uint64_t)a3.x | ((uint64_t)a3.y << 32), (uint64_t)a3.z | ((uint64_t)a3.w << 32)};
  int res = -1;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint64_t m = below_mask(w[k], v);
    const int lo = 8 * k;
    if (cnt <= lo) m = 0;
    else if (cnt < lo + 8) m &= (1ull << (8 * (cnt - lo))) - 1ull;
    if (m) res = lo + ((63 - __clzll(m)) >> 3);
  }
  return res;
}

// Previous smaller value: the largest j < i with lam[j] < v, or -1.
__device__ int32_t psv(const MinTree& t, int32_t i, uint32_t v) {
  int k = 0;
  int32_t pos = i;  // search level-k entries [pos & ~63, pos)
  while (true) {
    const int32_t g0 = pos & ~63;
    const int j = pos > g0 ? last_below(t.lv[k] + g0, pos - g0, v) : -1;
    if (j >= 0) {
      pos = g0 + j;
      break;
    }
    if (g0 == 0 || k == t.top) return -1;
    pos = g0 >> 6;
    ++k;
  }
  while (k > 0) {  // descend into the last group holding a smaller value
    --k;
    const int32_t g0 = pos << 6;
    const int32_t cnt = t.len[k] - g0 < 64 ? t.len[k] - g0 : 64;
    const int j = last_below(t.lv[k] + g0, cnt, v);
    if (j < 0) return -1;  // only if the tree is inconsistent
    pos = g0 + j;
  }
  return pos;
}

struct LinkArgs {
  int32_t n;
  int64_t cap;
  KeyView K;
  int32_t* coff;
  const int32_t* bpre;    // exclusive prefix of each block's cell count
  const int32_t* clocal;  // exclusive prefix within the block
  const uint8_t* lam;
  const uint32_t* perm;
  const float* x;
  const float* y;
  const float* m;
  ChildRec* blk;
  uint64_t* cmeta;
  DevScalars* dev;
  const int32_t* inst_of;
  const int32_t* ioff;
  int32_t nroot;  // number of instances = root blocks; cell c is block nroot + c
  MinTree mt;
  const float4* ibbox;
  // the body state moved into tree order for the sort stage, which swaps it in
  const float* vx;
  const float* vy;
  const int32_t* id;
  float* x2;
  float* y2;
  float* m2;
  float* vx2;
  float* vy2;
  int32_t* id2;
};

// Id of the first cell pair a opens.
__device__ __forceinline__ int64_t coff_at(const LinkArgs& A, int32_t a) {
  return (int64_t)A.bpre[a / kScanBlock] + A.clocal[a];
}

__device__ __forceinline__ void put_cell(const LinkArgs& A, int64_t blk, int slot, int64_t child) {
  A.blk[4 * blk + slot] = ChildRec{0.f, 0.f, 0.f, (int32_t)child};
}

// The tree's node blocks and the cells' metadata; one thread per tree
// position s.  Nodes are the instance's root (level 0, block inst) and its
// cells (block nroot + id); their blocks were cleared by the lambda and
// minimum-tree kernels, so a slot nobody writes is empty.  With lp = lam(s-1)
// (0 at the instance's first position) and ls = lam(s), position s
//   A. creates the nodes that start at s -- levels lp+1 .. ls, and the root
//      at the instance's first position -- and writes each one's first child
//      (the next node created, or body s);
//   B. unless it is the instance's first position, starts a new child of the
//      node at level lp that holds s-1 and s (found by a previous-smaller
//      search over lam) and writes it: its first created cell, or body s.
// A slot's digit is the position's slot code one level down, or in a level-64
// cell (bodies closer than 64 levels resolve) its rank.
__global__ __launch_bounds__(kScanBlock) void link_kernel(LinkArgs A) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= A.n) return;
  // body s's state, read once through the permutation: its record values,
  // and the whole state written out in tree order for the sort stage (a
  // permutation even when the build fails, so the state stays whole)
  const uint32_t p = A.perm[s];
  const float bx = A.x[p], by = A.y[p], bm = A.m[p];
  A.x2[s] = bx;
  A.y2[s] = by;
  A.m2[s] = bm;
  A.vx2[s] = A.vx[p];
  A.vy2[s] = A.vy[p];
  A.id2[s] = A.id[p];
  const ChildRec body{bx, by, bm, -1 - s};
  const int32_t inst = A.inst_of[s];
  const int32_t i0 = A.ioff[inst];
  const bool first = s == i0;
  const int lp = first ? 0 : A.lam[s - 1];
  const int ls = A.lam[s];
  const int64_t cs = coff_at(A, s);
  A.coff[s] = (int32_t)cs;
  if (s == A.n - 1) {
    if (cs > A.cap) atomicOr(&A.dev->err, kErrPool);
    A.dev->ncells = (int32_t)(cs < A.cap ? cs : A.cap);
    // the centre-of-mass passes count their leftovers from zero, the walk
    // its interactions and work-queue claims
#pragma unroll
    for (int k = 0; k < 4; ++k) A.dev->com_len[k] = 0;
    A.dev->interactions = 0;
    A.dev->quad_interactions = 0;
#pragma unroll
    for (int k = 0; k < kXcds; ++k) A.dev->wq[k] = 0u;
  }
  // block of the node created at s at level L (-1: beyond the pool)
  auto created = [&](int L) -> int64_t {
    if (L == 0) return inst;
    const int64_t id = cs + (L - lp - 1);
    return id < A.cap ? id + A.nroot : -1;
  };

  // B's node: level lp, holding s-1 and s, starting at position a (its
  // opening pair), la = lam(a-1)
  int32_t a = i0;
  int la = 0;
  int64_t pblk = inst;
  if (!first && lp > 0) {
    a = psv(A.mt, s - 1, (uint32_t)lp) + 1;
    la = a > 0 ? A.lam[a - 1] : 0;
    const int64_t pid = coff_at(A, a) + (lp - la - 1);
    pblk = pid < A.cap ? pid + A.nroot : -1;
  }

  // slot codes of s from level lp + 1 on (level 1 for the instance's first
  // position); B's slot is the first
  SlotWalk w = slot_walk(A.K, s, first ? 1 : lp + 1);
  const int dcur = first || pblk < 0 ? 0 : lp < kMaxLevel ? slot_code(A.K, s, w) : s - a;

  // A: nodes created at s
  if (first || ls > lp) {
    for (int L = first ? 0 : lp + 1; L <= ls; ++L) {
      const int64_t b = created(L);
      if (b < 0) break;
      if (L > 0) {
        A.cmeta[b] = cell_meta(s, (int32_t)(L == lp + 1 && !first ? pblk : created(L - 1)),
                               slot_code(A.K, s, w), L);
        slot_next(A.K, w);
      }
      const int d = L < kMaxLevel ? slot_code(A.K, s, w) : 0;
      if (L < ls) {
        const int64_t cb = created(L + 1);
        if (cb >= 0) put_cell(A, b, d, cb);
      } else {
        A.blk[4 * b + d] = body;
      }
    }
  }

  // B: s starts a new child of the node at level lp
  if (!first && pblk >= 0) {
    if (dcur > 3) {
      atomicOr(&A.dev->err, kErrBucket);
    } else {
      if (ls > lp) {
        const int64_t cb = created(lp + 1);
        if (cb >= 0) put_cell(A, pblk, dcur, cb);
      } else {
        A.blk[4 * pblk + dcur] = body;
      }
    }
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
  size_t a2 = 0;
  (void)rocprim::radix_sort_pairs(nullptr, a, kb, vb, (size_t)n, 0, 64);
  (void)rocprim::radix_sort_pairs<MergeSortConfig>(nullptr, a2, kb, vb, (size_t)n, 0, 64);
  (void)rocprim::exclusive_scan(nullptr, b, q, q, 0, (size_t)n,
                                rocprim::plus<int32_t>());
  if (a2 > a) a = a2;
  return (a > b ? a : b) + 256;
}

void launch_init_scalars(stellar_ctx* c) {
  init_scalars_kernel<<<1, 1, 0, c->stream>>>(c->b.dev);
}

void launch_bbox(stellar_ctx* c) {
  if (c->n == 0) return;
  Buffers& B = c->b;
  bbox_kernel<<<(unsigned)c->ntiles, kBboxThreads, 0, c->stream>>>(B.x, B.y, B.tstart, c->n,
                                                                    B.bbox_part);
  const unsigned waves_per_block = kBboxThreads / 64;
  bbox_final_kernel<<<(unsigned)((c->ninst + waves_per_block - 1) / waves_per_block),
                      kBboxThreads, 0, c->stream>>>(B.bbox_part, B.itile, c->ninst, B.ibbox);
}

// Instance bits of the sort key (even, so that slot codes stay 2-bit aligned).
int key_instance_bits(int32_t ninst) {
  int b = 0;
  while (b < 32 && ((int64_t)1 << b) < ninst) b += 2;
  return ninst > 1 ? b : 0;
}

void launch_build(stellar_ctx* c) {
  Buffers& B = c->b;
  const int32_t n = c->n;
  if (n == 0) return;
  const int ibits = key_instance_bits(c->ninst);
  const int l1 = 32 - ibits / 2;
  size_t bytes = B.tmp_bytes;
  const bool ss = sample_sort_fits(n) && !c->radix_sort;
  // with the previous sort's splitters at hand, the keys kernel of the key
  // sort also counts the keys per bucket
  const bool fused = ss && c->ss_have_spl;
  const bool hilbert = (c->p.flags & STELLAR_FLAG_HILBERT_ORDER) != 0;
  if (fused)
    launch_keys_count(c, l1, ibits, hilbert);
  else
    (hilbert ? keys_kernel<true> : keys_kernel<false>)<<<grid_for(n, 256), 256, 0, c->stream>>>(
        B.x, B.y, n, B.inst_of, B.ibbox, l1, ibits, B.key, B.perm);
  if (ss) {
    launch_sample_sort(c, fused);  // sorted data back in (key, perm)
  } else {
    rocprim::double_buffer<uint64_t> kb(B.key, B.key_alt);
    rocprim::double_buffer<uint32_t> vb(B.perm, B.perm_alt);
    if (c->capturing)
      (void)rocprim::radix_sort_pairs<MergeSortConfig>(B.tmp, bytes, kb, vb, (size_t)n, 0, 64,
                                                       c->stream);
    else
      (void)rocprim::radix_sort_pairs(B.tmp, bytes, kb, vb, (size_t)n, 0, 64, c->stream);
    // keep the sorted data in the primary buffers
    if (kb.current() != B.key) std::swap(B.key, B.key_alt);
    if (vb.current() != B.perm) std::swap(B.perm, B.perm_alt);
  }
  ties_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(B.x, B.y, B.id, n, B.key, B.inst_of,
                                                        B.ibbox, l1, B.perm, B.key2, B.key3,
                                                        B.dev, ss ? sample_sort_buckets(n) : 0,
                                                        ss ? B.ss_spl : nullptr);
  const KeyView K{B.key, B.key2, B.key3, l1, ibits, hilbert ? 1 : 0};
  int64_t mtb = 0;
  const MinTree mt = mintree_layout(n, B.lam, B.lmin, &mtb);
  const int32_t nblk = (int32_t)grid_for(n, kScanBlock);
  const int64_t max_blocks = c->cap + c->ninst;
  const int64_t nclear = (int64_t)n < max_blocks ? (int64_t)n : max_blocks;
  lambda_kernel<<<nblk, kScanBlock, 0, c->stream>>>(
      n, K, B.inst_of, B.ncell, B.bsum, B.lam,
      mt.top >= 1 ? const_cast<uint8_t*>(mt.lv[1]) : nullptr, reinterpret_cast<float4*>(B.blk),
      nclear);
  mintree_kernel<<<1, 1024, 0, c->stream>>>(mt, B.bsum, nblk, reinterpret_cast<float4*>(B.blk),
                                            nclear, max_blocks, c->ninst);
  LinkArgs A{n,         c->cap,   K,       B.coff,    B.bsum,   B.ncell,  B.lam,
             B.perm,    B.x,      B.y,     B.m,       B.blk,    B.cmeta,
             B.dev,     B.inst_of, B.ioff,
             c->ninst,  mt,        B.ibbox,
             B.vx,      B.vy,     B.id,
             B.x2,      B.y2,     B.m2,      B.vx2,    B.vy2,    B.id2};
  link_kernel<<<nblk, kScanBlock, 0, c->stream>>>(A);
}

size_t mintree_bytes(int32_t n) {
  int64_t b = 0;
  (void)mintree_layout(n, nullptr, nullptr, &b);
  return (size_t)b;
}

// ---- allocation-free bounding box (stellar_bbox_noalloc) -----------------
namespace {

constexpr int kNoAllocTiles = STELLAR_BBOX_NOALLOC_MAX_TILES;
static_assert((int64_t)STELLAR_BBOX_NOALLOC_MAX_BODIES == (int64_t)kNoAllocTiles * kBboxTile,
              "the header's body limit is whole tiles");

// Per-tile partial boxes of stellar_bbox_noalloc: the library's fixed device
// array (one copy on each GPU, at its own address).  Host code never passes
// this name to a kernel: it resolves each device's address once and uses that.
static __device__ __align__(256) float4 scratch[kNoAllocTiles];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
float4* scratch_ptr[kMaxDevices] = {};

// Address of `scratch` on the current device (resolved once per device).
int scratch_on_device(float4** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return STELLAR_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return STELLAR_ERR_DEVICE;
    scratch_ptr[dev] = static_cast<float4*>(p);
  }
  *out = scratch_ptr[dev];
  return STELLAR_OK;
}

// One wavefront folds ntiles partial boxes (none: the origin alone) into out4.
__global__ __launch_bounds__(64) void bbox_fold_kernel(const float4* __restrict__ part,
                                                       int32_t ntiles, float* __restrict__ out4) {
  const int lane = threadIdx.x;
  float l = 0.f, r = 0.f, b = 0.f, t = 0.f;
  for (int32_t i = lane; i < ntiles; i += 64) {
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
  if (lane == 0) {
    out4[0] = l;
    out4[1] = r;
    out4[2] = b;
    out4[3] = t;
  }
}

}  // namespace

}  // namespace stellar

extern "C" int stellar_bbox_noalloc(const float* x, const float* y, int32_t n, float* bbox4,
                                    void* stream) {
  using namespace stellar;
  if (n < 0 || !bbox4 || (n > 0 && (!x || !y))) return STELLAR_ERR_ARG;
  if ((int64_t)n > (int64_t)STELLAR_BBOX_NOALLOC_MAX_BODIES) return STELLAR_ERR_SIZE;
  float4* part = nullptr;
  const int rc = scratch_on_device(&part);
  if (rc != STELLAR_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int32_t ntiles = (int32_t)(((int64_t)n + kBboxTile - 1) / kBboxTile);
  if (ntiles > 0) bbox_kernel<<<(unsigned)ntiles, kBboxThreads, 0, s>>>(x, y, nullptr, n, part);
  bbox_fold_kernel<<<1, 64, 0, s>>>(part, ntiles, bbox4);
  return hipGetLastError() == hipSuccess ? STELLAR_OK : STELLAR_ERR_DEVICE;
}
