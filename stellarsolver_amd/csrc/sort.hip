// Sort of the bodies' path keys for gfx950: a sample sort that exploits the
// order the bodies are already in.
//
// Between steps the bodies are stored in the previous step's tree order, so
// the new keys are nearly sorted: half the bodies move fewer than 16
// positions, 99% fewer than ~2000, and a few tenths of a percent jump far
// (they crossed a high-level cell boundary, or the box changed).  A general
// merge or radix sort pays full price for that; here
//   1. the splitters (nb - 1 keys cutting buckets of about kSsBucket keys)
//      are the previous sort's quantiles: the sorted keys at every
//      (n/nb)-th position, read off after each sort.  The keys move a
//      little between steps, so the buckets stay balanced; after an upload
//      one workgroup sorts a strided sample of the keys (up to 4 per
//      bucket) instead.  Splitters only balance the buckets: any sorted set
//      gives the same result;
//   2. every key finds its bucket by binary search over the splitters (in
//      LDS); each workgroup counts its tile's keys per bucket and adds the
//      counts to the bucket totals (global atomics: a tile of nearly sorted
//      keys touches a few buckets).  Between steps this is done by the kernel
//      that computes the keys;
//   3. every scattering workgroup scans the totals into the buckets' starts,
//      reserves a range in each bucket it writes to (one atomic per bucket
//      and workgroup) and scatters its keys and values there;
//   4. one workgroup per bucket sorts it (a block radix sort, stable, with
//      the bucket's keys in their index order ahead of the padding) and writes
//      the final order.  Keys of a bucket agree above the highest bit in
//      which its two splitters differ; when more than 32 bits remain, the
//      block sort takes the 32-bit window below that bit (four 8-bit passes
//      instead of up to eight) and runs of equal windows -- bodies closer
//      than the window resolves, a few pairs per step -- are put in full-key
//      order by insertion afterwards (a run longer than 32 sends the bucket
//      to the full-key sort).
// Equal keys always land in the same bucket (the bucket is the number of
// splitters below the key), so runs of equal keys end up contiguous, which is
// all the later tie refinement needs (the order inside a run is not used).
// A bucket larger than one block sort (cap keys) is sorted in runs that one
// workgroup then merges (merge path) through the spare buffers: correct for
// any input, and only reached when the sample misjudges the distribution
// badly (a first step from an arbitrary body order, say).
#include <rocprim/block/block_radix_sort.hpp>

#include "engine.h"

namespace stellar {

namespace {

constexpr int kSsOver = 4;              // samples per bucket
constexpr int kSsMaxSamples = 8192;     // at most: beyond 8192 buckets some splitters repeat
constexpr int kSsThreads = 256;       // threads of the sample workgroup
constexpr int kSsBucketThreads = 512;  // threads of a bucket's workgroup
constexpr int kSsTileThreads = 1024;  // threads of the counting / scattering workgroups
constexpr int kSsTileMin = 2048;  // keys per counting / scattering workgroup (at least)
constexpr int kSsMaxTiles = 1024;  // counting / scattering workgroups (at most)
// Bucket capacities of the counting and scattering kernels' LDS (splitters
// and one counter per bucket): the smallest that holds the sort's buckets,
// so that small sorts keep their LDS, and with it the workgroups per CU, small.
constexpr int kSsBucketsSmall = 1024;  // n <= 1.5M
constexpr int kSsBucketsMid = 4096;    // n <= 6.3M
static_assert(kSsBucketsMid < kSsMaxBuckets, "three capacity classes");
constexpr int kSsRunItems = kSsCapMax / kSsBucketThreads;   // 4 per thread
constexpr int kSsSampleItems = kSsMaxSamples / kSsThreads;  // 32 per thread
template <int IPT>
using RunSort = rocprim::block_radix_sort<uint64_t, kSsBucketThreads, IPT, uint32_t>;
template <int IPT>
using SampleSort = rocprim::block_radix_sort<uint64_t, kSsThreads, IPT>;
// 32-bit window of the keys with the run position as value (one-run buckets)
using WinSort = rocprim::block_radix_sort<uint32_t, kSsBucketThreads, kSsRunItems, uint32_t>;
constexpr int kSsFixMax = 32;  // longest run of equal windows fixed by one thread
template <class A, class B>
constexpr size_t max_size() {
  return sizeof(A) > sizeof(B) ? sizeof(A) : sizeof(B);
}

// Bucket of a key: the number of splitters below it (0 .. nb-1).
__device__ __forceinline__ int bucket_of(const uint64_t* sp, int nb, uint64_t key) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (sp[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// 1. strided sample -> sorted -> splitters spl[0 .. nb-2].  The block sort
// holds IPT samples per thread (the smallest of 2, 4, 8 that fits).
template <int IPT>
__device__ void sort_samples(const uint64_t* key, int32_t n, int s, uint64_t* sorted,
                             void* storage) {
  uint64_t k[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = threadIdx.x * IPT + j;  // blocked arrangement
    k[j] = i < s ? key[(int64_t)i * n / s] : ~0ull;
  }
  SampleSort<IPT>().sort(k, *reinterpret_cast<typename SampleSort<IPT>::storage_type*>(storage));
#pragma unroll
  for (int j = 0; j < IPT; ++j) sorted[threadIdx.x * IPT + j] = k[j];
}

__global__ __launch_bounds__(kSsThreads) void ss_sample_kernel(const uint64_t* __restrict__ key,
                                                               int32_t n, int32_t nb,
                                                               uint64_t* __restrict__ spl) {
  __shared__ __align__(16) unsigned char store[max_size<typename SampleSort<kSsSampleItems>::storage_type,
                                                        typename SampleSort<8>::storage_type>()];
  __shared__ uint64_t sorted[kSsMaxSamples];
  const int s = nb * kSsOver < kSsMaxSamples ? nb * kSsOver : kSsMaxSamples;
  if (s <= 8 * kSsThreads)
    sort_samples<8>(key, n, s, sorted, store);
  else if (s <= 16 * kSsThreads)
    sort_samples<16>(key, n, s, sorted, store);
  else
    sort_samples<kSsSampleItems>(key, n, s, sorted, store);
  __syncthreads();
  for (int j = threadIdx.x; j < nb - 1; j += blockDim.x) {
    const int q = (int)(((int64_t)(j + 1) * s) / nb) - 1;
    spl[j] = sorted[q > 0 ? q : 0];
  }
}

// Splitters into LDS and zeroed per-bucket counts (nb <= kSsMaxBuckets).
__device__ __forceinline__ void load_splitters(const uint64_t* __restrict__ spl, int32_t nb,
                                               uint64_t* sp, uint32_t* cnt) {
  for (int j = threadIdx.x; j < nb; j += blockDim.x) {
    if (j < nb - 1) sp[j] = spl[j];
    cnt[j] = 0;
  }
  __syncthreads();
}

// The tile's per-bucket counts (LDS) added to the bucket totals.
__device__ __forceinline__ void add_counts(const uint32_t* cnt, int32_t nb,
                                           uint32_t* __restrict__ tot) {
  __syncthreads();
  for (int j = threadIdx.x; j < nb; j += blockDim.x)
    if (cnt[j]) atomicAdd(&tot[j], cnt[j]);
}

// 2. bucket totals of the keys (the first sort after an upload, whose
// splitters are sampled from keys computed beforehand).  MAXB >= nb.
template <int MAXB>
__global__ __launch_bounds__(kSsTileThreads) void ss_count_kernel(const uint64_t* __restrict__ key,
                                                              int32_t n, int32_t nb, int32_t tile,
                                                              const uint64_t* __restrict__ spl,
                                                              uint32_t* __restrict__ tot) {
  __shared__ uint64_t sp[MAXB];
  __shared__ uint32_t cnt[MAXB];
  load_splitters(spl, nb, sp, cnt);
  const int64_t t0 = (int64_t)blockIdx.x * tile;
  const int64_t t1 = t0 + tile < n ? t0 + tile : n;
  for (int64_t e = t0 + threadIdx.x; e < t1; e += blockDim.x)
    atomicAdd(&cnt[bucket_of(sp, nb, key[e])], 1u);
  add_counts(cnt, nb, tot);
}

}  // namespace

// 1+2. keys of a tile (as the build's keys kernel computes them) and their
// bucket totals, when the splitters are already known (every sort after the
// first).  MAXB >= nb.
template <int MAXB, bool HILBERT>
__global__ __launch_bounds__(kSsTileThreads) void ss_keys_count_kernel(
    const float* __restrict__ x, const float* __restrict__ y, int32_t n,
    const int32_t* __restrict__ inst_of, const float4* __restrict__ ibbox, int l1, int ibits,
    int32_t nb, int32_t tile, const uint64_t* __restrict__ spl, uint64_t* __restrict__ key,
    uint32_t* __restrict__ perm, uint32_t* __restrict__ tot) {
  __shared__ uint64_t sp[MAXB];
  __shared__ uint32_t cnt[MAXB];
  load_splitters(spl, nb, sp, cnt);
  const int64_t t0 = (int64_t)blockIdx.x * tile;
  const int64_t t1 = t0 + tile < n ? t0 + tile : n;
  for (int64_t e = t0 + threadIdx.x; e < t1; e += blockDim.x) {
    const int32_t inst = inst_of[e];
    const uint64_t k = path_key<HILBERT>(x[e], y[e], ibbox[inst], inst, l1, ibits);
    key[e] = k;
    perm[e] = (uint32_t)e;
    atomicAdd(&cnt[bucket_of(sp, nb, k)], 1u);
  }
  add_counts(cnt, nb, tot);
}

namespace {

// 3. scatter to the buckets (order inside a bucket is arbitrary).  Every
// workgroup scans the totals itself (workgroup 0 also writes the bucket
// starts for the bucket sort), then reserves its range in each bucket its
// tile writes to.  cur[] (the buckets' fill levels) starts at zero.
// MAXB >= nb.
template <int MAXB>
__global__ __launch_bounds__(kSsTileThreads) void ss_scatter_kernel(
    const uint64_t* __restrict__ key, const uint32_t* __restrict__ val, int32_t n, int32_t nb,
    int32_t tile, const uint64_t* __restrict__ spl, const uint32_t* __restrict__ tot,
    uint32_t* __restrict__ cur, uint32_t* __restrict__ bstart, uint64_t* __restrict__ key_out,
    uint32_t* __restrict__ val_out) {
  __shared__ uint64_t sp[MAXB];
  __shared__ uint32_t pos[MAXB];  // tile counts, then write cursors
  __shared__ uint32_t part[kSsTileThreads / 64];
  load_splitters(spl, nb, sp, pos);
  const int64_t t0 = (int64_t)blockIdx.x * tile;
  const int64_t t1 = t0 + tile < n ? t0 + tile : n;
  for (int64_t e = t0 + threadIdx.x; e < t1; e += blockDim.x)
    atomicAdd(&pos[bucket_of(sp, nb, key[e])], 1u);
  // exclusive scan of the totals: thread t sums buckets [t*per, (t+1)*per)
  const int per = (nb + (int)blockDim.x - 1) / (int)blockDim.x;
  const int b0 = (int)threadIdx.x * per;
  const int b1 = b0 + per < nb ? b0 + per : nb;
  uint32_t mine = 0;
  for (int j = b0; j < b1; ++j) mine += tot[j];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t incl = mine;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = __shfl_up(incl, off);
    if (lane >= off) incl += o;
  }
  if (lane == 63) part[wv] = incl;
  __syncthreads();
  uint32_t run = incl - mine;
  for (int w = 0; w < wv; ++w) run += part[w];
  for (int j = b0; j < b1; ++j) {
    if (blockIdx.x == 0) bstart[j] = run;
    const uint32_t c = pos[j];
    pos[j] = c ? run + atomicAdd(&cur[j], c) : 0u;
    run += tot[j];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) bstart[nb] = (uint32_t)n;
  __syncthreads();
  for (int64_t e = t0 + threadIdx.x; e < t1; e += blockDim.x) {
    const uint64_t k = key[e];
    const uint32_t p = atomicAdd(&pos[bucket_of(sp, nb, k)], 1u);
    key_out[p] = k;
    val_out[p] = val[e];
  }
}

// Merges sorted runs [a, a+len) and [a+len, b) of (ks, vs) into (kd, vd) over
// [a, b), by the whole workgroup: each thread finds where its slice of the
// output splits the two runs (merge path) and merges it.
__device__ void wg_merge(const uint64_t* ks, const uint32_t* vs, uint64_t* kd, uint32_t* vd,
                         int64_t a, int64_t len, int64_t b) {
  const int64_t mid = a + len < b ? a + len : b;
  const int64_t na = mid - a, nbb = b - mid, tot = b - a;
  const int64_t per = (tot + blockDim.x - 1) / blockDim.x;
  const int64_t o0 = (int64_t)threadIdx.x * per;
  if (o0 >= tot) return;
  const int64_t o1 = o0 + per < tot ? o0 + per : tot;
  // i + j = o0 with A[i-1] <= B[j] and B[j-1] < A[i]
  int64_t lo = o0 > nbb ? o0 - nbb : 0, hi = o0 < na ? o0 : na;
  while (lo < hi) {
    const int64_t i = (lo + hi) >> 1, j = o0 - i - 1;
    if (ks[a + i] <= ks[mid + j])
      lo = i + 1;
    else
      hi = i;
  }
  int64_t i = lo, j = o0 - lo;
  for (int64_t o = o0; o < o1; ++o) {
    const bool takea = j >= nbb || (i < na && ks[a + i] <= ks[mid + j]);
    if (takea) {
      kd[a + o] = ks[a + i];
      vd[a + o] = vs[a + i];
      ++i;
    } else {
      kd[a + o] = ks[mid + j];
      vd[a + o] = vs[mid + j];
      ++j;
    }
  }
}

// 4. one workgroup per bucket: (kb, vb) hold the buckets; the sorted result
// goes to (ko, vo).  Buckets up to `cap` keys take one block sort; larger
// ones are sorted in runs of `cap` into (ko, vo) and merged back and forth
// with (kb, vb), ending in (ko, vo).  A run is loaded in index order
// (blocked) with padding keys ~0 after it; the sort is stable, so a real key
// ~0 stays ahead of the padding.
// One block sort of keys [r0, r0 + rm) of (kb, vb) into (ko, vo), rm <= IPT
// per thread.
template <int IPT>
__device__ void sort_run(const uint64_t* kb, const uint32_t* vb, int64_t r0, int rm,
                         uint64_t* ko, uint32_t* vo, void* storage, unsigned end_bit) {
  uint64_t k[IPT];
  uint32_t v[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = threadIdx.x * IPT + j;
    k[j] = i < rm ? kb[r0 + i] : ~0ull;
    v[j] = i < rm ? vb[r0 + i] : 0u;
  }
  RunSort<IPT>().sort_to_striped(
      k, v, *reinterpret_cast<typename RunSort<IPT>::storage_type*>(storage), 0u, end_bit);
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = j * kSsBucketThreads + threadIdx.x;  // striped arrangement
    if (i < rm) {
      ko[r0 + i] = k[j];
      vo[r0 + i] = v[j];
    }
  }
}

// LDS of a bucket workgroup: the full keys of a one-run bucket, then either
// block sort's storage or the sorted windows and positions.
struct BucketLds {
  uint64_t key[kSsCapMax];
  union {
    typename RunSort<kSsRunItems>::storage_type run;
    typename WinSort::storage_type win;
    struct {
      uint32_t w[kSsCapMax];    // sorted windows
      uint32_t pos[kSsCapMax];  // their positions in the bucket
    } post;
  } u;
  int32_t fallback;
};

// One block sort of a bucket [r0, r0 + rm), rm <= kSsCapMax, whose keys agree
// above end_bit > 32: sort the 32-bit windows key >> (end_bit - 32) with the
// keys' positions, then put each run of equal windows (keys that differ only
// below the window: bodies closer than a cell of the window's last level)
// in full-key order by insertion.  Four radix passes instead of up to eight.
// Returns false, with nothing written, if a run is longer than kSsFixMax
// (a dense cluster): the caller then sorts the full keys.
__device__ bool sort_run_window(const uint64_t* kb, const uint32_t* vb, int64_t r0, int rm,
                                uint64_t* ko, uint32_t* vo, BucketLds& L, unsigned end_bit) {
  constexpr int IPT = kSsRunItems;
  const unsigned shift = end_bit - 32u;
  uint32_t w[IPT], v[IPT];
  if (threadIdx.x == 0) L.fallback = 0;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = threadIdx.x * IPT + j;  // blocked: stable order = position order
    const uint64_t k = i < rm ? kb[r0 + i] : ~0ull;
    L.key[i] = k;
    w[j] = (uint32_t)(k >> shift);  // padding ~0 sorts after every real key
    v[j] = (uint32_t)i;
  }
  WinSort().sort(w, v, L.u.win, 0u, 32u);
  __syncthreads();  // the sort's storage is reused below
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = threadIdx.x * IPT + j;
    L.u.post.w[i] = w[j];
    L.u.post.pos[i] = v[j];
  }
  __syncthreads();
  // the thread holding a run's first entry sorts the whole run
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int p = threadIdx.x * IPT + j;
    if (p + 1 >= rm) continue;
    const uint32_t wp = L.u.post.w[p];
    if (L.u.post.w[p + 1] != wp || (p > 0 && L.u.post.w[p - 1] == wp)) continue;
    int q = p + 2;
    while (q < rm && L.u.post.w[q] == wp && q - p <= kSsFixMax) ++q;
    if (q - p > kSsFixMax) {
      atomicOr(&L.fallback, 1);
      continue;
    }
    for (int a = p + 1; a < q; ++a) {  // insertion by full key (stable)
      const uint32_t pa = L.u.post.pos[a];
      const uint64_t ka = L.key[pa];
      int b = a;
      while (b > p && L.key[L.u.post.pos[b - 1]] > ka) {
        L.u.post.pos[b] = L.u.post.pos[b - 1];
        --b;
      }
      L.u.post.pos[b] = pa;
    }
  }
  __syncthreads();
  if (L.fallback) return false;
  for (int i = threadIdx.x; i < rm; i += kSsBucketThreads) {
    const uint32_t src = L.u.post.pos[i];
    ko[r0 + i] = L.key[src];
    vo[r0 + i] = vb[r0 + src];
  }
  return true;
}

// The bucket's totals and fill level are cleared for the next sort here (the
// scatter that read them has finished).
__global__ __launch_bounds__(kSsBucketThreads) void ss_bucket_kernel(
    uint64_t* __restrict__ kb, uint32_t* __restrict__ vb, int32_t nb,
    const uint32_t* __restrict__ bstart, uint32_t* __restrict__ tot, uint32_t* __restrict__ cur,
    const uint64_t* __restrict__ spl, int32_t cap, uint64_t* __restrict__ ko,
    uint32_t* __restrict__ vo) {
  __shared__ BucketLds L;
  const int b = blockIdx.x;
  const int64_t s = bstart[b];
  const int64_t e = bstart[b + 1];
  if (threadIdx.x == 0) {
    tot[b] = 0u;
    cur[b] = 0u;
  }
  const int64_t m = e - s;
  if (m <= 0) return;
  // every key of the bucket lies in [lo, hi] and so shares their common
  // leading bits, which need no sorting (the padding key ~0 sorts last
  // either way: its low bits are all ones)
  const uint64_t lo = b > 0 ? spl[b - 1] : 0ull;
  const uint64_t hi = b + 1 < nb ? spl[b] : ~0ull;
  const uint64_t dif = lo ^ hi;
  const unsigned end_bit = dif ? 64u - (unsigned)__builtin_clzll(dif) : 1u;
  const int64_t nruns = (m + cap - 1) / cap;
  if (nruns == 1 && end_bit > 32u) {
    if (sort_run_window(kb, vb, s, (int)m, ko, vo, L, end_bit)) return;
    __syncthreads();
  }
  for (int64_t r = 0; r < nruns; ++r) {
    const int64_t r0 = s + r * cap;
    const int rm = (int)((r0 + cap < e ? r0 + cap : e) - r0);
    sort_run<kSsRunItems>(kb, vb, r0, rm, ko, vo, &L.u.run, end_bit);
    __syncthreads();
  }
  if (nruns == 1) return;
  // merge runs of length len: (ko, vo) -> (kb, vb) -> (ko, vo) ...
  bool in_out = true;  // current runs live in (ko, vo)
  for (int64_t len = cap; len < m; len *= 2) {
    const uint64_t* ks = in_out ? ko : kb;
    const uint32_t* vs = in_out ? vo : vb;
    uint64_t* kd = in_out ? kb : ko;
    uint32_t* vd = in_out ? vb : vo;
    for (int64_t a = s; a < e; a += 2 * len) {
      const int64_t bnd = a + 2 * len < e ? a + 2 * len : e;
      wg_merge(ks, vs, kd, vd, a, len, bnd);
      __syncthreads();
    }
    in_out = !in_out;
    __threadfence_block();
    __syncthreads();
  }
  if (!in_out) {
    for (int64_t i = s + threadIdx.x; i < e; i += blockDim.x) {
      ko[i] = kb[i];
      vo[i] = vb[i];
    }
  }
}

}  // namespace

bool sample_sort_fits(int32_t n) { return n >= 1 && (int64_t)n <= kSsMaxN; }

int32_t sample_sort_buckets(int32_t n) {
  const int64_t nb = ((int64_t)n + kSsBucket - 1) / kSsBucket;
  return (int32_t)(nb < 1 ? 1 : nb);
}

// Keys per counting / scattering workgroup: kSsTileMin, or more so that at
// most kSsMaxTiles workgroups (and bucket-count columns) are needed.
static int32_t sample_sort_tile(int32_t n) {
  const int64_t per = ((int64_t)n + kSsMaxTiles - 1) / kSsMaxTiles;
  const int64_t t = (per + 1023) / 1024 * 1024;
  return (int32_t)(t > kSsTileMin ? t : kSsTileMin);
}

int32_t sample_sort_tiles(int32_t n) {
  const int32_t t = sample_sort_tile(n);
  return (int32_t)(((int64_t)n + t - 1) / t);
}

// Sorts (B.key, B.perm) by key; the result is in B.key, B.perm.  keys_done:
// the keys were computed, and their bucket totals added, by
// ss_keys_count_kernel (launch_build does that when splitters are known).
void launch_sample_sort(stellar_ctx* c, bool keys_counted) {
  Buffers& B = c->b;
  const int32_t n = c->n;
  const int32_t nb = sample_sort_buckets(n);
  const int32_t nwg = sample_sort_tiles(n);
  const int32_t tile = sample_sort_tile(n);
  uint32_t* tot = B.ss_tot;
  uint32_t* cur = B.ss_tot + kSsMaxBuckets;
  if (!keys_counted) {
    if (!c->ss_have_spl) ss_sample_kernel<<<1, kSsThreads, 0, c->stream>>>(B.key, n, nb, B.ss_spl);
    auto count = nb <= kSsBucketsSmall ? ss_count_kernel<kSsBucketsSmall>
                 : nb <= kSsBucketsMid ? ss_count_kernel<kSsBucketsMid>
                                       : ss_count_kernel<kSsMaxBuckets>;
    count<<<nwg, kSsTileThreads, 0, c->stream>>>(B.key, n, nb, tile, B.ss_spl, tot);
  }
  auto scatter = nb <= kSsBucketsSmall ? ss_scatter_kernel<kSsBucketsSmall>
                 : nb <= kSsBucketsMid ? ss_scatter_kernel<kSsBucketsMid>
                                       : ss_scatter_kernel<kSsMaxBuckets>;
  scatter<<<nwg, kSsTileThreads, 0, c->stream>>>(B.key, B.perm, n, nb, tile, B.ss_spl, tot, cur,
                                                 B.ss_bstart, B.key_alt, B.perm_alt);
  // buckets in (key_alt, perm_alt); the sorted order goes back to (key, perm)
  ss_bucket_kernel<<<nb, kSsBucketThreads, 0, c->stream>>>(B.key_alt, B.perm_alt, nb, B.ss_bstart, tot, cur,
                                                     B.ss_spl, c->sort_cap, B.key, B.perm);
  // the next sort's splitters are read off the sorted keys by the build's
  // tie kernel, which runs next
  c->ss_have_spl = true;
}

void launch_keys_count(stellar_ctx* c, int l1, int ibits, bool hilbert) {
  Buffers& B = c->b;
  const int32_t n = c->n;
  const int32_t nb = sample_sort_buckets(n);
  auto k = hilbert ? (nb <= kSsBucketsSmall ? ss_keys_count_kernel<kSsBucketsSmall, true>
                      : nb <= kSsBucketsMid ? ss_keys_count_kernel<kSsBucketsMid, true>
                                            : ss_keys_count_kernel<kSsMaxBuckets, true>)
                   : (nb <= kSsBucketsSmall ? ss_keys_count_kernel<kSsBucketsSmall, false>
                      : nb <= kSsBucketsMid ? ss_keys_count_kernel<kSsBucketsMid, false>
                                            : ss_keys_count_kernel<kSsMaxBuckets, false>);
  k<<<sample_sort_tiles(n), kSsTileThreads, 0, c->stream>>>(B.x, B.y, n, B.inst_of, B.ibbox, l1,
                                                            ibits, nb, sample_sort_tile(n),
                                                            B.ss_spl, B.key, B.perm, B.ss_tot);
}

}  // namespace stellar
