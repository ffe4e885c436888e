// Bottom-up centre of mass (and optional quadrupole moments) for gfx950.
//
// The reference accumulates sum(m*x), sum(m*y), sum(m) into every cell on a
// body's insertion path with FP32 atomics (kernels.cu:180-183, 221-250) and
// divides afterwards (centre_of_mass_kernel, kernels.cu:275-288); the order
// of the atomics, and so the rounding, changes from run to run, and every
// level-1 cell takes about N/4 atomics.
//
// Here each cell is summed once from its four children, in slot order, in
// FP64, and the FP32 centre of mass is the FP64 quotient rounded once, so the
// result does not depend on scheduling.
//
// Scheduling.  Cells are numbered in DFS pre-order, so a cell's subtree is a
// contiguous id range that ends at the next cell of the same or a lower level,
// and its children have larger ids.  A pass
// takes a sorted list of unfinished cells in chunks of kChunk; a workgroup
// stages its chunk's child records and child moments in LDS and finishes
// every cell whose unfinished descendants all lie in the chunk, level by
// level from the deepest (a cell's children are one level deeper), the
// cells of one level spread densely over the workgroup's threads (no
// device-scope fences: on gfx950 each one writes back the issuing XCD's
// L2).  The cells it cannot
// finish are ancestors of the first cell after the chunk -- at most one per
// level -- and form the next pass's list.  Pass 1 runs over all cells; after
// a few passes the list is short and one workgroup finishes it, walking its
// chunks from the last to the first so that every descendant is done before
// its ancestors.
#include "engine.h"

namespace stellar {

namespace {

constexpr int kChunk = 256;          // cells per workgroup and chunk
constexpr int kComThreads = 256;
constexpr int kSlots = kMaxLevel;    // leftover cells per chunk (one per level)
constexpr int kMom = kSumStride;     // FP64 moments kept in global memory
// In LDS a finished cell also keeps its FP64 centre of mass (the quotient its
// parent would otherwise recompute): m, mx, my, Qxx, Qxy, Qyy, bmax, cx, cy.
constexpr int kMomL = kMom + 2;

struct ComArgs {
  ChildRec* blk;
  float4* qblk;
  const uint64_t* cmeta;
  double* csum;  // kSumStride per block
  DevScalars* dev;
  int quad;
  int32_t nroot;  // root blocks (instances); cell c is block nroot + c
  int bmax;       // per-cell opening thresholds bmax^2/t2 + eps2 in qblk .w
  double t2;      // theta^2
  double eps2;
  // mixed precision: qblk .x/.y hold Qh = Q / (m ws^2) as FP16 bits (Qxx | Qxy
  // << 16, Qyy), ws = S / 2^L with S the larger extent of the instance's box
  int mixed;
  const int32_t* inst_of;
  const float4* ibbox;
  // the walk's level threshold of a cell, stored beside its quadrupole
  // moments (walk_level_threshold)
  float theta_div;
  float eps2f;
  int unsoft;
};

// FP16 bits of v (round to nearest even), clamped to the finite range.
__device__ __forceinline__ uint32_t half_bits(float v) {
  v = v > 65504.f ? 65504.f : (v < -65504.f ? -65504.f : v);
  return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v);
}

template <bool EXT>
struct ChunkLds {
  int32_t ids[kChunk];          // block ids, ascending
  int32_t pidx[kChunk];         // chunk index of the parent, -1 if none
  int16_t bylev[kChunk];        // local entries grouped by level, deepest first
  uint8_t lev[kChunk];          // level of each entry
  uint8_t local[kChunk];        // every unfinished descendant is in the chunk
  int16_t cidx[kChunk][4];      // chunk index of each cell child, -1 if none
  ChildRec rec[kChunk][4];
  double mom[kChunk][kMomL];    // moments of the entry once finished
  // moments of cell children finished before this chunk (EXT only)
  double ext[EXT ? kChunk : 1][4][kMomL];
  int32_t wcnt[kComThreads / 64];
  int32_t wmin[kComThreads / 64];
  int32_t nout;
  int32_t lvcnt[kMaxLevel + 1];  // local entries per level (levels 1..kMaxLevel)
  int32_t lvoff[kMaxLevel + 1];  // their offset in bylev
  uint8_t lvlist[kMaxLevel + 1]; // levels present, deepest first
  int32_t nlev;
};
static_assert(kChunk == kComThreads, "one chunk entry per thread");
static_assert(kChunk <= 32767, "chunk indices are int16");

// Index of block id b among ids[lo, hi), or -1.
__device__ __forceinline__ int find_id(const int32_t* ids, int lo, int hi, int32_t b) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int32_t v = ids[mid];
    if (v == b) return mid;
    if (v < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  return -1;
}

// Moments of entry i from its child slots (slot order, FP64) into L.mom[i]
// (publish_cell writes them out).  Every slot is read the same way,
// without branches: a cell child's moments come from L.mom (finished in this
// chunk) or L.ext (before it), a body's are (m, m x, m y, 0, 0, 0, 0, x, y)
// from its record, and an empty slot, whose record is all zeros, gives
// zeros, which leave every sum unchanged (no sum here is ever -0).
template <bool EXT>
__device__ void finish_cell(const ComArgs& A, ChunkLds<EXT>& L, int i) {
  const int32_t b = L.ids[i];
  double cm[4], cmx[4], cmy[4], cqxx[4], cqxy[4], cqyy[4], cbm[4], ccx[4], ccy[4];
  bool used[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const ChildRec r = L.rec[i][k];
    const int ci = L.cidx[i][k];
    const double* src = (!EXT || ci >= 0) ? L.mom[ci < 0 ? 0 : ci] : L.ext[EXT ? i : 0][k];
    const bool cell = r.code > 0;
    used[k] = r.code != 0;
    const double m = (double)r.m, x = (double)r.x, y = (double)r.y;
    cm[k] = cell ? src[0] : m;
    cmx[k] = cell ? src[1] : m * x;
    cmy[k] = cell ? src[2] : m * y;
    cqxx[k] = cell ? src[3] : 0.0;
    cqxy[k] = cell ? src[4] : 0.0;
    cqyy[k] = cell ? src[5] : 0.0;
    cbm[k] = cell ? src[6] : 0.0;
    ccx[k] = cell ? src[7] : x;
    ccy[k] = cell ? src[8] : y;
  }
  double sm = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    sm += cm[k];
    sx += cmx[k];
    sy += cmy[k];
  }
  double out[kMomL];
  out[0] = sm;
  out[1] = sx;
  out[2] = sy;
  // centre of mass as sums times the reciprocal mass (one FP64 division)
  const double inv = sm != 0.0 ? 1.0 / sm : 0.0;
  const double ox = sx * inv, oy = sy * inv;
  double qxx = 0.0, qxy = 0.0, qyy = 0.0;
  if (A.quad) {
    // traceless in-plane quadrupole about the FP64 centre of mass with the
    // parallel-axis shift of each child (operation order of the CPU
    // restatement, so the FP32 results agree bit for bit)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double dx = ccx[k] - ox, dy = ccy[k] - oy;
      const double t1 = dx * dx, t2 = dy * dy;
      qxx += cqxx[k] + cm[k] * (2.0 * t1 - t2);
      qxy += cqxy[k] + cm[k] * (3.0 * (dx * dy));
      qyy += cqyy[k] + cm[k] * (2.0 * t2 - t1);
    }
  }
  double bm = 0.0;
  if (A.bmax) {
    // bmax bound: the largest distance from the FP64 centre of mass to a
    // body, through the children (bodies contribute their distance)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double dx = ccx[k] - ox, dy = ccy[k] - oy;
      const double d = sqrt(__builtin_fma(dx, dx, dy * dy)) + cbm[k];
      bm = used[k] && d > bm ? d : bm;
    }
  }
  out[3] = qxx;
  out[4] = qxy;
  out[5] = qyy;
  out[6] = bm;
  out[7] = ox;
  out[8] = oy;
#pragma unroll
  for (int t = 0; t < kMomL; ++t) L.mom[i][t] = out[t];
}

// Writes finished entry i out, by the thread that staged it (block b, child
// slot s of parent block p): its child record {CoM, m, b} in the parent's
// block, {Qxx, Qxy, Qyy} beside it (with the bmax threshold in bmax mode;
// otherwise the level threshold the build put there stays), and its FP64 moments
// when a parent finishing in a later pass or chunk will read them (a parent
// finished in this chunk took the LDS copy).
template <bool EXT>
__device__ void publish_cell(const ComArgs& A, const ChunkLds<EXT>& L, int i, int32_t b,
                             int32_t p, int s) {
  const double* o = L.mom[i];
  A.blk[4 * (int64_t)p + s] = ChildRec{(float)o[7], (float)o[8], (float)o[0], b};
  if (A.quad || A.bmax) {
    // second block: the moments the walk needs and the cell's opening
    // threshold, one 16-byte store
    const float4 bb = A.ibbox[A.inst_of[meta_start(A.cmeta[b])]];
    const int lev = L.lev[i];
    const float thr =
        A.bmax ? (float)((o[6] * o[6]) / A.t2 + A.eps2)
               : walk_level_threshold(walk_root_depth(bb, A.theta_div), lev, A.eps2f, A.unsoft != 0);
    float4 q = make_float4((float)o[3], (float)o[4], (float)o[5], thr);
    if (A.mixed) {
      // Qh of a cell at level lev: Q / (m ws^2), ws^2 = S^2 / 4^lev in FP32
      const float w = bb.y - bb.x, h = bb.w - bb.z;
      const float sz = w > h ? w : h;
      const float ws2 = __builtin_ldexpf(sz * sz, -2 * lev);
      const double den = o[0] * (double)ws2;
      const float qxx = den != 0.0 ? (float)(o[3] / den) : 0.f;
      const float qxy = den != 0.0 ? (float)(o[4] / den) : 0.f;
      const float qyy = den != 0.0 ? (float)(o[5] / den) : 0.f;
      q.x = __uint_as_float(half_bits(qxx) | (half_bits(qxy) << 16));
      q.y = __uint_as_float(half_bits(qyy));
      q.z = 0.f;
    }
    A.qblk[4 * (int64_t)p + s] = q;
  }
  const int pi = L.pidx[i];
  if (pi < 0 || !L.local[pi]) {
    double* g = A.csum + kSumStride * (int64_t)b;
#pragma unroll
    for (int t = 0; t < kMom; ++t) g[t] = o[t];
  }
}

// One chunk: list positions [p0, p1) of `list` (nullptr: block id = pos + nroot),
// `next` = block id at p1 (INT32_MAX if none or if every later cell is
// already finished).  Leftover cells go to out_ids[0..*out_n).
//
// Every list is closed under taking parents (a cell is unfinished while any
// descendant is), and in such a subsequence of pre-order the first entry
// after a cell's descendants has a level no deeper than the cell's.  So the
// unfinished descendants of entry i all lie in the chunk exactly when some
// later entry of the chunk, or `next`, has level <= level(i).
//
// Everything the finishing steps need is staged in LDS first, so they issue
// global stores only and never wait on global memory.
template <bool EXT>
__device__ void process_chunk(const ComArgs& A, ChunkLds<EXT>& L, const int32_t* list,
                              int32_t p0, int32_t p1, int32_t next,
                              int32_t* out_ids, int32_t* out_n) {
  const int ne = p1 - p0;
  const int i = threadIdx.x;  // one entry per thread (kChunk == blockDim)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int32_t pb = 0;  // parent block and slot of entry i
  int sl = 0;
  {
    int lev = 255;
    if (i < ne) {
      const int32_t b = list ? list[p0 + i] : p0 + i + A.nroot;
      L.ids[i] = b;
      const uint64_t meta = A.cmeta[b];
      lev = meta_level(meta);
      L.lev[i] = (uint8_t)lev;
      pb = meta_parent(meta);
      sl = meta_slot(meta);
      const ChildRec* r = A.blk + 4 * (int64_t)b;
#pragma unroll
      for (int k = 0; k < 4; ++k) L.rec[i][k] = r[k];
    }
    // suffix minimum of the levels: within the wavefront, then across
    int suf = lev;
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_down(suf, off);
      if (lane + off < 64) suf = min(suf, o);
    }
    if (lane == 0) L.wmin[wv] = suf;
    int after = __shfl_down(suf, 1);  // entries i+1 .. end of wavefront
    if (lane == 63) after = 255;
    __syncthreads();
    for (int w = wv + 1; w < kComThreads / 64; ++w) after = min(after, L.wmin[w]);
    const int nlev = next == INT32_MAX ? -1 : meta_level(A.cmeta[next]);
    after = min(after, nlev);
    if (i < ne) L.local[i] = after <= lev ? 1 : 0;
  }
  __syncthreads();
  {
    const bool live = i < ne;
    const int32_t b = live ? L.ids[i] : 0;
    bool left = false;
    if (live) {
      // a contiguous chunk (the first pass) maps block ids to indices directly
      const int32_t base = p0 + A.nroot;
      const int32_t pr = pb - base;
      L.pidx[i] = list ? find_id(L.ids, 0, i, pb) : (pr >= 0 && pr < i ? pr : -1);
      if (L.local[i]) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int32_t code = L.rec[i][k].code;
          int ci = -1;
          if (code > 0) {
            const int32_t cr = code - base;
            ci = list ? find_id(L.ids, i + 1, ne, code) : (cr > i && cr < ne ? cr : -1);
            if (ci < 0 && EXT) {
              // finished by an earlier pass or a later chunk
              const double* g = A.csum + kSumStride * (int64_t)code;
              double* e = L.ext[EXT ? i : 0][k];
#pragma unroll
              for (int t = 0; t < kMom; ++t) e[t] = g[t];
              // its centre of mass, as it computed it: sums times 1/m
              const double inv = g[0] != 0.0 ? 1.0 / g[0] : 0.0;
              e[7] = g[1] * inv;
              e[8] = g[2] * inv;
            }
          }
          L.cidx[i][k] = (int16_t)ci;
        }
      } else {
        left = true;
      }
    }
    // leftovers keep ascending id order: rank = leftovers before entry i
    const uint64_t bal = __builtin_amdgcn_ballot_w64(left);
    if (lane == 0) L.wcnt[wv] = __popcll(bal);
    __syncthreads();
    int rank = __popcll(bal & ((1ull << lane) - 1ull));
    for (int w = 0; w < wv; ++w) rank += L.wcnt[w];
    if (left && rank < kSlots) out_ids[rank] = b;
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int w = 0; w < kComThreads / 64; ++w) tot += L.wcnt[w];
      L.nout = tot;
    }
    for (int l = threadIdx.x; l <= kMaxLevel; l += kComThreads) L.lvcnt[l] = 0;
  }
  __syncthreads();
  // group the local entries by level (counting sort in LDS)
  int rk = 0;
  const bool fin = i < ne && L.local[i];
  if (fin) rk = atomicAdd(&L.lvcnt[L.lev[i]], 1);
  __syncthreads();
  static_assert(kMaxLevel == 64, "one lane per cell level 1..64");
  if (threadIdx.x < 64) {
    // lane j holds level 64 - j (cells have levels 1..64): offsets are an
    // exclusive scan in lane order, the present levels a ballot
    const int l = kMaxLevel - lane;
    const int c = L.lvcnt[l];
    int incl = c;
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(incl, off);
      if (lane >= off) incl += o;
    }
    L.lvoff[l] = incl - c;
    const uint64_t present = __builtin_amdgcn_ballot_w64(c > 0);
    if (c > 0) L.lvlist[__popcll(present & ((1ull << lane) - 1ull))] = (uint8_t)l;
    if (lane == 0) L.nlev = __popcll(present);
  }
  __syncthreads();
  if (fin) L.bylev[L.lvoff[L.lev[i]] + rk] = (int16_t)i;
  __syncthreads();
  // deepest level first: every child of a cell is then finished
  const int nl = L.nlev;
  for (int k = 0; k < nl; ++k) {
    const int l = L.lvlist[k];
    const int cnt = L.lvcnt[l], off = L.lvoff[l];
    for (int t = threadIdx.x; t < cnt; t += kComThreads) finish_cell<EXT>(A, L, L.bylev[off + t]);
    __syncthreads();
  }
  if (fin) publish_cell<EXT>(A, L, i, L.ids[i], pb, sl);
  if (threadIdx.x == 0) *out_n = L.nout < kSlots ? L.nout : kSlots;
}

// A later pass (and its compaction) is skipped once an earlier list fits in
// one chunk: the last pass then takes that list (com_final_kernel).
__device__ __forceinline__ bool list_done(const int32_t* a, const int32_t* b) {
  return (a && *a <= kChunk) || (b && *b <= kChunk);
}

// Pass over list positions: one chunk per workgroup.  list == nullptr: all
// cells (length dev->ncells); else the list compacted by the previous pass.
template <bool EXT>
__global__ __launch_bounds__(kComThreads) void com_pass_kernel(
    ComArgs A, const int32_t* __restrict__ list, const int32_t* __restrict__ len,
    int32_t* __restrict__ out_slots, int32_t* __restrict__ out_counts,
    const int32_t* __restrict__ skip_a, const int32_t* __restrict__ skip_b) {
  __shared__ ChunkLds<EXT> L;
  if (A.dev->err & kBuildErrors) return;  // no complete tree to sum
  if (list_done(skip_a, skip_b)) return;
  const int32_t m = list ? *len : A.dev->ncells;
  const int32_t p0 = blockIdx.x * kChunk;
  if (p0 >= m) return;
  const int32_t p1 = p0 + kChunk < m ? p0 + kChunk : m;
  const int32_t next = p1 < m ? (list ? list[p1] : p1 + A.nroot) : INT32_MAX;
  process_chunk<EXT>(A, L, list, p0, p1, next, out_slots + (int64_t)blockIdx.x * kSlots,
                out_counts + blockIdx.x);
}

// Compaction of a pass's leftovers into a dense list, in two launches:
// offsets (one workgroup scans the per-chunk counts) and gather (one
// wavefront per chunk).  Chunks are in ascending id order and each chunk's
// leftovers were written in ascending order, so the list is sorted.
__global__ __launch_bounds__(1024) void com_offsets_kernel(
    const ComArgs A, const int32_t* __restrict__ len_in, int32_t nchunks_max,
    const int32_t* __restrict__ counts, int32_t* __restrict__ offs,
    int32_t* __restrict__ len_out, const int32_t* __restrict__ skip_a,
    const int32_t* __restrict__ skip_b) {
  __shared__ int32_t part[1024];
  if (A.dev->err & kBuildErrors) return;
  if (list_done(skip_a, skip_b)) return;
  const int32_t m = len_in ? *len_in : A.dev->ncells;
  int32_t nchunks = (m + kChunk - 1) / kChunk;
  if (nchunks > nchunks_max) nchunks = nchunks_max;
  const int32_t per = (nchunks + blockDim.x - 1) / blockDim.x;
  const int32_t c0 = threadIdx.x * per;
  const int32_t c1 = c0 + per < nchunks ? c0 + per : nchunks;
  int32_t tot = 0;
  for (int32_t c = c0; c < c1; ++c) tot += counts[c];
  part[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan
    const int32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int32_t pos = part[threadIdx.x] - tot;
  for (int32_t c = c0; c < c1; ++c) {
    offs[c] = pos;
    pos += counts[c];
  }
  if (threadIdx.x == blockDim.x - 1) *len_out = part[threadIdx.x];
}

__global__ __launch_bounds__(64) void com_gather_kernel(
    const ComArgs A, const int32_t* __restrict__ len_in,
    const int32_t* __restrict__ slots, const int32_t* __restrict__ counts,
    const int32_t* __restrict__ offs, int32_t* __restrict__ list,
    const int32_t* __restrict__ skip_a, const int32_t* __restrict__ skip_b) {
  if (A.dev->err & kBuildErrors) return;
  if (list_done(skip_a, skip_b)) return;
  const int32_t m = len_in ? *len_in : A.dev->ncells;
  const int32_t c = blockIdx.x;
  if (c * kChunk >= m) return;
  if ((int)threadIdx.x < counts[c])
    list[offs[c] + threadIdx.x] = slots[(int64_t)c * kSlots + threadIdx.x];
}

// Last pass: one workgroup finishes the whole list, chunks from last to
// first (every descendant of a chunk's cells is then in the chunk or done).
// The list is the first of the pass-1, pass-2 leftovers that fits in one
// chunk (the passes after it were skipped), else the pass-3 leftovers.
__global__ __launch_bounds__(kComThreads) void com_final_kernel(
    ComArgs A, const int32_t* __restrict__ l0, const int32_t* __restrict__ l1,
    const int32_t* __restrict__ lens, int32_t* __restrict__ scratch_slots,
    int32_t* __restrict__ scratch_count) {
  __shared__ ChunkLds<true> L;
  if (A.dev->err & kBuildErrors) return;
  const bool short0 = lens[0] <= kChunk, short1 = lens[1] <= kChunk;
  const int32_t* list = short0 ? l0 : short1 ? l1 : l0;
  const int32_t m = short0 ? lens[0] : short1 ? lens[1] : lens[2];
  const int32_t nch = (m + kChunk - 1) / kChunk;
  for (int32_t c = nch - 1; c >= 0; --c) {
    const int32_t p0 = c * kChunk;
    const int32_t p1 = p0 + kChunk < m ? p0 + kChunk : m;
    process_chunk<true>(A, L, list, p0, p1, INT32_MAX, scratch_slots, scratch_count);
    __syncthreads();
  }
}

}  // namespace

// Worst-case list lengths: pass 1 leaves at most kSlots cells per chunk.
int64_t com_list_capacity(int64_t cap) {
  return ((cap + kChunk - 1) / kChunk) * (int64_t)kSlots + kChunk;
}

void launch_com(stellar_ctx* c) {
  if (c->n == 0 || c->cap == 0) return;
  Buffers& B = c->b;
  ComArgs A{B.blk,  B.qblk, B.cmeta, B.csum, B.dev,
            (c->p.flags & STELLAR_FLAG_QUADRUPOLE) ? 1 : 0, c->ninst,
            (c->p.flags & STELLAR_FLAG_BMAX_OPENING) ? 1 : 0,
            (double)c->p.theta * (double)c->p.theta, (double)c->p.eps2,
            c->p.precision == STELLAR_PRECISION_MIXED ? 1 : 0, B.inst_of, B.ibbox,
            (c->p.flags & STELLAR_FLAG_TEXTBOOK_OPENING) ? c->p.theta * c->p.theta : c->p.theta,
            c->p.eps2, (c->p.flags & STELLAR_FLAG_TEXTBOOK_OPENING) ? 1 : 0};
  // pass 1 over all cells; each later list holds at most kSlots cells per
  // chunk of the pass before, so grids are sized from that bound and chunks
  // beyond the actual length exit at once
  const int64_t ch1 = (c->cap + kChunk - 1) / kChunk;
  const int64_t ch2 = (ch1 * kSlots + kChunk - 1) / kChunk;
  const int64_t ch3 = (ch2 * kSlots + kChunk - 1) / kChunk;
  int32_t* l0 = B.clist;
  int32_t* l1 = B.clist + com_list_capacity(c->cap);
  hipStream_t s = c->stream;
  // a pass runs only while every list before it is longer than one chunk
  auto compact = [&](const int32_t* len_in, int64_t ch, int32_t* list, int32_t* len_out,
                     const int32_t* ska, const int32_t* skb) {
    com_offsets_kernel<<<1, 1024, 0, s>>>(A, len_in, (int32_t)ch, B.scount, B.choff, len_out,
                                          ska, skb);
    com_gather_kernel<<<(unsigned)ch, 64, 0, s>>>(A, len_in, B.spine, B.scount, B.choff, list,
                                                  ska, skb);
  };
  const int32_t* c0 = B.clen + 0;
  const int32_t* c1 = B.clen + 1;
  com_pass_kernel<false><<<(unsigned)ch1, kComThreads, 0, s>>>(A, nullptr, nullptr, B.spine,
                                                               B.scount, nullptr, nullptr);
  compact(nullptr, ch1, l0, B.clen + 0, nullptr, nullptr);
  com_pass_kernel<true><<<(unsigned)ch2, kComThreads, 0, s>>>(A, l0, c0, B.spine, B.scount,
                                                              c0, nullptr);
  compact(c0, ch2, l1, B.clen + 1, c0, nullptr);
  com_pass_kernel<true><<<(unsigned)ch3, kComThreads, 0, s>>>(A, l1, c1, B.spine, B.scount,
                                                              c0, c1);
  compact(c1, ch3, l0, B.clen + 2, c0, c1);
  com_final_kernel<<<1, kComThreads, 0, s>>>(A, l0, l1, B.clen, B.spine, B.scount);
}

}  // namespace stellar
