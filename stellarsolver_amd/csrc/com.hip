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

constexpr int kChunk = 256;          // cells per workgroup and chunk (at most)
constexpr int kComThreads = 256;
// Chunking.  A pass cuts its list into chunks of at most kChunk entries.
// Chunk w starts near position kStride w: at the first entry of the lowest
// level among positions [kStride w, kStride w + kWin), so that few of its
// cells are ancestors of the entry after it (the chunk's leftovers: at most
// one per level above that entry's).  Neighbouring chunks choose their
// shared boundary from the same positions.
constexpr int kStride = 248;
constexpr int kWin = 8;
static_assert(kStride + kWin == 256, "a chunk and its next boundary window fit one window");
static_assert(kWin <= 32, "one half-wavefront per boundary window");
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
  // the walk's level threshold of a cell (walk_level_threshold), stored
  // beside its moments in mixed precision
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
  uint8_t lev[kChunk];          // level of each entry
  uint8_t local[kChunk];        // every unfinished descendant is in the chunk
  int16_t bylev[kChunk];        // local entries grouped by level, deepest first
  int16_t cidx[kChunk][4];      // chunk index of each cell child, -1 if none
  ChildRec rec[kChunk][4];
  union {
    double mom[kChunk][kMomL];  // moments of the entry once finished
    uint64_t wmeta[kChunk];     // metadata of the window's entries (staging)
  };
  // moments of cell children finished before this chunk (EXT only)
  double ext[EXT ? kChunk : 1][4][kMomL];
  int32_t wcnt[kComThreads / 64];
  int32_t wmin[kComThreads / 64];
  int32_t nout;
  uint32_t levmask[2];           // levels of the local entries (bit l-1 for level l)
  int32_t lvcnt[kMaxLevel + 1];  // local entries per level (levels 1..kMaxLevel)
  int32_t lvoff[kMaxLevel + 1];  // their offset in bylev
  int32_t blo, bhi;              // the chunk within its window
};
static_assert(kChunk == kComThreads, "one chunk entry per thread");

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

// Indices of four block ids among ids[lo, hi) (ascending, distinct), -1 for
// an absent one: four branch-free binary searches side by side, so that each
// step issues four LDS reads and waits once.
__device__ __forceinline__ void find_ids4(const int32_t* ids, int lo, int hi, const int32_t key[4],
                                          int out[4]) {
  int base[4] = {lo, lo, lo, lo};
  int n = hi - lo;
  if (n <= 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = -1;
    return;
  }
  while (n > 1) {
    const int half = n >> 1;
    int32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = ids[base[k] + half];
#pragma unroll
    for (int k = 0; k < 4; ++k) base[k] = v[k] < key[k] ? base[k] + half : base[k];
    n -= half;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int idx = ids[base[k]] < key[k] ? base[k] + 1 : base[k];
    out[k] = idx < hi && ids[idx] == key[k] ? idx : -1;
  }
}

// Moments of entry i from its child slots (slot order, FP64) into L.mom[i]
// (publish_cell writes them out): a cell child's moments come from L.mom
// (finished in this chunk) or L.ext (before it), a body's are
// (m, m x, m y, 0, 0, 0, 0, x, y) from its record, and an empty slot, whose
// record is all zeros, gives zeros, which leave every sum unchanged (no sum
// here is ever -0).
template <bool EXT>
__device__ void finish_cell(const ComArgs& A, ChunkLds<EXT>& L, int i, const ChildRec rr[4],
                            const int cis[4]) {
  double cm[4], cmx[4], cmy[4], cqxx[4], cqxy[4], cqyy[4], cbm[4], ccx[4], ccy[4];
  bool used[4];
  // The entry's child records and chunk indices come in registers; the nine
  // moments of every slot are read before any is used (one wait), each from
  // a row chosen without branching (a cell child's row, finished in this
  // chunk (L.mom) or before it (L.ext); row 0, unused, for bodies and empty
  // slots).
  double mv[4][kMomL];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int ci = cis[k];
    const double* src = (!EXT || ci >= 0 || rr[k].code <= 0) ? L.mom[ci < 0 ? 0 : ci]
                                                             : L.ext[EXT ? i : 0][k];
#pragma unroll
    for (int t = 0; t < kMomL; ++t) mv[k][t] = src[t];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const ChildRec r = rr[k];
    const bool cell = r.code > 0;
    used[k] = r.code != 0;
    if (cell) {
      cm[k] = mv[k][0];
      cmx[k] = mv[k][1];
      cmy[k] = mv[k][2];
      cqxx[k] = mv[k][3];
      cqxy[k] = mv[k][4];
      cqyy[k] = mv[k][5];
      cbm[k] = mv[k][6];
      ccx[k] = mv[k][7];
      ccy[k] = mv[k][8];
    } else {
      // a body (an empty slot's record is all zeros and gives zeros)
      const double m = (double)r.m, x = (double)r.x, y = (double)r.y;
      cm[k] = m;
      cmx[k] = m * x;
      cmy[k] = m * y;
      cqxx[k] = 0.0;
      cqxy[k] = 0.0;
      cqyy[k] = 0.0;
      cbm[k] = 0.0;
      ccx[k] = x;
      ccy[k] = y;
    }
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
// block, {Qxx, Qxy, Qyy, w} beside it (w: the quadrupole's softening term, or
// the cell's opening threshold with bmax or mixed precision), and its FP64 moments
// when a parent finishing in a later pass or chunk will read them (a parent
// finished in this chunk took the LDS copy).
template <bool EXT>
__device__ void publish_cell(const ComArgs& A, const ChunkLds<EXT>& L, int i, int32_t b,
                             int32_t p, int s, float4 bb) {
  const double* o = L.mom[i];
  A.blk[4 * (int64_t)p + s] = ChildRec{(float)o[7], (float)o[8], (float)o[0], b};
  if (A.quad || A.bmax) {
    // second block: the moments the walk needs and, with per-cell opening
    // thresholds (bmax) or mixed precision, the cell's threshold (bb: the box
    // of the cell's instance), one 16-byte store; an FP32 quadrupole walk
    // with level thresholds (it reads those from a per-level table) finds
    // the quadrupole's softening term -2.5 eps2 (Qxx + Qyy) there instead
    const int lev = L.lev[i];
    float4 q = make_float4((float)o[3], (float)o[4], (float)o[5], 0.f);
    if (A.bmax || A.mixed)
      q.w = A.bmax ? (float)((o[6] * o[6]) / A.t2 + A.eps2)
                   : walk_level_threshold(walk_root_depth(bb, A.theta_div), lev, A.eps2f,
                                          A.unsoft != 0);
    else
      q.w = (q.x + q.z) * (-2.5f * A.eps2f);
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

// Block ids of list positions [W0, W0 + kChunk) below m into win[], from a
// list held as a pass's per-chunk leftovers (kSlots per chunk, in chunk
// order; nch chunks): the per-chunk counts are scanned, four per thread per
// tile, until the window is passed, and each chunk copies the entries that
// fall in the window.
__device__ void gather_window(const int32_t* __restrict__ slots,
                              const int32_t* __restrict__ counts, int32_t nch, int32_t W0,
                              int32_t* win, int32_t* wsum) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int32_t base = 0;  // entries of the chunks before the tile (uniform)
  for (int32_t c0 = 0; c0 < nch && base < W0 + kChunk; c0 += 4 * kComThreads) {
    const int32_t c = c0 + 4 * (int32_t)threadIdx.x;
    int32_t n[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) n[j] = c + j < nch ? counts[c + j] : 0;
    const int32_t tot = n[0] + n[1] + n[2] + n[3];
    int32_t incl = tot;
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t o = __shfl_up(incl, off);
      if (lane >= off) incl += o;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int32_t pos = base + incl - tot, tile = 0;
    for (int w = 0; w < kComThreads / 64; ++w) {
      if (w < wv) pos += wsum[w];
      tile += wsum[w];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int32_t lo = pos > W0 ? pos : W0;
      const int32_t hi = pos + n[j] < W0 + kChunk ? pos + n[j] : W0 + kChunk;
      for (int32_t q = lo; q < hi; ++q) win[q - W0] = slots[(int64_t)(c + j) * kSlots + (q - pos)];
      pos += n[j];
    }
    base += tile;
    __syncthreads();  // wsum is rewritten by the next tile
  }
}

// A pass's input: every cell in id order (slots == nullptr), or the
// leftovers of pass k (len = k, 0-based): held per chunk of that pass, which
// cut a list of prod entries (the leftovers of pass prod, -1: every cell).
struct ListIn {
  const int32_t* slots;
  const int32_t* counts;
  int len;
  int prod;
};

// The device scalars the passes decide on, read together (one memory round
// trip at a kernel's start): error bits, cell count, pass list lengths.
struct ComState {
  uint32_t err;
  int32_t ncells;
  int32_t len[4];
  // length of list k (-1: every cell); k is uniform, so no indexed access
  __device__ __forceinline__ int32_t at(int k) const {
    return k < 0 ? ncells : k == 0 ? len[0] : k == 1 ? len[1] : k == 2 ? len[2] : len[3];
  }
  // a later pass is skipped once list a or b (-1: none) fits in one chunk:
  // the last pass then takes that list (com_final_kernel)
  __device__ __forceinline__ bool done(int a, int b) const {
    return (a >= 0 && at(a) <= kChunk) || (b >= 0 && at(b) <= kChunk);
  }
};
static_assert(offsetof(DevScalars, com_len) % 16 == 0, "one 16-byte load");

__device__ __forceinline__ ComState com_state(const DevScalars* d) {
  const int4 h = *reinterpret_cast<const int4*>(d);  // err, pad, ncells, pad
  const int4 l = *reinterpret_cast<const int4*>(d->com_len);
  return ComState{(uint32_t)h.x, h.z, {l.x, l.y, l.z, l.w}};
}

// Block ids of window positions [W0, W0 + kChunk) below m into L.pidx (used
// as scratch until the chunk is staged), their metadata into L.wmeta, their
// child records into L.rec and their levels into L.local (255 past the end).
template <bool EXT>
__device__ void load_window(const ComArgs& A, ChunkLds<EXT>& L, const ListIn& in, int32_t m,
                            int32_t pm, int32_t W0) {
  const int t = threadIdx.x;
  // the first pass (EXT false) always runs over every cell
  if (EXT && in.slots) {
    gather_window(in.slots, in.counts, (pm + kStride - 1) / kStride, W0, L.pidx, L.wcnt);
    __syncthreads();
  }
  const bool valid = W0 + t < m;
  const int32_t b = !valid ? 0 : (EXT && in.slots) ? L.pidx[t] : W0 + t + A.nroot;
  const uint64_t meta = valid ? A.cmeta[b] : 0;
  if (valid) {
    const ChildRec* r = A.blk + 4 * (int64_t)b;
#pragma unroll
    for (int k = 0; k < 4; ++k) L.rec[t][k] = r[k];
  }
  L.pidx[t] = b;
  L.wmeta[t] = meta;
  L.local[t] = valid ? (uint8_t)meta_level(meta) : (uint8_t)255;
  __syncthreads();
}

// One chunk, staged: thread i < ne holds entry i (block b, metadata meta),
// whose child records are L.rec[roff + i]; `contiguous` when the entries are consecutive block ids (the first pass);
// next_lev = level of the list entry after the chunk (-1 if none, or if
// every later cell is already finished).  Leftover cells go to
// out_ids[0..*out_n), and their number is added to *out_len.
//
// Every list is closed under taking parents (a cell is unfinished while any
// descendant is), and in such a subsequence of pre-order the first entry
// after a cell's descendants has a level no deeper than the cell's.  So the
// unfinished descendants of entry i all lie in the chunk exactly when some
// later entry of the chunk, or the next entry, has level <= level(i).
//
// Everything the finishing steps need is staged in LDS first, so they issue
// global stores only and never wait on global memory.
template <bool EXT>
__device__ void process_chunk(const ComArgs& A, ChunkLds<EXT>& L, int ne, int roff,
                              bool contiguous, int32_t b, uint64_t meta, int next_lev,
                              int32_t* out_ids, int32_t* out_n, int32_t* out_len) {
  const int i = threadIdx.x;  // one entry per thread (kChunk == blockDim)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int32_t pb = 0;  // parent block and slot of entry i
  int sl = 0;
  int rank = -1;   // rank among the leftovers, -1 if finished here
  float4 bb = make_float4(0.f, 0.f, 0.f, 0.f);

  {
    int lev = 255;
    if (i < ne) {
      L.ids[i] = b;
      lev = meta_level(meta);
      L.lev[i] = (uint8_t)lev;
      pb = meta_parent(meta);
      sl = meta_slot(meta);
      // the instance's box, for the mixed-precision cell's level threshold
      // and size (loaded here, so that the first barrier covers its latency)
      if (A.mixed) bb = A.ibbox[A.nroot == 1 ? 0 : A.inst_of[meta_start(meta)]];
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
    after = min(after, next_lev);
    if (i < ne) L.local[i] = after <= lev ? 1 : 0;
  }
  __syncthreads();
  {
    const bool live = i < ne;
    bool left = false;
    if (live) {
      // consecutive block ids map to chunk indices directly
      const int32_t base = L.ids[0];
      const int32_t pr = pb - base;
      L.pidx[i] = contiguous ? (pr >= 0 && pr < i ? pr : -1) : find_id(L.ids, 0, i, pb);
      if (L.local[i]) {
        bool ext[4];
        int32_t codes[4];
        int cis[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) codes[k] = L.rec[roff + i][k].code;
        if (contiguous) {
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int32_t cr = codes[k] - base;
            cis[k] = codes[k] > 0 && cr > i && cr < ne ? cr : -1;
          }
        } else {
          // bodies and empty slots (codes <= 0) match no block id
          find_ids4(L.ids, i + 1, ne, codes, cis);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int ci = codes[k] > 0 ? cis[k] : -1;
          // finished by an earlier pass or a later chunk
          ext[k] = EXT && codes[k] > 0 && ci < 0;
          L.cidx[i][k] = (int16_t)ci;
        }
        if (EXT && (ext[0] || ext[1] || ext[2] || ext[3])) {
          // the moments of such children, all four loads issued before any
          // is used
          double gv[4][kMom];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            if (ext[k]) {
              const double* g = A.csum + kSumStride * (int64_t)codes[k];
#pragma unroll
              for (int t = 0; t < kMom; ++t) gv[k][t] = g[t];
            }
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            if (ext[k]) {
              double* e = L.ext[EXT ? i : 0][k];
#pragma unroll
              for (int t = 0; t < kMom; ++t) e[t] = gv[k][t];
              // its centre of mass, as it computed it: sums times 1/m
              const double inv = gv[k][0] != 0.0 ? 1.0 / gv[k][0] : 0.0;
              e[7] = gv[k][1] * inv;
              e[8] = gv[k][2] * inv;
            }
          }
        }
      } else {
        left = true;
      }
    }
    // leftovers keep ascending id order: rank = leftovers before entry i
    // (written out at the end: a global store before a barrier would make
    // every wavefront wait for it)
    const uint64_t bal = __builtin_amdgcn_ballot_w64(left);
    if (lane == 0) L.wcnt[wv] = __popcll(bal);
    __syncthreads();
    rank = __popcll(bal & ((1ull << lane) - 1ull));
    for (int w = 0; w < wv; ++w) rank += L.wcnt[w];
    rank = left ? rank : -1;
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int w = 0; w < kComThreads / 64; ++w) tot += L.wcnt[w];
      L.nout = tot;
    }
    if (threadIdx.x < 2) L.levmask[threadIdx.x] = 0u;
    for (int l = threadIdx.x; l <= kMaxLevel; l += kComThreads) L.lvcnt[l] = 0;
  }
  __syncthreads();
  // group the local entries by level, deepest first (counting sort in LDS),
  // and note the levels present
  const bool fin = i < ne && L.local[i];
  const int mylev = fin ? (int)L.lev[i] : 0;
  int rk = 0;
  if (fin) {
    rk = atomicAdd(&L.lvcnt[mylev], 1);
    atomicOr(&L.levmask[(mylev - 1) >> 5], 1u << ((mylev - 1) & 31));
  }
  __syncthreads();
  static_assert(kMaxLevel == 64, "one lane per cell level 1..64");
  if (threadIdx.x < 64) {
    // lane j holds level 64 - j (cells have levels 1..64): offsets are an
    // exclusive scan in lane order
    const int l = kMaxLevel - lane;
    const int c = L.lvcnt[l];
    int incl = c;
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(incl, off);
      if (lane >= off) incl += o;
    }
    L.lvoff[l] = incl - c;
    if (lane == 63) L.lvcnt[0] = incl;  // every local entry
  }
  __syncthreads();
  if (fin) L.bylev[L.lvoff[mylev] + rk] = (int16_t)i;
  __syncthreads();
  // Thread t finishes the t-th entry of that order at its level's step, so
  // that the cells of one level sit on consecutive threads (few wavefronts
  // per step); its child records and child indices are read once, here.
  int e = -1, le = 0;
  ChildRec rr[4] = {};
  int cs[4] = {-1, -1, -1, -1};
  if (i < L.lvcnt[0]) {
    e = L.bylev[i];
    le = L.lev[e];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      rr[k] = L.rec[roff + e][k];
      cs[k] = L.cidx[e][k];
    }
  }
  uint64_t levs = (uint64_t)L.levmask[0] | ((uint64_t)L.levmask[1] << 32);
  // deepest level first: every child of a cell is then finished
  while (levs) {
    const int l = 64 - __clzll(levs);
    levs &= ~(1ull << (l - 1));
    if (e >= 0 && le == l) finish_cell<EXT>(A, L, e, rr, cs);
    __syncthreads();
  }
  if (fin) publish_cell<EXT>(A, L, i, L.ids[i], pb, sl, bb);
  if (rank >= 0 && rank < kSlots) out_ids[rank] = b;
  if (threadIdx.x == 0) {
    const int nout = L.nout < kSlots ? L.nout : kSlots;
    *out_n = nout;
    if (out_len && nout > 0) atomicAdd(out_len, nout);
  }
}

// Pass over a list: chunk w (one workgroup at a time, grid-stride) covers
// the window of positions [kStride w, kStride w + kChunk), cut as described
// at kStride.  Its leftovers go to out_slots[w kSlots ..], their number to
// out_counts[w] and the pass's total to *out_len.
template <bool EXT>
__device__ void pass_chunk(const ComArgs& A, ChunkLds<EXT>& L, const ListIn& in, int32_t m,
                           int32_t pm, int32_t w, int32_t* out_slots, int32_t* out_counts,
                           int32_t* out_len) {
  const int32_t W0 = w * kStride;
  load_window<EXT>(A, L, in, m, pm, W0);
  // chunk boundaries: the first position of the lowest level in each
  // boundary window (half a wavefront each)
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x, j = lane & 31;
    const int t = lane < 32 ? j : kStride + j;
    int key = ((j < kWin ? (int)L.local[t] : 255) << 8) | j;
    for (int off = 1; off < 32; off <<= 1) key = min(key, __shfl_xor(key, off));
    if (lane == 0) L.blo = w == 0 ? 0 : (key & 31);
    if (lane == 32) L.bhi = W0 + kStride >= m ? m - W0 : kStride + (key & 31);
  }
  __syncthreads();
  const int blo = L.blo, bhi = L.bhi;
  const int ne = bhi - blo;
  const int next_lev = W0 + bhi < m ? (int)L.local[bhi] : -1;
  const bool live = (int)threadIdx.x < ne;
  const int32_t b = live ? L.pidx[blo + threadIdx.x] : 0;
  const uint64_t meta = live ? L.wmeta[blo + threadIdx.x] : 0;
  __syncthreads();  // the window's LDS is reused by the chunk
  process_chunk<EXT>(A, L, ne, blo, !EXT, b, meta, next_lev, out_slots + (int64_t)w * kSlots,
                     out_counts + w, out_len);
}

// Pass over a list: chunk w covers the window of positions
// [kStride w, kStride w + kChunk), cut as described at kStride.  Its
// leftovers go to out_slots[w kSlots ..], their number to out_counts[w] and
// the pass's total to *out_len.  The first pass (every cell) runs one chunk
// per workgroup; the list passes, whose LDS allows one workgroup per CU,
// loop over chunks with a grid of the resident workgroups.
template <bool EXT>
__global__ __launch_bounds__(kComThreads) void com_pass_kernel(
    ComArgs A, ListIn in, int32_t* __restrict__ out_slots, int32_t* __restrict__ out_counts,
    int32_t* __restrict__ out_len, int skip_a, int skip_b) {
  __shared__ ChunkLds<EXT> L;
  const ComState S = com_state(A.dev);
  if (S.err & kBuildErrors) return;  // no complete tree to sum
  if (S.done(skip_a, skip_b)) return;
  const int32_t m = in.slots ? S.at(in.len) : S.ncells;
  const int32_t pm = S.at(in.prod);
  const int32_t nch = (m + kStride - 1) / kStride;
  if (!EXT) {
    if ((int32_t)blockIdx.x < nch)
      pass_chunk<EXT>(A, L, in, m, pm, blockIdx.x, out_slots, out_counts, out_len);
    return;
  }
  for (int32_t w = blockIdx.x; w < nch; w += gridDim.x) {
    pass_chunk<EXT>(A, L, in, m, pm, w, out_slots, out_counts, out_len);
    __syncthreads();
  }
}

// Last pass: one workgroup finishes a whole list, in windows of kChunk from
// the last to the first (every descendant of a window's cells is then in the
// window or done).  The list is the first of the pass-1, pass-2 leftovers
// that fits in one chunk (the passes after it were skipped), else the pass-3
// leftovers.
__global__ __launch_bounds__(kComThreads) void com_final_kernel(
    ComArgs A, ListIn l1, ListIn l2, ListIn l3, int32_t* __restrict__ scratch_slots,
    int32_t* __restrict__ scratch_count) {
  __shared__ ChunkLds<true> L;
  const ComState S = com_state(A.dev);
  if (S.err & kBuildErrors) return;
  const ListIn& in = S.len[0] <= kChunk ? l1 : S.len[1] <= kChunk ? l2 : l3;
  const int32_t m = S.at(in.len), pm = S.at(in.prod);
  const int32_t nw = (m + kChunk - 1) / kChunk;
  for (int32_t c = nw - 1; c >= 0; --c) {
    const int32_t W0 = c * kChunk;
    load_window<true>(A, L, in, m, pm, W0);
    const int ne = m - W0 < kChunk ? m - W0 : kChunk;
    const bool live = (int)threadIdx.x < ne;
    const int32_t b = live ? L.pidx[threadIdx.x] : 0;
    const uint64_t meta = live ? L.wmeta[threadIdx.x] : 0;
    __syncthreads();
    process_chunk<true>(A, L, ne, 0, false, b, meta, -1, scratch_slots, scratch_count, nullptr);
    __syncthreads();
  }
}

}  // namespace

// Leftover slots of the three list passes (kSlots per chunk) and their
// per-chunk counts, for a pool of cap cells.
ComSlots com_slots(int64_t cap) {
  ComSlots s;
  s.ch[0] = (cap + kStride - 1) / kStride;
  s.ch[1] = (s.ch[0] * kSlots + kStride - 1) / kStride;
  s.ch[2] = (s.ch[1] * kSlots + kStride - 1) / kStride;
  s.slots = (s.ch[0] + s.ch[1] + s.ch[2]) * kSlots + kSlots;  // + the final pass's scratch
  s.counts = s.ch[0] + s.ch[1] + s.ch[2] + 1;
  return s;
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
  const ComSlots cs = com_slots(c->cap);
  int32_t* sl[3] = {B.spine, B.spine + cs.ch[0] * kSlots,
                    B.spine + (cs.ch[0] + cs.ch[1]) * kSlots};
  int32_t* ct[3] = {B.scount, B.scount + cs.ch[0], B.scount + cs.ch[0] + cs.ch[1]};
  // the passes' list lengths (device scalars, zeroed by the build's link kernel)
  int32_t* len = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(B.dev) +
                                            offsetof(DevScalars, com_len));
  hipStream_t s = c->stream;
  // grids: pass 1 one workgroup per chunk of the pool; the list passes the
  // workgroups resident at once, grid-stride over their chunks
  const int64_t r2 = resident_blocks(reinterpret_cast<const void*>(&com_pass_kernel<true>),
                                     kComThreads);
  const unsigned g1 = (unsigned)cs.ch[0];
  const unsigned g2 = (unsigned)std::min<int64_t>(cs.ch[1], r2);
  const unsigned g3 = (unsigned)std::min<int64_t>(cs.ch[2], r2);
  const ListIn all{nullptr, nullptr, -1, -1};
  const ListIn p1{sl[0], ct[0], 0, -1};
  const ListIn p2{sl[1], ct[1], 1, 0};
  const ListIn p3{sl[2], ct[2], 2, 1};
  com_pass_kernel<false><<<g1, kComThreads, 0, s>>>(A, all, sl[0], ct[0], len + 0, -1, -1);
  com_pass_kernel<true><<<g2, kComThreads, 0, s>>>(A, p1, sl[1], ct[1], len + 1, 0, -1);
  com_pass_kernel<true><<<g3, kComThreads, 0, s>>>(A, p2, sl[2], ct[2], len + 2, 0, 1);
  com_final_kernel<<<1, kComThreads, 0, s>>>(A, p1, p2, p3, sl[2] + cs.ch[2] * kSlots,
                                             ct[2] + cs.ch[2]);
}

}  // namespace stellar
