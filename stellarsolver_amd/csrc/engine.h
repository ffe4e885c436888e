// Internal definitions shared by the libstellar_hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "stellar.h"

namespace stellar {

// Sticky device error bits (ctx->dev_err), mapped to status codes on sync.
enum : uint32_t {
  kErrPool = 1u,    // cell count exceeded the node capacity
  kErrStack = 2u,   // walk stack overflow
  kErrTieRun = 4u,  // more than kMaxTieRun bodies share a 32-level key
  kErrBucket = 8u,  // more than 4 bodies share a level-64 cell
};
// Errors after which the current tree must not be walked or summed.
constexpr uint32_t kBuildErrors = kErrPool | kErrTieRun | kErrBucket;

constexpr int kMaxLevel = STELLAR_MAX_LEVEL;
constexpr int kBboxMaxBlocks = 1024;
constexpr int64_t kMaxBlocks = int64_t(1) << 26;  // root + cells
constexpr int kMaxTieRun = 256;
// Key sort (sort.hip): sample sort for up to kSsMaxN bodies, buckets of at
// most kSsCapMax keys sorted in one LDS pass.
constexpr int64_t kSsMaxN = (int64_t)1 << 24;
constexpr int kSsBucket = 1536;  // target keys per bucket
constexpr int kSsMaxBuckets = (int)((kSsMaxN + kSsBucket - 1) / kSsBucket);
constexpr int kSsCapMax = 2048;
constexpr int kWalkStack = 256;  // >= 4 + 3*(kMaxLevel-1)
constexpr int kWalkBlock = 256;  // threads per walk workgroup (4 wavefronts)

// Opening thresholds of the group walk (kernels.cu:349,376,394-395): the
// root's children are pushed with depth0 = radius^2/theta_div, radius = half
// the x extent of the instance's box, and each level down multiplies it by
// 0.25, so a cell at level L >= 2 is accepted by a body when
// depth0 * 0.25^(L-1) <= d^2 + eps2 (theta_div = theta; theta^2 for the
// textbook test, which compares the unsoftened distance: threshold + eps2).
// The build and the centre-of-mass pass store this value beside each child
// record for the walks that read a second block per node; the arithmetic is
// the walk's own, so the values are the same bits.
__device__ __forceinline__ float walk_root_depth(float4 bb, float theta_div) {
  const float radius = (float)(0.5 * (double)(bb.y - bb.x));
  return (float)((double)(radius * radius) / (double)theta_div);
}
__device__ __forceinline__ float walk_level_threshold(float depth0, int level, float eps2,
                                                      bool unsoft) {
  float thr = depth0;
  for (int i = 1; i < level; ++i) thr = 0.25f * thr;
  return unsoft ? thr + eps2 : thr;
}

// One level of box subdivision along one axis (kernels.cu:144-157): the
// FP32 sum lo+hi is halved in double, p < mid is decided in double and the
// new bound is mid rounded to FP32.  When |lo+hi| >= 2^-125 the halving is
// exact in FP32, so the FP32 path gives the identical decision and bound;
// the double path runs only for the lanes that need it, behind a
// wavefront-uniform test (a scalar branch, no exec-mask bookkeeping on the
// common path).
__device__ __forceinline__ uint32_t split1(float p, float& lo, float& hi) {
  const float s = lo + hi;
  float mid = s * 0.5f;
  bool lt = p < mid;
  const bool exact = fabsf(s) >= 0x1p-125f;  // false for NaN too
  if (__builtin_expect(__builtin_amdgcn_ballot_w64(!exact) != 0, 0)) {
    if (!exact) {
      const double md = 0.5 * (double)s;
      lt = (double)p < md;
      mid = (float)md;
    }
  }
  hi = lt ? mid : hi;
  lo = lt ? lo : mid;
  return lt ? 1u : 0u;
}

// Slot code of the next level: +1 west (x < mid), +2 south (y < mid).
__device__ __forceinline__ uint32_t split2(float px, float py, float& l,
                                           float& r, float& b, float& t) {
  const uint32_t cx = split1(px, l, r);
  const uint32_t cy = split1(py, b, t);
  return cx | (cy << 1);
}

// Hilbert order of a node's children (STELLAR_FLAG_HILBERT_ORDER).  Frames:
// base curve SW, NW, NE, SE; frame bit 0 transposes, bit 1 turns by 180
// degrees (the two commute, so frames compose by XOR); the child at curve
// position h gets frame ^ child(h), child = {1, 0, 0, 3}.
//   kHilbertNext: 4 bits per (frame f, slot code c) at 4 (4f + c): the
//     curve position h (bits 0-1) and the child's frame (bits 2-3);
//   kHilbertSlot: 2 bits per (frame f, curve position h) at 2 (4f + h): the
//     slot code;
//   kHilbertChild: 2 bits per h: the frame change.
constexpr uint64_t kHilbertNext = 0xe3d8a97c05b64f12ull;
constexpr uint32_t kHilbertSlot = 0xb4784b87u;
constexpr uint32_t kHilbertChild = 0xc1u;
__host__ __device__ inline uint32_t hilbert_step(uint32_t& frame, uint32_t slot) {
  const uint32_t v = (uint32_t)(kHilbertNext >> (4 * ((frame << 2) | slot))) & 15u;
  frame = v >> 2;
  return v & 3u;
}
__host__ __device__ inline uint32_t hilbert_slot(uint32_t frame, uint32_t h) {
  return (kHilbertSlot >> (2 * ((frame << 2) | h))) & 3u;
}
__host__ __device__ inline uint32_t hilbert_child_frame(uint32_t frame, uint32_t h) {
  return frame ^ ((kHilbertChild >> (2 * h)) & 3u);
}
// Levels whose children follow the Hilbert curve (deeper ones: slot order).
// Every context's key holds at least 16 levels (l1 >= 16), so a tree's order
// does not depend on the number of instances beside it; below the walk
// groups' scale the order changes no group's shape.
constexpr int kHilbertLevels = 16;

// Sort key of one body: the slot codes of levels 1..l1 below the instance's
// box (HILBERT: those of levels 1..kHilbertLevels replaced by their positions
// along each node's Hilbert curve, so that sorting orders the bodies along a
// Hilbert curve of the same tree), with
// the instance index in the top ibits bits (ibits = 0: one instance, l1 = 32).
template <bool HILBERT>
__device__ __forceinline__ uint64_t path_key(float px, float py, float4 bb, int32_t inst, int l1,
                                             int ibits) {
  float l = bb.x, r = bb.y, b = bb.z, t = bb.w;
  uint64_t k = 0;
  uint32_t frame = 0;
#pragma unroll 4
  for (int lev = 0; lev < l1; ++lev) {
    uint32_t c = split2(px, py, l, r, b, t);
    if (HILBERT && lev < kHilbertLevels) c = hilbert_step(frame, c);
    k = (k << 2) | c;
  }
  if (ibits > 0) k |= (uint64_t)inst << (64 - ibits);
  return k;
}

// MI355X dispatches the workgroups of a launch round-robin over its 8 XCDs
// (workgroup b to XCD b % 8), each XCD with its own L2.  xcd_block maps
// hardware workgroup ids to logical ones so that each XCD gets one
// contiguous range of logical workgroups: neighbouring walk groups (nearby
// bodies, mostly the same tree nodes) then share one L2.  A bijection on
// [0, nb) for any nb.
constexpr int kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t q = nb / kXcds, r = nb % kXcds;
  const uint32_t x = b % kXcds, i = b / kXcds;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Metadata of a cell, one 64-bit word per cell block (written with one store
// by the build): its first tree position (30 bits; n <= 2^30), its parent's
// block (26 bits; blocks < kMaxBlocks = 2^26), its slot in the parent (2 bits)
// and its level - 1 (6 bits; levels 1 .. 64).
__host__ __device__ inline uint64_t cell_meta(int32_t start, int32_t parent, int slot,
                                              int level) {
  return (uint64_t)(uint32_t)start | ((uint64_t)(uint32_t)parent << 30) |
         ((uint64_t)(slot & 3) << 56) | ((uint64_t)(level - 1) << 58);
}
__host__ __device__ inline int32_t meta_start(uint64_t m) { return (int32_t)(m & 0x3fffffffu); }
__host__ __device__ inline int32_t meta_parent(uint64_t m) {
  return (int32_t)((m >> 30) & 0x3ffffffu);
}
__host__ __device__ inline int meta_slot(uint64_t m) { return (int)((m >> 56) & 3u); }
__host__ __device__ inline int meta_level(uint64_t m) { return (int)(m >> 58) + 1; }

// FP64 moments per cell: sum m, sum m x, sum m y, Qxx, Qxy, Qyy, bmax
constexpr int kSumStride = 7;

// Child record of a tree block: the walk's only view of the tree.  A block is
// the four child slots of one node (block 0 = root, block 1+c = cell c),
// 64 bytes, loaded whole with one scalar load per visited node.
//   x, y : centre of mass (cell) or position (body)
//   m    : mass
//   code : 0 empty, -1-s body at tree position s, 1+c cell c
struct alignas(16) ChildRec {
  float x, y, m;
  int32_t code;
};

// Device scalars, one 256-byte line.
struct DevScalars {
  uint32_t err;          // sticky error bits
  uint32_t pad0;
  int32_t ncells;        // cells of the current tree (all instances)
  int32_t pad1;
  float bbox[4];         // unused (per-instance boxes live in Buffers::ibbox)
  unsigned long long interactions;       // body-node interactions of the last walk
  unsigned long long quad_interactions;  // of which with a cell's quadrupole term
  uint32_t wq[8];  // walk work queues: next unit of each XCD's range (zeroed before each walk)
  int32_t com_len[4];  // centre-of-mass pass leftovers (zeroed by each build)
  uint32_t pad[40];
};
static_assert(sizeof(DevScalars) == 256, "one 256-byte line");

// Positions per bounding-box tile (one workgroup; tiles never straddle two
// instances).
constexpr int kBboxTile = 4096;
// FP64 energy: slices of 256 targets per bounding-box tile.
constexpr int kEnergySlices = kBboxTile / 256;
// FP64 sums per slice: potential, kinetic, px, py, Lz.
constexpr int kEnergyVals = 5;

struct Buffers {
  // body state in tree order of the last sort (double-buffered for the sort)
  float *x = nullptr, *y = nullptr, *vx = nullptr, *vy = nullptr,
        *m = nullptr;
  int32_t* id = nullptr;
  float *x2 = nullptr, *y2 = nullptr, *vx2 = nullptr, *vy2 = nullptr,
        *m2 = nullptr;
  int32_t* id2 = nullptr;
  float *ax = nullptr, *ay = nullptr;
  int32_t* inter = nullptr;  // per-body interaction count (tree order)
  // build
  uint64_t *key = nullptr, *key_alt = nullptr;  // level 1..32 codes
  uint32_t *perm = nullptr, *perm_alt = nullptr;
  uint64_t* ss_spl = nullptr;   // sample sort: splitters
  uint32_t* ss_tot = nullptr;     // sample sort: bucket totals, then fill levels
  uint32_t* ss_bstart = nullptr;  // bucket starts (nb + 1)
  uint64_t* key2 = nullptr;   // 32 deeper levels (valid inside tie runs)
  uint32_t* key3 = nullptr;   // the deepest levels when the key holds instances
  int32_t* ncell = nullptr;   // cells opened by each adjacent pair
  int32_t* coff = nullptr;    // exclusive scan of ncell
  int32_t* bsum = nullptr;    // cells opened per block of 256 positions, then their scan
  uint8_t* lam = nullptr;     // common-prefix depth of each adjacent pair
  uint8_t* lmin = nullptr;    // minimum tree over lam (levels 1..top)
  void* tmp = nullptr;        // rocprim scratch
  size_t tmp_bytes = 0;
  // cells
  ChildRec* blk = nullptr;    // (1+cap) blocks of 4 records
  uint64_t* cmeta = nullptr;  // per cell block: cell_meta (roots hold none)
  double* csum = nullptr;     // kSumStride FP64 moments per cell
  float4* qblk = nullptr;     // per child record: Qxx, Qxy, Qyy, 0
  int32_t* spine = nullptr;   // CoM pass leftovers, kMaxLevel slots per chunk
  int32_t* scount = nullptr;  // leftovers per chunk
  float* out = nullptr;       // 2*(n+cap) interleaved output buffer
  float4* bbox_part = nullptr;  // per-tile bounding-box partials
  // instances: a context advances ninst independent simulations whose
  // bodies occupy consecutive position ranges [ioff[i], ioff[i+1]); block i
  // is the root of instance i and block ninst + c is cell c.
  int32_t* ioff = nullptr;      // ninst + 1 offsets
  int32_t* inst_of = nullptr;   // instance of each position (and body id)
  float4* ibbox = nullptr;      // per instance: left, right, bottom, top
  int32_t* tstart = nullptr;    // bbox tiles: ntiles + 1 starts
  int32_t* itile = nullptr;     // first tile of each instance, ninst + 1
  int32_t* gtab[3] = {};        // walk groups of 64, 32, 16: ngroups + 1 starts
  DevScalars* dev = nullptr;
  DevScalars* host_dev = nullptr;  // pinned mirror
};

enum Stage : int {
  kNone = 0,
  kBboxDone = 1,
  kBuildDone = 2,
  kComDone = 3,
  kSortDone = 4,
  kForcesDone = 5,
  kUpdateDone = 6,
};

}  // namespace stellar

struct stellar_ctx {
  stellar_params p;
  int32_t n = 0;
  int64_t cap = 0;
  int32_t ninst = 1;
  std::vector<int32_t> ioff;    // host copy of the instance offsets
  std::vector<int32_t> itile;   // host copy of each instance's first tile
  int32_t ntiles = 0;
  int64_t ngroups[3] = {};      // walk groups of 64, 32, 16
  int device = 0;
  hipStream_t stream = nullptr;
  stellar::Buffers b;
  int stage = stellar::kNone;
  bool sorted_storage = false;  // storage order == tree order of the build
  bool uploaded = false;
  hipEvent_t ev[8] = {};
  // captured two-step graphs, keyed by the buffer pointers at capture time
  // (every step swaps the body, key and permutation buffers once, so two
  // steps return them to where they were)
  struct Graph {
    const void* key[3] = {};
    hipGraphExec_t exec = nullptr;
    uint64_t used = 0;
  } graphs[4];
  uint64_t graph_clock = 0;
  bool capturing = false;
  // walk launch shape (group 64): nodes popped per memory round trip and
  // wavefronts per workgroup (stellar_autotune)
  int walk_kpop = 0;   // 0: default (2 one-block nodes, 1 two-block nodes)
  int walk_waves = 0;  // 0: default (8)
  // walk scheduling: 0 one launch over static XCD ranges; 1 persistent
  // workgroups claiming units of one group per wavefront from per-XCD queues;
  // 2 the same with each wavefront claiming its own groups
  int walk_queue = 0;
  int32_t sort_cap = stellar::kSsCapMax;  // keys per block pass of a sort bucket
  bool ss_have_spl = false;  // splitters from the previous sort are in place
  bool radix_sort = false;   // testing: the device-wide radix sort at any size
  // per-step stage events of the current stellar_run (8 per step)
  std::vector<hipEvent_t> evrun;
  int32_t evrun_steps = 0;
  float stage_ms[7] = {};
  // stellar_render's device image (counts, then greys), grown on demand
  void* render_buf = nullptr;
  int64_t render_bytes = 0;
  uint8_t stage_empty = 0;  // stages of the last step that launched no kernel (bit k)
  std::string err;
};

namespace stellar {

// Kernel launchers (stream-ordered, no host synchronisation).
void launch_bbox(stellar_ctx* c);
void launch_build(stellar_ctx* c);
void launch_com(stellar_ctx* c);
void launch_sort(stellar_ctx* c);
// The walk; with_update: also the step's update, fused into the walk when the
// walk allows it (returns true then; false: launch_update is still needed).
bool launch_forces(stellar_ctx* c, bool zero_counter, bool with_update = false);
// Key sort (sort.hip).
bool sample_sort_fits(int32_t n);
int32_t sample_sort_buckets(int32_t n);
int32_t sample_sort_tiles(int32_t n);
void launch_sample_sort(stellar_ctx* c, bool keys_counted);
void launch_keys_count(stellar_ctx* c, int l1, int ibits, bool hilbert);
void launch_update(stellar_ctx* c);
void launch_output(stellar_ctx* c);
void launch_zero_counter(stellar_ctx* c);
// Per-pixel body counts of the current positions into dcounts (width x height,
// cleared first) and, when dgrey is given, their 8-bit greys.
void launch_render(stellar_ctx* c, const float view[4], int32_t width, int32_t height,
                   uint32_t* dcounts, uint8_t* dgrey);
void launch_init_scalars(stellar_ctx* c);
size_t build_scratch_bytes(int32_t n);
size_t mintree_bytes(int32_t n);
void launch_direct_fp64(stellar_ctx* c, int32_t nt, const int32_t* dtargets,
                        double eps2, double* dax, double* day);
void launch_energy_fp64(stellar_ctx* c, double eps2, double half_dt, bool sync,
                        double* dpart);
int energy_parts(stellar_ctx* c);
// Leftover slots and per-chunk counts of the centre-of-mass passes for a
// pool of cap cells (chunks of each of the three list passes in ch).
struct ComSlots {
  int64_t ch[3];
  int64_t slots;
  int64_t counts;
};
ComSlots com_slots(int64_t cap);
// Workgroups of a kernel that are resident at once on the whole device.
int resident_blocks(const void* kernel, int threads);

}  // namespace stellar
