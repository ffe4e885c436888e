// libstellar_hip C ABI: context, buffers, stage sequencing, status codes.
//
// Replaces the launch wrappers of Particle_cuda.cu:8-58 and the buffer
// management of BarnesHutParticleSystem.cpp:28-84, 249-292.
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

#include "engine.h"

using namespace stellar;

namespace {

const char* status_text(int s) {
  switch (s) {
    case STELLAR_OK: return "ok";
    case STELLAR_ERR_ARG: return "invalid argument";
    case STELLAR_ERR_OOM: return "out of memory";
    case STELLAR_ERR_POOL: return "node pool overflow (more cells than node_capacity)";
    case STELLAR_ERR_STACK: return "walk stack overflow";
    case STELLAR_ERR_DEPTH: return "too many bodies share one level-32 cell";
    case STELLAR_ERR_DEVICE: return "HIP runtime error";
    case STELLAR_ERR_STATE: return "stage called out of order";
    case STELLAR_ERR_BUCKET: return "more than 4 bodies share a level-64 cell";
    case STELLAR_ERR_SIZE: return "input larger than the allocation-free entry point's maximum";
    default: return "unknown status";
  }
}

int fail(stellar_ctx* c, int status, const char* what) {
  if (c) {
    c->err = std::string(status_text(status)) + ": " + what;
  }
  return status;
}

int hip_fail(stellar_ctx* c, hipError_t e, const char* what) {
  if (c) {
    c->err = std::string("HIP runtime error in ") + what + ": " +
             hipGetErrorString(e);
  }
  return STELLAR_ERR_DEVICE;
}

#define HIP_TRY(ctx, expr)                                   \
  do {                                                       \
    hipError_t e_ = (expr);                                  \
    if (e_ != hipSuccess) return hip_fail((ctx), e_, #expr); \
  } while (0)

bool params_ok(const stellar_params* p) {
  if (!p) return false;
  if (!(p->theta > 0.f) || !(p->eps2 > 0.f)) return false;
  if (p->group != 64 && p->group != 32 && p->group != 16) return false;
  if (p->precision != STELLAR_PRECISION_FP32 && p->precision != STELLAR_PRECISION_MIXED)
    return false;
  // the mixed far field is the quadrupole term of group-64 walks
  if (p->precision == STELLAR_PRECISION_MIXED &&
      (p->group != 64 || !(p->flags & STELLAR_FLAG_QUADRUPOLE)))
    return false;
  // the reference's lone-root-body handling is a level-threshold walk
  if ((p->flags & STELLAR_FLAG_REFERENCE_ROOT) &&
      (p->flags & (STELLAR_FLAG_QUADRUPOLE | STELLAR_FLAG_BMAX_OPENING)))
    return false;
  return true;
}

template <typename T>
hipError_t dmalloc(T** p, size_t count) {
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T) + 256);
}

void free_all(stellar_ctx* c) {
  Buffers& B = c->b;
  void* ptrs[] = {B.x,     B.y,      B.vx,      B.vy,     B.m,      B.id,
                  B.x2,    B.y2,     B.vx2,     B.vy2,    B.m2,     B.id2,
                  B.ax,    B.ay,     B.inter,   B.key,    B.key_alt, B.perm,
                  B.perm_alt, B.key2, B.ncell,  B.coff,   B.tmp,    B.blk, B.bsum,
                  B.cmeta, B.lam,
                  B.csum,  B.out,    B.dev,
                  B.bbox_part, B.scount,
                  B.lmin,  B.qblk,  B.spine,
                  B.key3,  B.ioff,  B.inst_of, B.ibbox, B.tstart, B.itile,
                  B.gtab[0], B.gtab[1], B.gtab[2], B.ss_spl, B.ss_tot, B.ss_bstart};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (B.host_dev) (void)hipHostFree(B.host_dev);
  B = Buffers{};
}

// Read the sticky error word (synchronises the stream).
int check_device(stellar_ctx* c) {
  HIP_TRY(c, hipMemcpyAsync(c->b.host_dev, c->b.dev, sizeof(DevScalars),
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const uint32_t e = c->b.host_dev->err;
  if (e == 0) return STELLAR_OK;
  // clear so that a later, independent call is not reported again
  HIP_TRY(c, hipMemsetAsync(&c->b.dev->err, 0, sizeof(uint32_t), c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->stage = kNone;
  if (e & kErrPool) return fail(c, STELLAR_ERR_POOL, "raise node_capacity");
  if (e & kErrTieRun) return fail(c, STELLAR_ERR_DEPTH, "tie run too long");
  if (e & kErrBucket) return fail(c, STELLAR_ERR_BUCKET, "coincident bodies");
  if (e & kErrStack) return fail(c, STELLAR_ERR_STACK, "walk stack");
  return fail(c, STELLAR_ERR_DEVICE, "unknown device error");
}

bool timing(const stellar_ctx* c) {
  return (c->p.flags & (STELLAR_FLAG_STAGE_TIMING | STELLAR_FLAG_FORCES_TIMING)) != 0;
}

// Only the events around the forces stage (marks 4 and 5) are recorded.
bool forces_only(const stellar_ctx* c) {
  return (c->p.flags & STELLAR_FLAG_STAGE_TIMING) == 0 &&
         (c->p.flags & STELLAR_FLAG_FORCES_TIMING) != 0;
}

// Events of step `step` of the current run (step < 0: the single-step set).
hipEvent_t* step_events(stellar_ctx* c, int32_t step) {
  return step < 0 ? c->ev : c->evrun.data() + 8 * (size_t)step;
}

void mark(stellar_ctx* c, hipEvent_t* ev, int k) {
  if (ev && (!forces_only(c) || k == 4 || k == 5)) (void)hipEventRecord(ev[k], c->stream);
}

int run_step(stellar_ctx* c, int32_t step) {
  hipEvent_t* ev = timing(c) && !c->capturing ? step_events(c, step) : nullptr;
  mark(c, ev, 0);
  launch_bbox(c);
  mark(c, ev, 1);
  launch_build(c);
  mark(c, ev, 2);
  launch_com(c);
  mark(c, ev, 3);
  launch_sort(c);
  mark(c, ev, 4);
  // the build zeroed the interaction counter; the update runs inside the
  // walk when the walk allows it
  const bool updated = launch_forces(c, false, true);
  mark(c, ev, 5);
  if (!updated) launch_update(c);
  mark(c, ev, 6);
  // stages that launched nothing report zero time (the sort stage swaps
  // buffers, the output buffer is filled when read)
  c->stage_empty = (uint8_t)((1u << 3) | (updated ? 1u << 5 : 0u) | (1u << 6));
  // the output buffer is filled when it is read (stellar_download_positions)
  mark(c, ev, 7);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(c, e, "kernel launch");
  c->stage = kUpdateDone;
  c->sorted_storage = true;
  return STELLAR_OK;
}

void drop_graphs(stellar_ctx* c) {
  for (auto& g : c->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = stellar_ctx::Graph{};
  }
}

// The two-step graph for the current buffer pointers, captured on first use.
int two_step_graph(stellar_ctx* c, hipGraphExec_t* out) {
  const Buffers& B = c->b;
  const void* key[3] = {B.x, B.key, B.perm};
  stellar_ctx::Graph* slot = &c->graphs[0];
  for (auto& g : c->graphs) {
    if (g.exec && g.key[0] == key[0] && g.key[1] == key[1] && g.key[2] == key[2]) {
      g.used = ++c->graph_clock;
      *out = g.exec;
      return STELLAR_OK;
    }
    if (!g.exec || g.used < slot->used) slot = &g;
  }
  if (slot->exec) (void)hipGraphExecDestroy(slot->exec);
  *slot = stellar_ctx::Graph{};
  HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  int rc = run_step(c, -1);
  if (rc == STELLAR_OK) rc = run_step(c, -1);
  c->capturing = false;
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
  // two steps swapped every double buffer twice: the pointers are back
  if (B.x != key[0] || B.key != key[1] || B.perm != key[2]) {
    (void)hipGraphDestroy(graph);
    return fail(c, STELLAR_ERR_DEVICE, "buffers did not return after two captured steps");
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return hip_fail(c, ei, "hipGraphInstantiate");
  for (int k = 0; k < 3; ++k) slot->key[k] = key[k];
  slot->exec = exec;
  slot->used = ++c->graph_clock;
  *out = exec;
  return STELLAR_OK;
}

// Make room for `steps` sets of stage events.
int reserve_run_events(stellar_ctx* c, int32_t steps) {
  const size_t need = 8 * (size_t)steps;
  while (c->evrun.size() < need) {
    hipEvent_t e;
    HIP_TRY(c, hipEventCreate(&e));
    c->evrun.push_back(e);
  }
  return STELLAR_OK;
}

// Mean per-stage milliseconds over steps [0, steps) (steps < 0: single set).
int collect_times(stellar_ctx* c, int32_t steps) {
  if (!timing(c)) return STELLAR_OK;
  const int32_t nsets = steps < 0 ? 1 : steps;
  if (nsets == 0) return STELLAR_OK;
  double acc[7] = {};
  const bool fo = forces_only(c);
  for (int32_t s = 0; s < nsets; ++s) {
    hipEvent_t* ev = step_events(c, steps < 0 ? -1 : s);
    for (int k = 0; k < 7; ++k) {
      if (fo && k != 4) continue;
      float ms = 0.f;
      HIP_TRY(c, hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
      acc[k] += ms;
    }
  }
  for (int k = 0; k < 7; ++k)
    c->stage_ms[k] = (c->stage_empty >> k) & 1u ? 0.f : (float)(acc[k] / nsets);
  return STELLAR_OK;
}

}  // namespace

extern "C" {

void stellar_default_params(stellar_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->theta = 0.5f;
  p->eps2 = 0.025f;
  p->timestep = 0.001f;
  p->dampening = 1.0f;
  p->gravity = 1.0f;
  p->group = 64;
  p->precision = STELLAR_PRECISION_FP32;
  p->flags = 0;
}

const char* stellar_status_string(int status) { return status_text(status); }

const char* stellar_last_error(const stellar_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

}  // extern "C"

namespace {

// Host tables of a context's instance layout: positions of instance i are
// [ioff[i], ioff[i+1]); bounding-box tiles of at most kBboxTile positions and
// walk groups of G positions never straddle two instances.
struct Layout {
  std::vector<int32_t> inst_of, tstart, itile, gtab[3];
};

Layout make_layout(const std::vector<int32_t>& ioff) {
  Layout L;
  const int32_t ninst = (int32_t)ioff.size() - 1;
  const int32_t n = ioff[(size_t)ninst];
  L.inst_of.resize((size_t)(n > 0 ? n : 1), 0);
  const int sizes[3] = {64, 32, 16};
  for (int32_t i = 0; i < ninst; ++i) {
    const int32_t lo = ioff[(size_t)i], hi = ioff[(size_t)i + 1];
    for (int32_t s = lo; s < hi; ++s) L.inst_of[(size_t)s] = i;
    L.itile.push_back((int32_t)L.tstart.size());
    for (int32_t s = lo; s < hi; s += kBboxTile) L.tstart.push_back(s);
    for (int k = 0; k < 3; ++k)
      for (int32_t s = lo; s < hi; s += sizes[k]) L.gtab[k].push_back(s);
  }
  L.itile.push_back((int32_t)L.tstart.size());
  L.tstart.push_back(n);
  for (int k = 0; k < 3; ++k) L.gtab[k].push_back(n);
  return L;
}

int create_ctx(const stellar_params* p, int32_t ninst, const int32_t* counts,
               int64_t node_capacity, stellar_ctx** out) {
  if (!out || ninst < 1 || ninst > 65536 || !counts || !params_ok(p)) return STELLAR_ERR_ARG;
  *out = nullptr;
  std::vector<int32_t> ioff(1, 0);
  for (int32_t i = 0; i < ninst; ++i) {
    if (counts[i] < (ninst > 1 ? 1 : 0)) return STELLAR_ERR_ARG;
    const int64_t next = (int64_t)ioff.back() + counts[i];
    if (next > INT32_MAX / 2) return STELLAR_ERR_ARG;
    ioff.push_back((int32_t)next);
  }
  const int32_t n = ioff.back();
  stellar_ctx* c = new (std::nothrow) stellar_ctx();
  if (!c) return STELLAR_ERR_OOM;
  c->p = *p;
  c->n = n;
  c->ninst = ninst;
  c->ioff = ioff;
  c->cap = node_capacity > 0 ? node_capacity : 2 * (int64_t)n + 12000 * (int64_t)ninst;
  // the walk's stack words keep a block index in 26 bits
  if (c->cap + ninst >= kMaxBlocks) {
    delete c;
    return STELLAR_ERR_ARG;
  }
  const Layout lay = make_layout(ioff);
  c->ntiles = (int32_t)lay.tstart.size() - 1;
  c->itile = lay.itile;
  for (int k = 0; k < 3; ++k) c->ngroups[k] = (int64_t)lay.gtab[k].size() - 1;
  (void)hipGetDevice(&c->device);
  Buffers& B = c->b;
  const size_t N = (size_t)(n > 0 ? n : 1);
  const size_t C = (size_t)c->cap;
  const size_t NB = C + (size_t)ninst;  // root blocks + cells
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t r) {
    if (e == hipSuccess && r != hipSuccess) e = r;
  };
  chk(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  chk(dmalloc(&B.x, N));
  chk(dmalloc(&B.y, N));
  chk(dmalloc(&B.vx, N));
  chk(dmalloc(&B.vy, N));
  chk(dmalloc(&B.m, N));
  chk(dmalloc(&B.id, N));
  chk(dmalloc(&B.x2, N));
  chk(dmalloc(&B.y2, N));
  chk(dmalloc(&B.vx2, N));
  chk(dmalloc(&B.vy2, N));
  chk(dmalloc(&B.m2, N));
  chk(dmalloc(&B.id2, N));
  chk(dmalloc(&B.ax, N));
  chk(dmalloc(&B.ay, N));
  chk(dmalloc(&B.inter, N));
  chk(dmalloc(&B.key, N));
  chk(dmalloc(&B.key_alt, N));
  chk(dmalloc(&B.perm, N));
  chk(dmalloc(&B.perm_alt, N));
  if (sample_sort_fits((int32_t)N)) {
    chk(dmalloc(&B.ss_spl, (size_t)kSsMaxBuckets));
    chk(dmalloc(&B.ss_tot, 2 * (size_t)kSsMaxBuckets));
    chk(dmalloc(&B.ss_bstart, (size_t)kSsMaxBuckets + 1));
  }
  chk(dmalloc(&B.key2, N));
  chk(dmalloc(&B.key3, N));
  chk(dmalloc(&B.ncell, N));
  chk(dmalloc(&B.coff, N));
  chk(dmalloc(&B.bsum, N / 256 + 2));
  chk(dmalloc(&B.lam, (size_t)N + 64));
  chk(dmalloc(&B.lmin, mintree_bytes((int32_t)N)));
  B.tmp_bytes = build_scratch_bytes((int32_t)N);
  chk(hipMalloc(&B.tmp, B.tmp_bytes));
  chk(dmalloc(&B.blk, 4 * NB));
  chk(dmalloc(&B.cmeta, NB));
  chk(dmalloc(&B.csum, kSumStride * NB));
  chk(dmalloc(&B.qblk, 4 * NB));
  const ComSlots cs = com_slots(c->cap);
  chk(dmalloc(&B.spine, (size_t)cs.slots));
  chk(dmalloc(&B.scount, (size_t)cs.counts));
  chk(dmalloc(&B.out, 2 * (N + C)));
  chk(dmalloc(&B.dev, 1));
  chk(dmalloc(&B.bbox_part, (size_t)c->ntiles));
  chk(dmalloc(&B.ioff, (size_t)ninst + 1));
  chk(dmalloc(&B.inst_of, N));
  chk(dmalloc(&B.ibbox, (size_t)ninst));
  chk(dmalloc(&B.tstart, lay.tstart.size()));
  chk(dmalloc(&B.itile, lay.itile.size()));
  for (int k = 0; k < 3; ++k) chk(dmalloc(&B.gtab[k], lay.gtab[k].size()));
  chk(hipHostMalloc(reinterpret_cast<void**>(&B.host_dev), sizeof(DevScalars),
                    hipHostMallocDefault));
  for (int k = 0; k < 8; ++k) chk(hipEventCreate(&c->ev[k]));
  if (e != hipSuccess) {
    free_all(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    for (int k = 0; k < 8; ++k)
      if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
    delete c;
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? STELLAR_ERR_OOM : STELLAR_ERR_DEVICE;
  }
  auto up = [&](void* dst, const void* src, size_t bytes) {
    chk(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  };
  up(B.ioff, ioff.data(), ioff.size() * 4);
  up(B.inst_of, lay.inst_of.data(), lay.inst_of.size() * 4);
  up(B.tstart, lay.tstart.data(), lay.tstart.size() * 4);
  up(B.itile, lay.itile.data(), lay.itile.size() * 4);
  for (int k = 0; k < 3; ++k) up(B.gtab[k], lay.gtab[k].data(), lay.gtab[k].size() * 4);
  // clear the output buffer, the tree records and the boxes once
  chk(hipMemsetAsync(B.out, 0, 2 * (N + C) * sizeof(float), c->stream));
  chk(hipMemsetAsync(B.blk, 0, 4 * NB * sizeof(ChildRec), c->stream));
  chk(hipMemsetAsync(B.qblk, 0, 4 * NB * sizeof(float4), c->stream));
  chk(hipMemsetAsync(B.ibbox, 0, (size_t)ninst * sizeof(float4), c->stream));
  // the key sort's bucket totals and fill levels start at zero (its bucket
  // kernel clears them again after each use)
  if (B.ss_tot) chk(hipMemsetAsync(B.ss_tot, 0, 2 * (size_t)kSsMaxBuckets * sizeof(uint32_t), c->stream));
  chk(hipMemsetAsync(B.ax, 0, N * sizeof(float), c->stream));
  chk(hipMemsetAsync(B.ay, 0, N * sizeof(float), c->stream));
  chk(hipMemsetAsync(B.inter, 0, N * sizeof(int32_t), c->stream));
  launch_init_scalars(c);
  chk(hipGetLastError());
  chk(hipStreamSynchronize(c->stream));
  if (e != hipSuccess) {
    int rc = hip_fail(c, e, "stellar_create");
    stellar_destroy(c);
    return rc;
  }
  *out = c;
  return STELLAR_OK;
}

}  // namespace

extern "C" {

int stellar_create(const stellar_params* p, int32_t n, int64_t node_capacity,
                   stellar_ctx** out) {
  if (n < 0) return STELLAR_ERR_ARG;
  return create_ctx(p, 1, &n, node_capacity, out);
}

int stellar_batch_create(const stellar_params* p, int32_t ninst, const int32_t* counts,
                         int64_t node_capacity, stellar_ctx** out) {
  return create_ctx(p, ninst, counts, node_capacity, out);
}

int32_t stellar_num_instances(const stellar_ctx* c) { return c ? c->ninst : -1; }

int stellar_instance_offsets(const stellar_ctx* c, int32_t* offsets) {
  if (!c || !offsets) return STELLAR_ERR_ARG;
  std::memcpy(offsets, c->ioff.data(), c->ioff.size() * sizeof(int32_t));
  return STELLAR_OK;
}

int stellar_destroy(stellar_ctx* c) {
  if (!c) return STELLAR_ERR_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graphs(c);
  free_all(c);
  if (c->render_buf) (void)hipFree(c->render_buf);
  for (int k = 0; k < 8; ++k)
    if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
  for (hipEvent_t e : c->evrun) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return STELLAR_OK;
}

int stellar_set_params(stellar_ctx* c, const stellar_params* p) {
  if (!c) return STELLAR_ERR_ARG;
  if (!params_ok(p)) return fail(c, STELLAR_ERR_ARG, "parameters");
  // the build and centre-of-mass stages store per-child opening thresholds,
  // the quadrupole softening term and moments that depend on these: a tree
  // built under the old values must not be walked under the new ones, so the
  // stages start again from bbox
  constexpr int32_t kTreeFlags = STELLAR_FLAG_QUADRUPOLE | STELLAR_FLAG_TEXTBOOK_OPENING |
                                 STELLAR_FLAG_BMAX_OPENING | STELLAR_FLAG_REFERENCE_ROOT |
                                 STELLAR_FLAG_HILBERT_ORDER;
  // the sample sort's splitters are keys of the previous order
  if ((p->flags ^ c->p.flags) & STELLAR_FLAG_HILBERT_ORDER) c->ss_have_spl = false;
  if (((p->flags ^ c->p.flags) & kTreeFlags) || p->theta != c->p.theta ||
      p->eps2 != c->p.eps2 || p->precision != c->p.precision)
    c->stage = kNone;
  c->p = *p;
  drop_graphs(c);  // captured kernels carry the old parameters
  return STELLAR_OK;
}

int stellar_get_params(const stellar_ctx* c, stellar_params* p) {
  if (!c || !p) return STELLAR_ERR_ARG;
  *p = c->p;
  return STELLAR_OK;
}

int32_t stellar_num_bodies(const stellar_ctx* c) { return c ? c->n : -1; }

int stellar_upload(stellar_ctx* c, const float* mass, const float* x,
                   const float* y, const float* vx, const float* vy) {
  if (!c) return STELLAR_ERR_ARG;
  const int32_t n = c->n;
  if (n > 0 && (!mass || !x || !y || !vx || !vy))
    return fail(c, STELLAR_ERR_ARG, "null body array");
  Buffers& B = c->b;
  const size_t bytes = (size_t)n * sizeof(float);
  std::vector<int32_t> ids((size_t)n);
  for (int32_t i = 0; i < n; ++i) ids[(size_t)i] = i;
  if (n > 0) {
    HIP_TRY(c, hipMemcpyAsync(B.m, mass, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(B.x, x, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(B.y, y, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(B.vx, vx, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(B.vy, vy, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(B.id, ids.data(), bytes, hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipMemsetAsync(B.ax, 0, bytes, c->stream));
    HIP_TRY(c, hipMemsetAsync(B.ay, 0, bytes, c->stream));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->uploaded = true;
  c->stage = kNone;
  c->sorted_storage = false;
  c->ss_have_spl = false;  // the key sort samples the new state
  return STELLAR_OK;
}

int stellar_download(stellar_ctx* c, float* mass, float* x, float* y,
                     float* vx, float* vy, float* ax, float* ay) {
  if (!c) return STELLAR_ERR_ARG;
  if (!c->uploaded) return fail(c, STELLAR_ERR_STATE, "no state uploaded");
  const int32_t n = c->n;
  if (n == 0) return STELLAR_OK;
  Buffers& B = c->b;
  const size_t bytes = (size_t)n * sizeof(float);
  std::vector<int32_t> id((size_t)n);
  std::vector<float> tmp((size_t)n);
  HIP_TRY(c, hipMemcpyAsync(id.data(), B.id, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const float* src[7] = {B.m, B.x, B.y, B.vx, B.vy, B.ax, B.ay};
  float* dst[7] = {mass, x, y, vx, vy, ax, ay};
  for (int k = 0; k < 7; ++k) {
    if (!dst[k]) continue;
    HIP_TRY(c, hipMemcpyAsync(tmp.data(), src[k], bytes, hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int32_t s = 0; s < n; ++s) dst[k][id[(size_t)s]] = tmp[(size_t)s];
  }
  return STELLAR_OK;
}

int stellar_download_positions(stellar_ctx* c, float* out_xy) {
  if (!c || !out_xy) return STELLAR_ERR_ARG;
  if (!c->uploaded) return fail(c, STELLAR_ERR_STATE, "no state uploaded");
  if (c->n == 0) return STELLAR_OK;
  launch_output(c);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipMemcpyAsync(out_xy, c->b.out, 2 * (size_t)c->n * sizeof(float),
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return STELLAR_OK;
}

int stellar_render(stellar_ctx* c, const float* view, int32_t width, int32_t height,
                   uint32_t* counts, uint8_t* grey) {
  if (!c || !view) return STELLAR_ERR_ARG;
  if (width <= 0 || height <= 0 || (int64_t)width * height > ((int64_t)1 << 28) ||
      !(view[1] > view[0]) || !(view[3] > view[2]))
    return fail(c, STELLAR_ERR_ARG, "render view or size");
  if (!c->uploaded) return fail(c, STELLAR_ERR_STATE, "no state uploaded");
  const int64_t px = (int64_t)width * height;
  const int64_t need = 4 * px + ((px + 255) / 256) * 256;
  if (need > c->render_bytes) {
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->render_buf) (void)hipFree(c->render_buf);
    c->render_buf = nullptr;
    c->render_bytes = 0;
    HIP_TRY(c, hipMalloc(&c->render_buf, (size_t)need));
    c->render_bytes = need;
  }
  uint32_t* dcounts = static_cast<uint32_t*>(c->render_buf);
  uint8_t* dgrey = reinterpret_cast<uint8_t*>(dcounts + px);
  launch_render(c, view, width, height, dcounts, grey ? dgrey : nullptr);
  HIP_TRY(c, hipGetLastError());
  if (counts)
    HIP_TRY(c, hipMemcpyAsync(counts, dcounts, 4 * (size_t)px, hipMemcpyDeviceToHost, c->stream));
  if (grey)
    HIP_TRY(c, hipMemcpyAsync(grey, dgrey, (size_t)px, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return STELLAR_OK;
}

int stellar_bbox(stellar_ctx* c) {
  if (!c) return STELLAR_ERR_ARG;
  if (!c->uploaded) return fail(c, STELLAR_ERR_STATE, "no state uploaded");
  launch_bbox(c);
  HIP_TRY(c, hipGetLastError());
  c->stage = kBboxDone;
  return check_device(c);
}

int stellar_build(stellar_ctx* c) {
  if (!c) return STELLAR_ERR_ARG;
  if (c->stage != kBboxDone) return fail(c, STELLAR_ERR_STATE, "build needs bbox");
  launch_build(c);
  HIP_TRY(c, hipGetLastError());
  c->stage = kBuildDone;
  c->sorted_storage = false;
  return check_device(c);
}

int stellar_com(stellar_ctx* c) {
  if (!c) return STELLAR_ERR_ARG;
  if (c->stage != kBuildDone) return fail(c, STELLAR_ERR_STATE, "com needs build");
  launch_com(c);
  HIP_TRY(c, hipGetLastError());
  c->stage = kComDone;
  return check_device(c);
}

int stellar_sort(stellar_ctx* c) {
  if (!c) return STELLAR_ERR_ARG;
  if (c->stage != kComDone) return fail(c, STELLAR_ERR_STATE, "sort needs com");
  launch_sort(c);
  HIP_TRY(c, hipGetLastError());
  c->stage = kSortDone;
  c->sorted_storage = true;
  return check_device(c);
}

int stellar_forces(stellar_ctx* c) {
  if (!c) return STELLAR_ERR_ARG;
  if (c->stage != kSortDone && c->stage != kForcesDone)
    return fail(c, STELLAR_ERR_STATE, "forces need sort");
  launch_forces(c, true);
  HIP_TRY(c, hipGetLastError());
  c->stage = kForcesDone;
  return check_device(c);
}

int stellar_update(stellar_ctx* c) {
  if (!c) return STELLAR_ERR_ARG;
  if (c->stage != kForcesDone) return fail(c, STELLAR_ERR_STATE, "update needs forces");
  launch_update(c);
  HIP_TRY(c, hipGetLastError());
  c->stage = kUpdateDone;
  return check_device(c);
}

int stellar_step(stellar_ctx* c) {
  if (!c) return STELLAR_ERR_ARG;
  if (!c->uploaded) return fail(c, STELLAR_ERR_STATE, "no state uploaded");
  int rc = run_step(c, -1);
  if (rc) return rc;
  rc = check_device(c);
  if (rc) return rc;
  return collect_times(c, -1);
}

int stellar_run(stellar_ctx* c, int32_t steps) {
  if (!c || steps < 0) return STELLAR_ERR_ARG;
  if (!c->uploaded) return fail(c, STELLAR_ERR_STATE, "no state uploaded");
  if (c->p.flags & STELLAR_FLAG_GRAPH) {
    int32_t k = 0;
    while (k + 2 <= steps) {
      hipGraphExec_t exec = nullptr;
      int rc = two_step_graph(c, &exec);
      if (rc) return rc;
      HIP_TRY(c, hipGraphLaunch(exec, c->stream));
      c->stage = kUpdateDone;
      c->sorted_storage = true;
      k += 2;
    }
    for (; k < steps; ++k) {
      int rc = run_step(c, -1);
      if (rc) return rc;
    }
    return check_device(c);
  }
  if (timing(c)) {
    int rc = reserve_run_events(c, steps);
    if (rc) return rc;
  }
  for (int32_t k = 0; k < steps; ++k) {
    int rc = run_step(c, k);
    if (rc) return rc;
  }
  int rc = check_device(c);
  if (rc) return rc;
  return collect_times(c, steps);
}

int stellar_sync(stellar_ctx* c) {
  if (!c) return STELLAR_ERR_ARG;
  return check_device(c);
}

int stellar_get_bbox_instance(stellar_ctx* c, int32_t inst, float* bbox4) {
  if (!c || !bbox4 || inst < 0 || inst >= c->ninst) return STELLAR_ERR_ARG;
  int rc = check_device(c);
  if (rc) return rc;
  float4 b;
  HIP_TRY(c, hipMemcpy(&b, c->b.ibbox + inst, sizeof(float4), hipMemcpyDeviceToHost));
  bbox4[0] = b.x;
  bbox4[1] = b.y;
  bbox4[2] = b.z;
  bbox4[3] = b.w;
  return STELLAR_OK;
}

int stellar_get_bbox(stellar_ctx* c, float* bbox4) {
  return stellar_get_bbox_instance(c, 0, bbox4);
}

int stellar_num_cells(stellar_ctx* c, int64_t* ncells) {
  if (!c || !ncells) return STELLAR_ERR_ARG;
  int rc = check_device(c);
  if (rc) return rc;
  *ncells = c->b.host_dev->ncells;
  return STELLAR_OK;
}

int stellar_last_quad_interactions(stellar_ctx* c, int64_t* total) {
  if (!c || !total) return STELLAR_ERR_ARG;
  int rc = check_device(c);
  if (rc) return rc;
  *total = (int64_t)c->b.host_dev->quad_interactions;
  return STELLAR_OK;
}

int stellar_last_interactions(stellar_ctx* c, int64_t* total) {
  if (!c || !total) return STELLAR_ERR_ARG;
  int rc = check_device(c);
  if (rc) return rc;
  *total = (int64_t)c->b.host_dev->interactions;
  return STELLAR_OK;
}

int stellar_stage_times(stellar_ctx* c, float* ms7) {
  if (!c || !ms7) return STELLAR_ERR_ARG;
  std::memcpy(ms7, c->stage_ms, sizeof(c->stage_ms));
  return STELLAR_OK;
}

// The reference's numbering of a tree stored in another depth-first order
// (STELLAR_FLAG_HILBERT_ORDER): a depth-first pass in slot order numbers the
// cells and lists the bodies as the reference's sort does; the tree itself
// (slots, levels, moments) is read as stored.  srt[s]: body id at storage
// position s; blk: the C cell blocks of the instance (device cell q at
// blk[4q]).
static int export_canonical(stellar_ctx* c, int32_t inst, int32_t n, int32_t nr, int64_t c0,
                            int64_t C, const std::vector<int32_t>& srt,
                            const std::vector<ChildRec>& rootb, const std::vector<ChildRec>& blk,
                            int32_t* root4, int32_t* child, int32_t* count, int32_t* start,
                            int32_t* level, int32_t* sorted, float* com_x, float* com_y,
                            float* com_m) {
  Buffers& B = c->b;
  std::vector<int64_t> q2k((size_t)(C > 0 ? C : 1), -1);
  std::vector<int32_t> body_order;
  body_order.reserve((size_t)(n > 0 ? n : 1));
  std::vector<int32_t> kstart((size_t)(C > 0 ? C : 1), 0), kcount((size_t)(C > 0 ? C : 1), 0);
  std::vector<int64_t> k2q((size_t)(C > 0 ? C : 1), 0);
  // device cell q of a code, -1 for a body or an empty slot, -2 out of range
  auto cell_of = [&](int32_t code) -> int64_t {
    if (code <= 0) return -1;
    const int64_t q = (int64_t)code - nr - c0;
    return q >= 0 && q < C ? q : -2;
  };
  int64_t next = 0;
  // explicit stack: (device cell, or -1 for the root; child slot to visit next)
  std::vector<std::pair<int64_t, int>> stk;
  stk.push_back({-1, 0});
  while (!stk.empty()) {
    auto& top = stk.back();
    if (top.second == 4) {
      if (top.first >= 0) {
        const int64_t k = q2k[(size_t)top.first];
        kcount[(size_t)k] = (int32_t)body_order.size() - kstart[(size_t)k];
      }
      stk.pop_back();
      continue;
    }
    const int slot = top.second++;
    const ChildRec r = top.first < 0 ? rootb[(size_t)slot] : blk[(size_t)(4 * top.first + slot)];
    if (r.code == 0) continue;
    if (r.code < 0) {
      const int64_t sp = -1 - (int64_t)r.code - (int64_t)c->ioff[(size_t)inst];
      if (sp < 0 || sp >= n) return fail(c, STELLAR_ERR_DEVICE, "tree body out of range");
      body_order.push_back(srt[(size_t)sp]);
      continue;
    }
    const int64_t q = cell_of(r.code);
    if (q < 0 || q2k[(size_t)q] >= 0 || next >= C)
      return fail(c, STELLAR_ERR_DEVICE, "tree cell out of range");
    q2k[(size_t)q] = next;
    k2q[(size_t)next] = q;
    kstart[(size_t)next] = (int32_t)body_order.size();
    ++next;
    if (stk.size() > (size_t)(4 * kMaxLevel)) return fail(c, STELLAR_ERR_DEVICE, "tree depth");
    stk.push_back({q, 0});
  }
  if (next != C || (int64_t)body_order.size() != n)
    return fail(c, STELLAR_ERR_DEVICE, "tree does not span the instance");
  // canonical codes: -1 empty, body id, n + canonical cell
  auto conv = [&](int32_t code) -> int32_t {
    if (code == 0) return -1;
    if (code < 0) {
      const int64_t sp = -1 - (int64_t)code - (int64_t)c->ioff[(size_t)inst];
      return sp >= 0 && sp < n ? srt[(size_t)sp] : -2;
    }
    const int64_t q = cell_of(code);
    return q >= 0 ? (int32_t)(n + q2k[(size_t)q]) : -2;
  };
  if (root4)
    for (int k = 0; k < 4; ++k) root4[k] = conv(rootb[(size_t)k].code);
  if (child)
    for (int64_t k = 0; k < C; ++k)
      for (int sl = 0; sl < 4; ++sl)
        child[4 * k + sl] = conv(blk[(size_t)(4 * k2q[(size_t)k] + sl)].code);
  if (count)
    for (int64_t k = 0; k < C; ++k) count[k] = kcount[(size_t)k];
  if (start)
    for (int64_t k = 0; k < C; ++k) start[k] = kstart[(size_t)k];
  if (sorted && n > 0) std::memcpy(sorted, body_order.data(), (size_t)n * 4);
  std::vector<uint64_t> meta((size_t)(C > 0 ? C : 1));
  if (C > 0 && (level || com_x || com_y || com_m))
    HIP_TRY(c, hipMemcpy(meta.data(), B.cmeta + nr + c0, (size_t)C * 8, hipMemcpyDeviceToHost));
  if (C > 0 && level)
    for (int64_t k = 0; k < C; ++k) level[k] = meta_level(meta[(size_t)k2q[(size_t)k]]);
  if (C > 0 && (com_x || com_y || com_m)) {
    if (c->stage < kComDone) return fail(c, STELLAR_ERR_STATE, "no centre of mass yet");
    // a cell's centre of mass and mass are its record in the parent's block
    for (int64_t k = 0; k < C; ++k) {
      const uint64_t m = meta[(size_t)k2q[(size_t)k]];
      const int64_t pb = meta_parent(m);
      const int sl = meta_slot(m);
      ChildRec r{0.f, 0.f, 0.f, 0};
      if (pb == inst)
        r = rootb[(size_t)sl];
      else if (pb - nr - c0 >= 0 && pb - nr - c0 < C)
        r = blk[(size_t)(4 * (pb - nr - c0) + sl)];
      if (com_x) com_x[k] = r.x;
      if (com_y) com_y[k] = r.y;
      if (com_m) com_m[k] = r.m;
    }
  }
  return STELLAR_OK;
}

int stellar_export_tree_instance(stellar_ctx* c, int32_t inst, int32_t* root4,
                                 int32_t* child, int32_t* count, int32_t* start,
                                 int32_t* level, int32_t* sorted, float* com_x,
                                 float* com_y, float* com_m, int64_t capacity,
                                 int64_t* ncells) {
  if (!c || inst < 0 || inst >= c->ninst) return STELLAR_ERR_ARG;
  if (c->stage < kBuildDone) return fail(c, STELLAR_ERR_STATE, "no tree built");
  int rc = check_device(c);
  if (rc) return rc;
  Buffers& B = c->b;
  const int32_t lo = c->ioff[(size_t)inst], hi = c->ioff[(size_t)inst + 1];
  const int32_t n = hi - lo;  // bodies of this instance
  const int64_t ctot = c->b.host_dev->ncells;
  // cells of the instance: those opened by its positions, ids [c0, c1)
  int64_t c0 = 0, c1 = ctot;
  if (n > 0) {
    int32_t v = 0;
    HIP_TRY(c, hipMemcpy(&v, B.coff + lo, 4, hipMemcpyDeviceToHost));
    c0 = v;
    if (inst + 1 < c->ninst) {
      HIP_TRY(c, hipMemcpy(&v, B.coff + hi, 4, hipMemcpyDeviceToHost));
      c1 = v;
    }
  }
  if (c0 > ctot) c0 = ctot;
  if (c1 > ctot) c1 = ctot;
  const int64_t C = c1 - c0;
  if (ncells) *ncells = C;
  if (capacity < C && (child || count || start || level || com_x || com_y || com_m))
    return fail(c, STELLAR_ERR_ARG, "export capacity below cell count");
  const int32_t nr = c->ninst;  // block of cell q of this instance: nr + c0 + q
  const size_t nn = (size_t)(n > 0 ? n : 1);
  std::vector<int32_t> id(nn), perm(nn), srt(nn);
  if (n > 0) {
    HIP_TRY(c, hipMemcpy(id.data(), B.id + lo, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (c->sorted_storage) {
      for (int32_t s = 0; s < n; ++s) srt[(size_t)s] = id[(size_t)s] - lo;
    } else {
      std::vector<int32_t> idall((size_t)c->n);
      HIP_TRY(c, hipMemcpy(idall.data(), B.id, (size_t)c->n * 4, hipMemcpyDeviceToHost));
      HIP_TRY(c, hipMemcpy(perm.data(), B.perm + lo, (size_t)n * 4, hipMemcpyDeviceToHost));
      for (int32_t s = 0; s < n; ++s) srt[(size_t)s] = idall[(size_t)perm[(size_t)s]] - lo;
    }
  }
  if (sorted && n > 0) std::memcpy(sorted, srt.data(), (size_t)n * 4);
  std::vector<ChildRec> rootb(4), blk((size_t)(4 * (C > 0 ? C : 1)));
  HIP_TRY(c, hipMemcpy(rootb.data(), B.blk + 4 * (int64_t)inst, 4 * sizeof(ChildRec),
                       hipMemcpyDeviceToHost));
  if (C > 0)
    HIP_TRY(c, hipMemcpy(blk.data(), B.blk + 4 * (nr + c0), (size_t)(4 * C) * sizeof(ChildRec),
                         hipMemcpyDeviceToHost));
  auto conv = [&](int32_t code) -> int32_t {
    if (code == 0) return -1;
    if (code < 0) {
      const int64_t s = -1 - (int64_t)code - lo;
      return s >= 0 && s < n ? srt[(size_t)s] : -2;
    }
    const int64_t q = (int64_t)code - nr - c0;
    return q >= 0 && q < C ? (int32_t)(n + q) : -2;
  };
  if (c->p.flags & STELLAR_FLAG_HILBERT_ORDER)
    return export_canonical(c, inst, n, nr, c0, C, srt, rootb, blk, root4, child, count, start,
                            level, sorted, com_x, com_y, com_m);
  if (root4)
    for (int k = 0; k < 4; ++k) root4[k] = conv(rootb[(size_t)k].code);
  if (child)
    for (int64_t q = 0; q < C; ++q)
      for (int k = 0; k < 4; ++k) child[4 * q + k] = conv(blk[(size_t)(4 * q + k)].code);
  if (C > 0 && count) {
    // bodies below each cell, summed up the tree (children have larger ids)
    for (int64_t q = C - 1; q >= 0; --q) {
      int32_t t = 0;
      for (int k = 0; k < 4; ++k) {
        const int32_t code = blk[(size_t)(4 * q + k)].code;
        const int64_t cq = (int64_t)code - nr - c0;
        if (code < 0) t += 1;
        else if (code > 0 && cq > q && cq < C) t += count[cq];
      }
      count[q] = t;
    }
  }
  std::vector<uint64_t> meta((size_t)(C > 0 ? C : 1));
  if (C > 0 && (start || level || com_x || com_y || com_m))
    HIP_TRY(c, hipMemcpy(meta.data(), B.cmeta + nr + c0, (size_t)C * 8, hipMemcpyDeviceToHost));
  if (C > 0 && start)
    for (int64_t q = 0; q < C; ++q) start[q] = meta_start(meta[(size_t)q]) - lo;
  if (C > 0 && level)
    for (int64_t q = 0; q < C; ++q) level[q] = meta_level(meta[(size_t)q]);
  if (C > 0 && (com_x || com_y || com_m)) {
    if (c->stage < kComDone) return fail(c, STELLAR_ERR_STATE, "no centre of mass yet");
    // a cell's centre of mass and mass are its record in the parent's block
    for (int64_t q = 0; q < C; ++q) {
      const int64_t pb = meta_parent(meta[(size_t)q]);
      const int k = meta_slot(meta[(size_t)q]);
      ChildRec r{0.f, 0.f, 0.f, 0};
      if (pb == inst)
        r = rootb[(size_t)k];
      else if (pb - nr - c0 >= 0 && pb - nr - c0 < q)
        r = blk[(size_t)(4 * (pb - nr - c0) + k)];
      if (com_x) com_x[q] = r.x;
      if (com_y) com_y[q] = r.y;
      if (com_m) com_m[q] = r.m;
    }
  }
  return STELLAR_OK;
}

int stellar_export_tree(stellar_ctx* c, int32_t* root4, int32_t* child,
                        int32_t* count, int32_t* start, int32_t* level,
                        int32_t* sorted, float* com_x, float* com_y,
                        float* com_m, int64_t capacity, int64_t* ncells) {
  return stellar_export_tree_instance(c, 0, root4, child, count, start, level, sorted,
                                      com_x, com_y, com_m, capacity, ncells);
}

int stellar_export_forces(stellar_ctx* c, float* ax, float* ay,
                          int32_t* interactions) {
  if (!c) return STELLAR_ERR_ARG;
  if (c->stage < kForcesDone) return fail(c, STELLAR_ERR_STATE, "no forces yet");
  int rc = check_device(c);
  if (rc) return rc;
  const int32_t n = c->n;
  if (n == 0) return STELLAR_OK;
  Buffers& B = c->b;
  std::vector<int32_t> id((size_t)n), it((size_t)n);
  std::vector<float> tmp((size_t)n);
  HIP_TRY(c, hipMemcpy(id.data(), B.id, (size_t)n * 4, hipMemcpyDeviceToHost));
  if (ax) {
    HIP_TRY(c, hipMemcpy(tmp.data(), B.ax, (size_t)n * 4, hipMemcpyDeviceToHost));
    for (int32_t s = 0; s < n; ++s) ax[id[(size_t)s]] = tmp[(size_t)s];
  }
  if (ay) {
    HIP_TRY(c, hipMemcpy(tmp.data(), B.ay, (size_t)n * 4, hipMemcpyDeviceToHost));
    for (int32_t s = 0; s < n; ++s) ay[id[(size_t)s]] = tmp[(size_t)s];
  }
  if (interactions) {
    HIP_TRY(c, hipMemcpy(it.data(), B.inter, (size_t)n * 4, hipMemcpyDeviceToHost));
    for (int32_t s = 0; s < n; ++s) interactions[id[(size_t)s]] = it[(size_t)s];
  }
  return STELLAR_OK;
}

int stellar_direct_fp64(stellar_ctx* c, int32_t ntargets, const int32_t* targets,
                        double eps2, double* ax, double* ay) {
  if (!c || !ax || !ay || ntargets < 0) return STELLAR_ERR_ARG;
  if (!c->uploaded) return fail(c, STELLAR_ERR_STATE, "no state uploaded");
  if (!targets && ntargets != c->n) return fail(c, STELLAR_ERR_ARG, "ntargets");
  const int32_t n = c->n;
  if (ntargets == 0 || n == 0) return STELLAR_OK;
  if (targets)
    for (int32_t q = 0; q < ntargets; ++q)
      if (targets[q] < 0 || targets[q] >= n)
        return fail(c, STELLAR_ERR_ARG, "target id out of range");
  int32_t* dt = nullptr;
  double *dax = nullptr, *day = nullptr;
  HIP_TRY(c, hipMalloc(&dax, (size_t)ntargets * sizeof(double)));
  if (hipMalloc(&day, (size_t)ntargets * sizeof(double)) != hipSuccess) {
    (void)hipFree(dax);
    return fail(c, STELLAR_ERR_OOM, "direct sum buffers");
  }
  if (targets) {
    if (hipMalloc(&dt, (size_t)ntargets * sizeof(int32_t)) != hipSuccess) {
      (void)hipFree(dax);
      (void)hipFree(day);
      return fail(c, STELLAR_ERR_OOM, "direct sum buffers");
    }
    (void)hipMemcpyAsync(dt, targets, (size_t)ntargets * 4, hipMemcpyHostToDevice,
                         c->stream);
  }
  launch_direct_fp64(c, ntargets, dt, eps2, dax, day);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(ax, dax, (size_t)ntargets * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ay, day, (size_t)ntargets * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dax);
  (void)hipFree(day);
  if (dt) (void)hipFree(dt);
  if (e != hipSuccess) return hip_fail(c, e, "stellar_direct_fp64");
  return STELLAR_OK;
}

// Per-instance FP64 sums of the energy kernel's slices, in a fixed order
// (the tiles of an instance in position order), so results do not depend
// on scheduling.
static int diagnostics(stellar_ctx* c, double eps2, double half_dt, bool sync,
                       stellar_diagnostics* out) {
  const int32_t ni = c->ninst;
  for (int32_t i = 0; i < ni; ++i) out[i] = stellar_diagnostics{0.0, 0.0, 0.0, 0.0, 0.0};
  if (c->n == 0) return STELLAR_OK;
  const int parts = energy_parts(c);
  double* dpart = nullptr;
  HIP_TRY(c, hipMalloc(&dpart, (size_t)parts * kEnergyVals * sizeof(double)));
  launch_energy_fp64(c, eps2, half_dt, sync, dpart);
  std::vector<double> h((size_t)parts * kEnergyVals);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), dpart, h.size() * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dpart);
  if (e != hipSuccess) return hip_fail(c, e, "FP64 diagnostics");
  for (int32_t i = 0; i < ni; ++i) {
    stellar_diagnostics& d = out[i];
    for (int64_t q = (int64_t)c->itile[(size_t)i] * kEnergySlices;
         q < (int64_t)c->itile[(size_t)i + 1] * kEnergySlices; ++q) {
      const double* v = &h[(size_t)(kEnergyVals * q)];
      d.potential += v[0];
      d.kinetic += v[1];
      d.px += v[2];
      d.py += v[3];
      d.lz += v[4];
    }
  }
  return STELLAR_OK;
}

int stellar_diagnostics_fp64(stellar_ctx* c, double eps2, double half_dt,
                             stellar_diagnostics* out) {
  if (!c || !out) return STELLAR_ERR_ARG;
  if (!c->uploaded) return fail(c, STELLAR_ERR_STATE, "no state uploaded");
  const bool sync = half_dt != 0.0;
  if (sync && c->stage != kForcesDone)
    return fail(c, STELLAR_ERR_STATE, "synchronised velocities need forces at the current positions");
  return diagnostics(c, eps2, half_dt, sync, out);
}

int stellar_energy_fp64_instances(stellar_ctx* c, double eps2, double* kinetic,
                                  double* potential, double* px, double* py) {
  if (!c) return STELLAR_ERR_ARG;
  if (!c->uploaded) return fail(c, STELLAR_ERR_STATE, "no state uploaded");
  std::vector<stellar_diagnostics> d((size_t)c->ninst);
  const int rc = diagnostics(c, eps2, 0.0, false, d.data());
  if (rc) return rc;
  for (int32_t i = 0; i < c->ninst; ++i) {
    if (kinetic) kinetic[i] = d[(size_t)i].kinetic;
    if (potential) potential[i] = d[(size_t)i].potential;
    if (px) px[i] = d[(size_t)i].px;
    if (py) py[i] = d[(size_t)i].py;
  }
  return STELLAR_OK;
}

int stellar_energy_fp64(stellar_ctx* c, double eps2, double* kinetic,
                        double* potential, double* px, double* py) {
  if (!c) return STELLAR_ERR_ARG;
  std::vector<double> k((size_t)c->ninst), u((size_t)c->ninst), mx((size_t)c->ninst),
      my((size_t)c->ninst);
  const int rc = stellar_energy_fp64_instances(c, eps2, k.data(), u.data(), mx.data(), my.data());
  if (rc) return rc;
  double tk = 0, tu = 0, tx = 0, ty = 0;
  for (int32_t i = 0; i < c->ninst; ++i) {
    tk += k[(size_t)i];
    tu += u[(size_t)i];
    tx += mx[(size_t)i];
    ty += my[(size_t)i];
  }
  if (kinetic) *kinetic = tk;
  if (potential) *potential = tu;
  if (px) *px = tx;
  if (py) *py = ty;
  return STELLAR_OK;
}

int stellar_debug_sort_cap(stellar_ctx* c, int32_t cap) {
  if (!c) return STELLAR_ERR_ARG;
  if (cap != -1 && (cap < 1 || cap > kSsCapMax)) return fail(c, STELLAR_ERR_ARG, "sort cap");
  c->radix_sort = cap == -1;
  if (cap > 0) c->sort_cap = cap;
  drop_graphs(c);
  return STELLAR_OK;
}

int stellar_set_walk_shape(stellar_ctx* c, int32_t nodes_per_pop, int32_t block_threads) {
  if (!c) return STELLAR_ERR_ARG;
  if (nodes_per_pop != 1 && nodes_per_pop != 2 && nodes_per_pop != 4)
    return fail(c, STELLAR_ERR_ARG, "nodes_per_pop");
  if (block_threads != 64 && block_threads != 128 && block_threads != 256 && block_threads != 512 &&
      block_threads != 1024)
    return fail(c, STELLAR_ERR_ARG, "block_threads");
  c->walk_kpop = nodes_per_pop;
  c->walk_waves = block_threads / 64;
  drop_graphs(c);
  return STELLAR_OK;
}

int stellar_set_walk_schedule(stellar_ctx* c, int32_t schedule) {
  if (!c) return STELLAR_ERR_ARG;
  if (schedule != STELLAR_WALK_STATIC && schedule != STELLAR_WALK_QUEUE &&
      schedule != STELLAR_WALK_QUEUE_WAVE)
    return fail(c, STELLAR_ERR_ARG, "walk schedule");
  c->walk_queue = schedule;
  drop_graphs(c);
  return STELLAR_OK;
}

int stellar_autotune(stellar_ctx* c, int32_t group_mask, int32_t reps, stellar_tuning* out) {
  if (!c || reps < 1 || group_mask < 0 || group_mask > 7) return STELLAR_ERR_ARG;
  if (!c->uploaded) return fail(c, STELLAR_ERR_STATE, "no state uploaded");
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int, int>, stellar_tuning> cache;
  int lg = 0;
  for (int64_t per = c->n / c->ninst; per > 1; per >>= 1) ++lg;
  int li = 0;
  for (int32_t k = c->ninst; k > 1; k >>= 2) ++li;
  const int fl = c->p.flags & (STELLAR_FLAG_QUADRUPOLE | STELLAR_FLAG_TEXTBOOK_OPENING |
                               STELLAR_FLAG_HILBERT_ORDER);
  const auto key = std::make_tuple(lg, li, fl, group_mask, (int)c->p.precision);
  auto apply = [&](const stellar_tuning& t) {
    c->p.group = t.group;
    c->walk_kpop = t.nodes_per_pop;
    c->walk_waves = t.block_threads / 64;
    c->walk_queue = t.schedule;
    drop_graphs(c);
    if (out) *out = t;
  };
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      apply(it->second);
      return STELLAR_OK;
    }
  }
  // a tree to walk
  launch_bbox(c);
  launch_build(c);
  launch_com(c);
  launch_sort(c);
  HIP_TRY(c, hipGetLastError());
  c->stage = kSortDone;
  c->sorted_storage = true;
  int rc = check_device(c);
  if (rc) return rc;
  const bool quad = (c->p.flags & STELLAR_FLAG_QUADRUPOLE) != 0;
  std::vector<stellar_tuning> cand;
  // the mixed far field exists for group-64 walks only
  const int mask = group_mask == 0 || c->p.precision == STELLAR_PRECISION_MIXED ? 1 : group_mask;
  if (mask & 1)
    for (int kp : {1, 2, 4})
      for (int w : {1, 2, 4, 8, 16})
        if (!((quad || (c->p.flags & STELLAR_FLAG_BMAX_OPENING)) && kp == 4)) {
          cand.push_back(stellar_tuning{64, kp, 64 * w, 0.f, STELLAR_WALK_STATIC});
          if (w == 4 || w == 8)
            cand.push_back(stellar_tuning{64, kp, 64 * w, 0.f, STELLAR_WALK_QUEUE});
        }
  if (mask & 2) cand.push_back(stellar_tuning{32, 1, 256, 0.f, STELLAR_WALK_STATIC});
  if (mask & 4) cand.push_back(stellar_tuning{16, 1, 256, 0.f, STELLAR_WALK_STATIC});
  const stellar_params saved = c->p;
  const int skp = c->walk_kpop, sw = c->walk_waves, sq = c->walk_queue;
  stellar_tuning best{};
  best.walk_ms = 3.4e38f;
  for (const stellar_tuning& t : cand) {
    c->p.group = t.group;
    c->walk_kpop = t.nodes_per_pop;
    c->walk_waves = t.block_threads / 64;
    c->walk_queue = t.schedule;
    launch_forces(c, true);  // untimed
    HIP_TRY(c, hipEventRecord(c->ev[0], c->stream));
    for (int32_t r = 0; r < reps; ++r) launch_forces(c, true);
    HIP_TRY(c, hipEventRecord(c->ev[1], c->stream));
    HIP_TRY(c, hipGetLastError());
    rc = check_device(c);
    if (rc) {
      c->p = saved;
      c->walk_kpop = skp;
      c->walk_waves = sw;
      c->walk_queue = sq;
      return rc;
    }
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    ms /= (float)reps;
    if (ms < best.walk_ms) {
      best = t;
      best.walk_ms = ms;
    }
  }
  c->stage = kForcesDone;
  {
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = best;
  }
  apply(best);
  return STELLAR_OK;
}

}  // extern "C"
