// dc_hip.cpp — C ABI of the MI355X dual-contouring backend: context, HBM buffers, pipeline.
//
// One extraction (or one batch) is a fixed sequence of launches on the context's stream with a
// single host synchronization at the end (the counters readback). Output sizes are data
// dependent: buffers are sized from the previous call (or a fraction of the cell count), every
// kernel bounds its writes by the capacity it was given, and a run whose counts exceed the
// capacities is repeated once with exact capacities.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dc_decode.h"
#include "dc_hip.h"
#include "dc_internal.h"
#include "dc_kernels.h"
#include "dc_m2s.h"
#include "dc_sparse.h"
#include "dc_post.h"

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

constexpr int kNumEvents = 10;
// Pinned staging of one grid's descriptor and bounds (dc_hip_extract_device_noalloc).
constexpr size_t kStageBytes = 1024;  // 0..7 forward passes, 8..9 upload / backward
// Adjoint: active cells per CU up to which the second per-cell kernel runs 4 lanes per cell
// (its crossing edges split over the lanes) rather than one thread per cell.
constexpr int64_t kWideCellsPerCu = 1024;

// A launch sequence that is captured as a hipGraph the second time in a row it runs with the same
// key (one-off sequences stay eager) and replayed while the key stays the same.
struct Captured {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string key;       // what `exec` was captured for
  std::string last_key;  // the previous run's key
  int mode = 0;          // the last run: 0 eager, 1 captured, 2 replayed
  void drop() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key.clear();
  }
};

// The device buffers of one forward launch sequence: the context's grow-only buffers, or
// carvings of the static scratch array (dc_hip_extract_device_noalloc).
struct Bufs {
  DcGridDev* grids;
  int64_t* bounds;      // 7 * (n + 1): tiles, act words, pencils, samples, gather rows, face
                        // words, cell-pass waves
  uint64_t* act;
  uint4* sg;
  uint32_t* wprefix;
  uint64_t* ft;
  uint32_t* pcount;
  uint32_t* pbase;
  uint32_t* scan_sum;
  uint64_t* scan_base;
  DcCounters* ctr;
  uint64_t* alist;
  double* v64;
  float* v32;
  uint32_t* flags;
  uint32_t* fblist;
  int32_t* faces;
  uint32_t* pw;
  void* fnrm;
  float* fnrm32;  // FP32 mode: FP32 copy of the face normal table
  uint64_t* flist;
  uint32_t* goff;
  uint64_t* reorient;  // FP32 mode: near-tie quads (capF list entries), or null
  float* qstate;      // FP32 mode's adjoint inputs, or null
  uint32_t* rows;
  double* qstate64;
  uint64_t capA, capF, cap64;
  bool timing;        // per-pass events
};

}  // namespace

// ---------------------------------------------------------------------------------------------
// Static scratch of dc_hip_extract_device_noalloc: every intermediate buffer of one single-grid
// forward (sign and face planes, the active plane, scan tiles, lists, face normals, counters and
// descriptors) is carved from this one file-scope device array, so that entry point calls no
// allocator. Each device holds its own copy of kScratchBytes for the life of the process; host
// code never uses the name `scratch` as an address (on the host it is a stand-in): the device
// address comes from hipGetSymbolAddress, once per device, at dc_hip_create with that device
// current. Calls on one device share the array (and keep the last call's contents): calls that
// overlap on one device race.
// ---------------------------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_rc[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current), looked up once.
void init_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) (void)hipGetLastError();  // leave no sticky error for later launches
  g_scratch_rc[device] = e;
  g_scratch_ptr[device] = e == hipSuccess ? static_cast<unsigned char*>(p) : nullptr;
  g_scratch_done[device] = true;
}

// Consecutive 256-byte aligned pieces of a base address (null base: sizes only).
struct Carver {
  unsigned char* base;
  size_t off = 0;
  template <typename T>
  T* take(size_t bytes) {
    off = (off + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += std::max<size_t>(bytes, 256);
    return p;
  }
};

}  // namespace

struct dc_hip_context {
  int device = 0;
  hipStream_t own = nullptr;     // captures sequences meant for the null stream
  hipStream_t stream = nullptr;
  std::string err;
  // sign-pass tile (option "mark_rows"): 0 picks per plan (see plan), kMarkWideRows whole rows,
  // else that many 64-sample rows per wave (4 waves per block by default)
  int mark_rows = 0;
  int eff_mark_rows = 16;  // the plan's choice
  int mark_waves = 4;
  int gather_rows = 1;
  int64_t gather_blocks = 0;  // option "gather_blocks"
  int eff_gather_rows = 1;  // the value the last forward's row groups were laid out with
  int gather_lds_row = 128;
  bool use_graph = true;
  uint64_t alloc_gen = 0;
  Captured fwd, bwd;     // forward and adjoint launch sequences
  std::string dev_key;   // plan whose descriptors and bounds are in the device buffers
  bool timing = true;
  int64_t qef_blocks = 0;
  int cu_count = 256;
  DevBuf grids, bounds, act, sg, ft, pcount, pbase, wprefix, alist, v64, v32, faces, scan_sum, scan_base,
      counters, upload, flags, bwd_contrib, pw, fnrm, ubar, ubar_e, bwd_rows, fblist, flist, goff,
      qstate, qstate64, tris, reorient;
  DcCounters* h_counters = nullptr;
  unsigned char* h_stage = nullptr;  // pinned: descriptors of the allocation-free entry point
  uint32_t* h_goff = nullptr;  // pinned: per-grid first vertex / first face of the last batch
  char* h_ring[2] = {nullptr, nullptr};  // pinned: staging chunks of a large host-grid upload
  hipEvent_t ring_ev[2] = {nullptr, nullptr};
  int h_goff_cap = 0;
  hipEvent_t ev[kNumEvents] = {};
  hipStream_t side = nullptr;    // second stream for independent work beside the main one
  hipEvent_t ev_fork = nullptr, ev_side = nullptr;
  // last forward
  bool have = false;
  int precision = 0;
  int64_t nv = 0, nf = 0;
  uint64_t capA = 0, capF = 0;
  uint64_t cap64 = 0;  // FP32 mode: FP64 cells whose solve state is kept for the adjoint
  int64_t cap64_opt = -1;  // option "fallback_state" (-1: capA / 4 + 64)
  int bwd_lanes1 = 0, bwd_lanes2 = 0;  // options "adjoint_lanes", "adjoint_face_lanes" (0: auto)
  int bwd_lanes64 = 0;                 // option "adjoint_fallback_lanes" (0: auto)
  std::vector<DcGridDev> hgrids;
  DcPost* post = nullptr;  // mesh post-processing state and its last result
  DcM2s* m2s = nullptr;    // mesh-to-SDF state
  DcSparse* sparse = nullptr;  // sparse brick extraction state and its last mesh
  std::vector<int64_t> tile_b, act_b, pen_b, samp_b, sr_b, ft_b, cw_b, hbounds;
  int64_t n_words = 0, n_ft = 0, n_pen = 0, n_tiles = 0, n_samples = 0, n_sr = 0, n_cw = 0,
          n_sg = 0;
};

namespace {

int fail(dc_hip_context* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}
// A literal message: assigned within the error string's reserved capacity (no allocation).
int fail(dc_hip_context* c, int code, const char* msg) {
  if (c) c->err.assign(msg);
  return code;
}

int hip_fail(dc_hip_context* c, hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory)
    return fail(c, DC_HIP_ERR_OUT_OF_MEMORY, std::string(what) + ": out of device memory");
  return fail(c, DC_HIP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define DC_TRY(ctx, expr, what)                  \
  do {                                           \
    hipError_t e_ = (expr);                      \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, what); \
  } while (0)

// Grows a device buffer. forward: the buffer is used by the forward launch sequence, so moving it
// makes a captured sequence stale.
int ensure(dc_hip_context* c, DevBuf& b, size_t bytes, const char* name, bool forward = true) {
  if (bytes == 0) bytes = 256;
  if (b.cap >= bytes) return DC_HIP_OK;
  if (forward) ++c->alloc_gen;
  if (b.p) {
    hipError_t e = hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    if (e != hipSuccess) return hip_fail(c, e, name);
  }
  const size_t want = bytes + bytes / 8;
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    b.p = nullptr;
    return hip_fail(c, e, name);
  }
  b.cap = want;
  return DC_HIP_OK;
}

template <typename T>
T* as(DevBuf& b) {
  return reinterpret_cast<T*>(b.p);
}

// Layout of a batch; returns an error for degenerate or oversized grids.
int plan(dc_hip_context* c, const dc_hip_grid* grids, int n) {
  {
    // adjoint gather: LDS row length for the widest grid (at most 16 chunks per step)
    int max_ws = 1;
    for (int q = 0; q < n; ++q) max_ws = std::max(max_ws, std::min((grids[q].nx + 63) / 64, 16));
    c->gather_lds_row = 64 * max_ws + 64;
    c->eff_gather_rows = c->gather_rows;
    // The sign pass: the whole-row form reads 8 samples per lane as two 16-byte loads (rows of
    // 8k samples, 16-byte aligned values). By default it is used where the 16-row form would
    // have fewer than 8 waves per CU (below ~512^3: at 64^3-256^3 it takes the sign pass from
    // 43-46 us to 13-16 us, at 384^3 the two are equal, at 512^3 and 1024^3 the 16-row form is
    // 16% and 7% faster).
    bool wide_ok = true;
    int64_t t16 = 0;
    for (int q = 0; q < n; ++q) {
      const dc_hip_grid& s = grids[q];
      wide_ok = wide_ok && s.nx % 8 == 0 && ((uintptr_t)s.values & 15) == 0;
      t16 += (int64_t)((s.nx + 63) / 64) * ((s.ny + 15) / 16) * ((s.nz + 63) / 64);
    }
    int rows = c->mark_rows;
    if (rows == 0) rows = t16 < 8ll * c->cu_count ? kMarkWideRows : 16;
    c->eff_mark_rows = rows == kMarkWideRows && !wide_ok ? 16 : rows;
  }
  c->hgrids.resize(n);
  c->tile_b.assign(n + 1, 0);
  c->act_b.assign(n + 1, 0);
  c->pen_b.assign(n + 1, 0);
  c->samp_b.assign(n + 1, 0);
  c->sr_b.assign(n + 1, 0);
  c->ft_b.assign(n + 1, 0);
  c->cw_b.assign(n + 1, 0);
  int64_t words = 0, ftw = 0, pen = 0, tiles = 0, samp = 0, sr = 0, cw = 0, sgn = 0;
  for (int q = 0; q < n; ++q) {
    const dc_hip_grid& s = grids[q];
    if (s.nx < 2 || s.ny < 2 || s.nz < 2)
      return fail(c, DC_HIP_ERR_INVALID_GRID, "grid dimensions must be >= 2");
    if (s.nx > 65536 || s.ny > 65536 || s.nz > 65536)
      return fail(c, DC_HIP_ERR_INVALID_GRID, "grid dimension above 65536");
    if (!s.values) return fail(c, DC_HIP_ERR_INVALID_ARG, "grid values pointer is NULL");
    // the sign pass addresses one sample layer with 32-bit byte offsets
    if ((int64_t)s.nx * s.ny > (1ll << 30) - 64)
      return fail(c, DC_HIP_ERR_INVALID_GRID, "grid layer too large (nx * ny above 2^30 samples)");
    DcGridDev& g = c->hgrids[q];
    g.val = s.values;
    g.nx = s.nx;
    g.ny = s.ny;
    g.nz = s.nz;
    g.cx = s.nx - 1;
    g.cy = s.ny - 1;
    g.cz = s.nz - 1;
    g.W = (g.cx + 63) / 64;
    g.Wk = (g.cz + 63) / 64;
    g.Ws = (g.nx + 63) / 64;
    g.Wks = (g.nz + 63) / 64;
    g.vx = s.vx;
    g.vy = s.vy;
    g.vz = s.vz;
    g.ox = s.ox;
    g.oy = s.oy;
    g.oz = s.oz;
    g.act_off = words;
    g.sg_off = sgn;
    g.ft_off = ftw;
    g.pen_off = pen;
    g.samp_off = samp;
    g.sr_off = sr;
    c->tile_b[q] = tiles;
    c->act_b[q] = words;
    c->pen_b[q] = pen;
    c->samp_b[q] = samp;
    c->sr_b[q] = sr;
    c->ft_b[q] = ftw;
    c->cw_b[q] = cw;
    words += (int64_t)g.W * g.cy * g.cz;
    ftw += 3ll * g.Wk * g.cy * g.cx;
    pen += 3ll * g.cy * g.cx;
    tiles += dc_mark_tiles(g, c->eff_mark_rows);
    cw += dc_cell_waves(g);
    sgn += (int64_t)g.Wks * g.ny * 64 * g.Ws;
    samp += (int64_t)g.nx * g.ny * g.nz;
    sr += ((int64_t)g.ny * g.nz + c->eff_gather_rows - 1) / c->eff_gather_rows;  // gather groups
  }
  c->tile_b[n] = tiles;
  c->act_b[n] = words;
  c->pen_b[n] = pen;
  c->samp_b[n] = samp;
  c->sr_b[n] = sr;
  c->ft_b[n] = ftw;
  c->cw_b[n] = cw;
  if ((uint64_t)words * 64ull >= (1ull << 32) * 4ull)
    return fail(c, DC_HIP_ERR_INVALID_GRID, "batch has too many cells for 32-bit vertex ids");
  c->n_words = words;
  c->n_ft = ftw;
  c->n_pen = pen;
  c->n_tiles = tiles;
  c->n_samples = samp;
  c->n_sr = sr;
  c->n_cw = cw;
  c->n_sg = sgn;
  return DC_HIP_OK;
}

float elapsed(dc_hip_context* c, int a, int b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) != hipSuccess) return 0.f;
  return ms;
}

// The launch sequence after the buffers and descriptors are in place: kernels from the counter
// clear to the counters' (and batch offsets') read-back into pinned memory. No host
// synchronisation and no memset nodes, so it can be captured in a graph and replayed.
int enqueue(dc_hip_context* c, int precision, hipStream_t s, const Bufs& b);

// The context's own buffers as a launch sequence's Bufs.
Bufs context_bufs(dc_hip_context* c, int precision);

void drop_graph(dc_hip_context* c) {
  c->fwd.drop();
  c->bwd.drop();
}

// Runs `enqueue(s)` (a launch sequence without host synchronisation) on stream s: eagerly, or
// through cap's graph (captured now when `key` repeats the previous run's, replayed while it
// matches the captured one). graphs = false forces an eager run. The null stream cannot be
// captured: its sequence is captured on the context's own stream (capture records, it runs
// nothing) and the graph is launched on the null stream.
template <class Enqueue>
int run_captured(dc_hip_context* c, Captured& cap, std::string key, bool graphs, hipStream_t s,
                 Enqueue&& enqueue) {
  const bool replay = graphs && cap.exec && key == cap.key;
  const bool capture = graphs && !replay && key == cap.last_key;
  cap.last_key = std::move(key);
  cap.mode = replay ? 2 : capture ? 1 : 0;
  if (!replay && !capture) return enqueue(s);
  if (capture) {
    cap.drop();
    const hipStream_t cs = s ? s : c->own;
    DC_TRY(c, hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
    const int rc = enqueue(cs);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(c, e, "end capture");
    cap.graph = g;
    DC_TRY(c, hipGraphInstantiate(&cap.exec, g, nullptr, nullptr, 0), "instantiate graph");
    cap.key = cap.last_key;
  }
  DC_TRY(c, hipGraphLaunch(cap.exec, s), "graph launch");
  return DC_HIP_OK;
}

// The stream the calls run on: the caller's (set by dc_hip_set_stream), or the null stream.
int work_stream(dc_hip_context* c, hipStream_t* s) {
  *s = c->stream;
  return DC_HIP_OK;
}

// Everything a captured sequence depends on: descriptors (sizes, spacing, value pointers), the
// per-grid bounds, precision, launch options, capacities, the stream and the buffer addresses.
std::string plan_key(dc_hip_context* c, int precision, hipStream_t s) {
  std::string k;
  auto put = [&](const void* p, size_t n) { k.append(reinterpret_cast<const char*>(p), n); };
  const int n = (int)c->hgrids.size();
  put(&n, sizeof(n));
  put(c->hgrids.data(), n * sizeof(DcGridDev));
  put(c->hbounds.data(), c->hbounds.size() * sizeof(int64_t));
  const int64_t v[] = {precision, c->mark_rows, c->mark_waves, c->qef_blocks, c->timing,
                       (int64_t)c->capA, (int64_t)c->capF, c->cap64_opt, (int64_t)c->alloc_gen,
                       (int64_t)(uintptr_t)s, (int64_t)(uintptr_t)c->h_goff};
  put(v, sizeof(v));
  return k;
}

// One pass: buffers, descriptor upload, then the launch sequence (replayed from the captured
// graph when the plan is unchanged). Returns DC_HIP_OK after the counters have been read back.
int run_once(dc_hip_context* c, int precision) {
  const int n = (int)c->hgrids.size();
  int rc;
  hipStream_t s;
  if ((rc = work_stream(c, &s))) return rc;
  if ((rc = ensure(c, c->grids, n * sizeof(DcGridDev), "grid descriptors"))) return rc;
  if ((rc = ensure(c, c->bounds, 7 * (n + 1) * sizeof(int64_t), "grid bounds"))) return rc;
  // act and wprefix hold one word past their end: quad emission loads two words at a time
  if ((rc = ensure(c, c->act, (c->n_words + 1) * 8, "active plane"))) return rc;
  if ((rc = ensure(c, c->sg, c->n_sg * 16, "sign planes"))) return rc;
  if ((rc = ensure(c, c->wprefix, (c->n_words + 1) * 4, "word prefix"))) return rc;
  if ((rc = ensure(c, c->ft, c->n_ft * 8, "face planes"))) return rc;
  if ((rc = ensure(c, c->pcount, c->n_pen * 4, "pencil counts"))) return rc;
  if ((rc = ensure(c, c->pbase, c->n_pen * 4, "pencil bases"))) return rc;
  const int64_t t_a = dc_scan_tiles(c->n_words), t_p = dc_scan_tiles(c->n_pen),
                t_f = dc_scan_tiles(c->n_ft);
  if ((rc = ensure(c, c->scan_sum, (t_a + t_p + t_f) * 4, "scan tile sums"))) return rc;
  if ((rc = ensure(c, c->scan_base, (t_a + t_p + t_f) * 8, "scan tile bases"))) return rc;
  if ((rc = ensure(c, c->counters, sizeof(DcCounters), "counters"))) return rc;
  if ((rc = ensure(c, c->alist, c->capA * 8, "active list"))) return rc;
  if (precision == DC_HIP_PLAIN_FP64 &&
      (rc = ensure(c, c->v64, c->capA * 24, "fp64 vertices")))
    return rc;
  if ((rc = ensure(c, c->v32, c->capA * 12, "vertices"))) return rc;
  if ((rc = ensure(c, c->flags, c->capA * 4, "vertex flags"))) return rc;
  if ((rc = ensure(c, c->fblist, c->capA * 4, "fallback list"))) return rc;
  if ((rc = ensure(c, c->faces, c->capF * 16, "faces"))) return rc;
  if ((rc = ensure(c, c->pw, c->n_ft * 4, "face word prefix"))) return rc;
  // FP64 table (32 bytes per face), then in FP32 mode its FP32 copy (16 bytes per face)
  if ((rc = ensure(c, c->fnrm, c->capF * (precision == DC_HIP_FP32_FALLBACK ? 48 : 32),
                   "face normals")))
    return rc;
  if ((rc = ensure(c, c->flist, c->capF * 8, "face list"))) return rc;
  if ((rc = ensure(c, c->goff, 8 * (size_t)n, "grid offsets"))) return rc;
  // FP32 mode: the adjoint's inputs from the forward. The FP64 cells' solve states are kept for
  // up to cap64 of them (a few percent of the cells are typical on fine grids; with more, the
  // adjoint solves every FP64 cell again).
  c->cap64 = c->cap64_opt >= 0 ? (uint64_t)c->cap64_opt : c->capA / 4 + 64;
  if (precision == DC_HIP_FP32_FALLBACK &&
      ((rc = ensure(c, c->reorient, c->capF * 8, "near-tie quads")) ||
       (rc = ensure(c, c->qstate, c->capA * 4 * 24, "solve state")) ||
       (rc = ensure(c, c->bwd_rows, c->capA * 4 * 12, "adjoint row kinds")) ||
       (rc = ensure(c, c->qstate64, c->cap64 * 8 * 24, "fallback solve state"))))
    return rc;
  if (n > c->h_goff_cap) {
    if (c->h_goff) hipHostFree(c->h_goff);
    c->h_goff = nullptr;
    c->h_goff_cap = 0;
    if (hipHostMalloc((void**)&c->h_goff, 8 * (size_t)n, hipHostMallocDefault) != hipSuccess)
      return fail(c, DC_HIP_ERR_OUT_OF_MEMORY, "pinned grid offsets");
    c->h_goff_cap = n;
  }

  // host staging stays alive in the context until the copies have run
  std::vector<int64_t>& bounds = c->hbounds;
  bounds.clear();
  bounds.insert(bounds.end(), c->tile_b.begin(), c->tile_b.end());
  bounds.insert(bounds.end(), c->act_b.begin(), c->act_b.end());
  bounds.insert(bounds.end(), c->pen_b.begin(), c->pen_b.end());
  bounds.insert(bounds.end(), c->samp_b.begin(), c->samp_b.end());
  bounds.insert(bounds.end(), c->sr_b.begin(), c->sr_b.end());
  bounds.insert(bounds.end(), c->ft_b.begin(), c->ft_b.end());
  bounds.insert(bounds.end(), c->cw_b.begin(), c->cw_b.end());
  // A plan is captured the second time in a row it runs (one-off plans stay eager) and replayed
  // while it stays unchanged. Per-pass timing needs events between the kernels, which a replayed
  // graph does not order reliably: with timing on, every run is eager.
  std::string key = plan_key(c, precision, s);
  if (key != c->dev_key) {
    // the device copies of the descriptors and bounds (read by every kernel, not part of a
    // captured sequence) are rewritten only when they belong to another plan
    DC_TRY(c, hipMemcpyAsync(c->grids.p, c->hgrids.data(), n * sizeof(DcGridDev),
                             hipMemcpyHostToDevice, s), "upload descriptors");
    DC_TRY(c, hipMemcpyAsync(c->bounds.p, bounds.data(), bounds.size() * sizeof(int64_t),
                             hipMemcpyHostToDevice, s), "upload bounds");
    c->dev_key = key;
  }
  const Bufs b = context_bufs(c, precision);
  rc = run_captured(c, c->fwd, std::move(key), c->use_graph && !c->timing, s,
                    [&](hipStream_t st) { return enqueue(c, precision, st, b); });
  if (rc) return rc;
  DC_TRY(c, hipStreamSynchronize(s), "pipeline");
  return DC_HIP_OK;
}

Bufs context_bufs(dc_hip_context* c, int precision) {
  Bufs b;
  b.grids = as<DcGridDev>(c->grids);
  b.bounds = as<int64_t>(c->bounds);
  b.act = as<uint64_t>(c->act);
  b.sg = as<uint4>(c->sg);
  b.wprefix = as<uint32_t>(c->wprefix);
  b.ft = as<uint64_t>(c->ft);
  b.pcount = as<uint32_t>(c->pcount);
  b.pbase = as<uint32_t>(c->pbase);
  b.scan_sum = as<uint32_t>(c->scan_sum);
  b.scan_base = as<uint64_t>(c->scan_base);
  b.ctr = as<DcCounters>(c->counters);
  b.alist = as<uint64_t>(c->alist);
  b.v64 = as<double>(c->v64);
  b.v32 = as<float>(c->v32);
  b.flags = as<uint32_t>(c->flags);
  b.fblist = as<uint32_t>(c->fblist);
  b.faces = as<int32_t>(c->faces);
  b.pw = as<uint32_t>(c->pw);
  b.fnrm = c->fnrm.p;
  b.flist = as<uint64_t>(c->flist);
  b.goff = as<uint32_t>(c->goff);
  const bool fp32 = precision == DC_HIP_FP32_FALLBACK;
  b.fnrm32 = fp32 ? reinterpret_cast<float*>(static_cast<char*>(c->fnrm.p) + c->capF * 32)
                  : nullptr;
  b.reorient = fp32 ? as<uint64_t>(c->reorient) : nullptr;
  b.qstate = fp32 ? as<float>(c->qstate) : nullptr;
  b.rows = fp32 ? as<uint32_t>(c->bwd_rows) : nullptr;
  b.qstate64 = fp32 ? as<double>(c->qstate64) : nullptr;
  b.capA = c->capA;
  b.capF = c->capF;
  b.cap64 = c->cap64;
  b.timing = c->timing;
  return b;
}

int enqueue(dc_hip_context* c, int precision, hipStream_t s, const Bufs& b) {
  const int n = (int)c->hgrids.size();
  const int64_t t_a = dc_scan_tiles(c->n_words), t_p = dc_scan_tiles(c->n_pen);
  const int64_t face_blocks =
      std::max<int64_t>(1, std::min<int64_t>(((int64_t)b.capF + 255) / 256,
                                             (int64_t)c->cu_count * 16));
  const int64_t* d_tile_b = b.bounds;
  const int64_t* d_act_b = d_tile_b + (n + 1);
  const int64_t* d_pen_b = d_act_b + (n + 1);
  const int64_t* d_ft_b = d_tile_b + 5 * (n + 1);
  const DcGridDev* d_grids = b.grids;
  DcCounters* d_ctr = b.ctr;
  uint32_t* tsum = b.scan_sum;
  uint64_t* tbase = b.scan_base;

  // clears: kernels rather than memset nodes, so the sequence can be captured in a graph
  DC_TRY(c, dc_launch_fill_u32(reinterpret_cast<uint32_t*>(d_ctr),
                               (int64_t)(sizeof(DcCounters) / 4), 0u, s), "clear counters");

  if (b.timing) DC_TRY(c, hipEventRecord(c->ev[0], s), "event");
  DcMarkArgs ma;
  ma.grids = d_grids;
  ma.n_grids = n;
  ma.tile_bounds = d_tile_b;
  ma.n_tiles = c->n_tiles;
  ma.rows = c->eff_mark_rows;
  ma.waves_per_block = c->mark_waves;
  ma.sg = b.sg;
  ma.cw_bounds = d_tile_b + 6 * (n + 1);
  ma.n_cw = c->n_cw;
  ma.act = b.act;
  ma.ft = b.ft;
  ma.pcount = b.pcount;
  ma.ctr = d_ctr;
  DC_TRY(c, dc_launch_mark(ma, s), "mark");
  if (b.timing) DC_TRY(c, hipEventRecord(c->ev[7], s), "event");
  DC_TRY(c, dc_launch_mark_finish(ma, s), "cell pass");
  if (b.timing) DC_TRY(c, hipEventRecord(c->ev[1], s), "event");
  // With per-pass timing off the three scans (active words, pencil counts, face words) share
  // their launches (reduce, tile scan, apply: 3 launches instead of 9); with it on each runs by
  // itself between its pass's events.
  const bool shared_scans = !b.timing;
  if (shared_scans) {
    DcScan3Args sa;
    sa.act = b.act;
    sa.n_words = c->n_words;
    sa.wprefix = b.wprefix;
    sa.alist = b.alist;
    sa.capA = b.capA;
    sa.pcount = b.pcount;
    sa.n_pen = c->n_pen;
    sa.pbase = b.pbase;
    sa.ft = b.ft;
    sa.n_ft = c->n_ft;
    sa.flist = b.flist;
    sa.capF = b.capF;
    sa.tiles_a = t_a;
    sa.tiles_p = t_p;
    sa.tiles_f = dc_scan_tiles(c->n_ft);
    sa.tile_sum = tsum;
    sa.tile_base = tbase;
    sa.ctr = d_ctr;
    DC_TRY(c, dc_launch_scan3(sa, s), "scans");
  } else {
    DC_TRY(c, dc_launch_scan_act(b.act, c->n_words, tsum, tbase, b.wprefix, b.alist, b.capA,
                                 &d_ctr->n_active, s),
           "active scan");
    if (b.timing) DC_TRY(c, hipEventRecord(c->ev[2], s), "event");
    DC_TRY(c, dc_launch_scan_pen(b.pcount, c->n_pen, tsum + t_a, tbase + t_a, b.pbase,
                                 &d_ctr->n_faces, s),
           "pencil scan");
  }
  if (b.timing) DC_TRY(c, hipEventRecord(c->ev[3], s), "event");
  DcFaceNormalArgs fa;
  fa.grids = d_grids;
  fa.n_grids = n;
  fa.pen_bounds = d_pen_b;
  fa.n_pen = c->n_pen;
  fa.ft_bounds = d_ft_b;
  fa.n_ft = c->n_ft;
  fa.ft = b.ft;
  fa.pbase = b.pbase;
  fa.pw = b.pw;
  fa.flist = b.flist;
  fa.tile_sum = tsum + t_a + t_p;
  fa.tile_base = tbase + t_a + t_p;
  fa.ctr = d_ctr;
  fa.capF = b.capF;
  fa.face_blocks = face_blocks;
  fa.precision = precision;
  fa.nrm = b.fnrm;
  fa.nrm32 = b.fnrm32;
  fa.list_done = shared_scans ? 1 : 0;
  DC_TRY(c, dc_launch_face_normals(fa, s), "face normals");
  if (b.timing) DC_TRY(c, hipEventRecord(c->ev[4], s), "event");
  DcQefArgs qa;
  qa.grids = d_grids;
  qa.n_grids = n;
  qa.act_bounds = d_act_b;
  qa.act = b.act;
  qa.wprefix = b.wprefix;
  qa.alist = b.alist;
  qa.ctr = d_ctr;
  qa.capA = b.capA;
  qa.blocks = c->qef_blocks > 0 ? c->qef_blocks : (int64_t)c->cu_count * 8;
  qa.precision = precision;
  qa.v64 = b.v64;
  qa.v32 = b.v32;
  qa.flags = b.flags;
  qa.fblist = b.fblist;
  qa.ctr_rw = d_ctr;
  qa.ft = b.ft;
  qa.pw = b.pw;
  qa.nrm = b.fnrm32 ? static_cast<const void*>(b.fnrm32) : b.fnrm;
  qa.nrm64 = reinterpret_cast<const double*>(b.fnrm);
  qa.capF = b.capF;
  qa.qstate = b.qstate;
  qa.rows = b.rows;
  qa.qstate64 = b.qstate64;
  qa.cap64 = b.qstate64 ? b.cap64 : 0;
  DC_TRY(c, dc_launch_qef(qa, s), "qef");
  if (b.timing) DC_TRY(c, hipEventRecord(c->ev[5], s), "event");
  DcEmitArgs ea;
  ea.grids = d_grids;
  ea.n_grids = n;
  ea.ft_bounds = d_ft_b;
  ea.flist = b.flist;
  ea.face_blocks = face_blocks;
  ea.ft = b.ft;
  ea.pw = b.pw;
  ea.act = b.act;
  ea.wprefix = b.wprefix;
  ea.v64 = b.v64;
  ea.v32 = b.v32;
  ea.precision = precision;
  ea.ctr = d_ctr;
  ea.capF = b.capF;
  ea.faces = b.faces;
  ea.capA = b.capA;
  ea.reorient = b.reorient;
  ea.nrm64 = reinterpret_cast<const double*>(b.fnrm);
  DC_TRY(c, dc_launch_emit(ea, s), "emit");
  if (b.timing) DC_TRY(c, hipEventRecord(c->ev[6], s), "event");
  if (n > 1) {
    DC_TRY(c, dc_launch_grid_offsets(d_act_b, d_pen_b, n, b.wprefix, b.pbase, b.goff, s),
           "grid offsets");
    DC_TRY(c, hipMemcpyAsync(c->h_goff, b.goff, 8 * (size_t)n, hipMemcpyDeviceToHost, s),
           "read grid offsets");
  }
  DC_TRY(c, hipMemcpyAsync(c->h_counters, d_ctr, sizeof(DcCounters), hipMemcpyDeviceToHost, s),
         "read counters");
  return DC_HIP_OK;
}

int run_pipeline(dc_hip_context* c, const dc_hip_grid* grids, int n, int precision,
                 dc_hip_stats* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  c->have = false;
  if (precision != DC_HIP_PLAIN_FP64 && precision != DC_HIP_FP32_FALLBACK)
    return fail(c, DC_HIP_ERR_INVALID_ARG, "unknown precision");
  int rc = plan(c, grids, n);
  if (rc) return rc;
  int64_t cells = 0;
  for (const auto& g : c->hgrids) cells += (int64_t)g.cx * g.cy * g.cz;
  if (c->capA == 0) {
    c->capA = (uint64_t)std::min<int64_t>(cells, std::max<int64_t>(1 << 14, cells / 16));
    c->capF = c->capA * 2;
  }
  for (int attempt = 0; attempt < 2; ++attempt) {
    rc = run_once(c, precision);
    if (rc) return rc;
    const DcCounters& h = *c->h_counters;
    if (h.error)
      return fail(c, DC_HIP_ERR_INTERNAL,
                  std::string("device consistency check failed:") +
                      ((h.error & dcdecode::kErrCellRange) ? " an active-cell entry outside its grid" : "") +
                      ((h.error & dcdecode::kErrFaceRange) ? " a face entry outside its grid" : ""));
    if (h.n_list != h.n_faces) return fail(c, DC_HIP_ERR_INTERNAL, "face list count mismatch");
    if (h.nonfinite) return fail(c, DC_HIP_ERR_NONFINITE, "grid contains NaN or Inf samples");
    if (h.n_active >= (1ull << 31) || h.n_faces >= (1ull << 31))
      return fail(c, DC_HIP_ERR_INVALID_GRID, "mesh exceeds 32-bit vertex or face ids");
    if (h.n_active <= c->capA && h.n_faces <= c->capF) {
      c->nv = (int64_t)h.n_active;
      c->nf = (int64_t)h.n_faces;
      c->precision = precision;
      c->have = true;
      break;
    }
    c->capA = std::max<uint64_t>(c->capA, h.n_active + h.n_active / 16 + 64);
    c->capF = std::max<uint64_t>(c->capF, h.n_faces + h.n_faces / 16 + 64);
  }
  if (!c->have) return fail(c, DC_HIP_ERR_INTERNAL, "output capacity not reached after resize");
  if (stats) {
    dc_hip_stats& st = *stats;
    float up = st.upload_ms;
    std::memset(&st, 0, sizeof(st));
    st.upload_ms = up;
    st.nx = c->hgrids[0].nx;
    st.ny = c->hgrids[0].ny;
    st.nz = c->hgrids[0].nz;
    st.total_cells = cells;
    st.active_cells = c->nv;
    st.vertex_count = c->nv;
    st.face_count = c->nf;
    st.qef_fallback_count = (int64_t)c->h_counters->fallback;
    st.graph_mode = c->fwd.mode;
    if (c->timing) {
      st.marking_ms = elapsed(c, 0, 1);
      st.compaction_ms = elapsed(c, 1, 2);
      st.face_prefix_sum_ms = elapsed(c, 2, 3);
      st.normals_ms = elapsed(c, 3, 4);
      st.mark_kernel_ms = elapsed(c, 0, 7);
      st.qef_ms = elapsed(c, 4, 5);
      st.face_fill_ms = elapsed(c, 5, 6);
      st.face_emission_ms = st.face_prefix_sum_ms + st.face_fill_ms;
    }
    st.total_ms =
        std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count() +
        up;
  }
  return DC_HIP_OK;
}

}  // namespace

extern "C" {

const char* dc_hip_status_string(int status) {
  switch (status) {
    case DC_HIP_OK: return "ok";
    case DC_HIP_ERR_NO_DEVICE: return "no HIP device";
    case DC_HIP_ERR_INVALID_ARG: return "invalid argument";
    case DC_HIP_ERR_INVALID_GRID: return "invalid grid";
    case DC_HIP_ERR_NONFINITE: return "non-finite sample";
    case DC_HIP_ERR_OUT_OF_MEMORY: return "out of device memory";
    case DC_HIP_ERR_HIP: return "HIP runtime error";
    case DC_HIP_ERR_NO_FORWARD: return "no forward result";
    case DC_HIP_ERR_INTERNAL: return "internal error";
    default: return "unknown status";
  }
}

int dc_hip_create(int device, dc_hip_context** out) {
  if (!out) return DC_HIP_ERR_INVALID_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DC_HIP_ERR_NO_DEVICE;
  if (device < 0 || device >= count) return DC_HIP_ERR_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return DC_HIP_ERR_NO_DEVICE;
  dc_hip_context* c = new dc_hip_context();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    c->cu_count = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return DC_HIP_ERR_HIP;
  }
  c->stream = nullptr;  // the null stream's order (work_stream)
  for (int q = 0; q < kNumEvents; ++q) {
    if (hipEventCreate(&c->ev[q]) != hipSuccess) {
      dc_hip_destroy(c);
      return DC_HIP_ERR_HIP;
    }
  }
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_side, hipEventDisableTiming) != hipSuccess) {
    dc_hip_destroy(c);
    return DC_HIP_ERR_HIP;
  }
  if (hipHostMalloc((void**)&c->h_counters, sizeof(DcCounters), hipHostMallocDefault) !=
      hipSuccess) {
    dc_hip_destroy(c);
    return DC_HIP_ERR_OUT_OF_MEMORY;
  }
  std::memset(c->h_counters, 0, sizeof(DcCounters));
  if (hipHostMalloc((void**)&c->h_stage, kStageBytes, hipHostMallocDefault) != hipSuccess) {
    dc_hip_destroy(c);
    return DC_HIP_ERR_OUT_OF_MEMORY;
  }
  // the allocation-free entry point's set-up: the device address of the static scratch, and
  // host-side capacity for a single-grid plan and error messages, so that it allocates nothing
  init_scratch(device);
  c->hgrids.reserve(1);
  for (auto* v : {&c->tile_b, &c->act_b, &c->pen_b, &c->samp_b, &c->sr_b, &c->ft_b, &c->cw_b})
    v->reserve(2);
  c->err.reserve(512);
  *out = c;
  return DC_HIP_OK;
}

void dc_hip_destroy(dc_hip_context* c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->own) hipStreamSynchronize(c->own);
  if (c->stream) hipStreamSynchronize(c->stream);
  DevBuf* bufs[] = {&c->grids, &c->bounds, &c->act, &c->sg, &c->ft, &c->pcount, &c->pbase, &c->wprefix,
                    &c->alist, &c->v64, &c->v32, &c->faces, &c->scan_sum, &c->scan_base,
                    &c->counters, &c->upload, &c->flags, &c->bwd_contrib, &c->pw, &c->fnrm,
                    &c->ubar, &c->ubar_e, &c->bwd_rows, &c->fblist, &c->flist, &c->goff,
                    &c->qstate, &c->qstate64, &c->tris, &c->reorient};
  for (DevBuf* b : bufs)
    if (b->p) hipFree(b->p);
  delete c->post;
  delete c->m2s;
  delete c->sparse;
  for (int q = 0; q < kNumEvents; ++q)
    if (c->ev[q]) hipEventDestroy(c->ev[q]);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_side) hipEventDestroy(c->ev_side);
  drop_graph(c);
  if (c->h_counters) hipHostFree(c->h_counters);
  if (c->h_stage) hipHostFree(c->h_stage);
  if (c->h_goff) hipHostFree(c->h_goff);
  for (int q = 0; q < 2; ++q) {
    if (c->h_ring[q]) hipHostFree(c->h_ring[q]);
    if (c->ring_ev[q]) hipEventDestroy(c->ring_ev[q]);
  }
  if (c->side) {
    hipStreamSynchronize(c->side);
    hipStreamDestroy(c->side);
  }
  if (c->own) hipStreamDestroy(c->own);
  delete c;
}

int dc_hip_set_stream(dc_hip_context* c, void* stream) {
  if (!c) return DC_HIP_ERR_INVALID_ARG;
  c->stream = reinterpret_cast<hipStream_t>(stream);
  return DC_HIP_OK;
}

int dc_hip_set_option(dc_hip_context* c, const char* key, int64_t value) {
  if (!c || !key) return DC_HIP_ERR_INVALID_ARG;
  const std::string k(key);
  if (k == "mark_rows") {
    if (value != 0 && value != kMarkWideRows && value != 4 && value != 8 && value != 16 &&
        value != 32)
      return fail(c, DC_HIP_ERR_INVALID_ARG,
                  "mark_rows must be 0 (by grid size), 1 (whole rows), 4, 8, 16 or 32");
    c->mark_rows = (int)value;
  } else if (k == "mark_waves") {
    if (value != 1 && value != 2 && value != 4)
      return fail(c, DC_HIP_ERR_INVALID_ARG, "mark_waves must be 1, 2 or 4");
    c->mark_waves = (int)value;
  } else if (k == "gather_blocks") {
    if (value < 0) return fail(c, DC_HIP_ERR_INVALID_ARG, "gather_blocks must be >= 0");
    c->gather_blocks = value;
  } else if (k == "gather_rows") {
    if (value != 1 && value != 2 && value != 4 && value != 8)
      return fail(c, DC_HIP_ERR_INVALID_ARG, "gather_rows must be 1, 2, 4 or 8");
    c->gather_rows = (int)value;
  } else if (k == "graph") {
    c->use_graph = value != 0;
    if (!c->use_graph) drop_graph(c);
  } else if (k == "timing") {
    c->timing = value != 0;
  } else if (k == "adjoint_lanes" || k == "adjoint_face_lanes") {
    if (value != 0 && value != 1 && value != 4 && value != 16)
      return fail(c, DC_HIP_ERR_INVALID_ARG, k + " must be 0, 1, 4 or 16");
    (k == "adjoint_lanes" ? c->bwd_lanes1 : c->bwd_lanes2) = (int)value;
  } else if (k == "adjoint_fallback_lanes") {
    if (value != 0 && value != 1 && value != 16)
      return fail(c, DC_HIP_ERR_INVALID_ARG, "adjoint_fallback_lanes must be 0, 1 or 16");
    c->bwd_lanes64 = (int)value;
  } else if (k == "fallback_state") {
    if (value < -1) return fail(c, DC_HIP_ERR_INVALID_ARG, "fallback_state must be >= -1");
    c->cap64_opt = value;
  } else if (k == "qef_blocks") {
    if (value < 0) return fail(c, DC_HIP_ERR_INVALID_ARG, "qef_blocks must be >= 0");
    c->qef_blocks = value;
  } else {
    return fail(c, DC_HIP_ERR_INVALID_ARG, "unknown option " + k);
  }
  return DC_HIP_OK;
}

int dc_hip_extract_device(dc_hip_context* c, const dc_hip_grid* grid, int precision,
                          dc_hip_stats* stats) {
  if (!c || !grid) return fail(c, DC_HIP_ERR_INVALID_ARG, "NULL argument");
  DC_TRY(c, hipSetDevice(c->device), "set device");
  if (stats) stats->upload_ms = 0.f;
  return run_pipeline(c, grid, 1, precision, stats);
}

namespace {
// A large host grid is uploaded through two pinned chunks: host threads copy chunk q into one
// while the copy engine moves chunk q - 1 from the other, so the upload runs at the lower of the
// two rates instead of the runtime's single-threaded staging of pageable memory.
constexpr size_t kRingChunk = size_t(64) << 20;

void copy_threads(char* dst, const char* src, size_t len) {
  const unsigned hw = std::thread::hardware_concurrency();
  const unsigned t = std::max(1u, std::min(8u, hw ? hw / 2 : 1u));
  const size_t per = ((len + t - 1) / t + 63) & ~size_t(63);
  std::vector<std::thread> pool;
  for (unsigned q = 1; q < t && q * per < len; ++q)
    pool.emplace_back([=] { std::memcpy(dst + q * per, src + q * per, std::min(per, len - q * per)); });
  std::memcpy(dst, src, std::min(per, len));
  for (auto& th : pool) th.join();
}

int upload_pipelined(dc_hip_context* c, const float* host, size_t bytes) {
  for (int q = 0; q < 2; ++q) {
    if (!c->h_ring[q] && hipHostMalloc((void**)&c->h_ring[q], kRingChunk, hipHostMallocDefault) !=
                             hipSuccess) {
      c->h_ring[q] = nullptr;
      return fail(c, DC_HIP_ERR_OUT_OF_MEMORY, "pinned upload chunks");
    }
    if (!c->ring_ev[q])
      DC_TRY(c, hipEventCreateWithFlags(&c->ring_ev[q], hipEventDisableTiming), "upload event");
  }
  const char* src = reinterpret_cast<const char*>(host);
  char* dst = static_cast<char*>(c->upload.p);
  size_t q = 0;
  for (size_t off = 0; off < bytes; off += kRingChunk, ++q) {
    const size_t len = std::min(kRingChunk, bytes - off);
    if (q >= 2) DC_TRY(c, hipEventSynchronize(c->ring_ev[q & 1]), "upload");  // chunk q - 2 sent
    copy_threads(c->h_ring[q & 1], src + off, len);
    DC_TRY(c, hipMemcpyAsync(dst + off, c->h_ring[q & 1], len, hipMemcpyHostToDevice, c->stream),
           "upload");
    DC_TRY(c, hipEventRecord(c->ring_ev[q & 1], c->stream), "upload");
  }
  return DC_HIP_OK;
}
}  // namespace

int dc_hip_extract(dc_hip_context* c, const dc_hip_grid* grid, int precision,
                   dc_hip_stats* stats) {
  if (!c || !grid || !grid->values) return fail(c, DC_HIP_ERR_INVALID_ARG, "NULL argument");
  // everything the upload's size depends on is checked before the host buffer is read
  if (grid->nx < 2 || grid->ny < 2 || grid->nz < 2)
    return fail(c, DC_HIP_ERR_INVALID_GRID, "grid dimensions must be >= 2");
  if (grid->nx > 65536 || grid->ny > 65536 || grid->nz > 65536)
    return fail(c, DC_HIP_ERR_INVALID_GRID, "grid dimension above 65536");
  if (precision != DC_HIP_PLAIN_FP64 && precision != DC_HIP_FP32_FALLBACK)
    return fail(c, DC_HIP_ERR_INVALID_ARG, "unknown precision");
  DC_TRY(c, hipSetDevice(c->device), "set device");
  const size_t bytes = (size_t)grid->nx * grid->ny * grid->nz * sizeof(float);
  int rc = ensure(c, c->upload, bytes, "upload buffer", false);  // in the plan key as values
  if (rc) return rc;
  if (c->timing) DC_TRY(c, hipEventRecord(c->ev[8], c->stream), "event");
  if (bytes < 2 * kRingChunk) {
    DC_TRY(c, hipMemcpyAsync(c->upload.p, grid->values, bytes, hipMemcpyHostToDevice, c->stream),
           "upload");
  } else if ((rc = upload_pipelined(c, grid->values, bytes))) {
    return rc;
  }
  if (c->timing) DC_TRY(c, hipEventRecord(c->ev[9], c->stream), "event");
  DC_TRY(c, hipStreamSynchronize(c->stream), "upload");
  dc_hip_grid dg = *grid;
  dg.values = as<float>(c->upload);
  if (stats) stats->upload_ms = c->timing ? elapsed(c, 8, 9) : 0.f;
  return run_pipeline(c, &dg, 1, precision, stats);
}

int dc_hip_extract_batch(dc_hip_context* c, const dc_hip_grid* grids, int32_t n_grids,
                         int precision, int64_t* vertex_offsets, int64_t* face_offsets,
                         dc_hip_stats* stats) {
  if (!c || !grids || n_grids <= 0) return fail(c, DC_HIP_ERR_INVALID_ARG, "bad batch");
  DC_TRY(c, hipSetDevice(c->device), "set device");
  if (stats) stats->upload_ms = 0.f;
  int rc = run_pipeline(c, grids, n_grids, precision, stats);
  if (rc) return rc;
  if (vertex_offsets || face_offsets) {
    // per-grid offsets, read back with the counters (one grid: zero)
    for (int q = 0; q < n_grids; ++q) {
      if (vertex_offsets) vertex_offsets[q] = n_grids > 1 ? c->h_goff[2 * q] : 0;
      if (face_offsets) face_offsets[q] = n_grids > 1 ? c->h_goff[2 * q + 1] : 0;
    }
    if (vertex_offsets) vertex_offsets[n_grids] = c->nv;
    if (face_offsets) face_offsets[n_grids] = c->nf;
  }
  return DC_HIP_OK;
}

int dc_hip_extract_device_noalloc(dc_hip_context* c, const dc_hip_grid* grid, int precision,
                                  float* vertices, double* vertices_f64, int64_t vertex_capacity,
                                  int32_t* faces, int64_t face_capacity, int64_t* n_vertices,
                                  int64_t* n_faces) {
  if (!c || !grid || !n_vertices || !n_faces)
    return fail(c, DC_HIP_ERR_INVALID_ARG, "NULL argument");
  *n_vertices = 0;
  *n_faces = 0;
  c->have = false;  // this call's result lives in the caller's buffers, not in the context
  if (precision != DC_HIP_PLAIN_FP64 && precision != DC_HIP_FP32_FALLBACK)
    return fail(c, DC_HIP_ERR_INVALID_ARG, "unknown precision");
  if (vertex_capacity < 0 || face_capacity < 0)
    return fail(c, DC_HIP_ERR_INVALID_ARG, "negative output capacity");
  if ((vertex_capacity > 0 && !vertices) || (face_capacity > 0 && !faces))
    return fail(c, DC_HIP_ERR_INVALID_ARG, "NULL output buffer with a nonzero capacity");
  if (vertices_f64 && precision != DC_HIP_PLAIN_FP64)
    return fail(c, DC_HIP_ERR_INVALID_ARG, "FP64 vertices exist only in plain FP64 mode");
  if (c->device < 0 || c->device >= kMaxDevices || !g_scratch_ptr[c->device])
    return fail(c, DC_HIP_ERR_HIP, "static scratch unavailable on this device");
  DC_TRY(c, hipSetDevice(c->device), "set device");
  int rc = plan(c, grid, 1);
  if (rc) return rc;
  const DcGridDev& g = c->hgrids[0];
  const int64_t cells = (int64_t)g.cx * g.cy * g.cz;
  const uint64_t capA = (uint64_t)std::min<int64_t>(vertex_capacity, cells);
  const uint64_t capF = (uint64_t)std::min<int64_t>(face_capacity, 3 * cells);
  const int64_t t_a = dc_scan_tiles(c->n_words), t_p = dc_scan_tiles(c->n_pen),
                t_f = dc_scan_tiles(c->n_ft);
  const bool plain = precision == DC_HIP_PLAIN_FP64;
  // the layout: sized with a null base first, then carved from this device's copy
  auto layout = [&](unsigned char* base, Bufs& b) {
    Carver cv{base};
    b.grids = cv.take<DcGridDev>(sizeof(DcGridDev));
    b.bounds = cv.take<int64_t>(7 * 2 * sizeof(int64_t));
    b.ctr = cv.take<DcCounters>(sizeof(DcCounters));
    b.act = cv.take<uint64_t>((c->n_words + 1) * 8);  // one word past the end, as above
    b.sg = cv.take<uint4>(c->n_sg * 16);
    b.wprefix = cv.take<uint32_t>((c->n_words + 1) * 4);
    b.ft = cv.take<uint64_t>(c->n_ft * 8);
    b.pcount = cv.take<uint32_t>(c->n_pen * 4);
    b.pbase = cv.take<uint32_t>(c->n_pen * 4);
    b.scan_sum = cv.take<uint32_t>((t_a + t_p + t_f) * 4);
    b.scan_base = cv.take<uint64_t>((t_a + t_p + t_f) * 8);
    b.alist = cv.take<uint64_t>(capA * 8);
    b.v64 = (plain && !vertices_f64) ? cv.take<double>(capA * 24) : vertices_f64;
    b.reorient = plain ? nullptr : cv.take<uint64_t>(capF * 8);
    b.flags = cv.take<uint32_t>(plain ? 0 : capA * 4);
    b.fblist = cv.take<uint32_t>(plain ? 0 : capA * 4);
    b.pw = cv.take<uint32_t>(c->n_ft * 4);
    b.fnrm = cv.take<unsigned char>(capF * 32);
    b.fnrm32 = plain ? nullptr : cv.take<float>(capF * 16);
    b.flist = cv.take<uint64_t>(capF * 8);
    return cv.off;
  };
  Bufs b{};
  const size_t need = layout(nullptr, b);
  if (need > kScratchBytes) {
    char msg[160];
    std::snprintf(msg, sizeof(msg),
                  "grid and output capacities need %zu bytes of scratch; the static scratch "
                  "holds %zu",
                  need, kScratchBytes);
    return fail(c, DC_HIP_ERR_INVALID_GRID, (const char*)msg);
  }
  layout(g_scratch_ptr[c->device], b);
  b.v32 = vertices;
  b.faces = faces;
  b.goff = nullptr;  // one grid: no batch offsets
  b.qstate = nullptr;  // no adjoint state: the adjoint needs a regular forward
  b.rows = nullptr;
  b.qstate64 = nullptr;
  b.capA = capA;
  b.capF = capF;
  b.cap64 = 0;
  b.timing = false;
  // descriptor and bounds through the pinned staging (no host allocation, no pageable copy)
  const hipStream_t s = c->stream;
  std::memcpy(c->h_stage, &g, sizeof(DcGridDev));
  int64_t* hb = reinterpret_cast<int64_t*>(c->h_stage + 256);
  const std::vector<int64_t>* parts[7] = {&c->tile_b, &c->act_b, &c->pen_b, &c->samp_b,
                                          &c->sr_b, &c->ft_b, &c->cw_b};
  for (int q = 0; q < 7; ++q) {
    hb[2 * q] = (*parts[q])[0];
    hb[2 * q + 1] = (*parts[q])[1];
  }
  DC_TRY(c, hipMemcpyAsync(b.grids, c->h_stage, sizeof(DcGridDev), hipMemcpyHostToDevice, s),
         "upload descriptor");
  DC_TRY(c, hipMemcpyAsync(b.bounds, hb, 14 * sizeof(int64_t), hipMemcpyHostToDevice, s),
         "upload bounds");
  rc = enqueue(c, precision, s, b);
  if (rc) return rc;
  DC_TRY(c, hipStreamSynchronize(s), "pipeline");
  const DcCounters& h = *c->h_counters;
  if (h.error) return fail(c, DC_HIP_ERR_INTERNAL, "device consistency check failed");
  if (h.n_list != h.n_faces) return fail(c, DC_HIP_ERR_INTERNAL, "face list count mismatch");
  if (h.nonfinite) return fail(c, DC_HIP_ERR_NONFINITE, "grid contains NaN or Inf samples");
  *n_vertices = (int64_t)h.n_active;
  *n_faces = (int64_t)h.n_faces;
  if (h.n_active > (uint64_t)vertex_capacity || h.n_faces > (uint64_t)face_capacity)
    return fail(c, DC_HIP_ERR_INVALID_ARG,
                "output capacity too small for the mesh (required counts returned)");
  return DC_HIP_OK;
}

int dc_hip_autotune(dc_hip_context* c, const dc_hip_grid* grids, int32_t n_grids, int precision,
                    int32_t repeats, int32_t* best_rows, int32_t* best_waves, float* table_ms) {
  if (!c || !grids || n_grids <= 0 || repeats < 1)
    return fail(c, DC_HIP_ERR_INVALID_ARG, "bad autotune arguments");
  DC_TRY(c, hipSetDevice(c->device), "set device");
  static const int kRows[4] = {4, 8, 16, 32};
  static const int kWaves[3] = {1, 2, 4};
  const bool timing = c->timing;
  const int rows0 = c->mark_rows, waves0 = c->mark_waves;
  c->timing = true;
  float best = 0.f;
  int bl = rows0, bw = waves0, rc = DC_HIP_OK;
  for (int a = 0; a < 4 && rc == DC_HIP_OK; ++a) {
    for (int b = 0; b < 3 && rc == DC_HIP_OK; ++b) {
      c->mark_rows = kRows[a];
      c->mark_waves = kWaves[b];
      float t = 0.f;
      for (int r = 0; r <= repeats && rc == DC_HIP_OK; ++r) {  // run 0 warms up
        dc_hip_stats st{};
        rc = run_pipeline(c, grids, n_grids, precision, &st);
        if (r == 1 || (r > 1 && st.marking_ms < t)) t = st.marking_ms;
      }
      if (rc != DC_HIP_OK) break;
      if (table_ms) table_ms[3 * a + b] = t;
      if ((a == 0 && b == 0) || t < best) {
        best = t;
        bl = kRows[a];
        bw = kWaves[b];
      }
    }
  }
  c->timing = timing;
  if (rc != DC_HIP_OK) {
    c->mark_rows = rows0;
    c->mark_waves = waves0;
    return rc;
  }
  c->mark_rows = bl;
  c->mark_waves = bw;
  if (best_rows) *best_rows = bl;
  if (best_waves) *best_waves = bw;
  return DC_HIP_OK;
}

int dc_hip_triangulate(dc_hip_context* c, int32_t* triangles) {
  if (!c || !triangles) return fail(c, DC_HIP_ERR_INVALID_ARG, "NULL argument");
  if (!c->have) return fail(c, DC_HIP_ERR_NO_FORWARD, "no successful extraction");
  DC_TRY(c, hipSetDevice(c->device), "set device");
  if (c->nf == 0) return DC_HIP_OK;
  int rc;
  if ((rc = ensure(c, c->tris, (size_t)c->nf * 24, "triangles", false))) return rc;
  hipStream_t s;
  if ((rc = work_stream(c, &s))) return rc;
  const int n = (int)c->hgrids.size();
  const int64_t blocks =
      std::max<int64_t>(1, std::min<int64_t>((c->nf + 255) / 256, (int64_t)c->cu_count * 16));
  DC_TRY(c, dc_launch_triangulate(as<int32_t>(c->faces), (uint64_t)c->nf, n,
                                  as<uint32_t>(c->goff), as<double>(c->v64), as<float>(c->v32),
                                  c->precision == DC_HIP_PLAIN_FP64 ? 1 : 0, as<int32_t>(c->tris),
                                  blocks, s),
         "triangulate");
  // hipMemcpyDefault: the destination may be host or device memory (unified addressing)
  DC_TRY(c, hipMemcpyAsync(triangles, c->tris.p, (size_t)c->nf * 24, hipMemcpyDefault, s),
         "copy triangles");
  DC_TRY(c, hipStreamSynchronize(s), "triangulate");
  return DC_HIP_OK;
}

int dc_hip_result_counts(dc_hip_context* c, int64_t* nv, int64_t* nf) {
  if (!c) return DC_HIP_ERR_INVALID_ARG;
  if (!c->have) return fail(c, DC_HIP_ERR_NO_FORWARD, "no successful extraction");
  if (nv) *nv = c->nv;
  if (nf) *nf = c->nf;
  return DC_HIP_OK;
}

int dc_hip_copy_result(dc_hip_context* c, float* vertices, double* vertices_f64, int32_t* faces) {
  if (!c) return DC_HIP_ERR_INVALID_ARG;
  if (!c->have) return fail(c, DC_HIP_ERR_NO_FORWARD, "no successful extraction");
  DC_TRY(c, hipSetDevice(c->device), "set device");
  // hipMemcpyDefault: destinations may be host or device memory (unified addressing)
  if (vertices && c->nv)
    DC_TRY(c, hipMemcpyAsync(vertices, c->v32.p, c->nv * 12, hipMemcpyDefault, c->stream),
           "copy vertices");
  if (vertices_f64 && c->nv) {
    if (c->precision != DC_HIP_PLAIN_FP64)
      return fail(c, DC_HIP_ERR_INVALID_ARG, "FP64 vertices exist only in plain FP64 mode");
    DC_TRY(c, hipMemcpyAsync(vertices_f64, c->v64.p, c->nv * 24, hipMemcpyDefault, c->stream),
           "copy fp64 vertices");
  }
  if (faces && c->nf)
    DC_TRY(c, hipMemcpyAsync(faces, c->faces.p, c->nf * 16, hipMemcpyDefault, c->stream),
           "copy faces");
  DC_TRY(c, hipStreamSynchronize(c->stream), "download");
  return DC_HIP_OK;
}

int dc_hip_result_device(dc_hip_context* c, const float** vertices, const double** vertices_f64,
                         const int32_t** faces) {
  if (!c) return DC_HIP_ERR_INVALID_ARG;
  if (!c->have) return fail(c, DC_HIP_ERR_NO_FORWARD, "no successful extraction");
  if (vertices) *vertices = as<float>(c->v32);
  if (vertices_f64) *vertices_f64 = (c->precision == DC_HIP_PLAIN_FP64) ? as<double>(c->v64) : nullptr;
  if (faces) *faces = as<int32_t>(c->faces);
  return DC_HIP_OK;
}

int dc_hip_postprocess(dc_hip_context* c, const float* vertices, int64_t n_vertices,
                       const int32_t* faces, int64_t n_faces, int close_holes,
                       int remove_floaters, int64_t* out_n_vertices, int64_t* out_n_faces,
                       dc_hip_post_stats* stats) {
  if (!c) return DC_HIP_ERR_INVALID_ARG;
  DC_TRY(c, hipSetDevice(c->device), "set device");
  if (!vertices && !faces && n_vertices == 0 && n_faces == 0) {  // the last extraction's mesh
    if (!c->have) return fail(c, DC_HIP_ERR_NO_FORWARD, "no successful extraction");
    if (c->hgrids.size() != 1)
      return fail(c, DC_HIP_ERR_INVALID_ARG,
                  "postprocess of the last result needs a single-grid extraction (a batch's face "
                  "indices are local to each grid)");
    vertices = as<float>(c->v32);
    faces = as<int32_t>(c->faces);
    n_vertices = c->nv;
    n_faces = c->nf;
  }
  if (!c->post) c->post = new DcPost();
  hipStream_t s;
  int rc;
  if ((rc = work_stream(c, &s))) return rc;
  rc = c->post->run(vertices, n_vertices, faces, n_faces, close_holes != 0, remove_floaters != 0,
                    c->cu_count, s, stats, &c->err);
  if (rc != DC_HIP_OK) return rc;
  if (out_n_vertices) *out_n_vertices = c->post->n_vertices();
  if (out_n_faces) *out_n_faces = c->post->n_faces();
  return DC_HIP_OK;
}

int dc_hip_mesh_to_sdf(dc_hip_context* c, const float* vertices, int64_t num_vertices,
                       const int32_t* faces, int64_t num_faces, int32_t nx, int32_t ny,
                       int32_t nz, float ox, float oy, float oz, float vx, float vy, float vz,
                       int voxelize_first, int32_t voxel_res, float* values,
                       dc_hip_m2s_stats* stats) {
  if (!c) return DC_HIP_ERR_INVALID_ARG;
  DC_TRY(c, hipSetDevice(c->device), "set device");
  if (!c->m2s) c->m2s = new DcM2s();
  hipStream_t s;
  int rc;
  if ((rc = work_stream(c, &s))) return rc;
  const int32_t dims[3] = {nx, ny, nz};
  const float origin[3] = {ox, oy, oz}, spacing[3] = {vx, vy, vz};
  return c->m2s->run(vertices, num_vertices, faces, num_faces, dims, origin, spacing,
                     voxelize_first != 0, voxel_res, values, c->cu_count, s, stats, &c->err);
}

namespace {
int run_sparse(dc_hip_context* c, const dc_hip_grid* grid, bool on_device, int32_t brick_size,
               int normal_mode, int multi_vertex, dc_hip_sparse_stats* stats) {
  if (!c || !grid) return fail(c, DC_HIP_ERR_INVALID_ARG, "NULL argument");
  DC_TRY(c, hipSetDevice(c->device), "set device");
  if (!c->sparse) c->sparse = new DcSparse();
  hipStream_t s;
  int rc;
  if ((rc = work_stream(c, &s))) return rc;
  const bool sane = grid->values && grid->nx >= 2 && grid->ny >= 2 && grid->nz >= 2 &&
                    grid->nx <= 65536 && grid->ny <= 65536 && grid->nz <= 65536;
  const size_t bytes = sane ? (size_t)grid->nx * grid->ny * grid->nz * sizeof(float) : 0;
  if (on_device || bytes < 2 * kRingChunk)
    return c->sparse->run(*grid, on_device, brick_size, normal_mode, multi_vertex != 0,
                          c->cu_count, s, stats, &c->err);
  // a large host grid: the pipelined upload into the context's upload buffer, then the device path
  const auto t0 = std::chrono::steady_clock::now();
  if ((rc = ensure(c, c->upload, bytes, "upload buffer", false))) return rc;
  if ((rc = upload_pipelined(c, grid->values, bytes))) return rc;
  DC_TRY(c, hipStreamSynchronize(c->stream), "upload");
  const float up_ms =
      std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  dc_hip_grid dg = *grid;
  dg.values = as<float>(c->upload);
  rc = c->sparse->run(dg, true, brick_size, normal_mode, multi_vertex != 0, c->cu_count, s, stats,
                      &c->err);
  if (rc == DC_HIP_OK && stats) {
    stats->upload_ms = up_ms;
    stats->total_ms += up_ms;
  }
  return rc;
}
}  // namespace

int dc_hip_extract_sparse(dc_hip_context* c, const dc_hip_grid* grid, int32_t brick_size,
                          int normal_mode, int multi_vertex, dc_hip_sparse_stats* stats) {
  return run_sparse(c, grid, false, brick_size, normal_mode, multi_vertex, stats);
}

int dc_hip_extract_sparse_device(dc_hip_context* c, const dc_hip_grid* grid, int32_t brick_size,
                                 int normal_mode, int multi_vertex, dc_hip_sparse_stats* stats) {
  return run_sparse(c, grid, true, brick_size, normal_mode, multi_vertex, stats);
}

int dc_hip_sparse_result_device(dc_hip_context* c, int64_t* n_vertices, int64_t* n_faces,
                                const float** vertices, const int32_t** faces) {
  if (!c) return DC_HIP_ERR_INVALID_ARG;
  if (!c->sparse || !c->sparse->have())
    return fail(c, DC_HIP_ERR_NO_FORWARD, "no successful sparse extraction");
  if (n_vertices) *n_vertices = c->sparse->n_vertices();
  if (n_faces) *n_faces = c->sparse->n_faces();
  if (vertices) *vertices = c->sparse->vertices();
  if (faces) *faces = c->sparse->faces();
  return DC_HIP_OK;
}

int dc_hip_sparse_copy_result(dc_hip_context* c, float* vertices, int32_t* faces) {
  if (!c) return DC_HIP_ERR_INVALID_ARG;
  if (!c->sparse || !c->sparse->have())
    return fail(c, DC_HIP_ERR_NO_FORWARD, "no successful sparse extraction");
  DC_TRY(c, hipSetDevice(c->device), "set device");
  const int64_t nv = c->sparse->n_vertices(), nf = c->sparse->n_faces();
  // hipMemcpyDefault: destinations may be host or device memory (unified addressing)
  if (vertices && nv)
    DC_TRY(c, hipMemcpyAsync(vertices, c->sparse->vertices(), nv * 12, hipMemcpyDefault,
                             c->stream),
           "copy vertices");
  if (faces && nf)
    DC_TRY(c, hipMemcpyAsync(faces, c->sparse->faces(), nf * 16, hipMemcpyDefault, c->stream),
           "copy faces");
  DC_TRY(c, hipStreamSynchronize(c->stream), "download");
  return DC_HIP_OK;
}

int dc_hip_malloc(dc_hip_context* c, int64_t bytes, void** ptr) {
  if (!c || !ptr || bytes < 0) return fail(c, DC_HIP_ERR_INVALID_ARG, "NULL argument or bytes < 0");
  *ptr = nullptr;
  DC_TRY(c, hipSetDevice(c->device), "set device");
  const hipError_t e = hipMalloc(ptr, (size_t)std::max<int64_t>(bytes, 1));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return fail(c, e == hipErrorOutOfMemory ? DC_HIP_ERR_OUT_OF_MEMORY : DC_HIP_ERR_HIP,
                std::string("device allocation: ") + hipGetErrorString(e));
  }
  return DC_HIP_OK;
}

int dc_hip_free(void* ptr) {
  if (!ptr) return DC_HIP_OK;
  return hipFree(ptr) == hipSuccess ? DC_HIP_OK : DC_HIP_ERR_HIP;
}

int dc_hip_postprocess_copy(dc_hip_context* c, float* vertices, int32_t* faces) {
  if (!c) return DC_HIP_ERR_INVALID_ARG;
  if (!c->post || !c->post->have())
    return fail(c, DC_HIP_ERR_NO_FORWARD, "no successful postprocess");
  DC_TRY(c, hipSetDevice(c->device), "set device");
  const int64_t nv = c->post->n_vertices(), nf = c->post->n_faces();
  // hipMemcpyDefault: destinations may be host or device memory (unified addressing)
  if (vertices && nv)
    DC_TRY(c, hipMemcpyAsync(vertices, c->post->vertices(), nv * 12, hipMemcpyDefault, c->stream),
           "copy vertices");
  if (faces && nf)
    DC_TRY(c, hipMemcpyAsync(faces, c->post->faces(), nf * 16, hipMemcpyDefault, c->stream),
           "copy faces");
  DC_TRY(c, hipStreamSynchronize(c->stream), "download");
  return DC_HIP_OK;
}

int dc_hip_postprocess_device(dc_hip_context* c, const float** vertices, const int32_t** faces) {
  if (!c) return DC_HIP_ERR_INVALID_ARG;
  if (!c->post || !c->post->have())
    return fail(c, DC_HIP_ERR_NO_FORWARD, "no successful postprocess");
  if (vertices) *vertices = c->post->vertices();
  if (faces) *faces = c->post->faces();
  return DC_HIP_OK;
}

int dc_hip_backward(dc_hip_context* c, const float* grad_vertices, float* grad_values,
                    dc_hip_stats* stats) {
  if (!c || !grad_values) return fail(c, DC_HIP_ERR_INVALID_ARG, "NULL argument");
  if (!c->have) return fail(c, DC_HIP_ERR_NO_FORWARD, "no successful extraction");
  if (c->nv > 0 && !grad_vertices) return fail(c, DC_HIP_ERR_INVALID_ARG, "NULL grad_vertices");
  // the adjoint reads the forward's samples again while it writes the gradient: the two must not
  // share memory
  {
    const uintptr_t g0 = (uintptr_t)grad_values, g1 = g0 + (uintptr_t)c->n_samples * 4;
    for (const DcGridDev& g : c->hgrids) {
      const uintptr_t v0 = (uintptr_t)g.val, v1 = v0 + (uintptr_t)g.nx * g.ny * g.nz * 4;
      if (g0 < v1 && v0 < g1)
        return fail(c, DC_HIP_ERR_INVALID_ARG,
                    "grad_values overlaps the samples of the last forward (which the adjoint "
                    "reads again)");
    }
  }
  DC_TRY(c, hipSetDevice(c->device), "set device");
  const auto t0 = std::chrono::steady_clock::now();
  const int n = (int)c->hgrids.size();
  int rc;
  const size_t nv1 = (size_t)std::max<int64_t>(c->nv, 1);
  if ((rc = ensure(c, c->bwd_contrib, nv1 * 8 * 4, "adjoint contributions", false))) return rc;
  if ((rc = ensure(c, c->ubar, (size_t)std::max<int64_t>(c->nf, 1) * 48, "face adjoint slots",
                   false)))
    return rc;
  if ((rc = ensure(c, c->ubar_e, nv1 * 12 * 12, "edge adjoint slots", false))) return rc;
  if ((rc = ensure(c, c->bwd_rows, nv1 * 12 * 4, "adjoint row kinds", false))) return rc;
  hipStream_t s;
  if ((rc = work_stream(c, &s))) return rc;
  const int64_t* d_tile_b = as<int64_t>(c->bounds);
  DcBwdArgs a;
  std::memset(&a, 0, sizeof(a));  // its bytes are part of the capture key
  a.grids = as<DcGridDev>(c->grids);
  a.n_grids = n;
  a.act_bounds = d_tile_b + (n + 1);
  a.sr_bounds = d_tile_b + 4 * (n + 1);
  a.n_sr = c->n_sr;
  a.gather_rows = c->eff_gather_rows;
  a.gather_blocks = c->gather_blocks;
  a.gather_lds_row = c->gather_lds_row;
  a.act = as<uint64_t>(c->act);
  a.wprefix = as<uint32_t>(c->wprefix);
  a.alist = as<uint64_t>(c->alist);
  a.flags = as<uint32_t>(c->flags);
  a.fblist = as<uint32_t>(c->fblist);
  a.precision = c->precision;
  a.ctr = as<DcCounters>(c->counters);
  a.capA = c->capA;
  const int64_t n_fb = (int64_t)c->h_counters->fallback;
  a.fb_state = n_fb <= (int64_t)c->cap64 ? 1 : 0;
  a.n_cells = c->nv;
  a.n_fallback = n_fb;
  // both phases edge-parallel (4 lanes per cell) while one thread per cell would leave the
  // device idle (at 128^3: phase 1 25 -> 14 us, phase 2 30 -> 12 us)
  const bool small = c->nv <= (int64_t)c->cu_count * kWideCellsPerCu;
  a.lanes1 = c->bwd_lanes1 ? c->bwd_lanes1 : (small ? 4 : 1);
  a.lanes2 = c->bwd_lanes2 ? c->bwd_lanes2 : (small ? 4 : 1);
  // the FP64 cells: 16 lanes while few (they finish early on the side stream); when many, the
  // 16-fold per-cell setup competes with the FP32 cells' kernel for the device
  a.lanes64 = c->bwd_lanes64 ? c->bwd_lanes64 : (n_fb <= (int64_t)c->cu_count * 32 ? 16 : 1);
  a.grad_vertices = grad_vertices;
  a.ft = as<uint64_t>(c->ft);
  a.pw = as<uint32_t>(c->pw);
  a.nrm = c->fnrm.p;
  a.nrm32 = c->precision == DC_HIP_FP32_FALLBACK
                ? reinterpret_cast<const float*>(static_cast<char*>(c->fnrm.p) + c->capF * 32)
                : nullptr;
  a.capF = c->capF;
  a.ubar = as<float>(c->ubar);
  a.ubar_e = as<float>(c->ubar_e);
  a.rows = as<uint32_t>(c->bwd_rows);
  a.face_blocks = std::max<int64_t>(1, (c->nf + 255) / 256);
  a.contrib = as<float>(c->bwd_contrib);
  a.qstate = c->precision == DC_HIP_FP32_FALLBACK ? as<float>(c->qstate) : nullptr;
  a.qstate64 = c->precision == DC_HIP_FP32_FALLBACK ? as<double>(c->qstate64) : nullptr;
  a.grad_values = grad_values;
  if (c->timing) DC_TRY(c, hipEventRecord(c->ev[8], s), "event");
  // capture key: every launch argument (pointers, sizes, block counts), the stream and the plan
  // whose descriptors are on the device
  std::string key(reinterpret_cast<const char*>(&a), sizeof(a));
  key.append(reinterpret_cast<const char*>(&s), sizeof(s));
  key += c->dev_key;
  rc = run_captured(c, c->bwd, std::move(key), c->use_graph && !c->timing, s,
                    [&](hipStream_t st) -> int {
                      // FP32 mode: phase 1 of the FP64 cells runs beside the FP32 cells' on the
                      // side stream (disjoint cells and slots), joined before the face pass;
                      // phase 2 is one kernel over all cells. (Zeroing the gradient on the side
                      // stream beside the per-cell phases, with the gather storing only rows with
                      // contributions, was measured slower: the phases slowed 137 -> 226 us at
                      // 512^3 for a gather 31 us shorter.)
                      for (int phase = 1; phase <= 2; ++phase) {
                        const bool fork = a.precision != DC_PREC_FP64 && phase == 1;
                        if (fork) {
                          DC_TRY(c, hipEventRecord(c->ev_fork, st), "fork");
                          DC_TRY(c, hipStreamWaitEvent(c->side, c->ev_fork, 0), "fork");
                        }
                        DC_TRY(c, dc_launch_bwd_phase(a, phase, st, fork ? c->side : st),
                               "backward cells");
                        if (fork) {
                          DC_TRY(c, hipEventRecord(c->ev_side, c->side), "join");
                          DC_TRY(c, hipStreamWaitEvent(st, c->ev_side, 0), "join");
                        }
                        if (phase == 1)
                          DC_TRY(c, dc_launch_bwd_face_adjoint(a, st), "backward faces");
                      }
                      DC_TRY(c, dc_launch_gather(a, st), "backward gather");
                      return DC_HIP_OK;
                    });
  if (rc) return rc;
  // Without per-pass timing the work stays enqueued, on the null stream as on a caller's: stream
  // order covers every later use of the gradient, and the host can queue the next forward while
  // the adjoint runs (at 512^3 this hid the ~13 us between a synchronised adjoint and the next
  // forward's first kernel). With timing on the call waits for its events.
  if (c->timing) DC_TRY(c, hipEventRecord(c->ev[9], s), "event");
  if (c->timing) DC_TRY(c, hipStreamSynchronize(s), "backward");
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->vertex_count = c->nv;
    stats->face_count = c->nf;
    stats->backward_ms = c->timing ? elapsed(c, 8, 9) : 0.f;
    stats->graph_mode = c->bwd.mode;
    stats->total_ms =
        std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return DC_HIP_OK;
}

const char* dc_hip_last_error(dc_hip_context* c) {
  if (!c) return "no context";
  return c->err.c_str();
}

}  // extern "C"
