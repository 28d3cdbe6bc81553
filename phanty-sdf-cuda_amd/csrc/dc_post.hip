// dc_post.hip — mesh post-processing on the GPU (the reference's postprocess_mesh,
// src/cpp/dc_backend.cpp:127-558): vertex weld, floater removal and hole closing.
//
// Every O(mesh) step is a data-parallel pass whose result does not depend on thread order:
//   weld      quantised keys in an open-addressing table that keeps the lowest vertex index per key
//             (the reference's first-seen vertex), representatives numbered by an exclusive scan;
//             faces with fewer than 3 distinct vertices dropped, the rest compacted in order;
//   edges     undirected edges in a 64-bit-key table with a use count and the first two uses;
//   floaters  union-find over the faces joined by edges used exactly twice, hooking the larger
//             root under the smaller one, so each root is its component's lowest face, which is
//             the reference's component order (components numbered by their first face); the
//             largest component (ties: lowest first face) by one 64-bit atomic maximum;
//   renumber  each vertex's first position in the final face list by an atomic minimum, positions
//             that are a first use numbered by an exclusive scan (the reference's first-use order).
// The boundary-loop walk and fill (dc_backend.cpp:395-534) depend on the order edges are taken
// from each vertex's list, so they run on the host over the boundary edges alone, which the GPU
// compacts (typically a few thousand of a mesh's million edges). Host arithmetic is built without
// contraction, so each FP32 operation rounds where the reference's does.

#include "dc_post.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "dc_common.h"
#include "dc_flagscan.h"
#include "dc_unionfind.h"

namespace {

using dccommon::wave_scan_incl;
using namespace dcscan;
using namespace dcuf;

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr unsigned long long kEmptyKey = ~0ull;
constexpr int64_t kMaxFaces = int64_t(1) << 28;     // 4 * faces positions fit in 31 bits
constexpr int64_t kMaxVertices = int64_t(1) << 30;
constexpr size_t kMaxLoop = 4096;                   // longest loop the reference fills (:508, :525)
constexpr int64_t kMaxRings = int64_t(1) << 20;     // ring-fill iterations before giving up

// Totals read back to the host (slots of kCnt)
enum { kCntWeldV = 0, kCntWeldF, kCntKeep, kCntBnd, kCntFirst, kCntErr, kCntComp, kCntSlots = 8 };

int64_t blocks_for(int64_t n, int cu) {
  return std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, (int64_t)cu * 8));
}
size_t pow2_at_least(size_t n) {
  size_t p = 1024;
  while (p < n) p <<= 1;
  return p;
}

__global__ void k_fill32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill64(unsigned long long* __restrict__ p, int64_t n, unsigned long long v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

#define GRID_STRIDE(i, n)                                                          \
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)(n); \
       i += (int64_t)gridDim.x * blockDim.x)

// ---------------------------------------------------------------------------------------------
// Weld (merge_vertices_in_place, dc_backend.cpp:127-215)
// ---------------------------------------------------------------------------------------------
// Per-block bounding boxes: 6 floats per block (min x, y, z, max x, y, z).
__global__ void __launch_bounds__(kThreads)
    k_bbox(const float* __restrict__ v, int64_t n, float* __restrict__ part) {
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  GRID_STRIDE(i, n) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float x = v[3 * i + c];
      if (x < mn[c]) mn[c] = x;
      if (x > mx[c]) mx[c] = x;
    }
  }
  __shared__ float s[6][kThreads];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    s[c][threadIdx.x] = mn[c];
    s[3 + c][threadIdx.x] = mx[c];
  }
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        s[c][threadIdx.x] = fminf(s[c][threadIdx.x], s[c][threadIdx.x + w]);
        s[3 + c][threadIdx.x] = fmaxf(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// The quantisation step 1 / eps, eps = 1e-5 x the box diagonal (1e-5 for a single point), in the
// reference's FP32 operations (dc_backend.cpp:144-148). The square root and the division are done
// in FP64 and rounded once to FP32: FP64 carries more than 2 x 24 + 2 bits, so that equals the
// correctly rounded FP32 result.
__global__ void __launch_bounds__(kThreads)
    k_weld_step(const float* __restrict__ part, int nb, float* __restrict__ inv) {
  __shared__ float s[6][kThreads];
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < nb; b += kThreads)
    for (int c = 0; c < 3; ++c) {
      mn[c] = fminf(mn[c], part[6 * b + c]);
      mx[c] = fmaxf(mx[c], part[6 * b + 3 + c]);
    }
  for (int c = 0; c < 3; ++c) {
    s[c][threadIdx.x] = mn[c];
    s[3 + c][threadIdx.x] = mx[c];
  }
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 3; ++c) {
        s[c][threadIdx.x] = fminf(s[c][threadIdx.x], s[c][threadIdx.x + w]);
        s[3 + c][threadIdx.x] = fmaxf(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const float dx = s[3][0] - s[0][0], dy = s[4][0] - s[1][0], dz = s[5][0] - s[2][0];
  const float dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
  const float sum = (dx2 + dy2) + dz2;
  const float diag = (float)sqrt((double)sum);
  const float eps = diag > 0.0f ? diag * 1e-5f : 1e-5f;
  *inv = (float)(1.0 / (double)eps);
}

struct QKey {
  long long x, y, z;
};
__device__ __forceinline__ QKey quantize(const float* __restrict__ v, int64_t i, float inv) {
  // std::round: half away from zero, as roundf
  return QKey{(long long)roundf(v[3 * i] * inv), (long long)roundf(v[3 * i + 1] * inv),
              (long long)roundf(v[3 * i + 2] * inv)};
}
__device__ __forceinline__ uint64_t mix64(uint64_t h) {
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return h;
}
__device__ __forceinline__ uint64_t hash_key(const QKey& q) {
  return mix64((uint64_t)q.x * 0x9E3779B97F4A7C15ull ^ (uint64_t)q.y * 0xC2B2AE3D27D4EB4Full ^
               (uint64_t)q.z * 0x165667B19E3779F9ull);
}

// Each vertex finds the table slot of its key (claiming an empty one) and lowers the slot's
// vertex to its own index if smaller: afterwards every slot holds its key's lowest vertex. A
// slot's key never changes once claimed, so a stale read costs only a retry of the CAS.
__global__ void __launch_bounds__(kThreads)
    k_weld_insert(const float* __restrict__ v, int64_t n, const float* __restrict__ inv_p,
                  uint32_t* __restrict__ table, uint64_t mask, uint32_t* __restrict__ slot) {
  const float inv = *inv_p;
  GRID_STRIDE(i, n) {
    const QKey q = quantize(v, i, inv);
    uint64_t h = hash_key(q) & mask;
    while (true) {
      uint32_t c = table[h];
      if (c == kNone) {
        c = atomicCAS(table + h, kNone, (uint32_t)i);
        if (c == kNone) break;
      }
      const QKey o = quantize(v, c, inv);
      if (o.x == q.x && o.y == q.y && o.z == q.z) {
        atomicMin(table + h, (uint32_t)i);
        break;
      }
      h = (h + 1) & mask;
    }
    slot[i] = (uint32_t)h;
  }
}

__global__ void k_weld_flags(const uint32_t* __restrict__ table, const uint32_t* __restrict__ slot,
                             int64_t n, uint32_t* __restrict__ flag) {
  GRID_STRIDE(i, n) flag[i] = table[slot[i]] == (uint32_t)i ? 1u : 0u;
}

// remap[i] = new index of vertex i's representative; representatives copied to their place.
__global__ void k_weld_out(const float* __restrict__ v, int64_t n, const uint32_t* __restrict__ table,
                           const uint32_t* __restrict__ slot, const uint32_t* __restrict__ idx,
                           uint32_t* __restrict__ remap, float* __restrict__ vout) {
  GRID_STRIDE(i, n) {
    const uint32_t r = table[slot[i]];
    remap[i] = idx[r];
    if (r == (uint32_t)i) {
      const uint32_t o = idx[i];
      vout[3 * (int64_t)o] = v[3 * i];
      vout[3 * (int64_t)o + 1] = v[3 * i + 1];
      vout[3 * (int64_t)o + 2] = v[3 * i + 2];
    }
  }
}

// Faces through the remap; a face keeps unless fewer than 3 of its indices differ
// (dc_backend.cpp:193-207). An index outside [0, nv) sets *err and drops the face.
__global__ void k_weld_faces(const int4* __restrict__ f, int64_t nf, int64_t nv,
                             const uint32_t* __restrict__ remap, int4* __restrict__ ftmp,
                             uint32_t* __restrict__ keep, uint32_t* __restrict__ err) {
  GRID_STRIDE(q, nf) {
    const int4 a = f[q];
    if (a.x < 0 || a.y < 0 || a.z < 0 || a.w < 0 || a.x >= nv || a.y >= nv || a.z >= nv ||
        a.w >= nv) {
      atomicOr(err, 1u);
      keep[q] = 0;
      continue;
    }
    const int i0 = (int)remap[a.x], i1 = (int)remap[a.y], i2 = (int)remap[a.z],
              i3 = (int)remap[a.w];
    // distinct values among the four
    const int distinct = 1 + (i1 != i0) + (i2 != i0 && i2 != i1) + (i3 != i0 && i3 != i1 && i3 != i2);
    ftmp[q] = make_int4(i0, i1, i2, i3);
    keep[q] = distinct >= 3 ? 1u : 0u;
  }
}

__global__ void k_compact4(const int4* __restrict__ in, int64_t n, const uint32_t* __restrict__ keep,
                           const uint32_t* __restrict__ pos, int4* __restrict__ out) {
  GRID_STRIDE(q, n) if (keep[q]) out[pos[q]] = in[q];
}

// ---------------------------------------------------------------------------------------------
// Undirected edges (dc_backend.cpp:235-311)
// ---------------------------------------------------------------------------------------------
// Edge slot s = 4 f + e joins faces[f][e] and faces[f][(e + 1) % 4]; slots with equal endpoints
// take no edge (eslot = kNone). Each edge's table entry counts its slots and records the first two.
__global__ void __launch_bounds__(kThreads)
    k_edge_insert(const int4* __restrict__ f, int64_t nslots, unsigned long long* __restrict__ etab,
                  uint32_t* __restrict__ ecnt, uint32_t* __restrict__ epair, uint64_t mask,
                  uint32_t* __restrict__ eslot) {
  GRID_STRIDE(s, nslots) {
    const int4 a = f[s >> 2];
    const int e = (int)(s & 3);
    const int u = e == 0 ? a.x : e == 1 ? a.y : e == 2 ? a.z : a.w;
    const int w = e == 0 ? a.y : e == 1 ? a.z : e == 2 ? a.w : a.x;
    if (u == w) {
      eslot[s] = kNone;
      continue;
    }
    const unsigned long long key =
        ((unsigned long long)(uint32_t)min(u, w) << 32) | (uint32_t)max(u, w);
    uint64_t h = mix64(key) & mask;
    while (true) {
      unsigned long long c = etab[h];
      if (c == kEmptyKey) {
        c = atomicCAS(etab + h, kEmptyKey, key);
        if (c == kEmptyKey || c == key) break;
      }
      if (c == key) break;
      h = (h + 1) & mask;
    }
    const uint32_t k = atomicAdd(ecnt + h, 1u);
    if (k < 2) epair[2 * h + k] = (uint32_t)s;
    eslot[s] = (uint32_t)h;
  }
}

// Side classes, one word per edge slot (written over the slot's table index): the other slot of
// an edge used exactly twice, or one of these.
constexpr uint32_t kSideMany = 0xFFFFFFFDu;  // edge used by more than two slots
constexpr uint32_t kSideOne = 0xFFFFFFFEu;   // edge used by this slot alone
// (kNone: the slot's two ends are one vertex)
__global__ void k_edge_classify(uint32_t* __restrict__ eslot, int64_t nslots,
                                const uint32_t* __restrict__ ecnt,
                                const uint32_t* __restrict__ epair) {
  GRID_STRIDE(s, nslots) {
    const uint32_t h = eslot[s];
    if (h == kNone) continue;
    const uint32_t c = ecnt[h];
    uint32_t v;
    if (c == 1)
      v = kSideOne;
    else if (c > 2)
      v = kSideMany;
    else {
      const uint32_t a = epair[2 * h], b = epair[2 * h + 1];
      v = a == (uint32_t)s ? b : a;
    }
    eslot[s] = v;
  }
}

// ---------------------------------------------------------------------------------------------
// Components (dc_backend.cpp:313-370): faces joined by edges used exactly twice
// ---------------------------------------------------------------------------------------------
// Each face starts under its lowest neighbour across an edge used exactly twice (or itself): the
// pointers still decrease, so this is a forest of faces in one component each, and its trees are
// already shallow before the unions.
__global__ void k_uf_init(const uint32_t* __restrict__ side, uint32_t* __restrict__ parent,
                          int64_t n) {
  GRID_STRIDE(f, n) {
    const uint4 w = reinterpret_cast<const uint4*>(side)[f];
    uint32_t p = (uint32_t)f;
    if (w.x < kSideMany) p = min(p, w.x >> 2);
    if (w.y < kSideMany) p = min(p, w.y >> 2);
    if (w.z < kSideMany) p = min(p, w.z >> 2);
    if (w.w < kSideMany) p = min(p, w.w >> 2);
    parent[f] = p;
  }
}

// Each edge used by exactly two slots joins their faces (taken once, by the lower slot).
__global__ void k_uf_union(const uint32_t* __restrict__ side, int64_t nslots, uint32_t* parent) {
  GRID_STRIDE(s, nslots) {
    const uint32_t o = side[s];
    if (o >= kSideMany || o < (uint32_t)s) continue;
    uf_union(parent, (uint32_t)(s >> 2), o >> 2);
  }
}

// Shortens every path (halving). Its final pointers need not be roots: a thread halving through x
// may store an older ancestor after x's own store; the labels are taken read-only below.
__global__ void k_uf_compress(uint32_t* parent, int64_t n) {
  GRID_STRIDE(i, n) (void)uf_find(parent, (uint32_t)i);
}
// label[i] = root of i; nothing writes parent meanwhile, so every walk ends at the true root.
__global__ void k_uf_label(const uint32_t* __restrict__ parent, int64_t n,
                           uint32_t* __restrict__ label) {
  GRID_STRIDE(i, n) {
    uint32_t x = (uint32_t)i, p = parent[x];
    while (p != x) {
      x = p;
      p = parent[x];
    }
    label[i] = x;
  }
}

// Component sizes and count from the root labels. Nearly every face of a mesh can belong to one
// component, so adding 1 per face at its root would serialise on one address: each wave first
// merges its lanes by root (one ballot per distinct root), each block adds the merged counts into
// an LDS table of roots (a contiguous share of the faces holds few roots), and the table is
// flushed with one global atomic per entry. Roots beyond the table's capacity go straight to
// global memory.
constexpr int kRootSlots = 256;
__global__ void __launch_bounds__(kThreads)
    k_comp_size(const uint32_t* __restrict__ label, int64_t n, uint32_t* __restrict__ csize,
                uint32_t* __restrict__ ncomp) {
  __shared__ uint32_t s_key[kRootSlots], s_cnt[kRootSlots];
  for (int k = threadIdx.x; k < kRootSlots; k += kThreads) {
    s_key[k] = kNone;
    s_cnt[k] = 0;
  }
  __syncthreads();
  uint64_t lo, hi;
  dccommon::block_share((uint64_t)n, lo, hi);
  const int lane = threadIdx.x & 63;
  uint32_t roots = 0;
  for (uint64_t base = lo; base < hi; base += kThreads) {
    const uint64_t i = base + threadIdx.x;
    const bool valid = i < hi;
    const uint32_t r = valid ? label[i] : kNone;
    roots += (valid && r == (uint32_t)i) ? 1u : 0u;
    uint64_t pending = __ballot(valid);
    while (pending) {  // uniform across the wave
      const int leader = __ffsll((unsigned long long)pending) - 1;
      const uint32_t lr = (uint32_t)__shfl((int)r, leader);
      const uint64_t m = __ballot(valid && r == lr) & pending;
      if (lane == leader) {
        const uint32_t c = (uint32_t)__popcll(m);
        uint32_t h = (lr * 2654435761u) & (kRootSlots - 1);
        bool done = false;
        for (int probe = 0; probe < kRootSlots && !done; ++probe) {
          const uint32_t old = atomicCAS(s_key + h, kNone, lr);
          if (old == kNone || old == lr) {
            atomicAdd(s_cnt + h, c);
            done = true;
          }
          h = (h + 1) & (kRootSlots - 1);
        }
        if (!done) atomicAdd(csize + lr, c);
      }
      pending &= ~m;
    }
  }
  // roots counted per wave, one atomic per wave that saw any
  for (int off = 32; off > 0; off >>= 1) roots += (uint32_t)__shfl_down((int)roots, off);
  if (lane == 0 && roots) atomicAdd(ncomp, roots);
  __syncthreads();
  for (int k = threadIdx.x; k < kRootSlots; k += kThreads)
    if (s_key[k] != kNone) atomicAdd(csize + s_key[k], s_cnt[k]);
}

// The largest component, ties to the lowest first face (the reference keeps the first component
// of maximum size in first-face order, dc_backend.cpp:351-358).
__global__ void k_uf_best(const uint32_t* __restrict__ label, const uint32_t* __restrict__ csize,
                          int64_t n, unsigned long long* __restrict__ best) {
  GRID_STRIDE(i, n) if (label[i] == (uint32_t)i) {
    const unsigned long long key =
        ((unsigned long long)csize[i] << 32) | (unsigned long long)(0xFFFFFFFFu - (uint32_t)i);
    atomicMax(best, key);
  }
}
__global__ void k_uf_keep(const uint32_t* __restrict__ label, int64_t n,
                          const unsigned long long* __restrict__ best, uint32_t* __restrict__ keep) {
  const uint32_t root = 0xFFFFFFFFu - (uint32_t)(*best & 0xFFFFFFFFull);
  GRID_STRIDE(i, n) keep[i] = label[i] == root ? 1u : 0u;
}

// ---------------------------------------------------------------------------------------------
// Boundary edges (dc_backend.cpp:280-310, 393-398)
// ---------------------------------------------------------------------------------------------
struct BEdge {
  int32_t u, v;    // directed as in its face
  float pu[3], pv[3];
};

// An edge slot is a boundary edge when its edge has one slot, or (with floater removal) more than
// two; with floater removal only slots of kept faces count.
__global__ void k_bnd_flags(const uint32_t* __restrict__ side, int64_t nslots,
                            const uint32_t* __restrict__ keep, int floaters,
                            uint32_t* __restrict__ flag) {
  GRID_STRIDE(s, nslots) {
    const uint32_t o = side[s];
    uint32_t b = (o == kSideOne || (o == kSideMany && floaters)) ? 1u : 0u;
    if (b && floaters && !keep[s >> 2]) b = 0;
    flag[s] = b;
  }
}
__global__ void k_bnd_emit(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ idx,
                           int64_t nslots, const int4* __restrict__ f, const float* __restrict__ v,
                           BEdge* __restrict__ out) {
  GRID_STRIDE(s, nslots) if (flag[s]) {
    const int4 a = f[s >> 2];
    const int e = (int)(s & 3);
    BEdge b;
    b.u = e == 0 ? a.x : e == 1 ? a.y : e == 2 ? a.z : a.w;
    b.v = e == 0 ? a.y : e == 1 ? a.z : e == 2 ? a.w : a.x;
    for (int c = 0; c < 3; ++c) {
      b.pu[c] = v[3 * (int64_t)b.u + c];
      b.pv[c] = v[3 * (int64_t)b.v + c];
    }
    out[idx[s]] = b;
  }
}

// ---------------------------------------------------------------------------------------------
// Renumbering by first use (dc_backend.cpp:537-552)
// ---------------------------------------------------------------------------------------------
__global__ void k_first(const int32_t* __restrict__ af, int64_t npos, uint32_t* __restrict__ first) {
  GRID_STRIDE(p, npos) atomicMin(first + af[p], (uint32_t)p);
}
__global__ void k_first_flags(const int32_t* __restrict__ af, int64_t npos,
                              const uint32_t* __restrict__ first, uint32_t* __restrict__ flag) {
  GRID_STRIDE(p, npos) flag[p] = first[af[p]] == (uint32_t)p ? 1u : 0u;
}
// Vertex ids below nw are welded vertices (vw), the rest hole-fill vertices (vnew).
__global__ void k_renumber(const int32_t* __restrict__ af, int64_t npos,
                           const uint32_t* __restrict__ first, const uint32_t* __restrict__ nidx,
                           const float* __restrict__ vw, int64_t nw, const float* __restrict__ vnew,
                           float* __restrict__ vout, int32_t* __restrict__ fout) {
  GRID_STRIDE(p, npos) {
    const int32_t v = af[p];
    const uint32_t fp = first[v];
    fout[p] = (int32_t)nidx[fp];
    if (fp == (uint32_t)p) {
      const float* src = v < nw ? vw + 3 * (int64_t)v : vnew + 3 * ((int64_t)v - nw);
      float* dst = vout + 3 * (int64_t)nidx[p];
      dst[0] = src[0];
      dst[1] = src[1];
      dst[2] = src[2];
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Host: boundary-loop walk and fill (dc_backend.cpp:393-534)
// ---------------------------------------------------------------------------------------------
struct P3 {
  float x, y, z;
};

struct HoleFill {
  int64_t n_welded;            // ids below: welded vertices; from here: vertices made here
  std::vector<P3> made;        // coordinates of the vertices made here
  std::vector<int32_t> faces;  // 4 ids per face
  int64_t loops = 0;
  bool runaway = false;

  int32_t add(const P3& p) {
    made.push_back(p);
    return (int32_t)(n_welded + (int64_t)made.size() - 1);
  }
  void face(int32_t a, int32_t b, int32_t c, int32_t d) {
    faces.push_back(a);
    faces.push_back(b);
    faces.push_back(c);
    faces.push_back(d);
  }
  static float len3(float dx, float dy, float dz) { return std::sqrt(dx * dx + dy * dy + dz * dz); }
  static P3 centroid(const std::vector<P3>& pts) {
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    for (const P3& p : pts) {
      cx += p.x;
      cy += p.y;
      cz += p.z;
    }
    const float n = (float)pts.size();
    return P3{cx / n, cy / n, cz / n};
  }

  // Closes one loop (ids with their coordinates, in fill order): rings stepped towards the
  // centroid by the mean edge length while the loop is longer than 4 and some vertex is farther
  // than 1.25 mean edges from it, then one quad, one triangle or a centre fan.
  void fill(std::vector<int32_t> ids, std::vector<P3> pts) {
    const size_t n = ids.size();
    if (n < 3) return;
    ++loops;
    float step_len = 0.0f;
    for (size_t k = 0; k < n; ++k) {
      const P3& a = pts[k];
      const P3& b = pts[(k + 1) % n];
      step_len += len3(b.x - a.x, b.y - a.y, b.z - a.z);
    }
    step_len = step_len / (float)n;
    if (step_len <= 0.0001f) step_len = 1.0f;
    int64_t rings = 0;
    while (ids.size() > 4) {
      const P3 c = centroid(pts);
      float far = 0.0f;
      for (const P3& p : pts) far = std::max(far, len3(c.x - p.x, c.y - p.y, c.z - p.z));
      if (far <= 1.25f * step_len) break;
      if (++rings > kMaxRings) {
        runaway = true;
        return;
      }
      std::vector<int32_t> in_ids(n);
      std::vector<P3> in_pts(n);
      for (size_t k = 0; k < n; ++k) {
        const P3& p = pts[k];
        float dx = c.x - p.x, dy = c.y - p.y, dz = c.z - p.z;
        const float d = len3(dx, dy, dz);
        const float st = d > step_len ? step_len : d * 0.5f;
        if (d > 0.0001f) {
          dx /= d;
          dy /= d;
          dz /= d;
        } else {
          dx = dy = dz = 0.0f;
        }
        in_pts[k] = P3{p.x + dx * st, p.y + dy * st, p.z + dz * st};
        in_ids[k] = add(in_pts[k]);
      }
      for (size_t k = 0; k < n; ++k)
        face(ids[k], ids[(k + 1) % n], in_ids[(k + 1) % n], in_ids[k]);
      ids.swap(in_ids);
      pts.swap(in_pts);
    }
    const size_t m = ids.size();
    if (m == 3) {
      face(ids[0], ids[1], ids[2], ids[2]);
    } else if (m == 4) {
      face(ids[0], ids[1], ids[2], ids[3]);
    } else {
      const int32_t ctr = add(centroid(pts));
      for (size_t k = 0; k < m; ++k) face(ids[k], ids[(k + 1) % m], ctr, ctr);
    }
  }
};

// Walks the boundary edges (sorted by undirected key, the reference's order) into loops: from each
// source vertex in increasing order, repeatedly follow the last unused edge of the current vertex;
// returning to the start closes a loop, meeting a vertex already on the walk cuts off the loop
// since it; both are filled reversed when they have 3..4096 vertices.
void walk_and_fill(std::vector<BEdge>& be, HoleFill& hf) {
  std::sort(be.begin(), be.end(), [](const BEdge& a, const BEdge& b) {
    const int32_t alo = std::min(a.u, a.v), ahi = std::max(a.u, a.v);
    const int32_t blo = std::min(b.u, b.v), bhi = std::max(b.u, b.v);
    if (alo != blo) return alo < blo;
    if (ahi != bhi) return ahi < bhi;
    return a.u < b.u;  // equal keys: either order gives the same lists
  });
  // local ids of the vertices on boundary edges, ascending with the global ids
  std::vector<int32_t> gid;
  gid.reserve(2 * be.size());
  for (const BEdge& e : be) {
    gid.push_back(e.u);
    gid.push_back(e.v);
  }
  std::sort(gid.begin(), gid.end());
  gid.erase(std::unique(gid.begin(), gid.end()), gid.end());
  const size_t nl = gid.size();
  auto local = [&](int32_t g) {
    return (size_t)(std::lower_bound(gid.begin(), gid.end(), g) - gid.begin());
  };
  std::vector<P3> pos(nl);
  std::vector<size_t> start(nl + 1, 0), top(nl);
  std::vector<size_t> lu(be.size()), lv(be.size());
  for (size_t q = 0; q < be.size(); ++q) {
    lu[q] = local(be[q].u);
    lv[q] = local(be[q].v);
    pos[lu[q]] = P3{be[q].pu[0], be[q].pu[1], be[q].pu[2]};
    pos[lv[q]] = P3{be[q].pv[0], be[q].pv[1], be[q].pv[2]};
    ++start[lu[q] + 1];
  }
  for (size_t k = 0; k < nl; ++k) start[k + 1] += start[k];
  std::vector<size_t> adj(be.size());
  for (size_t k = 0; k < nl; ++k) top[k] = start[k];
  for (size_t q = 0; q < be.size(); ++q) adj[top[lu[q]]++] = lv[q];  // push order = sorted order
  // top[k] is now the end of k's list; taking an edge pops its last entry
  std::vector<int64_t> at(nl, -1);  // position on the current walk, or -1
  std::vector<size_t> walk;
  auto fill_reversed = [&](size_t from) {  // walk[from..] reversed
    std::vector<int32_t> ids;
    std::vector<P3> pts;
    for (size_t k = walk.size(); k > from; --k) {
      ids.push_back(gid[walk[k - 1]]);
      pts.push_back(pos[walk[k - 1]]);
    }
    hf.fill(std::move(ids), std::move(pts));
  };
  for (size_t s0 = 0; s0 < nl; ++s0) {
    while (top[s0] > start[s0]) {
      walk.assign(1, s0);
      at[s0] = 0;
      size_t cur = s0;
      bool closed = false;
      while (top[cur] > start[cur]) {
        const size_t nx = adj[--top[cur]];
        if (nx == s0) {
          closed = true;
          break;
        }
        if (at[nx] >= 0) {
          const size_t k = (size_t)at[nx];
          const size_t len = walk.size() - k;
          if (len >= 3 && len <= kMaxLoop) fill_reversed(k);
          for (size_t q = k; q < walk.size(); ++q) at[walk[q]] = -1;
          walk.resize(k);
        }
        at[nx] = (int64_t)walk.size();
        walk.push_back(nx);
        cur = nx;
      }
      if (closed && walk.size() >= 3 && walk.size() <= kMaxLoop) fill_reversed(0);
      for (size_t q : walk) at[q] = -1;
      if (hf.runaway) return;
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// DcPost
// ---------------------------------------------------------------------------------------------
DcPost::~DcPost() {
  for (Buf& b : b_)
    if (b.p) (void)hipFree(b.p);
}

int DcPost::hip_error(hipError_t e, const char* what) {
  if (err_) *err_ = std::string("postprocess: ") + what + ": " + hipGetErrorString(e);
  (void)hipGetLastError();
  return e == hipErrorOutOfMemory ? DC_HIP_ERR_OUT_OF_MEMORY : DC_HIP_ERR_HIP;
}

#define POST_TRY(expr, what)                       \
  do {                                             \
    const hipError_t e_ = (expr);                  \
    if (e_ != hipSuccess) return hip_error(e_, what); \
  } while (0)
#define POST_RC(expr)           \
  do {                          \
    const int rc_ = (expr);     \
    if (rc_ != DC_HIP_OK) return rc_; \
  } while (0)
#define LAUNCH(kernel, n, ...)                                                                  \
  do {                                                                                          \
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks_for((n), cu_)), dim3(kThreads), 0, s_,     \
                       __VA_ARGS__);                                                            \
    POST_TRY(hipGetLastError(), #kernel);                                                       \
  } while (0)

int DcPost::grow(int which, size_t bytes) {
  Buf& b = b_[which];
  if (bytes == 0) bytes = 256;
  if (b.cap >= bytes) return DC_HIP_OK;
  if (b.p) {
    POST_TRY(hipStreamSynchronize(s_), "synchronize");  // earlier work may still use it
    POST_TRY(hipFree(b.p), "free");
    b.p = nullptr;
    b.cap = 0;
  }
  const size_t want = bytes + bytes / 8;
  POST_TRY(hipMalloc(&b.p, want), "allocate");
  b.cap = want;
  return DC_HIP_OK;
}

template <typename T>
static T* at(void* p) {
  return static_cast<T*>(p);
}

int DcPost::scan(const uint32_t* flags, int64_t n, uint32_t* out, int slot) {
  POST_RC(grow(kTSum, (size_t)std::max<int64_t>(1, scan_tiles(n)) * 4));
  POST_TRY(exclusive_scan(flags, n, out, at<uint32_t>(b_[kTSum].p),
                          at<uint32_t>(b_[kCnt].p) + slot, s_),
           "scan");
  return DC_HIP_OK;
}

int DcPost::read_counts(uint32_t* h, int n) {
  POST_TRY(hipMemcpyAsync(h, b_[kCnt].p, (size_t)n * 4, hipMemcpyDeviceToHost, s_), "counts");
  POST_TRY(hipStreamSynchronize(s_), "counts");
  return DC_HIP_OK;
}

int DcPost::weld(const float* v, int64_t nv, const int32_t* f, int64_t nf, int ov, int of,
                 int64_t* nvo, int64_t* nfo) {
  const int nb = (int)blocks_for(nv, cu_);
  const size_t tsize = pow2_at_least((size_t)nv * 2);
  POST_RC(grow(kPart, (size_t)nb * 24 + 256));  // partial boxes, then the step
  POST_RC(grow(kTable, tsize * 4));
  POST_RC(grow(kSlot, (size_t)nv * 4));
  POST_RC(grow(kRemap, (size_t)nv * 4));
  POST_RC(grow(kFTmp, (size_t)nf * 16));
  POST_RC(grow(kFlag, (size_t)std::max(nv, nf) * 4));
  POST_RC(grow(kIdx, (size_t)std::max(nv, nf) * 4));
  POST_RC(grow(ov, (size_t)nv * 12));
  POST_RC(grow(of, (size_t)nf * 16));
  float* part = at<float>(b_[kPart].p);
  float* inv = part + 6 * nb;
  uint32_t* table = at<uint32_t>(b_[kTable].p);
  uint32_t* slot = at<uint32_t>(b_[kSlot].p);
  uint32_t* remap = at<uint32_t>(b_[kRemap].p);
  uint32_t* flag = at<uint32_t>(b_[kFlag].p);
  uint32_t* idx = at<uint32_t>(b_[kIdx].p);
  uint32_t* cnt = at<uint32_t>(b_[kCnt].p);
  hipLaunchKernelGGL(k_bbox, dim3((unsigned)nb), dim3(kThreads), 0, s_, v, nv, part);
  POST_TRY(hipGetLastError(), "k_bbox");
  hipLaunchKernelGGL(k_weld_step, dim3(1), dim3(kThreads), 0, s_, part, nb, inv);
  POST_TRY(hipGetLastError(), "k_weld_step");
  LAUNCH(k_fill32, (int64_t)tsize, table, (int64_t)tsize, kNone);
  LAUNCH(k_weld_insert, nv, v, nv, inv, table, (uint64_t)(tsize - 1), slot);
  LAUNCH(k_weld_flags, nv, table, slot, nv, flag);
  POST_RC(scan(flag, nv, idx, kCntWeldV));
  LAUNCH(k_weld_out, nv, v, nv, table, slot, idx, remap, at<float>(b_[ov].p));
  LAUNCH(k_weld_faces, nf, reinterpret_cast<const int4*>(f), nf, nv, remap,
         at<int4>(b_[kFTmp].p), flag, cnt + kCntErr);
  POST_RC(scan(flag, nf, idx, kCntWeldF));
  LAUNCH(k_compact4, nf, at<const int4>(b_[kFTmp].p), nf, flag, idx, at<int4>(b_[of].p));
  uint32_t h[kCntSlots];
  POST_RC(read_counts(h, kCntSlots));
  if (h[kCntErr]) {
    if (err_) *err_ = "postprocess: a face index is outside [0, n_vertices)";
    return DC_HIP_ERR_INVALID_ARG;
  }
  *nvo = h[kCntWeldV];
  *nfo = h[kCntWeldF];
  return DC_HIP_OK;
}

int DcPost::run(const float* vertices, int64_t nv, const int32_t* faces, int64_t nf,
                bool close_holes, bool remove_floaters, int cu_count, hipStream_t s,
                dc_hip_post_stats* st, std::string* err) {
  const auto t0 = std::chrono::steady_clock::now();
  s_ = s;
  cu_ = cu_count > 0 ? cu_count : 256;
  err_ = err;
  have_ = false;
  dc_hip_post_stats S{};
  if (nv < 0 || nf < 0 || nv > kMaxVertices || nf > kMaxFaces) {
    if (err_) *err_ = "postprocess: vertex or face count out of range (at most 2^30 and 2^28)";
    return DC_HIP_ERR_INVALID_ARG;
  }
  if ((nv > 0 && !vertices) || (nf > 0 && !faces)) {
    if (err_) *err_ = "postprocess: NULL vertices or faces";
    return DC_HIP_ERR_INVALID_ARG;
  }
  POST_RC(grow(kCnt, kCntSlots * 4));
  POST_RC(grow(kInV, (size_t)nv * 12));
  POST_RC(grow(kInF, (size_t)nf * 16));
  // hipMemcpyDefault: the input may be host or device memory (unified addressing)
  if (nv) POST_TRY(hipMemcpyAsync(b_[kInV].p, vertices, (size_t)nv * 12, hipMemcpyDefault, s_), "upload");
  if (nf) POST_TRY(hipMemcpyAsync(b_[kInF].p, faces, (size_t)nf * 16, hipMemcpyDefault, s_), "upload");
  LAUNCH(k_fill32, (int64_t)kCntSlots, at<uint32_t>(b_[kCnt].p), (int64_t)kCntSlots, 0u);
  auto finish = [&](int vbuf, int fbuf, int64_t v_n, int64_t f_n) {
    out_v_ = at<float>(b_[vbuf].p);
    out_f_ = at<int32_t>(b_[fbuf].p);
    nv_out_ = v_n;
    nf_out_ = f_n;
    have_ = true;
    S.vertex_count = v_n;
    S.face_count = f_n;
    S.total_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (st) *st = S;
    return DC_HIP_OK;
  };
  // nothing to do: the input as it is (dc_backend.cpp:218-220)
  if (nv == 0 || nf == 0 || (!close_holes && !remove_floaters)) {
    POST_TRY(hipStreamSynchronize(s_), "upload");
    return finish(kInV, kInF, nv, nf);
  }
  const float* v0 = at<float>(b_[kInV].p);
  const int32_t* f0 = at<int32_t>(b_[kInF].p);
  int64_t nw = 0, nfw = 0;
  POST_RC(weld(v0, nv, f0, nf, kWV, kWF, &nw, &nfw));
  S.welded_vertices = nw;
  S.welded_faces = nfw;
  if (nw == 0 || nfw == 0) return finish(kInV, kInF, nv, nf);  // (:226) the unwelded input
  const float* vw = at<float>(b_[kWV].p);
  const int4* fw = at<int4>(b_[kWF].p);
  uint32_t* cnt = at<uint32_t>(b_[kCnt].p);

  // undirected edges
  const int64_t nslots = 4 * nfw;
  const size_t esize = pow2_at_least((size_t)nslots * 2);
  POST_RC(grow(kETab, esize * 8));
  POST_RC(grow(kECnt, esize * 4));
  POST_RC(grow(kEPair, esize * 8));
  POST_RC(grow(kESlot, (size_t)nslots * 4));
  unsigned long long* etab = at<unsigned long long>(b_[kETab].p);
  uint32_t* ecnt = at<uint32_t>(b_[kECnt].p);
  uint32_t* epair = at<uint32_t>(b_[kEPair].p);
  uint32_t* eslot = at<uint32_t>(b_[kESlot].p);
  LAUNCH(k_fill64, (int64_t)esize, etab, (int64_t)esize, kEmptyKey);
  LAUNCH(k_fill32, (int64_t)esize, ecnt, (int64_t)esize, 0u);
  LAUNCH(k_edge_insert, nslots, fw, nslots, etab, ecnt, epair, (uint64_t)(esize - 1), eslot);
  LAUNCH(k_edge_classify, nslots, eslot, nslots, ecnt, epair);
  const uint32_t* side = eslot;

  // floaters: keep flags of the largest component
  POST_RC(grow(kFlag, (size_t)std::max<int64_t>(nslots, nfw) * 4));
  POST_RC(grow(kIdx, (size_t)std::max<int64_t>(nslots, nfw) * 4));
  uint32_t* flag = at<uint32_t>(b_[kFlag].p);
  uint32_t* idx = at<uint32_t>(b_[kIdx].p);
  int64_t nk = nfw;
  uint32_t* keep = nullptr;
  if (remove_floaters) {
    POST_RC(grow(kParent, (size_t)nfw * 4));
    POST_RC(grow(kCSize, (size_t)nfw * 4));
    POST_RC(grow(kBest, 256));
    POST_RC(grow(kKeep, (size_t)nfw * 4));
    uint32_t* parent = at<uint32_t>(b_[kParent].p);
    uint32_t* csize = at<uint32_t>(b_[kCSize].p);
    unsigned long long* best = at<unsigned long long>(b_[kBest].p);
    keep = at<uint32_t>(b_[kKeep].p);
    LAUNCH(k_uf_init, nfw, side, parent, nfw);
    LAUNCH(k_fill32, nfw, csize, nfw, 0u);
    LAUNCH(k_fill64, (int64_t)1, best, (int64_t)1, 0ull);
    LAUNCH(k_uf_compress, nfw, parent, nfw);  // short paths before the unions walk them
    LAUNCH(k_uf_union, nslots, side, nslots, parent);
    LAUNCH(k_uf_compress, nfw, parent, nfw);
    uint32_t* label = flag;  // free until the boundary flags
    LAUNCH(k_uf_label, nfw, parent, nfw, label);
    // a few hundred blocks: one flush per block and root
    hipLaunchKernelGGL(k_comp_size, dim3((unsigned)std::min<int64_t>(blocks_for(nfw, cu_), 2 * cu_)),
                       dim3(kThreads), 0, s_, label, nfw, csize, cnt + kCntComp);
    POST_TRY(hipGetLastError(), "k_comp_size");
    LAUNCH(k_uf_best, nfw, label, csize, nfw, best);
    LAUNCH(k_uf_keep, nfw, label, nfw, best, keep);
  }

  // boundary edges, then the loops on the host
  HoleFill hf;
  hf.n_welded = nw;
  if (close_holes) {
    LAUNCH(k_bnd_flags, nslots, side, nslots, keep, remove_floaters ? 1 : 0, flag);
    POST_RC(scan(flag, nslots, idx, kCntBnd));
    uint32_t h[kCntSlots];
    POST_RC(read_counts(h, kCntSlots));
    const int64_t nb = h[kCntBnd];
    POST_RC(grow(kBnd, (size_t)nb * sizeof(BEdge)));
    LAUNCH(k_bnd_emit, nslots, flag, idx, nslots, fw, vw, at<BEdge>(b_[kBnd].p));
    std::vector<BEdge> be((size_t)nb);
    if (nb) POST_TRY(hipMemcpyAsync(be.data(), b_[kBnd].p, (size_t)nb * sizeof(BEdge),
                                    hipMemcpyDeviceToHost, s_), "boundary edges");
    POST_TRY(hipStreamSynchronize(s_), "boundary edges");
    S.boundary_edges = nb;
    const auto w0 = std::chrono::steady_clock::now();
    walk_and_fill(be, hf);
    S.walk_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - w0).count();
    if (hf.runaway) {
      if (err_) *err_ = "postprocess: a hole fill did not converge";
      return DC_HIP_ERR_INTERNAL;
    }
    if (nw + (int64_t)hf.made.size() > kMaxVertices ||
        nfw + (int64_t)(hf.faces.size() / 4) > kMaxFaces) {
      if (err_) *err_ = "postprocess: the hole fill exceeds the mesh size limits";
      return DC_HIP_ERR_INVALID_ARG;
    }
  }
  if (remove_floaters) {
    uint32_t h[kCntSlots];
    POST_RC(read_counts(h, kCntSlots));
    S.components = h[kCntComp];
  }
  S.loops_filled = hf.loops;
  S.added_vertices = (int64_t)hf.made.size();
  S.added_faces = (int64_t)hf.faces.size() / 4;

  // kept faces in order, then the hole faces
  if (remove_floaters) {
    POST_RC(scan(keep, nfw, idx, kCntKeep));
    uint32_t h[kCntSlots];
    POST_RC(read_counts(h, kCntSlots));
    nk = h[kCntKeep];
  }
  S.kept_faces = nk;
  const int64_t nh = S.added_faces, na = nk + nh, npos = 4 * na;
  const int64_t nall = nw + S.added_vertices;
  POST_RC(grow(kAF, (size_t)na * 16));
  POST_RC(grow(kNewV, (size_t)S.added_vertices * 12));
  POST_RC(grow(kFirst, (size_t)nall * 4));
  int4* af = at<int4>(b_[kAF].p);
  if (remove_floaters)
    LAUNCH(k_compact4, nfw, fw, nfw, keep, idx, af);
  else
    POST_TRY(hipMemcpyAsync(af, fw, (size_t)nfw * 16, hipMemcpyDeviceToDevice, s_), "faces");
  if (nh)
    POST_TRY(hipMemcpyAsync(af + nk, hf.faces.data(), (size_t)nh * 16, hipMemcpyHostToDevice, s_),
             "hole faces");
  if (S.added_vertices)
    POST_TRY(hipMemcpyAsync(b_[kNewV].p, hf.made.data(), (size_t)S.added_vertices * 12,
                            hipMemcpyHostToDevice, s_),
             "hole vertices");

  // renumber by first use
  POST_RC(grow(kFlag, (size_t)npos * 4));
  POST_RC(grow(kIdx, (size_t)npos * 4));
  POST_RC(grow(kRV, (size_t)nall * 12));
  POST_RC(grow(kRF, (size_t)npos * 4));
  flag = at<uint32_t>(b_[kFlag].p);
  idx = at<uint32_t>(b_[kIdx].p);
  uint32_t* first = at<uint32_t>(b_[kFirst].p);
  const int32_t* afi = at<int32_t>(b_[kAF].p);
  LAUNCH(k_fill32, nall, first, nall, kNone);
  LAUNCH(k_first, npos, afi, npos, first);
  LAUNCH(k_first_flags, npos, afi, npos, first, flag);
  POST_RC(scan(flag, npos, idx, kCntFirst));
  LAUNCH(k_renumber, npos, afi, npos, first, idx, vw, nw, at<float>(b_[kNewV].p),
         at<float>(b_[kRV].p), at<int32_t>(b_[kRF].p));
  uint32_t h[kCntSlots];
  POST_RC(read_counts(h, kCntSlots));
  const int64_t nvr = h[kCntFirst];
  if (!close_holes) return finish(kRV, kRF, nvr, na);
  // with hole closing the result is welded again (dc_backend.cpp:554-556)
  int64_t nv2 = 0, nf2 = 0;
  POST_RC(weld(at<float>(b_[kRV].p), nvr, at<int32_t>(b_[kRF].p), na, kW2V, kW2F, &nv2, &nf2));
  return finish(kW2V, kW2F, nv2, nf2);
}
