// post_oracle.cpp — CPU restatement of the reference's mesh post-processing, postprocess_mesh
// (src/cpp/dc_backend.cpp:127-558), used ONLY as the parity checker of the GPU post-processing.
//
// TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library. The product (phanty-sdf-cuda_amd/) never links, loads or calls it.
//
// Sequential, in the reference's order, with its FP32 arithmetic (built with -ffp-contract=off):
//   weld        bounding-box diagonal, eps = 1e-5 x diag (1e-5 if 0), keys round(x / eps) per
//               coordinate, first vertex of each key kept, faces with < 3 distinct ids dropped
//                                                                    dc_backend.cpp:127-215
//   edges       (min, max) per face side with distinct ends, sorted; runs of equal keys:
//               1 -> boundary edge, 2 -> neighbours, > 2 -> boundary (hole closing with floater
//               removal) or all-pairs neighbours (without floater removal)  :235-311
//   floaters    breadth-first components in face order; the first of largest size kept  :313-370
//   holes       per source vertex in order, follow the last unused boundary edge; back at the
//               start: a loop; meeting the walk: the loop since then; loops of 3..4096 filled
//               reversed by centroid rings, then a quad, a triangle or a centre fan       :393-534
//   output      vertices renumbered by first use, welded again after hole closing     :537-557
// The std::sort of the edges (:255) is not stable; entries with equal keys only ever push equal
// values onto one vertex's boundary list, so their order does not change the result.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

namespace {

struct Mesh {
  std::vector<float> v;
  std::vector<int> f;
};

void weld(Mesh& m) {
  const int nv = (int)(m.v.size() / 3), nf = (int)(m.f.size() / 4);
  if (nv <= 0 || nf <= 0) return;
  float lo[3] = {m.v[0], m.v[1], m.v[2]}, hi[3] = {m.v[0], m.v[1], m.v[2]};
  for (int i = 0; i < nv; ++i)
    for (int c = 0; c < 3; ++c) {
      const float x = m.v[3 * i + c];
      if (x < lo[c]) lo[c] = x;
      if (x > hi[c]) hi[c] = x;
    }
  const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
  const float diag = std::sqrt(ex * ex + ey * ey + ez * ez);
  const float eps = diag > 0.0f ? diag * 1e-5f : 1e-5f;
  const float inv = 1.0f / eps;
  std::map<std::tuple<int64_t, int64_t, int64_t>, int> first;
  std::vector<int> to(nv);
  std::vector<float> out;
  for (int i = 0; i < nv; ++i) {
    const float x = m.v[3 * i], y = m.v[3 * i + 1], z = m.v[3 * i + 2];
    const auto key = std::make_tuple((int64_t)std::round(x * inv), (int64_t)std::round(y * inv),
                                     (int64_t)std::round(z * inv));
    auto it = first.find(key);
    if (it == first.end()) {
      const int id = (int)(out.size() / 3);
      first.emplace(key, id);
      to[i] = id;
      out.push_back(x);
      out.push_back(y);
      out.push_back(z);
    } else {
      to[i] = it->second;
    }
  }
  std::vector<int> faces;
  for (int q = 0; q < nf; ++q) {
    int id[4];
    for (int k = 0; k < 4; ++k) id[k] = to[m.f[4 * q + k]];
    int distinct = 0;
    for (int k = 0; k < 4; ++k) {
      bool seen = false;
      for (int j = 0; j < k; ++j) seen = seen || id[j] == id[k];
      distinct += seen ? 0 : 1;
    }
    if (distinct < 3) continue;
    faces.insert(faces.end(), id, id + 4);
  }
  m.v.swap(out);
  m.f.swap(faces);
}

struct Fill {
  std::vector<float>& v;
  std::vector<int>& faces;
  long loops = 0;
  bool runaway = false;

  float px(int i, int c) const { return v[3 * i + c]; }
  int add(float x, float y, float z) {
    v.push_back(x);
    v.push_back(y);
    v.push_back(z);
    return (int)(v.size() / 3) - 1;
  }
  void quad(int a, int b, int c, int d) {
    faces.push_back(a);
    faces.push_back(b);
    faces.push_back(c);
    faces.push_back(d);
  }
  void run(std::vector<int> loop) {
    if (loop.size() < 3) return;
    ++loops;
    float mean = 0.0f;
    for (size_t k = 0; k < loop.size(); ++k) {
      const int a = loop[k], b = loop[(k + 1) % loop.size()];
      const float dx = px(b, 0) - px(a, 0), dy = px(b, 1) - px(a, 1), dz = px(b, 2) - px(a, 2);
      mean += std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    mean /= loop.size();
    if (mean <= 0.0001f) mean = 1.0f;
    long rings = 0;
    while (loop.size() > 4) {
      const size_t n = loop.size();
      float cx = 0.0f, cy = 0.0f, cz = 0.0f;
      for (int id : loop) {
        cx += px(id, 0);
        cy += px(id, 1);
        cz += px(id, 2);
      }
      cx /= n;
      cy /= n;
      cz /= n;
      float far = 0.0f;
      for (int id : loop) {
        const float dx = cx - px(id, 0), dy = cy - px(id, 1), dz = cz - px(id, 2);
        const float d = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (d > far) far = d;
      }
      if (far <= 1.25f * mean) break;
      if (++rings > (1L << 20)) {
        runaway = true;
        return;
      }
      std::vector<int> inner;
      for (int id : loop) {
        const float x = px(id, 0), y = px(id, 1), z = px(id, 2);
        float dx = cx - x, dy = cy - y, dz = cz - z;
        const float d = std::sqrt(dx * dx + dy * dy + dz * dz);
        const float step = d > mean ? mean : d * 0.5f;
        if (d > 0.0001f) {
          dx /= d;
          dy /= d;
          dz /= d;
        } else {
          dx = dy = dz = 0.0f;
        }
        inner.push_back(add(x + dx * step, y + dy * step, z + dz * step));
      }
      for (size_t k = 0; k < n; ++k)
        quad(loop[k], loop[(k + 1) % n], inner[(k + 1) % n], inner[k]);
      loop = inner;
    }
    if (loop.size() == 3) {
      quad(loop[0], loop[1], loop[2], loop[2]);
    } else if (loop.size() == 4) {
      quad(loop[0], loop[1], loop[2], loop[3]);
    } else {
      const size_t n = loop.size();
      float cx = 0.0f, cy = 0.0f, cz = 0.0f;
      for (int id : loop) {
        cx += px(id, 0);
        cy += px(id, 1);
        cz += px(id, 2);
      }
      cx /= n;
      cy /= n;
      cz /= n;
      const int c = add(cx, cy, cz);
      for (size_t k = 0; k < n; ++k) quad(loop[k], loop[(k + 1) % n], c, c);
    }
  }
};

struct Side {
  int lo, hi, face, slot;
};

// info: welded vertices, welded faces, components, kept faces, boundary edges, loops filled,
// vertices added, faces added
int postprocess(const Mesh& in, bool close_holes, bool floaters, Mesh& out, int64_t info[8]) {
  std::memset(info, 0, 8 * sizeof(int64_t));
  if (in.v.empty() || in.f.empty() || (!close_holes && !floaters)) {
    out = in;
    return 0;
  }
  Mesh m = in;
  weld(m);
  const int nv = (int)(m.v.size() / 3), nf = (int)(m.f.size() / 4);
  info[0] = nv;
  info[1] = nf;
  if (nv == 0 || nf == 0) {
    out = in;
    return 0;
  }
  std::vector<Side> sides;
  for (int q = 0; q < nf; ++q)
    for (int e = 0; e < 4; ++e) {
      const int a = m.f[4 * q + e], b = m.f[4 * q + (e + 1) % 4];
      if (a == b) continue;
      sides.push_back(Side{std::min(a, b), std::max(a, b), q, e});
    }
  std::sort(sides.begin(), sides.end(), [](const Side& x, const Side& y) {
    return x.lo != y.lo ? x.lo < y.lo : x.hi < y.hi;
  });
  std::vector<std::vector<int>> nbr(nf);
  struct Directed {
    int from, to, face;
  };
  std::vector<Directed> boundary;
  auto directed = [&](const Side& s) {
    return Directed{m.f[4 * s.face + s.slot], m.f[4 * s.face + (s.slot + 1) % 4], s.face};
  };
  for (size_t i = 0; i < sides.size();) {
    size_t j = i + 1;
    while (j < sides.size() && sides[j].lo == sides[i].lo && sides[j].hi == sides[i].hi) ++j;
    const size_t run = j - i;
    if (run == 1) {
      boundary.push_back(directed(sides[i]));
    } else if (run == 2) {
      nbr[sides[i].face].push_back(sides[i + 1].face);
      nbr[sides[i + 1].face].push_back(sides[i].face);
    } else {
      if (!floaters)
        for (size_t x = i; x < j; ++x)
          for (size_t y = x + 1; y < j; ++y) {
            nbr[sides[x].face].push_back(sides[y].face);
            nbr[sides[y].face].push_back(sides[x].face);
          }
      if (close_holes && floaters)
        for (size_t x = i; x < j; ++x) boundary.push_back(directed(sides[x]));
    }
    i = j;
  }
  std::vector<int> comp(nf, 0);
  int keep_comp = 0;
  std::vector<int> kept;
  if (floaters) {
    std::fill(comp.begin(), comp.end(), -1);
    std::vector<int> size;
    for (int q = 0; q < nf; ++q) {
      if (comp[q] != -1) continue;
      const int id = (int)size.size();
      std::vector<int> queue{q};
      comp[q] = id;
      for (size_t h = 0; h < queue.size(); ++h)
        for (int g : nbr[queue[h]])
          if (comp[g] == -1) {
            comp[g] = id;
            queue.push_back(g);
          }
      size.push_back((int)queue.size());
    }
    info[2] = (int64_t)size.size();
    int best = -1;
    for (size_t c = 0; c < size.size(); ++c)
      if (best < 0 || size[c] > size[best]) best = (int)c;
    keep_comp = best;
    for (int q = 0; q < nf; ++q)
      if (comp[q] == keep_comp) kept.push_back(q);
  } else {
    for (int q = 0; q < nf; ++q) kept.push_back(q);
  }
  info[3] = (int64_t)kept.size();
  std::vector<int> faces;
  for (int q : kept) faces.insert(faces.end(), &m.f[4 * q], &m.f[4 * q] + 4);
  if (close_holes && !boundary.empty()) {
    std::vector<std::vector<int>> from(nv);
    for (const Directed& d : boundary)
      if (!floaters || comp[d.face] == keep_comp) {
        from[d.from].push_back(d.to);
        ++info[4];
      }
    Fill fill{m.v, faces};
    for (int s0 = 0; s0 < nv; ++s0) {
      while (!from[s0].empty()) {
        std::vector<int> walk{s0};
        int cur = s0;
        bool closed = false;
        while (!from[cur].empty()) {
          const int nx = from[cur].back();
          from[cur].pop_back();
          if (nx == s0) {
            closed = true;
            break;
          }
          auto it = std::find(walk.begin(), walk.end(), nx);
          if (it != walk.end()) {
            std::vector<int> sub(it, walk.end());
            walk.erase(it, walk.end());
            if (sub.size() >= 3 && sub.size() <= 4096)
              fill.run(std::vector<int>(sub.rbegin(), sub.rend()));
          }
          walk.push_back(nx);
          cur = nx;
        }
        if (closed && walk.size() >= 3 && walk.size() <= 4096)
          fill.run(std::vector<int>(walk.rbegin(), walk.rend()));
        if (fill.runaway) return 2;
      }
    }
    info[5] = fill.loops;
    info[6] = (int64_t)(m.v.size() / 3) - nv;
    info[7] = (int64_t)(faces.size() / 4) - (int64_t)kept.size();
  }
  std::vector<int> renum(m.v.size() / 3, -1);
  out.v.clear();
  out.f.clear();
  for (int id : faces) {
    if (renum[id] == -1) {
      renum[id] = (int)(out.v.size() / 3);
      out.v.insert(out.v.end(), &m.v[3 * id], &m.v[3 * id] + 3);
    }
    out.f.push_back(renum[id]);
  }
  if (close_holes) weld(out);
  return 0;
}

struct Handle {
  Mesh out;
  int64_t info[8];
};

}  // namespace

extern "C" {

// Post-processes (vertices: 3 * nv floats, faces: 4 * nf int32). Returns a handle (NULL on a
// hole fill that does not converge); read it with oracle_post_counts / oracle_post_copy.
void* oracle_postprocess(const float* v, int64_t nv, const int32_t* f, int64_t nf, int close_holes,
                         int remove_floaters) {
  Mesh in;
  in.v.assign(v, v + 3 * nv);
  in.f.assign(f, f + 4 * nf);
  Handle* h = new Handle();
  if (postprocess(in, close_holes != 0, remove_floaters != 0, h->out, h->info) != 0) {
    delete h;
    return nullptr;
  }
  return h;
}
void oracle_post_counts(void* handle, int64_t* nv, int64_t* nf, int64_t* info) {
  const Handle* h = static_cast<Handle*>(handle);
  *nv = (int64_t)(h->out.v.size() / 3);
  *nf = (int64_t)(h->out.f.size() / 4);
  std::memcpy(info, h->info, sizeof(h->info));
}
void oracle_post_copy(void* handle, float* v, int32_t* f) {
  const Handle* h = static_cast<Handle*>(handle);
  std::memcpy(v, h->out.v.data(), h->out.v.size() * sizeof(float));
  std::memcpy(f, h->out.f.data(), h->out.f.size() * sizeof(int32_t));
}
void oracle_post_free(void* handle) { delete static_cast<Handle*>(handle); }

// The weld alone (merge_vertices_in_place, dc_backend.cpp:127-215), as a handle like the above.
void* oracle_weld(const float* v, int64_t nv, const int32_t* f, int64_t nf) {
  Handle* h = new Handle();
  h->out.v.assign(v, v + 3 * nv);
  h->out.f.assign(f, f + 4 * nf);
  weld(h->out);
  std::memset(h->info, 0, sizeof(h->info));
  return h;
}

}  // extern "C"
