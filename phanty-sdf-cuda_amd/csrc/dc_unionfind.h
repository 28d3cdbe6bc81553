// dc_unionfind.h — lock-free union-find over uint32 ids for the mesh post-processing (faces) and
// the mesh-to-SDF flood (voxels). The larger root always hooks under the smaller one, so every
// root is the lowest id of its set and parent[x] < x for every other x.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dcuf {

// Parent pointers are read and halved with relaxed, wave-scoped (cached) accesses; a stale value
// is an earlier ancestor, which only costs a retry of the hooking CAS (performed at the L2).
__device__ __forceinline__ uint32_t ld_rlx(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ __forceinline__ void st_rlx(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
// parent[x] < x for every non-root x (roots hook under smaller roots, halving moves a pointer to
// an ancestor), so the walk ends and each root is the lowest id of its tree.
__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
  while (true) {
    const uint32_t p = ld_rlx(parent + x);
    if (p == x) return x;
    const uint32_t g = ld_rlx(parent + p);
    if (g == p) return p;
    st_rlx(parent + x, g);  // path halving
    x = g;
  }
}
__device__ __forceinline__ void uf_union(uint32_t* parent, uint32_t a, uint32_t b) {
  while (true) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    if (a > b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const uint32_t old = atomicCAS(parent + b, b, a);  // hook the larger root under the smaller
    if (old == b) return;
    b = old;  // b was hooked meanwhile: continue from its new parent
  }
}

}  // namespace dcuf
