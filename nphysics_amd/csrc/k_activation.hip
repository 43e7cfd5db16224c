// nphysics_amd/csrc/k_activation.hip -- sleeping and activation islands.
//
// Restates ActivationManager::update (src/detection/activation_manager.rs:61-201) for rigid
// bodies without joints or kinematic bodies, then the manifold collection of
// MechanicalWorld::step (src/world/mechanical_world.rs:287-300) that hands the solver only the
// manifolds with an awake dynamic side:
//   k_act_wake     (before the step's dynamics update, after the caller set the bodies' state)
//                  sleeping bodies woken with 2 x threshold, as RigidBody::update_dynamics does
//                  for a body whose pose or velocity changed (src/object/rigid_body.rs:558-565);
//   k_act_energy   the energy running average e <- (1 - 0.01) e + 0.01 |v|^2 capped at
//                  4 x threshold (activation_manager.rs:47-57) of every awake body that has a
//                  threshold; union-find parents reset;
//   k_act_union    islands: union over the manifolds whose bodies are both dynamic (:143-147),
//                  lock-free (a root is linked under the smaller root by compare-and-swap),
//                  restricted to the manifolds between bodies that could sleep (see there);
//   k_act_flags    every body stores its root (read-only walks: no other body overwrites it), and
//                  an island cannot sleep as soon as one of its bodies has no threshold or an
//                  energy >= its threshold (:160-172);
//   k_act_apply    islands that can sleep are put to sleep (energy and velocity zeroed,
//                  rigid_body.rs:396-400), the others woken with 2 x threshold (body.rs:338-342);
//                  counts the awake dynamic bodies;
//   k_act_manifolds, k_act_compact  the solved-manifold list (stable compaction of the manifolds
//                  with an awake dynamic side) and, for the other manifolds, the impulse cache
//                  carried over from the previous step: the reference keeps a contact's cached
//                  impulse as long as the contact exists (signorini_coulomb_pyramid_model.rs:226-261
//                  only overwrites the entries of solved contacts).
// Activation state: one float2 per body (energy, threshold; threshold < 0 means the body never
// sleeps, energy 0 means asleep). Every kernel is one pass over bodies or manifolds; union-find
// walks are bounded by the body count (parents only point to smaller indices) and halve the
// paths they walk, which keeps the trees shallow when one island spans the whole world.
#include <hip/hip_runtime.h>

#include "devworld.hpp"
#include "kernels.hpp"
#include "stages.hpp"

namespace nph {

namespace {

constexpr float kMix = 0.01f;  // ActivationManager::new(0.01) (mechanical_world.rs:80)

__device__ inline int uf_load(const int* p, int x) {
  return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving: every visited node is pointed at its grandparent, which is also
// one of its ancestors, so concurrent walks and links stay valid.
__device__ inline int uf_root(int* p, int x, int n) {
  for (int it = 0; it < n; ++it) {
    const int y = uf_load(p, x);
    if (y == x) break;
    const int z = uf_load(p, y);
    if (z != y) __hip_atomic_store(p + x, z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = z;
  }
  return x;
}

// Sum of v over the 256-thread block, added to *out by one atomic (one address shared by every
// block: per-wave atomics serialise there).
__device__ inline void block_add(uint32_t v, uint32_t* out) {
  __shared__ uint32_t part[4];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t s = (part[0] + part[1]) + (part[2] + part[3]);
    if (s) atomicAdd(out, s);
  }
}

// Root of x without writes: used once the trees are final, where every body stores its own root
// and no walk may overwrite that with an older ancestor it read before.
__device__ inline int uf_root_ro(const int* p, int x, int n) {
  for (int it = 0; it < n; ++it) {
    const int y = uf_load(p, x);
    if (y == x) break;
    x = y;
  }
  return x;
}

__global__ void __launch_bounds__(256) k_act_wake(DevBodies b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n || b.status[i] != STATUS_DYNAMIC) return;
  const float2 a = b.act[i];
  if (a.x == 0.0f && a.y >= 0.0f) b.act[i] = make_float2(a.y * 2.0f, a.y);
}

__global__ void __launch_bounds__(256) k_act_energy(DevBodies b, int* parent, uint32_t* can, uint32_t* touch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  parent[i] = i;
  can[i] = 1u;
  touch[i] = 0u;
  if (b.status[i] != STATUS_DYNAMIC) return;
  float2 a = b.act[i];
  if (a.x != 0.0f && a.y >= 0.0f) {
    const float4 v = b.linvel[i], w = b.angvel[i];
    const float nsq = ((((v.x * v.x + v.y * v.y) + v.z * v.z) + w.x * w.x) + w.y * w.y) + w.z * w.z;
    const float e = (1.0f - kMix) * a.x + kMix * nsq;
    const float cap = a.y * 4.0f;
    a.x = e < cap ? e : cap;
  }
  b.act[i] = a;
}

// A body that keeps its island awake: no threshold, or an energy at or above it.
__device__ inline bool keeps_awake(float2 a) { return !(a.y >= 0.0f && a.x < a.y); }

// An island can sleep iff none of its bodies keeps it awake. Take the components of the bodies
// that could sleep under the manifolds between two such bodies: if a member of one touches a body
// that keeps its island awake, its island contains that body; if none does, every neighbour of
// the component is in it and the component is the whole island. So only those manifolds are
// united, and a manifold between a body that could sleep and one that keeps its island awake just
// marks the former (`touch`): the answer for every body is the one of the full islands, with far
// fewer unions (and smaller trees) while much of a world is moving.
__global__ void __launch_bounds__(256) k_act_union(DevBodies b, DevContacts c, int M, int* parent, uint32_t* touch) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int2 bb = c.mbody[m];
  if (b.status[bb.x] != STATUS_DYNAMIC || b.status[bb.y] != STATUS_DYNAMIC) return;
  const bool h1 = keeps_awake(b.act[bb.x]), h2 = keeps_awake(b.act[bb.y]);
  if (h1 || h2) {
    if (!h1) touch[bb.x] = 1u;
    if (!h2) touch[bb.y] = 1u;
    return;
  }
  const int n = b.n;
  int a = uf_root(parent, bb.x, n), d = uf_root(parent, bb.y, n);
  // every failed exchange means another thread linked root a, which happens at most n - 1 times
  for (int it = 0; it < 2 * n && a != d; ++it) {
    if (a < d) {
      const int t = a;
      a = d;
      d = t;
    }
    const int old = atomicCAS(parent + a, a, d);
    if (old == a) break;
    a = uf_root(parent, old, n);
    d = uf_root(parent, d, n);
  }
}

__global__ void __launch_bounds__(256) k_act_flags(DevBodies b, int* parent, uint32_t* can,
                                                  const uint32_t* __restrict__ touch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  if (b.status[i] != STATUS_DYNAMIC) return;
  // read-only walk: a halving write by another body's walk could land after this body stored its
  // root and leave it pointing at a non-root ancestor, whose flag k_act_apply would then read
  const int r = uf_root_ro(parent, i, b.n);
  __hip_atomic_store(parent + i, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (keeps_awake(b.act[i]) || touch[i]) can[r] = 0u;
}

__global__ void __launch_bounds__(256) k_act_apply(DevBodies b, const int* __restrict__ parent,
                                                  const uint32_t* __restrict__ can, uint32_t* awake) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool on = false;
  if (i < b.n && b.status[i] == STATUS_DYNAMIC) {
    float2 a = b.act[i];
    if (can[parent[i]]) {
      if (a.x != 0.0f) {
        a.x = 0.0f;
        b.linvel[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b.angvel[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
    } else if (a.x == 0.0f && a.y >= 0.0f) {
      a.x = a.y * 2.0f;
    }
    b.act[i] = a;
    on = a.x != 0.0f;
  }
  block_add(on ? 1u : 0u, awake);
}

__global__ void __launch_bounds__(256) k_act_manifolds(DevBodies b, DevContacts c, int M, uint32_t* flag) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int2 bb = c.mbody[m];
  const bool on1 = b.status[bb.x] == STATUS_DYNAMIC && b.act[bb.x].x != 0.0f;
  const bool on2 = b.status[bb.y] == STATUS_DYNAMIC && b.act[bb.y].x != 0.0f;
  flag[m] = (on1 || on2) ? 1u : 0u;
}

// Solved manifolds to `out` in manifold order; the other manifolds' contacts get the cached
// impulse of the same contact in the previous step's manifolds (w = -1 marks "no entry").
// totals[0] = solved manifolds, totals[1] = their contacts.
__global__ void __launch_bounds__(256) k_act_compact(DevContacts c, int M, const uint32_t* __restrict__ flag,
                                                    const uint32_t* __restrict__ pos, DevContacts out,
                                                    DevContacts prev, int prevM, uint32_t* totals) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t cnt = 0;
  if (m < M) {
    const int first = c.mfirst[m], count = c.mcount[m];
    if (flag[m]) {
      const uint32_t p = pos[m];
      out.mbody[p] = c.mbody[m];
      out.mfirst[p] = first;
      out.mcount[p] = count;
      cnt = uint32_t(count);
    } else {
      const int2 bb = c.mbody[m];
      int pf = 0, pc = 0;
      const int j = find_manifold(prev, prevM, bb.x, bb.y);
      if (j >= 0) {
        pf = prev.mfirst[j];
        pc = prev.mcount[j];
      }
      for (int k = 0; k < count; ++k) {
        const uint32_t feat = c.feature(first + k);
        float4 imp = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        for (int kk = 0; kk < pc; ++kk)
          if (prev.feature(pf + kk) == feat) {
            imp = prev.imp[pf + kk];
            break;
          }
        c.imp[first + k] = imp;
      }
    }
    if (m == M - 1) totals[0] = pos[m] + flag[m];
  }
  block_add(cnt, totals + 1);
}

int nb(long n) { return int((n + 255) / 256); }

}  // namespace

void launch_wake(const DevBodies& b, hipStream_t s) {
  if (b.n > 0) hipLaunchKernelGGL(k_act_wake, dim3(nb(b.n)), dim3(256), 0, s, b);
}

void launch_activation(const DevBodies& b, const DevContacts& c, int M, const ActivationBufs& a, const JointBufs* j,
                       hipStream_t s) {
  if (b.n > 0) {
    hipLaunchKernelGGL(k_act_energy, dim3(nb(b.n)), dim3(256), 0, s, b, a.parent, a.can, a.touch);
    if (M > 0) hipLaunchKernelGGL(k_act_union, dim3(nb(M)), dim3(256), 0, s, b, c, M, a.parent, a.touch);
    if (j) launch_joint_union(b, *j, a.parent, a.touch, s);  // joints join islands (:154-159)
    hipLaunchKernelGGL(k_act_flags, dim3(nb(b.n)), dim3(256), 0, s, b, a.parent, a.can, a.touch);
    hipLaunchKernelGGL(k_act_apply, dim3(nb(b.n)), dim3(256), 0, s, b, a.parent, a.can, a.totals + 2);
  }
  if (M > 0) {
    hipLaunchKernelGGL(k_act_manifolds, dim3(nb(M)), dim3(256), 0, s, b, c, M, a.mflag);
    scan_exclusive_u32(a.mflag, a.pos, M, a.tmp, s);
  }
}

void launch_solve_list(const DevContacts& c, int M, const ActivationBufs& a, const DevContacts& out,
                       const DevContacts& prev, int prevM, hipStream_t s) {
  if (M > 0)
    hipLaunchKernelGGL(k_act_compact, dim3(nb(M)), dim3(256), 0, s, c, M, a.mflag, a.pos, out, prev, prevM,
                       a.totals);
}

}  // namespace nph
