// nphysics_amd/csrc/k_broadphase.hip -- broad phase: loose AABBs, Morton LBVH, pair emission.
//
// Replaces ncollide's DBVTBroadPhase behind GeometricalWorld::perform_broad_phase
// (src/world/geometrical_world.rs:60, 285-312) and its DefaultCollisionFilter (:926-971):
//   * loose AABB rule (k_loose_update): tight = shape AABB at the collider position, merged with
//     the one at the predicted position when the CCD-phase sync sweeps the collider
//     (geometrical_world.rs:250-261, position0 -> position), loosened by margin + linear
//     prediction (collider.rs:542-549); the loose AABB is replaced by tight loosened by 0.01 only
//     when it no longer contains tight (the DBVT margin of geometrical_world.rs:60).
//   * pairs = every two bodies whose loose AABBs overlap (closed intervals), not the same body and
//     not both static; stored as (first, second) with first dynamic and second static or larger,
//     sorted.
// The pair set is found with a linear BVH rebuilt whenever a loose AABB changed: Morton codes of
// the loose-AABB centroids, stable LDS radix sort, Karras (HPG 2012) radix-tree construction, an
// atomic min/max refit with early termination, then one thread per dynamic body (in Morton order, so
// a wavefront walks neighbouring subtrees) traverses the tree and writes its partners into its
// own fixed-capacity segment; segments are sorted and compacted with a scan. The result is
// deterministic and independent of thread timing. Algorithmic bytes: 124 per body + 8 per pair.
// When only a few dynamic bodies' loose AABBs were replaced since the last pair build (the common
// case of a resting pile), the segments are updated instead (see the incremental kernels below):
// a few thousand traversals and one pass over the segments in place of a tree build and n
// traversals; the pair list is the same.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "stages.hpp"

namespace nph {

namespace {

constexpr int kStack = 64;

__device__ __forceinline__ uint32_t f2ord(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void __launch_bounds__(256) k_loose_update(DevBodies b, int mode, uint32_t* changed,
                                                     uint32_t* dirty) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  if (mode != 0 && b.status[i] != STATUS_DYNAMIC) return;
  const int kind = b.shape[i];
  const f3 he = xyz(b.dims[i]);
  aabb a = shape_aabb(kind, he, xyz(b.pos(i)), toq(b.rot(i)));
  if (mode == 2) {  // safe_position (rigid_body.rs:904-910): position0, with linear motion interpolation
                   // the current rotation
    const q4 r0 = (b.ccd && (b.ccd[i] & 2)) ? toq(b.rot(i)) : toq(b.rot0[i]);
    a = merged(shape_aabb(kind, he, xyz(b.pos0[i]), r0), a);
  }
  else a = merged(a, a);
  const aabb tight = loosened(a, b.qlim[i]);
  const aabb loose{xyz(b.loose_mn[i]), xyz(b.loose_mx[i])};
  if (!contains(loose, tight)) {
    const aabb nl = loosened(tight, 0.01f);
    b.loose_mn[i] = f4(nl.mn, 0.0f);
    b.loose_mx[i] = f4(nl.mx, 0.0f);
    dirty[i] = 1u;
    atomicAdd(&changed[0], 1u);
    if (b.status[i] != STATUS_DYNAMIC) atomicAdd(&changed[1], 1u);
  }
}

__global__ void __launch_bounds__(256) k_bounds(DevBodies b, uint32_t* bounds) {
  __shared__ uint32_t s_b[6][256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = threadIdx.x;
  uint32_t v[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  if (i < b.n) {
    const float4 mn = b.loose_mn[i], mx = b.loose_mx[i];
    const float cx = (mn.x + mx.x) * 0.5f, cy = (mn.y + mx.y) * 0.5f, cz = (mn.z + mx.z) * 0.5f;
    v[0] = v[3] = f2ord(cx);
    v[1] = v[4] = f2ord(cy);
    v[2] = v[5] = f2ord(cz);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) s_b[k][t] = v[k];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        s_b[k][t] = s_b[k][t] < s_b[k][t + off] ? s_b[k][t] : s_b[k][t + off];
        s_b[k + 3][t] = s_b[k + 3][t] > s_b[k + 3][t + off] ? s_b[k + 3][t] : s_b[k + 3][t + off];
      }
    }
    __syncthreads();
  }
  if (t < 3) atomicMin(&bounds[t], s_b[t][0]);
  else if (t < 6) atomicMax(&bounds[t], s_b[t][0]);
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Coordinate quantised to `levels` (2^bits - 1) steps of the centroid bounds.
__device__ __forceinline__ uint32_t quant(float c, float lo, float hi, float levels, uint32_t top) {
  const float ext = hi - lo;
  float u = ext > 0.0f ? (c - lo) / ext : 0.5f;
  u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
  const uint32_t q = uint32_t(u * levels);
  return q > top ? top : q;
}

// Key = scene << 3m | Morton code of the loose-AABB centroid with m bits per axis,
// m = (30 - scene_bits) / 3 (10 for a single scene).
__global__ void __launch_bounds__(256) k_morton(DevBodies b, const uint32_t* bounds, int scene_bits,
                                               uint32_t* keys, uint32_t* vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  const int m = (30 - scene_bits) / 3;
  const uint32_t top = (1u << m) - 1u;
  const float levels = float(top);
  const float lx = ord2f(bounds[0]), ly = ord2f(bounds[1]), lz = ord2f(bounds[2]);
  const float hx = ord2f(bounds[3]), hy = ord2f(bounds[4]), hz = ord2f(bounds[5]);
  const float4 mn = b.loose_mn[i], mx = b.loose_mx[i];
  const float cx = (mn.x + mx.x) * 0.5f, cy = (mn.y + mx.y) * 0.5f, cz = (mn.z + mx.z) * 0.5f;
  const uint32_t x = quant(cx, lx, hx, levels, top), y = quant(cy, ly, hy, levels, top),
                 z = quant(cz, lz, hz, levels, top);
  const uint32_t morton = (expand10(x) << 2) | (expand10(y) << 1) | expand10(z);
  keys[i] = scene_bits ? (uint32_t(b.scene[i]) << (3 * m)) | morton : morton;
  vals[i] = uint32_t(i);
}

__device__ __forceinline__ int delta(const uint32_t* __restrict__ keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = keys[i], c = keys[j];
  if (a == c) return 32 + __clz(uint32_t(i ^ j));
  return __clz(a ^ c);
}

// Karras 2012: internal node i spans a key range with one split; children < n-1 are internal,
// children >= n-1 are leaves (leaf t is node n-1+t).
__global__ void __launch_bounds__(256) k_lbvh(const uint32_t* __restrict__ keys, int n, int scene_shift,
                                             int* child, int* parent, int2* node_scene) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(keys, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + (d < 0 ? d : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const int left = lo == gamma ? (n - 1 + gamma) : gamma;
  const int right = hi == gamma + 1 ? (n - 1 + gamma + 1) : gamma + 1;
  child[2 * i] = left;
  child[2 * i + 1] = right;
  node_scene[i] = make_int2(int(keys[lo] >> scene_shift), int(keys[hi] >> scene_shift));
  parent[left] = i;
  parent[right] = i;
  if (i == 0) parent[0] = -1;
}

// Leaf nodes hold the loose AABBs in Morton order.
__global__ void __launch_bounds__(256) k_leaf_nodes(DevBodies b, const uint32_t* __restrict__ vals, int n,
                                                   float4* node_mn, float4* node_mx, int* leaf_of) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int body = int(vals[t]);
  node_mn[n - 1 + t] = b.loose_mn[body];
  node_mx[n - 1 + t] = b.loose_mx[body];
  leaf_of[body] = t;
}

// Internal-node bounds by atomic min/max on order-preserving integer encodings: every leaf climbs
// towards the root and stops at the first ancestor none of whose six bounds it changes (a value
// that does not change a node is dominated by one whose writer keeps climbing). No arrival
// counters or fences are needed and the bounds are exact (min/max of the leaves below).
__global__ void __launch_bounds__(256) k_refit_atomic(DevBodies b, const uint32_t* __restrict__ vals, int n,
                                                     const int* __restrict__ parent, uint32_t* inode) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int body = int(vals[t]);
  const float4 mn = b.loose_mn[body], mx = b.loose_mx[body];
  const uint32_t e[6] = {f2ord(mn.x), f2ord(mn.y), f2ord(mn.z), f2ord(mx.x), f2ord(mx.y), f2ord(mx.z)};
  int p = parent[n - 1 + t];
  while (p >= 0) {
    uint32_t* q = inode + 8 * long(p);
    bool changed = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (atomicMin(&q[k], e[k]) > e[k]) changed = true;
      if (atomicMax(&q[4 + k], e[3 + k]) < e[3 + k]) changed = true;
    }
    if (!changed) break;
    p = parent[p];
  }
}

__global__ void __launch_bounds__(256) k_decode_nodes(int n, const uint32_t* __restrict__ inode, float4* node_mn,
                                                     float4* node_mx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const uint32_t* q = inode + 8 * long(i);
  node_mn[i] = make_float4(ord2f(q[0]), ord2f(q[1]), ord2f(q[2]), 0.0f);
  node_mx[i] = make_float4(ord2f(q[4]), ord2f(q[5]), ord2f(q[6]), 0.0f);
}

__global__ void __launch_bounds__(256) k_init_inodes(int n, uint32_t* inode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  uint4* q = reinterpret_cast<uint4*>(inode + 8 * long(i));
  q[0] = make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0u);
  q[1] = make_uint4(0u, 0u, 0u, 0u);
}

// Packed internal nodes for the traversals: one 64-byte record per internal node with both
// children's bounds and references, so descending one level is a single 64-byte load:
//   [0] (child0 min, ref0)  [1] (child0 max, ref1)  [2] (child1 min, first scene)
//   [3] (child1 max, last scene)
// ref = internal node index, or kLeafRef | dynamic << 30 | body for a leaf; the scene range is
// that of the node's own leaves.
constexpr uint32_t kLeafRef = 0x80000000u;

__device__ __forceinline__ uint32_t child_ref(DevBodies b, const uint32_t* __restrict__ vals, int n, int c) {
  if (c < n - 1) return uint32_t(c);
  const int o = int(vals[c - (n - 1)]);
  return kLeafRef | (b.status[o] == STATUS_DYNAMIC ? 1u << 30 : 0u) | uint32_t(o);
}

__global__ void __launch_bounds__(256) k_pack_nodes(DevBodies b, const uint32_t* __restrict__ vals, int n,
                                                   const int* __restrict__ child,
                                                   const float4* __restrict__ node_mn,
                                                   const float4* __restrict__ node_mx,
                                                   const int2* __restrict__ node_scene, float4* pnode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int c0 = child[2 * i], c1 = child[2 * i + 1];
  const int2 sr = node_scene[i];
  float4* r = pnode + 4 * long(i);
  r[0] = f4(xyz(node_mn[c0]), __uint_as_float(child_ref(b, vals, n, c0)));
  r[1] = f4(xyz(node_mx[c0]), __uint_as_float(child_ref(b, vals, n, c1)));
  r[2] = f4(xyz(node_mn[c1]), __int_as_float(sr.x));
  r[3] = f4(xyz(node_mx[c1]), __int_as_float(sr.y));
}

// Walks the packed tree for the loose AABB of body q and calls leaf(o, dynamic) for every other
// body of q's scene whose loose AABB overlaps it and whose collision groups can interact with q's.
// Returns false on a stack overflow.
// The traversal stack lives in LDS, interleaved across the workgroup's lanes (entry k of lane t at
// stack[k * kTraverseBlock + t]): a stack in private memory would be scratch, one memory round trip
// per push and pop on the traversal's dependent chain.
constexpr int kTraverseBlock = 64;

template <bool kScenes, class F>
__device__ __forceinline__ bool walk_tree(const DevBodies& b, const float4* __restrict__ pnode, int q,
                                          int* __restrict__ stack, F&& leaf) {
  const aabb Q{xyz(b.loose_mn[q]), xyz(b.loose_mx[q])};
  const int qs = kScenes ? b.scene[q] : 0;
  const uint4 qg = b.groups ? b.groups[q] : uint4{};
  int sp = 0;
  int node = 0;
  bool ok = true;
  while (true) {
    const float4* r = pnode + 4 * long(node);
    const float4 a0 = r[0], a1 = r[1], a2 = r[2], a3 = r[3];
    if (!kScenes || (qs >= __float_as_int(a2.w) && qs <= __float_as_int(a3.w))) {
      const uint32_t ref[2] = {__float_as_uint(a0.w), __float_as_uint(a1.w)};
      const aabb B[2] = {{xyz(a0), xyz(a1)}, {xyz(a2), xyz(a3)}};
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (!intersects(Q, B[k])) continue;
        const uint32_t c = ref[k];
        if (c & kLeafRef) {
          const int o = int(c & 0x3fffffffu);
          if (o != q && (!kScenes || b.scene[o] == qs) && (!b.groups || groups_interact(qg, b.groups[o])))
            leaf(o, ((c >> 30) & 1u) != 0);
        } else if (sp < kStack) {
          stack[(sp++) * kTraverseBlock] = int(c);
        } else {
          ok = false;
        }
      }
    }
    if (sp == 0) break;
    node = stack[(--sp) * kTraverseBlock];
  }
  return ok;
}

template <bool kScenes>
__global__ void __launch_bounds__(kTraverseBlock) k_traverse(DevBodies b, const uint32_t* __restrict__ vals, int n,
                                                 const float4* __restrict__ pnode, int* cand,
                                                 uint32_t* cand_count, int cap, uint32_t* stats) {
  __shared__ int s_stack[kStack * kTraverseBlock];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int q = int(vals[t]);
  if (b.status[q] != STATUS_DYNAMIC || n < 2) {
    cand_count[q] = 0;
    return;
  }
  const long base = long(q) * cap;
  uint32_t cnt = 0;
  const bool ok = walk_tree<kScenes>(b, pnode, q, s_stack + threadIdx.x, [&](int o, bool dyn) {
    if (!dyn || o > q) {
      if (cnt < uint32_t(cap)) cand[base + cnt] = o;
      ++cnt;
    }
  });
  cand_count[q] = cnt;
  atomicMax(&stats[0], cnt);
  if (!ok) atomicOr(&stats[1], 1u);
}

// ---- incremental pair update -------------------------------------------------------------------
// Used when only dynamic bodies' loose AABBs were replaced since the last pair build (dirty[i]).
// Pairs of two clean bodies cannot have changed, so: (1) every clean body's segment drops its
// dirty partners (order kept: segments are sorted); (2) the dirty leaves get their new bounds and
// expand their ancestors by atomic min/max (node bounds stay a superset of the leaves below, which
// is all a traversal needs; a full build restores tight bounds); (3) each dirty body traverses the
// tree: a partner that is static or has a larger index goes into its own segment, a clean dynamic
// partner with a smaller index gets the dirty body appended to its segment (atomic count), a dirty
// one finds it itself. The segments then hold exactly what a full traversal would write, up to
// order, and are sorted by the compaction as before.

__global__ void __launch_bounds__(256) k_seg_filter(DevBodies b, const uint32_t* __restrict__ dirty, int* cand,
                                                   uint32_t* cand_count, int cap) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= b.n) return;
  if (b.status[q] != STATUS_DYNAMIC) return;
  if (dirty[q]) {
    cand_count[q] = 0;
    return;
  }
  const int c = int(cand_count[q]);
  int* seg = cand + long(q) * cap;
  int k = 0;
  for (int i = 0; i < c; ++i) {
    const int o = seg[i];
    if (!dirty[o]) seg[k++] = o;
  }
  if (k != c) cand_count[q] = uint32_t(k);
}

// Dirty dynamic bodies (in arbitrary order: the segments are sorted afterwards).
__global__ void __launch_bounds__(256) k_dirty_list(DevBodies b, const uint32_t* __restrict__ vals, int n,
                                                   const uint32_t* __restrict__ dirty, int* list,
                                                   uint32_t* list_count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int q = int(vals[t]);
  if (dirty[q] && b.status[q] == STATUS_DYNAMIC) list[atomicAdd(list_count, 1u)] = q;
}

__global__ void __launch_bounds__(256) k_refit_dirty(DevBodies b, const int* __restrict__ list,
                                                    const uint32_t* __restrict__ list_count, int n,
                                                    const int* __restrict__ leaf_of, const int* __restrict__ parent,
                                                    float4* node_mn, float4* node_mx, uint32_t* inode) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= int(*list_count)) return;
  const int body = list[t];
  const int leaf = n - 1 + leaf_of[body];
  const float4 mn = b.loose_mn[body], mx = b.loose_mx[body];
  node_mn[leaf] = mn;
  node_mx[leaf] = mx;
  const uint32_t e[6] = {f2ord(mn.x), f2ord(mn.y), f2ord(mn.z), f2ord(mx.x), f2ord(mx.y), f2ord(mx.z)};
  int p = parent[leaf];
  while (p >= 0) {
    uint32_t* q = inode + 8 * long(p);
    bool changed = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (atomicMin(&q[k], e[k]) > e[k]) changed = true;
      if (atomicMax(&q[4 + k], e[3 + k]) < e[3 + k]) changed = true;
    }
    if (!changed) break;
    p = parent[p];
  }
}

template <bool kScenes>
__global__ void __launch_bounds__(kTraverseBlock) k_traverse_dirty(DevBodies b, const int* __restrict__ list,
                                                       const uint32_t* __restrict__ list_count,
                                                       const float4* __restrict__ pnode,
                                                       const uint32_t* __restrict__ dirty, int* cand,
                                                       uint32_t* cand_count, int cap, uint32_t* stats) {
  __shared__ int s_stack[kStack * kTraverseBlock];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= int(*list_count)) return;
  const int q = list[t];
  const long base = long(q) * cap;
  uint32_t cnt = 0, most = 0;
  const bool ok = walk_tree<kScenes>(b, pnode, q, s_stack + threadIdx.x, [&](int o, bool dyn) {
    if (!dyn || o > q) {
      if (cnt < uint32_t(cap)) cand[base + cnt] = o;
      ++cnt;
    } else if (!dirty[o]) {
      const uint32_t pos = atomicAdd(&cand_count[o], 1u);
      if (pos < uint32_t(cap)) cand[long(o) * cap + pos] = q;
      most = pos + 1 > most ? pos + 1 : most;
    }
  });
  cand_count[q] = cnt;
  atomicMax(&stats[0], cnt > most ? cnt : most);
  if (!ok) atomicOr(&stats[1], 1u);
}

__global__ void __launch_bounds__(256) k_clamp_counts(const uint32_t* cnt, uint32_t* out, int n, int cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = cnt[i];
  out[i] = c < uint32_t(cap) ? c : uint32_t(cap);
}

__global__ void __launch_bounds__(256) k_sort_compact(int n, int* cand, const uint32_t* __restrict__ cnt,
                                                     const uint32_t* __restrict__ off, int cap,
                                                     int2* pairs) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uint32_t c0 = cnt[q];
  const int c = int(c0 < uint32_t(cap) ? c0 : uint32_t(cap));
  if (c == 0) return;
  int* seg = cand + long(q) * cap;
  for (int i = 1; i < c; ++i) {
    const int v = seg[i];
    int j = i - 1;
    while (j >= 0 && seg[j] > v) {
      seg[j + 1] = seg[j];
      --j;
    }
    seg[j + 1] = v;
  }
  const uint32_t o = off[q];
  for (int i = 0; i < c; ++i) pairs[o + i] = make_int2(q, seg[i]);
}

inline unsigned nb(long n) { return unsigned((n + 255) / 256); }

}  // namespace

void launch_loose_update(const DevBodies& b, int mode, uint32_t* changed, uint32_t* dirty, hipStream_t s) {
  if (b.n == 0) return;
  hipLaunchKernelGGL(k_loose_update, dim3(nb(b.n)), dim3(256), 0, s, b, mode, changed, dirty);
}

void launch_build_lbvh(const DevBodies& b, BroadPhaseBufs& bp, hipStream_t s) {
  const int n = b.n;
  if (n < 2) return;
  const uint32_t init[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  (void)hipMemcpyAsync(bp.bounds, init, sizeof(init), hipMemcpyHostToDevice, s);
  hipLaunchKernelGGL(k_bounds, dim3(nb(n)), dim3(256), 0, s, b, bp.bounds);
  hipLaunchKernelGGL(k_morton, dim3(nb(n)), dim3(256), 0, s, b, bp.bounds, bp.scene_bits, bp.keys, bp.vals);
  radix_sort_pairs(bp.keys, bp.vals, bp.keys_alt, bp.vals_alt, n, 30, bp.tmp, s);
  const int scene_shift = 3 * ((30 - bp.scene_bits) / 3);
  hipLaunchKernelGGL(k_lbvh, dim3(nb(n - 1)), dim3(256), 0, s, bp.keys, n, scene_shift, bp.child, bp.parent,
                     bp.node_scene);
  hipLaunchKernelGGL(k_init_inodes, dim3(nb(n - 1)), dim3(256), 0, s, n, bp.visit);
  hipLaunchKernelGGL(k_leaf_nodes, dim3(nb(n)), dim3(256), 0, s, b, bp.vals, n, bp.node_mn, bp.node_mx,
                     bp.leaf_of);
  hipLaunchKernelGGL(k_refit_atomic, dim3(nb(n)), dim3(256), 0, s, b, bp.vals, n, bp.parent, bp.visit);
  hipLaunchKernelGGL(k_decode_nodes, dim3(nb(n - 1)), dim3(256), 0, s, n, bp.visit, bp.node_mn, bp.node_mx);
  hipLaunchKernelGGL(k_pack_nodes, dim3(nb(n - 1)), dim3(256), 0, s, b, bp.vals, n, bp.child, bp.node_mn,
                     bp.node_mx, bp.node_scene, bp.pnode);
}

void launch_traverse(const DevBodies& b, BroadPhaseBufs& bp, int cap, hipStream_t s) {
  const int n = b.n;
  if (n == 0) return;
  fill_u32(bp.stats, 0u, 4, s);
  if (n < 2) {
    fill_u32(bp.cand_count, 0u, size_t(n), s);
    return;
  }
  if (bp.scene_bits)
    hipLaunchKernelGGL(k_traverse<true>, dim3(unsigned((n + kTraverseBlock - 1) / kTraverseBlock)), dim3(kTraverseBlock), 0, s, b, bp.vals, n, bp.pnode, bp.cand,
                       bp.cand_count, cap, bp.stats);
  else
    hipLaunchKernelGGL(k_traverse<false>, dim3(unsigned((n + kTraverseBlock - 1) / kTraverseBlock)), dim3(kTraverseBlock), 0, s, b, bp.vals, n, bp.pnode, bp.cand,
                       bp.cand_count, cap, bp.stats);
}

void launch_incremental_pairs(const DevBodies& b, BroadPhaseBufs& bp, int cap, int max_dirty, hipStream_t s) {
  const int n = b.n;
  if (n < 2) return;
  fill_u32(bp.stats, 0u, 4, s);
  fill_u32(bp.list_count, 0u, 1, s);
  hipLaunchKernelGGL(k_seg_filter, dim3(nb(n)), dim3(256), 0, s, b, bp.dirty, bp.cand, bp.cand_count, cap);
  hipLaunchKernelGGL(k_dirty_list, dim3(nb(n)), dim3(256), 0, s, b, bp.vals, n, bp.dirty, bp.list, bp.list_count);
  if (max_dirty <= 0) return;
  hipLaunchKernelGGL(k_refit_dirty, dim3(nb(max_dirty)), dim3(256), 0, s, b, bp.list, bp.list_count, n,
                     bp.leaf_of, bp.parent, bp.node_mn, bp.node_mx, bp.visit);
  hipLaunchKernelGGL(k_decode_nodes, dim3(nb(n - 1)), dim3(256), 0, s, n, bp.visit, bp.node_mn, bp.node_mx);
  hipLaunchKernelGGL(k_pack_nodes, dim3(nb(n - 1)), dim3(256), 0, s, b, bp.vals, n, bp.child, bp.node_mn,
                     bp.node_mx, bp.node_scene, bp.pnode);
  if (bp.scene_bits)
    hipLaunchKernelGGL(k_traverse_dirty<true>, dim3(unsigned((max_dirty + kTraverseBlock - 1) / kTraverseBlock)), dim3(kTraverseBlock), 0, s, b, bp.list, bp.list_count,
                       bp.pnode, bp.dirty, bp.cand, bp.cand_count, cap, bp.stats);
  else
    hipLaunchKernelGGL(k_traverse_dirty<false>, dim3(unsigned((max_dirty + kTraverseBlock - 1) / kTraverseBlock)), dim3(kTraverseBlock), 0, s, b, bp.list, bp.list_count,
                       bp.pnode, bp.dirty, bp.cand, bp.cand_count, cap, bp.stats);
}

void launch_compact_pairs(const DevBodies& b, BroadPhaseBufs& bp, int cap, hipStream_t s) {
  const int n = b.n;
  if (n == 0) return;
  hipLaunchKernelGGL(k_sort_compact, dim3(nb(n)), dim3(256), 0, s, n, bp.cand, bp.cand_count, bp.cand_off,
                     cap, bp.pairs);
}

// Exclusive scan of min(count, cap) into cand_off; the clamped counts reuse keys_alt.
void launch_pair_offsets(const DevBodies& b, BroadPhaseBufs& bp, int cap, hipStream_t s) {
  const int n = b.n;
  if (n == 0) return;
  hipLaunchKernelGGL(k_clamp_counts, dim3(nb(n)), dim3(256), 0, s, bp.cand_count, bp.keys_alt, n, cap);
  scan_exclusive_u32(bp.keys_alt, bp.cand_off, n, bp.tmp, s);
}

}  // namespace nph
