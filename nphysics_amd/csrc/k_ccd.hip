// nphysics_amd/csrc/k_ccd.hip -- continuous collision detection: times of impact, the impact
// queue, and the substep island.
//
// Restates MechanicalWorld::predict_next_impacts and the island traversal of solve_ccd
// (src/world/mechanical_world.rs:400-557, 630-861) with TOIEntry::try_from_colliders (:946-1022):
//   * k_ccd_first_pass, one lane per broad-phase pair with a CCD-enabled side: the pair gets a CCD
//     slot and its time of impact up to the step's end (conservative advancement, np_geometry.hpp);
//     impacts within the window go to the queue, later ones lower the earliest-later time;
//   * k_ccd_events, one workgroup: wave 0 pops the earliest impact (ties by pair index) from an
//     indexed 64-ary heap, confirms pairs of two frozen bodies only at the later freeze time, and
//     freezes the two bodies (their keys: a static collider shares its static body's); then every
//     lane re-sweeps the CCD slots of the two bodies' dynamic colliders with the frozen bodies held
//     at their freeze poses, and wave 0 replaces, removes or adds those slots' entries;
//   * the island of a substep grows in rounds: the dynamic bodies of every pair touching a
//     traversed collider are moved to their impact time (k_ccd_touch, k_ccd_prepare; the dynamics
//     update of the substep, no gravity, forces cleared), those pairs' manifolds are computed by
//     the narrow phase at the moved poses, and the dynamic bodies of the non-empty ones are
//     traversed from the next round (k_ccd_grow, k_ccd_merge);
//   * k_ccd_validate (validate_advancement of the impacting bodies) and k_ccd_clamp (bodies still
//     predicted to impact go back to the start of their motion, rigid_body.rs:530-537).
// The queue is serial by nature; it holds the few impacts of a step, while the time-of-impact
// queries (the arithmetic) run one per lane.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "np_geometry.hpp"
#include "stages.hpp"

namespace nph {

namespace {

constexpr int kEventBlock = 512;  // eight waves: one freeze re-sweeps a few hundred swept-AABB pairs in one round

__device__ __forceinline__ unsigned long long toi_key(float toi, int pair) {
  return (static_cast<unsigned long long>(__float_as_uint(toi)) << 32) | static_cast<unsigned long long>(uint32_t(pair));
}
__device__ __forceinline__ float key_toi(unsigned long long k) { return __uint_as_float(uint32_t(k >> 32)); }
__device__ __forceinline__ int key_pair(unsigned long long k) { return int(uint32_t(k & 0xffffffffull)); }

// Body::part_motion (rigid_body.rs:539-553): static bodies at their pose; with linear motion
// interpolation from (position0, current rotation); else from position0 with rotation.
__device__ __forceinline__ CcdMotion part_motion(const DevBodies& b, int i, float t0) {
  CcdMotion m;
  m.t0 = t0;
  m.v = xyz(b.linvel[i]);
  m.w = xyz(b.angvel[i]);
  if (b.status[i] != STATUS_DYNAMIC) {
    m.kind = MOTION_STATIC;
    m.t0p = xyz(b.pos(i));
    m.q0 = toq(b.rot(i));
  } else if (b.ccd[i] & 2) {
    m.kind = MOTION_LINEAR;
    m.t0p = xyz(b.pos0[i]);
    m.q0 = toq(b.rot(i));
  } else {
    m.kind = MOTION_ROTATING;
    m.t0p = xyz(b.pos0[i]);
    m.q0 = toq(b.rot0[i]);
  }
  return m;
}

// TOIEntry::try_from_colliders for bodies i, j with freeze times f1, f2 (< 0: not frozen): the
// global time of impact up to end_time.
__device__ bool pair_toi(const DevBodies& b, const CcdBufs& c, int i, int j, float f1, float f2, float end_time,
                         float target, int on_penetration, float* out) {
  const float bt1 = c.body_time[i], bt2 = c.body_time[j];
  const float time1 = f1 >= 0.0f ? f1 : bt1;
  const float time2 = f2 >= 0.0f ? f2 : bt2;
  const float start = time1 > time2 ? time1 : time2;
  const float o1 = time1 - start, o2 = time2 - start;
  CcdMotion m1 = part_motion(b, i, o1), m2 = part_motion(b, j, o2);
  if (f1 >= 0.0f) {
    m1.at((f1 + o1) - bt1, m1.t0p, m1.q0);
    m1.kind = MOTION_STATIC;
  }
  if (f2 >= 0.0f) {
    m2.at((f2 + o2) - bt2, m2.t0p, m2.q0);
    m2.kind = MOTION_STATIC;
  }
  float toi = 0.0f;
  int status = 0;
  if (!time_of_impact(b.shape[i], xyz(b.dims[i]), m1, b.shape[j], xyz(b.dims[j]), m2, end_time - start, target, &toi,
                      &status))
    return false;
  if (!on_penetration && status == TOI_PENETRATING) return false;
  *out = start + toi;
  return true;
}

__global__ void k_ccd_reset(uint32_t* ctl) {
  const int t = threadIdx.x;
  if (t < kCcdWords) ctl[t] = t == kCcdMinAfter ? 0x7f800000u : 0u;  // +inf
}

__global__ void __launch_bounds__(128) k_ccd_first_pass(DevBodies b, const int2* __restrict__ pairs, int npairs,
                                                        CcdBufs c, float end_time, float dt0, float target,
                                                        int on_penetration) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int2 pr = pairs[p];
  if (!((b.ccd[pr.x] | b.ccd[pr.y]) & 1)) return;
  const uint32_t slot = atomicAdd(&c.ctl[kCcdSlots], 1u);
  c.slot_pair[slot] = p;
  c.slot_bodies[slot] = int4{pr.x, pr.y, b.ccd_key[pr.x] | (b.status[pr.x] != STATUS_DYNAMIC ? kSlotStatic : 0),
                             b.ccd_key[pr.y] | (b.status[pr.y] != STATUS_DYNAMIC ? kSlotStatic : 0)};
  c.pos[slot] = -1;
  float toi;
  if (!pair_toi(b, c, pr.x, pr.y, -1.0f, -1.0f, dt0, target, on_penetration, &toi)) return;
  if (toi > end_time) {
    atomicMin(&c.ctl[kCcdMinAfter], __float_as_uint(toi));  // toi > 0: the bits order as the values
  } else {
    atomicOr(&c.ctl[kCcdAnyIn], 1u);
    const uint32_t e = atomicAdd(&c.ctl[kCcdEntries], 1u);
    if (int(e) < c.cap) c.heap[e] = ToiEntry{toi_key(toi, p), int(slot)};
    else atomicOr(&c.ctl[kCcdOverflow], 1u);
  }
}

// The CCD slots of every dynamic body (adjacency in slot order is not needed: the re-sweep's
// results go into the queue, whose order is by key).
__global__ void __launch_bounds__(256) k_ccd_degree(int npairs, CcdBufs c) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= npairs || slot >= int(c.ctl[kCcdSlots])) return;
  const int4 sb = c.slot_bodies[slot];
  if (!(sb.z & kSlotStatic)) atomicAdd(&c.deg[sb.x], 1u);
  if (!(sb.w & kSlotStatic)) atomicAdd(&c.deg[sb.y], 1u);
}

__global__ void __launch_bounds__(256) k_ccd_fill(int npairs, CcdBufs c) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= npairs || slot >= int(c.ctl[kCcdSlots])) return;
  const int4 sb = c.slot_bodies[slot];
  if (!(sb.z & kSlotStatic)) c.adj[c.off[sb.x] + atomicAdd(&c.fill[sb.x], 1u)] = slot;
  if (!(sb.w & kSlotStatic)) c.adj[c.off[sb.y] + atomicAdd(&c.fill[sb.y], 1u)] = slot;
}

// ---- the impact queue: an indexed 64-ary heap, its top three levels in LDS, operated by wave 0 ----
// Each CCD slot holds at most one entry (its latest time of impact), so a re-sweep replaces or
// removes it in place instead of leaving a superseded entry behind: the pop sequence of the valid
// entries is the reference's, without popping stale ones. A node's 64 children are compared across
// the wave in one step; nodes 0..4160 (levels 0-2) live in LDS and deeper ones in HBM, so a node's
// children are all in one of the two. Keys are unique (they hold the pair index): the order is
// total. pos[slot] (HBM) is the node of each slot's entry, -1 when it has none.
constexpr int kHeapD = 64;
constexpr int kTop = 1 + kHeapD + kHeapD * kHeapD;  // 4,161 nodes, 12 B each in LDS
constexpr unsigned long long kNoKey = ~0ull;

struct Queue {
  unsigned long long* lk;  // LDS keys, nodes < kTop
  int* ls;                 // LDS slots
  ToiEntry* g;             // HBM entries, nodes >= kTop
  int* pos;
  __device__ unsigned long long key(int i) const { return i < kTop ? lk[i] : g[i].key; }
  __device__ int slot(int i) const { return i < kTop ? ls[i] : g[i].slot; }
  __device__ void set(int i, unsigned long long k, int s) const {  // one lane
    if (i < kTop) {
      lk[i] = k;
      ls[i] = s;
    } else {
      g[i] = ToiEntry{k, s};
    }
    pos[s] = i;
  }
};

// Smallest value across the wave, in every lane: DPP steps within quads, half rows and rows, then
// row broadcasts; lane 63 ends with the minimum.
template <int Ctrl, int RowMask>
__device__ __forceinline__ uint32_t dpp_min_step(uint32_t v) {
  const uint32_t o = uint32_t(__builtin_amdgcn_update_dpp(int(v), int(v), Ctrl, RowMask, 0xf, false));
  return o < v ? o : v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  v = dpp_min_step<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v = dpp_min_step<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v = dpp_min_step<0x141, 0xf>(v);  // row_half_mirror
  v = dpp_min_step<0x140, 0xf>(v);  // row_mirror
  v = dpp_min_step<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_min_step<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return uint32_t(__builtin_amdgcn_readlane(int(v), 63));
}

// The smallest of the wave's 64 keys (uniform) and the lane holding it: the time bits first, then
// the pair index among the lanes with that time.
__device__ __forceinline__ int wave_argmin(unsigned long long k, unsigned long long& mk) {
  const uint32_t hi = uint32_t(k >> 32), lo = uint32_t(k & 0xffffffffull);
  const uint32_t mh = wave_min_u32(hi);
  const uint32_t ml = wave_min_u32(hi == mh ? lo : 0xffffffffu);
  mk = (static_cast<unsigned long long>(mh) << 32) | ml;
  return __ffsll(static_cast<long long>(__ballot(hi == mh && lo == ml))) - 1;
}

// Every lane of wave 0 runs the queue operations with the same values (lane 0 writes); a lane may
// track one slot's node (tslot, tnode), kept current as entries move.
__device__ void sift_down(const Queue& q, int n, int k, unsigned long long ek, int es, int lane, int tslot,
                          int& tnode) {
  for (;;) {
    const int c0 = kHeapD * k + 1;
    if (c0 >= n) break;
    const int ch = c0 + lane;
    unsigned long long ck = kNoKey;
    int cs = -1;
    if (ch < n) {
      ck = q.key(ch);
      cs = q.slot(ch);
    }
    unsigned long long mk;
    const int wl = wave_argmin(ck, mk);
    if (!(mk < ek)) break;
    const int bs = __builtin_amdgcn_readlane(cs, wl);
    if (lane == 0) q.set(k, mk, bs);
    if (tslot == bs) tnode = k;
    k = c0 + wl;
  }
  if (lane == 0) q.set(k, ek, es);
  if (tslot == es) tnode = k;
}

__device__ void sift_up(const Queue& q, int k, unsigned long long ek, int es, int lane, int tslot, int& tnode) {
  while (k > 0) {
    const int par = (k - 1) / kHeapD;
    const unsigned long long pk = q.key(par);
    if (!(ek < pk)) break;
    const int ps = q.slot(par);
    if (lane == 0) q.set(k, pk, ps);
    if (tslot == ps) tnode = k;
    k = par;
  }
  if (lane == 0) q.set(k, ek, es);
  if (tslot == es) tnode = k;
}

// The entry at node p takes key (ek, es): up or down as the key decreased or grew.
__device__ void place(const Queue& q, int n, int p, unsigned long long ek, int es, int lane, int tslot, int& tnode) {
  if (ek < q.key(p)) sift_up(q, p, ek, es, lane, tslot, tnode);
  else sift_down(q, n, p, ek, es, lane, tslot, tnode);
}

// Freeze times held in LDS for up to this many body keys (the rest of the 160 KB beside the queue).
constexpr int kFzCap = 27648;

// predict_next_impacts' event loop (mechanical_world.rs:453-557) in one workgroup; with LdsFrozen
// the freeze times live in LDS for the loop and are written back at its end.
template <bool LdsFrozen>
__global__ void __launch_bounds__(kEventBlock) k_ccd_events(DevBodies b, CcdBufs c, float end_time, float dt0,
                                                            float target, int on_penetration) {
  __shared__ unsigned long long s_hk[kTop];
  __shared__ int s_hs[kTop];
  __shared__ float s_fz[LdsFrozen ? kFzCap : 1];
  __shared__ int s_go, s_i, s_j, s_nstage, s_hsize, s_nimp;
  float* const fz = LdsFrozen ? s_fz : c.frozen;
  constexpr int kKey = kSlotStatic - 1;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int nent = int(c.ctl[kCcdEntries]);
  const Queue q{s_hk, s_hs, c.heap, c.pos};
  if (t == 0) {
    const bool any = c.ctl[kCcdAnyIn] != 0u;
    // min_overstep (mechanical_world.rs:412-451): the step's end lowered to the earliest impact
    // after the window, or the window's end when an impact lies within it
    float min_over = dt0;
    const float after = __uint_as_float(c.ctl[kCcdMinAfter]);
    if (after < min_over) min_over = after;
    if (any) min_over = end_time;
    int kind = PREDICT_IMPACTS;
    if (min_over == dt0 && !any) kind = PREDICT_NONE;
    else if (min_over > end_time) kind = PREDICT_AFTER_END;
    c.ctl[kCcdKind] = uint32_t(kind);
    c.ctl[kCcdMinOver] = __float_as_uint(min_over);
    s_go = (kind == PREDICT_IMPACTS && c.ctl[kCcdOverflow] == 0u) ? 1 : 0;
    s_hsize = nent;
    s_nimp = 0;
  }
  __syncthreads();
  if (!s_go) {
    if (t == 0) c.ctl[kCcdImpacts] = 0u;
    return;
  }
  // ---- the first pass's entries into a heap: the LDS levels copied in, then the internal nodes
  // sifted down level by level from the deepest (the subtrees of one level are disjoint: one wave
  // per node), then every entry's node recorded
  for (int k = t; k < nent && k < kTop; k += kEventBlock) {
    s_hk[k] = c.heap[k].key;
    s_hs[k] = c.heap[k].slot;
  }
  if (LdsFrozen)
    for (int k = t; k < b.n; k += kEventBlock) s_fz[k] = c.frozen[k];
  __syncthreads();
  if (nent > 1) {
    const int last = (nent - 2) / kHeapD;  // the last internal node
    int first = 0, width = 1;
    while (first + width <= last) {  // the first node of the deepest internal level
      first += width;
      width *= kHeapD;
    }
    for (;;) {
      const int end = first + width - 1 < last ? first + width - 1 : last;
      int none = -1;
      for (int k = first + wave; k <= end; k += kEventBlock / 64) sift_down(q, nent, k, q.key(k), q.slot(k), lane, -1, none);
      __syncthreads();
      if (first == 0) break;
      width /= kHeapD;
      first -= width;
    }
  }
  for (int k = t; k < nent; k += kEventBlock) c.pos[q.slot(k)] = k;
  __syncthreads();
  for (;;) {
    // ---- pop to the next impact that freezes a body
    if (t < 64) {
      int n = s_hsize, nimp = s_nimp, found = 0, none = -1;
      while (n > 0) {
        if (nimp >= c.cap) {  // the impact list is full: reported as an overflow
          if (lane == 0) c.ctl[kCcdOverflow] = 1u;
          break;
        }
        const unsigned long long ek = s_hk[0];
        const int es = s_hs[0];
        --n;
        if (n > 0) sift_down(q, n, 0, q.key(n), q.slot(n), lane, -1, none);
        int d = 0;  // 0 next entry, 1 freeze
        if (lane == 0) {
          c.pos[es] = -1;
          const float toi = key_toi(ek);
          const int4 sb = c.slot_bodies[es];
          const int ki = sb.z & kKey, kj = sb.w & kKey;
          const float t1 = fz[ki], t2 = fz[kj];
          if (t1 >= 0.0f && t2 >= 0.0f) {
            if (toi == (t1 > t2 ? t1 : t2)) c.impacts[nimp++] = ek;
          } else {
            if (t1 < 0.0f) fz[ki] = toi;
            if (t2 < 0.0f) fz[kj] = toi;
            c.impacts[nimp++] = ek;
            s_i = sb.x;
            s_j = sb.y;
            d = 1;
          }
        }
        d = __shfl(d, 0);
        nimp = __shfl(nimp, 0);
        if (d) {
          found = 1;
          break;
        }
      }
      if (lane == 0) {
        s_hsize = n;
        s_nimp = nimp;
        s_go = found;
        s_nstage = 0;
      }
    }
    __syncthreads();
    if (!s_go) break;
    // ---- re-sweep the CCD slots of the two bodies' dynamic colliders (mechanical_world.rs:503-551);
    // the event's own pair is in both lists and skipped there (both bodies are frozen). A slot's
    // entry is replaced by its new time, or removed when there is none.
    const int si = s_i, sj = s_j;
    const int ni = int(c.deg[si]), nj = int(c.deg[sj]);  // 0 for a body that is not dynamic
    for (int k = t; k < ni + nj; k += kEventBlock) {
      const int slot = k < ni ? c.adj[c.off[si] + k] : c.adj[c.off[sj] + (k - ni)];
      const int4 sb = c.slot_bodies[slot];
      const int p = c.slot_pair[slot];
      const int had = c.pos[slot];
      const float fa = fz[sb.z & kKey], fb = fz[sb.w & kKey];
      if ((fa >= 0.0f || (sb.z & kSlotStatic)) && (fb >= 0.0f || (sb.w & kSlotStatic))) continue;
      float toi;
      const bool hit = pair_toi(b, c, sb.x, sb.y, fa, fb, end_time, target, on_penetration, &toi);
      if (hit || had >= 0) {
        const int st = atomicAdd(&s_nstage, 1);
        c.stage[st] = ToiEntry{hit ? toi_key(toi, p) : kNoKey, slot};
      }
    }
    __syncthreads();
    // ---- the re-sweep's results into the queue, 64 at a time: each lane loads one entry and its
    // slot's node, and keeps the node current while the entries before it move
    if (t < 64) {
      int n = s_hsize, ovf = 0;
      const int ns = s_nstage;
      for (int base = 0; base < ns && !ovf; base += 64) {
        const int cnt = ns - base < 64 ? ns - base : 64;
        unsigned long long mkey = kNoKey;
        int tslot = -1, tnode = -1;
        if (lane < cnt) {
          const ToiEntry u = c.stage[base + lane];
          mkey = u.key;
          tslot = u.slot;
          tnode = c.pos[u.slot];
        }
        for (int k = 0; k < cnt; ++k) {
          const uint32_t khi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(mkey >> 32)), k));
          const uint32_t klo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(mkey & 0xffffffffull)), k));
          const unsigned long long uk = (static_cast<unsigned long long>(khi) << 32) | klo;
          const int us = __builtin_amdgcn_readlane(tslot, k);
          const int p = __builtin_amdgcn_readlane(tnode, k);
          if (lane == k) tslot = -1;  // this lane's slot is handled now
          if (uk == kNoKey) {  // removed: the last entry takes its node (a removal always has one)
            if (p >= 0) {
              --n;
              if (p < n) place(q, n, p, q.key(n), q.slot(n), lane, tslot, tnode);
              if (lane == 0) c.pos[us] = -1;
            }
          } else if (p >= 0) {
            place(q, n, p, uk, us, lane, tslot, tnode);
          } else if (n < c.cap) {
            sift_up(q, n, uk, us, lane, tslot, tnode);
            ++n;
          } else {  // the queue is full: reported as an overflow
            if (lane == 0) c.ctl[kCcdOverflow] = 1u;
            ovf = 1;
            break;
          }
        }
      }
      if (lane == 0) {
        s_hsize = n;
        if (ovf) s_go = 0;
      }
    }
    __syncthreads();
    if (!s_go) break;
  }
  if (LdsFrozen) {
    __syncthreads();
    for (int k = t; k < b.n; k += kEventBlock) c.frozen[k] = s_fz[k];
  }
  if (t == 0) {
    const int nimp = s_nimp;
    c.ctl[kCcdImpacts] = uint32_t(nimp);
    c.ctl[kCcdLastToi] = nimp > 0 ? uint32_t(c.impacts[nimp - 1] >> 32) : 0u;
  }
}

__global__ void __launch_bounds__(256) k_ccd_seed(DevBodies b, const int2* __restrict__ pairs, CcdBufs c, int nimp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nimp) return;
  const int2 pr = pairs[key_pair(c.impacts[k])];
  c.in_t[b.ccd_key[pr.x]] = 1u;
  c.in_t[b.ccd_key[pr.y]] = 1u;
}

__global__ void __launch_bounds__(256) k_ccd_touch(DevBodies b, const int2* __restrict__ pairs, int npairs,
                                                   CcdBufs c) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int2 pr = pairs[p];
  const bool on = !c.evaluated[p] && (c.in_t[b.ccd_key[pr.x]] || c.in_t[b.ccd_key[pr.y]]);
  c.cand[p] = on ? 1u : 0u;
  if (!on) return;
  if (b.status[pr.x] == STATUS_DYNAMIC) c.prep[pr.x] = 1u;
  if (b.status[pr.y] == STATUS_DYNAMIC) c.prep[pr.y] = 1u;
}

// prepare_body (mechanical_world.rs:653-681): forces cleared, the dynamics update of the substep
// (dt_sub, no gravity) at the current pose, then RigidBody::advance to the body's freeze time (or
// the last impact time) along its motion and clamp_advancement.
__global__ void __launch_bounds__(256) k_ccd_prepare(DevBodies b, CcdBufs c, float dt_sub, float last_toi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n || !c.prep[i] || c.prepared[i]) return;
  c.prepared[i] = 1u;
  if (b.f_lin) {
    b.f_lin[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    b.f_ang[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  body_prestep(b, i, toq(b.rot(i)), xyz(b.angvel[i]), false, false, f3{0.0f, 0.0f, 0.0f}, dt_sub, dt_sub,
               f3{0.0f, 0.0f, 0.0f});
  const float f = c.frozen[b.ccd_key[i]];
  const float target = f >= 0.0f ? f : last_toi;
  const CcdMotion m = part_motion(b, i, 0.0f);
  f3 tp;
  q4 q;
  m.at(target - c.body_time[i], tp, q);
  b.pos0[i] = f4(tp, 0.0f);
  b.rot0[i] = fromq(q);
  b.pos(i) = f4(tp, 0.0f);
  if (!(b.ccd[i] & 2)) b.rot(i) = fromq(q);
  c.body_time[i] = last_toi;
}

__global__ void __launch_bounds__(256) k_ccd_scatter(const int2* __restrict__ pairs, int npairs,
                                                     const uint32_t* __restrict__ flag, CcdBufs c) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const uint32_t at = c.cpos[p];
  if (p == npairs - 1) c.ctl[kCcdList] = at + flag[p];
  if (!flag[p]) return;
  c.list[at] = pairs[p];
  c.list_pair[at] = p;
}

__global__ void __launch_bounds__(256) k_ccd_grow(DevBodies b, CcdBufs c, int nlist, const uint32_t* __restrict__ count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nlist) return;
  const int p = c.list_pair[k];
  c.evaluated[p] = 1u;
  if (count[k] == 0u) return;
  c.kept[p] = 1u;
  const int2 pr = c.list[k];
  if (b.status[pr.x] == STATUS_DYNAMIC && !c.in_t[b.ccd_key[pr.x]]) c.new_t[b.ccd_key[pr.x]] = 1u;
  if (b.status[pr.y] == STATUS_DYNAMIC && !c.in_t[b.ccd_key[pr.y]]) c.new_t[b.ccd_key[pr.y]] = 1u;
}

__global__ void __launch_bounds__(256) k_ccd_merge(int n, CcdBufs c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !c.new_t[i]) return;
  c.in_t[i] = 1u;
  c.new_t[i] = 0u;
}

__global__ void __launch_bounds__(256) k_ccd_validate(DevBodies b, const int2* __restrict__ pairs, CcdBufs c, int nimp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nimp) return;
  const int2 pr = pairs[key_pair(c.impacts[k])];
  if (b.status[pr.x] == STATUS_DYNAMIC) {
    b.pos0[pr.x] = b.pos(pr.x);
    b.rot0[pr.x] = b.rot(pr.x);
  }
  if (b.status[pr.y] == STATUS_DYNAMIC) {
    b.pos0[pr.y] = b.pos(pr.y);
    b.rot0[pr.y] = b.rot(pr.y);
  }
}

// The bodies still frozen when the CCD ends go back to their safe position; their last dynamics
// update (after the substep) used the rotation they have now, kept for the next one.
__global__ void __launch_bounds__(256) k_ccd_clamp(DevBodies b, CcdBufs c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n || b.status[i] != STATUS_DYNAMIC || !(c.frozen[b.ccd_key[i]] >= 0.0f)) return;
  b.ccd_rot[i] = b.rot(i);
  b.w_dyn[i].w = 2.0f;
  b.pos(i) = b.pos0[i];
  if (!(b.ccd[i] & 2)) b.rot(i) = b.rot0[i];
}

int blocks_of(int n, int bs) { return (n + bs - 1) / bs; }

}  // namespace

void launch_ccd_predict(const DevBodies& b, const int2* pairs, int npairs, CcdBufs& c, float end_time, float dt0,
                        float target, int on_penetration, hipStream_t s) {
  hipLaunchKernelGGL(k_ccd_reset, dim3(1), dim3(64), 0, s, c.ctl);
  fill_u32(reinterpret_cast<uint32_t*>(c.frozen), 0xbf800000u, size_t(b.n), s);  // -1
  fill_u32(c.deg, 0u, size_t(b.n), s);
  fill_u32(c.fill, 0u, size_t(b.n), s);
  if (npairs > 0) {
    hipLaunchKernelGGL(k_ccd_first_pass, dim3(blocks_of(npairs, 128)), dim3(128), 0, s, b, pairs, npairs, c,
                       end_time, dt0, target, on_penetration);
    hipLaunchKernelGGL(k_ccd_degree, dim3(blocks_of(npairs, 256)), dim3(256), 0, s, npairs, c);
    scan_exclusive_u32(c.deg, c.off, b.n, c.tmp, s);
    hipLaunchKernelGGL(k_ccd_fill, dim3(blocks_of(npairs, 256)), dim3(256), 0, s, npairs, c);
  }
  if (b.n <= kFzCap)
    hipLaunchKernelGGL(k_ccd_events<true>, dim3(1), dim3(kEventBlock), 0, s, b, c, end_time, dt0, target,
                       on_penetration);
  else
    hipLaunchKernelGGL(k_ccd_events<false>, dim3(1), dim3(kEventBlock), 0, s, b, c, end_time, dt0, target,
                       on_penetration);
}

void launch_ccd_seed(const DevBodies& b, const int2* pairs, CcdBufs& c, int nimp, hipStream_t s) {
  if (nimp <= 0) return;
  hipLaunchKernelGGL(k_ccd_seed, dim3(blocks_of(nimp, 256)), dim3(256), 0, s, b, pairs, c, nimp);
}

void launch_ccd_touch(const DevBodies& b, const int2* pairs, int npairs, CcdBufs& c, hipStream_t s) {
  if (npairs <= 0) return;
  hipLaunchKernelGGL(k_ccd_touch, dim3(blocks_of(npairs, 256)), dim3(256), 0, s, b, pairs, npairs, c);
}

void launch_ccd_prepare(const DevBodies& b, CcdBufs& c, float dt_sub, float last_toi, hipStream_t s) {
  if (b.n <= 0) return;
  hipLaunchKernelGGL(k_ccd_prepare, dim3(blocks_of(b.n, 256)), dim3(256), 0, s, b, c, dt_sub, last_toi);
}

void launch_ccd_compact(const int2* pairs, int npairs, const uint32_t* flag, CcdBufs& c, hipStream_t s) {
  if (npairs <= 0) {
    fill_u32(c.ctl + kCcdList, 0u, 1, s);
    return;
  }
  scan_exclusive_u32(flag, c.cpos, npairs, c.tmp, s);
  hipLaunchKernelGGL(k_ccd_scatter, dim3(blocks_of(npairs, 256)), dim3(256), 0, s, pairs, npairs, flag, c);
}

void launch_ccd_grow(const DevBodies& b, CcdBufs& c, int nlist, const uint32_t* count, hipStream_t s) {
  if (nlist <= 0) return;
  hipLaunchKernelGGL(k_ccd_grow, dim3(blocks_of(nlist, 256)), dim3(256), 0, s, b, c, nlist, count);
}

void launch_ccd_merge(int n, CcdBufs& c, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_ccd_merge, dim3(blocks_of(n, 256)), dim3(256), 0, s, n, c);
}

void launch_ccd_validate(const DevBodies& b, const int2* pairs, CcdBufs& c, int nimp, hipStream_t s) {
  if (nimp <= 0) return;
  hipLaunchKernelGGL(k_ccd_validate, dim3(blocks_of(nimp, 256)), dim3(256), 0, s, b, pairs, c, nimp);
}

void launch_ccd_clamp(const DevBodies& b, CcdBufs& c, hipStream_t s) {
  if (b.n <= 0) return;
  hipLaunchKernelGGL(k_ccd_clamp, dim3(blocks_of(b.n, 256)), dim3(256), 0, s, b, c);
}

}  // namespace nph
