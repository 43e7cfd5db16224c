// nphysics_amd/csrc/k_solver.hip -- contact assembly and the graph-coloured SOR-Prox solve.
//
// Colouring. Manifolds conflict when they share a dynamic body. A deterministic Jones-Plassmann
// greedy colouring runs one kernel per round: every uncoloured manifold whose priority
// (hash(m), m) beats all neighbours that are uncoloured or were coloured in this same round takes
// the smallest colour unused by neighbours coloured in earlier rounds. The state word packs
// (round << 16 | colour), so a neighbour coloured concurrently in the current round is always
// seen as "this round" whichever value the load returns; the result is independent of timing.
// Colours persist across steps: manifolds whose body pair was solved in the previous step keep
// their colour and only new manifolds run the rounds, so at rest colouring costs one pass.
// Manifolds are then sorted by colour (stable radix sort) and row records laid out in that order.
//
// Assembly (k_assemble, one lane per manifold) restates SignoriniCoulombPyramidModel::constraints
// (src/solver/signorini_coulomb_pyramid_model.rs:56-224), SignoriniModel::build_velocity_constraint
// (src/solver/signorini_model.rs:37-138), helper::constraint_pair_geometry
// (src/solver/helper.rs:53-135) and RigidBody::fill_constraint_geometry
// (src/object/rigid_body.rs:672-722): per contact one normal row along -n at
// world1 + n margin1 / world2 - n margin2 and two friction rows along the tangents; material
// Average combine (src/material/material.rs:72-86); warm start from the previous step's impulse
// of the contact with the same (body1, body2, feature) key (the ImpulseCache lookup at
// signorini_coulomb_pyramid_model.rs:104-108).
//
// Solve (k_warmstart / k_pgs, one lane per manifold of one colour) restates SORProx
// (src/solver/sor_prox.rs:19-436): warm start dv += l W J, then per sweep the manifold's friction
// rows (bounded by mu times the current normal impulse, zeroed when it is zero) and then its normal
// rows (projected on l >= 0). Manifolds of one colour share no dynamic body, so one launch per
// colour performs exactly the sequential Gauss-Seidel updates of the colour order.
//
// Row record (16 float4 per contact, colour order): 0 n|mu, 1 t1|im1, 2 t2|im2,
// 3..5 body-1 angular jacobian of rows (n, t1, t2)|r, 6..8 body-2 angular jacobian|rhs,
// 9..11 I1^-1 J1_ang, 12..14 I2^-1 J2_ang, 15 impulses (n, t1, t2). Linear jacobians are the
// row direction d (body 1) and -d (body 2), weighted by the inverse mass.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "stages.hpp"

namespace nph {

namespace {

constexpr uint32_t kUncoloured = 0xffffffffu;

__device__ __forceinline__ uint32_t colour_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__global__ void __launch_bounds__(256) k_degree(DevBodies b, DevContacts c, int M, uint32_t* deg) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int2 bb = c.mbody[m];
  if (b.status[bb.x] == STATUS_DYNAMIC) atomicAdd(&deg[bb.x], 1u);
  if (b.status[bb.y] == STATUS_DYNAMIC) atomicAdd(&deg[bb.y], 1u);
}

// Colours persist across steps: a manifold whose body pair was a manifold of the previous solve
// keeps that colour (marked as coloured in round 0); the others start uncoloured.
__global__ void __launch_bounds__(256) k_inherit(DevContacts c, int M, DevContacts prev, int prevM,
                                                const uint32_t* __restrict__ prev_state, uint32_t* state) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int2 key = c.mbody[m];
  int lo = 0, hi = prevM;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int2 k = prev.mbody[mid];
    if (k.x < key.x || (k.x == key.x && k.y < key.y)) lo = mid + 1;
    else hi = mid;
  }
  uint32_t st = kUncoloured;
  if (lo < prevM) {
    const int2 k = prev.mbody[lo];
    if (k.x == key.x && k.y == key.y) st = prev_state[lo] & 0xffffu;
  }
  state[m] = st;
}

__global__ void __launch_bounds__(256) k_adjacency(DevBodies b, DevContacts c, int M,
                                                  const uint32_t* __restrict__ off, uint32_t* fill,
                                                  int* adj) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int2 bb = c.mbody[m];
  if (b.status[bb.x] == STATUS_DYNAMIC) adj[off[bb.x] + atomicAdd(&fill[bb.x], 1u)] = m;
  if (b.status[bb.y] == STATUS_DYNAMIC) adj[off[bb.y] + atomicAdd(&fill[bb.y], 1u)] = m;
}

__global__ void __launch_bounds__(256) k_colour_round(DevBodies b, DevContacts c, int M, uint32_t round,
                                                     const uint32_t* __restrict__ off,
                                                     const uint32_t* __restrict__ deg,
                                                     const int* __restrict__ adj, uint32_t* state,
                                                     uint32_t* remaining) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  if (state[m] != kUncoloured) return;
  const uint32_t pm = colour_hash(uint32_t(m));
  const int2 bb = c.mbody[m];
  const int bs[2] = {bb.x, bb.y};
  bool dyn[2];
  dyn[0] = b.status[bs[0]] == STATUS_DYNAMIC;
  dyn[1] = b.status[bs[1]] == STATUS_DYNAMIC;
  for (int s = 0; s < 2; ++s) {
    if (!dyn[s]) continue;
    const uint32_t o0 = off[bs[s]], o1 = o0 + deg[bs[s]];
    for (uint32_t k = o0; k < o1; ++k) {
      const int o = adj[k];
      if (o == m) continue;
      const uint32_t so = state[o];
      if (so == kUncoloured || (so >> 16) == round) {
        const uint32_t po = colour_hash(uint32_t(o));
        if (po > pm || (po == pm && o > m)) {
          *remaining = 1u;  // any uncoloured manifold left; a plain store, every writer stores 1
          return;
        }
      }
    }
  }
  uint32_t colour = 0;
  for (uint32_t base = 0;; base += 64) {
    uint64_t mask = 0;
    for (int s = 0; s < 2; ++s) {
      if (!dyn[s]) continue;
      const uint32_t o0 = off[bs[s]], o1 = o0 + deg[bs[s]];
      for (uint32_t k = o0; k < o1; ++k) {
        const int o = adj[k];
        if (o == m) continue;
        const uint32_t so = state[o];
        if (so == kUncoloured || (so >> 16) == round) continue;
        const uint32_t cc = so & 0xffffu;
        if (cc >= base && cc < base + 64) mask |= 1ull << (cc - base);
      }
    }
    if (mask != ~0ull) {
      colour = base + uint32_t(__ffsll((long long)(~mask))) - 1u;
      break;
    }
  }
  state[m] = (round << 16) | colour;
}

__global__ void __launch_bounds__(256) k_colour_keys(int M, const uint32_t* __restrict__ state,
                                                    uint32_t* keys, uint32_t* order, uint32_t* ncol) {
  __shared__ uint32_t s_max[256];
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t c1 = 0;
  if (m < M) {
    const uint32_t c = state[m] & 0xffffu;
    keys[m] = c;
    order[m] = uint32_t(m);
    c1 = c + 1u;
  }
  s_max[threadIdx.x] = c1;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) s_max[threadIdx.x] = s_max[threadIdx.x] > s_max[threadIdx.x + off] ? s_max[threadIdx.x] : s_max[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(ncol, s_max[0]);
}

__global__ void __launch_bounds__(256) k_colour_starts(DevContacts c, int M,
                                                      const uint32_t* __restrict__ keys,
                                                      const uint32_t* __restrict__ order,
                                                      uint32_t* col_start, uint32_t* mcount_ord) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= M) return;
  const uint32_t k = keys[s];
  if (s == 0 || keys[s - 1] != k) col_start[k] = uint32_t(s);
  if (s == M - 1) col_start[k + 1] = uint32_t(M);
  mcount_ord[s] = uint32_t(c.mcount[order[s]]);
}

// One fill_constraint_geometry call: body-side jacobian (dir, p x dir), weighted angular part,
// inv_r and rhs accumulation with J.v then J.ext.
struct BodyView {
  f3 com, v, w, el, ea;
  m33 inv;
  float im;
};

__device__ __forceinline__ void fill_geom(const BodyView& bv, f3 point, f3 dir, f3& ang, f3& wang,
                                          float& inv_r, float& rhs) {
  const f3 p = sub(point, bv.com);
  ang = cross(p, dir);
  wang = mulv(bv.inv, ang);
  inv_r = inv_r + (bv.im + dot(ang, wang));
  const float J[6] = {dir.x, dir.y, dir.z, ang.x, ang.y, ang.z};
  const float V[6] = {bv.v.x, bv.v.y, bv.v.z, bv.w.x, bv.w.y, bv.w.z};
  const float E[6] = {bv.el.x, bv.el.y, bv.el.z, bv.ea.x, bv.ea.y, bv.ea.z};
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 6; ++k) s = s + J[k] * V[k];
  rhs = rhs + s;
  float e = 0.0f;
#pragma unroll
  for (int k = 0; k < 6; ++k) e = e + J[k] * E[k];
  rhs = rhs + e;
}

__device__ __forceinline__ BodyView load_body(const DevBodies& b, int i) {
  BodyView v;
  v.com = xyz(b.pos[i]);
  v.v = xyz(b.linvel[i]);
  v.w = xyz(b.angvel[i]);
  v.el = xyz(b.ext_lin[i]);
  v.ea = xyz(b.ext_ang[i]);
  v.inv = load_inv(b, i, v.im);
  return v;
}

__device__ __forceinline__ void tangents(f3 n, f3& t1, f3& t2) {
  f3 a = fabsf(n.x) > fabsf(n.y) ? f3{n.z, 0.0f, -n.x} : f3{0.0f, -n.z, n.y};
  const float l = sqrtf(dot(a, a));
  a = f3{a.x / l, a.y / l, a.z / l};
  t1 = cross(a, n);
  t2 = a;
}

__global__ void __launch_bounds__(256) k_assemble(DevBodies b, DevContacts cur, DevContacts prev, int prevM,
                                                 int M, SimParams p, SolverBufs sb, int C) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= M) return;
  const int m = int(sb.order[s]);
  const int2 bb = cur.mbody[m];
  const int b1 = bb.x, b2 = bb.y;
  const bool d1 = b.status[b1] == STATUS_DYNAMIC;
  const bool d2 = b.status[b2] == STATUS_DYNAMIC;
  const float mu = (b.friction[b1] + b.friction[b2]) * 0.5f;
  const float e = (b.restitution[b1] + b.restitution[b2]) * 0.5f;
  const float m1 = b.margin[b1], m2 = b.margin[b2];
  BodyView v1, v2;
  if (d1) v1 = load_body(b, b1);
  if (d2) v2 = load_body(b, b2);
  const float im1 = d1 ? v1.im : 0.0f;
  const float im2 = d2 ? v2.im : 0.0f;
  // previous manifold with the same bodies (lexicographic binary search)
  int pf = 0, pc = 0;
  {
    int lo = 0, hi = prevM;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      const int2 k = prev.mbody[mid];
      if (k.x < b1 || (k.x == b1 && k.y < b2)) lo = mid + 1;
      else hi = mid;
    }
    if (lo < prevM) {
      const int2 k = prev.mbody[lo];
      if (k.x == b1 && k.y == b2) {
        pf = prev.mfirst[lo];
        pc = prev.mcount[lo];
      }
    }
  }
  const int first = cur.mfirst[m], cnt = cur.mcount[m];
  const uint32_t r0 = sb.soff[s];
  for (int k = 0; k < cnt; ++k) {
    const int c = first + k;
    const uint32_t r = r0 + k;
    const float4 w14 = cur.w1[c];
    const f3 n = xyz(cur.n[c]);
    const f3 w1 = xyz(w14), w2 = xyz(cur.w2[c]);
    const float depth = w14.w;
    const uint32_t feat = cur.feature[c];
    float4 imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int kk = 0; kk < pc; ++kk)
      if (prev.feature[pf + kk] == feat) {
        imp = prev.imp[pf + kk];
        break;
      }
    const f3 center1 = add(w1, scale(n, m1));
    const f3 center2 = sub(w2, scale(n, m2));
    f3 t1, t2;
    tangents(n, t1, t2);
    float4 rA1[3], rA2[3], rW1[3], rW2[3];
    float lam[3];
#pragma unroll
    for (int row = 0; row < 3; ++row) {
      const f3 dir = row == 0 ? neg(n) : (row == 1 ? t1 : t2);
      float inv_r = 0.0f, rhs = 0.0f;
      f3 a1{0.0f, 0.0f, 0.0f}, wa1{0.0f, 0.0f, 0.0f}, a2{0.0f, 0.0f, 0.0f}, wa2{0.0f, 0.0f, 0.0f};
      if (d1) fill_geom(v1, center1, dir, a1, wa1, inv_r, rhs);
      if (d2) fill_geom(v2, center2, neg(dir), a2, wa2, inv_r, rhs);
      const float rr = inv_r != 0.0f ? 1.0f / inv_r : 1.0f;
      if (row == 0) {
        if (rhs <= -p.restitution_threshold) rhs = rhs + e * rhs;
        const float de = (depth + m1) + m2;
        if (de < 0.0f) rhs = rhs + (-de) * p.inv_dt;
      }
      rA1[row] = f4(a1, rr);
      rA2[row] = f4(a2, rhs);
      rW1[row] = f4(wa1, 0.0f);
      rW2[row] = f4(wa2, 0.0f);
    }
    lam[0] = imp.x * p.warmstart_coeff;
    lam[1] = imp.y * p.warmstart_coeff;
    lam[2] = imp.z * p.warmstart_coeff;
    float4* R = sb.rec;
    R[0 * long(C) + r] = f4(n, mu);
    R[1 * long(C) + r] = f4(t1, im1);
    R[2 * long(C) + r] = f4(t2, im2);
#pragma unroll
    for (int row = 0; row < 3; ++row) {
      R[(3 + row) * long(C) + r] = rA1[row];
      R[(6 + row) * long(C) + r] = rA2[row];
      R[(9 + row) * long(C) + r] = rW1[row];
      R[(12 + row) * long(C) + r] = rW2[row];
    }
    R[15 * long(C) + r] = make_float4(lam[0], lam[1], lam[2], 0.0f);
    sb.rec2c[r] = uint32_t(c);
  }
}

// Per-lane solver context for one manifold.
struct Dv {
  f3 l, a;
};

__device__ __forceinline__ float dot6(f3 dl, f3 ang, const Dv& v) {
  float s = 0.0f;
  s = s + dl.x * v.l.x;
  s = s + dl.y * v.l.y;
  s = s + dl.z * v.l.z;
  s = s + ang.x * v.a.x;
  s = s + ang.y * v.a.y;
  s = s + ang.z * v.a.z;
  return s;
}

__device__ __forceinline__ void axpy6(float alpha, f3 wl, f3 wa, Dv& v) {
  v.l.x = v.l.x + alpha * wl.x;
  v.l.y = v.l.y + alpha * wl.y;
  v.l.z = v.l.z + alpha * wl.z;
  v.a.x = v.a.x + alpha * wa.x;
  v.a.y = v.a.y + alpha * wa.y;
  v.a.z = v.a.z + alpha * wa.z;
}

struct RowView {
  f3 d;            // body-1 linear direction
  f3 a1, a2;       // angular jacobians
  f3 wa1, wa2;     // weighted angular jacobians
  float im1, im2, r, rhs;
};

__device__ __forceinline__ RowView load_row(const float4* __restrict__ R, long C, uint32_t r, int row) {
  RowView v;
  const float4 n4 = R[0 * C + r];
  const float4 t14 = R[1 * C + r];
  const float4 t24 = R[2 * C + r];
  v.im1 = t14.w;
  v.im2 = t24.w;
  v.d = row == 0 ? neg(xyz(n4)) : (row == 1 ? xyz(t14) : xyz(t24));
  const float4 a1 = R[(3 + row) * C + r];
  const float4 a2 = R[(6 + row) * C + r];
  v.a1 = xyz(a1);
  v.r = a1.w;
  v.a2 = xyz(a2);
  v.rhs = a2.w;
  v.wa1 = xyz(R[(9 + row) * C + r]);
  v.wa2 = xyz(R[(12 + row) * C + r]);
  return v;
}

__device__ __forceinline__ float row_w(const RowView& v, bool d1, bool d2, const Dv& x1, const Dv& x2) {
  if (d1 && d2) return (dot6(v.d, v.a1, x1) + dot6(neg(v.d), v.a2, x2)) + v.rhs;
  if (d1) return dot6(v.d, v.a1, x1) + v.rhs;
  return dot6(neg(v.d), v.a2, x2) + v.rhs;
}

__device__ __forceinline__ void row_add(const RowView& v, float alpha, bool d1, bool d2, Dv& x1, Dv& x2) {
  if (d1) axpy6(alpha, scale(v.d, v.im1), v.wa1, x1);
  if (d2) axpy6(alpha, scale(neg(v.d), v.im2), v.wa2, x2);
}

struct SlotCtx {
  int b1, b2, cnt;
  bool d1, d2;
  uint32_t r0;
};

__device__ __forceinline__ SlotCtx slot_ctx(const DevBodies& b, const DevContacts& c, const SolverBufs& sb,
                                            int s) {
  SlotCtx x;
  const int m = int(sb.order[s]);
  const int2 bb = c.mbody[m];
  x.b1 = bb.x;
  x.b2 = bb.y;
  x.d1 = b.status[bb.x] == STATUS_DYNAMIC;
  x.d2 = b.status[bb.y] == STATUS_DYNAMIC;
  x.cnt = c.mcount[m];
  x.r0 = sb.soff[s];
  return x;
}

__global__ void __launch_bounds__(256) k_warmstart(DevBodies b, DevContacts c, int first_slot, int nslots,
                                                  SolverBufs sb, int C) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  const SlotCtx x = slot_ctx(b, c, sb, first_slot + t);
  Dv v1{}, v2{};
  if (x.d1) { v1.l = xyz(b.dv_lin[x.b1]); v1.a = xyz(b.dv_ang[x.b1]); }
  if (x.d2) { v2.l = xyz(b.dv_lin[x.b2]); v2.a = xyz(b.dv_ang[x.b2]); }
  const float4* R = sb.rec;
  for (int k = 0; k < x.cnt; ++k) {
    const uint32_t r = x.r0 + k;
    const float ln = R[15 * long(C) + r].x;
    if (ln != 0.0f) row_add(load_row(R, C, r, 0), ln, x.d1, x.d2, v1, v2);
  }
  for (int k = 0; k < x.cnt; ++k) {
    const uint32_t r = x.r0 + k;
    const float4 lam = R[15 * long(C) + r];
    if (lam.y != 0.0f) row_add(load_row(R, C, r, 1), lam.y, x.d1, x.d2, v1, v2);
    if (lam.z != 0.0f) row_add(load_row(R, C, r, 2), lam.z, x.d1, x.d2, v1, v2);
  }
  if (x.d1) { b.dv_lin[x.b1] = f4(v1.l, 0.0f); b.dv_ang[x.b1] = f4(v1.a, 0.0f); }
  if (x.d2) { b.dv_lin[x.b2] = f4(v2.l, 0.0f); b.dv_ang[x.b2] = f4(v2.a, 0.0f); }
}

__global__ void __launch_bounds__(256) k_pgs(DevBodies b, DevContacts c, int first_slot, int nslots,
                                            SolverBufs sb, int C) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  const SlotCtx x = slot_ctx(b, c, sb, first_slot + t);
  Dv v1{}, v2{};
  if (x.d1) { v1.l = xyz(b.dv_lin[x.b1]); v1.a = xyz(b.dv_ang[x.b1]); }
  if (x.d2) { v2.l = xyz(b.dv_lin[x.b2]); v2.a = xyz(b.dv_ang[x.b2]); }
  float4* R = sb.rec;
  const float4* Rc = sb.rec;
  // friction rows of every contact, bounded by the current normal impulse
  for (int k = 0; k < x.cnt; ++k) {
    const uint32_t r = x.r0 + k;
    float4 lam = R[15 * long(C) + r];
    const float mu = R[0 * long(C) + r].w;
    const float ln = lam.x;
#pragma unroll
    for (int f = 1; f <= 2; ++f) {
      float li = f == 1 ? lam.y : lam.z;
      const RowView v = load_row(Rc, C, r, f);
      if (ln == 0.0f) {
        if (li != 0.0f) {
          row_add(v, -li, x.d1, x.d2, v1, v2);
          li = 0.0f;
        }
      } else {
        const float mx = mu * ln;
        const float mn = -mx;
        const float w = row_w(v, x.d1, x.d2, v1, v2);
        const float val = li - v.r * w;
        const float ni = val > mn ? (val < mx ? val : mx) : mn;
        const float dl = ni - li;
        li = ni;
        row_add(v, dl, x.d1, x.d2, v1, v2);
      }
      if (f == 1) lam.y = li;
      else lam.z = li;
    }
    R[15 * long(C) + r] = lam;
  }
  // normal rows
  for (int k = 0; k < x.cnt; ++k) {
    const uint32_t r = x.r0 + k;
    float4 lam = R[15 * long(C) + r];
    const RowView v = load_row(Rc, C, r, 0);
    const float w = row_w(v, x.d1, x.d2, v1, v2);
    const float a = lam.x - v.r * w;
    const float ni = a > 0.0f ? a : 0.0f;
    const float dl = ni - lam.x;
    lam.x = ni;
    row_add(v, dl, x.d1, x.d2, v1, v2);
    R[15 * long(C) + r] = lam;
  }
  if (x.d1) { b.dv_lin[x.b1] = f4(v1.l, 0.0f); b.dv_ang[x.b1] = f4(v1.a, 0.0f); }
  if (x.d2) { b.dv_lin[x.b2] = f4(v2.l, 0.0f); b.dv_ang[x.b2] = f4(v2.a, 0.0f); }
}

// Natural-map residual per row; per block (max |phi|, sum phi^2) in double, reduced on the host
// in block order.
__global__ void __launch_bounds__(256) k_residual(DevBodies b, DevContacts c, int M, SolverBufs sb, int C) {
  __shared__ double s_max[256];
  __shared__ double s_sum[256];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double mx = 0.0, sum = 0.0;
  if (s < M) {
    const SlotCtx x = slot_ctx(b, c, sb, s);
    Dv v1{}, v2{};
    if (x.d1) { v1.l = xyz(b.dv_lin[x.b1]); v1.a = xyz(b.dv_ang[x.b1]); }
    if (x.d2) { v2.l = xyz(b.dv_lin[x.b2]); v2.a = xyz(b.dv_ang[x.b2]); }
    const float4* R = sb.rec;
    for (int k = 0; k < x.cnt; ++k) {
      const uint32_t r = x.r0 + k;
      const float4 lam = R[15 * long(C) + r];
      const float mu = R[0 * long(C) + r].w;
      {
        const RowView v = load_row(R, C, r, 0);
        const float w = row_w(v, x.d1, x.d2, v1, v2);
        const float a = lam.x - v.r * w;
        const float pr = a > 0.0f ? a : 0.0f;
        const double phi = fabs(double((lam.x - pr) / v.r));
        mx = phi > mx ? phi : mx;
        sum += phi * phi;
      }
#pragma unroll
      for (int f = 1; f <= 2; ++f) {
        const float li = f == 1 ? lam.y : lam.z;
        const RowView v = load_row(R, C, r, f);
        const float w = row_w(v, x.d1, x.d2, v1, v2);
        const float mxv = mu * lam.x, mnv = -mxv;
        const float val = li - v.r * w;
        const float pr = val > mnv ? (val < mxv ? val : mxv) : mnv;
        const double phi = fabs(double((li - pr) / v.r));
        mx = phi > mx ? phi : mx;
        sum += phi * phi;
      }
    }
  }
  s_max[threadIdx.x] = mx;
  s_sum[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      s_max[threadIdx.x] = s_max[threadIdx.x] > s_max[threadIdx.x + off] ? s_max[threadIdx.x] : s_max[threadIdx.x + off];
      s_sum[threadIdx.x] += s_sum[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sb.red[2 * blockIdx.x] = s_max[0];
    sb.red[2 * blockIdx.x + 1] = s_sum[0];
  }
}

__global__ void __launch_bounds__(256) k_cache(DevContacts c, SolverBufs sb, int C) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= C) return;
  c.imp[sb.rec2c[r]] = sb.rec[15 * long(C) + r];
}

inline unsigned nb(long n) { return unsigned((n + 255) / 256); }

}  // namespace

void launch_colour_prepare(const DevBodies& b, const DevContacts& c, int M, const DevContacts& prev,
                           int prevM, const uint32_t* prev_state, SolverBufs& sb, hipStream_t s) {
  fill_u32(sb.deg, 0u, size_t(b.n), s);
  fill_u32(sb.adj_fill, 0u, size_t(b.n), s);
  if (M == 0) return;
  hipLaunchKernelGGL(k_inherit, dim3(nb(M)), dim3(256), 0, s, c, M, prev, prevM, prev_state, sb.state);
  hipLaunchKernelGGL(k_degree, dim3(nb(M)), dim3(256), 0, s, b, c, M, sb.deg);
  scan_exclusive_u32(sb.deg, sb.adj_off, b.n, sb.tmp, s);
  hipLaunchKernelGGL(k_adjacency, dim3(nb(M)), dim3(256), 0, s, b, c, M, sb.adj_off, sb.adj_fill, sb.adj);
}

void launch_colour_round(const DevBodies& b, const DevContacts& c, int M, int round, SolverBufs& sb,
                         hipStream_t s) {
  if (M == 0) return;
  fill_u32(sb.remaining, 0u, 1, s);
  hipLaunchKernelGGL(k_colour_round, dim3(nb(M)), dim3(256), 0, s, b, c, M, uint32_t(round), sb.adj_off,
                     sb.deg, sb.adj, sb.state, sb.remaining);
}

void launch_colour_keys(int M, SolverBufs& sb, uint32_t* ncol, hipStream_t s) {
  fill_u32(ncol, 0u, 1, s);
  if (M == 0) return;
  hipLaunchKernelGGL(k_colour_keys, dim3(nb(M)), dim3(256), 0, s, M, sb.state, sb.ckeys, sb.order, ncol);
}

void launch_colour_order(const DevContacts& c, int M, int ncolours, SolverBufs& sb, hipStream_t s) {
  if (M == 0) return;
  const int bits = ncolours <= 256 ? 8 : 16;
  radix_sort_pairs(sb.ckeys, sb.order, sb.ckeys_alt, sb.order_alt, M, bits, sb.tmp, s);
  fill_u32(sb.col_start, 0xffffffffu, size_t(ncolours) + 1, s);
  hipLaunchKernelGGL(k_colour_starts, dim3(nb(M)), dim3(256), 0, s, c, M, sb.ckeys, sb.order, sb.col_start,
                     sb.mcount_ord);
  scan_exclusive_u32(sb.mcount_ord, sb.soff, M, sb.tmp, s);
}

void launch_assemble(const DevBodies& b, const DevContacts& cur, const DevContacts& prev, int prevM,
                     int M, const SimParams& p, SolverBufs& sb, int C, hipStream_t s) {
  if (M == 0) return;
  hipLaunchKernelGGL(k_assemble, dim3(nb(M)), dim3(256), 0, s, b, cur, prev, prevM, M, p, sb, C);
}

void launch_warmstart_colour(const DevBodies& b, const DevContacts& c, int first_slot, int nslots,
                             SolverBufs& sb, int C, hipStream_t s) {
  if (nslots <= 0) return;
  hipLaunchKernelGGL(k_warmstart, dim3(nb(nslots)), dim3(256), 0, s, b, c, first_slot, nslots, sb, C);
}

void launch_pgs_colour(const DevBodies& b, const DevContacts& c, int first_slot, int nslots,
                       SolverBufs& sb, int C, hipStream_t s) {
  if (nslots <= 0) return;
  hipLaunchKernelGGL(k_pgs, dim3(nb(nslots)), dim3(256), 0, s, b, c, first_slot, nslots, sb, C);
}

int residual_blocks(int M) { return int(nb(M > 0 ? M : 1)); }

void launch_residual(const DevBodies& b, const DevContacts& c, int M, SolverBufs& sb, int C,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_residual, dim3(nb(M > 0 ? M : 1)), dim3(256), 0, s, b, c, M, sb, C);
}

void launch_cache_impulses(const DevContacts& c, SolverBufs& sb, int C, hipStream_t s) {
  if (C == 0) return;
  hipLaunchKernelGGL(k_cache, dim3(nb(C)), dim3(256), 0, s, c, sb, C);
}

}  // namespace nph
