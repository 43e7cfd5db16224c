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
// Solve (k_warmstart / k_pgs, all of the manifold's data loaded into registers before its serial
// chain) restates SORProx (src/solver/sor_prox.rs:19-436): warm start dv += l W J, then per sweep
// and contact the normal row (projected on l >= 0) and the two friction rows (bounded by mu times
// the normal impulse just computed, zeroed when it is zero). Manifolds of one colour share no
// dynamic body, so one launch per colour performs exactly the sequential Gauss-Seidel updates of
// the colour order. k_pgs gives each manifold two lanes, one per body: each lane forms its body's
// half of a row's J dv and its body's dv update, and the halves meet by a lane exchange, so the
// serial chain per row is half as long and twice as many lanes are in flight.
//
// Row record (4 contiguous float4 per contact, colour order): 0 n|mu, 1 lever arm p1 of body 1
// (contact point minus centre of mass, zero for a static body)|r_n, 2 p2|r_t1,
// 3 rhs (n, t1, t2)|r_t2. The impulses live in their own array, one float4 per record (n, t1, t2
// | 1 if the warm start came from the cache, else 0), so a sweep's impulse writes are contiguous
// 16-byte runs rather than one dirty line per contact.
// Linear jacobians are the row direction d (body 1) and -d (body 2) weighted by the inverse
// mass; the tangents, the angular jacobians p1 x d and p2 x (-d) and I^-1 J_ang are recomputed
// in the sweep kernels with the assembly's own expressions, so they carry the same FP32 values.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

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
  const int j = find_manifold(prev, prevM, key.x, key.y);
  state[m] = j >= 0 ? (prev_state[j] & 0xffffu) : kUncoloured;
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
                                                     uint32_t* remaining, uint32_t* last_round) {
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
  atomicMax(last_round, round);
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

// Per slot (colour order): bodies with their dynamic flag in bit 30, first row record, contact
// count and first contact index, packed into one int4.
__global__ void __launch_bounds__(256) k_slot_info(DevBodies b, DevContacts c, int M,
                                                  const uint32_t* __restrict__ order,
                                                  const uint32_t* __restrict__ soff, int4* sinfo,
                                                  uint32_t* slot_of) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= M) return;
  const int m = int(order[s]);
  slot_of[m] = uint32_t(s);
  const int2 bb = c.mbody[m];
  const int d1 = b.status[bb.x] == STATUS_DYNAMIC ? 1 : 0;
  const int d2 = b.status[bb.y] == STATUS_DYNAMIC ? 1 : 0;
  sinfo[s] = make_int4(bb.x | (d1 << 30), bb.y | (d2 << 30), int(soff[s]), c.mcount[m] | (c.mfirst[m] << 3));
}

// One fill_constraint_geometry call: body-side jacobian (dir, p x dir), weighted angular part,
// inv_r and rhs accumulation with J.v then J.ext.
struct BodyView {
  f3 com, v, w, el, ea;
  f3 ml, ma;  // jacobian mask
  m33 inv;
  float im;
};

// The jacobian [dir; p x dir] is masked; the velocity term uses the unmasked force, so kinematic
// dofs stay controllable at the velocity level, and the external term the masked one
// (rigid_body.rs:685-718).
__device__ __forceinline__ void fill_geom(const BodyView& bv, f3 point, f3 dir, f3& ang, f3& wang,
                                          float& inv_r, float& rhs) {
  const f3 p = sub(point, bv.com);
  const f3 ang_u = cross(p, dir);
  const f3 lin = cmul(dir, bv.ml);
  ang = cmul(ang_u, bv.ma);
  wang = mulv(bv.inv, ang);
  inv_r = inv_r + (bv.im + dot(ang, wang));
  const float F[6] = {dir.x, dir.y, dir.z, ang_u.x, ang_u.y, ang_u.z};
  const float J[6] = {lin.x, lin.y, lin.z, ang.x, ang.y, ang.z};
  const float V[6] = {bv.v.x, bv.v.y, bv.v.z, bv.w.x, bv.w.y, bv.w.z};
  const float E[6] = {bv.el.x, bv.el.y, bv.el.z, bv.ea.x, bv.ea.y, bv.ea.z};
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 6; ++k) s = s + F[k] * V[k];
  rhs = rhs + s;
  float e = 0.0f;
#pragma unroll
  for (int k = 0; k < 6; ++k) e = e + J[k] * E[k];
  rhs = rhs + e;
}

__device__ __forceinline__ BodyView load_body(const DevBodies& b, int i) {
  BodyView v;
  v.com = xyz(b.pos(i));
  v.v = xyz(b.linvel[i]);
  v.w = xyz(b.angvel[i]);
  v.el = xyz(b.ext_lin[i]);
  v.ea = xyz(b.ext_ang[i]);
  v.inv = load_inv(b, i, v.im);
  b.mask(i, v.ml, v.ma);
  return v;
}

__device__ __forceinline__ void tangents(f3 n, f3& t1, f3& t2) {
  f3 a = fabsf(n.x) > fabsf(n.y) ? f3{n.z, 0.0f, -n.x} : f3{0.0f, -n.z, n.y};
  const float l = sqrtf(dot(a, a));
  a = f3{a.x / l, a.y / l, a.z / l};
  t1 = cross(a, n);
  t2 = a;
}

// Material::combine (material.rs:134-178): the table's coefficient when both materials have an
// id and the table holds the pair, else MaterialCombineMode::combine of the two.
__device__ __forceinline__ void material_coefficients(const DevBodies& b, const MatTable& mt, int b1, int b2,
                                                      float& mu, float& e) {
  const int f1 = b.flags ? b.flags[b1] : 0, f2 = b.flags ? b.flags[b2] : 0;
  mu = combine_coeff(b.friction[b1], fric_mode(f1), b.friction[b2], fric_mode(f2));
  e = combine_coeff(b.restitution[b1], rest_mode(f1), b.restitution[b2], rest_mode(f2));
  if (!b.mat_id) return;
  const int i1 = b.mat_id[b1], i2 = b.mat_id[b2];
  if (i1 < 0 || i2 < 0) return;
  const unsigned long long lo = unsigned(i1 < i2 ? i1 : i2), hi = unsigned(i1 < i2 ? i2 : i1);
  const unsigned long long key = (lo << 32) | hi;
  float v;
  if (mat_lookup(mt.fkey, mt.fval, mt.nf, key, v)) mu = v;
  if (mat_lookup(mt.rkey, mt.rval, mt.nr, key, v)) e = v;
}

__global__ void __launch_bounds__(256) k_assemble(DevBodies b, DevContacts cur, DevContacts prev, int prevM,
                                                 int M, SimParams p, SolverBufs sb, MatTable mt) {
  // manifold order (neighbouring lanes read neighbouring bodies and contacts); the records go to
  // the manifold's slot in colour order
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int s = int(sb.slot_of[m]);
  const int2 bb = cur.mbody[m];
  const int b1 = bb.x, b2 = bb.y;
  const bool d1 = b.status[b1] == STATUS_DYNAMIC;
  const bool d2 = b.status[b2] == STATUS_DYNAMIC;
  float mu, e;
  material_coefficients(b, mt, b1, b2, mu, e);
  // surface_velocity = position1 * v1 - position2 * v2 (material.rs:175, basic_material.rs:47-56)
  f3 sv{0.0f, 0.0f, 0.0f};
  if (b.surf_vel) sv = sub(qrot(toq(b.rot(b1)), xyz(b.surf_vel[b1])), qrot(toq(b.rot(b2)), xyz(b.surf_vel[b2])));
  const float m1 = b.margin[b1], m2 = b.margin[b2];
  BodyView v1, v2;
  if (d1) v1 = load_body(b, b1);
  if (d2) v2 = load_body(b, b2);
  // previous manifold with the same bodies
  int pf = 0, pc = 0;
  {
    const int j = find_manifold(prev, prevM, b1, b2);
    if (j >= 0) {
      pf = prev.mfirst[j];
      pc = prev.mcount[j];
    }
  }
  const int first = cur.mfirst[m], cnt = cur.mcount[m];
  const uint32_t r0 = sb.soff[s];
  // the position solver's inputs in slot / record order (see k_position.hip)
  const int k1 = b.shape[b1], k2 = b.shape[b2];
  sb.pgeo[2 * long(s)] = f4(xyz(b.dims[b1]), m1);
  sb.pgeo[2 * long(s) + 1] = f4(xyz(b.dims[b2]), m2);
  for (int k = 0; k < cnt; ++k) {
    const int c = first + k;
    const uint32_t r = r0 + k;
    const float4* K = cur.contact(c);
    const float4 w14 = K[0], w24 = K[1], n4 = K[2];
    float4* P = sb.prec + 3 * long(r);
    P[0] = f4(xyz(K[3]), n4.w);  // the kinematic kind's bits
    P[1] = f4(xyz(K[4]), __int_as_float(k1));
    P[2] = f4(xyz(K[5]), __int_as_float(k2));
    const f3 n = xyz(n4);
    const f3 w1 = xyz(w14), w2 = xyz(w24);
    const float depth = w14.w;
    const uint32_t feat = __float_as_uint(w24.w);
    float4 imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float has = 0.0f;
    for (int kk = 0; kk < pc; ++kk)
      if (prev.feature(pf + kk) == feat) {
        const float4 q = prev.imp[pf + kk];
        if (q.w >= 0.0f) {  // w < 0: a sleeping contact that had no cached impulse
          imp = q;
          has = 1.0f;
        }
        break;
      }
    const f3 center1 = add(w1, scale(n, m1));
    const f3 center2 = sub(w2, scale(n, m2));
    f3 t1, t2;
    tangents(n, t1, t2);
    float rr3[3], rhs3[3];
    float lam[3];
#pragma unroll
    for (int row = 0; row < 3; ++row) {
      const f3 dir = row == 0 ? neg(n) : (row == 1 ? t1 : t2);
      // n . surface_velocity for the normal row, t . surface_velocity for the friction rows
      float inv_r = 0.0f, rhs = b.surf_vel ? dot(row == 0 ? n : dir, sv) : 0.0f;
      f3 a1{0.0f, 0.0f, 0.0f}, wa1{0.0f, 0.0f, 0.0f}, a2{0.0f, 0.0f, 0.0f}, wa2{0.0f, 0.0f, 0.0f};
      if (d1) fill_geom(v1, center1, dir, a1, wa1, inv_r, rhs);
      if (d2) fill_geom(v2, center2, neg(dir), a2, wa2, inv_r, rhs);
      const float rr = inv_r != 0.0f ? 1.0f / inv_r : 1.0f;
      if (row == 0) {
        if (rhs <= -p.restitution_threshold) rhs = rhs + e * rhs;
        const float de = (depth + m1) + m2;
        if (de < 0.0f) rhs = rhs + (-de) * p.inv_dt;
      }
      rr3[row] = rr;
      rhs3[row] = rhs;
    }
    // lever arms of the contact point from each dynamic body's centre (fill_geom's p); the sweeps
    // rebuild every row's angular jacobian from them with fill_geom's own cross product
    const f3 zero{0.0f, 0.0f, 0.0f};
    const f3 p1 = d1 ? sub(center1, v1.com) : zero;
    const f3 p2 = d2 ? sub(center2, v2.com) : zero;
    lam[0] = imp.x * p.warmstart_coeff;
    lam[1] = imp.y * p.warmstart_coeff;
    lam[2] = imp.z * p.warmstart_coeff;
    float4* R = sb.rec + long(r) * kRecFields;
    R[0] = f4(n, mu);
    R[1] = f4(p1, rr3[0]);
    R[2] = f4(p2, rr3[1]);
    R[3] = make_float4(rhs3[0], rhs3[1], rhs3[2], rr3[2]);
    sb.lam[r] = make_float4(lam[0], lam[1], lam[2], has);
    sb.rec2c[r] = uint32_t(c);
  }
}

// Per-lane solver state of one manifold.
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

struct SlotCtx {
  int b1, b2, cnt, first;
  bool d1, d2;
  uint32_t r0;
};

// One 16-byte load gives everything a lane needs about its manifold (see k_slot_info).
__device__ __forceinline__ SlotCtx slot_ctx(const SolverBufs& sb, int s) {
  const int4 q = sb.sinfo[s];
  SlotCtx x;
  x.b1 = q.x & 0x3fffffff;
  x.b2 = q.y & 0x3fffffff;
  x.d1 = (q.x >> 30) & 1;
  x.d2 = (q.y >> 30) & 1;
  x.r0 = uint32_t(q.z);
  x.cnt = q.w & 7;
  x.first = int(uint32_t(q.w) >> 3);
  return x;
}

// A contact's record in registers. The tangents and the inverse-inertia weighted angular
// jacobians are not stored: they are recomputed with the expressions k_assemble evaluates
// (tangents(n) and mulv(I^-1, J_ang)), so they carry the same FP32 values.
struct CRec {
  f3 n;
  float mu;
  f3 p1, p2;  // lever arms (zero for a static body)
  float r[3], rhs[3];
  float4 lam;
};

// Angular jacobians of one row along d (d for body 1, -d for body 2), as fill_geom forms them.
__device__ __forceinline__ f3 ang1(const CRec& c, f3 d) { return cross(c.p1, d); }
__device__ __forceinline__ f3 ang2(const CRec& c, f3 d) { return cross(c.p2, neg(d)); }

__device__ __forceinline__ void load_rec(const SolverBufs& sb, uint32_t r, CRec& c) {
  const float4* p = sb.rec + long(r) * kRecFields;
  const float4 v0 = p[0], v1 = p[1], v2 = p[2], v3 = p[3];
  c.n = xyz(v0);
  c.mu = v0.w;
  c.p1 = xyz(v1);
  c.p2 = xyz(v2);
  c.r[0] = v1.w;
  c.r[1] = v2.w;
  c.r[2] = v3.w;
  c.rhs[0] = v3.x;
  c.rhs[1] = v3.y;
  c.rhs[2] = v3.z;
  c.lam = sb.lam[r];
}

struct Inv {
  m33 m;
  float im;
};

// Jacobian masks of a manifold's two bodies (only the masked kernel variants carry them).
template <bool kMask>
struct Masks {
  __device__ void load(const DevBodies&, const SlotCtx&) {}
  __device__ f3 l1(f3 d) const { return d; }
  __device__ f3 l2(f3 d) const { return d; }
  __device__ f3 a1(f3 a) const { return a; }
  __device__ f3 a2(f3 a) const { return a; }
};
template <>
struct Masks<true> {
  f3 ml1, ma1, ml2, ma2;
  __device__ void load(const DevBodies& b, const SlotCtx& x) {
    b.mask(x.b1, ml1, ma1);
    b.mask(x.b2, ml2, ma2);
  }
  __device__ f3 l1(f3 d) const { return cmul(d, ml1); }
  __device__ f3 l2(f3 d) const { return cmul(d, ml2); }
  __device__ f3 a1(f3 a) const { return cmul(a, ma1); }
  __device__ f3 a2(f3 a) const { return cmul(a, ma2); }
};

template <bool kMask>
struct Manifold {
  SlotCtx x;
  Inv i1, i2;
  CRec c[4];
  f3 t1[4], t2[4];
  Masks<kMask> mk;
};

// Everything a lane needs for its manifold, loaded up front: the loads are independent of the
// Gauss-Seidel chain, so they are all in flight together.
template <bool kMask>
__device__ __forceinline__ void load_manifold(const DevBodies& b, const SolverBufs& sb, int s, Manifold<kMask>& m) {
  m.x = slot_ctx(sb, s);
  m.i1.im = m.i2.im = 0.0f;
  m.i1.m = m33{};
  m.i2.m = m33{};
  if (m.x.d1) m.i1.m = load_inv(b, m.x.b1, m.i1.im);
  if (m.x.d2) m.i2.m = load_inv(b, m.x.b2, m.i2.im);
  m.mk.load(b, m.x);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < m.x.cnt) {
      load_rec(sb, m.x.r0 + k, m.c[k]);
      tangents(m.c[k].n, m.t1[k], m.t2[k]);
    }
}

// Row `row` (0 normal, 1 and 2 friction) of contact k: w = J dv + rhs and the dv update. Both
// bodies' terms are always evaluated, a static body having zero inverse mass and inertia and a
// zero dv: its terms are then exact zeros, so w and every dv component equal those of the
// one-body expression (up to the sign of a zero), and a wavefront mixing ground and body-body
// manifolds runs one code path. The masked variant multiplies each body's jacobian by its mask,
// as k_assemble's fill_geom does.
struct Row {
  f3 d1, d2, a1, a2, wa1, wa2;
  float r, rhs;
};

template <bool kMask>
__device__ __forceinline__ Row row_of(const Manifold<kMask>& m, int k, int row) {
  Row v;
  const CRec& c = m.c[k];
  const f3 d = row == 0 ? neg(c.n) : (row == 1 ? m.t1[k] : m.t2[k]);
  v.d1 = m.mk.l1(d);
  v.d2 = m.mk.l2(neg(d));
  v.a1 = m.mk.a1(ang1(c, d));
  v.a2 = m.mk.a2(ang2(c, d));
  v.r = c.r[row];
  v.rhs = c.rhs[row];
  v.wa1 = mulv(m.i1.m, v.a1);
  v.wa2 = mulv(m.i2.m, v.a2);
  return v;
}

__device__ __forceinline__ float row_w(const Row& v, const Dv& x1, const Dv& x2) {
  return (dot6(v.d1, v.a1, x1) + dot6(v.d2, v.a2, x2)) + v.rhs;
}

template <bool kMask>
__device__ __forceinline__ void row_add(const Row& v, float alpha, const Manifold<kMask>& m, Dv& x1, Dv& x2) {
  axpy6(alpha, scale(v.d1, m.i1.im), v.wa1, x1);
  axpy6(alpha, scale(v.d2, m.i2.im), v.wa2, x2);
}

template <bool kMask>
__device__ __forceinline__ void load_dv(const DevBodies& b, const Manifold<kMask>& m, Dv& v1, Dv& v2) {
  if (m.x.d1) b.load_dv(m.x.b1, v1.l, v1.a);
  if (m.x.d2) b.load_dv(m.x.b2, v2.l, v2.a);
}

template <bool kMask>
__device__ __forceinline__ void store_dv(const DevBodies& b, const Manifold<kMask>& m, const Dv& v1, const Dv& v2) {
  if (m.x.d1) b.store_dv(m.x.b1, v1.l, v1.a);
  if (m.x.d2) b.store_dv(m.x.b2, v2.l, v2.a);
}

// Slot range of the class visited k-th: one load (the host writes the ranges with the visiting
// order), not a visit -> class start chain.
__device__ __forceinline__ void colour_range(const SolverBufs& sb, int k, int& first, int& n) {
  const int2 r = sb.vrange[k];
  first = r.x;
  n = r.y;
}

// Solver velocity increments of the manifold's bodies in HBM (the per-colour kernels) ...
struct GlobalDv {
  DevBodies b;
  template <bool kMask>
  __device__ void load(const Manifold<kMask>& m, Dv& v1, Dv& v2) const { load_dv(b, m, v1, v2); }
  template <bool kMask>
  __device__ void store(const Manifold<kMask>& m, const Dv& v1, const Dv& v2) const { store_dv(b, m, v1, v2); }
  __device__ void load1(int body, Dv& v) const { b.load_dv(body, v.l, v.a); }
  __device__ void store1(int body, const Dv& v) const { b.store_dv(body, v.l, v.a); }
};

// ... or in LDS, indexed by the body's rank among its scene's dynamic bodies (the scene-blocked
// kernel). The values are the same FP32 numbers either way.
struct LdsDv {
  float4* lin;
  float4* ang;
  const int* local;
  template <bool kMask>
  __device__ void load(const Manifold<kMask>& m, Dv& v1, Dv& v2) const {
    if (m.x.d1) {
      const int i = local[m.x.b1];
      v1.l = xyz(lin[i]);
      v1.a = xyz(ang[i]);
    }
    if (m.x.d2) {
      const int i = local[m.x.b2];
      v2.l = xyz(lin[i]);
      v2.a = xyz(ang[i]);
    }
  }
  template <bool kMask>
  __device__ void store(const Manifold<kMask>& m, const Dv& v1, const Dv& v2) const {
    if (m.x.d1) {
      const int i = local[m.x.b1];
      lin[i] = f4(v1.l, 0.0f);
      ang[i] = f4(v1.a, 0.0f);
    }
    if (m.x.d2) {
      const int i = local[m.x.b2];
      lin[i] = f4(v2.l, 0.0f);
      ang[i] = f4(v2.a, 0.0f);
    }
  }
  __device__ void load1(int body, Dv& v) const {
    const int i = local[body];
    v.l = xyz(lin[i]);
    v.a = xyz(ang[i]);
  }
  __device__ void store1(int body, const Dv& v) const {
    const int i = local[body];
    lin[i] = f4(v.l, 0.0f);
    ang[i] = f4(v.a, 0.0f);
  }
};

// Warm start of one manifold (SORProx::warmstart_set, sor_prox.rs:19-45): dv += l W J for every
// row with l != 0, normal rows first, then friction rows.
template <bool kMask, class A>
__device__ __forceinline__ void warm_manifold(const A& acc, const DevBodies& b, const SolverBufs& sb, int slot) {
  Manifold<kMask> m;
  load_manifold(b, sb, slot, m);
  Dv v1{}, v2{};
  acc.load(m, v1, v2);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k >= m.x.cnt) break;
    if (m.c[k].lam.x != 0.0f) row_add(row_of(m, k, 0), m.c[k].lam.x, m, v1, v2);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k >= m.x.cnt) break;
    if (m.c[k].lam.y != 0.0f) row_add(row_of(m, k, 1), m.c[k].lam.y, m, v1, v2);
    if (m.c[k].lam.z != 0.0f) row_add(row_of(m, k, 2), m.c[k].lam.z, m, v1, v2);
  }
  acc.store(m, v1, v2);
}

// One SOR-Prox step of one manifold (sor_prox.rs:159-343), contact by contact: the normal row,
// projected on l >= 0, then the two friction rows, bounded by mu times that normal impulse and
// zeroed when it is zero.
template <bool kMask, class A>
__device__ __forceinline__ void pgs_manifold(const A& acc, const DevBodies& b, const SolverBufs& sb, int slot) {
  Manifold<kMask> m;
  load_manifold(b, sb, slot, m);
  Dv v1{}, v2{};
  acc.load(m, v1, v2);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k >= m.x.cnt) break;
    {
      const Row v = row_of(m, k, 0);
      const float w = row_w(v, v1, v2);
      const float a = m.c[k].lam.x - v.r * w;
      const float ni = a > 0.0f ? a : 0.0f;
      const float dl = ni - m.c[k].lam.x;
      m.c[k].lam.x = ni;
      row_add(v, dl, m, v1, v2);
    }
    const float mu = m.c[k].mu;
    const float ln = m.c[k].lam.x;
#pragma unroll
    for (int f = 1; f <= 2; ++f) {
      float li = f == 1 ? m.c[k].lam.y : m.c[k].lam.z;
      const Row v = row_of(m, k, f);
      if (ln == 0.0f) {
        if (li != 0.0f) {
          row_add(v, -li, m, v1, v2);
          li = 0.0f;
        }
      } else {
        const float mx = mu * ln;
        const float mn = -mx;
        const float w = row_w(v, v1, v2);
        const float val = li - v.r * w;
        const float ni = val > mn ? (val < mx ? val : mx) : mn;
        const float dl = ni - li;
        li = ni;
        row_add(v, dl, m, v1, v2);
      }
      if (f == 1) m.c[k].lam.y = li;
      else m.c[k].lam.z = li;
    }
  }
  acc.store(m, v1, v2);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < m.x.cnt) sb.lam[m.x.r0 + k] = m.c[k].lam;
}

// One colour of the warm start / of one sweep: one lane per manifold of the colour visited kv-th.
template <bool kMask>
__global__ void __launch_bounds__(64) k_warmstart(DevBodies b, DevContacts c, SolverBufs sb, int first, int n) {
  const GlobalDv acc{b};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    warm_manifold<kMask>(acc, b, sb, first + t);
}

// One SOR-Prox step of one manifold on two lanes (lane h = 0 body 1, h = 1 body 2), the same
// arithmetic as pgs_manifold: each lane evaluates its body's dot6 of the row, the two partial sums
// are exchanged between the lanes (adjacent lanes of one wavefront), both lanes form
// w = (s1 + s2) + rhs in that order and the same new impulse, and each lane updates its own body.
// A contact's row data as one lane of the pair needs it: the normal, mu, its body's lever arm,
// the three r and rhs, the impulses.
struct CHalf {
  f3 n, p;
  float mu;
  float r[3], rhs[3];
  float4 lam;
};

__device__ __forceinline__ void load_half(const SolverBufs& sb, uint32_t r, int h, CHalf& c) {
  const float4* q = sb.rec + long(r) * kRecFields;
  const float4 v0 = q[0], v1 = q[1], v2 = q[2], v3 = q[3];
  c.n = xyz(v0);
  c.mu = v0.w;
  c.p = h ? xyz(v2) : xyz(v1);
  c.r[0] = v1.w;
  c.r[1] = v2.w;
  c.r[2] = v3.w;
  c.rhs[0] = v3.x;
  c.rhs[1] = v3.y;
  c.rhs[2] = v3.z;
  c.lam = sb.lam[r];
}

// One SOR-Prox step of one manifold on two lanes (lane h = 0 body 1, h = 1 body 2), the same
// arithmetic as pgs_manifold: each lane evaluates its body's dot6 of the row, the two partial sums
// are exchanged between the lanes (adjacent lanes of one wavefront), both lanes form
// w = (s1 + s2) + rhs in that order and the same new impulse, and each lane updates its own body.
template <bool kMask, class A>
__device__ __forceinline__ void pgs_manifold_pair(const A& acc, const DevBodies& b, const SolverBufs& sb, int slot,
                                                  int h) {
  const SlotCtx x = slot_ctx(sb, slot);
  const int body = h ? x.b2 : x.b1;
  const bool dyn = h ? x.d2 : x.d1;
  float im = 0.0f;
  m33 I{};
  if (dyn) I = load_inv(b, body, im);
  f3 ml{1.0f, 1.0f, 1.0f}, ma{1.0f, 1.0f, 1.0f};
  if (kMask) b.mask(body, ml, ma);
  Dv v{};
  if (dyn) acc.load1(body, v);
  CHalf c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < x.cnt) load_half(sb, x.r0 + k, h, c[k]);
  // body 1 takes d and p1 x d, body 2 -d and p2 x -d (ang1 / ang2), each masked
  auto half = [&](f3 p, f3 d, f3& d_h, f3& a_h, f3& wa_h) {
    const f3 dd = h ? neg(d) : d;
    const f3 a = cross(p, dd);
    d_h = kMask ? cmul(dd, ml) : dd;
    a_h = kMask ? cmul(a, ma) : a;
    wa_h = mulv(I, a_h);
  };
  auto w_of = [&](const f3& d_h, const f3& a_h, float rhs) {
    const float s_h = dot6(d_h, a_h, v);
    const float s_o = __shfl_xor(s_h, 1);
    const float s1 = h == 0 ? s_h : s_o;
    const float s2 = h == 0 ? s_o : s_h;
    return (s1 + s2) + rhs;
  };
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k >= x.cnt) break;
    {
      f3 d_h, a_h, wa_h;
      half(c[k].p, neg(c[k].n), d_h, a_h, wa_h);
      const float w = w_of(d_h, a_h, c[k].rhs[0]);
      const float a = c[k].lam.x - c[k].r[0] * w;
      const float ni = a > 0.0f ? a : 0.0f;
      const float dl = ni - c[k].lam.x;
      c[k].lam.x = ni;
      axpy6(dl, scale(d_h, im), wa_h, v);
    }
    const float mu = c[k].mu;
    const float ln = c[k].lam.x;
    f3 t[2];
    tangents(c[k].n, t[0], t[1]);
#pragma unroll
    for (int f = 1; f <= 2; ++f) {
      float li = f == 1 ? c[k].lam.y : c[k].lam.z;
      f3 d_h, a_h, wa_h;
      half(c[k].p, t[f - 1], d_h, a_h, wa_h);
      if (ln == 0.0f) {
        if (li != 0.0f) {
          axpy6(-li, scale(d_h, im), wa_h, v);
          li = 0.0f;
        }
      } else {
        const float mx = mu * ln;
        const float mn = -mx;
        const float w = w_of(d_h, a_h, c[k].rhs[f]);
        const float val = li - c[k].r[f] * w;
        const float ni = val > mn ? (val < mx ? val : mx) : mn;
        const float dl = ni - li;
        li = ni;
        axpy6(dl, scale(d_h, im), wa_h, v);
      }
      if (f == 1) c[k].lam.y = li;
      else c[k].lam.z = li;
    }
  }
  if (dyn) acc.store1(body, v);
  if (h == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < x.cnt) sb.lam[x.r0 + k] = c[k].lam;
  }
}

// One colour of one sweep, two lanes per manifold (the grid covers both lanes of every pair in
// one wavefront: blockDim and the stride are even).
template <bool kMask>
__global__ void __launch_bounds__(64) k_pgs(DevBodies b, DevContacts c, SolverBufs sb, int first, int n) {
  const GlobalDv acc{b};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; (t >> 1) < n; t += gridDim.x * blockDim.x)
    pgs_manifold_pair<kMask>(acc, b, sb, first + (t >> 1), t & 1);
}

// A run of consecutive small colour classes of the visiting order in one workgroup, a barrier
// between classes: per body the same updates in the same order as one launch per class, without
// a launch per class. Each class fits one pass of the workgroup.
template <bool kWarm, bool kMask>
__global__ void __launch_bounds__(kRunBlock) k_colour_run(DevBodies b, SolverBufs sb, int kv0, int nk) {
  const GlobalDv acc{b};
  for (int kv = kv0; kv < kv0 + nk; ++kv) {
    int first, n;
    colour_range(sb, kv, first, n);
    for (int t = threadIdx.x; t < n; t += kRunBlock) {
      if (kWarm) warm_manifold<kMask>(acc, b, sb, first + t);
      else pgs_manifold<kMask>(acc, b, sb, first + t);
    }
    __syncthreads();
  }
}

// Scene-blocked sweeps: one workgroup per scene runs the warm start (when warm, one lane per
// manifold) and nsweeps sweeps (two lanes per manifold, as k_pgs) over all colours in the visiting
// order, with the scene's dynamic-body velocity increments resident in LDS and a workgroup barrier
// between colours. Every body sees the same sequence of row updates as under the per-colour
// launches (manifolds of one colour share no dynamic body; colours and sweeps keep their order),
// so the results are identical.
template <int kBlock, int kCap, bool kMask>
__global__ void __launch_bounds__(kBlock) k_sweeps_blocked(DevBodies b, SolverBufs sb, SceneBlocks sc,
                                                         const int* __restrict__ scene_list, int ncolours,
                                                         int nsweeps, int warm) {
  __shared__ float4 s_lin[kCap];
  __shared__ float4 s_ang[kCap];
  const int s = scene_list[blockIdx.x];
  const int d0 = sc.dyn_start[s], nd = sc.dyn_start[s + 1] - d0;
  for (int i = threadIdx.x; i < nd; i += kBlock) {
    const int body = sc.dyn_list[d0 + i];
    f3 l, a;
    b.load_dv(body, l, a);
    s_lin[i] = f4(l, 0.0f);
    s_ang[i] = f4(a, 0.0f);
  }
  __syncthreads();
  const LdsDv acc{s_lin, s_ang, sc.local};
  const uint32_t* st = sc.start + long(s) * (ncolours + 1);
  for (int pass = warm ? -1 : 0; pass < nsweeps; ++pass) {
    for (int k = 0; k < ncolours; ++k) {
      const int col = sb.visit[k];
      const int s0 = int(st[col]), e = int(st[col + 1]);
      if (pass < 0) {
        for (int i = s0 + threadIdx.x; i < e; i += kBlock) warm_manifold<kMask>(acc, b, sb, int(sc.slot[i]));
      } else {
        // both lanes of a manifold are adjacent lanes of one wavefront (kBlock and 2 * s0 are even)
        for (int i = 2 * s0 + threadIdx.x; i < 2 * e; i += kBlock)
          pgs_manifold_pair<kMask>(acc, b, sb, int(sc.slot[i >> 1]), i & 1);
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < nd; i += kBlock) {
    const int body = sc.dyn_list[d0 + i];
    b.store_dv(body, xyz(s_lin[i]), xyz(s_ang[i]));
  }
}

// Scene-major slot order for the blocked kernel: key = scene << 12 | colour per slot.
__global__ void __launch_bounds__(256) k_scene_keys(DevBodies b, int M, SolverBufs sb, uint32_t* keys,
                                                   uint32_t* slots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const int b1 = sb.sinfo[i].x & 0x3fffffff;
  keys[i] = (uint32_t(b.scene[b1]) << 12) | sb.ckeys[i];
  slots[i] = uint32_t(i);
}

// start[s][c] = first position of (scene s, colour >= c) in the sorted keys (lower bound).
__global__ void __launch_bounds__(256) k_scene_starts(const uint32_t* __restrict__ keys, int M, int nscenes,
                                                     int ncolours, uint32_t* start) {
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  const long total = long(nscenes) * (ncolours + 1);
  if (t >= total) return;
  const int s = int(t / (ncolours + 1)), c = int(t % (ncolours + 1));
  const uint32_t key = (uint32_t(s) << 12) | uint32_t(c);
  int lo = 0, hi = M;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  start[t] = uint32_t(lo);
}

// Colours left empty (inherited colours can leave gaps) start where the next one does.
// Also writes the first row record of every class (class_rec[k], class_rec[ncolours] = C),
// which gives the host the contacts per class.
__global__ void k_fix_colour_starts(int ncolours, int M, int C, uint32_t* col_start, const uint32_t* __restrict__ soff,
                                    uint32_t* class_rec) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  col_start[ncolours] = uint32_t(M);
  class_rec[ncolours] = uint32_t(C);
  for (int k = ncolours - 1; k >= 0; --k) {
    if (col_start[k] == 0xffffffffu) col_start[k] = col_start[k + 1];
    class_rec[k] = col_start[k] < uint32_t(M) ? soff[col_start[k]] : uint32_t(C);
  }
}

// Natural-map residual per row; per block (max |phi|, sum phi^2) in double, reduced on the host
// in block order.
template <bool kMask>
__global__ void __launch_bounds__(256) k_residual(DevBodies b, DevContacts c, int M, SolverBufs sb) {
  __shared__ double s_max[256];
  __shared__ double s_sum[256];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double mx = 0.0, sum = 0.0;
  if (s < M) {
    Manifold<kMask> m;
    load_manifold(b, sb, s, m);
    Dv v1{}, v2{};
    load_dv(b, m, v1, v2);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k >= m.x.cnt) break;
      const float4 lam = m.c[k].lam;
      {
        const Row v = row_of(m, k, 0);
        const float w = row_w(v, v1, v2);
        const float a = lam.x - v.r * w;
        const float pr = a > 0.0f ? a : 0.0f;
        const double phi = fabs(double((lam.x - pr) / v.r));
        mx = phi > mx ? phi : mx;
        sum += phi * phi;
      }
#pragma unroll
      for (int f = 1; f <= 2; ++f) {
        const float li = f == 1 ? lam.y : lam.z;
        const Row v = row_of(m, k, f);
        const float w = row_w(v, v1, v2);
        const float mxv = m.c[k].mu * lam.x, mnv = -mxv;
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
  c.imp[sb.rec2c[r]] = sb.lam[r];
}

// ------------------------------------------------------------------ learned warm start
// Features of every contact row record (the network's 24 inputs, in the order documented with
// nphys_mlp_create), computed after assembly from the record, the bodies and the contact:
// velocities over |g| dt, depth over |g| dt^2, mass terms over the row's inverse effective mass,
// cached impulses over ls = r_n |g| dt (the impulse that cancels |g| dt along the normal), which
// is also written per record as the network's output unit.
__global__ void __launch_bounds__(256) k_features(DevBodies b, DevContacts cur, SimParams p, int M, SolverBufs sb,
                                                  float* __restrict__ X, float* __restrict__ ls) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= M) return;
  const SlotCtx x = slot_ctx(sb, s);
  float im1 = 0.0f, im2 = 0.0f;
  m33 I1, I2;
  if (x.d1) I1 = load_inv(b, x.b1, im1);
  if (x.d2) I2 = load_inv(b, x.b2, im2);
  const float g = sqrtf(p.gx * p.gx + p.gy * p.gy + p.gz * p.gz);
  const float gdt = (g > 0.0f ? g : 1.0f) * p.dt;
  const float e = (b.restitution[x.b1] + b.restitution[x.b2]) * 0.5f;
  const float mm = b.margin[x.b1] + b.margin[x.b2];
  for (int k = 0; k < x.cnt; ++k) {
    const uint32_t r = x.r0 + k;
    CRec c;
    load_rec(sb, r, c);
    f3 t1, t2;
    tangents(c.n, t1, t2);
    const float depth = cur.contact(x.first + k)[0].w + mm;
    const float lsc = c.r[0] * gdt;
    float* f = X + long(r) * 24;
#pragma unroll
    for (int row = 0; row < 3; ++row) {
      f[row] = c.rhs[row] / gdt;
      const f3 d = row == 0 ? neg(c.n) : (row == 1 ? t1 : t2);
      const f3 j1 = ang1(c, d), j2 = ang2(c, d);
      const float a1 = x.d1 ? dot(j1, mulv(I1, j1)) : 0.0f;
      const float a2 = x.d2 ? dot(j2, mulv(I2, j2)) : 0.0f;
      f[9 + row] = a1 * c.r[row];
      f[12 + row] = a2 * c.r[row];
    }
    f[3] = c.n.x;
    f[4] = c.n.y;
    f[5] = c.n.z;
    f[6] = depth / (gdt * p.dt);
    f[7] = im1 * c.r[0];
    f[8] = im2 * c.r[0];
    f[15] = float(x.cnt) * 0.25f;
    f[16] = c.lam.x / lsc;
    f[17] = c.lam.y / lsc;
    f[18] = c.lam.z / lsc;
    f[19] = c.mu;
    f[20] = e;
    f[21] = gdt;
    f[22] = (x.d1 && x.d2) ? 0.0f : 1.0f;
    f[23] = c.lam.w;
    ls[r] = lsc;
  }
}

// The network predicts the correction of the cache warm start in units of ls: the warm start
// becomes cache + y ls (cache 0 for a new contact), projected on the admissible set (normal >= 0,
// each tangent within +-mu normal, the pyramid of the friction rows).
__global__ void __launch_bounds__(256) k_apply_prediction(SolverBufs sb, int C, const float* __restrict__ y,
                                                          const float* __restrict__ ls) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= C) return;
  const float s = ls[r];
  const float mu = sb.rec[long(r) * kRecFields].w;
  float4* L = sb.lam + r;
  const float4 c = *L;
  const float ln = fmaxf(c.x + y[3 * r] * s, 0.0f);
  const float lim = mu * ln;
  const float t1 = fminf(fmaxf(c.y + y[3 * r + 1] * s, -lim), lim);
  const float t2 = fminf(fmaxf(c.z + y[3 * r + 2] * s, -lim), lim);
  *L = make_float4(ln, t1, t2, c.w);
}

// Training targets of the last solve: the impulses in units of ls.
__global__ void __launch_bounds__(256) k_sample_targets(SolverBufs sb, int C, const float* __restrict__ ls,
                                                        float* __restrict__ t) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= C) return;
  const float4 l = sb.lam[r];
  const float s = ls[r];
  t[3 * r] = l.x / s;
  t[3 * r + 1] = l.y / s;
  t[3 * r + 2] = l.z / s;
}

// Impulses of all records to / from a side buffer (the judge's saved solves).
__global__ void __launch_bounds__(256) k_lam_copy(SolverBufs sb, int C, float4* __restrict__ buf, int to_buf) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= C) return;
  float4* L = sb.lam + r;
  if (to_buf) buf[r] = *L;
  else *L = buf[r];
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
                     sb.deg, sb.adj, sb.state, sb.remaining, sb.remaining + 1);
}

void launch_colour_keys(int M, SolverBufs& sb, uint32_t* ncol, hipStream_t s) {
  fill_u32(ncol, 0u, 1, s);
  if (M == 0) return;
  hipLaunchKernelGGL(k_colour_keys, dim3(nb(M)), dim3(256), 0, s, M, sb.state, sb.ckeys, sb.order, ncol);
}

void launch_colour_order(const DevBodies& b_view_for_slots, const DevContacts& c, int M, int ncolours,
                         SolverBufs& sb, hipStream_t s) {
  if (M == 0) return;
  const int bits = ncolours <= 256 ? 8 : 16;
  radix_sort_pairs(sb.ckeys, sb.order, sb.ckeys_alt, sb.order_alt, M, bits, sb.tmp, s);
  fill_u32(sb.col_start, 0xffffffffu, size_t(ncolours) + 1, s);
  hipLaunchKernelGGL(k_colour_starts, dim3(nb(M)), dim3(256), 0, s, c, M, sb.ckeys, sb.order, sb.col_start,
                     sb.mcount_ord);
  scan_exclusive_u32(sb.mcount_ord, sb.soff, M, sb.tmp, s);
  hipLaunchKernelGGL(k_slot_info, dim3(nb(M)), dim3(256), 0, s, b_view_for_slots, c, M, sb.order, sb.soff, sb.sinfo,
                     sb.slot_of);
}

void launch_assemble(const DevBodies& b, const DevContacts& cur, const DevContacts& prev, int prevM,
                     int M, const SimParams& p, SolverBufs& sb, const MatTable& mt, hipStream_t s) {
  if (M == 0) return;
  hipLaunchKernelGGL(k_assemble, dim3(nb(M)), dim3(256), 0, s, b, cur, prev, prevM, M, p, sb, mt);
}

void launch_fix_colour_starts(int ncolours, int M, int C, SolverBufs& sb, uint32_t* class_rec, hipStream_t s) {
  hipLaunchKernelGGL(k_fix_colour_starts, dim3(1), dim3(1), 0, s, ncolours, M, C, sb.col_start, sb.soff, class_rec);
}

// One lane per manifold: the cap on a colour launch's grid, one 64-lane workgroup per 64
// manifolds of the largest colour class rounded up to a power of two, at least 64 and at most
// 8,192; each class is launched with the workgroups it needs under that cap, and larger classes
// are walked by the grid-stride loop.
int sweep_blocks(int max_class) {
  int need = (max_class + 63) / 64;
  int blocks = 64;
  while (blocks < need && blocks < 8192) blocks *= 2;
  return blocks;
}

// The masked kernel variants run only in worlds where some dynamic body has a kinematic dof.
void launch_warmstart_colour(const DevBodies& b, const DevContacts& c, int first, int n, int blocks, SolverBufs& sb,
                             hipStream_t s) {
  if (b.kmask) hipLaunchKernelGGL(k_warmstart<true>, dim3(blocks), dim3(64), 0, s, b, c, sb, first, n);
  else hipLaunchKernelGGL(k_warmstart<false>, dim3(blocks), dim3(64), 0, s, b, c, sb, first, n);
}

void launch_pgs_colour(const DevBodies& b, const DevContacts& c, int first, int n, int blocks, SolverBufs& sb,
                       hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
  if (b.kmask) {
    if (ev0) hipExtLaunchKernelGGL(k_pgs<true>, dim3(blocks), dim3(64), 0, s, ev0, ev1, 0, b, c, sb, first, n);
    else hipLaunchKernelGGL(k_pgs<true>, dim3(blocks), dim3(64), 0, s, b, c, sb, first, n);
    return;
  }
  if (ev0) {
    // the dispatch packet itself records both events, timing the kernel alone
    hipExtLaunchKernelGGL(k_pgs<false>, dim3(blocks), dim3(64), 0, s, ev0, ev1, 0, b, c, sb, first, n);
  } else {
    hipLaunchKernelGGL(k_pgs<false>, dim3(blocks), dim3(64), 0, s, b, c, sb, first, n);
  }
}

void launch_colour_run(const DevBodies& b, SolverBufs& sb, int kv0, int nk, bool warm, hipStream_t s) {
  if (b.kmask) {
    if (warm) hipLaunchKernelGGL((k_colour_run<true, true>), dim3(1), dim3(kRunBlock), 0, s, b, sb, kv0, nk);
    else hipLaunchKernelGGL((k_colour_run<false, true>), dim3(1), dim3(kRunBlock), 0, s, b, sb, kv0, nk);
    return;
  }
  if (warm) hipLaunchKernelGGL((k_colour_run<true, false>), dim3(1), dim3(kRunBlock), 0, s, b, sb, kv0, nk);
  else hipLaunchKernelGGL((k_colour_run<false, false>), dim3(1), dim3(kRunBlock), 0, s, b, sb, kv0, nk);
}

int residual_blocks(int M) { return int(nb(M > 0 ? M : 1)); }

void launch_residual(const DevBodies& b, const DevContacts& c, int M, SolverBufs& sb, hipStream_t s) {
  if (b.kmask) hipLaunchKernelGGL(k_residual<true>, dim3(nb(M > 0 ? M : 1)), dim3(256), 0, s, b, c, M, sb);
  else hipLaunchKernelGGL(k_residual<false>, dim3(nb(M > 0 ? M : 1)), dim3(256), 0, s, b, c, M, sb);
}

void launch_cache_impulses(const DevContacts& c, SolverBufs& sb, int C, hipStream_t s) {
  if (C == 0) return;
  hipLaunchKernelGGL(k_cache, dim3(nb(C)), dim3(256), 0, s, c, sb, C);
}

void launch_features(const DevBodies& b, const DevContacts& c, const SimParams& p, int M, SolverBufs& sb, float* X,
                     float* ls, hipStream_t s) {
  if (M == 0) return;
  hipLaunchKernelGGL(k_features, dim3(nb(M)), dim3(256), 0, s, b, c, p, M, sb, X, ls);
}

void launch_apply_prediction(SolverBufs& sb, int C, const float* y, const float* ls, hipStream_t s) {
  if (C == 0) return;
  hipLaunchKernelGGL(k_apply_prediction, dim3(nb(C)), dim3(256), 0, s, sb, C, y, ls);
}

void launch_sample_targets(SolverBufs& sb, int C, const float* ls, float* t, hipStream_t s) {
  if (C == 0) return;
  hipLaunchKernelGGL(k_sample_targets, dim3(nb(C)), dim3(256), 0, s, sb, C, ls, t);
}

void launch_lam_copy(SolverBufs& sb, int C, float4* buf, bool to_buf, hipStream_t s) {
  if (C == 0) return;
  hipLaunchKernelGGL(k_lam_copy, dim3(nb(C)), dim3(256), 0, s, sb, C, buf, to_buf ? 1 : 0);
}

void launch_scene_order(const DevBodies& b, int M, int nscenes, int ncolours, SolverBufs& sb, SceneBlocks& sc,
                        hipStream_t s) {
  if (M == 0) return;
  hipLaunchKernelGGL(k_scene_keys, dim3(nb(M)), dim3(256), 0, s, b, M, sb, sc.keys, sc.slot);
  radix_sort_pairs(sc.keys, sc.slot, sc.keys_alt, sc.slot_alt, M, 28, sb.tmp, s);
  const long total = long(nscenes) * (ncolours + 1);
  hipLaunchKernelGGL(k_scene_starts, dim3(nb(total)), dim3(256), 0, s, sc.keys, M, nscenes, ncolours, sc.start);
}

// Two launches: the scenes with more than kSmallScene dynamic bodies (`block` lanes, 64 KB of LDS)
// on s_large, the others (64-lane workgroups, 8 KB of LDS, many resident per CU) on s_small, so the
// few large scenes run beside the many small ones.
template <bool kMask>
void sweeps_blocked(const DevBodies& b, SolverBufs& sb, const SceneBlocks& sc, int ncolours, int nsweeps, bool warm,
                    int block, hipStream_t s_small, hipStream_t s_large) {
  const int w = warm ? 1 : 0;
  if (kMask && block == 1024) block = 512;  // the masked kernel would spill at 1024 lanes (128 VGPRs)
  if (sc.n_large > 0) {
    if (block == 256)
      hipLaunchKernelGGL((k_sweeps_blocked<256, kLdsBodies, kMask>), dim3(sc.n_large), dim3(256), 0, s_large, b, sb,
                         sc, sc.large_list, ncolours, nsweeps, w);
    else if (kMask || block == 512)
      hipLaunchKernelGGL((k_sweeps_blocked<512, kLdsBodies, kMask>), dim3(sc.n_large), dim3(512), 0, s_large, b, sb,
                         sc, sc.large_list, ncolours, nsweeps, w);
    else
      hipLaunchKernelGGL((k_sweeps_blocked<1024, kLdsBodies, false>), dim3(sc.n_large), dim3(1024), 0, s_large, b,
                         sb, sc, sc.large_list, ncolours, nsweeps, w);
  }
  if (sc.n_small > 0)
    hipLaunchKernelGGL((k_sweeps_blocked<64, kSmallScene, kMask>), dim3(sc.n_small), dim3(64), 0, s_small, b, sb, sc,
                       sc.small_list, ncolours, nsweeps, w);
}

void launch_sweeps_blocked(const DevBodies& b, SolverBufs& sb, const SceneBlocks& sc, int ncolours, int nsweeps,
                           bool warm, int block, hipStream_t s_small, hipStream_t s_large) {
  if (b.kmask) sweeps_blocked<true>(b, sb, sc, ncolours, nsweeps, warm, block, s_small, s_large);
  else sweeps_blocked<false>(b, sb, sc, ncolours, nsweeps, warm, block, s_small, s_large);
}

}  // namespace nph
