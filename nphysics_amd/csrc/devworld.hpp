// nphysics_amd/csrc/devworld.hpp -- device-side views of one world, passed to kernels by value.
//
// HBM layout (one float4 per 3-vector so every access is a 16-byte load):
//   bodies   a 32-byte pose block per body (pos, rot as quaternion xyzw) and a 64-byte solver
//            block holding exactly what a velocity sweep reads and writes per body, so a body
//            costs one 64-byte line per manifold:
//              [0] (I00, I01, I02, 1/m)  [1] (I10, I11, I12, I20)
//              [2] (I21, I22, dv_lin.x, dv_lin.y)  [3] (dv_lin.z, dv_ang.x, dv_ang.y, dv_ang.z)
//            with I the inverse augmented inertia. Structure of arrays for the rest:
//            pos0/rot0/linvel/angvel, external velocity increment dt*acc (lin, ang), loose AABB
//            (mn, mx), shape data, and where bodies can sleep (energy, threshold) per body.
//   contacts one 96-byte record per contact slot, each broad-phase pair owning 4 consecutive slots
//            (a manifold's contacts are slots 4 p .. 4 p + count - 1 of its pair p):
//            [0] world1 | depth  [1] world2 | feature  [2] normal | kinematic kind
//            [3] local point 1  [4] local point 2  [5] local normal (the position solver's);
//            the solved impulse per slot; per manifold: bodies, first slot, count.
//   rows     per contact in colour order: a 4-float4 row record and a float4 of impulses (see
//            k_solver.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "devmath.hpp"

namespace nph {

struct SimParams {
  float dt, inv_dt, erp, warmstart_coeff, restitution_threshold;
  float allowed_linear_error, max_linear_correction;
  float gx, gy, gz;
  float allowed_angular_error, max_angular_correction;
  // time step of the bodies' last dynamics update (dt, or a CCD substep's after one)
  float dyn_dt;
};

// float4 per pose block and per solver block
constexpr int kPoseBlock = 2;
constexpr int kSolverBlock = 4;

// CollisionGroups::can_interact_with_groups (ncollide3d 0.31, unvendored): each collider is a
// member of some group the other whitelists and of no group the other blacklists.
NPH_HD bool groups_interact(uint4 a, uint4 c) {
  return (a.x & c.z) == 0u && (c.x & a.z) == 0u && (a.x & c.y) != 0u && (c.x & a.y) != 0u;
}

struct DevBodies {
  int n;
  const int* status;
  const int* shape;
  const int* scene;      // independent-world index; bodies of different scenes never interact
  const float4* dims;    // half extents; balls (r, r, r)
  const float* margin;
  const float* qlim;     // margin + linear prediction
  const float* friction;
  const float* restitution;
  const float* mass;
  const float4* local_i; // diagonal of the local angular inertia
  float4* pose;          // kPoseBlock float4 per body
  float4* sol;           // kSolverBlock float4 per body
  float4* pos0;
  float4* rot0;
  float4* linvel;
  float4* angvel;
  float4* ext_lin;
  float4* ext_ang;
  float4* loose_mn;
  float4* loose_mx;
  float2* act;           // (activation energy, sleep threshold); null when no body can sleep
  // Optional per-body attributes, null when every body has the default:
  const float4* damp;    // (linear damping, angular damping, max |v|, max |w|)
  const int* flags;      // kinematic dofs (bits 0-5), gravity disabled (bit 6), friction and
                         // restitution combine modes (bits 8-9, 10-11); null: all defaults
  const int* kmask;      // = flags when some dynamic body has a kinematic dof, else null
  const int* mat_id;     // material id for the coefficient table (< 0: none)
  const float4* surf_vel;  // surface velocity in the collider frame
  float4* f_lin;         // accumulated external forces, null when none are pending
  float4* f_ang;
  float4* w_dyn;         // angular velocity of the last dynamics update (written by k_integrate);
                         // w = 2 marks a body put back by the CCD, whose last update used ccd_rot
  // CCD, null when no body is CCD-enabled: flags (bit 0 ccd_enabled, bit 1 linear motion
  // interpolation), the key of the body a collider belongs to (itself, or the first collider of its
  // static body), and the rotation of the last dynamics update of a body the CCD put back
  const int* ccd;
  const int* ccd_key;
  float4* ccd_rot;
  // collision groups (membership, whitelist, blacklist, -), null when every collider has the default
  const uint4* groups;

  // Jacobian mask of body i (1 free, 0 kinematic), linear and angular (rigid_body.rs:46,103-118).
  NPH_HD void mask(int i, f3& ml, f3& ma) const {
    const int k = kmask ? kmask[i] : 0;
    ml = f3{(k & 1) ? 0.0f : 1.0f, (k & 2) ? 0.0f : 1.0f, (k & 4) ? 0.0f : 1.0f};
    ma = f3{(k & 8) ? 0.0f : 1.0f, (k & 16) ? 0.0f : 1.0f, (k & 32) ? 0.0f : 1.0f};
  }

  NPH_HD float4& pos(int i) const { return pose[long(i) * kPoseBlock]; }
  NPH_HD float4& rot(int i) const { return pose[long(i) * kPoseBlock + 1]; }
  NPH_HD float4* solver(int i) const { return sol + long(i) * kSolverBlock; }

  // Inverse augmented inertia and inverse mass of body i.
  NPH_HD m33 inv_inertia(int i, float& inv_mass) const {
    const float4* q = solver(i);
    const float4 r0 = q[0], r1 = q[1], r2 = q[2];
    m33 m;
    m.m[0][0] = r0.x; m.m[0][1] = r0.y; m.m[0][2] = r0.z;
    m.m[1][0] = r1.x; m.m[1][1] = r1.y; m.m[1][2] = r1.z;
    m.m[2][0] = r1.w; m.m[2][1] = r2.x; m.m[2][2] = r2.y;
    inv_mass = r0.w;
    return m;
  }
  // Writes the inverse inertia and mass of body i and clears its velocity increment.
  NPH_HD void set_inv_inertia(int i, const m33& m, float inv_mass) const {
    float4* q = solver(i);
    q[0] = make_float4(m.m[0][0], m.m[0][1], m.m[0][2], inv_mass);
    q[1] = make_float4(m.m[1][0], m.m[1][1], m.m[1][2], m.m[2][0]);
    q[2] = make_float4(m.m[2][1], m.m[2][2], 0.0f, 0.0f);
    q[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  NPH_HD void load_dv(int i, f3& lin, f3& ang) const {
    const float4* q = solver(i);
    const float2 a = *reinterpret_cast<const float2*>(&q[2].z);
    const float4 c = q[3];
    lin = f3{a.x, a.y, c.x};
    ang = f3{c.y, c.z, c.w};
  }
  NPH_HD void store_dv(int i, f3 lin, f3 ang) const {
    float4* q = solver(i);
    *reinterpret_cast<float2*>(&q[2].z) = make_float2(lin.x, lin.y);
    q[3] = make_float4(lin.z, ang.x, ang.y, ang.z);
  }
};

constexpr int kContactFields = 6;  // float4 per contact record
constexpr int kPairSlots = 4;      // contact slots per broad-phase pair

struct DevContacts {
  int2* mbody;       // (body1, body2) per manifold, sorted
  int* mfirst;       // first contact slot
  int* mcount;
  float4* rec;       // kContactFields float4 per contact slot, layout above
  float4* imp;       // per slot: (normal, tangent 1, tangent 2) impulses after the solve
  int* bfirst;       // per body below bcount: its first manifold as body 1, or -1; null: none built
  int bcount;

  NPH_HD float4* contact(long slot) const { return rec + slot * kContactFields; }
  NPH_HD uint32_t feature(long slot) const { return __float_as_uint(rec[slot * kContactFields + 1].w); }
};

// Index of the manifold of body pair (b1, b2) among the first M manifolds of c (sorted by pair),
// or -1: through the per-body index (the manifolds of one body 1 are contiguous and sorted by
// body 2) when there is one, else by binary search.
__device__ inline int find_manifold(const DevContacts& c, int M, int b1, int b2) {
  if (c.bfirst) {
    if (b1 >= c.bcount) return -1;
    for (int i = c.bfirst[b1]; i >= 0 && i < M; ++i) {
      const int2 k = c.mbody[i];
      if (k.x != b1 || k.y > b2) return -1;
      if (k.y == b2) return i;
    }
    return -1;
  }
  int lo = 0, hi = M;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int2 k = c.mbody[mid];
    if (k.x < b1 || (k.x == b1 && k.y < b2)) lo = mid + 1;
    else hi = mid;
  }
  if (lo < M) {
    const int2 k = c.mbody[lo];
    if (k.x == b1 && k.y == b2) return lo;
  }
  return -1;
}

NPH_HD m33 load_inv(const DevBodies& b, int i, float& inv_mass) { return b.inv_inertia(i, inv_mass); }

// Component-wise product (the jacobian mask, rigid_body.rs:615,690).
NPH_HD f3 cmul(f3 a, f3 m) { return f3{a.x * m.x, a.y * m.y, a.z * m.z}; }

// Body flag bits (DevBodies::flags).
constexpr int kFlagNoGravity = 64;
NPH_HD int fric_mode(int flags) { return (flags >> 8) & 3; }

// RigidBody::update_dynamics (rigid_body.rs:558-588) at rotation q with angular velocity w over
// dyn_dt: I_w = R I_l R^T, augmented inertia I_w + [w dt]x I_w - [(I_w w) dt]x and its inverse
// (stored with 1/m), or, for a body that keeps the augmented mass of an earlier update, that one
// (`stored`; wd: the angular velocity of a stale update when given); then update_acceleration
// (:590-619): gyroscopic term, gravity g unless disabled, external forces, the jacobian mask; and
// the external velocity increment dt * acc (moreau_jean_solver.rs:166-174) of dynamic body i.
__device__ __forceinline__ void body_prestep(const DevBodies& b, int i, q4 q, f3 w, bool stored, bool stale, f3 wd,
                                             float dyn_dt, float dt, f3 g) {
  const m33 R = qmat(q);
  const float4 li = b.local_i[i];
  m33 L;
  L.m[0][0] = li.x; L.m[0][1] = 0.0f; L.m[0][2] = 0.0f;
  L.m[1][0] = 0.0f; L.m[1][1] = li.y; L.m[1][2] = 0.0f;
  L.m[2][0] = 0.0f; L.m[2][1] = 0.0f; L.m[2][2] = li.z;
  const m33 Iw = mmul(mmul(R, L), mtrans(R));
  const f3 iw = mulv(Iw, w);
  float inv_mass;
  m33 inv;
  if (stored) {
    inv = b.inv_inertia(i, inv_mass);
  } else {
    const f3 wu = stale ? wd : w;
    const f3 iwd = stale ? mulv(Iw, wd) : iw;
    const m33 wx = cross_matrix(scale(wu, dyn_dt));
    const m33 iwx = cross_matrix(scale(iwd, dyn_dt));
    const m33 wxI = mmul(wx, Iw);
    m33 aug;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) aug.m[r][c] = Iw.m[r][c] + (wxI.m[r][c] - iwx.m[r][c]);
    const float mass = b.mass[i];
    inv_mass = mass == 0.0f ? 0.0f : 1.0f / mass;
    try_inverse(aug, inv);
  }
  b.set_inv_inertia(i, inv, inv_mass);
  const f3 gyro = neg(cross(w, iw));
  f3 acc_ang = mulv(inv, gyro);
  const bool grav = !(b.flags && (b.flags[i] & kFlagNoGravity));
  f3 acc_lin = (inv_mass != 0.0f && grav) ? g : f3{0.0f, 0.0f, 0.0f};
  if (b.f_lin) {  // acceleration += inv_augmented_mass * external_forces
    acc_lin = add(acc_lin, scale(xyz(b.f_lin[i]), inv_mass));
    acc_ang = add(acc_ang, mulv(inv, xyz(b.f_ang[i])));
  }
  if (b.kmask) {  // (.) jacobian_mask
    f3 ml, ma;
    b.mask(i, ml, ma);
    acc_lin = cmul(acc_lin, ml);
    acc_ang = cmul(acc_ang, ma);
  }
  b.ext_lin[i] = make_float4(dt * acc_lin.x, dt * acc_lin.y, dt * acc_lin.z, 0.0f);
  b.ext_ang[i] = make_float4(dt * acc_ang.x, dt * acc_ang.y, dt * acc_ang.z, 0.0f);
}
NPH_HD int rest_mode(int flags) { return (flags >> 10) & 3; }

// MaterialCombineMode::combine (material.rs:72-86): Max > Multiply > Min > Average.
NPH_HD float combine_coeff(float a, int ma, float b, int mb) {
  if (ma == 3 || mb == 3) return a > b ? a : b;
  if (ma == 2 || mb == 2) return a * b;
  if (ma == 1 || mb == 1) return a < b ? a : b;
  return (a + b) * 0.5f;
}

// Sorted material-pair table (MaterialsCoefficientsTable): key (min id << 32 | max id).
struct MatTable {
  const unsigned long long* fkey;
  const float* fval;
  int nf;
  const unsigned long long* rkey;
  const float* rval;
  int nr;
};

__device__ inline bool mat_lookup(const unsigned long long* key, const float* val, int n, unsigned long long k,
                                  float& out) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (key[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  if (lo < n && key[lo] == k) {
    out = val[lo];
    return true;
  }
  return false;
}

}  // namespace nph
