// nphysics_amd/csrc/k_joints.hip -- joint constraints in the coloured solver.
//
// Restates the JointConstraint implementations BallConstraint (src/joint/ball_constraint.rs:64-178),
// FixedConstraint (src/joint/fixed_constraint.rs:91-249), RevoluteConstraint without limits or
// motor (src/joint/revolute_constraint.rs:172-349) and PrismaticConstraint without limits
// (src/joint/prismatic_constraint.rs:132-345: two linear rows orthogonal to the axis, then three
// angular rows; rotation cancelled, then anchor 2 projected onto the axis) and CylindricalConstraint
// without limits or motor (src/joint/cylindrical_constraint.rs:125-292: two linear and two angular
// rows orthogonal to the axis; axes aligned, then anchor 2 projected), with their solver helpers
// (src/solver/helper.rs:53-135, 167-296, 364-420, 422-612, 614-769), inside the per-step solve:
//   k_joint_assemble   velocity_constraints of every active joint (moreau_jean_solver.rs:211-223):
//                      bilateral rows with unbounded (Independent) limits, linear along x, y, z at
//                      the anchors, then angular (fixed: x, y, z; revolute: the two directions
//                      orthogonal to the world axis), warm-started with the joint's cached impulses;
//   joint phases       one lane per joint of one colour (joints of a colour share no dynamic body):
//                      the warm start after the contacts' (sor_prox.rs:56-58), every sweep's joint
//                      rows before the contacts' (:167-168), every position iteration's joint
//                      constraints before the contacts' (nonlinear_sor_prox.rs:33-38);
//   k_joint_cache      JointConstraint::cache_impulses: impulses kept per joint, the joint broken
//                      once |impulse|^2 / dt^2 exceeds the squared break force or torque;
//   k_joint_residual   the rows' natural-map residual (an unbounded row's is w itself);
//   k_joint_union      joints that are not broken join their bodies' activation islands
//                      (activation_manager.rs:154-159).
// The row arithmetic is the contact rows' (k_solver.hip): a row record keeps the direction, the
// lever arms, r and rhs; the jacobians and weighted jacobians are rebuilt in every phase with the
// assembly's own expressions, so every phase sees the same FP32 values as the CPU restatement.
// A static body enters a row with zero inverse mass, zero inverse inertia and zero velocity
// increment (its terms are exact zeros).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "devworld.hpp"
#include "stages.hpp"

namespace nph {

namespace {

constexpr int kJointBlock = 256;

struct JBody {
  f3 com, v, w, el, ea;
  m33 inv;
  float im;
  f3 ml, ma;
};

__device__ __forceinline__ JBody load_jbody(const DevBodies& b, int i) {
  JBody v;
  v.com = xyz(b.pos(i));
  v.v = xyz(b.linvel[i]);
  v.w = xyz(b.angvel[i]);
  v.el = xyz(b.ext_lin[i]);
  v.ea = xyz(b.ext_ang[i]);
  v.inv = load_inv(b, i, v.im);
  b.mask(i, v.ml, v.ma);
  return v;
}

// fill_constraint_geometry of one body (rigid_body.rs:672-722) for a linear direction at `point`
// or, with `angular`, an angular direction (force (0, dir)); inv_r gains the inverse mass either
// way (rigid_body.rs:707-708).
__device__ __forceinline__ void jfill(const JBody& bv, f3 point, f3 dir, bool angular, float& inv_r, float& rhs) {
  const f3 zero{0.0f, 0.0f, 0.0f};
  const f3 p = sub(point, bv.com);
  const f3 ang_u = angular ? dir : cross(p, dir);
  const f3 fl = angular ? zero : dir;
  const f3 lin = cmul(fl, bv.ml);
  const f3 ang = cmul(ang_u, bv.ma);
  const f3 wang = mulv(bv.inv, ang);
  inv_r = inv_r + (bv.im + dot(ang, wang));
  const float F[6] = {fl.x, fl.y, fl.z, ang_u.x, ang_u.y, ang_u.z};
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

__device__ __forceinline__ void tangents_of(f3 n, f3& t1, f3& t2) {
  f3 a = fabsf(n.x) > fabsf(n.y) ? f3{n.z, 0.0f, -n.x} : f3{0.0f, -n.z, n.y};
  const float l = sqrtf(dot(a, a));
  a = f3{a.x / l, a.y / l, a.z / l};
  t1 = cross(a, n);
  t2 = a;
}

__device__ __forceinline__ int joint_rows(int kind) {
  return kind == JOINT_BALL ? 3 : (kind == JOINT_FIXED ? 6 : (kind == JOINT_CYLINDRICAL ? 4 : 5));
}

__device__ __forceinline__ bool body_awake(const DevBodies& b, int i) {
  return b.status[i] == STATUS_DYNAMIC && (!b.act || b.act[i].x != 0.0f);
}

// World anchors of joint ji and, for the fixed joint, the frames' world rotations
// (position_at_material_point * ref_frame, rigid_body.rs:641-648).
__device__ __forceinline__ void joint_frames(const DevBodies& b, const JointBufs& j, int ji, int b1, int b2, f3& w1,
                                             f3& w2, q4& r1, q4& r2) {
  const q4 q1 = toq(b.rot(b1)), q2 = toq(b.rot(b2));
  w1 = add(qrot(q1, xyz(j.a1[ji])), xyz(b.pos(b1)));
  w2 = add(qrot(q2, xyz(j.a2[ji])), xyz(b.pos(b2)));
  r1 = qmul(q1, toq(j.f1[ji]));
  r2 = qmul(q2, toq(j.f2[ji]));
}

__global__ void __launch_bounds__(256) k_joint_assemble(DevBodies b, SimParams p, JointBufs j) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= j.nj) return;
  const int ji = j.order[s];
  const int4 inf = j.info[ji];
  const int kind = inf.x, b1 = inf.y, b2 = inf.z;
  float4* R = j.rec + long(s) * kJointRec;
  const bool active = j.state[ji] == 0 && (body_awake(b, b1) || body_awake(b, b2));
  if (!active) {
    R[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  const bool d1 = b.status[b1] == STATUS_DYNAMIC, d2 = b.status[b2] == STATUS_DYNAMIC;
  JBody v1, v2;
  if (d1) v1 = load_jbody(b, b1);
  if (d2) v2 = load_jbody(b, b2);
  f3 w1, w2;
  q4 r1, r2;
  joint_frames(b, j, ji, b1, b2, w1, w2, r1, r2);
  const int nrows = joint_rows(kind);
  f3 dirs[6];
  dirs[0] = f3{1.0f, 0.0f, 0.0f};
  dirs[1] = f3{0.0f, 1.0f, 0.0f};
  dirs[2] = f3{0.0f, 0.0f, 1.0f};
  if (kind == JOINT_FIXED) {
    dirs[3] = dirs[0];
    dirs[4] = dirs[1];
    dirs[5] = dirs[2];
  } else if (kind == JOINT_CYLINDRICAL) {
    // restrict_relative_linear_velocity_to_axis, then restrict_relative_angular_velocity_to_axis
    // (helper.rs:614-699): both the orthonormal_subspace_basis of the world axis
    const f3 axis = qrot(toq(b.rot(b1)), xyz(j.f1[ji]));
    tangents_of(axis, dirs[0], dirs[1]);
    dirs[2] = dirs[0];
    dirs[3] = dirs[1];
  } else if (kind == JOINT_PRISMATIC) {
    // restrict_relative_linear_velocity_to_axis (helper.rs:788-856): the orthonormal_subspace_basis
    // of the world axis, then cancel_relative_angular_velocity: x, y, z
    const f3 axis = qrot(toq(b.rot(b1)), xyz(j.f1[ji]));
    tangents_of(axis, dirs[0], dirs[1]);
    dirs[2] = f3{1.0f, 0.0f, 0.0f};
    dirs[3] = f3{0.0f, 1.0f, 0.0f};
    dirs[4] = f3{0.0f, 0.0f, 1.0f};
  } else if (kind == JOINT_REVOLUTE) {
    // axis1 = position_at_material_point(anchor1) * axis1 (revolute_constraint.rs:230), and the
    // orthonormal_subspace_basis of it in that order
    const f3 axis = qrot(toq(b.rot(b1)), xyz(j.f1[ji]));
    tangents_of(axis, dirs[3], dirs[4]);
  }
  const float4 il = j.imp_lin[ji], ia = j.imp_ang[ji];
  const int nlin = joint_linear_rows(kind);
  float imp[6] = {il.x, il.y, il.z, ia.x, ia.y, ia.z};
  if (nlin == 2) {  // prismatic: angular x, y, z; cylindrical: the two angular rows
    imp[2] = ia.x;
    imp[3] = ia.y;
    imp[4] = ia.z;
    imp[5] = 0.0f;
  }
  float rhs_all[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  for (int k = 0; k < nrows; ++k) {
    const bool angular = k >= nlin;
    float inv_r = 0.0f, rhs = 0.0f;
    if (d1) jfill(v1, w1, dirs[k], angular, inv_r, rhs);
    if (d2) jfill(v2, w2, neg(dirs[k]), angular, inv_r, rhs);
    const float rr = inv_r != 0.0f ? 1.0f / inv_r : 1.0f;
    R[2 + k] = f4(dirs[k], rr);
    rhs_all[k] = rhs;
  }
  const f3 zero{0.0f, 0.0f, 0.0f};
  const f3 p1 = d1 ? sub(w1, v1.com) : zero;
  const f3 p2 = d2 ? sub(w2, v2.com) : zero;
  R[0] = f4(p1, __int_as_float(nrows | (kind << 8)));
  R[1] = f4(p2, __int_as_float(1 | (int(d1) << 1) | (int(d2) << 2)));
  R[8] = make_float4(rhs_all[0], rhs_all[1], rhs_all[2], rhs_all[3]);
  R[9] = make_float4(rhs_all[4], rhs_all[5], 0.0f, 0.0f);
  j.lam[2 * long(s)] = make_float4(imp[0], imp[1], imp[2], imp[3]);
  j.lam[2 * long(s) + 1] = make_float4(imp[4], imp[5], 0.0f, 0.0f);
}

// One joint's rows in registers for the velocity phases.
struct JointCtx {
  int b1, b2, nrows, nlin;
  bool active, d1, d2;
  f3 p1, p2;
  f3 dir[6];
  float r[6], rhs[6], lam[6];
  m33 I1, I2;
  float im1, im2;
  f3 ml1, ma1, ml2, ma2;
  f3 x1l, x1a, x2l, x2a;  // velocity increments
};

__device__ __forceinline__ void load_ctx(const DevBodies& b, const JointBufs& j, int s, JointCtx& c) {
  const float4* R = j.rec + long(s) * kJointRec;
  const float4 r0 = R[0], r1 = R[1];
  const int fl = __float_as_int(r1.w);
  c.active = (fl & 1) != 0;
  if (!c.active) return;
  c.d1 = (fl >> 1) & 1;
  c.d2 = (fl >> 2) & 1;
  c.nrows = __float_as_int(r0.w) & 0xff;
  c.nlin = joint_linear_rows(__float_as_int(r0.w) >> 8);
  const int4 inf = j.info[j.order[s]];
  c.b1 = inf.y;
  c.b2 = inf.z;
  c.p1 = xyz(r0);
  c.p2 = xyz(r1);
  for (int k = 0; k < 6; ++k) {
    const float4 d = R[2 + k];
    c.dir[k] = xyz(d);
    c.r[k] = d.w;
  }
  const float4 h0 = R[8], h1 = R[9];
  c.rhs[0] = h0.x; c.rhs[1] = h0.y; c.rhs[2] = h0.z; c.rhs[3] = h0.w; c.rhs[4] = h1.x; c.rhs[5] = h1.y;
  const float4 l0 = j.lam[2 * long(s)], l1 = j.lam[2 * long(s) + 1];
  c.lam[0] = l0.x; c.lam[1] = l0.y; c.lam[2] = l0.z; c.lam[3] = l0.w; c.lam[4] = l1.x; c.lam[5] = l1.y;
  const f3 zero{0.0f, 0.0f, 0.0f};
  c.im1 = c.im2 = 0.0f;
  c.I1 = m33{};
  c.I2 = m33{};
  c.x1l = c.x1a = c.x2l = c.x2a = zero;
  if (c.d1) {
    c.I1 = load_inv(b, c.b1, c.im1);
    b.load_dv(c.b1, c.x1l, c.x1a);
  }
  if (c.d2) {
    c.I2 = load_inv(b, c.b2, c.im2);
    b.load_dv(c.b2, c.x2l, c.x2a);
  }
  b.mask(c.b1, c.ml1, c.ma1);
  b.mask(c.b2, c.ml2, c.ma2);
}

__device__ __forceinline__ void store_ctx(const DevBodies& b, const JointBufs& j, int s, const JointCtx& c,
                                          bool write_lam) {
  if (c.d1) b.store_dv(c.b1, c.x1l, c.x1a);
  if (c.d2) b.store_dv(c.b2, c.x2l, c.x2a);
  if (write_lam) {
    j.lam[2 * long(s)] = make_float4(c.lam[0], c.lam[1], c.lam[2], c.lam[3]);
    j.lam[2 * long(s) + 1] = make_float4(c.lam[4], c.lam[5], 0.0f, 0.0f);
  }
}

// Jacobians of row k: linear rows [d; p x d] on body 1 and [-d; p2 x -d] on body 2, angular rows
// [0; d] and [0; -d], each masked, with their inverse-inertia weighted angular parts.
struct JRow {
  f3 d1, d2, a1, a2, wa1, wa2;
};

__device__ __forceinline__ JRow jrow(const JointCtx& c, int k) {
  JRow v;
  const f3 zero{0.0f, 0.0f, 0.0f};
  const f3 d = c.dir[k];
  const f3 nd = neg(d);
  const bool angular = k >= c.nlin;
  v.d1 = cmul(angular ? zero : d, c.ml1);
  v.d2 = cmul(angular ? zero : nd, c.ml2);
  v.a1 = cmul(angular ? d : cross(c.p1, d), c.ma1);
  v.a2 = cmul(angular ? nd : cross(c.p2, nd), c.ma2);
  v.wa1 = mulv(c.I1, v.a1);
  v.wa2 = mulv(c.I2, v.a2);
  return v;
}

__device__ __forceinline__ float jdot6(f3 dl, f3 ang, f3 vl, f3 va) {
  float s = 0.0f;
  s = s + dl.x * vl.x;
  s = s + dl.y * vl.y;
  s = s + dl.z * vl.z;
  s = s + ang.x * va.x;
  s = s + ang.y * va.y;
  s = s + ang.z * va.z;
  return s;
}

__device__ __forceinline__ void jaxpy6(float alpha, f3 wl, f3 wa, f3& vl, f3& va) {
  vl.x = vl.x + alpha * wl.x;
  vl.y = vl.y + alpha * wl.y;
  vl.z = vl.z + alpha * wl.z;
  va.x = va.x + alpha * wa.x;
  va.y = va.y + alpha * wa.y;
  va.z = va.z + alpha * wa.z;
}

__device__ __forceinline__ void jrow_add(JointCtx& c, const JRow& v, float alpha) {
  jaxpy6(alpha, scale(v.d1, c.im1), v.wa1, c.x1l, c.x1a);
  jaxpy6(alpha, scale(v.d2, c.im2), v.wa2, c.x2l, c.x2a);
}

// warmstart_bilateral(_ground) for every row with a nonzero impulse (sor_prox.rs:345-435)
__device__ __forceinline__ void joint_warm_one(const DevBodies& b, const JointBufs& j, int s) {
  JointCtx c;
  load_ctx(b, j, s, c);
  if (!c.active) return;
  for (int k = 0; k < c.nrows; ++k)
    if (c.lam[k] != 0.0f) jrow_add(c, jrow(c, k), c.lam[k]);
  store_ctx(b, j, s, c, false);
}

// solve_bilateral(_ground) with Independent { -max, max } limits (sor_prox.rs:232-343)
__device__ __forceinline__ void joint_sweep_one(const DevBodies& b, const JointBufs& j, int s) {
  JointCtx c;
  load_ctx(b, j, s, c);
  if (!c.active) return;
  for (int k = 0; k < c.nrows; ++k) {
    const JRow v = jrow(c, k);
    const float w = (jdot6(v.d1, v.a1, c.x1l, c.x1a) + jdot6(v.d2, v.a2, c.x2l, c.x2a)) + c.rhs[k];
    const float lo = -FLT_MAX, hi = FLT_MAX;
    const float val = c.lam[k] - c.r[k] * w;
    const float ni = val > lo ? (val < hi ? val : hi) : lo;
    const float dl = ni - c.lam[k];
    c.lam[k] = ni;
    jrow_add(c, v, dl);
  }
  store_ctx(b, j, s, c, true);
}

// One GenericNonlinearConstraint (NonlinearSORProx::solve_generic, nonlinear_sor_prox.rs:85-119,
// clamp_rhs :296-309): when |error| exceeds the allowed error, the bodies move at once by their
// weighted jacobians along error / |error| times -rhs r.
__device__ __forceinline__ void generic_solve(const SimParams& p, bool angular, f3 error, bool d1, bool d2, f3 w1,
                                              f3 w2, f3& t1, q4& q1, f3& t2, q4& q2, const m33& I1, float im1,
                                              const m33& I2, float im2, f3 ml1, f3 ma1, f3 ml2, f3 ma2) {
  const float allowed = angular ? p.allowed_angular_error : p.allowed_linear_error;
  const float depth = sqrtf(dot(error, error));
  if (!(depth > allowed)) return;
  const f3 dir{error.x / depth, error.y / depth, error.z / depth};
  const float a = angular ? (-depth + p.allowed_angular_error) * p.erp : (-depth + p.allowed_linear_error) * p.erp;
  const float lim = angular ? -p.max_angular_correction : -p.max_linear_correction;
  const float rhs = a > lim ? a : lim;
  if (!(rhs < 0.0f)) return;
  const f3 zero{0.0f, 0.0f, 0.0f};
  float inv_r = 0.0f;
  f3 wl1 = zero, wa1 = zero, wl2 = zero, wa2 = zero;
  if (d1) {
    const f3 lin = angular ? zero : cmul(dir, ml1);
    const f3 ang = cmul(angular ? dir : cross(sub(w1, t1), dir), ma1);
    wl1 = scale(lin, im1);
    wa1 = mulv(I1, ang);
    inv_r = inv_r + (im1 + dot(ang, wa1));
  }
  if (d2) {
    const f3 nd = neg(dir);
    const f3 lin = angular ? zero : cmul(nd, ml2);
    const f3 ang = cmul(angular ? nd : cross(sub(w2, t2), nd), ma2);
    wl2 = scale(lin, im2);
    wa2 = mulv(I2, ang);
    inv_r = inv_r + (im2 + dot(ang, wa2));
  }
  const float r = inv_r != 0.0f ? 1.0f / inv_r : 1.0f;
  const float impulse = -rhs * r;
  if (d1) apply_displacement(t1, q1, scale(wl1, impulse), scale(wa1, impulse));
  if (d2) apply_displacement(t2, q2, scale(wl2, impulse), scale(wa2, impulse));
}

// NonlinearConstraintGenerator::position_constraint of one joint, its constraints in order, each at
// the poses the previous one left (ball: translation; fixed: rotation then translation; revolute:
// translation then axis alignment).
__device__ __forceinline__ void joint_position_one(const DevBodies& b, const SimParams& p, const JointBufs& j,
                                                   int s) {
  const float4* R = j.rec + long(s) * kJointRec;
  const int fl = __float_as_int(R[1].w);
  if (!(fl & 1)) return;
  const bool d1 = (fl >> 1) & 1, d2 = (fl >> 2) & 1;
  const int ji = j.order[s];
  const int4 inf = j.info[ji];
  const int kind = inf.x, b1 = inf.y, b2 = inf.z;
  f3 t1 = xyz(b.pos(b1)), t2 = xyz(b.pos(b2));
  q4 q1 = toq(b.rot(b1)), q2 = toq(b.rot(b2));
  float im1 = 0.0f, im2 = 0.0f;
  m33 I1{}, I2{};
  if (d1) I1 = load_inv(b, b1, im1);
  if (d2) I2 = load_inv(b, b2, im2);
  f3 ml1, ma1, ml2, ma2;
  b.mask(b1, ml1, ma1);
  b.mask(b2, ml2, ma2);
  const f3 la1 = xyz(j.a1[ji]), la2 = xyz(j.a2[ji]);
  auto anchors = [&](f3& w1, f3& w2) {
    w1 = add(qrot(q1, la1), t1);
    w2 = add(qrot(q2, la2), t2);
  };
  auto translation = [&]() {
    f3 w1, w2;
    anchors(w1, w2);
    generic_solve(p, false, sub(w2, w1), d1, d2, w1, w2, t1, q1, t2, q2, I1, im1, I2, im2, ml1, ma1, ml2, ma2);
  };
  if (kind == JOINT_FIXED) {
    f3 w1, w2;
    anchors(w1, w2);
    const q4 r1 = qmul(q1, toq(j.f1[ji])), r2 = qmul(q2, toq(j.f2[ji]));
    generic_solve(p, true, q_scaled_axis(qmul(r2, qconj(r1))), d1, d2, w1, w2, t1, q1, t2, q2, I1, im1, I2, im2,
                  ml1, ma1, ml2, ma2);
    translation();
  } else if (kind == JOINT_CYLINDRICAL) {
    // align_axis (helper.rs:701-769), then project_anchor_to_axis (helper.rs:858-912)
    f3 w1, w2;
    anchors(w1, w2);
    const f3 a1 = qrot(q1, xyz(j.f1[ji])), a2 = qrot(q2, xyz(j.f2[ji]));
    generic_solve(p, true, axis_alignment_error(a1, a2), d1, d2, w1, w2, t1, q1, t2, q2, I1, im1, I2, im2, ml1,
                  ma1, ml2, ma2);
    anchors(w1, w2);
    const f3 axis = qrot(q1, xyz(j.f1[ji]));
    const f3 proj = add(w1, scale(axis, dot(axis, sub(w2, w1))));
    generic_solve(p, false, sub(w2, proj), d1, d2, w1, w2, t1, q1, t2, q2, I1, im1, I2, im2, ml1, ma1, ml2, ma2);
  } else if (kind == JOINT_PRISMATIC) {
    // cancel_relative_rotation of the bodies' rotations (helper.rs:553-612), then
    // project_anchor_to_axis (helper.rs:858-912): anchor 2 onto the line through anchor 1 along the axis
    f3 w1, w2;
    anchors(w1, w2);
    generic_solve(p, true, q_scaled_axis(qmul(q2, qconj(q1))), d1, d2, w1, w2, t1, q1, t2, q2, I1, im1, I2, im2, ml1,
                  ma1, ml2, ma2);
    anchors(w1, w2);
    const f3 axis = qrot(q1, xyz(j.f1[ji]));
    const f3 proj = add(w1, scale(axis, dot(axis, sub(w2, w1))));
    generic_solve(p, false, sub(w2, proj), d1, d2, w1, w2, t1, q1, t2, q2, I1, im1, I2, im2, ml1, ma1, ml2, ma2);
  } else if (kind == JOINT_REVOLUTE) {
    translation();
    f3 w1, w2;
    anchors(w1, w2);
    const f3 a1 = qrot(q1, xyz(j.f1[ji])), a2 = qrot(q2, xyz(j.f2[ji]));
    generic_solve(p, true, axis_alignment_error(a1, a2), d1, d2, w1, w2, t1, q1, t2, q2, I1, im1, I2, im2, ml1,
                  ma1, ml2, ma2);
  } else {
    translation();
  }
  if (d1) {
    b.pos(b1) = f4(t1, 0.0f);
    b.rot(b1) = fromq(q1);
  }
  if (d2) {
    b.pos(b2) = f4(t2, 0.0f);
    b.rot(b2) = fromq(q2);
  }
}

enum { kPhaseWarm = 0, kPhaseSweep = 1, kPhasePosition = 2 };

__device__ __forceinline__ void joint_phase_one(int phase, const DevBodies& b, const SimParams& p, const JointBufs& j,
                                                int s) {
  if (phase == kPhaseWarm) joint_warm_one(b, j, s);
  else if (phase == kPhaseSweep) joint_sweep_one(b, j, s);
  else joint_position_one(b, p, j, s);
}

// Every joint colour in order in one workgroup, a barrier between colours: per body the same
// sequence of updates as one launch per colour.
__global__ void __launch_bounds__(kJointBlock) k_joint_run(int phase, DevBodies b, SimParams p, JointBufs j) {
  for (int c = 0; c < j.ncol; ++c) {
    const int2 r = j.crange[c];
    for (int t = threadIdx.x; t < r.y; t += kJointBlock) joint_phase_one(phase, b, p, j, r.x + t);
    __syncthreads();
  }
}

// One joint colour over the grid.
__global__ void __launch_bounds__(64) k_joint_colour(int phase, DevBodies b, SimParams p, JointBufs j, int c) {
  const int2 r = j.crange[c];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < r.y; t += gridDim.x * blockDim.x)
    joint_phase_one(phase, b, p, j, r.x + t);
}

__global__ void __launch_bounds__(256) k_joint_cache(DevBodies b, SimParams p, JointBufs j) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= j.nj) return;
  const float4* R = j.rec + long(s) * kJointRec;
  if (!(__float_as_int(R[1].w) & 1)) return;
  const int ji = j.order[s];
  const int nrows = __float_as_int(R[0].w) & 0xff;
  const float4 l0 = j.lam[2 * long(s)], l1 = j.lam[2 * long(s) + 1];
  // the rows' impulses in order: nlin linear ones, then the angular ones; missing components are
  // zero (the ball joint has no angular impulse, the revolute joint no third angular one, the
  // prismatic and cylindrical joints no third linear one)
  const int nlin = joint_linear_rows(__float_as_int(R[0].w) >> 8);
  const float l[6] = {l0.x, l0.y, l0.z, l0.w, l1.x, l1.y};
  float lv[3] = {0.0f, 0.0f, 0.0f}, av[3] = {0.0f, 0.0f, 0.0f};
  for (int k = 0; k < nrows; ++k) {
    if (k < nlin) lv[k] = l[k];
    else av[k - nlin] = l[k];
  }
  const f3 lin{lv[0], lv[1], lv[2]};
  const f3 ang{av[0], av[1], av[2]};
  j.imp_lin[ji] = f4(lin, 0.0f);
  j.imp_ang[ji] = f4(ang, 0.0f);
  const float inv_dt2 = p.inv_dt * p.inv_dt;
  const float2 br = j.brk[ji];
  if (dot(lin, lin) * inv_dt2 > br.x || dot(ang, ang) * inv_dt2 > br.y) j.state[ji] |= 1;
}

__global__ void __launch_bounds__(256) k_joint_residual(DevBodies b, JointBufs j, double* out) {
  __shared__ double smax[256], ssum[256];
  __shared__ int scnt[256];
  double mx = 0.0, ss = 0.0;
  int cnt = 0;
  for (int s = threadIdx.x; s < j.nj; s += 256) {
    JointCtx c;
    load_ctx(b, j, s, c);
    if (!c.active) continue;
    for (int k = 0; k < c.nrows; ++k) {
      const JRow v = jrow(c, k);
      const float w = (jdot6(v.d1, v.a1, c.x1l, c.x1a) + jdot6(v.d2, v.a2, c.x2l, c.x2a)) + c.rhs[k];
      const double a = fabs(double(w));
      mx = a > mx ? a : mx;
      ss += a * a;
      ++cnt;
    }
  }
  smax[threadIdx.x] = mx;
  ssum[threadIdx.x] = ss;
  scnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) {
      smax[threadIdx.x] = smax[threadIdx.x] > smax[threadIdx.x + o] ? smax[threadIdx.x] : smax[threadIdx.x + o];
      ssum[threadIdx.x] += ssum[threadIdx.x + o];
      scnt[threadIdx.x] += scnt[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = smax[0];
    out[1] = ssum[0];
    out[2] = double(scnt[0]);
  }
}

__device__ inline int uf_load_j(const int* p, int x) {
  return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline int uf_root_j(int* p, int x, int n) {
  for (int it = 0; it < n; ++it) {
    const int y = uf_load_j(p, x);
    if (y == x) break;
    const int z = uf_load_j(p, y);
    if (z != y) __hip_atomic_store(p + x, z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = z;
  }
  return x;
}

// The activation union of k_act_union (k_activation.hip) for the joints' body pairs: only joints
// between two bodies that could sleep are united; one between such a body and a body that keeps its
// island awake marks the former.
__global__ void __launch_bounds__(256) k_joint_union(DevBodies b, JointBufs j, int* parent, uint32_t* touch) {
  const int ji = blockIdx.x * blockDim.x + threadIdx.x;
  if (ji >= j.nj || j.state[ji] != 0) return;
  const int4 inf = j.info[ji];
  const int x = inf.y, y = inf.z;
  if (b.status[x] != STATUS_DYNAMIC || b.status[y] != STATUS_DYNAMIC) return;
  const float2 ax = b.act[x], ay = b.act[y];
  const bool h1 = !(ax.y >= 0.0f && ax.x < ax.y), h2 = !(ay.y >= 0.0f && ay.x < ay.y);
  if (h1 || h2) {
    if (!h1) touch[x] = 1u;
    if (!h2) touch[y] = 1u;
    return;
  }
  const int n = b.n;
  int a = uf_root_j(parent, x, n), d = uf_root_j(parent, y, n);
  for (int it = 0; it < 2 * n && a != d; ++it) {
    if (a < d) {
      const int t = a;
      a = d;
      d = t;
    }
    const int old = atomicCAS(parent + a, a, d);
    if (old == a) break;
    a = uf_root_j(parent, old, n);
    d = uf_root_j(parent, d, n);
  }
}

int nbj(long n) { return int((n + 255) / 256); }

// One workgroup runs every colour when the joints fit a few passes of it.
constexpr int kJointRunMax = 4 * kJointBlock;

void launch_phase(int phase, const DevBodies& b, const SimParams& p, JointBufs& j, hipStream_t s) {
  if (j.nj == 0) return;
  if (j.nj <= kJointRunMax) {
    hipLaunchKernelGGL(k_joint_run, dim3(1), dim3(kJointBlock), 0, s, phase, b, p, j);
    return;
  }
  const int blocks = std::min(4096, (j.nj + 63) / 64);
  for (int c = 0; c < j.ncol; ++c) hipLaunchKernelGGL(k_joint_colour, dim3(blocks), dim3(64), 0, s, phase, b, p, j, c);
}

}  // namespace

void launch_joint_assemble(const DevBodies& b, const SimParams& p, JointBufs& j, hipStream_t s) {
  if (j.nj > 0) hipLaunchKernelGGL(k_joint_assemble, dim3(nbj(j.nj)), dim3(256), 0, s, b, p, j);
}

void launch_joint_warm(const DevBodies& b, JointBufs& j, hipStream_t s) {
  launch_phase(kPhaseWarm, b, SimParams{}, j, s);
}

void launch_joint_sweep(const DevBodies& b, JointBufs& j, hipStream_t s) {
  launch_phase(kPhaseSweep, b, SimParams{}, j, s);
}

void launch_joint_position(const DevBodies& b, const SimParams& p, JointBufs& j, hipStream_t s) {
  launch_phase(kPhasePosition, b, p, j, s);
}

void launch_joint_cache(const DevBodies& b, const SimParams& p, JointBufs& j, hipStream_t s) {
  if (j.nj > 0) hipLaunchKernelGGL(k_joint_cache, dim3(nbj(j.nj)), dim3(256), 0, s, b, p, j);
}

void launch_joint_residual(const DevBodies& b, JointBufs& j, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_joint_residual, dim3(1), dim3(256), 0, s, b, j, out);
}

void launch_joint_union(const DevBodies& b, const JointBufs& j, int* parent, uint32_t* touch, hipStream_t s) {
  if (j.nj > 0) hipLaunchKernelGGL(k_joint_union, dim3(nbj(j.nj)), dim3(256), 0, s, b, j, parent, touch);
}

}  // namespace nph
