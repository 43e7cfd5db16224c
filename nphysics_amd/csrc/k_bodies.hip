// nphysics_amd/csrc/k_bodies.hip -- per-body kernels: pre-step dynamics and integration.
//
// k_prestep restates, one thread per body:
//   RigidBody::step_started (position0 = position, src/object/rigid_body.rs:517-519),
//   RigidBody::update_dynamics (:558-588): I_w = R I_l R^T, augmented inertia
//     I_w + [w dt]x I_w - [(I_w w) dt]x, inverse by cofactors,
//   RigidBody::update_acceleration (:590-619): gyroscopic term and gravity,
//   MoreauJeanSolver::assemble_system's ext_vels = dt * acc and the zeroed mj_lambda
//     (src/solver/moreau_jean_solver.rs:166-174, 322-326).
// update_dynamics is evaluated at the start of every step; the reference evaluates it after the
// previous step's solve on the same pose and velocity, so the values are identical. A sleeping
// body keeps the inverse inertia of the last step it moved in, as the reference does (its update
// flags stay clear while it sleeps).
//   with the external forces (acceleration += M^-1 F, :612) and the jacobian mask (:613-615),
//   and gravity only for bodies that have it enabled.
// k_integrate restates MoreauJeanSolver::update_velocities_and_integrate (:328-347) and
// RigidBody::integrate (rigid_body.rs:467-505): v += ext; v += dv; damping v /= 1 + dt d; the
// velocity clamps; pose <- displacement(v dt) * pose; and keeps w as the angular velocity of the
// body's dynamics update (the reference updates the dynamics after the position solve from that
// velocity, mechanical_world.rs:343-346; k_apply_force needs it between steps).
// k_apply_force restates Body::apply_force (rigid_body.rs:771-822) between steps.
// All are HBM bound: 112 B read + 160 B written per body for k_prestep, 96 B + 48 B for
// k_integrate (plus the optional attribute arrays when some body uses them).
#include <hip/hip_runtime.h>

#include "devworld.hpp"
#include "stages.hpp"

namespace nph {

namespace {

__global__ void __launch_bounds__(256) k_prestep(DevBodies b, SimParams p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  const float4 t = b.pos(i);
  const float4 q = b.rot(i);
  b.pos0[i] = t;
  b.rot0[i] = q;
  const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (b.status[i] != STATUS_DYNAMIC) {
    b.set_inv_inertia(i, m33{}, 0.0f);
    b.ext_lin[i] = z4;
    b.ext_ang[i] = z4;
    return;
  }
  // A body woken between steps by a force (no velocity or pose change) has no update flag, so the
  // reference keeps the augmented mass of its last update (body.rs:498-507), made with the angular
  // velocity it had before it slept: w_dyn, marked by k_apply_force. A body the CCD put back at the
  // end of the last step was last updated before that, at the rotation kept in ccd_rot.
  const bool asleep = b.act && b.act[i].x == 0.0f;
  const bool stale = b.act && b.w_dyn[i].w != 0.0f;
  const bool ccd_back = b.ccd_rot && b.w_dyn[i].w == 2.0f;
  body_prestep(b, i, ccd_back ? toq(b.ccd_rot[i]) : toq(q), xyz(b.angvel[i]), asleep, stale,
               stale ? xyz(b.w_dyn[i]) : f3{0.0f, 0.0f, 0.0f}, p.dyn_dt, p.dt, f3{p.gx, p.gy, p.gz});
}

// Augmented inertia I_w + [w dt]x I_w - [(I_w w) dt]x of a body (rigid_body.rs:572-580) at
// rotation q and angular velocity w; `aug` receives it, `inv` its inverse (zero if singular).
__device__ __forceinline__ void augmented_inertia(const DevBodies& b, int i, float4 q, f3 w, float dt, m33& aug,
                                                  m33& inv) {
  const m33 R = qmat(toq(q));
  const float4 li = b.local_i[i];
  m33 L;
  L.m[0][0] = li.x; L.m[0][1] = 0.0f; L.m[0][2] = 0.0f;
  L.m[1][0] = 0.0f; L.m[1][1] = li.y; L.m[1][2] = 0.0f;
  L.m[2][0] = 0.0f; L.m[2][1] = 0.0f; L.m[2][2] = li.z;
  const m33 Iw = mmul(mmul(R, L), mtrans(R));
  const f3 iw = mulv(Iw, w);
  const m33 wx = cross_matrix(scale(w, dt));
  const m33 iwx = cross_matrix(scale(iw, dt));
  const m33 wxI = mmul(wx, Iw);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) aug.m[r][c] = Iw.m[r][c] + (wxI.m[r][c] - iwx.m[r][c]);
  try_inverse(aug, inv);
}

// Body::apply_force for the entries of one body each thread, in the caller's order: entries
// order[off[u] .. off[u+1]) of unique body ubody[u]. forces: 6 floats per entry (linear,
// torque), or with points 3 floats of linear force at a world point (apply_force_at_point).
// Impulses and acceleration changes use the (inverse) augmented mass of the body's last dynamics
// update: rotation rot_dyn (the pose before a set_state of this interval, else the current one)
// and angular velocity w_dyn.
__global__ void __launch_bounds__(256) k_apply_force(DevBodies b, SimParams p, int nu, const int* __restrict__ ubody,
                                                     const int* __restrict__ off, const int* __restrict__ order,
                                                     const float* __restrict__ forces,
                                                     const float* __restrict__ points, int type, int wake,
                                                     const float4* __restrict__ rot_dyn) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nu) return;
  const int i = ubody[u];
  if (b.status[i] != STATUS_DYNAMIC) return;
  f3 ml{1.0f, 1.0f, 1.0f}, ma{1.0f, 1.0f, 1.0f};
  if (b.kmask) b.mask(i, ml, ma);
  const float mass = b.mass[i];
  const float inv_mass = mass == 0.0f ? 0.0f : 1.0f / mass;
  m33 aug, inv;
  if (type == 1 || type == 2) {
    // the rotation of the last dynamics update: before a set_state of this interval, the one kept
    // by the CCD for a body it put back, else the current one; p.dyn_dt: that update's time step
    const float4 rq = rot_dyn ? rot_dyn[i] : ((b.ccd_rot && b.w_dyn[i].w == 2.0f) ? b.ccd_rot[i] : b.rot(i));
    augmented_inertia(b, i, rq, xyz(b.w_dyn[i]), p.dyn_dt, aug, inv);
  }
  const f3 com = xyz(b.pos(i));
  f3 fl_acc = b.f_lin ? xyz(b.f_lin[i]) : f3{0.0f, 0.0f, 0.0f};
  f3 fa_acc = b.f_lin ? xyz(b.f_ang[i]) : f3{0.0f, 0.0f, 0.0f};
  f3 lv = xyz(b.linvel[i]), av = xyz(b.angvel[i]);
  for (int k = off[u]; k < off[u + 1]; ++k) {
    const int e = order[k];
    f3 fl, fa;
    if (points) {  // Force::linear_at_point(f, point - com)
      fl = f3{forces[3 * e], forces[3 * e + 1], forces[3 * e + 2]};
      const f3 pt{points[3 * e], points[3 * e + 1], points[3 * e + 2]};
      fa = cross(sub(pt, com), fl);
    } else {
      fl = f3{forces[6 * e], forces[6 * e + 1], forces[6 * e + 2]};
      fa = f3{forces[6 * e + 3], forces[6 * e + 4], forces[6 * e + 5]};
    }
    if (wake && b.act && b.act[i].y >= 0.0f) {  // Body::activate
      // a sleeping body woken by a force keeps the dynamics of its last update (k_prestep)
      if (b.act[i].x == 0.0f && (type == 0 || type == 2)) b.w_dyn[i].w = 1.0f;
      b.act[i].x = b.act[i].y * 2.0f;
    }
    if (type == 0) {
      fl_acc = add(cmul(fl, ml), fl_acc);
      fa_acc = add(cmul(fa, ma), fa_acc);
    } else if (type == 1) {
      lv = add(cmul(scale(fl, inv_mass), ml), lv);
      av = add(cmul(mulv(inv, fa), ma), av);
    } else if (type == 2) {
      fl_acc = add(cmul(scale(fl, mass), ml), fl_acc);
      fa_acc = add(cmul(mulv(aug, fa), ma), fa_acc);
    } else {
      lv = add(cmul(fl, ml), lv);
      av = add(cmul(fa, ma), av);
    }
  }
  if (type == 0 || type == 2) {
    b.f_lin[i] = f4(fl_acc, 0.0f);
    b.f_ang[i] = f4(fa_acc, 0.0f);
  } else {
    b.linvel[i] = f4(lv, 0.0f);
    b.angvel[i] = f4(av, 0.0f);
    // a changed velocity wakes a sleeping body at the next dynamics update (rigid_body.rs:563-565),
    // which then uses the new velocity; nothing reads the energy before then, so it is set here
    if (b.act) {
      if (b.act[i].x == 0.0f && b.act[i].y >= 0.0f) b.act[i].x = b.act[i].y * 2.0f;
      b.w_dyn[i].w = 0.0f;
    }
  }
}

__global__ void __launch_bounds__(256) k_integrate(DevBodies b, SimParams p, const uint32_t* __restrict__ only) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  if (b.status[i] != STATUS_DYNAMIC) return;
  if (only && !only[i]) return;  // a CCD substep integrates its island only
  if (b.act && b.act[i].x == 0.0f) return;  // asleep: not in the solver's island
  const float4 v = b.linvel[i], w = b.angvel[i];
  const float4 el = b.ext_lin[i], ea = b.ext_ang[i];
  f3 dl, da;
  b.load_dv(i, dl, da);
  f3 lv{(v.x + el.x) + dl.x, (v.y + el.y) + dl.y, (v.z + el.z) + dl.z};
  f3 av{(w.x + ea.x) + da.x, (w.y + ea.y) + da.y, (w.z + ea.z) + da.z};
  if (b.damp) {  // RigidBody::integrate: damping, then the velocity clamps (rigid_body.rs:467-501)
    const float4 d = b.damp[i];
    lv = scale(lv, 1.0f / (1.0f + p.dt * d.x));
    av = scale(av, 1.0f / (1.0f + p.dt * d.y));
    const float ln = sqrtf(dot(lv, lv));
    if (ln > d.z) lv = d.z == 0.0f ? f3{0.0f, 0.0f, 0.0f} : scale(lv, d.z / ln);
    const float an = sqrtf(dot(av, av));
    if (an > d.w) av = d.w == 0.0f ? f3{0.0f, 0.0f, 0.0f} : scale(av, d.w / an);
  }
  b.linvel[i] = f4(lv, 0.0f);
  b.angvel[i] = f4(av, 0.0f);
  b.w_dyn[i] = f4(av, 0.0f);
  f3 t = xyz(b.pos(i));
  q4 q = toq(b.rot(i));
  apply_displacement(t, q, scale(lv, p.dt), scale(av, p.dt));
  b.pos(i) = f4(t, 0.0f);
  b.rot(i) = fromq(q);
}

}  // namespace

void launch_prestep(const DevBodies& b, const SimParams& p, hipStream_t s) {
  if (b.n == 0) return;
  hipLaunchKernelGGL(k_prestep, dim3((b.n + 255) / 256), dim3(256), 0, s, b, p);
}

void launch_integrate(const DevBodies& b, const SimParams& p, hipStream_t s, const uint32_t* only) {
  if (b.n == 0) return;
  hipLaunchKernelGGL(k_integrate, dim3((b.n + 255) / 256), dim3(256), 0, s, b, p, only);
}

// The rotation each body's last dynamics update used (the CCD's kept one for a body it put back).
__global__ void __launch_bounds__(256) k_save_rot_dyn(DevBodies b, float4* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < b.n) out[i] = (b.ccd_rot && b.w_dyn[i].w == 2.0f) ? b.ccd_rot[i] : b.rot(i);
}

void launch_save_rot_dyn(const DevBodies& b, float4* out, hipStream_t s) {
  if (b.n == 0) return;
  hipLaunchKernelGGL(k_save_rot_dyn, dim3((b.n + 255) / 256), dim3(256), 0, s, b, out);
}

// set_state flags every body for a dynamics update from its new state: no body keeps an old one.
__global__ void __launch_bounds__(256) k_clear_dyn_marks(DevBodies b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < b.n) b.w_dyn[i].w = 0.0f;
}

void launch_clear_dyn_marks(const DevBodies& b, hipStream_t s) {
  if (b.n == 0) return;
  hipLaunchKernelGGL(k_clear_dyn_marks, dim3((b.n + 255) / 256), dim3(256), 0, s, b);
}

void launch_apply_force(const DevBodies& b, const SimParams& p, int nu, const int* ubody, const int* off,
                        const int* order, const float* forces, const float* points, int type, int wake,
                        const float4* rot_dyn, hipStream_t s) {
  if (nu == 0) return;
  hipLaunchKernelGGL(k_apply_force, dim3((nu + 255) / 256), dim3(256), 0, s, b, p, nu, ubody, off, order, forces,
                     points, type, wake, rot_dyn);
}

}  // namespace nph
