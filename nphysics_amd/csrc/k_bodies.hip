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
// k_integrate restates MoreauJeanSolver::update_velocities_and_integrate (:328-347) and
// RigidBody::integrate (rigid_body.rs:467-505) with zero damping and unbounded velocities (the
// RigidBodyDesc defaults), i.e. v += ext; v += dv; pose <- displacement(v dt) * pose.
// Both are HBM bound: 112 B read + 160 B written per body for k_prestep, 96 B + 32 B for
// k_integrate.
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
  const float dt = p.dt;
  const m33 R = qmat(toq(q));
  const float4 li = b.local_i[i];
  m33 L;
  L.m[0][0] = li.x; L.m[0][1] = 0.0f; L.m[0][2] = 0.0f;
  L.m[1][0] = 0.0f; L.m[1][1] = li.y; L.m[1][2] = 0.0f;
  L.m[2][0] = 0.0f; L.m[2][1] = 0.0f; L.m[2][2] = li.z;
  const m33 Iw = mmul(mmul(R, L), mtrans(R));
  const f3 w = xyz(b.angvel[i]);
  const f3 iw = mulv(Iw, w);
  float inv_mass;
  m33 inv;
  if (b.act && b.act[i].x == 0.0f) {
    inv = b.inv_inertia(i, inv_mass);
  } else {
    const m33 wx = cross_matrix(scale(w, dt));
    const m33 iwx = cross_matrix(scale(iw, dt));
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
  const f3 acc_ang = mulv(inv, gyro);
  const f3 acc_lin = inv_mass != 0.0f ? f3{p.gx, p.gy, p.gz} : f3{0.0f, 0.0f, 0.0f};
  b.ext_lin[i] = make_float4(dt * acc_lin.x, dt * acc_lin.y, dt * acc_lin.z, 0.0f);
  b.ext_ang[i] = make_float4(dt * acc_ang.x, dt * acc_ang.y, dt * acc_ang.z, 0.0f);
}

__global__ void __launch_bounds__(256) k_integrate(DevBodies b, SimParams p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  if (b.status[i] != STATUS_DYNAMIC) return;
  if (b.act && b.act[i].x == 0.0f) return;  // asleep: not in the solver's island
  const float4 v = b.linvel[i], w = b.angvel[i];
  const float4 el = b.ext_lin[i], ea = b.ext_ang[i];
  f3 dl, da;
  b.load_dv(i, dl, da);
  const f3 lv{(v.x + el.x) + dl.x, (v.y + el.y) + dl.y, (v.z + el.z) + dl.z};
  const f3 av{(w.x + ea.x) + da.x, (w.y + ea.y) + da.y, (w.z + ea.z) + da.z};
  b.linvel[i] = f4(lv, 0.0f);
  b.angvel[i] = f4(av, 0.0f);
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

void launch_integrate(const DevBodies& b, const SimParams& p, hipStream_t s) {
  if (b.n == 0) return;
  hipLaunchKernelGGL(k_integrate, dim3((b.n + 255) / 256), dim3(256), 0, s, b, p);
}

}  // namespace nph
