// nphysics_amd/csrc/k_position.hip -- coloured nonlinear position correction.
//
// Restates NonlinearSORProx::solve / solve_unilateral / update_contact_constraint / clamp_rhs
// (src/solver/nonlinear_sor_prox.rs:17-55, 121-309): per contact, re-evaluate the contact at the
// current poses (the kinematic of the narrow phase, np_geometry.hpp::reevaluate), dilate by both
// margins (src/solver/signorini_model.rs:174-181), rhs = max((-depth + allowed_linear_error) erp,
// -max_linear_correction); if rhs < 0 move each dynamic body by its weighted jacobian times
// -rhs / (J M^-1 J^T) (RigidBody::apply_displacement, src/object/rigid_body.rs:371-381),
// immediately, so later contacts see the corrected poses. One lane per manifold of one colour
// keeps both poses in registers across the manifold's contacts.
#include <hip/hip_runtime.h>

#include "np_geometry.hpp"
#include "stages.hpp"

namespace nph {

namespace {

__global__ void __launch_bounds__(256) k_position(DevBodies b, DevContacts c, SimParams p, int first_slot,
                                                 int nslots, SolverBufs sb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  const int m = int(sb.order[first_slot + t]);
  const int2 bb = c.mbody[m];
  const int b1 = bb.x, b2 = bb.y;
  const bool d1 = b.status[b1] == STATUS_DYNAMIC;
  const bool d2 = b.status[b2] == STATUS_DYNAMIC;
  const int k1 = b.shape[b1], k2 = b.shape[b2];
  const f3 h1 = xyz(b.dims[b1]), h2 = xyz(b.dims[b2]);
  const float m1 = b.margin[b1], m2 = b.margin[b2];
  f3 t1 = xyz(b.pos[b1]), t2 = xyz(b.pos[b2]);
  q4 q1 = toq(b.rot[b1]), q2 = toq(b.rot[b2]);
  float im1 = 0.0f, im2 = 0.0f;
  m33 I1, I2;
  if (d1) I1 = load_inv(b, b1, im1);
  if (d2) I2 = load_inv(b, b2, im2);
  const int first = c.mfirst[m], cnt = c.mcount[m];
  bool moved1 = false, moved2 = false;
  for (int k = 0; k < cnt; ++k) {
    const int ci = first + k;
    const Contact g = reevaluate(c.kind[ci], xyz(c.lp1[ci]), xyz(c.lp2[ci]), xyz(c.ln[ci]), k1, h1, t1, q1,
                                 k2, h2, t2, q2);
    const f3 w1 = add(g.w1, scale(g.n, m1));
    const f3 w2 = sub(g.w2, scale(g.n, m2));
    const float depth = (g.depth + m1) + m2;
    const float rhs = fmaxf_((-depth + p.allowed_linear_error) * p.erp, -p.max_linear_correction);
    if (rhs >= 0.0f) continue;
    float inv_r = 0.0f;
    f3 wl1{0.0f, 0.0f, 0.0f}, wa1{0.0f, 0.0f, 0.0f}, wl2{0.0f, 0.0f, 0.0f}, wa2{0.0f, 0.0f, 0.0f};
    if (d1) {
      const f3 dir = neg(g.n);
      const f3 ang = cross(sub(w1, t1), dir);
      wl1 = scale(dir, im1);
      wa1 = mulv(I1, ang);
      inv_r = inv_r + (im1 + dot(ang, wa1));
    }
    if (d2) {
      const f3 dir = g.n;
      const f3 ang = cross(sub(w2, t2), dir);
      wl2 = scale(dir, im2);
      wa2 = mulv(I2, ang);
      inv_r = inv_r + (im2 + dot(ang, wa2));
    }
    if (inv_r == 0.0f) continue;
    const float r = 1.0f / inv_r;
    const float impulse = -rhs * r;
    if (d1) {
      apply_displacement(t1, q1, scale(wl1, impulse), scale(wa1, impulse));
      moved1 = true;
    }
    if (d2) {
      apply_displacement(t2, q2, scale(wl2, impulse), scale(wa2, impulse));
      moved2 = true;
    }
  }
  if (moved1) {
    b.pos[b1] = f4(t1, 0.0f);
    b.rot[b1] = fromq(q1);
  }
  if (moved2) {
    b.pos[b2] = f4(t2, 0.0f);
    b.rot[b2] = fromq(q2);
  }
}

}  // namespace

void launch_position_colour(const DevBodies& b, const DevContacts& c, const SimParams& p,
                            int first_slot, int nslots, SolverBufs& sb, hipStream_t s) {
  if (nslots <= 0) return;
  hipLaunchKernelGGL(k_position, dim3(unsigned((nslots + 255) / 256)), dim3(256), 0, s, b, c, p, first_slot,
                     nslots, sb);
}

}  // namespace nph
