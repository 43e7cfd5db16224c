// nphysics_amd/csrc/devworld.hpp -- device-side views of one world, passed to kernels by value.
//
// HBM layout (structure of arrays, one float4 per 3-vector so every access is a 16-byte load):
//   bodies   pos/rot/pos0/rot0/linvel/angvel, inverse augmented inertia as three rows with the
//            inverse mass in row 0's w, external velocity increment dt*acc (lin, ang), solver
//            velocity increment dv (lin, ang), loose AABB (mn, mx), static shape data.
//   contacts per contact: world1 (w = depth), world2, normal, feature, kinematic kind, local
//            points/normal for the position solver, solved impulse; per manifold: bodies,
//            first contact, count.
//   rows     per contact in colour order: 8 float4 (see k_solver.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "devmath.hpp"

namespace nph {

struct SimParams {
  float dt, inv_dt, erp, warmstart_coeff, restitution_threshold;
  float allowed_linear_error, max_linear_correction;
  float gx, gy, gz;
};

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
  float4* pos;
  float4* rot;
  float4* pos0;
  float4* rot0;
  float4* linvel;
  float4* angvel;
  float4* ia0;
  float4* ia1;
  float4* ia2;
  float4* ext_lin;
  float4* ext_ang;
  float4* dv_lin;
  float4* dv_ang;
  float4* loose_mn;
  float4* loose_mx;
};

struct DevContacts {
  int2* mbody;       // (body1, body2) per manifold, sorted
  int* mfirst;
  int* mcount;
  float4* w1;        // xyz world point on collider 1, w = depth
  float4* w2;
  float4* n;
  uint32_t* feature;
  int* kind;
  float4* lp1;
  float4* lp2;
  float4* ln;
  float4* imp;       // (normal, tangent 1, tangent 2) impulses after the solve
};

NPH_HD m33 load_inv(const DevBodies& b, int i, float& inv_mass) {
  const float4 r0 = b.ia0[i], r1 = b.ia1[i], r2 = b.ia2[i];
  m33 m;
  m.m[0][0] = r0.x; m.m[0][1] = r0.y; m.m[0][2] = r0.z;
  m.m[1][0] = r1.x; m.m[1][1] = r1.y; m.m[1][2] = r1.z;
  m.m[2][0] = r2.x; m.m[2][1] = r2.y; m.m[2][2] = r2.z;
  inv_mass = r0.w;
  return m;
}

}  // namespace nph
