// nphysics_amd/csrc/k_position.hip -- coloured nonlinear position correction.
//
// Restates NonlinearSORProx::solve / solve_unilateral / update_contact_constraint / clamp_rhs
// (src/solver/nonlinear_sor_prox.rs:17-55, 121-309): per contact, re-evaluate the contact at the
// current poses (the kinematic of the narrow phase, np_geometry.hpp::reevaluate), dilate by both
// margins (src/solver/signorini_model.rs:174-181), rhs = max((-depth + allowed_linear_error) erp,
// -max_linear_correction); if rhs < 0 move each dynamic body by its weighted jacobian times
// -rhs / (J M^-1 J^T) (RigidBody::apply_displacement, src/object/rigid_body.rs:371-381),
// immediately, so later contacts see the corrected poses. One lane per manifold of one colour
// keeps both poses in registers across the manifold's contacts. Its inputs are laid out by the
// assembly in colour order: per slot the two bodies' half extents and margins, per contact a
// 48-byte position record (local points and normal, contact kind, shape kinds), so a manifold
// reads contiguous runs besides its two bodies' pose and solver blocks.
#include <hip/hip_runtime.h>

#include "np_geometry.hpp"
#include "stages.hpp"

namespace nph {

namespace {

// Poses of a manifold's bodies: in HBM (the per-colour kernels) ...
struct GlobalPose {
  DevBodies b;
  __device__ void load(int i, bool, f3& t, q4& q) const {
    t = xyz(b.pos(i));
    q = toq(b.rot(i));
  }
  __device__ void store(int i, f3 t, q4 q) const {
    b.pos(i) = f4(t, 0.0f);
    b.rot(i) = fromq(q);
  }
};

// ... or in LDS for the scene's dynamic bodies, indexed by their rank in the scene (the
// scene-blocked kernel); a static body's pose is read from HBM, and never written.
struct LdsPose {
  DevBodies b;
  float4* pos;
  float4* rot;
  const int* local;
  __device__ void load(int i, bool dyn, f3& t, q4& q) const {
    if (dyn) {
      const int k = local[i];
      t = xyz(pos[k]);
      q = toq(rot[k]);
    } else {
      t = xyz(b.pos(i));
      q = toq(b.rot(i));
    }
  }
  __device__ void store(int i, f3 t, q4 q) const {
    const int k = local[i];
    pos[k] = f4(t, 0.0f);
    rot[k] = fromq(q);
  }
};

template <bool kMask, class A>
__device__ __forceinline__ void position_one(const A& acc, const DevBodies& b, const SimParams& p,
                                             const SolverBufs& sb, int slot) {
  const int4 q = sb.sinfo[slot];
  const int b1 = q.x & 0x3fffffff, b2 = q.y & 0x3fffffff;
  const bool d1 = (q.x >> 30) & 1;
  const bool d2 = (q.y >> 30) & 1;
  const uint32_t r0 = uint32_t(q.z);
  const int cnt = q.w & 7;
  const float4 g1 = sb.pgeo[2 * long(slot)], g2 = sb.pgeo[2 * long(slot) + 1];
  const f3 h1 = xyz(g1), h2 = xyz(g2);
  const float m1 = g1.w, m2 = g2.w;
  f3 t1, t2;
  q4 q1, q2;
  acc.load(b1, d1, t1, q1);
  acc.load(b2, d2, t2, q2);
  float im1 = 0.0f, im2 = 0.0f;
  m33 I1, I2;
  if (d1) I1 = load_inv(b, b1, im1);
  if (d2) I2 = load_inv(b, b2, im2);
  // jacobian masks (fill_constraint_geometry masks the jacobian, rigid_body.rs:685-705)
  f3 ml1{1.0f, 1.0f, 1.0f}, ma1{1.0f, 1.0f, 1.0f}, ml2{1.0f, 1.0f, 1.0f}, ma2{1.0f, 1.0f, 1.0f};
  if (kMask) {
    b.mask(b1, ml1, ma1);
    b.mask(b2, ml2, ma2);
  }
  // the manifold's kinematics (at most 4 contacts, contiguous position records written by the
  // assembly in colour order) are loaded up front, independent of the poses
  int kind[4];
  f3 lp1[4], lp2[4], ln[4];
  int k1 = 0, k2 = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k < cnt) {
      const float4* P = sb.prec + 3 * long(r0 + k);
      const float4 a = P[0], c = P[1], e = P[2];
      kind[k] = __float_as_int(a.w);
      lp1[k] = xyz(a);
      lp2[k] = xyz(c);
      ln[k] = xyz(e);
      k1 = __float_as_int(c.w);
      k2 = __float_as_int(e.w);
    }
  }
  bool moved1 = false, moved2 = false;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k >= cnt) break;
    const Contact g = reevaluate(kind[k], lp1[k], lp2[k], ln[k], k1, h1, t1, q1, k2, h2, t2, q2);
    const f3 w1 = add(g.w1, scale(g.n, m1));
    const f3 w2 = sub(g.w2, scale(g.n, m2));
    const float depth = (g.depth + m1) + m2;
    const float rhs = fmaxf_((-depth + p.allowed_linear_error) * p.erp, -p.max_linear_correction);
    if (rhs >= 0.0f) continue;
    float inv_r = 0.0f;
    f3 wl1{0.0f, 0.0f, 0.0f}, wa1{0.0f, 0.0f, 0.0f}, wl2{0.0f, 0.0f, 0.0f}, wa2{0.0f, 0.0f, 0.0f};
    if (d1) {
      const f3 dir = neg(g.n);
      f3 ang = cross(sub(w1, t1), dir);
      if (kMask) ang = cmul(ang, ma1);
      wl1 = scale(kMask ? cmul(dir, ml1) : dir, im1);
      wa1 = mulv(I1, ang);
      inv_r = inv_r + (im1 + dot(ang, wa1));
    }
    if (d2) {
      const f3 dir = g.n;
      f3 ang = cross(sub(w2, t2), dir);
      if (kMask) ang = cmul(ang, ma2);
      wl2 = scale(kMask ? cmul(dir, ml2) : dir, im2);
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
  if (moved1) acc.store(b1, t1, q1);
  if (moved2) acc.store(b2, t2, q2);
}

template <bool kMask>
__global__ void __launch_bounds__(64) k_position(DevBodies b, SimParams p, SolverBufs sb, int first_slot,
                                                int nslots) {
  const GlobalPose acc{b};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nslots; t += gridDim.x * blockDim.x)
    position_one<kMask>(acc, b, p, sb, first_slot + t);
}

// Consecutive small classes of the visiting order in one workgroup, a barrier between classes
// (the same corrections in the same order per body as one launch per class).
template <bool kMask>
__global__ void __launch_bounds__(kRunBlock) k_position_run(DevBodies b, SimParams p, SolverBufs sb, int kv0, int nk) {
  const GlobalPose acc{b};
  for (int kv = kv0; kv < kv0 + nk; ++kv) {
    const int2 range = sb.vrange[kv];
    for (int t = threadIdx.x; t < range.y; t += kRunBlock) position_one<kMask>(acc, b, p, sb, range.x + t);
    __syncthreads();
  }
}

// Scene-blocked position correction: one workgroup per scene, the scene's dynamic poses in LDS,
// `iterations` passes over the colours in the visiting order with a barrier between colours.
template <int kBlock, int kCap, bool kMask>
__global__ void __launch_bounds__(kBlock) k_position_blocked(DevBodies b, SimParams p, SolverBufs sb, SceneBlocks sc,
                                                           const int* __restrict__ scene_list, int ncolours,
                                                           int iterations) {
  __shared__ float4 s_pos[kCap];
  __shared__ float4 s_rot[kCap];
  const int s = scene_list[blockIdx.x];
  const int d0 = sc.dyn_start[s], nd = sc.dyn_start[s + 1] - d0;
  for (int i = threadIdx.x; i < nd; i += kBlock) {
    const int body = sc.dyn_list[d0 + i];
    s_pos[i] = b.pos(body);
    s_rot[i] = b.rot(body);
  }
  __syncthreads();
  const LdsPose acc{b, s_pos, s_rot, sc.local};
  const uint32_t* st = sc.start + long(s) * (ncolours + 1);
  for (int it = 0; it < iterations; ++it) {
    for (int k = 0; k < ncolours; ++k) {
      const int col = sb.visit[k];
      const int e = int(st[col + 1]);
      for (int i = int(st[col]) + threadIdx.x; i < e; i += kBlock) position_one<kMask>(acc, b, p, sb, int(sc.slot[i]));
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < nd; i += kBlock) {
    const int body = sc.dyn_list[d0 + i];
    b.pos(body) = s_pos[i];
    b.rot(body) = s_rot[i];
  }
}

}  // namespace

// Sized like the sweep grid: from the largest colour class.
int position_blocks(int max_class) { return sweep_blocks(max_class); }

void launch_position_colour(const DevBodies& b, const DevContacts& c, const SimParams& p, int first, int n,
                            int blocks, SolverBufs& sb, hipStream_t s) {
  (void)c;
  if (b.kmask) hipLaunchKernelGGL(k_position<true>, dim3(blocks), dim3(64), 0, s, b, p, sb, first, n);
  else hipLaunchKernelGGL(k_position<false>, dim3(blocks), dim3(64), 0, s, b, p, sb, first, n);
}

template <bool kMask>
void position_blocked(const DevBodies& b, const SimParams& p, SolverBufs& sb, const SceneBlocks& sc, int ncolours,
                      int iterations, int block, hipStream_t s_small, hipStream_t s_large) {
  if (sc.n_large > 0) {
    if (block == 256)
      hipLaunchKernelGGL((k_position_blocked<256, kLdsBodies, kMask>), dim3(sc.n_large), dim3(256), 0, s_large, b, p,
                         sb, sc, sc.large_list, ncolours, iterations);
    else if (block == 512)
      hipLaunchKernelGGL((k_position_blocked<512, kLdsBodies, kMask>), dim3(sc.n_large), dim3(512), 0, s_large, b, p,
                         sb, sc, sc.large_list, ncolours, iterations);
    else
      hipLaunchKernelGGL((k_position_blocked<1024, kLdsBodies, kMask>), dim3(sc.n_large), dim3(1024), 0, s_large, b,
                         p, sb, sc, sc.large_list, ncolours, iterations);
  }
  if (sc.n_small > 0)
    hipLaunchKernelGGL((k_position_blocked<64, kSmallScene, kMask>), dim3(sc.n_small), dim3(64), 0, s_small, b, p, sb,
                       sc, sc.small_list, ncolours, iterations);
}

void launch_position_blocked(const DevBodies& b, const SimParams& p, SolverBufs& sb, const SceneBlocks& sc,
                             int ncolours, int iterations, int block, hipStream_t s_small, hipStream_t s_large) {
  if (b.kmask) position_blocked<true>(b, p, sb, sc, ncolours, iterations, block, s_small, s_large);
  else position_blocked<false>(b, p, sb, sc, ncolours, iterations, block, s_small, s_large);
}

void launch_position_run(const DevBodies& b, const SimParams& p, SolverBufs& sb, int kv0, int nk, hipStream_t s) {
  if (b.kmask) hipLaunchKernelGGL(k_position_run<true>, dim3(1), dim3(kRunBlock), 0, s, b, p, sb, kv0, nk);
  else hipLaunchKernelGGL(k_position_run<false>, dim3(1), dim3(kRunBlock), 0, s, b, p, sb, kv0, nk);
}

}  // namespace nph
