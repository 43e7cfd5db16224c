// nphysics_amd/csrc/k_narrowphase.hip -- narrow phase: one lane per broad-phase pair.
//
// Replaces GeometricalWorld::perform_narrow_phase (src/world/geometrical_world.rs:315-320) over
// the interaction graph and the manifold collection of MechanicalWorld::step
// (src/world/mechanical_world.rs:287-300, manifolds with at least one contact). Each lane runs the
// contact generator of its pair (np_geometry.hpp) with prediction = sum of both query limits and
// writes up to 4 contacts as 96-byte records into the pair's fixed slots, where they stay; a scan
// then lists the manifolds (pairs with contacts, in pair order) with their first slot and count.
// The contact order (pair order, then generator order) is the manifold list's. HBM bound: 100 B
// per pair + 48 B per contact of algorithmic traffic.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "np_geometry.hpp"
#include "stages.hpp"

namespace nph {

namespace {

// Lanes are binned by pair kind: cuboid-cuboid pairs (SAT + clipping, the expensive one) first,
// run by k_np_boxes with their clip polygons in LDS, all other pairs after, run by k_np_balls.
// Only the lane -> pair mapping changes; every pair still writes its own slots, so the output is
// the unbinned one.
__global__ void __launch_bounds__(256) k_np_flag(DevBodies b, const int2* __restrict__ pairs, int npairs,
                                                uint32_t* flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  const int2 pr = pairs[i];
  flag[i] = (b.shape[pr.x] == SHAPE_CUBOID && b.shape[pr.y] == SHAPE_CUBOID) ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_np_perm(int npairs, const uint32_t* __restrict__ flag,
                                                const uint32_t* __restrict__ before, uint32_t* perm,
                                                uint32_t* nbox) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  const uint32_t nbb = before[npairs - 1] + flag[npairs - 1];
  if (i == 0) *nbox = nbb;
  const uint32_t b = before[i];
  perm[flag[i] ? b : nbb + (uint32_t(i) - b)] = uint32_t(i);
}

// Writes contact k of pair i as a 96-byte record into the pair's slots.
__device__ __forceinline__ void put_contact(const DevContacts& out_c, int i, int k, const Contact& ct) {
  float4* r = out_c.contact(long(i) * kPairSlots + k);
  r[0] = f4(ct.w1, ct.depth);
  r[1] = f4(ct.w2, __uint_as_float(ct.feature));
  r[2] = f4(ct.n, __int_as_float(ct.kind));
  r[3] = f4(ct.lp1, 0.0f);
  r[4] = f4(ct.lp2, 0.0f);
  r[5] = f4(ct.ln, 0.0f);
}

// Cuboid-cuboid pairs (lanes [0, nbox) of the binned order): SAT + clipping, the clip polygons
// and separations of each lane interleaved in LDS.
constexpr int kBoxBlock = 64;
__global__ void __launch_bounds__(kBoxBlock) k_np_boxes(DevBodies b, const int2* __restrict__ pairs, int npairs,
                                                       const uint32_t* __restrict__ perm,
                                                       const uint32_t* __restrict__ nbox, NarrowPhaseBufs np,
                                                       DevContacts out_c) {
  __shared__ ClipVert s_poly[2 * 8 * kBoxBlock];
  __shared__ float s_sep[8 * kBoxBlock];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= int(*nbox)) return;
  const int i = int(perm[j]);
  const int2 pr = pairs[i];
  const int a = pr.x, c = pr.y;
  const float prediction = b.qlim[a] + b.qlim[c];
  const PolyBuf p0{s_poly + threadIdx.x, kBoxBlock}, p1{s_poly + 8 * kBoxBlock + threadIdx.x, kBoxBlock};
  const SepBuf sep{s_sep + threadIdx.x, kBoxBlock};
  const int nc = collide_box_box(xyz(b.pos(a)), toq(b.rot(a)), xyz(b.dims[a]), xyz(b.pos(c)), toq(b.rot(c)),
                                 xyz(b.dims[c]), prediction, p0, p1, sep,
                                 [&](int k, const Contact& ct) { put_contact(out_c, i, k, ct); });
  np.count[i] = uint32_t(nc);
  np.has[i] = nc > 0 ? 1u : 0u;
}

// Pairs with a ball (lanes [nbox, npairs)).
__global__ void __launch_bounds__(256) k_np_balls(DevBodies b, const int2* __restrict__ pairs, int npairs,
                                                 const uint32_t* __restrict__ perm,
                                                 const uint32_t* __restrict__ nbox, NarrowPhaseBufs np,
                                                 DevContacts out_c) {
  const int j = int(*nbox) + blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npairs) return;
  const int i = int(perm[j]);
  const int2 pr = pairs[i];
  const int a = pr.x, c = pr.y;
  const float prediction = b.qlim[a] + b.qlim[c];
  const int nc = collide_ball_pair(b.shape[a], xyz(b.dims[a]), xyz(b.pos(a)), toq(b.rot(a)), b.shape[c],
                                   xyz(b.dims[c]), xyz(b.pos(c)), toq(b.rot(c)), prediction,
                                   [&](int k, const Contact& ct) { put_contact(out_c, i, k, ct); });
  np.count[i] = uint32_t(nc);
  np.has[i] = nc > 0 ? 1u : 0u;
}

// The manifold list: one entry per pair with contacts, in pair order; its contacts stay in the
// pair's slots.
__global__ void __launch_bounds__(256) k_np_manifolds(const int2* __restrict__ pairs, int npairs,
                                                     NarrowPhaseBufs np, DevContacts out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  const uint32_t nc = np.count[i];
  if (i == npairs - 1) {
    np.totals[0] = np.coff[i] + nc;
    np.totals[1] = np.moff[i] + (nc > 0 ? 1u : 0u);
  }
  if (nc == 0) return;
  const uint32_t m = np.moff[i];
  out.mbody[m] = pairs[i];
  out.mfirst[m] = i * kPairSlots;
  out.mcount[m] = int(nc);
}

// Per-body index of the manifold list (sorted by body pair): the first manifold of each body 1.
__global__ void __launch_bounds__(256) k_np_index(int npairs, const uint32_t* __restrict__ totals, DevContacts out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= npairs || m >= int(totals[1])) return;
  const int b = out.mbody[m].x;
  if (m == 0 || out.mbody[m - 1].x != b) out.bfirst[b] = m;
}

__global__ void __launch_bounds__(256) k_gather_contacts(DevContacts c, int M, const uint32_t* __restrict__ moff,
                                                        float4* dense, float4* dense_imp) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int first = c.mfirst[m], cnt = c.mcount[m];
  const long d0 = moff[m];
  for (int k = 0; k < cnt; ++k) {
    const float4* r = c.contact(first + k);
    float4* o = dense + (d0 + k) * kContactFields;
    for (int f = 0; f < kContactFields; ++f) o[f] = r[f];
    if (dense_imp) dense_imp[d0 + k] = c.imp[first + k];
  }
}

}  // namespace

void launch_narrow_phase(const DevBodies& b, const int2* pairs, int npairs, NarrowPhaseBufs& np,
                         DevContacts out, hipStream_t s) {
  if (out.bfirst) fill_u32(reinterpret_cast<uint32_t*>(out.bfirst), 0xffffffffu, size_t(out.bcount), s);
  if (npairs <= 0) {
    fill_u32(np.totals, 0u, 2, s);
    return;
  }
  const unsigned g = unsigned((npairs + 255) / 256);
  // binning by kind: flags in np.has, their exclusive scan in np.moff (both rewritten below)
  hipLaunchKernelGGL(k_np_flag, dim3(g), dim3(256), 0, s, b, pairs, npairs, np.has);
  scan_exclusive_u32(np.has, np.moff, npairs, np.tmp, s);
  uint32_t* nbox = np.totals + 2;
  hipLaunchKernelGGL(k_np_perm, dim3(g), dim3(256), 0, s, npairs, np.has, np.moff, np.perm, nbox);
  hipLaunchKernelGGL(k_np_boxes, dim3(unsigned((npairs + kBoxBlock - 1) / kBoxBlock)), dim3(kBoxBlock), 0, s, b,
                     pairs, npairs, np.perm, nbox, np, out);
  hipLaunchKernelGGL(k_np_balls, dim3(g), dim3(256), 0, s, b, pairs, npairs, np.perm, nbox, np, out);
  scan_exclusive_u32(np.count, np.coff, npairs, np.tmp, s);
  scan_exclusive_u32(np.has, np.moff, npairs, np.tmp, s);
  hipLaunchKernelGGL(k_np_manifolds, dim3(g), dim3(256), 0, s, pairs, npairs, np, out);
  if (out.bfirst) hipLaunchKernelGGL(k_np_index, dim3(g), dim3(256), 0, s, npairs, np.totals, out);
}

void launch_gather_contacts(const DevContacts& c, int M, const uint32_t* moff, float4* dense,
                            float4* dense_imp, hipStream_t s) {
  if (M <= 0) return;
  hipLaunchKernelGGL(k_gather_contacts, dim3(unsigned((M + 255) / 256)), dim3(256), 0, s, c, M, moff, dense,
                     dense_imp);
}

}  // namespace nph
