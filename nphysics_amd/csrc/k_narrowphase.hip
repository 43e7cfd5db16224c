// nphysics_amd/csrc/k_narrowphase.hip -- narrow phase: one lane per broad-phase pair.
//
// Replaces GeometricalWorld::perform_narrow_phase (src/world/geometrical_world.rs:315-320) over
// the interaction graph and the manifold collection of MechanicalWorld::step
// (src/world/mechanical_world.rs:287-300, manifolds with at least one contact). Each lane runs the
// contact generator of its pair (np_geometry.hpp) with prediction = sum of both query limits and
// writes up to 4 contacts into the pair's fixed slots; two scans then compact the slots into the
// contact list (pair order, then generator order) and the manifold list. HBM bound: 100 B per
// pair + 48 B per contact of algorithmic traffic.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "np_geometry.hpp"
#include "stages.hpp"

namespace nph {

namespace {

// Lanes are binned by pair kind so a wavefront runs one contact generator: cuboid-cuboid pairs
// (SAT + clipping, the expensive one) first, all other pairs after. Only the lane -> pair mapping
// changes; every pair still writes its own slots, so the output is the unbinned one.
__global__ void __launch_bounds__(256) k_np_flag(DevBodies b, const int2* __restrict__ pairs, int npairs,
                                                uint32_t* flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  const int2 pr = pairs[i];
  flag[i] = (b.shape[pr.x] == SHAPE_CUBOID && b.shape[pr.y] == SHAPE_CUBOID) ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_np_perm(int npairs, const uint32_t* __restrict__ flag,
                                                const uint32_t* __restrict__ before, uint32_t* perm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  const uint32_t nbb = before[npairs - 1] + flag[npairs - 1];
  const uint32_t b = before[i];
  perm[flag[i] ? b : nbb + (uint32_t(i) - b)] = uint32_t(i);
}

__global__ void __launch_bounds__(256) k_np_pairs(DevBodies b, const int2* __restrict__ pairs,
                                                 int npairs, const uint32_t* __restrict__ perm,
                                                 NarrowPhaseBufs np) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npairs) return;
  const int i = int(perm[j]);
  const int2 pr = pairs[i];
  const int a = pr.x, c = pr.y;
  const float prediction = b.qlim[a] + b.qlim[c];
  Contact out[4];
  const int nc = collide_pair(b.shape[a], xyz(b.dims[a]), xyz(b.pos(a)), toq(b.rot(a)), b.shape[c],
                              xyz(b.dims[c]), xyz(b.pos(c)), toq(b.rot(c)), prediction, out);
  np.count[i] = uint32_t(nc);
  np.has[i] = nc > 0 ? 1u : 0u;
  for (int k = 0; k < nc; ++k) {
    const long s = long(i) * 4 + k;
    np.sw1[s] = f4(out[k].w1, out[k].depth);
    np.sw2[s] = f4(out[k].w2, 0.0f);
    np.sn[s] = f4(out[k].n, 0.0f);
    np.sfeat[s] = out[k].feature;
    np.skind[s] = out[k].kind;
    np.slp1[s] = f4(out[k].lp1, 0.0f);
    np.slp2[s] = f4(out[k].lp2, 0.0f);
    np.sln[s] = f4(out[k].ln, 0.0f);
  }
}

__global__ void __launch_bounds__(256) k_np_compact(const int2* __restrict__ pairs, int npairs,
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
  const uint32_t c0 = np.coff[i];
  out.mbody[m] = pairs[i];
  out.mfirst[m] = int(c0);
  out.mcount[m] = int(nc);
  for (uint32_t k = 0; k < nc; ++k) {
    const long s = long(i) * 4 + k;
    const uint32_t c = c0 + k;
    out.w1[c] = np.sw1[s];
    out.w2[c] = np.sw2[s];
    out.n[c] = np.sn[s];
    out.feature[c] = np.sfeat[s];
    out.kind[c] = np.skind[s];
    out.lp1[c] = np.slp1[s];
    out.lp2[c] = np.slp2[s];
    out.ln[c] = np.sln[s];
  }
}

}  // namespace

void launch_narrow_phase(const DevBodies& b, const int2* pairs, int npairs, NarrowPhaseBufs& np,
                         DevContacts out, hipStream_t s) {
  if (npairs <= 0) {
    fill_u32(np.totals, 0u, 2, s);
    return;
  }
  const unsigned g = unsigned((npairs + 255) / 256);
  // binning by kind: flags in np.has, their exclusive scan in np.moff (both rewritten below)
  hipLaunchKernelGGL(k_np_flag, dim3(g), dim3(256), 0, s, b, pairs, npairs, np.has);
  scan_exclusive_u32(np.has, np.moff, npairs, np.tmp, s);
  hipLaunchKernelGGL(k_np_perm, dim3(g), dim3(256), 0, s, npairs, np.has, np.moff, np.perm);
  hipLaunchKernelGGL(k_np_pairs, dim3(g), dim3(256), 0, s, b, pairs, npairs, np.perm, np);
  scan_exclusive_u32(np.count, np.coff, npairs, np.tmp, s);
  scan_exclusive_u32(np.has, np.moff, npairs, np.tmp, s);
  hipLaunchKernelGGL(k_np_compact, dim3(g), dim3(256), 0, s, pairs, npairs, np, out);
}

}  // namespace nph
