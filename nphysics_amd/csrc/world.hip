// nphysics_amd/csrc/world.hip -- host orchestration of one MechanicalWorld::step on one GPU and
// the C ABI of include/nphys.h.
//
// Step order (src/world/mechanical_world.rs:199-396):
//   1. collision detection on the current poses when bodies were added or moved by the caller
//      (:250-253);
//   2. per-body dynamics, accelerations, ext = dt acc, dv = 0 (k_prestep; :230-243 and
//      moreau_jean_solver.rs:137-174);
//   3. MoreauJeanSolver::step (moreau_jean_solver.rs:47-90): colouring, assembly with warm start,
//      coloured SOR-Prox sweeps, residual, impulse cache, integration, coloured position
//      correction;
//   4. the CCD-phase sync + broad phase over swept AABBs (:604-606), the final sync (a no-op for
//      the loose AABBs, checked anyway) and the broad + narrow phase that produce the next step's
//      manifolds (:375-385).
// The host reads back a handful of counters per step (loose-AABB changes, pair/contact totals,
// colouring progress, colour offsets, residual partials); everything else stays on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <initializer_list>
#include <limits>
#include <map>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nphys.h"
#include "common.hpp"
#include "devmath.hpp"
#include "kernels.hpp"
#include "mlp.hpp"
#include "stages.hpp"

namespace nph {

namespace {

// Manifold list plus the contact slots of every broad-phase pair (kPairSlots per pair).
struct ContactBufs {
  DBuf mbody, mfirst, mcount, rec, imp, bfirst;
  int M = 0, C = 0, nb = 0;  // nb: bodies covered by bfirst
  hipError_t ensure(size_t mcap, size_t slots) {
    hipError_t e;
    if ((e = mbody.ensure(mcap * sizeof(int2))) != hipSuccess) return e;
    if ((e = mfirst.ensure(mcap * 4)) != hipSuccess) return e;
    if ((e = mcount.ensure(mcap * 4)) != hipSuccess) return e;
    if ((e = rec.ensure(slots * kContactFields * sizeof(float4))) != hipSuccess) return e;
    if ((e = imp.ensure(slots * sizeof(float4))) != hipSuccess) return e;
    return hipSuccess;
  }
  DevContacts view() const {
    DevContacts d;
    d.mbody = mbody.as<int2>();
    d.mfirst = mfirst.as<int>();
    d.mcount = mcount.as<int>();
    d.rec = rec.as<float4>();
    d.imp = imp.as<float4>();
    d.bfirst = nb > 0 ? bfirst.as<int>() : nullptr;
    d.bcount = nb;
    return d;
  }
};

enum StageEv { EV_START, EV_PRE_CD, EV_ASM, EV_VEL, EV_INT, EV_POS, EV_CCD, EV_BP, EV_NP, EV_COUNT };

// Small device values to pinned host memory without a stream synchronisation: one workgroup copies
// up to four word ranges into the host buffer, fences at system scope and then writes the
// sequence number the host waits for. Launched last in stream order, so the flag also means that
// everything enqueued before it has completed.
struct RbJob {
  const uint32_t* src[4];
  int n[4];
  int dst[4];
};

__global__ void __launch_bounds__(256) k_readback(RbJob j, uint32_t* host, int flag, uint32_t seq) {
  for (int s = 0; s < 4; ++s)
    for (int i = threadIdx.x; i < j.n[s]; i += blockDim.x) host[j.dst[s] + i] = j.src[s][i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(host + flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace

struct World {
  hipStream_t stream = nullptr;
  // the scene-blocked solve runs its large scenes on stream_large beside the small ones on stream
  // (fork and join by events)
  hipStream_t stream_large = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  int device = 0;
  nphys_params prm{};
  SimParams sp{};
  int32_t contact_model = NPHYS_CONTACT_MODEL_SIGNORINI_COULOMB_PYRAMID;

  // host copies of the per-body static attributes
  std::vector<int> h_status, h_shape, h_scene;
  int scene_bits = 0;  // bits of the scene index above the Morton code (0: one scene)
  std::vector<float4> h_dims, h_local_i;
  std::vector<float> h_margin, h_qlim, h_friction, h_restitution, h_mass;
  int n = 0;

  // device body arrays
  DBuf status, shape, scene, dims, margin, qlim, friction, restitution, mass, local_i;
  DBuf pose;  // kPoseBlock float4 per body: position, rotation
  DBuf sol;   // kSolverBlock float4 per body: inverse inertia and mass, solver velocity increment
  DBuf pos0, rot0, linvel, angvel, ext_lin, ext_ang;
  DBuf loose_mn, loose_mx;
  DBuf act;  // float2 per body: activation energy (0: asleep), sleep threshold (< 0: never sleeps)
  // Optional per-body attributes (damping and clamps, flags, material ids, surface velocities);
  // each is passed to the kernels only when some body differs from the default.
  std::vector<float4> h_damp, h_surf;
  std::vector<uint4> h_groups;
  std::vector<int> h_flags, h_matid;
  bool any_damp = false, any_flags = false, any_mask = false, any_matid = false, any_surf = false;
  bool any_groups = false;
  DBuf damp, bflags, matid, surf, groups;
  // External forces accumulated for the next step (cleared after it), the angular velocity of the
  // last dynamics update, and the rotations from before a set_state of the current interval.
  DBuf f_lin, f_ang, w_dyn, rot_dyn, af_tmp;
  bool forces_pending = false, rot_dyn_saved = false;
  // material coefficient table (sorted keys min id << 32 | max id), friction and restitution
  std::vector<std::pair<unsigned long long, float>> mt_host[2];
  DBuf mt_key[2], mt_val[2];
  // Sleeping: on once a dynamic body with a threshold exists; a wake of every sleeping body is
  // due at the next step after the caller set the bodies' state.
  bool sleeping = false, wake_pending = false;
  int64_t n_dynamic = 0, active_bodies = 0;
  DBuf act_parent, act_can, act_touch, act_pos, act_tmp, act_mflag[2], sl_mbody[2], sl_mfirst[2], sl_mcount[2];

  // broad phase
  DBuf bp_dirty, bp_leaf_of, bp_list, bp_pnode;
  // Incremental pair updates: the tree and segments of the last pair build are valid, how many
  // incremental updates followed the last full build, and the loose-AABB replacements (all, of
  // non-dynamic bodies) since the last pair build.
  bool tree_valid = false;
  bool incremental_bp = true;
  int inc_since_full = 0;
  int64_t dirty_pending = 0, dirty_static_pending = 0;
  static constexpr int kMaxIncremental = 32;
  DBuf bp_small, keys, vals, keys_alt, vals_alt, child, parent, visit, node_mn, node_mx, node_scene, cand,
      cand_count, cand_off, pairs, bp_tmp;
  int cap = 32;
  int npairs = 0;
  bool pairs_dirty = true;
  bool moved = true;
  bool recollide = false;  // collision groups changed: the next step starts with collision detection

  // narrow phase
  DBuf np_count, np_coff, np_has, np_moff, np_perm, np_tmp;
  DBuf rd_moff, rd_rec, rd_imp;  // contact readback staging
  ContactBufs cb[2];
  int cur = 0;      // index of the manifolds produced by the last collision detection
  bool have_solved = false;

  // solver
  // The manifolds handed to the solver with cb[k]'s contacts: all of cb[k], or (sleeping) the
  // compacted list of those with an awake dynamic side in sl_*[k].
  DevContacts sview[2] = {};
  int sview_M[2] = {0, 0}, sview_C[2] = {0, 0};
  bool sview_compact[2] = {false, false};
  DBuf sv_state2[2];  // colouring state of the manifolds of sview[k]
  DBuf sv_deg, sv_adj_off, sv_adj_fill, sv_adj, sv_ckeys, sv_order, sv_ckeys_alt,
      sv_order_alt, sv_col_start, sv_mcount_ord, sv_soff, sv_slot_of, sv_sinfo, sv_rec2c, sv_rec, sv_lam, sv_prec, sv_pgeo, sv_red, sv_tmp;
  int ncolours = 0;
  long rec_stride = 0;
  int colour_rounds = 0;
  static constexpr int kRoundBatch = 4;  // colouring rounds launched between polls
  std::vector<int> colour_visit;  // colour classes in visiting order
  long solve_count = 0;    // solves performed, never reset
  DBuf sv_visit;           // device copy of colour_visit, then the classes' slot ranges in that order
  int* h_visit = nullptr;  // pinned staging of the same
  int solved_M = 0, solved_C = 0;
  double res_inf = 0.0, res_rms = 0.0;
  double* red_pinned = nullptr;  // deferred residual partials (pinned)
  size_t red_pinned_cap = 0;
  int res_pending_blocks = -1, res_pending_C = 0;
  // solve-to-residual mode (off when res_target <= 0): see nphys_set_residual_target
  double res_target = 0.0;
  int res_max_sweeps = 500, res_check_every = 1, res_norm = NPHYS_RESIDUAL_MAX;
  int last_sweeps = 0;
  int max_class = 0;             // largest colour class of the current solve
  std::vector<double> red_host;
  // learned warm start: network (FP32 parameters + BF16 images), per-record features, impulse
  // units, predictions and recorded targets; the judge's saved plain solve
  // scene-blocked solve: per-scene dynamic-body lists (host-built at insertion), the per-solve
  // scene-major slot order, and the variant chosen by the autotuner (0 per-colour launches,
  // 256/512/1024 the blocked kernels' workgroup size for the scenes of more than 256 dynamic bodies)
  int nscenes = 1, max_scene_dyn = 0;
  DBuf sc_dyn_start, sc_dyn_list, sc_local, sc_keys, sc_slot, sc_keys_alt, sc_slot_alt, sc_start, sc_small, sc_large;
  int n_small_scenes = 0, n_large_scenes = 0;
  int variant_request = -1;  // -1: autotune; else forced
  int cur_variant = 0;
  static constexpr int kTuneReps = 3;
  std::vector<int> tune_cands;
  std::vector<double> tune_time;
  std::vector<int> tune_count;
  bool tuned = false, tune_pending = false;
  int tune_M = 0;  // manifold count when the current pick was timed
  hipEvent_t tune_ev[2] = {nullptr, nullptr};
  bool ml_loaded = false, capture = false, judge = false;
  long samples = 0;
  DBuf ml_params, ml_img, ml_X, ml_ls, ml_y, ml_t, jd_lam, jd_lam0, jd_dv;
  std::vector<int> last_colours;
  bool keep_colours = false;

  // joints: static per-joint arrays (host copies in joint order), their greedy colouring with the
  // joints listed colour by colour, per-solve row records and impulses, and the judge's saved ones
  std::vector<int4> h_jinfo;
  std::vector<float4> h_ja1, h_ja2, h_jf1, h_jf2;
  std::vector<float2> h_jbrk;
  std::vector<int> h_jorder;
  std::vector<int2> h_jcrange;
  int nj = 0, njcol = 0;
  DBuf j_info, j_a1, j_a2, j_f1, j_f2, j_brk, j_imp_lin, j_imp_ang, j_state, j_order, j_crange, j_rec, j_lam, jd_jlam;

  // CCD, once some body is CCD-enabled: per-body flags (bit 0 ccd_enabled, bit 1 linear motion
  // interpolation) and keys (a static collider that shares a static body takes the index of that
  // body's first collider), the CcdBufs arrays, the scratch manifolds of the island rounds, and the
  // last prediction's results
  bool any_ccd = false;
  std::vector<int> h_ccd, h_ccd_key;
  std::map<std::pair<int, int>, int> static_owner;  // (scene, static body index) -> first collider
  DBuf ccd_flags, ccd_keys, ccd_rot, ccd_time, ccd_frozen, ccd_slot, ccd_heap, ccd_stage, ccd_imp, ccd_ctl,
      ccd_in, ccd_new, ccd_prep, ccd_prepared, ccd_eval, ccd_kept, ccd_cand, ccd_cpos, ccd_list, ccd_lpair, ccd_tmp,
      ccd_deg, ccd_off, ccd_fill, ccd_adj, ccd_pos, ccd_sbodies;
  ContactBufs ccd_cb;
  int ccd_cap = 0, ccd_nimp = 0;
  float ccd_last_toi = 0.0f, ccd_min_over = 0.0f;
  bool ccd_sub = false;        // a CCD substep is being solved: no joint rows, per-colour launches
  float dyn_dt_next = 0.0f;    // time step of the bodies' last dynamics update (a substep's after one)

  // Readback buffer (pinned, device-mapped, coherent): words [0, 64) small per-step values,
  // [kRbClass, +2 (kMaxColours + 2)) the class offsets, kRbFlag the sequence flag.
  static constexpr int kRbClass = 64, kRbFlag = kRbClass + 2 * (kMaxColours + 2) + 64, kRbWords = kRbFlag + 64;
  uint32_t* rb_host = nullptr;
  uint32_t* rb_dev = nullptr;
  uint32_t rb_seq = 0, class_seq = 0;
  struct RbSeg {
    const uint32_t* src;
    int n;
    int dst;
  };
  hipEvent_t ev[EV_COUNT] = {};
  nphys_counters counters{};
  bool counters_enabled = false;
  std::vector<hipEvent_t> kt_ev;  // event pairs around sweep-kernel launches
  int kt_used = 0;

  ~World() {
    if (h_visit) (void)hipHostFree(h_visit);
    if (red_pinned) (void)hipHostFree(red_pinned);
    if (rb_host) (void)hipHostFree(rb_host);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : kt_ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : tune_ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : ml_ev)
      if (e) (void)hipEventDestroy(e);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    if (stream_large) (void)hipStreamDestroy(stream_large);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void set_params(const nphys_params& p) {
    prm = p;
    sp.dt = float(p.dt);
    sp.inv_dt = sp.dt == 0.0f ? 0.0f : 1.0f / sp.dt;
    sp.erp = float(p.erp);
    sp.warmstart_coeff = float(p.warmstart_coeff);
    sp.restitution_threshold = float(p.restitution_velocity_threshold);
    sp.allowed_linear_error = float(p.allowed_linear_error);
    sp.max_linear_correction = float(p.max_linear_correction);
    sp.allowed_angular_error = float(p.allowed_angular_error);
    sp.max_angular_correction = float(p.max_angular_correction);
    sp.gx = float(p.gravity[0]);
    sp.gy = float(p.gravity[1]);
    sp.gz = float(p.gravity[2]);
    sp.dyn_dt = sp.dt;
    dyn_dt_next = sp.dt;
  }

  int init(const nphys_params& p) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      return set_err(NPHYS_ERR_NO_DEVICE, "no HIP device");
    HIPCK(hipGetDevice(&device));
    HIPCK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCK(hipStreamCreateWithFlags(&stream_large, hipStreamNonBlocking));
    HIPCK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    HIPCK(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming));
    HIPCK(hipHostMalloc(reinterpret_cast<void**>(&rb_host), kRbWords * sizeof(uint32_t),
                        hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(rb_host, 0, kRbWords * sizeof(uint32_t));
    HIPCK(hipHostGetDevicePointer(reinterpret_cast<void**>(&rb_dev), rb_host, 0));
    HIPCK(hipHostMalloc(reinterpret_cast<void**>(&h_visit), 3 * (kMaxColours + 2) * sizeof(int)));
    HIPCK(sv_visit.ensure(3 * (kMaxColours + 2) * sizeof(int)));
    for (auto& e : ev) HIPCK(hipEventCreate(&e));
    for (auto& e : tune_ev) HIPCK(hipEventCreate(&e));
    set_params(p);
    HIPCK(bp_small.ensure(64 * sizeof(uint32_t)));
    return NPHYS_OK;
  }

  DevBodies bodies() const {
    DevBodies b;
    b.n = n;
    b.status = status.as<int>();
    b.shape = shape.as<int>();
    b.scene = scene.as<int>();
    b.dims = dims.as<float4>();
    b.margin = margin.as<float>();
    b.qlim = qlim.as<float>();
    b.friction = friction.as<float>();
    b.restitution = restitution.as<float>();
    b.mass = mass.as<float>();
    b.local_i = local_i.as<float4>();
    b.pose = pose.as<float4>();
    b.sol = sol.as<float4>();
    b.pos0 = pos0.as<float4>();
    b.rot0 = rot0.as<float4>();
    b.linvel = linvel.as<float4>();
    b.angvel = angvel.as<float4>();
    b.ext_lin = ext_lin.as<float4>();
    b.ext_ang = ext_ang.as<float4>();
    b.loose_mn = loose_mn.as<float4>();
    b.loose_mx = loose_mx.as<float4>();
    b.act = sleeping ? act.as<float2>() : nullptr;
    b.damp = any_damp ? damp.as<float4>() : nullptr;
    b.flags = any_flags ? bflags.as<int>() : nullptr;
    b.kmask = any_mask ? bflags.as<int>() : nullptr;
    b.mat_id = any_matid ? matid.as<int>() : nullptr;
    b.surf_vel = any_surf ? surf.as<float4>() : nullptr;
    b.f_lin = forces_pending ? f_lin.as<float4>() : nullptr;
    b.f_ang = forces_pending ? f_ang.as<float4>() : nullptr;
    b.w_dyn = w_dyn.as<float4>();
    b.ccd = any_ccd ? ccd_flags.as<int>() : nullptr;
    b.ccd_key = any_ccd ? ccd_keys.as<int>() : nullptr;
    b.ccd_rot = any_ccd ? ccd_rot.as<float4>() : nullptr;
    b.groups = any_groups ? groups.as<uint4>() : nullptr;
    return b;
  }

  MatTable mat_table() const {
    MatTable t;
    t.fkey = mt_key[0].as<unsigned long long>();
    t.fval = mt_val[0].as<float>();
    t.nf = int(mt_host[0].size());
    t.rkey = mt_key[1].as<unsigned long long>();
    t.rval = mt_val[1].as<float>();
    t.nr = int(mt_host[1].size());
    return t;
  }

  // ------------------------------------------------------------------ scene
  JointBufs joint_view() {
    JointBufs j;
    j.nj = nj;
    j.ncol = njcol;
    j.info = j_info.as<int4>();
    j.a1 = j_a1.as<float4>();
    j.a2 = j_a2.as<float4>();
    j.f1 = j_f1.as<float4>();
    j.f2 = j_f2.as<float4>();
    j.brk = j_brk.as<float2>();
    j.imp_lin = j_imp_lin.as<float4>();
    j.imp_ang = j_imp_ang.as<float4>();
    j.state = j_state.as<int>();
    j.order = j_order.as<int>();
    j.crange = j_crange.as<int2>();
    j.rec = j_rec.as<float4>();
    j.lam = j_lam.as<float4>();
    return j;
  }

  // Greedy colouring of the joints in joint order: a joint takes the smallest colour that no
  // earlier joint sharing one of its dynamic bodies has (the CPU restatement colours the same way).
  void colour_joints() {
    std::vector<std::vector<int>> used(size_t(std::max(n, 1)));
    njcol = 0;
    for (int ji = 0; ji < nj; ++ji) {
      int4& inf = h_jinfo[size_t(ji)];
      std::vector<char> taken;
      for (int bi : {inf.y, inf.z}) {
        if (h_status[size_t(bi)] != NPHYS_STATUS_DYNAMIC) continue;
        for (int c : used[size_t(bi)]) {
          if (int(taken.size()) <= c) taken.resize(size_t(c) + 1, 0);
          taken[size_t(c)] = 1;
        }
      }
      int c = 0;
      while (c < int(taken.size()) && taken[size_t(c)]) ++c;
      inf.w = c;
      njcol = std::max(njcol, c + 1);
      for (int bi : {inf.y, inf.z})
        if (h_status[size_t(bi)] == NPHYS_STATUS_DYNAMIC) used[size_t(bi)].push_back(c);
    }
    h_jorder.clear();
    h_jcrange.assign(size_t(njcol), int2{0, 0});
    for (int c = 0; c < njcol; ++c) {
      h_jcrange[size_t(c)].x = int(h_jorder.size());
      for (int ji = 0; ji < nj; ++ji)
        if (h_jinfo[size_t(ji)].w == c) h_jorder.push_back(ji);
      h_jcrange[size_t(c)].y = int(h_jorder.size()) - h_jcrange[size_t(c)].x;
    }
  }

  // JointConstraintSet::insert for every joint of the descriptor (ball_constraint.rs:27-48,
  // fixed_constraint.rs:28-58, revolute_constraint.rs:37-72).
  int add_joints(const nphys_joint_desc& d, int32_t* first) {
    if (d.n < 0 || (d.n > 0 && (!d.kind || !d.body1 || !d.body2 || !d.anchor1 || !d.anchor2)))
      return set_err(NPHYS_ERR_INVALID, "bad joint descriptor");
    for (int i = 0; i < d.n; ++i) {
      const int b1 = d.body1[i], b2 = d.body2[i];
      if (b1 < 0 || b1 >= n || b2 < 0 || b2 >= n || b1 == b2) return set_err(NPHYS_ERR_INVALID, "bad joint body");
      if (d.kind[i] < NPHYS_JOINT_BALL || d.kind[i] > NPHYS_JOINT_CYLINDRICAL)
        return set_err(NPHYS_ERR_INVALID, "bad joint kind");
      if (h_status[size_t(b1)] != NPHYS_STATUS_DYNAMIC && h_status[size_t(b2)] != NPHYS_STATUS_DYNAMIC)
        return set_err(NPHYS_ERR_INVALID, "a joint needs a dynamic body");
    }
    if (first) *first = nj;
    if (d.n == 0) return NPHYS_OK;
    HIPCK(hipStreamSynchronize(stream));
    const int old = nj;
    for (int i = 0; i < d.n; ++i) {
      h_jinfo.push_back(make_int4(d.kind[i], d.body1[i], d.body2[i], 0));
      h_ja1.push_back(make_float4(d.anchor1[3 * i], d.anchor1[3 * i + 1], d.anchor1[3 * i + 2], 0.0f));
      h_ja2.push_back(make_float4(d.anchor2[3 * i], d.anchor2[3 * i + 1], d.anchor2[3 * i + 2], 0.0f));
      float4 f1 = make_float4(0.0f, 0.0f, 0.0f, 1.0f), f2 = f1;
      if (d.kind[i] == NPHYS_JOINT_FIXED) {
        if (d.frame1) f1 = make_float4(d.frame1[4 * i], d.frame1[4 * i + 1], d.frame1[4 * i + 2], d.frame1[4 * i + 3]);
        if (d.frame2) f2 = make_float4(d.frame2[4 * i], d.frame2[4 * i + 1], d.frame2[4 * i + 2], d.frame2[4 * i + 3]);
      } else if (d.kind[i] == NPHYS_JOINT_PRISMATIC) {
        f1 = d.axis1 ? make_float4(d.axis1[3 * i], d.axis1[3 * i + 1], d.axis1[3 * i + 2], 0.0f)
                     : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
      } else if (d.kind[i] == NPHYS_JOINT_REVOLUTE || d.kind[i] == NPHYS_JOINT_CYLINDRICAL) {
        f1 = d.axis1 ? make_float4(d.axis1[3 * i], d.axis1[3 * i + 1], d.axis1[3 * i + 2], 0.0f)
                     : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
        f2 = d.axis2 ? make_float4(d.axis2[3 * i], d.axis2[3 * i + 1], d.axis2[3 * i + 2], 0.0f)
                     : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
      }
      h_jf1.push_back(f1);
      h_jf2.push_back(f2);
      const float bf = d.break_force ? d.break_force[i] : -1.0f;
      const float bt = d.break_torque ? d.break_torque[i] : -1.0f;
      h_jbrk.push_back(make_float2(bf >= 0.0f ? bf * bf : FLT_MAX, bt >= 0.0f ? bt * bt : FLT_MAX));
    }
    nj += d.n;
    colour_joints();
    const size_t J = size_t(nj);
    HIPCK(j_info.ensure(J * sizeof(int4)));
    HIPCK(j_a1.ensure(J * sizeof(float4)));
    HIPCK(j_a2.ensure(J * sizeof(float4)));
    HIPCK(j_f1.ensure(J * sizeof(float4)));
    HIPCK(j_f2.ensure(J * sizeof(float4)));
    HIPCK(j_brk.ensure(J * sizeof(float2)));
    HIPCK(j_order.ensure(J * sizeof(int)));
    HIPCK(j_crange.ensure(size_t(std::max(njcol, 1)) * sizeof(int2)));
    HIPCK(j_rec.ensure(J * kJointRec * sizeof(float4)));
    HIPCK(j_lam.ensure(J * 2 * sizeof(float4)));
    // cached impulses and broken / removed flags live on the device: kept, the new ones cleared
    HIPCK(j_imp_lin.ensure(J * sizeof(float4), true, stream));
    HIPCK(j_imp_ang.ensure(J * sizeof(float4), true, stream));
    HIPCK(j_state.ensure(J * sizeof(int), true, stream));
    const size_t add = size_t(d.n);
    HIPCK(hipMemsetAsync(j_imp_lin.as<float4>() + old, 0, add * sizeof(float4), stream));
    HIPCK(hipMemsetAsync(j_imp_ang.as<float4>() + old, 0, add * sizeof(float4), stream));
    HIPCK(hipMemsetAsync(j_state.as<int>() + old, 0, add * sizeof(int), stream));
    HIPCK(hipMemcpyAsync(j_info.p, h_jinfo.data(), J * sizeof(int4), hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(j_a1.p, h_ja1.data(), J * sizeof(float4), hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(j_a2.p, h_ja2.data(), J * sizeof(float4), hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(j_f1.p, h_jf1.data(), J * sizeof(float4), hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(j_f2.p, h_jf2.data(), J * sizeof(float4), hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(j_brk.p, h_jbrk.data(), J * sizeof(float2), hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(j_order.p, h_jorder.data(), J * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(j_crange.p, h_jcrange.data(), size_t(njcol) * sizeof(int2), hipMemcpyHostToDevice, stream));
    HIPCK(hipStreamSynchronize(stream));
    return NPHYS_OK;
  }

  // JointConstraintSet::remove: the joint no longer takes part in the solve or the islands.
  int remove_joint(int ji) {
    if (ji < 0 || ji >= nj) return set_err(NPHYS_ERR_INVALID, "bad joint handle");
    int st = 0;
    HIPCK(hipMemcpyAsync(&st, j_state.as<int>() + ji, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCK(hipStreamSynchronize(stream));
    st |= 2;
    HIPCK(hipMemcpyAsync(j_state.as<int>() + ji, &st, sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCK(hipStreamSynchronize(stream));
    return NPHYS_OK;
  }

  int read_joints(float* lin, float* ang, int32_t* broken, int64_t cap) {
    if (cap < nj) return set_err(NPHYS_ERR_CAPACITY, "joint buffers too small");
    if (nj == 0) return NPHYS_OK;
    const size_t J = static_cast<size_t>(nj);
    std::vector<float4> l(J), a(J);
    std::vector<int> st(J);
    HIPCK(hipMemcpyAsync(l.data(), j_imp_lin.p, J * sizeof(float4), hipMemcpyDeviceToHost, stream));
    HIPCK(hipMemcpyAsync(a.data(), j_imp_ang.p, J * sizeof(float4), hipMemcpyDeviceToHost, stream));
    HIPCK(hipMemcpyAsync(st.data(), j_state.p, J * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCK(hipStreamSynchronize(stream));
    for (int i = 0; i < nj; ++i) {
      if (lin) {
        lin[3 * i] = l[size_t(i)].x;
        lin[3 * i + 1] = l[size_t(i)].y;
        lin[3 * i + 2] = l[size_t(i)].z;
      }
      if (ang) {
        ang[3 * i] = a[size_t(i)].x;
        ang[3 * i + 1] = a[size_t(i)].y;
        ang[3 * i + 2] = a[size_t(i)].z;
      }
      if (broken) broken[i] = st[size_t(i)] & 1;
    }
    return NPHYS_OK;
  }

  int add_bodies(const nphys_body_desc& d, int32_t* first) {
    if (d.n < 0) return set_err(NPHYS_ERR_INVALID, "negative body count");
    if (!d.status || !d.shape || !d.dims || !d.pos || !d.rot || !d.density || !d.margin ||
        !d.linear_prediction || !d.friction || !d.restitution)
      return set_err(NPHYS_ERR_INVALID, "missing body array");
    if (int64_t(n) + d.n >= (int64_t(1) << 24)) return set_err(NPHYS_ERR_INVALID, "too many bodies");
    if (first) *first = n;
    if (d.n == 0) return NPHYS_OK;
    const int n0 = n, n1 = n + d.n;
    std::vector<float4> hp(d.n), hr(d.n), hv(d.n), hw(d.n), hmn(d.n), hmx(d.n);
    std::vector<float2> ha(d.n);
    bool any_sleeper = false;
    int64_t new_dynamic = 0;
    for (int k = 0; k < d.n; ++k) {
      const int st = d.status[k], sh = d.shape[k];
      if (st != NPHYS_STATUS_DYNAMIC && st != NPHYS_STATUS_STATIC)
        return set_err(NPHYS_ERR_INVALID, "bad body status");
      if (sh != NPHYS_SHAPE_BALL && sh != NPHYS_SHAPE_CUBOID) return set_err(NPHYS_ERR_INVALID, "bad shape");
      if (d.scene && (d.scene[k] < 0 || d.scene[k] >= NPHYS_MAX_SCENES))
        return set_err(NPHYS_ERR_INVALID, "scene index out of range");
      const float* dm = d.dims + 3 * k;
      if (sh == NPHYS_SHAPE_BALL ? !(dm[0] > 0.0f) : !(dm[0] > 0.0f && dm[1] > 0.0f && dm[2] > 0.0f))
        return set_err(NPHYS_ERR_INVALID, "shape dimensions must be positive");
      if ((d.friction_combine && (d.friction_combine[k] < 0 || d.friction_combine[k] > 3)) ||
          (d.restitution_combine && (d.restitution_combine[k] < 0 || d.restitution_combine[k] > 3)))
        return set_err(NPHYS_ERR_INVALID, "bad combine mode");
    }
    for (int k = 0; k < d.n; ++k) {
      const int st = d.status[k], sh = d.shape[k];
      const float* dm = d.dims + 3 * k;
      const float4 he = sh == NPHYS_SHAPE_BALL ? make_float4(dm[0], dm[0], dm[0], 0.0f)
                                               : make_float4(dm[0], dm[1], dm[2], 0.0f);
      // mass properties (volumetric_ball.rs:9-44, volumetric_cuboid.rs:10-75, volumetric.rs:46-59)
      float m = 0.0f;
      float4 li = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const float density = d.density[k];
      if (st == NPHYS_STATUS_DYNAMIC && density != 0.0f) {
        if (sh == NPHYS_SHAPE_BALL) {
          const float r = dm[0];
          const float vol = ((((float)M_PI * r) * r) * r) * (float)(4.0 / 3.0);
          m = vol * density;
          const float diag = (r * r) * (float)(2.0 / 5.0);
          li = make_float4(diag * m, diag * m, diag * m, 0.0f);
        } else {
          float vol = 1.0f;
          vol = vol * dm[0] * 2.0f;
          vol = vol * dm[1] * 2.0f;
          vol = vol * dm[2] * 2.0f;
          m = vol * density;
          const float w = ((float)(1.0 / 12.0) * 2.0f) * 2.0f;
          const float ix = (w * dm[0]) * dm[0], iy = (w * dm[1]) * dm[1], iz = (w * dm[2]) * dm[2];
          li = make_float4((iy + iz) * m, (ix + iz) * m, (ix + iy) * m, 0.0f);
        }
      }
      h_status.push_back(st);
      h_shape.push_back(sh);
      h_scene.push_back(d.scene ? d.scene[k] : 0);
      h_dims.push_back(he);
      h_local_i.push_back(li);
      h_mass.push_back(m);
      h_margin.push_back(d.margin[k]);
      h_qlim.push_back(d.margin[k] + d.linear_prediction[k]);
      h_friction.push_back(d.friction[k]);
      h_restitution.push_back(d.restitution[k]);
      // RigidBodyDesc / BasicMaterial attributes (rigid_body.rs:972-990, basic_material.rs:33-43)
      const float fmax = 3.402823466e38f;
      const float4 dp = make_float4(d.linear_damping ? d.linear_damping[k] : 0.0f,
                                    d.angular_damping ? d.angular_damping[k] : 0.0f,
                                    d.max_linear_velocity ? d.max_linear_velocity[k] : fmax,
                                    d.max_angular_velocity ? d.max_angular_velocity[k] : fmax);
      h_damp.push_back(dp);
      any_damp = any_damp || dp.x != 0.0f || dp.y != 0.0f || dp.z != fmax || dp.w != fmax;
      const int kin = d.kinematic_dofs ? (d.kinematic_dofs[k] & 63) : 0;
      const int fl = kin | ((d.gravity_enabled && !d.gravity_enabled[k]) ? kFlagNoGravity : 0) |
                     ((d.friction_combine ? d.friction_combine[k] : 0) << 8) |
                     ((d.restitution_combine ? d.restitution_combine[k] : 0) << 10);
      h_flags.push_back(fl);
      any_flags = any_flags || fl != 0;
      any_mask = any_mask || (st == NPHYS_STATUS_DYNAMIC && kin != 0);
      const int mid = d.material_id ? (d.material_id[k] < 0 ? -1 : d.material_id[k]) : -1;
      h_matid.push_back(mid);
      any_matid = any_matid || mid >= 0;
      const float4 sv = d.surface_velocity
                            ? make_float4(d.surface_velocity[3 * k], d.surface_velocity[3 * k + 1],
                                          d.surface_velocity[3 * k + 2], 0.0f)
                            : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      h_surf.push_back(sv);
      any_surf = any_surf || sv.x != 0.0f || sv.y != 0.0f || sv.z != 0.0f;
      // ColliderDesc::collision_groups: groups 0-29 (CollisionGroups::new() without the array)
      constexpr uint32_t kAllGroups = (1u << 30) - 1u;
      const uint4 cg = d.collision_groups
                           ? make_uint4(d.collision_groups[3 * k] & kAllGroups, d.collision_groups[3 * k + 1] & kAllGroups,
                                        d.collision_groups[3 * k + 2] & kAllGroups, 0u)
                           : make_uint4(kAllGroups, kAllGroups, 0u, 0u);
      h_groups.push_back(cg);
      any_groups = any_groups || cg.x != kAllGroups || cg.y != kAllGroups || cg.z != 0u;
      // ColliderDesc::ccd_enabled, RigidBody::enable_linear_motion_interpolation, and the static body
      // a static collider belongs to
      const int cf = ((d.ccd_enabled && d.ccd_enabled[k]) ? 1 : 0) |
                     ((st == NPHYS_STATUS_DYNAMIC && d.linear_motion_interpolation && d.linear_motion_interpolation[k])
                          ? 2
                          : 0);
      h_ccd.push_back(cf);
      any_ccd = any_ccd || (cf & 1) != 0;
      int ckey = n0 + k;
      if (st != NPHYS_STATUS_DYNAMIC && d.static_body && d.static_body[k] >= 0)
        ckey = static_owner.emplace(std::make_pair(h_scene.back(), int(d.static_body[k])), n0 + k).first->second;
      h_ccd_key.push_back(ckey);
      const float* p = d.pos + 3 * k;
      const float* q = d.rot + 4 * k;
      hp[k] = make_float4(p[0], p[1], p[2], 0.0f);
      hr[k] = make_float4(q[0], q[1], q[2], q[3]);
      const bool dyn = st == NPHYS_STATUS_DYNAMIC;
      // ActivationStatus::new_active (body.rs:77-82): energy 4 x the default threshold 0.01
      const float thr = d.sleep_threshold ? d.sleep_threshold[k] : -1.0f;
      ha[k] = make_float2(0.01f * 4.0f, thr);
      any_sleeper = any_sleeper || (dyn && thr >= 0.0f);
      new_dynamic += dyn ? 1 : 0;
      hv[k] = (dyn && d.linvel) ? make_float4(d.linvel[3 * k], d.linvel[3 * k + 1], d.linvel[3 * k + 2], 0.0f)
                                : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      hw[k] = (dyn && d.angvel) ? make_float4(d.angvel[3 * k], d.angvel[3 * k + 1], d.angvel[3 * k + 2], 0.0f)
                                : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      // proxy creation: loose = aabb(pos) loosened by the query limit, then by 0.01
      const aabb a = shape_aabb(sh, xyz(he), xyz(hp[k]), toq(hr[k]));
      const aabb lo = loosened(loosened(a, h_qlim.back()), 0.01f);
      hmn[k] = f4(lo.mn, 0.0f);
      hmx[k] = f4(lo.mx, 0.0f);
    }
    const size_t f4b = sizeof(float4);
    // grow the per-body device arrays, keeping the existing bodies' state
    for (DBuf* b : {&linvel, &angvel, &loose_mn, &loose_mx})
      HIPCK(b->ensure(size_t(n1) * f4b, true, stream));
    HIPCK(pose.ensure(size_t(n1) * kPoseBlock * f4b, true, stream));
    // kept: a sleeping body's inverse inertia is not recomputed
    HIPCK(sol.ensure(size_t(n1) * kSolverBlock * f4b, true, stream));
    HIPCK(act.ensure(size_t(n1) * sizeof(float2), true, stream));
    for (DBuf* b : {&w_dyn, &f_lin, &f_ang}) HIPCK(b->ensure(size_t(n1) * f4b, true, stream));
    for (DBuf* b : {&pos0, &rot0, &ext_lin, &ext_ang, &dims, &local_i, &damp, &surf, &rot_dyn, &groups})
      HIPCK(b->ensure(size_t(n1) * f4b));
    for (DBuf* b : {&status, &shape, &scene, &margin, &qlim, &friction, &restitution, &mass, &bflags, &matid,
                    &ccd_flags, &ccd_keys})
      HIPCK(b->ensure(size_t(n1) * 4));
    HIPCK(ccd_rot.ensure(size_t(n1) * f4b));
    // scene bits above the Morton code: the BVH keeps every scene in its own key range
    int max_scene = 0;
    for (int sc : h_scene) max_scene = std::max(max_scene, sc);
    scene_bits = 0;
    while ((1 << scene_bits) <= max_scene) ++scene_bits;
    auto up = [&](DBuf& b, const void* src, size_t bytes, size_t off) {
      return hipMemcpyAsync(static_cast<char*>(b.p) + off, src, bytes, hipMemcpyHostToDevice, stream);
    };
    HIPCK(up(status, h_status.data(), size_t(n1) * 4, 0));
    HIPCK(up(shape, h_shape.data(), size_t(n1) * 4, 0));
    HIPCK(up(scene, h_scene.data(), size_t(n1) * 4, 0));
    HIPCK(up(dims, h_dims.data(), size_t(n1) * f4b, 0));
    HIPCK(up(local_i, h_local_i.data(), size_t(n1) * f4b, 0));
    HIPCK(up(mass, h_mass.data(), size_t(n1) * 4, 0));
    HIPCK(up(margin, h_margin.data(), size_t(n1) * 4, 0));
    HIPCK(up(qlim, h_qlim.data(), size_t(n1) * 4, 0));
    HIPCK(up(friction, h_friction.data(), size_t(n1) * 4, 0));
    HIPCK(up(restitution, h_restitution.data(), size_t(n1) * 4, 0));
    HIPCK(up(damp, h_damp.data(), size_t(n1) * f4b, 0));
    HIPCK(up(surf, h_surf.data(), size_t(n1) * f4b, 0));
    HIPCK(up(groups, h_groups.data(), size_t(n1) * f4b, 0));
    HIPCK(up(bflags, h_flags.data(), size_t(n1) * 4, 0));
    HIPCK(up(matid, h_matid.data(), size_t(n1) * 4, 0));
    HIPCK(up(ccd_flags, h_ccd.data(), size_t(n1) * 4, 0));
    HIPCK(up(ccd_keys, h_ccd_key.data(), size_t(n1) * 4, 0));
    // new bodies: no dynamics update yet (the augmented mass is the world inertia, w = 0) and no
    // pending force
    HIPCK(hipMemsetAsync(w_dyn.as<float4>() + n0, 0, size_t(d.n) * f4b, stream));
    HIPCK(hipMemsetAsync(f_lin.as<float4>() + n0, 0, size_t(d.n) * f4b, stream));
    HIPCK(hipMemsetAsync(f_ang.as<float4>() + n0, 0, size_t(d.n) * f4b, stream));
    HIPCK(hipMemsetAsync(sol.as<float4>() + size_t(n0) * kSolverBlock, 0, size_t(d.n) * kSolverBlock * f4b, stream));
    HIPCK(hipMemcpyAsync(act.as<float2>() + n0, ha.data(), size_t(d.n) * sizeof(float2), hipMemcpyHostToDevice,
                         stream));
    HIPCK(pose_copy(0, hp.data(), n0, d.n, hipMemcpyHostToDevice));
    HIPCK(pose_copy(1, hr.data(), n0, d.n, hipMemcpyHostToDevice));
    HIPCK(up(linvel, hv.data(), size_t(d.n) * f4b, size_t(n0) * f4b));
    HIPCK(up(angvel, hw.data(), size_t(d.n) * f4b, size_t(n0) * f4b));
    HIPCK(up(loose_mn, hmn.data(), size_t(d.n) * f4b, size_t(n0) * f4b));
    HIPCK(up(loose_mx, hmx.data(), size_t(d.n) * f4b, size_t(n0) * f4b));
    // per-scene dynamic-body lists of the scene-blocked sweeps
    nscenes = 1;
    for (int sc : h_scene) nscenes = std::max(nscenes, sc + 1);
    std::vector<int> dstart(size_t(nscenes) + 1, 0), dlist, local(size_t(n1), -1);
    for (int i = 0; i < n1; ++i)
      if (h_status[i] == NPHYS_STATUS_DYNAMIC) ++dstart[size_t(h_scene[i]) + 1];
    max_scene_dyn = 0;
    for (int sc = 0; sc < nscenes; ++sc) {
      max_scene_dyn = std::max(max_scene_dyn, dstart[size_t(sc) + 1]);
      dstart[size_t(sc) + 1] += dstart[size_t(sc)];
    }
    dlist.resize(size_t(dstart[size_t(nscenes)]));
    {
      std::vector<int> fill(dstart.begin(), dstart.end() - 1);
      for (int i = 0; i < n1; ++i)
        if (h_status[i] == NPHYS_STATUS_DYNAMIC) {
          const int sc = h_scene[i];
          local[size_t(i)] = fill[size_t(sc)] - dstart[size_t(sc)];
          dlist[size_t(fill[size_t(sc)]++)] = i;
        }
    }
    HIPCK(sc_dyn_start.ensure(dstart.size() * 4));
    HIPCK(sc_dyn_list.ensure(std::max<size_t>(dlist.size(), 1) * 4));
    HIPCK(sc_local.ensure(local.size() * 4));
    HIPCK(up(sc_dyn_start, dstart.data(), dstart.size() * 4, 0));
    if (!dlist.empty()) HIPCK(up(sc_dyn_list, dlist.data(), dlist.size() * 4, 0));
    HIPCK(up(sc_local, local.data(), local.size() * 4, 0));
    {
      std::vector<int> small, large;
      for (int sc = 0; sc < nscenes; ++sc) {
        const int nd = dstart[size_t(sc) + 1] - dstart[size_t(sc)];
        if (nd == 0) continue;
        (nd <= kSmallScene ? small : large).push_back(sc);
      }
      n_small_scenes = int(small.size());
      n_large_scenes = int(large.size());
      HIPCK(sc_small.ensure(std::max<size_t>(small.size(), 1) * 4));
      HIPCK(sc_large.ensure(std::max<size_t>(large.size(), 1) * 4));
      if (!small.empty()) HIPCK(up(sc_small, small.data(), small.size() * 4, 0));
      if (!large.empty()) HIPCK(up(sc_large, large.data(), large.size() * 4, 0));
    }
    tune_cands.clear();
    tuned = false;
    HIPCK(hipStreamSynchronize(stream));
    n = n1;
    n_dynamic += new_dynamic;
    sleeping = sleeping || any_sleeper;
    pairs_dirty = true;
    moved = true;
    return NPHYS_OK;
  }

  // Copies the positions (slot 0) or rotations (slot 1) of `count` pose blocks from body `first`
  // on, to or from a packed host array of float4.
  hipError_t pose_copy(int slot, float4* host, int first, int count, hipMemcpyKind kind) {
    if (count <= 0) return hipSuccess;
    const size_t pitch = kPoseBlock * sizeof(float4), w = sizeof(float4);
    float4* dev = pose.as<float4>() + size_t(first) * kPoseBlock + slot;
    if (kind == hipMemcpyHostToDevice)
      return hipMemcpy2DAsync(dev, pitch, host, w, w, size_t(count), kind, stream);
    return hipMemcpy2DAsync(host, w, dev, pitch, w, size_t(count), kind, stream);
  }

  int set_state(const float* p, const float* r, const float* lv, const float* av) {
    if (n == 0) return NPHYS_OK;
    if (r && !rot_dyn_saved) {
      // impulses applied before the next step use the augmented mass of the rotation the last
      // dynamics update saw (the reference recomputes it only in the next step)
      if (any_ccd) {
        launch_save_rot_dyn(bodies(), rot_dyn.as<float4>(), stream);
      } else {
        HIPCK(hipMemcpy2DAsync(rot_dyn.p, sizeof(float4), pose.as<float4>() + 1, kPoseBlock * sizeof(float4),
                               sizeof(float4), size_t(n), hipMemcpyDeviceToDevice, stream));
      }
      rot_dyn_saved = true;
    }
    std::vector<float4> hp(n), hr(n), hv(n), hw(n);
    const size_t f4b = sizeof(float4);
    HIPCK(pose_copy(0, hp.data(), 0, n, hipMemcpyDeviceToHost));
    HIPCK(pose_copy(1, hr.data(), 0, n, hipMemcpyDeviceToHost));
    HIPCK(hipMemcpyAsync(hv.data(), linvel.p, n * f4b, hipMemcpyDeviceToHost, stream));
    HIPCK(hipMemcpyAsync(hw.data(), angvel.p, n * f4b, hipMemcpyDeviceToHost, stream));
    HIPCK(hipStreamSynchronize(stream));
    for (int i = 0; i < n; ++i) {
      if (p) hp[i] = make_float4(p[3 * i], p[3 * i + 1], p[3 * i + 2], 0.0f);
      if (r) hr[i] = make_float4(r[4 * i], r[4 * i + 1], r[4 * i + 2], r[4 * i + 3]);
      if (h_status[i] == NPHYS_STATUS_DYNAMIC) {
        if (lv) hv[i] = make_float4(lv[3 * i], lv[3 * i + 1], lv[3 * i + 2], 0.0f);
        if (av) hw[i] = make_float4(av[3 * i], av[3 * i + 1], av[3 * i + 2], 0.0f);
      }
    }
    HIPCK(pose_copy(0, hp.data(), 0, n, hipMemcpyHostToDevice));
    HIPCK(pose_copy(1, hr.data(), 0, n, hipMemcpyHostToDevice));
    HIPCK(hipMemcpyAsync(linvel.p, hv.data(), n * f4b, hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(angvel.p, hw.data(), n * f4b, hipMemcpyHostToDevice, stream));
    if (sleeping || any_ccd) launch_clear_dyn_marks(bodies(), stream);
    HIPCK(hipStreamSynchronize(stream));
    moved = true;
    wake_pending = sleeping;
    return NPHYS_OK;
  }

  int set_collision_groups(int32_t count, const int32_t* bodies_, const uint32_t* g) {
    if (count < 0 || (count > 0 && (!bodies_ || !g))) return set_err(NPHYS_ERR_INVALID, "bad collision group arrays");
    for (int k = 0; k < count; ++k)
      if (bodies_[k] < 0 || bodies_[k] >= n) return set_err(NPHYS_ERR_INVALID, "bad body handle");
    if (count == 0) return NPHYS_OK;
    constexpr uint32_t kAllGroups = (1u << 30) - 1u;
    for (int k = 0; k < count; ++k)
      h_groups[size_t(bodies_[k])] = make_uint4(g[3 * k] & kAllGroups, g[3 * k + 1] & kAllGroups, g[3 * k + 2] & kAllGroups, 0u);
    any_groups = false;
    for (const uint4& cg : h_groups) any_groups = any_groups || cg.x != kAllGroups || cg.y != kAllGroups || cg.z != 0u;
    HIPCK(hipMemcpyAsync(groups.p, h_groups.data(), size_t(n) * sizeof(uint4), hipMemcpyHostToDevice, stream));
    HIPCK(hipStreamSynchronize(stream));
    pairs_dirty = true;
    recollide = true;
    return NPHYS_OK;
  }

  int read_activation(float* energy, int64_t cap) {
    if (cap < n) return set_err(NPHYS_ERR_CAPACITY, "energy buffer too small");
    if (n == 0) return NPHYS_OK;
    std::vector<float2> h(static_cast<size_t>(n));
    HIPCK(hipMemcpyAsync(h.data(), act.p, size_t(n) * sizeof(float2), hipMemcpyDeviceToHost, stream));
    HIPCK(hipStreamSynchronize(stream));
    for (int i = 0; i < n; ++i) energy[i] = h[size_t(i)].x;
    return NPHYS_OK;
  }

  int get_state(float* p, float* r, float* lv, float* av) {
    if (n == 0) return NPHYS_OK;
    std::vector<float4> h(n);
    const size_t f4b = sizeof(float4);
    struct Item { DBuf* b; int slot; float* out; int w; };
    const Item items[4] = {{nullptr, 0, p, 3}, {nullptr, 1, r, 4}, {&linvel, 0, lv, 3}, {&angvel, 0, av, 3}};
    for (const Item& it : items) {
      if (!it.out) continue;
      if (it.b)
        HIPCK(hipMemcpyAsync(h.data(), it.b->p, n * f4b, hipMemcpyDeviceToHost, stream));
      else
        HIPCK(pose_copy(it.slot, h.data(), 0, n, hipMemcpyDeviceToHost));
      HIPCK(hipStreamSynchronize(stream));
      for (int i = 0; i < n; ++i) {
        it.out[it.w * i + 0] = h[i].x;
        it.out[it.w * i + 1] = h[i].y;
        it.out[it.w * i + 2] = h[i].z;
        if (it.w == 4) it.out[4 * i + 3] = h[i].w;
      }
    }
    return NPHYS_OK;
  }

  // ------------------------------------------------------------------ broad phase
  int ensure_bp() {
    const size_t nn = size_t(std::max(n, 2));
    for (DBuf* b : {&keys, &vals, &keys_alt, &vals_alt, &cand_count, &cand_off})
      HIPCK(b->ensure(nn * 4));
    HIPCK(visit.ensure(8 * nn * 4));
    HIPCK(child.ensure(2 * nn * 4));
    HIPCK(parent.ensure(2 * nn * 4));
    HIPCK(node_mn.ensure(2 * nn * sizeof(float4)));
    HIPCK(node_mx.ensure(2 * nn * sizeof(float4)));
    HIPCK(node_scene.ensure(nn * sizeof(int2)));
    HIPCK(bp_pnode.ensure(4 * nn * sizeof(float4)));
    if (cand.bytes < nn * size_t(cap) * 4) tree_valid = false;  // segments are lost on growth
    HIPCK(cand.ensure(nn * size_t(cap) * 4));
    for (DBuf* b : {&bp_leaf_of, &bp_list}) HIPCK(b->ensure(nn * 4));
    if (bp_dirty.bytes < nn * 4) {
      tree_valid = false;
      HIPCK(bp_dirty.ensure(nn * 4));
      fill_u32(bp_dirty.as<uint32_t>(), 0u, nn, stream);
    }
    const size_t tw = std::max(radix_temp_words(int(nn)), scan_temp_words(int(nn)));
    HIPCK(bp_tmp.ensure(tw * 4));
    return NPHYS_OK;
  }

  BroadPhaseBufs bp_view() {
    BroadPhaseBufs v;
    uint32_t* sm = bp_small.as<uint32_t>();
    v.changed = sm;
    v.bounds = sm + 8;
    v.stats = sm + 16;
    v.keys = keys.as<uint32_t>();
    v.vals = vals.as<uint32_t>();
    v.keys_alt = keys_alt.as<uint32_t>();
    v.vals_alt = vals_alt.as<uint32_t>();
    v.child = child.as<int>();
    v.parent = parent.as<int>();
    v.visit = visit.as<uint32_t>();
    v.node_mn = node_mn.as<float4>();
    v.node_mx = node_mx.as<float4>();
    v.node_scene = node_scene.as<int2>();
    v.pnode = bp_pnode.as<float4>();
    v.scene_bits = scene_bits;
    v.cand = cand.as<int>();
    v.cand_count = cand_count.as<uint32_t>();
    v.cand_off = cand_off.as<uint32_t>();
    v.pairs = pairs.as<int2>();
    v.tmp = bp_tmp.as<uint32_t>();
    v.dirty = bp_dirty.as<uint32_t>();
    v.leaf_of = bp_leaf_of.as<int>();
    v.list = bp_list.as<int>();
    v.list_count = sm + 28;
    return v;
  }

  // Enqueues a readback of up to four device word ranges into rb_host; returns its sequence number.
  uint32_t rb_post(std::initializer_list<RbSeg> segs) {
    RbJob j{};
    int k = 0;
    for (const RbSeg& s : segs) {
      j.src[k] = s.src;
      j.n[k] = s.n;
      j.dst[k] = s.dst;
      ++k;
    }
    ++rb_seq;
    hipLaunchKernelGGL(k_readback, dim3(1), dim3(256), 0, stream, j, rb_dev, kRbFlag, rb_seq);
    return rb_seq;
  }

  // Waits until readback `seq` (and everything before it on the stream) has completed. After two
  // seconds the stream is synchronised instead, which reports a failed launch or fault.
  int rb_wait(uint32_t seq) {
    const volatile uint32_t* f = rb_host + kRbFlag;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t it = 1; int32_t(*f - seq) < 0; ++it) {
      if ((it & 0xfffu) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) {
        HIPCK(hipStreamSynchronize(stream));
        if (int32_t(*f - seq) >= 0) break;
        return set_err(NPHYS_ERR_HIP, "readback did not complete");
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return NPHYS_OK;
  }

  int readback(std::initializer_list<RbSeg> segs) { return rb_wait(rb_post(segs)); }

  // Loose-AABB pass; returns the number of replaced loose AABBs through *changed.
  int loose_pass(int mode, uint32_t* changed) {
    int rc = ensure_bp();  // the dirty flags
    if (rc) return rc;
    BroadPhaseBufs v = bp_view();
    fill_u32(v.changed, 0u, 2, stream);
    launch_loose_update(bodies(), mode, v.changed, v.dirty, stream);
    if ((rc = readback({{v.changed, 2, 0}}))) return rc;
    *changed = rb_host[0];
    dirty_pending += rb_host[0];
    dirty_static_pending += rb_host[1];
    return NPHYS_OK;
  }

  // The end-of-step syncs (CCD-phase swept AABBs, then the final poses) with one readback.
  // The swept pass is the first of solve_ccd's loop, which returns before it when dt or
  // max_ccd_substeps is zero (mechanical_world.rs:578-581).
  int loose_passes_end_of_step(uint32_t* ch_swept, uint32_t* ch_final) {
    int rc = ensure_bp();
    if (rc) return rc;
    BroadPhaseBufs v = bp_view();
    fill_u32(v.changed, 0u, 4, stream);
    if (sp.dt != 0.0f && prm.max_ccd_substeps != 0) launch_loose_update(bodies(), 2, v.changed, v.dirty, stream);
    launch_loose_update(bodies(), 1, v.changed + 2, v.dirty, stream);
    if ((rc = readback({{v.changed, 4, 0}}))) return rc;
    *ch_swept = rb_host[0];
    *ch_final = rb_host[2];
    dirty_pending += int64_t(rb_host[0]) + rb_host[2];
    dirty_static_pending += int64_t(rb_host[1]) + rb_host[3];
    return NPHYS_OK;
  }

  int rebuild_pairs() {
    if (n < 2) {
      npairs = 0;
      pairs_dirty = false;
      return NPHYS_OK;
    }
    int rc = ensure_bp();
    if (rc) return rc;
    DevBodies b = bodies();
    BroadPhaseBufs v = bp_view();
    bool done = false;
    if (tree_valid && incremental_bp && !pairs_dirty && dirty_static_pending == 0 &&
        inc_since_full < kMaxIncremental && dirty_pending <= n / 8) {
      // the pair offsets are computed right away and read back with the overflow check; they are
      // discarded if the update overflowed
      launch_incremental_pairs(b, v, cap, int(dirty_pending), stream);
      launch_pair_offsets(b, v, cap, stream);
      if ((rc = readback({{v.stats, 2, 0}, {v.cand_off + (n - 1), 1, 2}, {v.keys_alt + (n - 1), 1, 3}}))) return rc;
      done = rb_host[1] == 0 && int(rb_host[0]) <= cap;  // else the full build below
      if (done) {
        ++inc_since_full;
        ++counters.incremental_pair_updates;
        npairs = int(rb_host[2] + rb_host[3]);
      } else if (rb_host[1] == 0) {
        // a clean body's segment took more appended partners than it has room for: the full build
        // below gets segments that large, so the following incremental updates fit
        ++counters.incremental_overflows;
        cap = int(rb_host[0]) + 8;
        if ((rc = ensure_bp())) return rc;
        v = bp_view();
      }
    }
    if (!done) {
      if ((rc = full_traversal(b, v))) return rc;
      inc_since_full = 0;
      launch_pair_offsets(b, v, cap, stream);
      // total = off[n-1] + min(count[n-1], cap)
      if ((rc = readback({{v.cand_off + (n - 1), 1, 0}, {v.keys_alt + (n - 1), 1, 1}}))) return rc;
      npairs = int(rb_host[0] + rb_host[1]);
    }
    HIPCK(pairs.ensure(size_t(std::max(npairs, 1)) * sizeof(int2)));
    v = bp_view();
    launch_compact_pairs(b, v, cap, stream);
    fill_u32(v.dirty, 0u, size_t(n), stream);
    dirty_pending = dirty_static_pending = 0;
    tree_valid = true;
    pairs_dirty = false;
    return NPHYS_OK;
  }

  // Full build: the tree from scratch and every dynamic body traversed, the segment capacity grown
  // until no body overflows it.
  int full_traversal(DevBodies& b, BroadPhaseBufs& v) {
    int rc;
    launch_build_lbvh(b, v, stream);
    for (;;) {
      launch_traverse(b, v, cap, stream);
      if ((rc = readback({{v.stats, 2, 0}}))) return rc;
      if (rb_host[1]) return set_err(NPHYS_ERR_STATE, "BVH traversal stack overflow");
      if (int(rb_host[0]) <= cap) break;
      cap = int(rb_host[0]) + 8;
      if ((rc = ensure_bp())) return rc;
      v = bp_view();
    }
    return NPHYS_OK;
  }

  // The contacts of `c` in manifold order (dense records and, when imp is given, impulses), with
  // the manifolds' bodies and counts, copied to the host.
  int gather_contacts(ContactBufs& c, std::vector<int2>& mb, std::vector<int>& mc, std::vector<float4>& rec,
                      std::vector<float4>* imp) {
    const int M = c.M, C = c.C;
    mb.resize(size_t(M));
    mc.resize(size_t(M));
    HIPCK(hipMemcpyAsync(mb.data(), c.mbody.p, size_t(M) * sizeof(int2), hipMemcpyDeviceToHost, stream));
    HIPCK(hipMemcpyAsync(mc.data(), c.mcount.p, size_t(M) * 4, hipMemcpyDeviceToHost, stream));
    HIPCK(hipStreamSynchronize(stream));
    std::vector<uint32_t> off(static_cast<size_t>(M));
    uint32_t acc = 0;
    for (int m = 0; m < M; ++m) {
      off[size_t(m)] = acc;
      acc += uint32_t(mc[size_t(m)]);
    }
    if (int(acc) != C) return set_err(NPHYS_ERR_STATE, "contact count mismatch");
    HIPCK(rd_moff.ensure(size_t(std::max(M, 1)) * 4));
    HIPCK(rd_rec.ensure(size_t(std::max(C, 1)) * kContactFields * sizeof(float4)));
    if (imp) HIPCK(rd_imp.ensure(size_t(std::max(C, 1)) * sizeof(float4)));
    HIPCK(hipMemcpyAsync(rd_moff.p, off.data(), size_t(M) * 4, hipMemcpyHostToDevice, stream));
    launch_gather_contacts(c.view(), M, rd_moff.as<uint32_t>(), rd_rec.as<float4>(),
                           imp ? rd_imp.as<float4>() : nullptr, stream);
    rec.resize(size_t(C) * kContactFields);
    HIPCK(hipMemcpyAsync(rec.data(), rd_rec.p, rec.size() * sizeof(float4), hipMemcpyDeviceToHost, stream));
    if (imp) {
      imp->resize(size_t(C));
      HIPCK(hipMemcpyAsync(imp->data(), rd_imp.p, imp->size() * sizeof(float4), hipMemcpyDeviceToHost, stream));
    }
    HIPCK(hipStreamSynchronize(stream));
    return NPHYS_OK;
  }

  // ------------------------------------------------------------------ narrow phase
  // `done`, when given, is recorded after the narrow phase and before its readback (the step's last
  // stage event, complete once the readback is).
  int narrow_phase(hipEvent_t done = nullptr) { return narrow_phase_of(pairs.as<int2>(), npairs, cb[cur], done); }

  // The narrow phase of `np` pairs (in pair order) into `out`.
  int narrow_phase_of(const int2* list, int np, ContactBufs& out, hipEvent_t done = nullptr) {
    const size_t P = size_t(std::max(np, 1));
    // a manifold's first slot is packed in 28 bits of the solver's slot record
    if (kPairSlots * P >= (size_t(1) << 28)) return set_err(NPHYS_ERR_CAPACITY, "too many broad-phase pairs");
    for (DBuf* b : {&np_count, &np_coff, &np_has, &np_moff, &np_perm}) HIPCK(b->ensure(P * 4));
    HIPCK(np_tmp.ensure(scan_temp_words(int(P)) * 4 + 64));
    HIPCK(out.ensure(P, kPairSlots * P));
    HIPCK(out.bfirst.ensure(size_t(std::max(n, 1)) * 4));
    out.nb = n;
    NarrowPhaseBufs v;
    v.count = np_count.as<uint32_t>();
    v.coff = np_coff.as<uint32_t>();
    v.has = np_has.as<uint32_t>();
    v.perm = np_perm.as<uint32_t>();
    v.moff = np_moff.as<uint32_t>();
    v.totals = bp_small.as<uint32_t>() + 24;
    v.tmp = np_tmp.as<uint32_t>();
    launch_narrow_phase(bodies(), list, np, v, out.view(), stream);
    if (done) HIPCK(hipEventRecord(done, stream));
    const int rc = readback({{v.totals, 2, 0}});
    if (rc) return rc;
    out.C = int(rb_host[0]);
    out.M = int(rb_host[1]);
    return NPHYS_OK;
  }

  // sync_colliders + perform_broad_phase + perform_narrow_phase on the current poses.
  int collide_now(int mode) {
    uint32_t changed = 0;
    int rc = loose_pass(mode, &changed);
    if (rc) return rc;
    if (changed || pairs_dirty) {
      if ((rc = rebuild_pairs())) return rc;
    }
    return narrow_phase();
  }

  // ------------------------------------------------------------------ solver
  int ensure_solver(int M, int C) {
    const size_t nn = size_t(std::max(n, 1));
    const size_t m = size_t(std::max(M, 1)), c = size_t(std::max(C, 1));
    for (DBuf* b : {&sv_deg, &sv_adj_off, &sv_adj_fill}) HIPCK(b->ensure(nn * 4));
    HIPCK(sv_adj.ensure(2 * m * 4));
    for (DBuf* b : {&sv_ckeys, &sv_order, &sv_ckeys_alt, &sv_order_alt, &sv_mcount_ord, &sv_soff, &sv_slot_of})
      HIPCK(b->ensure(m * 4));
    HIPCK(sv_state2[cur].ensure(m * 4));
    HIPCK(sv_state2[cur ^ 1].ensure(4));
    HIPCK(sv_col_start.ensure(2 * (kMaxColours + 2) * 4));
    HIPCK(sv_sinfo.ensure(m * sizeof(int4)));
    HIPCK(sv_rec2c.ensure(c * 4));
    HIPCK(sv_rec.ensure(c * kRecFields * sizeof(float4)));
    HIPCK(sv_lam.ensure(c * sizeof(float4)));
    HIPCK(sv_prec.ensure(c * 3 * sizeof(float4)));
    HIPCK(sv_pgeo.ensure(m * 2 * sizeof(float4)));
    rec_stride = long(std::min({sv_rec.bytes / (kRecFields * sizeof(float4)), sv_lam.bytes / sizeof(float4),
                                sv_prec.bytes / (3 * sizeof(float4))}));
    HIPCK(sv_red.ensure((size_t(residual_blocks(M)) * 2 + 4) * sizeof(double)));  // + the joints' partial
    const size_t tw = std::max({radix_temp_words(int(m)), scan_temp_words(int(m)), scan_temp_words(int(nn))});
    HIPCK(sv_tmp.ensure(tw * 4 + 64));
    return NPHYS_OK;
  }

  SolverBufs sv_view() {
    SolverBufs v;
    v.deg = sv_deg.as<uint32_t>();
    v.adj_off = sv_adj_off.as<uint32_t>();
    v.adj_fill = sv_adj_fill.as<uint32_t>();
    v.adj = sv_adj.as<int>();
    v.state = sv_state2[cur].as<uint32_t>();
    v.remaining = bp_small.as<uint32_t>() + 32;
    v.ckeys = sv_ckeys.as<uint32_t>();
    v.order = sv_order.as<uint32_t>();
    v.ckeys_alt = sv_ckeys_alt.as<uint32_t>();
    v.order_alt = sv_order_alt.as<uint32_t>();
    v.col_start = sv_col_start.as<uint32_t>();
    v.mcount_ord = sv_mcount_ord.as<uint32_t>();
    v.soff = sv_soff.as<uint32_t>();
    v.sinfo = sv_sinfo.as<int4>();
    v.slot_of = sv_slot_of.as<uint32_t>();
    v.rec_stride = rec_stride;
    v.visit = sv_visit.as<int>();
    v.vrange = reinterpret_cast<int2*>(sv_visit.as<int>() + (kMaxColours + 2));
    v.rec2c = sv_rec2c.as<uint32_t>();
    v.rec = sv_rec.as<float4>();
    v.lam = sv_lam.as<float4>();
    v.prec = sv_prec.as<float4>();
    v.pgeo = sv_pgeo.as<float4>();
    v.red = sv_red.as<double>();
    v.tmp = sv_tmp.as<uint32_t>();
    return v;
  }

  // Visit order of the colour classes for one solve: Fisher-Yates driven by SplitMix64 seeded
  // with the solve index, so every step orders conflicting manifolds afresh while the colours
  // themselves persist between steps; then stably sorted by size bucket floor(log2(size)),
  // smallest first, so the largest classes end every sweep (the rows solved last are the most
  // numerous: a lower residual at equal iteration count).
  static int size_bucket(int n) {
    int l = 0;
    while (l < 30 && (2 << l) <= n) ++l;
    return l;
  }
  static void colour_visit_order(uint64_t solve_index, int k, const int* size, std::vector<int>& perm) {
    perm.resize(k);
    for (int c = 0; c < k; ++c) perm[c] = c;
    uint64_t st = solve_index * 0x9E3779B97F4A7C15ull;
    for (int i = k - 1; i > 0; --i) {
      st += 0x9E3779B97F4A7C15ull;
      uint64_t z = st;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      const int j = int(z % uint64_t(i + 1));
      std::swap(perm[i], perm[j]);
    }
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return size_bucket(size[a]) < size_bucket(size[b]); });
  }

  int colour(const DevBodies& b, const DevContacts& c, int M, int C, SolverBufs& v) {
    // colours are inherited from the manifolds of the previous solve
    const int prevM = have_solved ? sview_M[cur ^ 1] : 0;
    launch_colour_prepare(b, c, M, sview[cur ^ 1], prevM, sv_state2[cur ^ 1].as<uint32_t>(), v, stream);
    // Rounds in batches of four with one poll per batch: a round after the last one that colours
    // anything changes nothing, and the device records the last round that did.
    // The sort keys and colour count are computed after every batch and read back with the poll,
    // so the usual single batch costs one synchronisation; they are recomputed after a further one.
    fill_u32(v.remaining + 1, 0u, 1, stream);
    uint32_t* ncol = bp_small.as<uint32_t>() + 40;
    for (int round = 1;; round += kRoundBatch) {
      if (round + kRoundBatch >= 0xffff) return set_err(NPHYS_ERR_STATE, "colouring did not converge");
      for (int k = 0; k < kRoundBatch; ++k) launch_colour_round(b, c, M, round + k, v, stream);
      launch_colour_keys(M, v, ncol, stream);
      const int rc = readback({{v.remaining, 2, 0}, {ncol, 1, 2}});
      if (rc) return rc;
      if (rb_host[0] == 0) break;
    }
    ncolours = int(rb_host[2]);
    colour_rounds = int(rb_host[1]);
    if (ncolours > kMaxColours) return set_err(NPHYS_ERR_STATE, "too many colours");
    launch_colour_order(b, c, M, ncolours, v, stream);
    uint32_t* class_rec = v.col_start + (kMaxColours + 2);
    launch_fix_colour_starts(ncolours, M, C, v, class_rec, stream);
    // class offsets (slots and row records) to the host, read after the assembly is enqueued
    class_seq = rb_post({{v.col_start, ncolours + 1, kRbClass}, {class_rec, ncolours + 1, kRbClass + kMaxColours + 2}});
    return NPHYS_OK;
  }

  // Visit order of this solve's colour classes (the SplitMix64 Fisher-Yates shuffle seeded by the
  // solve index, computed on the host) and the largest class (sizes the sweep grids), from the
  // class offsets read back in colour().
  int visit_order() {
    const int rc = rb_wait(class_seq);
    if (rc) return rc;
    const uint32_t* h_colstart = rb_host + kRbClass;
    int csize[kMaxColours];
    for (int c = 0; c < ncolours; ++c) csize[c] = int(h_colstart[c + 1] - h_colstart[c]);
    colour_visit_order(uint64_t(solve_count), ncolours, csize, colour_visit);
    max_class = 0;
    int* h_range = h_visit + (kMaxColours + 2);
    for (int k = 0; k < ncolours; ++k) {
      const int c = colour_visit[k];
      const int first = int(h_colstart[c]), size = int(h_colstart[c + 1] - h_colstart[c]);
      max_class = std::max(max_class, size);
      h_visit[k] = c;
      h_range[2 * k] = first;
      h_range[2 * k + 1] = size;
    }
    // h_visit is rewritten only by the next solve, after these copies have completed (that solve's
    // class readback is enqueued after them in stream order)
    HIPCK(hipMemcpyAsync(sv_visit.p, h_visit, size_t(ncolours) * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(sv_visit.as<int>() + (kMaxColours + 2), h_range, size_t(ncolours) * 2 * sizeof(int),
                         hipMemcpyHostToDevice, stream));
    return NPHYS_OK;
  }

  // Warm start (when `warm`) and `nsweeps` coloured SOR-Prox sweeps. With counters enabled the
  // per-colour sweep kernels carry HIP events in their dispatch packets. Every launch is direct:
  // one per colour class, or one workgroup for a run of consecutive small classes.
  int sweeps(const DevBodies& b, const DevContacts& cv, SolverBufs& v, int sblocks, bool warm, int nsweeps) {
    if (cur_variant > 0) {
      // scene-blocked: one launch for the warm start and all nsweeps sweeps
      SceneBlocks sc = sc_view();
      if (counters_enabled) {
        while (kt_ev.size() < size_t(kt_used) + 2) {
          hipEvent_t e;
          HIPCK(hipEventCreate(&e));
          kt_ev.push_back(e);
        }
        HIPCK(hipEventRecord(kt_ev[kt_used], stream));
      }
      HIPCK(hipEventRecord(fork_ev, stream));
      HIPCK(hipStreamWaitEvent(stream_large, fork_ev, 0));
      launch_sweeps_blocked(b, v, sc, ncolours, nsweeps, warm, cur_variant, stream, stream_large);
      HIPCK(hipEventRecord(join_ev, stream_large));
      HIPCK(hipStreamWaitEvent(stream, join_ev, 0));
      if (counters_enabled) {
        HIPCK(hipEventRecord(kt_ev[kt_used + 1], stream));
        kt_used += 2;
      }
      return NPHYS_OK;
    }
    // Direct launches (the visiting order changes every solve, so the launch sequence does): two or
    // more consecutive small classes of the visiting order run back to back in one workgroup. With
    // counters enabled each k_pgs launch carries HIP events in its dispatch packet.
    if (counters_enabled) {
      const size_t need = size_t(kt_used) + size_t(2) * size_t(nsweeps) * size_t(ncolours);
      while (kt_ev.size() < need) {
        hipEvent_t e;
        HIPCK(hipEventCreate(&e));
        kt_ev.push_back(e);
      }
    }
    const int* h_range = h_visit + (kMaxColours + 2);
    const uint32_t* class_rec = rb_host + kRbClass + kMaxColours + 2;
    auto pass = [&](bool w) {
      for (int k = 0; k < ncolours;) {
        int e = k;
        while (e < ncolours && h_range[2 * e + 1] <= kRunBlock) ++e;
        if (e - k >= 2) {
          launch_colour_run(b, v, k, e - k, w, stream);
          k = e;
          continue;
        }
        // one 64-lane workgroup per 64 manifolds of this class (the grid-stride loop covers the
        // rest above the cap): no idle workgroups for the smaller classes
        const int cblocks = std::min(sblocks, std::max(1, (h_range[2 * k + 1] + 63) / 64));
        // the sweep kernel runs two lanes per manifold
        const int pblocks2 = std::min(2 * sblocks, std::max(1, (2 * h_range[2 * k + 1] + 63) / 64));
        if (w) {
          launch_warmstart_colour(b, cv, h_range[2 * k], h_range[2 * k + 1], cblocks, v, stream);
        } else if (counters_enabled) {
          launch_pgs_colour(b, cv, h_range[2 * k], h_range[2 * k + 1], pblocks2, v, stream, kt_ev[kt_used],
                            kt_ev[kt_used + 1]);
          kt_used += 2;
          const int c = h_visit[k];
          counters.sweep_kernel_contacts += int64_t(class_rec[c + 1]) - int64_t(class_rec[c]);
        } else {
          launch_pgs_colour(b, cv, h_range[2 * k], h_range[2 * k + 1], pblocks2, v, stream);
        }
        ++k;
      }
    };
    // the joints' rows: warm-started after the contacts' (sor_prox.rs:56-58), solved before the
    // contacts' in every sweep (sor_prox.rs:167-168)
    JointBufs jv = joint_view();
    const bool joints = nj > 0 && !ccd_sub;  // a CCD substep solves no joints (mechanical_world.rs:909)
    if (warm) {
      pass(true);
      if (joints) launch_joint_warm(b, jv, stream);
    }
    for (int it = 0; it < nsweeps; ++it) {
      if (joints) launch_joint_sweep(b, jv, stream);
      pass(false);
    }
    return NPHYS_OK;
  }

  // Residual of the fixed-count mode: the partials are copied into pinned memory without waiting
  // and reduced after the step's final synchronisation (finish_residual), since nothing in the step
  // needs them.
  int residual_deferred(const DevBodies& b, const DevContacts& cv, int M, int C, SolverBufs& v) {
    launch_residual(b, cv, M, v, stream);
    const int nbk = residual_blocks(M);
    JointBufs jv = joint_view();
    if (nj > 0) launch_joint_residual(b, jv, v.red + 2 * nbk, stream);
    const size_t need = size_t(nbk) * 2 + (nj > 0 ? 3 : 0);
    if (need > red_pinned_cap) {
      HIPCK(hipStreamSynchronize(stream));  // the previous buffer may be the target of a copy
      if (red_pinned) (void)hipHostFree(red_pinned);
      red_pinned = nullptr;
      red_pinned_cap = 0;
      HIPCK(hipHostMalloc(reinterpret_cast<void**>(&red_pinned), std::max<size_t>(need, 1024) * sizeof(double)));
      red_pinned_cap = std::max<size_t>(need, 1024);
    }
    HIPCK(hipMemcpyAsync(red_pinned, v.red, need * sizeof(double), hipMemcpyDeviceToHost, stream));
    res_pending_blocks = nbk;
    res_pending_C = C;
    res_pending_joints = nj > 0;
    return NPHYS_OK;
  }
  bool res_pending_joints = false;

  // Reduces a deferred residual; the stream must have passed its copy.
  void finish_residual() {
    if (res_pending_blocks < 0) return;
    double mx = 0.0, ss = 0.0, rows = double(3 * res_pending_C);
    for (int k = 0; k < res_pending_blocks; ++k) {
      mx = std::max(mx, red_pinned[2 * k]);
      ss += red_pinned[2 * k + 1];
    }
    if (res_pending_joints) {  // the joint rows' partial after the contacts' blocks
      const double* jp = red_pinned + 2 * res_pending_blocks;
      mx = std::max(mx, jp[0]);
      ss += jp[1];
      rows += jp[2];
    }
    res_inf = mx;
    res_rms = rows > 0.0 ? std::sqrt(ss / rows) : 0.0;
    res_pending_blocks = -1;
  }

  // Natural-map residual of the current impulses: per-block partials reduced on the host in
  // block order (max is exact; the sum of squares is the RMS).
  int residual(const DevBodies& b, const DevContacts& cv, int M, int C, SolverBufs& v) {
    launch_residual(b, cv, M, v, stream);
    const int nbk = residual_blocks(M);
    JointBufs jv = joint_view();
    if (nj > 0) launch_joint_residual(b, jv, v.red + 2 * nbk, stream);
    red_host.resize(size_t(nbk) * 2 + 3);
    HIPCK(hipMemcpyAsync(red_host.data(), v.red, (size_t(nbk) * 2 + (nj > 0 ? 3 : 0)) * sizeof(double),
                         hipMemcpyDeviceToHost, stream));
    HIPCK(hipStreamSynchronize(stream));
    double mx = 0.0, ss = 0.0, rows = double(3 * C);
    for (int k = 0; k < nbk; ++k) {
      mx = std::max(mx, red_host[2 * k]);
      ss += red_host[2 * k + 1];
    }
    if (nj > 0) {
      mx = std::max(mx, red_host[2 * nbk]);
      ss += red_host[2 * nbk + 1];
      rows += red_host[2 * nbk + 2];
    }
    res_inf = mx;
    res_rms = rows > 0.0 ? std::sqrt(ss / rows) : 0.0;
    return NPHYS_OK;
  }

  SceneBlocks sc_view() {
    SceneBlocks v;
    v.keys = sc_keys.as<uint32_t>();
    v.slot = sc_slot.as<uint32_t>();
    v.keys_alt = sc_keys_alt.as<uint32_t>();
    v.slot_alt = sc_slot_alt.as<uint32_t>();
    v.start = sc_start.as<uint32_t>();
    v.dyn_start = sc_dyn_start.as<int>();
    v.dyn_list = sc_dyn_list.as<int>();
    v.local = sc_local.as<int>();
    v.small_list = sc_small.as<int>();
    v.large_list = sc_large.as<int>();
    v.n_small = n_small_scenes;
    v.n_large = n_large_scenes;
    return v;
  }

  bool blocked_eligible(int ncol) const {
    return max_scene_dyn <= kLdsBodies && int64_t(nscenes) * (ncol + 1) <= (int64_t(1) << 24);
  }

  // Variant of this solve's sweeps: forced, or the autotuner's pick, or (while tuning) the next
  // candidate to time. All variants give identical results.
  void choose_variant(int M) {
    tune_pending = false;
    // the scene-blocked kernels solve contacts only: worlds with joints use per-colour launches
    if (nj > 0 || !blocked_eligible(ncolours)) {
      cur_variant = 0;
      return;
    }
    if (variant_request >= 0) {
      cur_variant = variant_request;
      return;
    }
    // a pick is kept while the problem stays of the size it was timed at (a batch of falling
    // scenes has few manifolds at first): retime when the manifold count leaves [2/3, 3/2] of it
    if (tuned && (2 * int64_t(M) > 3 * int64_t(tune_M) || 3 * int64_t(M) < 2 * int64_t(tune_M))) {
      tune_cands.clear();
      tuned = false;
    }
    if (tune_cands.empty()) {
      tune_M = M;
      tune_cands = {0, 256, 512, 1024};
      tune_time.assign(tune_cands.size(), 0.0);
      tune_count.assign(tune_cands.size(), 0);
    }
    if (!tuned) {
      size_t k = 0;
      while (k < tune_cands.size() && tune_count[k] >= kTuneReps) ++k;
      if (k < tune_cands.size()) {
        cur_variant = tune_cands[k];
        tune_pending = true;
        return;
      }
      size_t best = 0;
      for (size_t i = 1; i < tune_cands.size(); ++i)
        if (tune_time[i] < tune_time[best]) best = i;
      cur_variant = tune_cands[best];
      tuned = true;
    }
  }

  void tune_record() {
    if (!tune_pending) return;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, tune_ev[0], tune_ev[1]) != hipSuccess) return;
    for (size_t i = 0; i < tune_cands.size(); ++i)
      if (tune_cands[i] == cur_variant) {
        tune_time[i] += double(ms);
        ++tune_count[i];
      }
    tune_pending = false;
  }

  // Warm start and SOR-Prox sweeps: the fixed count of the reference (sor_prox.rs:56-79), or, in
  // solve-to-residual mode, sweeps until max |phi| <= res_target, checked after the warm start
  // and every res_check_every sweeps, at most res_max_sweeps. Leaves the final residual in
  // res_inf / res_rms and the sweep count in nsw.
  int solve_sweeps(const DevBodies& b, const DevContacts& cv, SolverBufs& v, int M, int C, int sblocks, int& nsw) {
    int rc;
    if (res_target <= 0.0) {
      if ((rc = sweeps(b, cv, v, sblocks, true, prm.max_velocity_iterations))) return rc;
      nsw = prm.max_velocity_iterations;
      return residual_deferred(b, cv, M, C, v);
    }
    if ((rc = sweeps(b, cv, v, sblocks, true, 0))) return rc;
    int done = 0;
    for (;;) {
      if ((rc = residual(b, cv, M, C, v))) return rc;
      if ((res_norm == NPHYS_RESIDUAL_RMS ? res_rms : res_inf) <= res_target || done >= res_max_sweeps) break;
      const int k = std::min(res_check_every, res_max_sweeps - done);
      if ((rc = sweeps(b, cv, v, sblocks, false, k))) return rc;
      done += k;
    }
    nsw = done;
    return NPHYS_OK;
  }

  // The network's impulses for every row record (features already computed) as the warm start.
  int predict(SolverBufs& v, int C) {
    HIPCK(ml_y.ensure(size_t(C) * kOut * sizeof(float)));
    if (counters_enabled) {
      for (auto& e : ml_ev)
        if (!e) HIPCK(hipEventCreate(&e));
      HIPCK(hipEventRecord(ml_ev[0], stream));
    }
    mlp_infer(MlpDev{ml_img.as<uint16_t>(), ml_params.as<float>()}, ml_X.as<float>(), C, ml_y.as<float>(), stream);
    if (counters_enabled) {
      HIPCK(hipEventRecord(ml_ev[1], stream));
      ml_ev_pending = true;
      ml_rows_pending = C;
    }
    launch_apply_prediction(v, C, ml_y.as<float>(), ml_ls.as<float>(), stream);
    return NPHYS_OK;
  }
  hipEvent_t ml_ev[2] = {nullptr, nullptr};
  bool ml_ev_pending = false;
  int ml_rows_pending = 0;

  // Activation islands and the solved-manifold list (sleeping worlds): sets cv, M and C to the
  // manifolds with an awake dynamic side; the others' cached impulses are carried over.
  int solve_list(const DevBodies& b, ContactBufs& cc, ContactBufs& pc, int prevM, DevContacts& cv, int& M, int& C) {
    const size_t m = size_t(std::max(cc.M, 1)), nn = size_t(std::max(n, 1));
    HIPCK(act_parent.ensure(nn * 4));
    HIPCK(act_can.ensure(nn * 4));
    HIPCK(act_touch.ensure(nn * 4));
    HIPCK(act_mflag[cur].ensure(m * 4));
    HIPCK(act_pos.ensure(m * 4));
    HIPCK(act_tmp.ensure(scan_temp_words(int(m)) * 4 + 64));
    HIPCK(sl_mbody[cur].ensure(m * sizeof(int2)));
    HIPCK(sl_mfirst[cur].ensure(m * 4));
    HIPCK(sl_mcount[cur].ensure(m * 4));
    ActivationBufs a;
    a.parent = act_parent.as<int>();
    a.can = act_can.as<uint32_t>();
    a.touch = act_touch.as<uint32_t>();
    a.mflag = act_mflag[cur].as<uint32_t>();
    a.pos = act_pos.as<uint32_t>();
    a.tmp = act_tmp.as<uint32_t>();
    a.totals = bp_small.as<uint32_t>() + 48;
    fill_u32(a.totals, 0u, 3, stream);
    const DevContacts full = cc.view();
    const JointBufs jv = joint_view();
    launch_activation(b, full, cc.M, a, nj > 0 ? &jv : nullptr, stream);
    DevContacts out = full;
    out.bfirst = nullptr;  // the solved list has no per-body index
    out.bcount = 0;
    out.mbody = sl_mbody[cur].as<int2>();
    out.mfirst = sl_mfirst[cur].as<int>();
    out.mcount = sl_mcount[cur].as<int>();
    launch_solve_list(full, cc.M, a, out, pc.view(), prevM, stream);
    const int rc = readback({{a.totals, 3, 0}});
    if (rc) return rc;
    M = int(rb_host[0]);
    C = int(rb_host[1]);
    active_bodies = int64_t(rb_host[2]);
    cv = out;
    return NPHYS_OK;
  }

  int solve() {
    ++solve_count;  // counts every solve, as the checker does
    ContactBufs& cc = cb[cur];
    ContactBufs& pc = cb[cur ^ 1];
    const int prevM = have_solved ? pc.M : 0;
    DevBodies b = bodies();
    DevContacts cv = cc.view();
    int M = cc.M, C = cc.C;
    if (sleeping) {
      const int rc = solve_list(b, cc, pc, prevM, cv, M, C);
      if (rc) return rc;
    } else {
      active_bodies = n_dynamic;
    }
    sview[cur] = cv;
    sview_M[cur] = M;
    sview_C[cur] = C;
    sview_compact[cur] = sleeping;
    solved_M = M;
    solved_C = C;
    res_inf = res_rms = 0.0;
    ncolours = 0;
    max_class = 0;
    if (M == 0 && nj == 0) {
      last_sweeps = res_target > 0.0 ? 0 : prm.max_velocity_iterations;
      counters.velocity_sweeps += last_sweeps;
      HIPCK(hipEventRecord(ev[EV_ASM], stream));
      HIPCK(hipEventRecord(ev[EV_VEL], stream));
      return NPHYS_OK;
    }
    int rc = ensure_solver(M, C);
    if (rc) return rc;
    SolverBufs v = sv_view();
    if (M > 0) {
      if ((rc = colour(b, cv, M, C, v))) return rc;
      launch_assemble(b, cv, pc.view(), prevM, M, sp, v, mat_table(), stream);
    }
    // the joints' rows (moreau_jean_solver.rs:211-223); the island joints are those neither broken
    // nor removed with an awake dynamic body, decided on the device
    JointBufs jv = joint_view();
    if (nj > 0) launch_joint_assemble(b, sp, jv, stream);
    HIPCK(hipEventRecord(ev[EV_ASM], stream));
    // class sizes (read back while the assembly runs) give the visit order and the sweep grids
    if (M > 0 && (rc = visit_order())) return rc;
    const int sblocks = sweep_blocks(max_class);
    kt_used = 0;
    choose_variant(M);
    if (tune_pending) HIPCK(hipEventRecord(tune_ev[0], stream));
    if (cur_variant > 0) {
      HIPCK(sc_keys.ensure(size_t(M) * 4));
      HIPCK(sc_slot.ensure(size_t(M) * 4));
      HIPCK(sc_keys_alt.ensure(size_t(M) * 4));
      HIPCK(sc_slot_alt.ensure(size_t(M) * 4));
      HIPCK(sc_start.ensure(size_t(nscenes) * size_t(ncolours + 1) * 4));
      SceneBlocks sc = sc_view();
      launch_scene_order(b, M, nscenes, ncolours, v, sc, stream);
    }
    const bool learned = contact_model == NPHYS_CONTACT_MODEL_LEARNED_WARMSTART && C > 0;
    if ((learned || capture) && C > 0) {
      HIPCK(ml_X.ensure(size_t(C) * kFeat * sizeof(float)));
      HIPCK(ml_ls.ensure(size_t(C) * sizeof(float)));
      launch_features(b, cv, sp, M, v, ml_X.as<float>(), ml_ls.as<float>(), stream);
    }
    if (learned && judge && res_target > 0.0) {
      // The plain cache warm start is the judge: both solves run from the same assembled system
      // (the same cached impulses, the prediction added to them for the second) and the
      // prediction is kept only when it reaches the residual target in fewer sweeps.
      HIPCK(jd_lam.ensure(size_t(C) * sizeof(float4)));
      HIPCK(jd_lam0.ensure(size_t(C) * sizeof(float4)));
      HIPCK(jd_dv.ensure(size_t(n) * 2 * sizeof(float4)));
      int n_plain = 0, n_learned = 0;
      launch_lam_copy(v, C, jd_lam0.as<float4>(), true, stream);  // the cache warm start
      // the joints' rows start from their cached impulses in both solves
      const size_t jl = size_t(nj) * 2 * sizeof(float4);
      if (nj > 0) {
        HIPCK(jd_jlam.ensure(2 * jl));
        HIPCK(hipMemcpyAsync(jd_jlam.p, j_lam.p, jl, hipMemcpyDeviceToDevice, stream));
      }
      if ((rc = solve_sweeps(b, cv, v, M, C, sblocks, n_plain))) return rc;
      const double rinf_p = res_inf, rrms_p = res_rms;
      launch_lam_copy(v, C, jd_lam.as<float4>(), true, stream);
      launch_lam_copy(v, C, jd_lam0.as<float4>(), false, stream);
      if (nj > 0) {
        HIPCK(hipMemcpyAsync(jd_jlam.as<char>() + jl, j_lam.p, jl, hipMemcpyDeviceToDevice, stream));
        HIPCK(hipMemcpyAsync(j_lam.p, jd_jlam.p, jl, hipMemcpyDeviceToDevice, stream));
      }
      // the velocity increment is the last 24 bytes of each 64-byte solver block
      char* dv0 = sol.as<char>() + 10 * sizeof(float);
      const size_t pitch = kSolverBlock * sizeof(float4), dvb = 6 * sizeof(float);
      HIPCK(hipMemcpy2DAsync(jd_dv.p, dvb, dv0, pitch, dvb, size_t(n), hipMemcpyDeviceToDevice, stream));
      HIPCK(hipMemset2DAsync(dv0, pitch, 0, dvb, size_t(n), stream));
      if ((rc = predict(v, C))) return rc;
      if ((rc = solve_sweeps(b, cv, v, M, C, sblocks, n_learned))) return rc;
      const bool keep = n_learned < n_plain;
      if (!keep) {
        launch_lam_copy(v, C, jd_lam.as<float4>(), false, stream);
        if (nj > 0) HIPCK(hipMemcpyAsync(j_lam.p, jd_jlam.as<char>() + jl, jl, hipMemcpyDeviceToDevice, stream));
        HIPCK(hipMemcpy2DAsync(dv0, pitch, jd_dv.p, dvb, dvb, size_t(n), hipMemcpyDeviceToDevice, stream));
        res_inf = rinf_p;
        res_rms = rrms_p;
      }
      counters.judge_plain_sweeps += n_plain;
      counters.judge_learned_sweeps += n_learned;
      counters.judge_kept += keep ? 1 : 0;
      last_sweeps = n_plain + n_learned;
    } else {
      if (learned && (rc = predict(v, C))) return rc;
      if ((rc = solve_sweeps(b, cv, v, M, C, sblocks, last_sweeps))) return rc;
    }
    if (capture && C > 0) {
      HIPCK(ml_t.ensure(size_t(C) * kOut * sizeof(float)));
      launch_sample_targets(v, C, ml_ls.as<float>(), ml_t.as<float>(), stream);
      samples = C;
    }
    counters.velocity_sweeps += last_sweeps;
    if (C > 0) launch_cache_impulses(cv, v, C, stream);
    if (nj > 0) launch_joint_cache(b, sp, jv, stream);
    HIPCK(hipEventRecord(ev[EV_VEL], stream));
    if (keep_colours && M > 0) {
      std::vector<uint32_t> st(M);
      HIPCK(hipMemcpyAsync(st.data(), sv_state2[cur].p, size_t(M) * 4, hipMemcpyDeviceToHost, stream));
      HIPCK(hipStreamSynchronize(stream));
      last_colours.resize(M);
      for (int m = 0; m < M; ++m) last_colours[m] = int(st[m] & 0xffffu);
    }
    return NPHYS_OK;
  }

  int position_correction(const SimParams& ps, int iterations) {
    const bool joints = nj > 0 && !ccd_sub;
    if (sview_M[cur] == 0 && !joints) return NPHYS_OK;
    DevBodies b = bodies();
    SolverBufs v = sv_view();
    JointBufs jv = joint_view();
    const DevContacts cv = sview[cur];
    if (cur_variant > 0 && !joints) {
      // scene-blocked, as the sweeps of this solve (worlds without joints)
      SceneBlocks sc = sc_view();
      HIPCK(hipEventRecord(fork_ev, stream));
      HIPCK(hipStreamWaitEvent(stream_large, fork_ev, 0));
      launch_position_blocked(b, ps, v, sc, ncolours, iterations, cur_variant, stream, stream_large);
      HIPCK(hipEventRecord(join_ev, stream_large));
      HIPCK(hipStreamWaitEvent(stream, join_ev, 0));
      return NPHYS_OK;
    }
    const int pblocks = position_blocks(max_class);
    // direct launches, runs of two or more consecutive small classes in one workgroup (as the
    // sweeps); the joints' constraints first in every iteration (nonlinear_sor_prox.rs:33-38)
    const int* h_range = h_visit + (kMaxColours + 2);
    const int ncol = sview_M[cur] > 0 ? ncolours : 0;
    for (int it = 0; it < iterations; ++it) {
      if (joints) launch_joint_position(b, ps, jv, stream);
      for (int k = 0; k < ncol;) {
        int e = k;
        while (e < ncolours && h_range[2 * e + 1] <= kRunBlock) ++e;
        if (e - k >= 2) {
          launch_position_run(b, ps, v, k, e - k, stream);
          k = e;
          continue;
        }
        launch_position_colour(b, cv, ps, h_range[2 * k], h_range[2 * k + 1],
                               std::min(pblocks, std::max(1, (h_range[2 * k + 1] + 63) / 64)), v, stream);
        ++k;
      }
    }
    return NPHYS_OK;
  }


  // ------------------------------------------------------------------ CCD
  int ensure_ccd() {
    if (n >= kSlotStatic) return set_err(NPHYS_ERR_CAPACITY, "CCD supports fewer than 2^30 bodies");
    const size_t nn = size_t(std::max(n, 1)), P = size_t(std::max(npairs, 1));
    for (DBuf* b : {&ccd_time, &ccd_frozen, &ccd_in, &ccd_new, &ccd_prep, &ccd_prepared, &ccd_deg, &ccd_off, &ccd_fill})
      HIPCK(b->ensure(nn * 4));
    HIPCK(ccd_adj.ensure(2 * P * 4));
    for (DBuf* b : {&ccd_slot, &ccd_pos, &ccd_eval, &ccd_kept, &ccd_cand, &ccd_cpos, &ccd_lpair})
      HIPCK(b->ensure(P * 4));
    HIPCK(ccd_list.ensure(P * sizeof(int2)));
    HIPCK(ccd_sbodies.ensure(P * sizeof(int4)));
    HIPCK(ccd_stage.ensure(P * sizeof(ToiEntry)));
    // the queue holds at most one entry per CCD slot (at most P); the impact list takes 3P + 4096
    // confirmed impacts (a full list is an error, never a silent loss)
    const size_t cap = 3 * P + 4096;
    HIPCK(ccd_heap.ensure(cap * sizeof(ToiEntry)));
    HIPCK(ccd_imp.ensure(cap * sizeof(unsigned long long)));
    ccd_cap = int(std::min<size_t>(ccd_heap.bytes / sizeof(ToiEntry), ccd_imp.bytes / sizeof(unsigned long long)));
    HIPCK(ccd_ctl.ensure(kCcdWords * 4));
    HIPCK(ccd_tmp.ensure(std::max(scan_temp_words(int(P)), scan_temp_words(int(nn))) * 4 + 64));
    return NPHYS_OK;
  }

  CcdBufs ccd_view() {
    CcdBufs c;
    c.body_time = ccd_time.as<float>();
    c.frozen = ccd_frozen.as<float>();
    c.slot_pair = ccd_slot.as<int>();
    c.slot_bodies = ccd_sbodies.as<int4>();
    c.pos = ccd_pos.as<int>();
    c.heap = ccd_heap.as<ToiEntry>();
    c.stage = ccd_stage.as<ToiEntry>();
    c.impacts = ccd_imp.as<unsigned long long>();
    c.ctl = ccd_ctl.as<uint32_t>();
    c.cap = ccd_cap;
    c.in_t = ccd_in.as<uint32_t>();
    c.new_t = ccd_new.as<uint32_t>();
    c.prep = ccd_prep.as<uint32_t>();
    c.prepared = ccd_prepared.as<uint32_t>();
    c.evaluated = ccd_eval.as<uint32_t>();
    c.kept = ccd_kept.as<uint32_t>();
    c.cand = ccd_cand.as<uint32_t>();
    c.cpos = ccd_cpos.as<uint32_t>();
    c.list = ccd_list.as<int2>();
    c.list_pair = ccd_lpair.as<int>();
    c.tmp = ccd_tmp.as<uint32_t>();
    c.deg = ccd_deg.as<uint32_t>();
    c.off = ccd_off.as<uint32_t>();
    c.fill = ccd_fill.as<uint32_t>();
    c.adj = ccd_adj.as<int>();
    return c;
  }

  // The CCD phase's sync and broad phase with swept AABBs (position0 -> position).
  int ccd_broad_phase() {
    uint32_t ch = 0;
    int rc = loose_pass(2, &ch);
    if (rc) return rc;
    counters.loose_changes += int64_t(ch);
    if (ch || pairs_dirty) {
      if ((rc = rebuild_pairs())) return rc;
      ++counters.pair_rebuilds;
    }
    return ensure_ccd();
  }

  // predict_next_impacts up to end_time: sets *kind (PREDICT_*), the impact count, the last
  // impact time and min_overstep.
  int ccd_predict(float end_time, float dt0, int* kind) {
    CcdBufs c = ccd_view();
    launch_ccd_predict(bodies(), pairs.as<int2>(), npairs, c, end_time, dt0, sp.allowed_linear_error,
                       prm.ccd_on_penetration_enabled, stream);
    const int rc = readback({{c.ctl, kCcdWords, 0}});
    if (rc) return rc;
    if (rb_host[kCcdOverflow]) return set_err(NPHYS_ERR_CAPACITY, "CCD impact queue overflow");
    *kind = int(rb_host[kCcdKind]);
    ccd_nimp = *kind == PREDICT_IMPACTS ? int(rb_host[kCcdImpacts]) : 0;
    uint32_t u = rb_host[kCcdLastToi];
    std::memcpy(&ccd_last_toi, &u, 4);
    u = rb_host[kCcdMinOver];
    std::memcpy(&ccd_min_over, &u, 4);
    return NPHYS_OK;
  }

  // One substep (mechanical_world.rs:630-935): the island moved to the last impact time, its
  // manifolds at the moved poses, then MoreauJeanSolver::step_ccd over dt0 - last_toi: assembly
  // without warm start (coefficient 0), max_ccd_position_iterations position iterations, the
  // impacting bodies' advancement validated, the velocity sweeps, integration of the island.
  int ccd_substep(float dt0) {
    int rc;
    ++counters.ccd_substeps;
    counters.ccd_impacts += ccd_nimp;
    const float last_toi = ccd_last_toi;
    const float dt_sub = dt0 - last_toi;
    const DevBodies b = bodies();
    CcdBufs c = ccd_view();
    const size_t nn = size_t(n), P = size_t(std::max(npairs, 1));
    for (uint32_t* a : {c.in_t, c.new_t, c.prep, c.prepared}) fill_u32(a, 0u, nn, stream);
    for (uint32_t* a : {c.evaluated, c.kept}) fill_u32(a, 0u, P, stream);
    launch_ccd_seed(b, pairs.as<int2>(), c, ccd_nimp, stream);
    for (;;) {  // island rounds
      launch_ccd_touch(b, pairs.as<int2>(), npairs, c, stream);
      launch_ccd_prepare(b, c, dt_sub, last_toi, stream);
      launch_ccd_compact(pairs.as<int2>(), npairs, c.cand, c, stream);
      if ((rc = readback({{c.ctl + kCcdList, 1, 0}}))) return rc;
      const int nl = int(rb_host[0]);
      if (nl == 0) break;
      if ((rc = narrow_phase_of(c.list, nl, ccd_cb))) return rc;
      launch_ccd_grow(b, c, nl, np_count.as<uint32_t>(), stream);
      launch_ccd_merge(n, c, stream);
    }
    // the island's manifolds, in pair order, into cb[cur] (the final narrow phase rewrites it)
    launch_ccd_compact(pairs.as<int2>(), npairs, c.kept, c, stream);
    if ((rc = readback({{c.ctl + kCcdList, 1, 0}}))) return rc;
    const int nk = int(rb_host[0]);
    ContactBufs& cc = cb[cur];
    if ((rc = narrow_phase_of(c.list, nk, cc))) return rc;
    const int M = cc.M, C = cc.C;
    SimParams ss = sp;
    ss.dt = dt_sub;
    ss.inv_dt = dt_sub == 0.0f ? 0.0f : 1.0f / dt_sub;
    ss.warmstart_coeff = 0.0f;
    ss.gx = ss.gy = ss.gz = 0.0f;
    ss.dyn_dt = dt_sub;
    const DevContacts cv = cc.view();
    ++solve_count;
    ccd_sub = true;
    const int saved_variant = cur_variant;
    cur_variant = 0;
    SolverBufs v{};
    if (M > 0) {
      if ((rc = ensure_solver(M, C))) return rc;
      v = sv_view();
      if ((rc = colour(b, cv, M, C, v))) return rc;
      const ContactBufs& pc = cb[cur ^ 1];
      launch_assemble(b, cv, pc.view(), have_solved ? pc.M : 0, M, ss, v, mat_table(), stream);
      if ((rc = visit_order())) return rc;
      sview[cur] = cv;
      sview_M[cur] = M;
      sview_C[cur] = C;
      if ((rc = position_correction(ss, prm.max_ccd_position_iterations))) return rc;
    }
    launch_ccd_validate(b, pairs.as<int2>(), c, ccd_nimp, stream);
    if (M > 0 && (rc = sweeps(b, cv, v, sweep_blocks(max_class), true, prm.max_velocity_iterations))) return rc;
    launch_integrate(b, ss, stream, c.prepared);
    ccd_sub = false;
    cur_variant = saved_variant;
    // every moved body's dynamics are updated with the substep's time step after it (:915-923)
    dyn_dt_next = dt_sub;
    return NPHYS_OK;
  }

  // MechanicalWorld::solve_ccd (mechanical_world.rs:559-943).
  int solve_ccd() {
    int rc;
    const float dt0 = sp.dt;
    if (dt0 == 0.0f || prm.max_ccd_substeps == 0) return NPHYS_OK;
    const float sub_len = dt0 / float(prm.max_ccd_substeps);
    if ((rc = ensure_ccd())) return rc;
    fill_u32(ccd_time.as<uint32_t>(), 0u, size_t(n), stream);
    float end_time = 0.0f;
    bool last_iter = false;
    for (;;) {
      if (end_time < dt0) end_time = fminf_(end_time + sub_len, dt0);
      else last_iter = true;
      if ((rc = ccd_broad_phase())) return rc;
      int kind = PREDICT_NONE;
      if ((rc = ccd_predict(end_time, dt0, &kind))) return rc;
      if (kind == PREDICT_AFTER_END) {
        end_time = std::floor(ccd_min_over / sub_len) * sub_len;
        end_time = fminf_(end_time + sub_len, dt0);
        if ((rc = ccd_predict(end_time, dt0, &kind))) return rc;
        if (kind != PREDICT_IMPACTS) return set_err(NPHYS_ERR_STATE, "CCD: no impact in the window of the earliest one");
      }
      if (!last_iter && ccd_nimp > 0) {
        if ((rc = ccd_substep(dt0))) return rc;
        continue;
      }
      CcdBufs c = ccd_view();
      launch_ccd_clamp(bodies(), c, stream);  // the last prediction's frozen bodies
      return NPHYS_OK;
    }
  }

  // ------------------------------------------------------------------ step
  int step_once() {
    int rc;
    HIPCK(hipEventRecord(ev[EV_START], stream));
    // bodies moved from the host are updated at their new state with this step's dt; the others
    // keep the time step of their last update (a CCD substep's after one)
    SimParams spp = sp;
    spp.dyn_dt = moved ? sp.dt : dyn_dt_next;
    dyn_dt_next = sp.dt;
    if (moved || recollide) {
      if ((rc = collide_now(0))) return rc;
      moved = false;
      recollide = false;
    }
    HIPCK(hipEventRecord(ev[EV_PRE_CD], stream));
    if (wake_pending) {
      launch_wake(bodies(), stream);
      wake_pending = false;
    }
    launch_prestep(bodies(), spp, stream);
    if ((rc = solve())) return rc;
    launch_integrate(bodies(), sp, stream);
    HIPCK(hipEventRecord(ev[EV_INT], stream));
    if ((rc = position_correction(sp, prm.max_position_iterations))) return rc;
    HIPCK(hipEventRecord(ev[EV_POS], stream));
    if (tune_pending) HIPCK(hipEventRecord(tune_ev[1], stream));  // the variant's sweeps and position passes
    // this step's manifolds and impulses become the warm-start cache
    cur ^= 1;
    have_solved = true;
    if (any_ccd) {
      // the CCD phase (its swept syncs and broad phases, impacts and substeps), then the final sync
      if ((rc = solve_ccd())) return rc;
      HIPCK(hipEventRecord(ev[EV_CCD], stream));
      uint32_t ch = 0;
      if ((rc = loose_pass(1, &ch))) return rc;
      counters.loose_changes += int64_t(ch);
      if (ch || pairs_dirty) {
        if ((rc = rebuild_pairs())) return rc;
        ++counters.pair_rebuilds;
      }
    } else {
      HIPCK(hipEventRecord(ev[EV_CCD], stream));
      uint32_t ch1 = 0, ch2 = 0;
      if ((rc = loose_passes_end_of_step(&ch1, &ch2))) return rc;  // CCD-phase sync, then the final one
      counters.loose_changes += int64_t(ch1) + int64_t(ch2);
      if (ch1 || ch2 || pairs_dirty) {
        if ((rc = rebuild_pairs())) return rc;
        ++counters.pair_rebuilds;
      }
    }
    HIPCK(hipEventRecord(ev[EV_BP], stream));
    // the narrow phase's readback completes the step: every earlier event and copy is done after it
    if ((rc = narrow_phase(ev[EV_NP]))) return rc;
    finish_residual();
    auto el = [&](int a, int c) {
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, ev[a], ev[c]);
      return double(ms) * 1e-3;
    };
    counters.step += el(EV_START, EV_NP);
    counters.collision_detection += el(EV_START, EV_PRE_CD) + el(EV_CCD, EV_NP);
    counters.broad_phase += el(EV_CCD, EV_BP);
    counters.ccd_time += el(EV_POS, EV_CCD);
    counters.narrow_phase += el(EV_BP, EV_NP);
    counters.assembly += el(EV_PRE_CD, EV_ASM);
    counters.velocity_resolution += el(EV_ASM, EV_VEL);
    counters.velocity_update += el(EV_VEL, EV_INT);
    counters.position_resolution += el(EV_INT, EV_POS);
    counters.nsteps += 1;
    counters.npairs = npairs;
    counters.ncontacts = solved_C;
    counters.nmanifolds = solved_M;
    counters.active_bodies = active_bodies;
    counters.nconstraints = int64_t(3) * solved_C;
    counters.ncolours = ncolours;
    counters.colour_rounds += colour_rounds;
    tune_record();
    counters.sweep_variant = cur_variant;
    for (int k = 0; k < kt_used; k += 2) {
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, kt_ev[k], kt_ev[k + 1]);
      counters.sweep_kernel_time += double(ms) * 1e-3;
      counters.sweep_kernel_launches += 1;
    }
    kt_used = 0;
    if (ml_ev_pending) {  // the step's final synchronisation has passed the inference
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ml_ev[0], ml_ev[1]) == hipSuccess) {
        counters.warmstart_inference_time += double(ms) * 1e-3;
        counters.warmstart_inference_rows += ml_rows_pending;
      }
      ml_ev_pending = false;
    }
    prm.t += prm.dt;
    // clear_forces (mechanical_world.rs:365): the next accumulation starts from zero
    if (forces_pending) {
      fill_u32(f_lin.as<uint32_t>(), 0u, size_t(n) * 4, stream);
      fill_u32(f_ang.as<uint32_t>(), 0u, size_t(n) * 4, stream);
      forces_pending = false;
    }
    rot_dyn_saved = false;
    return NPHYS_OK;
  }

  // Body::apply_force / apply_force_at_point for `count` entries: grouped by body on the host (a
  // stable sort keeps each body's entries in the caller's order), one thread per body.
  int apply_force(int32_t count, const int32_t* bodies_, const float* forces, const float* points, int32_t type,
                  int32_t wake) {
    if (count < 0 || type < 0 || type > 3) return set_err(NPHYS_ERR_INVALID, "bad force arguments");
    if (count == 0) return NPHYS_OK;
    if (!bodies_ || !forces) return set_err(NPHYS_ERR_INVALID, "null force arrays");
    for (int k = 0; k < count; ++k)
      if (bodies_[k] < 0 || bodies_[k] >= n) return set_err(NPHYS_ERR_INVALID, "bad body handle");
    std::vector<int> order(static_cast<size_t>(count));
    for (int k = 0; k < count; ++k) order[size_t(k)] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return bodies_[a] < bodies_[c]; });
    std::vector<int> ub, off;
    for (int k = 0; k < count; ++k) {
      const int bd = bodies_[order[size_t(k)]];
      if (ub.empty() || ub.back() != bd) {
        ub.push_back(bd);
        off.push_back(k);
      }
    }
    off.push_back(count);
    const int nu = int(ub.size());
    const int fw = points ? 3 : 6;
    const size_t words = size_t(nu) + size_t(nu + 1) + size_t(count) + size_t(count) * fw + (points ? size_t(count) * 3 : 0);
    HIPCK(af_tmp.ensure(words * 4));
    char* base = af_tmp.as<char>();
    int* d_ub = reinterpret_cast<int*>(base);
    int* d_off = d_ub + nu;
    int* d_order = d_off + nu + 1;
    float* d_f = reinterpret_cast<float*>(d_order + count);
    float* d_p = points ? d_f + size_t(count) * fw : nullptr;
    HIPCK(hipMemcpyAsync(d_ub, ub.data(), size_t(nu) * 4, hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(d_off, off.data(), size_t(nu + 1) * 4, hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(d_order, order.data(), size_t(count) * 4, hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(d_f, forces, size_t(count) * fw * 4, hipMemcpyHostToDevice, stream));
    if (points) HIPCK(hipMemcpyAsync(d_p, points, size_t(count) * 12, hipMemcpyHostToDevice, stream));
    if (type == NPHYS_FORCE || type == NPHYS_ACCELERATION_CHANGE) forces_pending = true;
    SimParams spf = sp;
    spf.dyn_dt = dyn_dt_next;  // the time step of the augmented mass the forces act through
    launch_apply_force(bodies(), spf, nu, d_ub, d_off, d_order, d_f, d_p, type, wake, rot_dyn_saved ? rot_dyn.as<float4>() : nullptr,
                       stream);
    HIPCK(hipStreamSynchronize(stream));  // the staging buffer is reused by the next call
    return NPHYS_OK;
  }

  // MaterialsCoefficientsTable::set_friction_coefficient / set_restitution_coefficient.
  int set_material_coefficient(int32_t which, int32_t id1, int32_t id2, float coeff) {
    if ((which != 0 && which != 1) || id1 < 0 || id2 < 0) return set_err(NPHYS_ERR_INVALID, "bad material pair");
    const unsigned long long lo = unsigned(std::min(id1, id2)), hi = unsigned(std::max(id1, id2));
    const unsigned long long key = (lo << 32) | hi;
    auto& t = mt_host[which];
    auto it = std::lower_bound(t.begin(), t.end(), std::make_pair(key, -std::numeric_limits<float>::infinity()),
                               [](const std::pair<unsigned long long, float>& a,
                                  const std::pair<unsigned long long, float>& c) { return a.first < c.first; });
    if (it != t.end() && it->first == key) {
      it->second = coeff;
    } else {
      if (t.size() >= NPHYS_MAX_MATERIAL_PAIRS) return set_err(NPHYS_ERR_CAPACITY, "material table full");
      t.insert(it, std::make_pair(key, coeff));
    }
    std::vector<unsigned long long> keys(t.size());
    std::vector<float> vals(t.size());
    for (size_t i = 0; i < t.size(); ++i) {
      keys[i] = t[i].first;
      vals[i] = t[i].second;
    }
    HIPCK(hipStreamSynchronize(stream));  // the previous table may be read by enqueued work
    HIPCK(mt_key[which].ensure(keys.size() * 8));
    HIPCK(mt_val[which].ensure(vals.size() * 4));
    HIPCK(hipMemcpyAsync(mt_key[which].p, keys.data(), keys.size() * 8, hipMemcpyHostToDevice, stream));
    HIPCK(hipMemcpyAsync(mt_val[which].p, vals.data(), vals.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCK(hipStreamSynchronize(stream));
    return NPHYS_OK;
  }
};

}  // namespace nph

using nph::World;
using nph::set_err;

struct nphys_world {
  World w;
};

extern "C" {

void nphys_params_default(nphys_params* out) {
  if (!out) return;
  std::memset(out, 0, sizeof(*out));
  out->t = 0.0;
  out->dt = 1.0 / 60.0;
  out->erp = 0.2;
  out->warmstart_coeff = 1.0;
  out->restitution_velocity_threshold = 1.0;
  out->allowed_linear_error = 0.001;
  out->allowed_angular_error = 0.001;
  out->max_linear_correction = 0.2;
  out->max_angular_correction = 0.2;
  out->max_stabilization_multiplier = 0.2;
  out->max_velocity_iterations = 8;
  out->max_position_iterations = 3;
  out->max_ccd_position_iterations = 10;
  out->max_ccd_substeps = 1;
  out->gravity[0] = 0.0;
  out->gravity[1] = -9.81;
  out->gravity[2] = 0.0;
}

static int check_params(const nphys_params* p) {
  if (!p) return set_err(NPHYS_ERR_INVALID, "null params");
  if (!(p->dt >= 0.0)) return set_err(NPHYS_ERR_INVALID, "dt must be non-negative");
  if (p->max_velocity_iterations < 0 || p->max_position_iterations < 0)
    return set_err(NPHYS_ERR_INVALID, "negative iteration count");
  return NPHYS_OK;
}

int nphys_world_create(const nphys_params* params, nphys_world** out) {
  if (!out) return set_err(NPHYS_ERR_INVALID, "null out");
  *out = nullptr;
  int rc = check_params(params);
  if (rc) return rc;
  nphys_world* w = new nphys_world();
  rc = w->w.init(*params);
  if (rc) {
    delete w;
    return rc;
  }
  *out = w;
  return NPHYS_OK;
}

int nphys_world_destroy(nphys_world* w) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  if (w->w.stream) (void)hipStreamSynchronize(w->w.stream);
  delete w;
  return NPHYS_OK;
}

int nphys_set_params(nphys_world* w, const nphys_params* params) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  const int rc = check_params(params);
  if (rc) return rc;
  w->w.set_params(*params);
  return NPHYS_OK;
}

int nphys_get_params(const nphys_world* w, nphys_params* out) {
  if (!w || !out) return set_err(NPHYS_ERR_INVALID, "null argument");
  *out = w->w.prm;
  return NPHYS_OK;
}

int nphys_add_bodies(nphys_world* w, const nphys_body_desc* desc, int32_t* first_handle) {
  if (!w || !desc) return set_err(NPHYS_ERR_INVALID, "null argument");
  return w->w.add_bodies(*desc, first_handle);
}

int nphys_add_joints(nphys_world* w, const nphys_joint_desc* desc, int32_t* first_handle) {
  if (!w || !desc) return set_err(NPHYS_ERR_INVALID, "null argument");
  return w->w.add_joints(*desc, first_handle);
}

int nphys_remove_joint(nphys_world* w, int32_t joint) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  return w->w.remove_joint(joint);
}

int nphys_num_joints(const nphys_world* w, int32_t* n) {
  if (!w || !n) return set_err(NPHYS_ERR_INVALID, "null argument");
  *n = w->w.nj;
  return NPHYS_OK;
}

int nphys_read_joints(nphys_world* w, float* lin_impulses, float* ang_impulses, int32_t* broken, int64_t cap) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  return w->w.read_joints(lin_impulses, ang_impulses, broken, cap);
}

int nphys_num_bodies(const nphys_world* w, int32_t* n) {
  if (!w || !n) return set_err(NPHYS_ERR_INVALID, "null argument");
  *n = w->w.n;
  return NPHYS_OK;
}

int nphys_set_state(nphys_world* w, const float* pos, const float* rot, const float* linvel,
                    const float* angvel) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  return w->w.set_state(pos, rot, linvel, angvel);
}

int nphys_set_collision_groups(nphys_world* w, int32_t count, const int32_t* bodies, const uint32_t* groups) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  return w->w.set_collision_groups(count, bodies, groups);
}

int nphys_get_state(nphys_world* w, float* pos, float* rot, float* linvel, float* angvel) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  return w->w.get_state(pos, rot, linvel, angvel);
}

int nphys_apply_force(nphys_world* w, int32_t count, const int32_t* bodies, const float* forces, int32_t type,
                      int32_t auto_wake_up) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  return w->w.apply_force(count, bodies, forces, nullptr, type, auto_wake_up);
}

int nphys_apply_force_at_point(nphys_world* w, int32_t count, const int32_t* bodies, const float* forces,
                               const float* points, int32_t type, int32_t auto_wake_up) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  if (count > 0 && !points) return set_err(NPHYS_ERR_INVALID, "null points");
  return w->w.apply_force(count, bodies, forces, points, type, auto_wake_up);
}

int nphys_set_material_coefficient(nphys_world* w, int32_t which, int32_t id1, int32_t id2, float coeff) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  return w->w.set_material_coefficient(which, id1, id2, coeff);
}

int nphys_step(nphys_world* w, int32_t nsteps) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  if (nsteps < 0) return set_err(NPHYS_ERR_INVALID, "negative step count");
  if (w->w.any_ccd && (w->w.sleeping || w->w.nscenes > 1))
    return set_err(NPHYS_ERR_INVALID, "CCD needs one scene and no sleep thresholds");
  for (int i = 0; i < nsteps; ++i) {
    const int rc = w->w.step_once();
    if (rc) return rc;
  }
  return NPHYS_OK;
}

int nphys_collide(nphys_world* w) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  const int rc = w->w.collide_now(0);
  if (rc) return rc;
  w->w.moved = false;
  w->w.recollide = false;
  if (hipStreamSynchronize(w->w.stream) != hipSuccess) return set_err(NPHYS_ERR_HIP, "sync failed");
  return NPHYS_OK;
}

int nphys_contact_model_set(nphys_world* w, int32_t model) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  if (model != NPHYS_CONTACT_MODEL_SIGNORINI_COULOMB_PYRAMID && model != NPHYS_CONTACT_MODEL_LEARNED_WARMSTART)
    return set_err(NPHYS_ERR_INVALID, "unknown contact model");
  if (model == NPHYS_CONTACT_MODEL_LEARNED_WARMSTART && !w->w.ml_loaded)
    return set_err(NPHYS_ERR_STATE, "the learned warm start needs a network (nphys_warmstart_load)");
  w->w.contact_model = model;
  return NPHYS_OK;
}

int nphys_warmstart_load(nphys_world* w, const float* params, int64_t n) {
  if (!w || !params) return set_err(NPHYS_ERR_INVALID, "null argument");
  if (n != nph::kNumParams) return set_err(NPHYS_ERR_INVALID, "parameter count mismatch");
  for (int i = 0; i < nph::kFeat; ++i)
    if (!(params[nph::kOffXScale + i] > 0.0f)) return set_err(NPHYS_ERR_INVALID, "feature scales must be > 0");
  for (int i = 0; i < nph::kOut; ++i)
    if (!(params[nph::kOffYScale + i] > 0.0f)) return set_err(NPHYS_ERR_INVALID, "output scales must be > 0");
  World& W = w->w;
  HIPCK(W.ml_params.ensure(size_t(nph::kNumParams) * 4));
  HIPCK(W.ml_img.ensure(size_t(nph::kImgElems) * 2));
  HIPCK(hipMemcpyAsync(W.ml_params.p, params, size_t(nph::kNumParams) * 4, hipMemcpyHostToDevice, W.stream));
  nph::mlp_pack_images(W.ml_params.as<float>(), W.ml_img.as<uint16_t>(), W.stream);
  HIPCK(hipStreamSynchronize(W.stream));
  W.ml_loaded = true;
  return NPHYS_OK;
}

int nphys_capture_samples(nphys_world* w, int32_t on) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  w->w.capture = on != 0;
  w->w.samples = 0;
  return NPHYS_OK;
}

int nphys_read_samples(nphys_world* w, float* features, float* targets, int64_t cap, int64_t* n) {
  if (!w || !n) return set_err(NPHYS_ERR_INVALID, "null argument");
  World& W = w->w;
  *n = W.samples;
  if (!features && !targets) return NPHYS_OK;
  if (cap < W.samples) return set_err(NPHYS_ERR_CAPACITY, "sample buffers too small");
  if (W.samples == 0) return NPHYS_OK;
  if (features)
    HIPCK(hipMemcpyAsync(features, W.ml_X.p, size_t(W.samples) * nph::kFeat * 4, hipMemcpyDeviceToHost, W.stream));
  if (targets)
    HIPCK(hipMemcpyAsync(targets, W.ml_t.p, size_t(W.samples) * nph::kOut * 4, hipMemcpyDeviceToHost, W.stream));
  HIPCK(hipStreamSynchronize(W.stream));
  return NPHYS_OK;
}

int nphys_set_sweep_variant(nphys_world* w, int32_t variant) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  if (variant != -1 && variant != 0 && variant != 256 && variant != 512 && variant != 1024)
    return set_err(NPHYS_ERR_INVALID, "variant must be -1 (autotune), 0, 256, 512 or 1024");
  w->w.variant_request = variant;
  w->w.tune_cands.clear();
  w->w.tuned = false;
  return NPHYS_OK;
}

int nphys_set_incremental_pairs(nphys_world* w, int32_t on) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  w->w.incremental_bp = on != 0;
  return NPHYS_OK;
}

int nphys_set_warmstart_judge(nphys_world* w, int32_t on) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  w->w.judge = on != 0;
  return NPHYS_OK;
}

int nphys_num_pairs(nphys_world* w, int64_t* n) {
  if (!w || !n) return set_err(NPHYS_ERR_INVALID, "null argument");
  *n = w->w.npairs;
  return NPHYS_OK;
}

int nphys_read_pairs(nphys_world* w, int32_t* pairs, int64_t cap) {
  if (!w || (!pairs && w->w.npairs)) return set_err(NPHYS_ERR_INVALID, "null argument");
  if (cap < w->w.npairs) return set_err(NPHYS_ERR_CAPACITY, "pair buffer too small");
  if (w->w.npairs == 0) return NPHYS_OK;
  if (hipMemcpy(pairs, w->w.pairs.p, size_t(w->w.npairs) * sizeof(int2), hipMemcpyDeviceToHost) != hipSuccess)
    return set_err(NPHYS_ERR_HIP, "pair copy failed");
  return NPHYS_OK;
}

int nphys_num_contacts(nphys_world* w, int64_t* n) {
  if (!w || !n) return set_err(NPHYS_ERR_INVALID, "null argument");
  *n = w->w.cb[w->w.cur].C;
  return NPHYS_OK;
}

int nphys_read_contacts(nphys_world* w, int32_t* bodies, uint32_t* feature, float* world1,
                        float* world2, float* normal, float* depth, int64_t cap) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  nph::ContactBufs& c = w->w.cb[w->w.cur];
  if (cap < c.C) return set_err(NPHYS_ERR_CAPACITY, "contact buffer too small");
  if (c.C == 0) return NPHYS_OK;
  std::vector<int2> mb;
  std::vector<int> mc;
  std::vector<float4> rec;
  int rc = w->w.gather_contacts(c, mb, mc, rec, nullptr);
  if (rc) return rc;
  long i = 0;
  for (size_t m = 0; m < mb.size(); ++m)
    for (int k = 0; k < mc[m]; ++k, ++i) {
      const float4* r = &rec[size_t(i) * nph::kContactFields];
      if (bodies) { bodies[2 * i] = mb[m].x; bodies[2 * i + 1] = mb[m].y; }
      if (feature) feature[i] = nph::bits_of(r[1].w);
      if (world1) { world1[3 * i] = r[0].x; world1[3 * i + 1] = r[0].y; world1[3 * i + 2] = r[0].z; }
      if (world2) { world2[3 * i] = r[1].x; world2[3 * i + 1] = r[1].y; world2[3 * i + 2] = r[1].z; }
      if (normal) { normal[3 * i] = r[2].x; normal[3 * i + 1] = r[2].y; normal[3 * i + 2] = r[2].z; }
      if (depth) depth[i] = r[0].w;
    }
  return NPHYS_OK;
}

int nphys_num_manifolds(nphys_world* w, int64_t* n) {
  if (!w || !n) return set_err(NPHYS_ERR_INVALID, "null argument");
  *n = w->w.cb[w->w.cur].M;
  return NPHYS_OK;
}

int nphys_read_manifolds(nphys_world* w, int32_t* bodies, int32_t* first, int32_t* count, int64_t cap) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  nph::ContactBufs& c = w->w.cb[w->w.cur];
  if (cap < c.M) return set_err(NPHYS_ERR_CAPACITY, "manifold buffer too small");
  if (c.M == 0) return NPHYS_OK;
  std::vector<int2> mb(static_cast<size_t>(c.M));
  std::vector<int> mc(static_cast<size_t>(c.M));
  HIPCK(hipMemcpyAsync(mb.data(), c.mbody.p, mb.size() * sizeof(int2), hipMemcpyDeviceToHost, w->w.stream));
  HIPCK(hipMemcpyAsync(mc.data(), c.mcount.p, mc.size() * 4, hipMemcpyDeviceToHost, w->w.stream));
  HIPCK(hipStreamSynchronize(w->w.stream));
  int32_t acc = 0;
  for (size_t m = 0; m < mb.size(); ++m) {
    if (bodies) {
      bodies[2 * m] = mb[m].x;
      bodies[2 * m + 1] = mb[m].y;
    }
    if (first) first[m] = acc;
    if (count) count[m] = mc[m];
    acc += mc[m];
  }
  return NPHYS_OK;
}

int nphys_num_solved_contacts(nphys_world* w, int64_t* n) {
  if (!w || !n) return set_err(NPHYS_ERR_INVALID, "null argument");
  *n = w->w.have_solved ? w->w.sview_C[w->w.cur ^ 1] : 0;
  return NPHYS_OK;
}

int nphys_read_impulses(nphys_world* w, uint64_t* keys, float* impulses, int64_t cap) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  if (!w->w.have_solved) return NPHYS_OK;
  const int par = w->w.cur ^ 1;
  nph::ContactBufs& c = w->w.cb[par];
  if (cap < w->w.sview_C[par]) return set_err(NPHYS_ERR_CAPACITY, "impulse buffer too small");
  if (w->w.sview_C[par] == 0) return NPHYS_OK;
  std::vector<int2> mb;
  std::vector<int> mc;
  std::vector<float4> rec, imp;
  int rc = w->w.gather_contacts(c, mb, mc, rec, &imp);
  if (rc) return rc;
  // sleeping worlds: only the manifolds handed to the solver
  std::vector<uint32_t> solved;
  if (w->w.sview_compact[par]) {
    solved.resize(mb.size());
    if (!solved.empty()) {
      HIPCK(hipMemcpyAsync(solved.data(), w->w.act_mflag[par].p, solved.size() * 4, hipMemcpyDeviceToHost,
                           w->w.stream));
      HIPCK(hipStreamSynchronize(w->w.stream));
    }
  }
  long i = 0, j = 0;
  for (size_t m = 0; m < mb.size(); ++m)
    for (int k = 0; k < mc[m]; ++k, ++j) {
      if (!solved.empty() && !solved[m]) continue;
      const uint32_t f = nph::bits_of(rec[size_t(j) * nph::kContactFields + 1].w);
      if (keys)
        keys[i] = (uint64_t(uint32_t(mb[m].x)) << 40) | (uint64_t(uint32_t(mb[m].y)) << 16) | uint64_t(f & 0xffffu);
      if (impulses) {
        impulses[3 * i] = imp[size_t(j)].x;
        impulses[3 * i + 1] = imp[size_t(j)].y;
        impulses[3 * i + 2] = imp[size_t(j)].z;
      }
      ++i;
    }
  return NPHYS_OK;
}

int nphys_residual(nphys_world* w, double* r_inf, double* r_rms) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  if (r_inf) *r_inf = w->w.res_inf;
  if (r_rms) *r_rms = w->w.res_rms;
  return NPHYS_OK;
}

int nphys_read_counters(nphys_world* w, nphys_counters* out) {
  if (!w || !out) return set_err(NPHYS_ERR_INVALID, "null argument");
  *out = w->w.counters;
  return NPHYS_OK;
}

int nphys_reset_counters(nphys_world* w) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  std::memset(&w->w.counters, 0, sizeof(w->w.counters));
  return NPHYS_OK;
}

int nphys_counters_enable(nphys_world* w, int32_t enabled) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  w->w.counters_enabled = enabled != 0;
  return NPHYS_OK;
}

int nphys_set_residual_target(nphys_world* w, double target, int32_t max_sweeps, int32_t check_every,
                              int32_t norm) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  if (!(target >= 0.0) || max_sweeps < 0 || check_every < 1 ||
      (norm != NPHYS_RESIDUAL_MAX && norm != NPHYS_RESIDUAL_RMS))
    return set_err(NPHYS_ERR_INVALID, "target >= 0, max_sweeps >= 0, check_every >= 1, norm MAX or RMS");
  w->w.res_target = target;
  w->w.res_max_sweeps = max_sweeps;
  w->w.res_check_every = check_every;
  w->w.res_norm = norm;
  return NPHYS_OK;
}

int nphys_read_activation(nphys_world* w, float* energy, int64_t cap) {
  if (!w || !energy) return set_err(NPHYS_ERR_INVALID, "null argument");
  return w->w.read_activation(energy, cap);
}

int nphys_read_colours(nphys_world* w, int32_t* colours, int64_t cap, int32_t* ncolours) {
  if (!w) return set_err(NPHYS_ERR_INVALID, "null world");
  w->w.keep_colours = true;
  const int M = int(w->w.last_colours.size());
  if (ncolours) *ncolours = w->w.ncolours;
  if (!colours) return NPHYS_OK;
  if (cap < M) return set_err(NPHYS_ERR_CAPACITY, "colour buffer too small");
  for (int m = 0; m < M; ++m) colours[m] = w->w.last_colours[m];
  return NPHYS_OK;
}

const char* nphys_last_error(void) { return nph::g_err.c_str(); }

}  // extern "C"
