// oracle/orc_world.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// A CPU restatement of one nphysics 0.23 MechanicalWorld::step over rigid balls and cuboids,
// templated on the scalar type (nphysics is generic over `N: RealField`). It is the checker for
// the HIP pipeline and the timed CPU baseline; nothing in the product library uses it.
//
// Stage order follows src/world/mechanical_world.rs:199-396 (step_with_filter):
//   step_started + update_dynamics (:230-234), update_acceleration (:241-243), pre-solve
//   sync/broad/narrow phase (:250-253), manifold collection (:287-300), MoreauJeanSolver::step
//   (:316-326), post-solve update_dynamics (:343-346), the no-op CCD sync + broad phase
//   (:604-606, swept AABBs), validate_advancement (:364-367), the real CD of the step (:375-385).
// Solver: src/solver/moreau_jean_solver.rs:47-90,129-347; contact model
//   src/solver/signorini_coulomb_pyramid_model.rs:56-261, src/solver/signorini_model.rs:37-197,
//   src/solver/helper.rs:53-135, src/object/rigid_body.rs:672-722 (fill_constraint_geometry);
//   SOR-Prox src/solver/sor_prox.rs:19-436; position correction
//   src/solver/nonlinear_sor_prox.rs:17-55,121-309.
// Bodies: src/object/rigid_body.rs:305-315 (set_position), :371-381 (apply_displacement),
//   :467-505 (integrate), :558-619 (update_dynamics / update_acceleration), :747-766
//   (add_local_inertia_and_com); mass properties src/volumetric/volumetric_ball.rs:9-44,
//   src/volumetric/volumetric_cuboid.rs:10-75; material combine src/material/material.rs:72-178.
//
// Collision groups in the pair rule restate ncollide3d 0.31's CollisionGroups (unvendored): pinned
// by a truth table and the reference's collision groups example; parity with ncollide's own code is
// unpinned.
//
// Two row orders are provided for the velocity and position sweeps:
//   ORDER_REFERENCE -- the reference's: friction rows (non-ground, then ground), then normal rows
//                      (non-ground, then ground), manifolds in pair order (sor_prox.rs:159-179);
//   ORDER_COLOURED  -- the same rows visited manifold by manifold in the deterministic greedy
//                      colouring order the GPU uses (friction rows of the manifold's contacts,
//                      then its normal rows). Used to check the GPU row arithmetic tightly.
#pragma once
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <queue>
#include <string>
#include <unordered_map>
#include <vector>

#include "orc_geometry.hpp"

namespace orc {

enum BodyStatus : int { STATUS_DYNAMIC = 0, STATUS_STATIC = 1 };
enum OrderMode : int { ORDER_REFERENCE = 0, ORDER_COLOURED = 1 };

// BodyUpdateStatus flags (src/object/body.rs:450-527).
enum : unsigned {
  F_POSITION = 1u,
  F_VELOCITY = 2u,
  F_LOCAL_INERTIA = 4u,
  F_LOCAL_COM = 8u,
  F_DAMPING = 16u,
  F_STATUS = 32u,
  F_ALL = 63u
};

template <class T>
struct Params {  // IntegrationParameters (src/solver/integration_parameters.rs:5-76)
  T t, dt, inv_dt;
  T erp, warmstart_coeff, restitution_velocity_threshold;
  T allowed_linear_error, allowed_angular_error, max_linear_correction, max_angular_correction;
  T max_stabilization_multiplier;
  int max_velocity_iterations, max_position_iterations, max_ccd_position_iterations;
  int max_ccd_substeps;
  int ccd_on_penetration = 0;  // ccd_on_penetration_enabled (integration_parameters.rs:56-69)
  void set_dt(T d) {
    dt = d;
    inv_dt = d == T(0) ? T(0) : T(1) / d;
  }
};

template <class T>
struct Body {
  int status, shape;
  int scene;  // bodies of different scenes never interact
  ShapeDesc<T> sd;
  T density, margin, lin_pred, friction, restitution;
  Iso<T> pos, pos0;
  V3<T> linvel, angvel;
  V3<T> acc_lin, acc_ang;
  T mass;
  M3<T> local_I, I_world, aug, inv_aug;
  T inv_mass;
  V3<T> com;
  bool gravity;
  T lin_damping, ang_damping, max_linvel, max_angvel;
  // jacobian_mask (rigid_body.rs:46,78,103-118): 1 for a free dof, 0 for a kinematic one
  V3<T> mask_lin, mask_ang;
  // external_forces (rigid_body.rs:771-822), cleared at the end of every step (mechanical_world.rs:365)
  V3<T> f_lin, f_ang;
  // BasicMaterial (basic_material.rs:14-63): combine modes, lookup id (-1: None), surface velocity
  // in the collider's frame
  int fric_mode, rest_mode, mat_id;
  V3<T> surf_vel;
  int companion;
  unsigned flags;
  // CCD: ColliderDesc::ccd_enabled (collider.rs:70), RigidBody::linear_motion_interpolation_enabled
  // (rigid_body.rs:44,91-101), and the key of the body the collider belongs to (the body itself, or
  // for a static collider that shares a static body with others, the first of them)
  bool ccd, lin_interp;
  int ccd_key;
  // ColliderDesc::collision_groups (collider.rs:444): membership, whitelist, blacklist masks
  uint32_t g_member, g_white, g_black;
  AABB<T> loose;
  T energy;     // ActivationStatus::energy (0: asleep)
  T sleep_thr;  // ActivationStatus::threshold, < 0 for None
  T query_limit() const { return margin + lin_pred; }
};

struct Manifold {
  int b1, b2;
  int first, count;
};

template <class T>
struct VelRow {  // Unilateral/Bilateral(+Ground)Constraint, src/solver/constraint.rs
  T impulse, r, rhs;
  uint64_t key;
  int a1, a2;
  size_t j1, j2, wj1, wj2;
  int dep;   // friction rows: index of the normal row in its list
  T mu;
  int ground;
};

template <class T>
struct PosRow {  // NonlinearUnilateralConstraint, src/solver/nonlinear_constraint.rs:74-105
  int b1, b2, ndofs1, ndofs2;
  Contact<T> kin;
  T r, rhs;
};

struct ContactRows {
  int ground;
  int normal;      // index in uni or uni_g
  int fric[2];     // index in bil or bil_g
};

// A joint constraint (src/joint/): BallConstraint (ball_constraint.rs), FixedConstraint
// (fixed_constraint.rs) or RevoluteConstraint without limits or motor (revolute_constraint.rs).
enum { JOINT_BALL = 0, JOINT_FIXED = 1, JOINT_REVOLUTE = 2, JOINT_PRISMATIC = 3, JOINT_CYLINDRICAL = 4 };
template <class T>
struct Joint {
  int kind, b1, b2;
  V3<T> a1, a2;       // anchors in the bodies' local frames
  Q<T> f1, f2;        // fixed: reference frames ref_frame1/2
  V3<T> ax1, ax2;     // revolute: unit axes in the local frames
  V3<T> lin_imp{T(0), T(0), T(0)}, ang_imp{T(0), T(0), T(0)};  // impulses cached by the last solve
  T break_force_sq, break_torque_sq;
  bool broken = false, removed = false;
  int colour = 0;
};

// Rows of one active joint: linear rows along x, y, z at the anchors (cancel_relative_linear_velocity,
// helper.rs:247-296), then angular rows (fixed: x, y, z, cancel_relative_angular_velocity,
// helper.rs:502-551; revolute: the two directions orthogonal to the world axis,
// restrict_relative_angular_velocity_to_axis, helper.rs:614-699).
struct JointRows {
  int joint, ground, n;
  int row[6];  // indices in jbil or jbil_g
};

struct Counters {  // stage names of src/counters/mod.rs:148-209
  double step = 0, collision_detection = 0, broad_phase = 0, narrow_phase = 0, assembly = 0,
         velocity_resolution = 0, velocity_update = 0, position_resolution = 0, ccd = 0;
  long nsteps = 0;
};

inline double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline uint64_t contact_key(int a, int b, uint32_t feature) {
  return (uint64_t(uint32_t(a)) << 40) | (uint64_t(uint32_t(b)) << 16) | uint64_t(feature & 0xffffu);
}

// 32-bit integer hash (bijective mixer) used as the colouring priority.
inline uint32_t colour_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// Order in which the colour classes of solve number `solve_index` are visited: a Fisher-Yates
// shuffle of 0..ncolours-1 driven by SplitMix64 seeded with the solve index, then stably sorted by
// size bucket floor(log2(class size)), smallest first. Visiting the classes in a new order every
// step re-randomises the relative order of conflicting manifolds of similar-sized classes, as a
// fresh colouring would, while colours themselves persist between steps; the largest classes come
// last in every sweep, so most rows end a sweep freshly solved (a lower residual at equal
// iteration count than a plain shuffle). `size` holds the manifolds per class.
inline int size_bucket(int n) {
  int l = 0;
  while (l < 30 && (2 << l) <= n) ++l;
  return l;
}
inline void colour_visit_order(uint64_t solve_index, int ncolours, const std::vector<int>& size,
                               std::vector<int>& perm) {
  perm.resize(ncolours);
  for (int c = 0; c < ncolours; ++c) perm[c] = c;
  uint64_t st = solve_index * 0x9E3779B97F4A7C15ull;
  for (int i = ncolours - 1; i > 0; --i) {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const int j = int(z % uint64_t(i + 1));
    std::swap(perm[i], perm[j]);
  }
  std::stable_sort(perm.begin(), perm.end(),
                   [&](int a, int b) { return size_bucket(size[size_t(a)]) < size_bucket(size[size_t(b)]); });
}

// Deterministic Jones-Plassmann greedy colouring of manifolds; two manifolds conflict when they
// share a dynamic body. Entries of `colour` that are >= 0 on entry are kept (colours inherited
// from the previous step). Priority (hash(m), m); in each round every uncoloured manifold that
// beats all its uncoloured neighbours takes the smallest colour unused by its coloured
// neighbours. The result depends only on the inputs, never on evaluation order. Returns one more
// than the largest colour.
inline int colour_manifolds(const std::vector<Manifold>& ms, const std::vector<int>& is_dyn,
                            int nbodies, std::vector<int>& colour) {
  const int M = int(ms.size());
  if (int(colour.size()) != M) colour.assign(M, -1);
  std::vector<int> off(nbodies + 1, 0);
  for (const auto& m : ms) {
    if (is_dyn[m.b1]) off[m.b1 + 1]++;
    if (is_dyn[m.b2]) off[m.b2 + 1]++;
  }
  for (int i = 0; i < nbodies; ++i) off[i + 1] += off[i];
  std::vector<int> adj(off[nbodies]);
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int m = 0; m < M; ++m) {
    if (is_dyn[ms[m].b1]) adj[fill[ms[m].b1]++] = m;
    if (is_dyn[ms[m].b2]) adj[fill[ms[m].b2]++] = m;
  }
  std::vector<uint32_t> pri(M);
  for (int m = 0; m < M; ++m) pri[m] = colour_hash(uint32_t(m));
  auto beats = [&](int x, int y) { return pri[x] > pri[y] || (pri[x] == pri[y] && x > y); };
  int remaining = 0, ncol = 0;
  for (int m = 0; m < M; ++m) {
    if (colour[m] < 0) ++remaining;
    else ncol = std::max(ncol, colour[m] + 1);
  }
  std::vector<int> sel;
  while (remaining > 0) {
    sel.clear();
    for (int m = 0; m < M; ++m) {
      if (colour[m] >= 0) continue;
      bool best = true;
      const int bs[2] = {ms[m].b1, ms[m].b2};
      for (int s = 0; s < 2 && best; ++s) {
        const int b = bs[s];
        if (!is_dyn[b]) continue;
        for (int k = off[b]; k < off[b + 1]; ++k) {
          const int o = adj[k];
          if (o != m && colour[o] < 0 && beats(o, m)) { best = false; break; }
        }
      }
      if (best) sel.push_back(m);
    }
    for (int m : sel) {
      // smallest colour not used by a coloured neighbour
      std::vector<char> used;
      const int bs[2] = {ms[m].b1, ms[m].b2};
      for (int s = 0; s < 2; ++s) {
        const int b = bs[s];
        if (!is_dyn[b]) continue;
        for (int k = off[b]; k < off[b + 1]; ++k) {
          const int c = colour[adj[k]];
          if (c >= 0) {
            if (c >= int(used.size())) used.resize(c + 1, 0);
            used[c] = 1;
          }
        }
      }
      int c = 0;
      while (c < int(used.size()) && used[c]) ++c;
      colour[m] = c;
      ncol = std::max(ncol, c + 1);
    }
    remaining -= int(sel.size());
  }
  return ncol;
}

enum CombineMode : int { COMBINE_AVERAGE = 0, COMBINE_MIN = 1, COMBINE_MULTIPLY = 2, COMBINE_MAX = 3 };

// MaterialCombineMode::combine (material.rs:72-86): the mode of higher precedence decides
// (Max > Multiply > Min > Average).
template <class T>
inline T combine_coeff(T a, int ma, T b, int mb) {
  if (ma == COMBINE_MAX || mb == COMBINE_MAX) return tmax(a, b);
  if (ma == COMBINE_MULTIPLY || mb == COMBINE_MULTIPLY) return a * b;
  if (ma == COMBINE_MIN || mb == COMBINE_MIN) return tmin(a, b);
  return (a + b) * T(0.5);
}

// Per-body attributes beyond shape, pose and BasicMaterial coefficients, with the
// RigidBodyDesc / BasicMaterial defaults (rigid_body.rs:972-990, basic_material.rs:33-43).
template <class T>
struct BodyExtras {
  T lin_damping = T(0), ang_damping = T(0);
  T max_linvel = std::numeric_limits<T>::max(), max_angvel = std::numeric_limits<T>::max();
  bool gravity = true;
  int kinematic_dofs = 0;  // bit k: dof k kinematic (0-2 translations, 3-5 rotations)
  int fric_mode = COMBINE_AVERAGE, rest_mode = COMBINE_AVERAGE, mat_id = -1;
  V3<T> surf_vel{T(0), T(0), T(0)};
  bool ccd = false, lin_interp = false;
  int static_body = -1;  // static colliders with one index >= 0 (per scene) belong to one body
  // CollisionGroups::new(): member of groups 0-29, whitelisting them all, blacklisting none
  uint32_t g_member = (1u << 30) - 1u, g_white = (1u << 30) - 1u, g_black = 0u;
};

template <class T>
class World {
 public:
  using Scalar = T;
  Params<T> prm;
  V3<T> gravity;
  std::vector<Body<T>> bodies;
  int order_mode = ORDER_REFERENCE;

  // geometrical world
  std::vector<std::pair<int, int>> pairs;
  std::vector<Manifold> manifolds;  // pairs with >= 1 contact, pair order
  std::vector<Contact<T>> contacts;
  std::unordered_map<uint64_t, V3<T>> cache;  // ImpulseCache: ContactId -> (n, t1, t2)

  // solver state of the last step
  std::vector<T> jacobians, mj_lambda, ext_vels;
  std::vector<VelRow<T>> uni, uni_g, bil, bil_g;
  std::vector<PosRow<T>> posc;
  std::vector<Manifold> solved;        // manifolds handed to the solver (first: row index)
  std::vector<Contact<T>> solved_contacts;  // their contacts, when a subset is solved
  std::vector<char> manifold_active;   // per manifold: handed to the solver this step
  bool any_sleeper = false;            // some dynamic body has a sleep threshold
  std::vector<ContactRows> crow;       // per contact of `solved`
  // joints (JointConstraintSet), their colouring (joints sharing a dynamic body conflict; fixed
  // greedy colours in joint order, recomputed when joints are added) and the rows of the joints
  // active in the last solve (island_joints, mechanical_world.rs:272-277)
  std::vector<Joint<T>> joints;
  int njcolours = 0;
  std::vector<int> active_joints;      // reference order: joint order
  std::vector<int> joint_order;        // colour order of the active joints
  std::vector<VelRow<T>> jbil, jbil_g;
  std::vector<JointRows> jrow;         // per active joint, in active_joints order
  std::vector<int> colours;
  int ncolours = 0;
  T res_inf = T(0), res_rms = T(0);
  long n_rows = 0;
  Counters counters;
  // MaterialsCoefficientsTable (materials_coefficients_table.rs:8-50), keyed by the sorted pair
  std::map<std::pair<uint32_t, uint32_t>, T> fric_table, rest_table;
  std::map<std::pair<int, int>, int> static_owner;  // (scene, static body index) -> its first collider
  bool any_surface_velocity = false;  // some body's material has a surface velocity

  World() {
    prm.t = T(0);
    prm.set_dt(T(1.0 / 60.0));
    prm.erp = T(0.2);
    prm.warmstart_coeff = T(1.0);
    prm.restitution_velocity_threshold = T(1.0);
    prm.allowed_linear_error = T(0.001);
    prm.allowed_angular_error = T(0.001);
    prm.max_linear_correction = T(0.2);
    prm.max_angular_correction = T(0.2);
    prm.max_stabilization_multiplier = T(0.2);
    prm.max_velocity_iterations = 8;
    prm.max_position_iterations = 3;
    prm.max_ccd_position_iterations = 10;
    prm.max_ccd_substeps = 1;
    gravity = V3<T>{T(0), T(-9.81), T(0)};
  }

  // ------------------------------------------------------------------ body insertion
  // RigidBodyDesc::build + ColliderDesc::build + GeometricalWorld::handle_insertions
  // (geometrical_world.rs:104-140): mass properties from the density, proxy creation with
  // loose AABB = aabb(pos) loosened by the query limit, then by the broad-phase margin 0.01.
  int add_body(int status, int shape, V3<T> dims, Iso<T> pos, V3<T> lv, V3<T> av, T density,
               T margin, T lin_pred, T friction, T restitution, int scene = 0, T sleep_thr = T(-1),
               const BodyExtras<T>& ex = BodyExtras<T>()) {
    Body<T> b;
    std::memset(static_cast<void*>(&b), 0, sizeof(b));
    b.scene = scene;
    // ActivationStatus::new_active (body.rs:77-82): 4 x the default threshold 0.01; the body's own
    // threshold (or None) is set afterwards by the descriptor (rigid_body.rs:1028)
    b.energy = T(0.01) * T(4.0);
    b.sleep_thr = sleep_thr;
    if (status == STATUS_DYNAMIC && sleep_thr >= T(0)) any_sleeper = true;
    b.status = status;
    b.shape = shape;
    b.sd.kind = shape;
    b.sd.he = shape == SHAPE_BALL ? V3<T>{dims.x, dims.x, dims.x} : dims;
    b.density = density;
    b.margin = margin;
    b.lin_pred = lin_pred;
    b.friction = friction;
    b.restitution = restitution;
    b.pos = pos;
    b.pos0 = pos;
    b.linvel = status == STATUS_DYNAMIC ? lv : V3<T>{T(0), T(0), T(0)};
    b.angvel = status == STATUS_DYNAMIC ? av : V3<T>{T(0), T(0), T(0)};
    b.gravity = ex.gravity;
    b.lin_damping = ex.lin_damping;
    b.ang_damping = ex.ang_damping;
    b.max_linvel = ex.max_linvel;
    b.max_angvel = ex.max_angvel;
    for (int k = 0; k < 3; ++k) {
      set_comp(b.mask_lin, k, (ex.kinematic_dofs >> k) & 1 ? T(0) : T(1));
      set_comp(b.mask_ang, k, (ex.kinematic_dofs >> (3 + k)) & 1 ? T(0) : T(1));
    }
    b.fric_mode = ex.fric_mode;
    b.rest_mode = ex.rest_mode;
    b.mat_id = ex.mat_id;
    b.surf_vel = ex.surf_vel;
    if (ex.surf_vel.x != T(0) || ex.surf_vel.y != T(0) || ex.surf_vel.z != T(0)) any_surface_velocity = true;
    b.flags = F_ALL;
    b.ccd = ex.ccd;
    b.lin_interp = status == STATUS_DYNAMIC && ex.lin_interp;
    if (ex.ccd) any_ccd = true;
    b.ccd_key = int(bodies.size());
    b.g_member = ex.g_member;
    b.g_white = ex.g_white;
    b.g_black = ex.g_black;
    if (status != STATUS_DYNAMIC && ex.static_body >= 0) {
      const auto key = std::make_pair(scene, ex.static_body);
      auto it = static_owner.find(key);
      if (it == static_owner.end()) static_owner.emplace(key, b.ccd_key);
      else b.ccd_key = it->second;
    }
    b.com = pos.t;
    b.mass = T(0);
    b.local_I = m3_zero<T>();
    if (status == STATUS_DYNAMIC && density != T(0)) {
      T mass;
      M3<T> ai;
      if (shape == SHAPE_BALL) {
        const T r = dims.x;
        const T vol = (((T(M_PI) * r) * r) * r) * T(4.0 / 3.0);
        mass = vol * density;
        const T diag = (r * r) * T(2.0 / 5.0);
        ai = m3_diag(diag * mass, diag * mass, diag * mass);
      } else {
        T vol = T(1);
        vol = vol * dims.x * T(2);
        vol = vol * dims.y * T(2);
        vol = vol * dims.z * T(2);
        mass = vol * density;
        const T w = (T(1.0 / 12.0) * T(2)) * T(2);
        const T ix = (w * dims.x) * dims.x, iy = (w * dims.y) * dims.y, iz = (w * dims.z) * dims.z;
        ai = m3_diag((iy + iz) * mass, (ix + iz) * mass, (ix + iy) * mass);
      }
      b.mass = mass;
      b.local_I = ai;
    }
    // RigidBody::new / update_inertia_from_local_inertia (rigid_body.rs:56-81,262-267): the
    // augmented mass starts as the world inertia, without gyroscopic terms
    b.inv_mass = b.mass == T(0) ? T(0) : T(1) / b.mass;
    if (status == STATUS_DYNAMIC) {
      const M3<T> R = qmat(pos.q);
      b.I_world = mul(mul(R, b.local_I), transpose(R));
      b.aug = b.I_world;
      try_inverse(b.aug, b.inv_aug);
    }
    const AABB<T> tight = loosened(shape_aabb(b.sd, pos), b.query_limit());
    b.loose = loosened(tight, T(0.01));
    bodies.push_back(b);
    pairs_dirty = true;
    return int(bodies.size()) - 1;
  }

  // Collider::set_collision_groups (collider.rs:386-388): the collider's pairs are re-filtered by
  // the collision detection at the start of the next step
  void set_collision_groups(int i, uint32_t member, uint32_t white, uint32_t black) {
    Body<T>& b = bodies[i];
    b.g_member = member;
    b.g_white = white;
    b.g_black = black;
    pairs_dirty = true;
    groups_changed = true;
  }

  void set_state(int i, Iso<T> pos, V3<T> lv, V3<T> av) {
    Body<T>& b = bodies[i];
    b.pos = pos;
    b.com = pos.t;
    if (b.status == STATUS_DYNAMIC) {
      b.linvel = lv;
      b.angvel = av;
    }
    b.flags |= F_POSITION | F_VELOCITY;
  }

  // ------------------------------------------------------------------ body dynamics
  void update_dynamics(Body<T>& b) {  // rigid_body.rs:558-588
    if (!(b.flags & F_ALL) || b.status != STATUS_DYNAMIC) return;
    if (b.energy == T(0) && b.sleep_thr >= T(0)) b.energy = b.sleep_thr * T(2.0);  // wake (:563-565)
    const M3<T> R = qmat(b.pos.q);
    b.I_world = mul(mul(R, b.local_I), transpose(R));
    b.aug = b.I_world;
    const V3<T> iw = mulv(b.I_world, b.angvel);
    const V3<T> w_dt = b.angvel * prm.dt;
    const M3<T> wx = cross_matrix(w_dt);
    const M3<T> iwx = cross_matrix(iw * prm.dt);
    b.aug = madd(b.aug, msub(mul(wx, b.I_world), iwx));
    b.inv_mass = b.mass == T(0) ? T(0) : T(1) / b.mass;
    try_inverse(b.aug, b.inv_aug);
  }

  void update_acceleration(Body<T>& b) {  // rigid_body.rs:590-619
    b.acc_lin = V3<T>{T(0), T(0), T(0)};
    b.acc_ang = V3<T>{T(0), T(0), T(0)};
    if (b.status != STATUS_DYNAMIC) return;
    const V3<T> iw = mulv(b.I_world, b.angvel);
    const V3<T> gyro = -cross(b.angvel, iw);
    b.acc_ang = mulv(b.inv_aug, gyro);
    if (b.inv_mass != T(0) && b.gravity) b.acc_lin = gravity;
    // acceleration += inv_augmented_mass * external_forces, then (.) jacobian_mask
    b.acc_lin = b.acc_lin + b.f_lin * b.inv_mass;
    b.acc_ang = b.acc_ang + mulv(b.inv_aug, b.f_ang);
    b.acc_lin = cmul(b.acc_lin, b.mask_lin);
    b.acc_ang = cmul(b.acc_ang, b.mask_ang);
  }
  static V3<T> cmul(V3<T> a, V3<T> m) { return V3<T>{a.x * m.x, a.y * m.y, a.z * m.z}; }

  // Body::apply_force (rigid_body.rs:771-822) for a world-frame force (linear, torque), with
  // the inverse augmented mass / augmented mass of the last update_dynamics. Types: 0 Force,
  // 1 Impulse, 2 AccelerationChange, 3 VelocityChange (ForceType, algebra/mod.rs:17-26).
  void apply_force(int i, V3<T> fl, V3<T> fa, int type, bool wake) {
    Body<T>& b = bodies[size_t(i)];
    if (b.status != STATUS_DYNAMIC) return;
    if (wake && b.sleep_thr >= T(0)) b.energy = b.sleep_thr * T(2.0);  // Body::activate (body.rs:338-342)
    if (type == 0) {
      b.f_lin = cmul(fl, b.mask_lin) + b.f_lin;
      b.f_ang = cmul(fa, b.mask_ang) + b.f_ang;
    } else if (type == 1) {
      b.linvel = cmul(fl * b.inv_mass, b.mask_lin) + b.linvel;
      b.angvel = cmul(mulv(b.inv_aug, fa), b.mask_ang) + b.angvel;
      b.flags |= F_VELOCITY;
    } else if (type == 2) {
      b.f_lin = cmul(fl * b.mass, b.mask_lin) + b.f_lin;
      b.f_ang = cmul(mulv(b.aug, fa), b.mask_ang) + b.f_ang;
    } else {
      b.linvel = cmul(fl, b.mask_lin) + b.linvel;
      b.angvel = cmul(fa, b.mask_ang) + b.angvel;
      b.flags |= F_VELOCITY;
    }
  }
  // apply_force_at_point (rigid_body.rs:835-845): Force::linear_at_point(f, point - com)
  void apply_force_at_point(int i, V3<T> f, V3<T> point, int type, bool wake) {
    const V3<T> p = point - bodies[size_t(i)].com;
    apply_force(i, f, cross(p, f), type, wake);
  }

  // Isometry composition of apply_displacement (rigid_body.rs:371-381):
  // new_pos = T(com) * Iso(lin, exp(ang)) * T(-com) * pos.
  // (the centre of mass is the pose's translation for every body here: com == pos.t)
  void apply_displacement(Body<T>& b, V3<T> lin, V3<T> ang) {
    b.pos = iso_displaced(b.pos, lin, ang);
    b.com = b.pos.t;
    b.flags |= F_POSITION;
  }

  void integrate(Body<T>& b) {  // rigid_body.rs:467-505
    b.linvel = b.linvel * (T(1) / (T(1) + prm.dt * b.lin_damping));
    b.angvel = b.angvel * (T(1) / (T(1) + prm.dt * b.ang_damping));
    const T ln = norm(b.linvel);
    if (ln > b.max_linvel) b.linvel = b.max_linvel == T(0) ? V3<T>{T(0), T(0), T(0)} : b.linvel * (b.max_linvel / ln);
    const T an = norm(b.angvel);
    if (an > b.max_angvel) b.angvel = b.max_angvel == T(0) ? V3<T>{T(0), T(0), T(0)} : b.angvel * (b.max_angvel / an);
    apply_displacement(b, b.linvel * prm.dt, b.angvel * prm.dt);
  }

  // ------------------------------------------------------------------ broad phase
  // Restated DBVT rule (ncollide DBVTBroadPhase, margin 0.01; geometrical_world.rs:60):
  // tight = shape AABB at the collider position merged with the one at the predicted position,
  // loosened by margin + linear_prediction; the proxy's loose AABB is replaced by tight + 0.01
  // only when it no longer contains tight. Pairs = overlapping loose AABBs that pass
  // DefaultCollisionFilter (geometrical_world.rs:933-971): not the same body, not both static.
  bool pairs_dirty = true;
  bool groups_changed = false;

  void sync_and_broad_phase(bool swept) {
    const double t0 = now_s();
    for (auto& b : bodies) {
      if (!(b.flags & F_POSITION)) continue;  // colliders_need_update (body.rs:510-512)
      AABB<T> a1 = shape_aabb(b.sd, swept ? safe_position(b) : b.pos);
      if (swept) a1 = merged(a1, shape_aabb(b.sd, b.pos));
      else a1 = merged(a1, a1);
      const AABB<T> tight = loosened(a1, b.query_limit());
      if (!contains(b.loose, tight)) {
        b.loose = loosened(tight, T(0.01));
        pairs_dirty = true;
      }
    }
    if (pairs_dirty) {
      recompute_pairs();
      pairs_dirty = false;
    }
    counters.broad_phase += now_s() - t0;
  }

  // DefaultCollisionFilter::is_pair_valid (geometrical_world.rs:933-971) for single-collider
  // bodies: at least one side has dofs; the collision groups ncollide's pipeline checks before it
  // (CollisionGroups::can_interact_with_groups, ncollide3d 0.31, unvendored: each collider is a
  // member of some group the other whitelists and of no group the other blacklists); plus the
  // scene rule (independent worlds in one batch).
  static bool groups_interact(const Body<T>& a, const Body<T>& b) {
    return (a.g_member & b.g_black) == 0u && (b.g_member & a.g_black) == 0u && (a.g_member & b.g_white) != 0u &&
           (b.g_member & a.g_white) != 0u;
  }
  static bool pair_allowed(const Body<T>& a, const Body<T>& b) {
    return (a.status == STATUS_DYNAMIC || b.status == STATUS_DYNAMIC) && a.scene == b.scene && groups_interact(a, b);
  }

  // canonical pair: first is dynamic; second is static or a higher index
  void push_pair(int i, int j) {
    const Body<T>& a = bodies[i];
    const Body<T>& b = bodies[j];
    if (!pair_allowed(a, b)) return;
    if (a.status == STATUS_DYNAMIC && (b.status != STATUS_DYNAMIC || j > i)) pairs.emplace_back(i, j);
    else if (b.status == STATUS_DYNAMIC && (a.status != STATUS_DYNAMIC || i > j)) pairs.emplace_back(j, i);
  }

  int brute_force_limit = 512;

  void recompute_pairs() {
    pairs.clear();
    const int n = int(bodies.size());
    if (n <= brute_force_limit) {
      for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
          if (intersects(bodies[i].loose, bodies[j].loose)) push_pair(i, j);
    } else {
      recompute_pairs_grid();
    }
    std::sort(pairs.begin(), pairs.end());
  }

  // Hashed uniform grid (keyed by scene and cell) for large scenes; returns exactly the
  // brute-force set.
  void recompute_pairs_grid() {
    const int n = int(bodies.size());
    std::vector<int> dyn;
    std::unordered_map<int, std::vector<int>> stat;
    V3<T> ext{T(0), T(0), T(0)};
    for (int i = 0; i < n; ++i) {
      if (bodies[i].status == STATUS_DYNAMIC) {
        dyn.push_back(i);
        ext = vmax(ext, bodies[i].loose.mx - bodies[i].loose.mn);
      } else {
        stat[bodies[i].scene].push_back(i);
      }
    }
    // statics against every dynamic body of their scene
    for (int d : dyn) {
      auto it = stat.find(bodies[d].scene);
      if (it == stat.end()) continue;
      for (int s : it->second)
        if (intersects(bodies[s].loose, bodies[d].loose)) push_pair(s, d);
    }
    const double cs = double(std::max(ext.x, std::max(ext.y, ext.z))) * 1.0001 + 1e-6;
    auto cell = [&](V3<T> p, int64_t& cx, int64_t& cy, int64_t& cz) {
      cx = int64_t(std::floor(double(p.x) / cs));
      cy = int64_t(std::floor(double(p.y) / cs));
      cz = int64_t(std::floor(double(p.z) / cs));
    };
    auto hkey = [](int scene, int64_t x, int64_t y, int64_t z) {
      return (uint64_t(x + (1 << 20)) << 42) ^ (uint64_t(y + (1 << 20)) << 21) ^ uint64_t(z + (1 << 20)) ^
             (uint64_t(uint32_t(scene)) * 0x9E3779B97F4A7C15ull);
    };
    std::unordered_map<uint64_t, std::vector<int>> grid;
    grid.reserve(dyn.size() * 2);
    for (int d : dyn) {
      int64_t cx, cy, cz;
      cell(bodies[d].loose.mn, cx, cy, cz);
      grid[hkey(bodies[d].scene, cx, cy, cz)].push_back(d);
    }
    for (int d : dyn) {
      int64_t cx, cy, cz;
      cell(bodies[d].loose.mn, cx, cy, cz);
      for (int64_t x = cx - 1; x <= cx + 1; ++x)
        for (int64_t y = cy - 1; y <= cy + 1; ++y)
          for (int64_t z = cz - 1; z <= cz + 1; ++z) {
            auto it = grid.find(hkey(bodies[d].scene, x, y, z));
            if (it == grid.end()) continue;
            for (int o : it->second)
              if (o > d && intersects(bodies[d].loose, bodies[o].loose)) push_pair(d, o);
          }
    }
  }

  // ------------------------------------------------------------------ narrow phase
  void narrow_phase() {
    const double t0 = now_s();
    manifolds.clear();
    contacts.clear();
    Contact<T> buf[8];
    for (const auto& p : pairs) {
      const Body<T>& a = bodies[p.first];
      const Body<T>& b = bodies[p.second];
      const T prediction = a.query_limit() + b.query_limit();
      const int nc = collide_pair(a.sd, a.pos, b.sd, b.pos, prediction, buf);
      if (nc == 0) continue;
      Manifold m{p.first, p.second, int(contacts.size()), nc};
      for (int k = 0; k < nc; ++k) contacts.push_back(buf[k]);
      manifolds.push_back(m);
    }
    counters.narrow_phase += now_s() - t0;
  }

  void collide() {  // sync_colliders + perform_broad_phase + perform_narrow_phase
    const double t0 = now_s();
    sync_and_broad_phase(false);
    narrow_phase();
    counters.collision_detection += now_s() - t0;
  }

  // ------------------------------------------------------------------ contact model
  // fill_constraint_geometry for one body: J = [d; (p - com) x d], W J = M^-1 J,
  // inv_r += m^-1 + (J_ang . I^-1 J_ang), out_vel += J . v + J . ext (rigid_body.rs:672-722).
  // With `angular` the force direction is ForceDirection::Angular (force (0, dir)); inv_r still
  // gains the inverse mass then (inv_mass.mass() at rigid_body.rs:707-708 is added for every
  // direction).
  void fill_geom(const Body<T>& b, V3<T> point, V3<T> dir, size_t j_id, size_t wj_id, T& inv_r,
                 const T* ext, T* out_vel, bool angular = false) {
    if (b.status != STATUS_DYNAMIC) return;
    const V3<T> zero{T(0), T(0), T(0)};
    const V3<T> p = point - b.com;
    const V3<T> ang_u = angular ? dir : cross(p, dir);
    if (angular) dir = zero;
    // the jacobian is masked; the velocity term uses the unmasked force so that kinematic dofs stay
    // controllable at the velocity level (rigid_body.rs:685-718)
    const V3<T> lin = cmul(dir, b.mask_lin);
    const V3<T> ang = cmul(ang_u, b.mask_ang);
    T* J = &jacobians[j_id];
    J[0] = lin.x; J[1] = lin.y; J[2] = lin.z; J[3] = ang.x; J[4] = ang.y; J[5] = ang.z;
    const V3<T> wl = lin * b.inv_mass;
    const V3<T> wa = mulv(b.inv_aug, ang);
    T* W = &jacobians[wj_id];
    W[0] = wl.x; W[1] = wl.y; W[2] = wl.z; W[3] = wa.x; W[4] = wa.y; W[5] = wa.z;
    inv_r = inv_r + (b.inv_mass + dot(ang, wa));
    if (out_vel) {
      const T v6[6] = {b.linvel.x, b.linvel.y, b.linvel.z, b.angvel.x, b.angvel.y, b.angvel.z};
      const T F[6] = {dir.x, dir.y, dir.z, ang_u.x, ang_u.y, ang_u.z};
      T s = T(0);
      for (int k = 0; k < 6; ++k) s = s + F[k] * v6[k];
      *out_vel = *out_vel + s;
      T e = T(0);
      for (int k = 0; k < 6; ++k) e = e + J[k] * ext[k];
      *out_vel = *out_vel + e;
    }
  }

  // helper::constraint_pair_geometry (helper.rs:53-135)
  struct Geom {
    size_t j1, j2, wj1, wj2;
    int nd1, nd2;
    T r;
  };
  Geom pair_geometry(const Body<T>& b1, const Body<T>& b2, V3<T> c1, V3<T> c2, V3<T> dir,
                     size_t& ground_j_id, size_t& j_id, T* out_vel, bool angular = false) {
    Geom g;
    g.nd1 = b1.status == STATUS_DYNAMIC ? 6 : 0;
    g.nd2 = b2.status == STATUS_DYNAMIC ? 6 : 0;
    size_t* out_j = (g.nd1 == 0 || g.nd2 == 0) ? &ground_j_id : &j_id;
    g.j1 = *out_j;
    g.j2 = g.j1 + g.nd1;
    g.wj1 = g.j2 + g.nd2;
    g.wj2 = g.wj1 + g.nd1;
    T inv_r = T(0);
    const T* e1 = g.nd1 ? &ext_vels[b1.companion] : nullptr;
    const T* e2 = g.nd2 ? &ext_vels[b2.companion] : nullptr;
    fill_geom(b1, c1, dir, g.j1, g.wj1, inv_r, e1, out_vel, angular);
    fill_geom(b2, c2, -dir, g.j2, g.wj2, inv_r, e2, out_vel, angular);
    g.r = inv_r != T(0) ? T(1) / inv_r : T(1);
    *out_j += size_t(g.nd1 + g.nd2) * 2;
    return g;
  }

  VelRow<T> make_row(const Geom& g, const Body<T>& b1, const Body<T>& b2, T rhs, T impulse,
                     uint64_t key) {
    VelRow<T> r;
    r.impulse = impulse;
    r.r = g.r;
    r.rhs = rhs;
    r.key = key;
    r.dep = -1;
    r.mu = T(0);
    r.ground = (g.nd1 == 0 || g.nd2 == 0);
    if (!r.ground) {
      r.a1 = b1.companion; r.a2 = b2.companion;
      r.j1 = g.j1; r.j2 = g.j2; r.wj1 = g.wj1; r.wj2 = g.wj2;
    } else if (g.nd1 == 0) {  // UnilateralGroundConstraint::new keeps the dynamic side
      r.a1 = b2.companion; r.a2 = -1;
      r.j1 = g.j2; r.wj1 = g.wj2; r.j2 = r.wj2 = 0;
    } else {
      r.a1 = b1.companion; r.a2 = -1;
      r.j1 = g.j1; r.wj1 = g.wj1; r.j2 = r.wj2 = 0;
    }
    return r;
  }

  // nalgebra orthonormal_subspace_basis for one 3D vector (unvendored; restated):
  // a = normalize(|n.x| > |n.y| ? (n.z, 0, -n.x) : (0, -n.z, n.y)); tangents a x n, then a.
  static void tangents(V3<T> n, V3<T>& t1, V3<T>& t2) {
    V3<T> a = std::fabs(n.x) > std::fabs(n.y) ? V3<T>{n.z, T(0), -n.x} : V3<T>{T(0), -n.z, n.y};
    const T l = norm(a);
    a = V3<T>{a.x / l, a.y / l, a.z / l};
    t1 = cross(a, n);
    t2 = a;
  }

  // ms[m].first indexes cs and, as row index, the assembled rows (contiguous over ms)
  void assemble(const std::vector<Manifold>& ms, const std::vector<Contact<T>>& cs) {
    // companion ids and buffers (moreau_jean_solver.rs:137-174, 322-326)
    int ndofs = 0;
    for (auto& b : bodies) {
      b.companion = 0;
      if (b.status == STATUS_DYNAMIC) {
        b.companion = ndofs;
        ndofs += 6;
      }
    }
    mj_lambda.assign(ndofs, T(0));
    ext_vels.assign(ndofs, T(0));
    for (auto& b : bodies) {
      if (b.status != STATUS_DYNAMIC) continue;
      const T acc[6] = {b.acc_lin.x, b.acc_lin.y, b.acc_lin.z, b.acc_ang.x, b.acc_ang.y, b.acc_ang.z};
      for (int k = 0; k < 6; ++k) ext_vels[b.companion + k] = prm.dt * acc[k];
    }
    // jacobian sizing: non-ground region first, ground region after (:181-224)
    size_t jac_sz = 0, gjac_sz = 0;
    for (int ji : active_joints) {
      const Joint<T>& jt = joints[size_t(ji)];
      const int nd1 = bodies[size_t(jt.b1)].status == STATUS_DYNAMIC ? 6 : 0;
      const int nd2 = bodies[size_t(jt.b2)].status == STATUS_DYNAMIC ? 6 : 0;
      const size_t sz = size_t(joint_num_rows(jt)) * 2 * size_t(nd1 + nd2);
      if (nd1 == 0 || nd2 == 0) gjac_sz += sz; else jac_sz += sz;
    }
    for (const auto& m : ms) {
      const int nd1 = bodies[m.b1].status == STATUS_DYNAMIC ? 6 : 0;
      const int nd2 = bodies[m.b2].status == STATUS_DYNAMIC ? 6 : 0;
      const size_t sz = size_t(3 * m.count) * size_t(nd1 + nd2) * 2;
      if (nd1 == 0 || nd2 == 0) gjac_sz += sz; else jac_sz += sz;
    }
    jacobians.assign(jac_sz + gjac_sz, T(0));
    size_t j_id = 0, ground_j_id = jac_sz;
    uni.clear(); uni_g.clear(); bil.clear(); bil_g.clear(); posc.clear(); crow.clear();
    jbil.clear(); jbil_g.clear(); jrow.clear();
    for (int ji : active_joints) joint_velocity_constraints(ji, ground_j_id, j_id);
    for (const auto& m : ms) {
      const Body<T>& b1 = bodies[m.b1];
      const Body<T>& b2 = bodies[m.b2];
      // Material::combine (material.rs:134-178): the lookup table when both materials have an id
      // and the table has the pair, else MaterialCombineMode::combine
      T mu, e;
      material_coefficients(b1, b2, mu, e);
      // surface_velocity = position1 * v1 - position2 * v2 (material.rs:175, basic_material.rs:47-56)
      V3<T> sv{T(0), T(0), T(0)};
      if (any_surface_velocity) sv = qrot(b1.pos.q, b1.surf_vel) - qrot(b2.pos.q, b2.surf_vel);
      for (int k = 0; k < m.count; ++k) {
        const Contact<T>& c = cs[size_t(m.first + k)];
        const uint64_t key = contact_key(m.b1, m.b2, c.feature);
        V3<T> imp{T(0), T(0), T(0)};
        auto it = cache.find(key);
        if (it != cache.end()) imp = it->second;
        // SignoriniModel::build_velocity_constraint (signorini_model.rs:37-138)
        const V3<T> center1 = c.world1 + c.normal * b1.margin;
        const V3<T> center2 = c.world2 - c.normal * b2.margin;
        T rhs = any_surface_velocity ? dot(c.normal, sv) : T(0);  // n . surface_velocity
        const Geom g = pair_geometry(b1, b2, center1, center2, -c.normal, ground_j_id, j_id, &rhs);
        if (rhs <= -prm.restitution_velocity_threshold) rhs = rhs + e * rhs;
        const T depth = (c.depth + b1.margin) + b2.margin;
        if (depth < T(0)) rhs = rhs + (-depth) * prm.inv_dt;
        VelRow<T> nr = make_row(g, b1, b2, rhs, imp.x * prm.warmstart_coeff, key);
        ContactRows cr;
        cr.ground = nr.ground;
        std::vector<VelRow<T>>& ulist = nr.ground ? uni_g : uni;
        ulist.push_back(nr);
        cr.normal = int(ulist.size()) - 1;
        // build_position_constraint (signorini_model.rs:153-197)
        PosRow<T> pr;
        pr.b1 = m.b1; pr.b2 = m.b2;
        pr.ndofs1 = b1.status == STATUS_DYNAMIC ? 6 : 0;
        pr.ndofs2 = b2.status == STATUS_DYNAMIC ? 6 : 0;
        pr.kin = c;
        pr.r = T(0); pr.rhs = T(0);
        posc.push_back(pr);
        // friction rows (signorini_coulomb_pyramid_model.rs:145-216)
        V3<T> t[2];
        tangents(c.normal, t[0], t[1]);
        for (int f = 0; f < 2; ++f) {
          T frhs = any_surface_velocity ? dot(t[f], sv) : T(0);  // t . surface_velocity
          const Geom fg = pair_geometry(b1, b2, center1, center2, t[f], ground_j_id, j_id, &frhs);
          VelRow<T> fr = make_row(fg, b1, b2, frhs, imp_comp(imp, f + 1) * prm.warmstart_coeff, key);
          fr.dep = cr.normal;
          fr.mu = mu;
          std::vector<VelRow<T>>& blist = fr.ground ? bil_g : bil;
          blist.push_back(fr);
          cr.fric[f] = int(blist.size()) - 1;
        }
        crow.push_back(cr);
      }
    }
    n_rows = long(uni.size() + uni_g.size() + bil.size() + bil_g.size() + jbil.size() + jbil_g.size());
  }

  // ------------------------------------------------------------------ joints
  static int joint_num_rows(const Joint<T>& j) {  // num_velocity_constraints
    return j.kind == JOINT_BALL ? 3 : (j.kind == JOINT_FIXED ? 6 : (j.kind == JOINT_CYLINDRICAL ? 4 : 5));
  }
  // World anchors and, for the fixed joint, the frames' world rotations
  // (position_at_material_point * ref_frame, rigid_body.rs:641-648, fixed_constraint.rs:173-177).
  void joint_frames(const Joint<T>& j, V3<T>& w1, V3<T>& w2, Q<T>& r1, Q<T>& r2) const {
    const Body<T>& b1 = bodies[size_t(j.b1)];
    const Body<T>& b2 = bodies[size_t(j.b2)];
    w1 = iso_apply(b1.pos, j.a1);
    w2 = iso_apply(b2.pos, j.a2);
    r1 = qmul(b1.pos.q, j.f1);
    r2 = qmul(b2.pos.q, j.f2);
  }
  // JointConstraint::velocity_constraints for one joint: bilateral rows with Independent
  // (unbounded) limits, warm-started with the joint's own cached impulses (no warm-start coefficient).
  void joint_velocity_constraints(int ji, size_t& ground_j_id, size_t& j_id) {
    const Joint<T>& jt = joints[size_t(ji)];
    const Body<T>& b1 = bodies[size_t(jt.b1)];
    const Body<T>& b2 = bodies[size_t(jt.b2)];
    V3<T> w1, w2;
    Q<T> r1, r2;
    joint_frames(jt, w1, w2, r1, r2);
    JointRows jr;
    jr.joint = ji;
    jr.n = 0;
    jr.ground = 0;
    auto push = [&](V3<T> dir, bool angular, T impulse) {
      T rhs = T(0);
      const Geom g = pair_geometry(b1, b2, w1, w2, dir, ground_j_id, j_id, &rhs, angular);
      VelRow<T> row = make_row(g, b1, b2, rhs, impulse, 0);
      jr.ground = row.ground;
      std::vector<VelRow<T>>& list = row.ground ? jbil_g : jbil;
      list.push_back(row);
      jr.row[jr.n++] = int(list.size()) - 1;
    };
    const V3<T> e[3] = {V3<T>{T(1), T(0), T(0)}, V3<T>{T(0), T(1), T(0)}, V3<T>{T(0), T(0), T(1)}};
    if (jt.kind == JOINT_CYLINDRICAL) {
      // CylindricalConstraint::velocity_constraints (cylindrical_constraint.rs:125-205):
      // restrict_relative_linear_velocity_to_axis, then restrict_relative_angular_velocity_to_axis
      // (helper.rs:614-699), both along orthonormal_subspace_basis(axis)
      const V3<T> axis = qrot(b1.pos.q, jt.ax1);
      V3<T> t[2];
      tangents(axis, t[0], t[1]);
      for (int i = 0; i < 2; ++i) push(t[i], false, comp(jt.lin_imp, i));
      for (int i = 0; i < 2; ++i) push(t[i], true, comp(jt.ang_imp, i));
      jrow.push_back(jr);
      return;
    }
    if (jt.kind == JOINT_PRISMATIC) {
      // PrismaticConstraint::velocity_constraints (prismatic_constraint.rs:132-206):
      // restrict_relative_linear_velocity_to_axis (helper.rs:788-856), the two directions of
      // orthonormal_subspace_basis(axis), axis = position_at_material_point(anchor1) * axis1; then
      // cancel_relative_angular_velocity (helper.rs:422-473)
      const V3<T> axis = qrot(b1.pos.q, jt.ax1);
      V3<T> t[2];
      tangents(axis, t[0], t[1]);
      for (int i = 0; i < 2; ++i) push(t[i], false, comp(jt.lin_imp, i));
      for (int i = 0; i < 3; ++i) push(e[i], true, comp(jt.ang_imp, i));
      jrow.push_back(jr);
      return;
    }
    for (int i = 0; i < 3; ++i) push(e[i], false, comp(jt.lin_imp, i));
    if (jt.kind == JOINT_FIXED) {
      for (int i = 0; i < 3; ++i) push(e[i], true, comp(jt.ang_imp, i));
    } else if (jt.kind == JOINT_REVOLUTE) {
      // axis1 = position_at_material_point(anchor1) * axis1 (revolute_constraint.rs:230)
      const V3<T> axis = qrot(b1.pos.q, jt.ax1);
      V3<T> t[2];
      tangents(axis, t[0], t[1]);  // orthonormal_subspace_basis(axis), in that order
      for (int i = 0; i < 2; ++i) push(t[i], true, comp(jt.ang_imp, i));
    }
    jrow.push_back(jr);
  }
  void solve_bilateral(VelRow<T>& c) {  // sor_prox.rs:232-343 with Independent limits
    const T w = row_w(c);
    const T lo = -std::numeric_limits<T>::max(), hi = std::numeric_limits<T>::max();
    const T v = c.impulse - c.r * w;
    const T ni = v > lo ? (v < hi ? v : hi) : lo;
    const T dl = ni - c.impulse;
    c.impulse = ni;
    row_add(c, dl);
  }
  VelRow<T>& jrow_at(const JointRows& jr, int k) { return jr.ground ? jbil_g[size_t(jr.row[k])] : jbil[size_t(jr.row[k])]; }
  // JointConstraint::cache_impulses (ball_constraint.rs:133-147, fixed_constraint.rs:194-218,
  // revolute_constraint.rs:265-289): impulses kept per joint; broken once |lin|^2 / dt^2 exceeds
  // the squared break force or |ang|^2 / dt^2 the squared break torque.
  void cache_joint_impulses() {
    const T inv_dt2 = prm.inv_dt * prm.inv_dt;
    for (const JointRows& jr : jrow) {
      Joint<T>& jt = joints[size_t(jr.joint)];
      const int nlin = jt.kind == JOINT_PRISMATIC || jt.kind == JOINT_CYLINDRICAL ? 2 : 3;  // linear rows first
      for (int k = 0; k < jr.n; ++k) {
        const T v = jrow_at(jr, k).impulse;
        if (k < nlin) set_comp(jt.lin_imp, k, v);
        else set_comp(jt.ang_imp, k - nlin, v);
      }
      if (norm2(jt.lin_imp) * inv_dt2 > jt.break_force_sq || norm2(jt.ang_imp) * inv_dt2 > jt.break_torque_sq)
        jt.broken = true;
    }
  }
  // One generic position constraint (GenericNonlinearConstraint, solved by
  // NonlinearSORProx::solve_generic, nonlinear_sor_prox.rs:85-119): rhs = -|error| when |error|
  // exceeds the allowed error, clamped by clamp_rhs, displacement W J (-rhs r) applied at once.
  void solve_joint_generic(Joint<T>& jt, V3<T> w1, V3<T> w2, V3<T> error, bool angular) {
    const T allowed = angular ? prm.allowed_angular_error : prm.allowed_linear_error;
    const T depth = norm(error);
    if (!(depth > allowed)) return;  // Unit::try_new_and_get(error, allowed)
    const V3<T> dir{error.x / depth, error.y / depth, error.z / depth};
    Body<T>& b1 = bodies[size_t(jt.b1)];
    Body<T>& b2 = bodies[size_t(jt.b2)];
    const T rhs = angular ? tmax((-depth + prm.allowed_angular_error) * prm.erp, -prm.max_angular_correction)
                          : tmax((-depth + prm.allowed_linear_error) * prm.erp, -prm.max_linear_correction);
    if (!(rhs < T(0))) return;
    T inv_r = T(0);
    V3<T> wl1{}, wa1{}, wl2{}, wa2{};
    const V3<T> zero{T(0), T(0), T(0)};
    if (b1.status == STATUS_DYNAMIC) {
      const V3<T> lin = angular ? zero : cmul(dir, b1.mask_lin);
      const V3<T> ang = cmul(angular ? dir : cross(w1 - b1.com, dir), b1.mask_ang);
      wl1 = lin * b1.inv_mass;
      wa1 = mulv(b1.inv_aug, ang);
      inv_r = inv_r + (b1.inv_mass + dot(ang, wa1));
    }
    if (b2.status == STATUS_DYNAMIC) {
      const V3<T> d2 = -dir;
      const V3<T> lin = angular ? zero : cmul(d2, b2.mask_lin);
      const V3<T> ang = cmul(angular ? d2 : cross(w2 - b2.com, d2), b2.mask_ang);
      wl2 = lin * b2.inv_mass;
      wa2 = mulv(b2.inv_aug, ang);
      inv_r = inv_r + (b2.inv_mass + dot(ang, wa2));
    }
    const T r = inv_r != T(0) ? T(1) / inv_r : T(1);
    const T impulse = -rhs * r;
    if (b1.status == STATUS_DYNAMIC) apply_displacement(b1, wl1 * impulse, wa1 * impulse);
    if (b2.status == STATUS_DYNAMIC) apply_displacement(b2, wl2 * impulse, wa2 * impulse);
  }
  // NonlinearConstraintGenerator::position_constraint of one joint, its constraints in order, each
  // evaluated at the poses the previous one left (ball: translation; fixed: rotation, then
  // translation; revolute: translation, then axis alignment).
  void joint_position(int ji) {
    Joint<T>& jt = joints[size_t(ji)];
    V3<T> w1, w2;
    Q<T> r1, r2;
    auto translation = [&]() {
      joint_frames(jt, w1, w2, r1, r2);
      solve_joint_generic(jt, w1, w2, w2 - w1, false);  // cancel_relative_translation (helper.rs:364-420)
    };
    if (jt.kind == JOINT_FIXED) {
      // cancel_relative_rotation (helper.rs:553-612): error = (rotation2 / rotation1).scaled_axis()
      joint_frames(jt, w1, w2, r1, r2);
      solve_joint_generic(jt, w1, w2, q_scaled_axis(qmul(r2, qconj(r1))), true);
      translation();
    } else if (jt.kind == JOINT_CYLINDRICAL) {
      // CylindricalConstraint::position_constraint (cylindrical_constraint.rs:256-292): i = 0
      // align_axis, i = 1 project_anchor_to_axis
      joint_frames(jt, w1, w2, r1, r2);
      const V3<T> a1 = qrot(bodies[size_t(jt.b1)].pos.q, jt.ax1);
      const V3<T> a2 = qrot(bodies[size_t(jt.b2)].pos.q, jt.ax2);
      solve_joint_generic(jt, w1, w2, axis_alignment_error(a1, a2), true);
      joint_frames(jt, w1, w2, r1, r2);
      const V3<T> axis = qrot(bodies[size_t(jt.b1)].pos.q, jt.ax1);
      const V3<T> proj = w1 + axis * dot(axis, w2 - w1);
      solve_joint_generic(jt, w1, w2, w2 - proj, false);
    } else if (jt.kind == JOINT_PRISMATIC) {
      // PrismaticConstraint::position_constraint (prismatic_constraint.rs:279-345): i = 0
      // cancel_relative_rotation of the bodies' rotations, i = 1 project_anchor_to_axis
      // (helper.rs:858-912): error = anchor2 - (anchor1 + axis (axis . (anchor2 - anchor1)))
      joint_frames(jt, w1, w2, r1, r2);
      const Q<T> q1 = bodies[size_t(jt.b1)].pos.q, q2 = bodies[size_t(jt.b2)].pos.q;
      solve_joint_generic(jt, w1, w2, q_scaled_axis(qmul(q2, qconj(q1))), true);
      joint_frames(jt, w1, w2, r1, r2);
      const V3<T> axis = qrot(bodies[size_t(jt.b1)].pos.q, jt.ax1);
      const V3<T> proj = w1 + axis * dot(axis, w2 - w1);
      solve_joint_generic(jt, w1, w2, w2 - proj, false);
    } else if (jt.kind == JOINT_REVOLUTE) {
      translation();
      // align_axis (helper.rs:701-769): the scaled axis of the rotation taking axis1 to axis2
      joint_frames(jt, w1, w2, r1, r2);
      const V3<T> a1 = qrot(bodies[size_t(jt.b1)].pos.q, jt.ax1);
      const V3<T> a2 = qrot(bodies[size_t(jt.b2)].pos.q, jt.ax2);
      solve_joint_generic(jt, w1, w2, axis_alignment_error(a1, a2), true);
    } else {
      translation();
    }
  }
  // Rotation::rotation_between_axis(a1, a2).scaled_axis() (nalgebra, unvendored; restated): the
  // axis a1 x a2 normalised times the angle atan2(|a1 x a2|, a1 . a2); for (anti)parallel axes
  // (|a1 x a2| at most epsilon) zero when a1 . a2 >= 0, otherwise the reference's pi fallback
  // (helper.rs:721-727): e_imin x a1 normalised times pi, imin the index of a1's smallest |component|.
  static V3<T> axis_alignment_error(V3<T> a1, V3<T> a2) {
    const V3<T> c = cross(a1, a2);
    const T sn = norm(c);
    const T cs = dot(a1, a2);
    if (sn > std::numeric_limits<T>::epsilon()) {
      const T ang = angle_of(sn, cs);
      return V3<T>{c.x / sn * ang, c.y / sn * ang, c.z / sn * ang};
    }
    if (cs >= T(0)) return V3<T>{T(0), T(0), T(0)};
    int imin = 0;
    if (std::fabs(a1.y) < std::fabs(comp(a1, imin))) imin = 1;
    if (std::fabs(a1.z) < std::fabs(comp(a1, imin))) imin = 2;
    V3<T> e{T(0), T(0), T(0)};
    set_comp(e, imin, T(1));
    const V3<T> d = cross(e, a1);
    const T l = norm(d);
    const T pi = T(3.14159265358979323846);
    return V3<T>{d.x / l * pi, d.y / l * pi, d.z / l * pi};
  }
  // Greedy colouring of the joints in joint order: a joint takes the smallest colour no earlier
  // joint sharing one of its dynamic bodies has. Joints between the same bodies conflict.
  void colour_joints() {
    std::vector<std::vector<int>> used(bodies.size());
    njcolours = 0;
    for (auto& jt : joints) {
      std::vector<char> taken;
      for (int b : {jt.b1, jt.b2}) {
        if (bodies[size_t(b)].status != STATUS_DYNAMIC) continue;
        for (int c : used[size_t(b)]) {
          if (int(taken.size()) <= c) taken.resize(size_t(c) + 1, 0);
          taken[size_t(c)] = 1;
        }
      }
      int c = 0;
      while (c < int(taken.size()) && taken[size_t(c)]) ++c;
      jt.colour = c;
      njcolours = std::max(njcolours, c + 1);
      for (int b : {jt.b1, jt.b2})
        if (bodies[size_t(b)].status == STATUS_DYNAMIC) used[size_t(b)].push_back(c);
    }
  }
  int add_joint(int kind, int b1, int b2, V3<T> a1, V3<T> a2, Q<T> f1, Q<T> f2, V3<T> ax1, V3<T> ax2,
                T break_force, T break_torque) {
    Joint<T> j;
    j.kind = kind;
    j.b1 = b1;
    j.b2 = b2;
    j.a1 = a1;
    j.a2 = a2;
    j.f1 = f1;
    j.f2 = f2;
    j.ax1 = ax1;
    j.ax2 = ax2;
    const T mx = std::numeric_limits<T>::max();
    j.break_force_sq = break_force >= T(0) ? break_force * break_force : mx;
    j.break_torque_sq = break_torque >= T(0) ? break_torque * break_torque : mx;
    joints.push_back(j);
    colour_joints();
    return int(joints.size()) - 1;
  }
  // island_joints: joints neither broken nor removed with a dynamic, awake body
  // (JointConstraint::is_active, joint_constraint.rs:218-227); `joint_order` lists them by colour.
  void collect_active_joints() {
    active_joints.clear();
    joint_order.clear();
    for (size_t i = 0; i < joints.size(); ++i) {
      const Joint<T>& jt = joints[i];
      if (jt.broken || jt.removed) continue;
      const Body<T>& b1 = bodies[size_t(jt.b1)];
      const Body<T>& b2 = bodies[size_t(jt.b2)];
      const bool a1 = b1.status == STATUS_DYNAMIC && b1.energy != T(0);
      const bool a2 = b2.status == STATUS_DYNAMIC && b2.energy != T(0);
      if (a1 || a2) active_joints.push_back(int(i));
    }
    for (int c = 0; c < njcolours; ++c)
      for (size_t k = 0; k < active_joints.size(); ++k)
        if (joints[size_t(active_joints[k])].colour == c) joint_order.push_back(int(k));
  }
  static T imp_comp(V3<T> v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

  void material_coefficients(const Body<T>& b1, const Body<T>& b2, T& mu, T& e) const {
    mu = combine_coeff(b1.friction, b1.fric_mode, b2.friction, b2.fric_mode);
    e = combine_coeff(b1.restitution, b1.rest_mode, b2.restitution, b2.rest_mode);
    if (b1.mat_id < 0 || b2.mat_id < 0) return;
    const uint32_t i1 = uint32_t(b1.mat_id), i2 = uint32_t(b2.mat_id);
    const std::pair<uint32_t, uint32_t> key{std::min(i1, i2), std::max(i1, i2)};
    const auto f = fric_table.find(key);
    if (f != fric_table.end()) mu = f->second;
    const auto r = rest_table.find(key);
    if (r != rest_table.end()) e = r->second;
  }

  // ------------------------------------------------------------------ SOR-Prox rows
  T dot6(size_t j, int a) const {
    T s = T(0);
    for (int k = 0; k < 6; ++k) s = s + jacobians[j + k] * mj_lambda[a + k];
    return s;
  }
  void axpy6(T alpha, size_t wj, int a) {
    for (int k = 0; k < 6; ++k) mj_lambda[a + k] = mj_lambda[a + k] + alpha * jacobians[wj + k];
  }
  T row_w(const VelRow<T>& c) const {
    if (!c.ground) return (dot6(c.j1, c.a1) + dot6(c.j2, c.a2)) + c.rhs;
    return dot6(c.j1, c.a1) + c.rhs;
  }
  void row_add(const VelRow<T>& c, T alpha) {
    axpy6(alpha, c.wj1, c.a1);
    if (!c.ground) axpy6(alpha, c.wj2, c.a2);
  }
  void warm_row(const VelRow<T>& c) {
    if (c.impulse != T(0)) row_add(c, c.impulse);
  }
  void solve_normal(VelRow<T>& c) {  // sor_prox.rs:181-230
    const T w = row_w(c);
    const T ni = tmax(c.impulse - c.r * w, T(0));
    const T dl = ni - c.impulse;
    c.impulse = ni;
    row_add(c, dl);
  }
  void solve_friction(VelRow<T>& c, const std::vector<VelRow<T>>& normals) {  // :232-343
    const T ln = normals[c.dep].impulse;
    if (ln == T(0)) {
      if (c.impulse != T(0)) {
        row_add(c, -c.impulse);
        c.impulse = T(0);
      }
      return;
    }
    const T mx = c.mu * ln;
    const T mn = -mx;
    const T w = row_w(c);
    const T v = c.impulse - c.r * w;
    const T ni = v > mn ? (v < mx ? v : mx) : mn;  // na::clamp
    const T dl = ni - c.impulse;
    c.impulse = ni;
    row_add(c, dl);
  }

  // SORProx::solve (sor_prox.rs:48-80), split into the warm start and one sweep.
  void warmstart() {  // sor_prox.rs:57-58, 19-45: the contacts' rows, then the joints'
    if (order_mode == ORDER_REFERENCE) {
      for (auto& c : uni) warm_row(c);
      for (auto& c : uni_g) warm_row(c);
      for (auto& c : bil) warm_row(c);
      for (auto& c : bil_g) warm_row(c);
      for (auto& c : jbil) warm_row(c);
      for (auto& c : jbil_g) warm_row(c);
      return;
    }
    warmstart_contacts_coloured();
    for (int k : joint_order) {
      const JointRows& jr = jrow[size_t(k)];
      for (int i = 0; i < jr.n; ++i) warm_row(jrow_at(jr, i));
    }
  }
  void warmstart_contacts_coloured() {
    for (int m : colour_order) {
      const Manifold& mf = solved[m];
      for (int k = 0; k < mf.count; ++k) {
        const ContactRows& cr = crow[mf.first + k];
        warm_row(cr.ground ? uni_g[cr.normal] : uni[cr.normal]);
      }
      for (int k = 0; k < mf.count; ++k) {
        const ContactRows& cr = crow[mf.first + k];
        for (int f = 0; f < 2; ++f) warm_row(cr.ground ? bil_g[cr.fric[f]] : bil[cr.fric[f]]);
      }
    }
  }
  void sweep() {  // SORProx::step (sor_prox.rs:159-179): joint rows first, then the contacts'
    if (order_mode == ORDER_REFERENCE) {
      for (auto& c : jbil) solve_bilateral(c);
      for (auto& c : jbil_g) solve_bilateral(c);
      for (auto& c : bil) solve_friction(c, uni);
      for (auto& c : bil_g) solve_friction(c, uni_g);
      for (auto& c : uni) solve_normal(c);
      for (auto& c : uni_g) solve_normal(c);
      return;
    }
    // colour order: the joints by joint colour, each its rows in order; then per contact its
    // normal row and its two friction rows bounded by the normal impulse just computed
    for (int k : joint_order) {
      const JointRows& jr = jrow[size_t(k)];
      for (int i = 0; i < jr.n; ++i) solve_bilateral(jrow_at(jr, i));
    }
    for (int m : colour_order) {
      const Manifold& mf = solved[m];
      for (int k = 0; k < mf.count; ++k) {
        const ContactRows& cr = crow[mf.first + k];
        solve_normal(cr.ground ? uni_g[cr.normal] : uni[cr.normal]);
        for (int f = 0; f < 2; ++f) {
          if (cr.ground) solve_friction(bil_g[cr.fric[f]], uni_g);
          else solve_friction(bil[cr.fric[f]], uni);
        }
      }
    }
  }
  // Fixed count (the reference), or, with res_target > 0, sweeps until the residual max |phi| is
  // at most res_target, checked after the warm start and every res_check_every sweeps.
  double res_target = 0.0;
  int res_max_sweeps = 500, res_check_every = 1, last_sweeps = 0, res_norm = 0;  // 0 max |phi|, 1 RMS
  long velocity_sweeps = 0;
  void sor_prox() {
    warmstart();
    if (res_target <= 0.0) {
      for (int it = 0; it < prm.max_velocity_iterations; ++it) sweep();
      last_sweeps = prm.max_velocity_iterations;
      compute_residual();
    } else {
      int done = 0;
      for (;;) {
        compute_residual();
        if (double(res_norm ? res_rms : res_inf) <= res_target || done >= res_max_sweeps) break;
        const int k = std::min(res_check_every, res_max_sweeps - done);
        for (int i = 0; i < k; ++i) sweep();
        done += k;
      }
      last_sweeps = done;
    }
    velocity_sweeps += last_sweeps;
  }

  // Natural-map residual of the LCP in velocity units, per row:
  //   normal:   (l - max(0, l - r w)) / r;  friction: (l - clamp(l - r w, -mu ln, mu ln)) / r
  void compute_residual() {
    double mx = 0, ss = 0;
    long n = 0;
    auto acc = [&](T phi) {
      const double a = std::fabs(double(phi));
      mx = std::max(mx, a);
      ss += a * a;
      ++n;
    };
    auto normal = [&](const VelRow<T>& c) {
      const T w = row_w(c);
      const T p = tmax(c.impulse - c.r * w, T(0));
      acc((c.impulse - p) / c.r);
    };
    auto fric = [&](const VelRow<T>& c, const std::vector<VelRow<T>>& ns) {
      const T ln = ns[c.dep].impulse;
      const T w = row_w(c);
      const T mxv = c.mu * ln, mnv = -mxv;
      const T v = c.impulse - c.r * w;
      const T p = v > mnv ? (v < mxv ? v : mxv) : mnv;
      acc((c.impulse - p) / c.r);
    };
    for (auto& c : uni) normal(c);
    for (auto& c : uni_g) normal(c);
    for (auto& c : bil) fric(c, uni);
    for (auto& c : bil_g) fric(c, uni_g);
    // joint rows (unbounded): l - (l - r w) over r, i.e. w
    for (auto& c : jbil) acc(row_w(c));
    for (auto& c : jbil_g) acc(row_w(c));
    res_inf = T(mx);
    res_rms = n ? T(std::sqrt(ss / double(n))) : T(0);
  }

  void cache_impulses() {  // signorini_coulomb_pyramid_model.rs:226-261
    // The reference's cache (a map keyed by the persistent contact id) keeps the entries of
    // contacts that still exist but were not solved (sleeping); those are carried over here.
    std::unordered_map<uint64_t, V3<T>> old;
    if (any_sleeper) old.swap(cache);
    cache.clear();
    if (any_sleeper)
      for (size_t m = 0; m < manifolds.size(); ++m) {
        if (manifold_active[m]) continue;
        const Manifold& mf = manifolds[m];
        for (int k = 0; k < mf.count; ++k) {
          const uint64_t key = contact_key(mf.b1, mf.b2, contacts[size_t(mf.first + k)].feature);
          const auto it = old.find(key);
          if (it != old.end()) cache[key] = it->second;
        }
      }
    for (const auto& c : uni) cache[c.key] = V3<T>{c.impulse, T(0), T(0)};
    for (const auto& c : uni_g) cache[c.key] = V3<T>{c.impulse, T(0), T(0)};
    for (size_t i = 0; i < crow.size(); ++i) {
      const ContactRows& cr = crow[i];
      const std::vector<VelRow<T>>& bl = cr.ground ? bil_g : bil;
      const uint64_t key = bl[cr.fric[0]].key;
      V3<T>& v = cache[key];
      v.y = bl[cr.fric[0]].impulse;
      v.z = bl[cr.fric[1]].impulse;
    }
  }

  // ------------------------------------------------------------------ position correction
  T clamp_rhs(T rhs) const {  // nonlinear_sor_prox.rs:296-309 (linear case)
    return tmax((rhs + prm.allowed_linear_error) * prm.erp, -prm.max_linear_correction);
  }

  void solve_position_row(PosRow<T>& pc) {  // nonlinear_sor_prox.rs:121-294
    Body<T>& b1 = bodies[pc.b1];
    Body<T>& b2 = bodies[pc.b2];
    Contact<T> g = reevaluate(pc.kin, b1.sd, b1.pos, b2.sd, b2.pos);
    // ContactKinematic dilations: margins on both sides (signorini_model.rs:174-181)
    const V3<T> w1 = g.world1 + g.normal * b1.margin;
    const V3<T> w2 = g.world2 - g.normal * b2.margin;
    const T depth = (g.depth + b1.margin) + b2.margin;
    pc.rhs = clamp_rhs(-depth);
    if (pc.rhs >= T(0)) return;
    T inv_r = T(0);
    V3<T> wl1{}, wa1{}, wl2{}, wa2{};
    if (pc.ndofs1) {  // fill_constraint_geometry with the masked jacobian (rigid_body.rs:672-705)
      const V3<T> dir = -g.normal;
      const V3<T> ang = cmul(cross(w1 - b1.com, dir), b1.mask_ang);
      wl1 = cmul(dir, b1.mask_lin) * b1.inv_mass;
      wa1 = mulv(b1.inv_aug, ang);
      inv_r = inv_r + (b1.inv_mass + dot(ang, wa1));
    }
    if (pc.ndofs2) {
      const V3<T> dir = g.normal;
      const V3<T> ang = cmul(cross(w2 - b2.com, dir), b2.mask_ang);
      wl2 = cmul(dir, b2.mask_lin) * b2.inv_mass;
      wa2 = mulv(b2.inv_aug, ang);
      inv_r = inv_r + (b2.inv_mass + dot(ang, wa2));
    }
    if (inv_r == T(0)) return;
    pc.r = T(1) / inv_r;
    const T impulse = -pc.rhs * pc.r;
    if (pc.ndofs1) apply_displacement(b1, wl1 * impulse, wa1 * impulse);
    if (pc.ndofs2) apply_displacement(b2, wl2 * impulse, wa2 * impulse);
  }

  void position_correction() {  // NonlinearSORProx::solve (nonlinear_sor_prox.rs:17-55)
    for (int it = 0; it < prm.max_position_iterations; ++it) {
      // the joints first (island_joints order, or by joint colour), then the contacts
      if (order_mode == ORDER_REFERENCE) {
        for (int ji : active_joints) joint_position(ji);
      } else {
        for (int k : joint_order) joint_position(active_joints[size_t(k)]);
      }
      if (order_mode == ORDER_REFERENCE) {
        for (auto& pc : posc) solve_position_row(pc);
      } else {
        for (int m : colour_order) {
          const Manifold& mf = solved[m];
          for (int k = 0; k < mf.count; ++k) solve_position_row(posc[mf.first + k]);
        }
      }
    }
  }

  // ------------------------------------------------------------------ colouring order
  std::vector<int> colour_order;
  std::vector<Manifold> prev_coloured;  // manifolds coloured by the previous step, pair order
  std::vector<int> prev_colours;

  // Colours persist across steps: a manifold whose body pair was also a manifold in the previous
  // step keeps that colour (the pair's conflicts are the same, so the partial colouring is valid),
  // and only new manifolds are coloured greedily around them.
  void build_colour_order() {
    std::vector<int> is_dyn(bodies.size());
    for (size_t i = 0; i < bodies.size(); ++i) is_dyn[i] = bodies[i].status == STATUS_DYNAMIC;
    colours.assign(solved.size(), -1);
    for (size_t m = 0; m < solved.size(); ++m) {
      const std::pair<int, int> key{solved[m].b1, solved[m].b2};
      size_t lo = 0, hi = prev_coloured.size();
      while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        const std::pair<int, int> k{prev_coloured[mid].b1, prev_coloured[mid].b2};
        if (k < key) lo = mid + 1;
        else hi = mid;
      }
      if (lo < prev_coloured.size() && prev_coloured[lo].b1 == key.first && prev_coloured[lo].b2 == key.second)
        colours[m] = prev_colours[lo];
    }
    ncolours = colour_manifolds(solved, is_dyn, int(bodies.size()), colours);
    std::vector<int> perm, size(size_t(ncolours), 0);
    for (int c : colours) size[size_t(c)]++;
    colour_visit_order(uint64_t(solve_count), ncolours, size, perm);
    colour_order.clear();
    for (int k = 0; k < ncolours; ++k)
      for (int m = 0; m < int(solved.size()); ++m)
        if (colours[m] == perm[k]) colour_order.push_back(m);
    prev_coloured = solved;
    prev_colours = colours;
  }
  long solve_count = 0;  // solves performed, never reset

  // ------------------------------------------------------------------ the step
  // ActivationManager::update (activation_manager.rs:61-201) without kinematic bodies or joints:
  // energy running average of the awake bodies, islands by union-find over the contact
  // manifolds between dynamic bodies, then every island sleeps (energy and velocity zeroed) when
  // all of its bodies have a threshold and an energy below it, and is woken otherwise.
  int uf_find(std::vector<int>& p, int x) {
    while (p[size_t(x)] != x) {
      p[size_t(x)] = p[size_t(p[size_t(x)])];
      x = p[size_t(x)];
    }
    return x;
  }
  void activation_update() {
    const T mix = T(0.01);
    for (auto& b : bodies) {
      if (b.status != STATUS_DYNAMIC || b.energy == T(0) || b.sleep_thr < T(0)) continue;
      // |generalized velocity|^2, components in order (linear, then angular)
      const T nsq = ((((b.linvel.x * b.linvel.x + b.linvel.y * b.linvel.y) + b.linvel.z * b.linvel.z) +
                      b.angvel.x * b.angvel.x) + b.angvel.y * b.angvel.y) + b.angvel.z * b.angvel.z;
      const T e = (T(1) - mix) * b.energy + mix * nsq;
      const T cap = b.sleep_thr * T(4.0);
      b.energy = e < cap ? e : cap;
    }
    std::vector<int> p(bodies.size());
    for (size_t i = 0; i < p.size(); ++i) p[i] = int(i);
    for (const auto& m : manifolds) {
      if (bodies[size_t(m.b1)].status != STATUS_DYNAMIC || bodies[size_t(m.b2)].status != STATUS_DYNAMIC) continue;
      const int a = uf_find(p, m.b1), c = uf_find(p, m.b2);
      if (a != c) p[size_t(std::max(a, c))] = std::min(a, c);
    }
    for (const auto& jt : joints) {  // joints that are not broken join their bodies' islands (:154-159)
      if (jt.broken || jt.removed) continue;
      if (bodies[size_t(jt.b1)].status != STATUS_DYNAMIC || bodies[size_t(jt.b2)].status != STATUS_DYNAMIC) continue;
      const int a = uf_find(p, jt.b1), c = uf_find(p, jt.b2);
      if (a != c) p[size_t(std::max(a, c))] = std::min(a, c);
    }
    std::vector<char> can(bodies.size(), 1);
    for (size_t i = 0; i < bodies.size(); ++i) {
      const Body<T>& b = bodies[i];
      if (b.status != STATUS_DYNAMIC) continue;
      if (!(b.sleep_thr >= T(0) && b.energy < b.sleep_thr)) can[size_t(uf_find(p, int(i)))] = 0;
    }
    for (size_t i = 0; i < bodies.size(); ++i) {
      Body<T>& b = bodies[i];
      if (b.status != STATUS_DYNAMIC) continue;
      if (can[size_t(uf_find(p, int(i)))]) {
        if (b.energy != T(0)) {  // RigidBody::deactivate (rigid_body.rs:396-400)
          b.energy = T(0);
          b.linvel = V3<T>{T(0), T(0), T(0)};
          b.angvel = V3<T>{T(0), T(0), T(0)};
        }
      } else if (b.energy == T(0) && b.sleep_thr >= T(0)) {
        b.energy = b.sleep_thr * T(2.0);  // Body::activate (body.rs:338-342)
      }
    }
  }

  void solver_step() {  // MoreauJeanSolver::step (moreau_jean_solver.rs:47-90)
    double t0 = now_s();
    // manifold collection (mechanical_world.rs:287-300): every kept manifold has >= 1 contact
    // and a dynamic side that is awake; narrow_phase() stores them with contiguous contacts.
    manifold_active.assign(manifolds.size(), 1);
    if (any_sleeper) activation_update();
    collect_active_joints();
    if (any_sleeper) {
      solved.clear();
      solved_contacts.clear();
      for (size_t m = 0; m < manifolds.size(); ++m) {
        const Body<T>& b1 = bodies[size_t(manifolds[m].b1)];
        const Body<T>& b2 = bodies[size_t(manifolds[m].b2)];
        const bool act = (b1.status == STATUS_DYNAMIC && b1.energy != T(0)) ||
                         (b2.status == STATUS_DYNAMIC && b2.energy != T(0));
        manifold_active[m] = act ? 1 : 0;
        if (!act) continue;
        Manifold s = manifolds[m];
        s.first = int(solved_contacts.size());
        for (int k = 0; k < s.count; ++k) solved_contacts.push_back(contacts[size_t(manifolds[m].first + k)]);
        solved.push_back(s);
      }
      assemble(solved, solved_contacts);
    } else {
      solved = manifolds;
      assemble(solved, contacts);
    }
    counters.assembly += now_s() - t0;
    t0 = now_s();
    ++solve_count;
    if (order_mode == ORDER_COLOURED) build_colour_order();
    sor_prox();
    cache_impulses();
    cache_joint_impulses();
    counters.velocity_resolution += now_s() - t0;
    t0 = now_s();
    for (auto& b : bodies) {  // update_velocities_and_integrate (:328-347), over the awake bodies
      if (b.status != STATUS_DYNAMIC || b.energy == T(0)) continue;
      const int id = b.companion;
      b.linvel = V3<T>{b.linvel.x + ext_vels[id + 0], b.linvel.y + ext_vels[id + 1], b.linvel.z + ext_vels[id + 2]};
      b.angvel = V3<T>{b.angvel.x + ext_vels[id + 3], b.angvel.y + ext_vels[id + 4], b.angvel.z + ext_vels[id + 5]};
      b.linvel = V3<T>{b.linvel.x + mj_lambda[id + 0], b.linvel.y + mj_lambda[id + 1], b.linvel.z + mj_lambda[id + 2]};
      b.angvel = V3<T>{b.angvel.x + mj_lambda[id + 3], b.angvel.y + mj_lambda[id + 4], b.angvel.z + mj_lambda[id + 5]};
      integrate(b);
    }
    counters.velocity_update += now_s() - t0;
    t0 = now_s();
    position_correction();
    counters.position_resolution += now_s() - t0;
  }

  // ------------------------------------------------------------------ CCD
  // MechanicalWorld::solve_ccd, predict_next_impacts and TOIEntry (mechanical_world.rs:400-1052)
  // with MoreauJeanSolver::step_ccd (moreau_jean_solver.rs:94-127), when some body has
  // ccd_enabled. Where the reference's order comes from hash maps or its interaction graph, this
  // restatement fixes one:
  //  * the first pass bounds every pair's time of impact by the step's end; the reference narrows
  //    the bound to the earliest impact seen beyond the window so far, which finds the same impacts
  //    within the window and the same earliest one beyond it;
  //  * the queue pops the earliest impact first and equal times by pair index (BinaryHeap leaves
  //    ties unordered);
  //  * a substep's island is every body reached from the impacting bodies' colliders through pairs
  //    whose manifold at the advanced poses is not empty, plus the dynamic neighbours of the
  //    traversed colliders (the reference skips a body that an earlier pair with an empty manifold
  //    already advanced, which depends on the graph's iteration order); its manifolds are solved
  //    in pair order.
  bool any_ccd = false;
  std::vector<T> body_time;  // SubstepState::body_times (0 for a body not advanced)
  std::vector<T> frozen;     // per ccd key: time the body was frozen at, < 0 when not frozen
  struct ToiEntry {
    T toi;
    int pair, ts;
  };
  std::vector<ToiEntry> impacts;
  std::string step_error;
  long ccd_substeps = 0, ccd_impacts = 0;

  // RigidBody::safe_position (rigid_body.rs:904-910)
  static Iso<T> safe_position(const Body<T>& b) { return b.lin_interp ? Iso<T>{b.pos0.t, b.pos.q} : b.pos0; }

  // Body::part_motion (rigid_body.rs:539-553; a static body: BodyPartMotion::Static at its pose)
  static CcdMotion<T> part_motion(const Body<T>& b, T t0) {
    CcdMotion<T> m;
    m.t0 = t0;
    m.v = b.linvel;
    m.w = b.angvel;
    if (b.status != STATUS_DYNAMIC) {
      m.kind = MOTION_STATIC;
      m.p0 = b.pos;
    } else if (b.lin_interp) {
      m.kind = MOTION_LINEAR;
      m.p0 = Iso<T>{b.pos0.t, b.pos.q};
    } else {
      m.kind = MOTION_ROTATING;
      m.p0 = b.pos0;
    }
    return m;
  }

  // RigidBody::clamp_advancement (rigid_body.rs:530-537)
  static void clamp_advancement(Body<T>& b) {
    b.pos = safe_position(b);
    b.com = b.pos.t;
    b.flags |= F_POSITION;
  }

  // TOIEntry::try_from_colliders (mechanical_world.rs:946-1022) for pair (i, j); f1, f2: the
  // bodies' freeze times (< 0: not frozen). The returned time is global (from the step's start).
  ToiResult<T> pair_toi(int i, int j, T f1, T f2, T end_time) const {
    const Body<T>& b1 = bodies[size_t(i)];
    const Body<T>& b2 = bodies[size_t(j)];
    const T bt1 = body_time[size_t(i)], bt2 = body_time[size_t(j)];
    const T time1 = f1 >= T(0) ? f1 : bt1;
    const T time2 = f2 >= T(0) ? f2 : bt2;
    const T start = tmax(time1, time2);
    const T o1 = time1 - start, o2 = time2 - start;
    CcdMotion<T> m1 = part_motion(b1, o1), m2 = part_motion(b2, o2);
    if (f1 >= T(0)) {
      m1.p0 = m1.at((f1 + o1) - bt1);
      m1.kind = MOTION_STATIC;
    }
    if (f2 >= T(0)) {
      m2.p0 = m2.at((f2 + o2) - bt2);
      m2.kind = MOTION_STATIC;
    }
    ToiResult<T> r = time_of_impact(b1.sd, m1, b2.sd, m2, end_time - start, prm.allowed_linear_error);
    if (r.found && prm.ccd_on_penetration == 0 && r.status == TOI_PENETRATING) r.found = false;
    if (r.found) r.toi = start + r.toi;
    return r;
  }

  int max_scene() const {
    int m = 0;
    for (const auto& b : bodies) m = std::max(m, b.scene);
    return m;
  }
  bool is_static(int i) const { return bodies[size_t(i)].status != STATUS_DYNAMIC; }
  int key_of(int i) const { return bodies[size_t(i)].ccd_key; }

  // predict_next_impacts (mechanical_world.rs:400-557): the impacts up to end_time in the order
  // they are confirmed, with the freeze times; returns PREDICT_*.
  enum { PREDICT_NONE = 0, PREDICT_AFTER_END = 1, PREDICT_IMPACTS = 2 };
  int predict_impacts(T end_time, T dt0, T& min_overstep) {
    impacts.clear();
    frozen.assign(bodies.size(), T(-1));
    std::vector<int> ccd_pairs;
    for (size_t p = 0; p < pairs.size(); ++p)
      if (bodies[size_t(pairs[p].first)].ccd || bodies[size_t(pairs[p].second)].ccd) ccd_pairs.push_back(int(p));
    // the CCD pairs of every dynamic body (the re-sweeps visit a frozen body's pairs only)
    std::vector<std::vector<int>> body_pairs(bodies.size());
    for (int p : ccd_pairs)
      for (const int s : {pairs[size_t(p)].first, pairs[size_t(p)].second})
        if (!is_static(s)) body_pairs[size_t(s)].push_back(p);
    std::vector<int> ts(pairs.size(), 0);
    auto later = [](const ToiEntry& a, const ToiEntry& b) {
      if (a.toi != b.toi) return a.toi > b.toi;
      if (a.pair != b.pair) return a.pair > b.pair;
      return a.ts > b.ts;
    };
    std::priority_queue<ToiEntry, std::vector<ToiEntry>, decltype(later)> q(later);
    min_overstep = dt0;
    for (int p : ccd_pairs) {
      const ToiResult<T> r = pair_toi(pairs[size_t(p)].first, pairs[size_t(p)].second, T(-1), T(-1), dt0);
      if (!r.found) continue;
      if (r.toi > end_time) {
        min_overstep = tmin(min_overstep, r.toi);
      } else {
        min_overstep = end_time;
        q.push(ToiEntry{r.toi, p, 0});
      }
    }
    if (min_overstep == dt0 && q.empty()) return PREDICT_NONE;
    if (min_overstep > end_time) return PREDICT_AFTER_END;
    while (!q.empty()) {
      const ToiEntry e = q.top();
      q.pop();
      if (e.ts != ts[size_t(e.pair)]) continue;  // superseded by a re-sweep
      const int i = pairs[size_t(e.pair)].first, j = pairs[size_t(e.pair)].second;
      const T t1 = frozen[size_t(key_of(i))], t2 = frozen[size_t(key_of(j))];
      if (t1 >= T(0) && t2 >= T(0)) {
        if (e.toi == tmax(t1, t2)) impacts.push_back(e);
        continue;
      }
      if (t1 < T(0)) frozen[size_t(key_of(i))] = e.toi;
      if (t2 < T(0)) frozen[size_t(key_of(j))] = e.toi;
      // re-sweep every pair of the two bodies' dynamic colliders with a CCD-enabled side
      for (const int s : {i, j}) {
        if (is_static(s)) continue;
        for (int p : body_pairs[size_t(s)]) {
          const int a = pairs[size_t(p)].first, b = pairs[size_t(p)].second;
          const T fa = frozen[size_t(key_of(a))], fb = frozen[size_t(key_of(b))];
          if ((fa >= T(0) || is_static(a)) && (fb >= T(0) || is_static(b))) continue;
          ts[size_t(p)] += 1;
          const ToiResult<T> r = pair_toi(a, b, fa, fb, end_time);
          if (r.found) q.push(ToiEntry{r.toi, p, ts[size_t(p)]});
        }
      }
      impacts.push_back(e);
    }
    return PREDICT_IMPACTS;
  }

  // prepare_body of solve_ccd (mechanical_world.rs:653-681) with the substep's parameters current
  void prepare_ccd_body(int s, T last_toi) {
    Body<T>& b = bodies[size_t(s)];
    b.f_lin = b.f_ang = V3<T>{T(0), T(0), T(0)};
    update_dynamics(b);
    update_acceleration(b);
    const T f = frozen[size_t(b.ccd_key)];
    const T target = f >= T(0) ? f : last_toi;
    b.pos0 = part_motion(b, T(0)).at(target - body_time[size_t(s)]);  // RigidBody::advance
    clamp_advancement(b);
    body_time[size_t(s)] = last_toi;
  }

  // One CCD substep (mechanical_world.rs:630-935): the island around the impacting bodies moved to
  // the last impact time, then MoreauJeanSolver::step_ccd over it with dt0 - last_toi, no warm
  // start, max_ccd_position_iterations, no joints and no gravity.
  void ccd_substep(T dt0) {
    ++ccd_substeps;
    ccd_impacts += long(impacts.size());
    const T last_toi = impacts.back().toi;
    const Params<T> saved_prm = prm;
    const V3<T> saved_g = gravity;
    prm.max_position_iterations = prm.max_ccd_position_iterations;
    prm.warmstart_coeff = T(0);
    prm.set_dt(dt0 - last_toi);
    gravity = V3<T>{T(0), T(0), T(0)};
    const size_t nb = bodies.size(), np = pairs.size();
    std::vector<char> in_t(nb, 0), prepared(nb, 0), evaluated(np, 0), kept(np, 0);
    for (const ToiEntry& e : impacts) {
      in_t[size_t(key_of(pairs[size_t(e.pair)].first))] = 1;
      in_t[size_t(key_of(pairs[size_t(e.pair)].second))] = 1;
    }
    Contact<T> buf[8];
    for (;;) {
      auto touches = [&](size_t p) { return in_t[size_t(key_of(pairs[p].first))] || in_t[size_t(key_of(pairs[p].second))]; };
      for (size_t p = 0; p < np; ++p) {
        if (evaluated[p] || !touches(p)) continue;
        for (const int s : {pairs[p].first, pairs[p].second})
          if (!is_static(s) && !prepared[size_t(s)]) {
            prepare_ccd_body(s, last_toi);
            prepared[size_t(s)] = 1;
          }
      }
      std::vector<int> grow;
      for (size_t p = 0; p < np; ++p) {
        if (evaluated[p] || !touches(p)) continue;
        evaluated[p] = 1;
        const Body<T>& a = bodies[size_t(pairs[p].first)];
        const Body<T>& b = bodies[size_t(pairs[p].second)];
        if (collide_pair(a.sd, a.pos, b.sd, b.pos, a.query_limit() + b.query_limit(), buf) == 0) continue;
        kept[p] = 1;
        for (const int s : {pairs[p].first, pairs[p].second})
          if (!is_static(s) && !in_t[size_t(key_of(s))]) grow.push_back(key_of(s));
      }
      if (grow.empty()) break;
      for (int k : grow) in_t[size_t(k)] = 1;
    }
    std::vector<Manifold> ms;
    std::vector<Contact<T>> cs;
    for (size_t p = 0; p < np; ++p) {
      if (!kept[p]) continue;
      const Body<T>& a = bodies[size_t(pairs[p].first)];
      const Body<T>& b = bodies[size_t(pairs[p].second)];
      const int nc = collide_pair(a.sd, a.pos, b.sd, b.pos, a.query_limit() + b.query_limit(), buf);
      ms.push_back(Manifold{pairs[p].first, pairs[p].second, int(cs.size()), nc});
      for (int k = 0; k < nc; ++k) cs.push_back(buf[k]);
    }
    // step_ccd on the island; the main solve's rows, colours and joint list are put back afterwards
    // (the next step's colours are inherited from the main solve)
    auto keep_uni = uni, keep_uni_g = uni_g, keep_bil = bil, keep_bil_g = bil_g;
    auto keep_jbil = jbil, keep_jbil_g = jbil_g;
    auto keep_posc = posc;
    auto keep_crow = crow;
    auto keep_jrow = jrow;
    auto keep_solved = solved;
    auto keep_solved_contacts = solved_contacts;
    auto keep_colours = colours;
    auto keep_order = colour_order;
    auto keep_prev = prev_coloured;
    auto keep_prev_colours = prev_colours;
    const int keep_ncolours = ncolours;
    const long keep_rows = n_rows;
    std::vector<int> keep_joints, keep_joint_order;
    keep_joints.swap(active_joints);
    keep_joint_order.swap(joint_order);
    solved = ms;
    solved_contacts = cs;
    assemble(ms, cs);
    ++solve_count;
    if (order_mode == ORDER_COLOURED) build_colour_order();
    position_correction();
    for (const ToiEntry& e : impacts)  // validate_advancement of the impacting bodies
      for (const int s : {pairs[size_t(e.pair)].first, pairs[size_t(e.pair)].second})
        if (!is_static(s)) bodies[size_t(s)].pos0 = bodies[size_t(s)].pos;
    warmstart();
    for (int it = 0; it < prm.max_velocity_iterations; ++it) sweep();
    for (size_t s = 0; s < nb; ++s) {  // update_velocities_and_integrate over the island
      if (!prepared[s]) continue;
      Body<T>& b = bodies[s];
      const int id = b.companion;
      b.linvel = V3<T>{b.linvel.x + ext_vels[id + 0], b.linvel.y + ext_vels[id + 1], b.linvel.z + ext_vels[id + 2]};
      b.angvel = V3<T>{b.angvel.x + ext_vels[id + 3], b.angvel.y + ext_vels[id + 4], b.angvel.z + ext_vels[id + 5]};
      b.linvel = V3<T>{b.linvel.x + mj_lambda[id + 0], b.linvel.y + mj_lambda[id + 1], b.linvel.z + mj_lambda[id + 2]};
      b.angvel = V3<T>{b.angvel.x + mj_lambda[id + 3], b.angvel.y + mj_lambda[id + 4], b.angvel.z + mj_lambda[id + 5]};
      integrate(b);
    }
    for (auto& b : bodies) update_dynamics(b);  // mechanical_world.rs:915-923
    keep_joints.swap(active_joints);
    keep_joint_order.swap(joint_order);
    uni = keep_uni; uni_g = keep_uni_g; bil = keep_bil; bil_g = keep_bil_g;
    jbil = keep_jbil; jbil_g = keep_jbil_g;
    posc = keep_posc; crow = keep_crow; jrow = keep_jrow;
    solved = keep_solved; solved_contacts = keep_solved_contacts;
    colours = keep_colours; colour_order = keep_order; ncolours = keep_ncolours; n_rows = keep_rows;
    prev_coloured = keep_prev; prev_colours = keep_prev_colours;
    prm = saved_prm;
    gravity = saved_g;
  }

  // MechanicalWorld::solve_ccd (mechanical_world.rs:559-943)
  void solve_ccd() {
    const T dt0 = prm.dt;
    if (dt0 == T(0) || prm.max_ccd_substeps == 0) return;
    const T sub_len = dt0 / T(prm.max_ccd_substeps);
    body_time.assign(bodies.size(), T(0));
    T end_time = T(0);
    bool last_iter = false;
    for (;;) {
      if (end_time < dt0) end_time = tmin(end_time + sub_len, dt0);
      else last_iter = true;
      sync_and_broad_phase(true);
      T min_ov = T(0);
      int kind = predict_impacts(end_time, dt0, min_ov);
      if (kind == PREDICT_AFTER_END) {
        end_time = std::floor(min_ov / sub_len) * sub_len;
        end_time = tmin(end_time + sub_len, dt0);
        kind = predict_impacts(end_time, dt0, min_ov);
        if (kind != PREDICT_IMPACTS) {
          step_error = "CCD: no impact in the window of the earliest one";
          return;
        }
      }
      if (kind == PREDICT_NONE) {
        impacts.clear();
        frozen.assign(bodies.size(), T(-1));
      }
      if (!last_iter && !impacts.empty()) {
        ccd_substep(dt0);
        continue;
      }
      for (auto& b : bodies)
        if (b.status == STATUS_DYNAMIC && frozen[size_t(b.ccd_key)] >= T(0)) clamp_advancement(b);
      break;
    }
  }

  void step() {  // MechanicalWorld::step_with_filter (mechanical_world.rs:199-396)
    const double ts = now_s();
    bool any_moved = false;
    for (auto& b : bodies) {
      b.pos0 = b.pos;  // step_started
      update_dynamics(b);
      if (b.flags & F_POSITION) any_moved = true;
    }
    for (auto& b : bodies) update_acceleration(b);
    // pre-solve CD: only when bodies were moved outside the solver or collision groups changed
    if (any_moved || groups_changed) collide();
    groups_changed = false;
    solver_step();
    for (auto& b : bodies) update_dynamics(b);  // post-solve update_dynamics (:343-346)
    // CCD phase: sync with (position0 -> position) and broad phase over swept AABBs (:604-606),
    // then with CCD-enabled bodies the impacts and substeps
    const double tccd = now_s();
    if (any_ccd) solve_ccd();
    else if (prm.dt != T(0) && prm.max_ccd_substeps != 0) sync_and_broad_phase(true);  // (:578-581)
    counters.ccd += now_s() - tccd;
    for (auto& b : bodies) b.pos0 = b.pos;  // validate_advancement
    const double tc = now_s();
    sync_and_broad_phase(false);
    narrow_phase();
    counters.collision_detection += now_s() - tc;
    prm.t = prm.t + prm.dt;
    for (auto& b : bodies) {
      b.flags = 0;
      b.f_lin = b.f_ang = V3<T>{T(0), T(0), T(0)};  // clear_forces (mechanical_world.rs:365)
    }
    counters.step += now_s() - ts;
    counters.nsteps++;
  }
};

}  // namespace orc
