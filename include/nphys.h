/* include/nphys.h -- C ABI of the MI355X-native nphysics contact pipeline (libnphys_hip.so).
 *
 * One handle (nphys_world) is one MechanicalWorld + GeometricalWorld + body/collider sets of
 * nphysics 0.23 restricted to rigid balls and cuboids. Every entry point below replaces one
 * reference call; the citation after each declaration names it (paths relative to the
 * nphysics source tree).
 *
 * Conventions
 *   - Return value: NPHYS_OK (0) or a negative NPHYS_ERR_* code; nphys_last_error() gives a
 *     message. Invariants the reference enforces with panics (negative dt, bad handles, a non
 *     dynamic body in the island, ...) are errors here, never silent skips.
 *   - Buffers are caller allocated host memory; the world owns all device memory.
 *   - A world handle is not reentrant and must be externally synchronised; distinct handles are
 *     independent. Each world owns one HIP stream on the current device.
 *   - Body handles are dense indices 0..n-1 in insertion order. Static bodies (ground, walls)
 *     have 0 status-dependent dofs; every body carries exactly one collider at the identity
 *     position_wrt_body.
 *   - Quaternions are (i, j, k, w), the nalgebra storage order.
 *   - Scalars on the device are FP32 (nphysics is generic over RealField; this build fixes f32).
 */
#ifndef NPHYS_H
#define NPHYS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NPHYS_OK 0
#define NPHYS_ERR_INVALID (-1)   /* bad argument or handle */
#define NPHYS_ERR_HIP (-2)       /* HIP runtime failure (message in nphys_last_error) */
#define NPHYS_ERR_CAPACITY (-3)  /* caller buffer too small */
#define NPHYS_ERR_STATE (-4)     /* call not valid in the current world state */
#define NPHYS_ERR_NO_DEVICE (-5) /* no usable gfx950 device */

enum { NPHYS_STATUS_DYNAMIC = 0, NPHYS_STATUS_STATIC = 1 };
enum { NPHYS_SHAPE_BALL = 0, NPHYS_SHAPE_CUBOID = 1 };
enum {
  NPHYS_CONTACT_MODEL_SIGNORINI_COULOMB_PYRAMID = 0, /* plain warm start from the impulse cache */
  NPHYS_CONTACT_MODEL_LEARNED_WARMSTART = 1          /* MLP-predicted initial impulses */
};

typedef struct nphys_world nphys_world;

/* IntegrationParameters (src/solver/integration_parameters.rs:5-76) field for field, plus the
 * world's gravity (src/world/mechanical_world.rs:67). inv_dt is derived from dt as set_dt does
 * (:142-157). Values are converted to the world's FP32 scalar on entry. */
typedef struct nphys_params {
  double t;
  double dt;
  double erp;
  double warmstart_coeff;
  double restitution_velocity_threshold;
  double allowed_linear_error;
  double allowed_angular_error;
  double max_linear_correction;
  double max_angular_correction;
  double max_stabilization_multiplier;
  int32_t max_velocity_iterations;
  int32_t max_position_iterations;
  int32_t max_ccd_position_iterations;
  int32_t max_ccd_substeps;
  int32_t return_after_ccd_substep;
  int32_t multiple_ccd_substep_sensor_events_enabled;
  int32_t ccd_on_penetration_enabled;
  int32_t reserved0;
  double gravity[3];
} nphys_params;

/* RigidBodyDesc + ColliderDesc for n bodies, structure of arrays
 * (src/object/rigid_body.rs:949-1086, src/object/collider.rs:455-572). */
typedef struct nphys_body_desc {
  int32_t n;
  const int32_t* status;            /* n:  NPHYS_STATUS_* */
  const int32_t* shape;             /* n:  NPHYS_SHAPE_* */
  const float* dims;                /* 3n: ball (r, -, -); cuboid half extents */
  const float* pos;                 /* 3n */
  const float* rot;                 /* 4n: unit quaternion (i, j, k, w) */
  const float* linvel;              /* 3n */
  const float* angvel;              /* 3n */
  const float* density;             /* n */
  const float* margin;              /* n: collider margin (default 0.01) */
  const float* linear_prediction;   /* n: default 0.001 */
  const float* friction;            /* n: BasicMaterial friction (default 0.5) */
  const float* restitution;         /* n: BasicMaterial restitution (default 0) */
  /* n: scene index in [0, NPHYS_MAX_SCENES), or NULL for scene 0. Bodies of different scenes
   * never interact: one handle then steps a batch of independent worlds (the reference needs one
   * MechanicalWorld per scene). */
  const int32_t* scene;
  /* n: sleep threshold of the energy running average (RigidBodyDesc::sleep_threshold,
   * rigid_body.rs:986,1028; activation_manager.rs:61-201), or NULL / a negative value for a body
   * that never sleeps (sleep_threshold(None)). With any threshold set the step first updates every
   * awake body's energy, groups the dynamic bodies into islands over the contact manifolds and puts
   * islands whose bodies all have energy below their threshold to sleep (zero velocity, excluded
   * from the solve and the integration) or wakes them. NULL differs from RigidBodyDesc's default
   * (Some(0.01)): pass 0.01 for the reference's default. nphys_set_state wakes every sleeping body
   * (2 x threshold) at the next step, as setting a pose or velocity does in the reference. */
  const float* sleep_threshold;
  /* Optional rigid-body and material attributes; NULL gives the RigidBodyDesc / BasicMaterial
   * default (rigid_body.rs:972-990, basic_material.rs:33-43). */
  const float* linear_damping;       /* n: default 0 (RigidBodyDesc::linear_damping) */
  const float* angular_damping;      /* n: default 0 */
  const float* max_linear_velocity;  /* n: default FLT_MAX (set_max_linear_velocity) */
  const float* max_angular_velocity; /* n: default FLT_MAX */
  const int32_t* gravity_enabled;    /* n: default 1 (RigidBodyDesc::gravity_enabled) */
  /* n: jacobian mask as kinematic dofs (set_translations_kinematic / set_rotations_kinematic,
   * rigid_body.rs:103-118): bit k set makes dof k kinematic, bits 0-2 the translations x, y, z,
   * bits 3-5 the rotations. Default 0 (every dof free). */
  const int32_t* kinematic_dofs;
  const int32_t* friction_combine;    /* n: NPHYS_COMBINE_* (BasicMaterial::friction_combine_mode) */
  const int32_t* restitution_combine; /* n: NPHYS_COMBINE_* */
  /* n: BasicMaterial::id for the coefficient lookup table (nphys_set_material_coefficient), a
   * negative value for None (the default) */
  const int32_t* material_id;
  /* 3n: BasicMaterial::surface_velocity in the collider's frame (default none = zero) */
  const float* surface_velocity;
  /* n: ColliderDesc::ccd_enabled (collider.rs:70,284-291), default 0. With any body CCD-enabled
   * every step runs MechanicalWorld::solve_ccd (mechanical_world.rs:559-943) after the position
   * correction: times of impact over the swept broad-phase pairs with a CCD-enabled side (bodies
   * frozen at their first impact, re-sweeps), then, when some impact falls in the step, a substep
   * (MoreauJeanSolver::step_ccd, moreau_jean_solver.rs:94-127) on the island around the impacting
   * bodies moved to the last impact time; bodies still predicted to impact are put back to their
   * pose at the start of their motion. One scene per world and no sleep thresholds with CCD. */
  const int32_t* ccd_enabled;
  /* n: RigidBody::enable_linear_motion_interpolation (rigid_body.rs:91-101), default 0: the CCD
   * motion of the body keeps its current rotation. */
  const int32_t* linear_motion_interpolation;
  /* n: for a static body, an index shared by the static colliders that belong to one static body
   * (the reference's Ground with several colliders, all at BodyPartHandle(ground, 0)); CCD freezes
   * and traverses that body as a whole. Negative or NULL: a body of its own. Ignored for dynamic
   * bodies. */
  const int32_t* static_body;
  /* 3n: ColliderDesc::collision_groups (collider.rs:444,506), ncollide3d 0.31's CollisionGroups as
   * (membership, whitelist, blacklist) bit masks over groups 0-29 (bits 30-31 are ignored): two
   * colliders interact only when each is a member of some group the other whitelists and of no
   * group the other blacklists. Pairs that cannot interact are never formed (no contacts, no
   * CCD). NULL: CollisionGroups::new() for every collider (member of all groups, whitelisting all,
   * blacklisting none). */
  const uint32_t* collision_groups;
} nphys_body_desc;

/* Joint constraints (src/joint/), structure of arrays for n joints. Kinds: BallConstraint
 * (ball_constraint.rs:14-178: the anchors coincide), FixedConstraint (fixed_constraint.rs:14-249:
 * anchors and reference frames coincide), RevoluteConstraint without limits or motor
 * (revolute_constraint.rs:14-349: anchors coincide, the bodies turn about the axis only) and
 * PrismaticConstraint without limits (prismatic_constraint.rs:14-345: the relative rotation is
 * cancelled and anchor 2 stays on the line through anchor 1 along axis1) and CylindricalConstraint
 * without limits or motor (cylindrical_constraint.rs:14-292: the axes stay aligned and anchor 2 on
 * the line along axis1; sliding along and turning about it are free). */
enum {
  NPHYS_JOINT_BALL = 0,
  NPHYS_JOINT_FIXED = 1,
  NPHYS_JOINT_REVOLUTE = 2,
  NPHYS_JOINT_PRISMATIC = 3,
  NPHYS_JOINT_CYLINDRICAL = 4
};
typedef struct nphys_joint_desc {
  int32_t n;
  const int32_t* kind;        /* n: NPHYS_JOINT_* */
  const int32_t* body1;       /* n: body handles, distinct, at least one of the two dynamic */
  const int32_t* body2;       /* n */
  const float* anchor1;       /* 3n: anchor in body 1's local frame */
  const float* anchor2;       /* 3n: anchor in body 2's local frame */
  const float* frame1;        /* 4n: fixed joints: ref_frame1 (i, j, k, w); NULL: identity */
  const float* frame2;        /* 4n: fixed joints: ref_frame2; NULL: identity */
  const float* axis1;         /* 3n: revolute, prismatic, cylindrical joints: unit axis in body 1's frame; NULL: (0, 0, 1) */
  const float* axis2;         /* 3n: revolute, cylindrical joints: unit axis in body 2's frame; NULL: (0, 0, 1) */
  const float* break_force;   /* n: set_break_force; NULL or a negative value: unbreakable */
  const float* break_torque;  /* n: set_break_torque (fixed, revolute); NULL or negative: unbreakable */
} nphys_joint_desc;

/* MaterialCombineMode (material.rs:46-66); the mode of higher precedence of the two colliders
 * decides: Max > Multiply > Min > Average. */
enum { NPHYS_COMBINE_AVERAGE = 0, NPHYS_COMBINE_MIN = 1, NPHYS_COMBINE_MULTIPLY = 2, NPHYS_COMBINE_MAX = 3 };
/* ForceType (src/algebra/mod.rs:17-26): a force, an impulse, an acceleration change
 * and a velocity change. */
enum { NPHYS_FORCE = 0, NPHYS_IMPULSE = 1, NPHYS_ACCELERATION_CHANGE = 2, NPHYS_VELOCITY_CHANGE = 3 };
#define NPHYS_MAX_SCENES 32768

/* Counters (src/counters/mod.rs:17-219): accumulated stage seconds plus sizes of the last step. */
typedef struct nphys_counters {
  double step;
  double collision_detection;
  double broad_phase;
  double narrow_phase;
  double assembly;
  double velocity_resolution;
  double velocity_update;
  double position_resolution;
  int64_t nsteps;
  int64_t npairs;
  int64_t ncontacts;
  int64_t nmanifolds;
  int64_t nconstraints;
  int64_t ncolours;
  int64_t colour_rounds;    /* accumulated greedy-colouring rounds (the last round that coloured) */
  /* Per-launch device time of the SOR-Prox sweep kernel over one colour class, summed over
   * launches, the launch count and the contacts those launches processed; measured with HIP events
   * carried by every per-colour sweep launch while counters are enabled (a scene-blocked variant
   * runs all its sweeps in one launch, counted as one). */
  double sweep_kernel_time;
  int64_t sweep_kernel_launches;
  int64_t sweep_kernel_contacts;
  int64_t velocity_sweeps;  /* accumulated SOR-Prox sweeps (fixed count, or to the residual target) */
  /* judge of the learned warm start (nphys_set_warmstart_judge): sweeps of the plain and of the
   * learned solve, and the steps whose learned solve was kept (it needed fewer sweeps) */
  int64_t judge_plain_sweeps;
  int64_t judge_learned_sweeps;
  int64_t judge_kept;
  int64_t sweep_variant;  /* solve of the last step: 0 per-colour launches, 256/512/1024 scene-blocked */
  int64_t loose_changes;  /* accumulated loose AABBs replaced by the end-of-step syncs */
  int64_t pair_rebuilds;  /* accumulated end-of-step broad-phase pair rebuilds */
  /* accumulated pair rebuilds done incrementally (only the bodies whose loose AABB was replaced
   * re-traversed, the tree refitted by expansion) instead of by a full tree build */
  int64_t incremental_pair_updates;
  int64_t active_bodies;  /* awake dynamic bodies after the last step's activation update */
  /* accumulated incremental pair updates that overflowed a clean body's candidate segment and
   * fell back to a full build with larger segments */
  int64_t incremental_overflows;
  /* learned warm start, while counters are enabled: device time of the network's inference
   * kernel (HIP events around its launch on the world's stream) and the rows it evaluated */
  double warmstart_inference_time;
  int64_t warmstart_inference_rows;
  /* CCD (counters.ccd, src/counters/ccd_counters.rs): time of the CCD phase of every step, the
   * substeps run and the impacts they resolved, accumulated */
  double ccd_time;
  int64_t ccd_substeps;
  int64_t ccd_impacts;
} nphys_counters;

/* ---------------------------------------------------------------- lifecycle */

/* IntegrationParameters::default() (integration_parameters.rs:169-189) and gravity (0,-9.81,0). */
void nphys_params_default(nphys_params* out);

/* MechanicalWorld::new + GeometricalWorld::new (mechanical_world.rs:76-102,
 * geometrical_world.rs:57-63). Selects the current HIP device; fails with NPHYS_ERR_NO_DEVICE
 * when there is none. */
int nphys_world_create(const nphys_params* params, nphys_world** out);
int nphys_world_destroy(nphys_world* w);

/* integration_parameters / gravity pub fields (mechanical_world.rs:55-68); set_timestep
 * (:108-110) is nphys_set_params with a new dt. A negative dt is NPHYS_ERR_INVALID
 * (the reference asserts, integration_parameters.rs:143-146). */
int nphys_set_params(nphys_world* w, const nphys_params* params);
int nphys_get_params(const nphys_world* w, nphys_params* out);

/* ---------------------------------------------------------------- scene */

/* BodySet::insert + ColliderSet::insert, then the insertion half of MechanicalWorld::maintain
 * (mechanical_world.rs:114-179 -> geometrical_world.rs:104-140): mass properties from the
 * density (volumetric.rs:46-59), proxy creation. *first_handle receives the first new index. */
int nphys_add_bodies(nphys_world* w, const nphys_body_desc* desc, int32_t* first_handle);
int nphys_num_bodies(const nphys_world* w, int32_t* n);
/* DefaultJointConstraintSet::insert (src/joint/joint_constraint.rs:86-107) of n joints
 * (nphys_joint_desc); joint handles are dense indices in insertion order, the first written to
 * *first_handle. Joints are solved with the contacts from the next step on: their rows before the
 * contacts' in every sweep, their position constraints before the contacts' in every position
 * iteration, and joints that are not broken join their bodies' activation islands. */
int nphys_add_joints(nphys_world* w, const nphys_joint_desc* desc, int32_t* first_handle);
/* JointConstraintSet::remove: the joint takes no further part (its handle stays valid). */
int nphys_remove_joint(nphys_world* w, int32_t joint);
int nphys_num_joints(const nphys_world* w, int32_t* n);
/* Impulses cached by each joint's last solve (JointConstraint::cache_impulses): linear [n][3]
 * along world x, y, z; angular [n][3] (fixed: about x, y, z; revolute: the two directions
 * orthogonal to the axis, then 0); broken [n] (set once a break force or torque was exceeded).
 * Any of the pointers may be NULL. */
int nphys_read_joints(nphys_world* w, float* lin_impulses, float* ang_impulses, int32_t* broken, int64_t cap);

/* RigidBody::set_position / set_velocity for every body (rigid_body.rs:305-324); marks the
 * bodies moved so the next step or nphys_collide re-runs collision detection. Static bodies
 * ignore velocities. NULL arrays are left unchanged. */
int nphys_set_state(nphys_world* w, const float* pos, const float* rot, const float* linvel,
                    const float* angvel);
int nphys_get_state(nphys_world* w, float* pos, float* rot, float* linvel, float* angvel);

/* Collider::set_collision_groups (collider.rs:386-388) for `count` bodies: groups holds 3 masks
 * (membership, whitelist, blacklist) per entry, as nphys_body_desc.collision_groups. The pair set
 * is rebuilt by the collision detection at the start of the next step (or by nphys_collide), as
 * ncollide re-filters the collider's pairs at its next broad phase. */
int nphys_set_collision_groups(nphys_world* w, int32_t count, const int32_t* bodies, const uint32_t* groups);

/* Body::apply_force (rigid_body.rs:771-822) for `count` entries: body handle, world-frame force
 * (fx, fy, fz, torque x, y, z) and the ForceType `type` (NPHYS_FORCE ...). Forces and acceleration
 * changes accumulate into the body's external forces, applied by the next step and then cleared
 * (mechanical_world.rs:365); impulses and velocity changes change the velocity now, impulses and
 * acceleration changes through the (inverse) augmented mass of the body's last dynamics update.
 * Each value is masked by the body's jacobian mask; non-dynamic bodies are ignored. With
 * auto_wake_up the body is activated (energy 2 x its sleep threshold, body.rs:338-342). Entries of
 * one body are applied in the given order. */
int nphys_apply_force(nphys_world* w, int32_t count, const int32_t* bodies, const float* forces,
                      int32_t type, int32_t auto_wake_up);
/* Body::apply_force_at_point (rigid_body.rs:835-845): linear force (3 per entry) at a world point
 * (3 per entry), i.e. the force (f, (point - com) x f). */
int nphys_apply_force_at_point(nphys_world* w, int32_t count, const int32_t* bodies, const float* forces,
                               const float* points, int32_t type, int32_t auto_wake_up);

/* MaterialsCoefficientsTable::set_friction_coefficient / set_restitution_coefficient
 * (materials_coefficients_table.rs:23-31): the coefficient used for a contact between colliders
 * whose materials have ids id1 and id2 (either order), instead of the combine modes. which: 0
 * friction, 1 restitution. At most NPHYS_MAX_MATERIAL_PAIRS pairs per kind. */
#define NPHYS_MAX_MATERIAL_PAIRS 65536
int nphys_set_material_coefficient(nphys_world* w, int32_t which, int32_t id1, int32_t id2, float coeff);

/* ---------------------------------------------------------------- stepping */

/* MechanicalWorld::step (mechanical_world.rs:182-196), nsteps times, on the world's stream;
 * returns after the last step is complete on the device. */
int nphys_step(nphys_world* w, int32_t nsteps);

/* GeometricalWorld::sync_colliders + perform_broad_phase + perform_narrow_phase
 * (geometrical_world.rs:230-320) on the current poses, as MechanicalWorld::step does before the
 * solver when bodies were moved by the user (mechanical_world.rs:250-253). */
int nphys_collide(nphys_world* w);

/* MoreauJeanSolver::set_contact_model (moreau_jean_solver.rs:42-44) restricted to the two models
 * this library provides. The learned model needs a network loaded with nphys_warmstart_load
 * first (NPHYS_ERR_STATE otherwise). */
int nphys_contact_model_set(nphys_world* w, int32_t model);

/* ---------------------------------------------------------------- readback */

/* Broad-phase pairs, sorted: (first, second) with first dynamic and second static or a larger
 * index (the interaction graph of geometrical_world.rs:301-311). pairs holds 2*cap int32. */
int nphys_num_pairs(nphys_world* w, int64_t* n);
int nphys_read_pairs(nphys_world* w, int32_t* pairs, int64_t cap);

/* Contact manifolds of the current interaction graph (GeometricalWorld::contact_pairs,
 * geometrical_world.rs:534-553): per contact its two bodies, feature id, world points (no
 * margin), normal (body 1 -> body 2) and depth (> 0 penetrating). Any pointer may be NULL. */
int nphys_num_contacts(nphys_world* w, int64_t* n);
int nphys_read_contacts(nphys_world* w, int32_t* bodies, uint32_t* feature, float* world1,
                        float* world2, float* normal, float* depth, int64_t cap);

/* The same contacts grouped as contact manifolds, one per body pair in contact (the manifolds
 * GeometricalWorld::contact_pairs yields, geometrical_world.rs:534-553, with
 * ColliderContactManifold's pair, collider_contact_manifold.rs:9-115): per manifold its two bodies
 * in the broad phase's pair order, the index of its first contact in nphys_read_contacts' order
 * and its contact count (1..4); manifolds in increasing (body 1, body 2) order. Any pointer may be
 * NULL. */
int nphys_num_manifolds(nphys_world* w, int64_t* n);
int nphys_read_manifolds(nphys_world* w, int32_t* bodies, int32_t* first, int32_t* count, int64_t cap);

/* Impulses (normal, tangent 1, tangent 2) of the contacts solved by the last step, in the order
 * and with the keys of that step's contacts (the ImpulseCache of
 * signorini_coulomb_pyramid_model.rs:226-261). keys may be NULL. */
int nphys_num_solved_contacts(nphys_world* w, int64_t* n);
int nphys_read_impulses(nphys_world* w, uint64_t* keys, float* impulses, int64_t cap);

/* Natural-map LCP residual after the last velocity solve: max and RMS over all rows of
 * |l - proj(l - r w)| / r, w = J dv + rhs (m/s). The reference computes no residual; this is the
 * measure both this library and its CPU checker report. */
int nphys_residual(nphys_world* w, double* r_inf, double* r_rms);

/* Counters::{step,...}_started/completed totals (counters/mod.rs:148-209). */
int nphys_read_counters(nphys_world* w, nphys_counters* out);
int nphys_reset_counters(nphys_world* w);

/* Counters::enable / disable (counters/mod.rs:51-63). Stage totals are always kept; enabling adds
 * the per-launch timing of the sweep kernel (one HIP event pair per launch). Off by default, as
 * MechanicalWorld::new creates Counters::new(false) (mechanical_world.rs:77). */
int nphys_counters_enable(nphys_world* w, int32_t enabled);

/* Solve-to-residual mode of the velocity solver. With target > 0 every step runs SOR-Prox sweeps
 * until the natural-map residual (see nphys_residual) is <= target, checked after the warm start
 * and after every check_every sweeps, at most max_sweeps sweeps; the residual is max |phi| with
 * norm == NPHYS_RESIDUAL_MAX, the root mean square with NPHYS_RESIDUAL_RMS. With target == 0 (the
 * default) it runs max_velocity_iterations sweeps as SORProx::solve does (sor_prox.rs:56-79).
 * No reference counterpart: the reference only has the fixed count. */
#define NPHYS_RESIDUAL_MAX 0
#define NPHYS_RESIDUAL_RMS 1
int nphys_set_residual_target(nphys_world* w, double target, int32_t max_sweeps, int32_t check_every,
                              int32_t norm);

/* Launch shape of the velocity sweeps and the position correction. 0: one launch per colour and
 * pass over all manifolds of the colour; 256, 512, 1024: per solve one launch for the warm start
 * and sweeps and one for the position passes, one workgroup per scene (the scene's velocity
 * increments, then its poses, in LDS; that many lanes for a scene of more than 256 dynamic bodies,
 * 64 for the others, which run beside them on a second stream). Scene-blocked only when every
 * scene has at most 2048 dynamic bodies and the world has no joints; otherwise 0 is used.
 * -1 (default): an autotuner times every eligible variant over the first steps and keeps the
 * fastest. All variants give bit-identical results. No reference counterpart (the reference has
 * one sequential loop). */
int nphys_set_sweep_variant(nphys_world* w, int32_t variant);

/* How the pair set follows replaced loose AABBs. 1 (default): when only dynamic bodies' loose
 * AABBs were replaced since the last pair build (at most n/8 of them, and a full build at least
 * every 32 updates), only those bodies are re-traversed in the previous tree refitted by
 * expansion; 0: the tree is rebuilt and every body traversed. The pair set is identical either
 * way. Replaces ncollide's incremental DBVT update (geometrical_world.rs:285-312; the DBVT itself
 * is not in the reference tree). */
int nphys_set_incremental_pairs(nphys_world* w, int32_t on);

/* Activation energy of every body (ActivationStatus::energy, body.rs:65-125): 0 for a sleeping
 * body; 0.04 (4 x the default threshold) for a new body. No counterpart entry point in the
 * reference, which exposes it per body (Body::activation_status); cap >= the body count. */
int nphys_read_activation(nphys_world* w, float* energy, int64_t cap);

/* Colour of every manifold solved by the last step, in manifold order. Recording starts with the
 * first call (a call with colours == NULL only enables it), so the first step after enabling is
 * the first one reported. */
int nphys_read_colours(nphys_world* w, int32_t* colours, int64_t cap, int32_t* ncolours);

/* ---------------------------------------------------------------- learned warm start
 *
 * No reference counterpart: the reference warm-starts every contact from the impulse cache
 * (signorini_coulomb_pyramid_model.rs:104-108). A network nphys_mlp (24 -> 128 -> 128 -> 3, GELU,
 * BF16 weights and activations, FP32 accumulation) maps a contact's 24 features to the correction
 * of its cached impulses (normal, tangent 1, tangent 2; the cache is 0 for a new contact) in units
 * of r_n |g| dt, r_n the effective mass of the contact's normal row. Parameters are FP32,
 * NPHYS_MLP_NUM_PARAMS of them, laid out as
 *   W1[24][128] b1[128] W2[128][128] b2[128] W3[128][3] b3[3] x_mean[24] x_scale[24] y_scale[3]
 * with inputs normalised as (x - x_mean) / x_scale and outputs multiplied by y_scale. */
#define NPHYS_MLP_NUM_FEATURES 24
#define NPHYS_MLP_NUM_PARAMS 20150
typedef struct nphys_mlp nphys_mlp;

/* He-initialised network on the current device; seed drives the initialisation and the
 * per-epoch shuffles. */
int nphys_mlp_create(uint64_t seed, nphys_mlp** out);
int nphys_mlp_destroy(nphys_mlp* m);
int nphys_mlp_num_params(int64_t* n);
int nphys_mlp_get_params(nphys_mlp* m, float* out, int64_t cap);
int nphys_mlp_set_params(nphys_mlp* m, const float* params, int64_t n);
/* Training set: n rows of 24 features and 3 target impulses (copied to the device). With
 * set_normalisation the feature means/scales and output scales are computed from this data. */
int nphys_mlp_set_data(nphys_mlp* m, const float* features, const float* targets, int64_t n,
                       int32_t set_normalisation);
/* Adam (beta 0.9 / 0.999) on the mean squared error in normalised output units, mini-batches of
 * `batch` rows (a multiple of 64) in a fresh shuffle per epoch; epoch_loss (may be NULL) receives
 * each epoch's mean batch loss. */
int nphys_mlp_train(nphys_mlp* m, int32_t epochs, int32_t batch, float lr, float* epoch_loss);
/* Inference of n rows: out[n][3] impulses. */
int nphys_mlp_forward(nphys_mlp* m, const float* features, int64_t n, float* out);
/* Gradient of the batch loss (mean over the n rows of the squared error summed over the 3
 * normalised outputs) with respect to the 20,099 trained parameters, in parameter order. */
int nphys_mlp_gradients(nphys_mlp* m, const float* features, const float* targets, int64_t n,
                        float* grad, double* loss);
/* nphys_mlp_gradients without any allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight training loops). No counterpart in the reference. All
 * pointers are DEVICE pointers on the current device (features [n][24], targets [n][3], grad
 * [20099], loss [1]); n is at most NPHYS_MLP_NOALLOC_MAX_ROWS (larger n: NPHYS_ERR_CAPACITY). The
 * batch buffers live in a fixed device array of the library (about 9.9 MB per device, held until
 * the process ends) shared by every call on that device: calls that overlap on one device race.
 * The work is enqueued on `stream` (a hipStream_t; NULL: the network's own stream, and the call
 * waits for it); with a caller's stream the call returns once enqueued. */
#define NPHYS_MLP_NOALLOC_MAX_ROWS 4096
int nphys_mlp_gradients_noalloc(nphys_mlp* m, const float* features, const float* targets, int64_t n,
                                float* grad, double* loss, void* stream);

/* Loads a network (NPHYS_MLP_NUM_PARAMS floats, the layout above) into the world; afterwards
 * nphys_contact_model_set(NPHYS_CONTACT_MODEL_LEARNED_WARMSTART) warm-starts every contact with
 * its cached impulses plus the network's correction (projected on lambda_n >= 0,
 * |lambda_t| <= mu lambda_n). Selecting that model without a network is NPHYS_ERR_STATE. */
int nphys_warmstart_load(nphys_world* w, const float* params, int64_t n);
/* With on, every step records its contacts' features (before the solve) and impulses (after it,
 * in units of r_n |g| dt) as training samples; nphys_read_samples returns the last step's
 * (features [n][24], targets [n][3]; with both NULL only *n is set). The network's training target
 * is targets minus features 16..18 (the cached impulses in the same unit). */
int nphys_capture_samples(nphys_world* w, int32_t on);
int nphys_read_samples(nphys_world* w, float* features, float* targets, int64_t cap, int64_t* n);
/* With the learned model and a residual target set: every step solves twice from the same
 * assembled system, from the cache warm start and from the prediction, and keeps the learned
 * solve only when it reached the target in fewer sweeps (counters judge_*). */
int nphys_set_warmstart_judge(nphys_world* w, int32_t on);

/* Message of the last error on this thread ("" if none). */
const char* nphys_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* NPHYS_H */
