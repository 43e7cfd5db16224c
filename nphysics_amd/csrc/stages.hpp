// nphysics_amd/csrc/stages.hpp -- launch wrappers of the per-stage kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "devworld.hpp"

namespace nph {

// k_bodies.hip
void launch_prestep(const DevBodies& b, const SimParams& p, hipStream_t s);
// only: when given, the bodies with only[i] != 0 (a CCD substep's island)
void launch_integrate(const DevBodies& b, const SimParams& p, hipStream_t s, const uint32_t* only = nullptr);
void launch_clear_dyn_marks(const DevBodies& b, hipStream_t s);
void launch_save_rot_dyn(const DevBodies& b, float4* out, hipStream_t s);
void launch_apply_force(const DevBodies& b, const SimParams& p, int nu, const int* ubody, const int* off,
                        const int* order, const float* forces, const float* points, int type, int wake,
                        const float4* rot_dyn, hipStream_t s);

// k_activation.hip ------------------------------------------------------------------------
struct ActivationBufs {
  int* parent;       // [n] union-find parents
  uint32_t* can;     // [n] per island root: every body may sleep
  uint32_t* touch;   // [n] body that could sleep touches one that keeps its island awake
  uint32_t* mflag;   // [M] manifold handed to the solver
  uint32_t* pos;     // [M] its index in the solved list
  uint32_t* tmp;     // scan temporary
  uint32_t* totals;  // [3] solved manifolds, their contacts, awake dynamic bodies (zeroed by the caller)
};
void launch_wake(const DevBodies& b, hipStream_t s);
struct JointBufs;
// j (may be null): joints whose bodies join the islands
void launch_activation(const DevBodies& b, const DevContacts& c, int M, const ActivationBufs& a, const JointBufs* j,
                       hipStream_t s);
void launch_solve_list(const DevContacts& c, int M, const ActivationBufs& a, const DevContacts& out,
                       const DevContacts& prev, int prevM, hipStream_t s);

// k_broadphase.hip ------------------------------------------------------------------------
struct BroadPhaseBufs {
  uint32_t* changed;      // [2] number of loose AABBs replaced: all, of non-dynamic bodies
  uint32_t* bounds;       // [6] ordered-int encoded centroid bounds
  uint32_t* keys;         // [n] Morton codes
  uint32_t* vals;         // [n] body indices
  uint32_t* keys_alt;
  uint32_t* vals_alt;
  int* child;             // [2(n-1)] children of internal nodes (>= n-1: leaf)
  int* parent;            // [2n-1]
  uint32_t* visit;        // [8(n-1)] internal-node bounds, order-preserving integer encoding
  float4* node_mn;        // [2n-1]
  float4* node_mx;
  int2* node_scene;       // [n-1] scene range (first, last) of every internal node's leaves
  float4* pnode;          // [4(n-1)] packed internal nodes for the traversals (k_broadphase.hip)
  int scene_bits;         // Morton keys carry the scene in their top scene_bits bits
  int* cand;              // [n * cap] candidate partners per body
  uint32_t* cand_count;   // [n] partner count (may exceed cap)
  uint32_t* cand_off;     // [n] exclusive scan of min(count, cap)
  uint32_t* stats;        // [4] 0: max count, 1: traversal stack overflow
  int2* pairs;            // [pair capacity]
  uint32_t* tmp;          // scan / radix scratch
  uint32_t* dirty;        // [n] 1 where the loose AABB was replaced since the last pair build
  int* leaf_of;           // [n] leaf of each body in the current tree
  int* list;              // [n] dirty dynamic bodies (incremental update)
  uint32_t* list_count;   // [1]
};

// mode 0: current pose of every body (static ones included); mode 1: dynamic bodies, current
// pose; mode 2: dynamic bodies, AABB swept from pos0/rot0 to pos/rot.
void launch_loose_update(const DevBodies& b, int mode, uint32_t* changed, uint32_t* dirty, hipStream_t s);
void launch_build_lbvh(const DevBodies& b, BroadPhaseBufs& bp, hipStream_t s);
void launch_traverse(const DevBodies& b, BroadPhaseBufs& bp, int cap, hipStream_t s);
void launch_pair_offsets(const DevBodies& b, BroadPhaseBufs& bp, int cap, hipStream_t s);
void launch_compact_pairs(const DevBodies& b, BroadPhaseBufs& bp, int cap, hipStream_t s);
// Pair segments updated for the dirty bodies only (the tree of the last full build refitted by
// expansion); max_dirty bounds the number of dirty bodies. stats[0] > cap or stats[1] means the
// caller must fall back to a full build.
void launch_incremental_pairs(const DevBodies& b, BroadPhaseBufs& bp, int cap, int max_dirty, hipStream_t s);

// k_narrowphase.hip -------------------------------------------------------------------------
struct NarrowPhaseBufs {
  uint32_t* count;        // [P] contacts per pair
  uint32_t* coff;         // [P] exclusive scan of count
  uint32_t* has;          // [P] count > 0
  uint32_t* moff;         // [P] exclusive scan of has
  uint32_t* perm;         // [P] pairs binned by kind (lane -> pair)
  uint32_t* totals;       // [3] contacts, manifolds, cuboid-cuboid pairs
  uint32_t* tmp;
};
void launch_narrow_phase(const DevBodies& b, const int2* pairs, int npairs, NarrowPhaseBufs& np,
                         DevContacts out, hipStream_t s);
// Contacts of all manifolds in manifold order, gathered from their slots into dense records
// (kContactFields float4 each) and impulses (dense_imp may be null); moff = exclusive scan of the
// manifolds' counts.
void launch_gather_contacts(const DevContacts& c, int M, const uint32_t* moff, float4* dense,
                            float4* dense_imp, hipStream_t s);

// k_solver.hip -----------------------------------------------------------------------------
struct SolverBufs {
  uint32_t* deg;          // [n] dynamic-body degree in the manifold graph
  uint32_t* adj_off;      // [n]
  uint32_t* adj_fill;     // [n]
  int* adj;               // [2M]
  uint32_t* state;        // [M] colouring state: 0xffffffff, or round << 16 | colour
  uint32_t* remaining;    // [2] 0: a manifold left uncoloured by the last round, 1: last round that coloured one
  uint32_t* ckeys;        // [M] colour keys
  uint32_t* order;        // [M] manifold ids in colour order
  uint32_t* ckeys_alt;
  uint32_t* order_alt;
  uint32_t* col_start;    // [max colours + 1]
  uint32_t* mcount_ord;   // [M] contact counts in colour order
  uint32_t* soff;         // [M] first row record of each slot
  int4* sinfo;            // [M] per slot: body1|dyn1<<30, body2|dyn2<<30, first record, count|first<<3
  uint32_t* slot_of;      // [M] slot of each manifold (inverse of order)
  uint32_t* rec2c;        // [C] contact index of each row record
  float4* rec;            // [kRecFields rec_stride] row records, field f of record r at kRecFields r + f
  float4* lam;            // [rec_stride] impulses of each row record (normal, tangent 1, tangent 2, has cache)
  float4* prec;           // [3 rec_stride] position records: (lp1 | contact kind), (lp2 | shape 1), (ln | shape 2)
  float4* pgeo;           // [2M] per slot: (half extents 1 | margin 1), (half extents 2 | margin 2)
  long rec_stride;        // record capacity (fixed between reallocations)
  int* visit;             // [max colours] colour classes in visiting order
  int2* vrange;           // [max colours] (first slot, slots) of the class visited k-th
  double* red;            // residual partials
  uint32_t* tmp;
};
constexpr int kMaxColours = 4096;
constexpr int kRecFields = 4;  // float4 per contact row record

void launch_colour_prepare(const DevBodies& b, const DevContacts& c, int M, const DevContacts& prev,
                           int prevM, const uint32_t* prev_state, SolverBufs& sb, hipStream_t s);
void launch_colour_round(const DevBodies& b, const DevContacts& c, int M, int round, SolverBufs& sb,
                         hipStream_t s);
void launch_colour_keys(int M, SolverBufs& sb, uint32_t* ncol, hipStream_t s);
void launch_colour_order(const DevBodies& b, const DevContacts& c, int M, int ncolours, SolverBufs& sb,
                         hipStream_t s);
void launch_fix_colour_starts(int ncolours, int M, int C, SolverBufs& sb, uint32_t* class_rec, hipStream_t s);
int residual_blocks(int M);
void launch_assemble(const DevBodies& b, const DevContacts& cur, const DevContacts& prev,
                     int prevM, int M, const SimParams& p, SolverBufs& sb, const MatTable& mt, hipStream_t s);
// Sweep kernels of the colour visited k-th; their slot ranges are read on the device, so the
// launches are independent of per-step sizes (grid = blocks from the largest colour class,
// grid-stride over the colour).
int sweep_blocks(int max_class);
int position_blocks(int max_class);
// (first, n): the class's slot range, known on the host from the visiting order
void launch_warmstart_colour(const DevBodies& b, const DevContacts& c, int first, int n, int blocks,
                             SolverBufs& sb, hipStream_t s);
// Consecutive visited classes kv0 .. kv0 + nk - 1, each of at most kRunBlock manifolds, in one
// workgroup (warm start or sweep).
constexpr int kRunBlock = 512;
void launch_colour_run(const DevBodies& b, SolverBufs& sb, int kv0, int nk, bool warm, hipStream_t s);
void launch_pgs_colour(const DevBodies& b, const DevContacts& c, int first, int n, int blocks, SolverBufs& sb,
                       hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
void launch_residual(const DevBodies& b, const DevContacts& c, int M, SolverBufs& sb, hipStream_t s);
void launch_cache_impulses(const DevContacts& c, SolverBufs& sb, int C, hipStream_t s);

// Scene-blocked sweeps (k_solver.hip): every scene's slots in (scene, colour) order, per scene
// and colour the first position, and the scene's dynamic bodies (CSR) with each body's rank.
constexpr int kLdsBodies = 2048;  // dynamic bodies of one scene the blocked kernel keeps in LDS
constexpr int kSmallScene = 256;  // scenes up to this many dynamic bodies use the small workgroups
struct SceneBlocks {
  uint32_t* keys;        // [M] scene << 12 | colour, sorted
  uint32_t* slot;        // [M] slots in that order
  uint32_t* keys_alt;
  uint32_t* slot_alt;
  uint32_t* start;       // [nscenes][ncolours + 1]
  const int* dyn_start;  // [nscenes + 1]
  const int* dyn_list;   // dynamic bodies scene by scene
  const int* local;      // [n] rank of a dynamic body within its scene
  const int* small_list; // scenes with at most kSmallScene dynamic bodies
  const int* large_list; // the other scenes
  int n_small, n_large;
};
void launch_scene_order(const DevBodies& b, int M, int nscenes, int ncolours, SolverBufs& sb, SceneBlocks& sc,
                        hipStream_t s);
// block: lanes per large scene (256, 512 or 1024); the large scenes go on s_large, the small on s_small
void launch_sweeps_blocked(const DevBodies& b, SolverBufs& sb, const SceneBlocks& sc, int ncolours, int nsweeps,
                           bool warm, int block, hipStream_t s_small, hipStream_t s_large);
// Scene-blocked position correction: `iterations` passes over the colours in the visiting order,
// one workgroup per scene with the scene's dynamic poses in LDS (the same corrections in the same
// order per body as the per-colour launches); streams and block as launch_sweeps_blocked.
void launch_position_blocked(const DevBodies& b, const SimParams& p, SolverBufs& sb, const SceneBlocks& sc,
                             int ncolours, int iterations, int block, hipStream_t s_small, hipStream_t s_large);

// learned warm start (k_solver.hip): features [C][24] and impulse units ls [C] per row record,
// the prediction y [C][3] applied as the warm start, post-solve targets t [C][3].
void launch_features(const DevBodies& b, const DevContacts& c, const SimParams& p, int M, SolverBufs& sb, float* X,
                     float* ls, hipStream_t s);
void launch_apply_prediction(SolverBufs& sb, int C, const float* y, const float* ls, hipStream_t s);
void launch_sample_targets(SolverBufs& sb, int C, const float* ls, float* t, hipStream_t s);
void launch_lam_copy(SolverBufs& sb, int C, float4* buf, bool to_buf, hipStream_t s);

// k_joints.hip -----------------------------------------------------------------------------
// Joint constraints (ball, fixed, revolute): static SoA per joint in joint order, a fixed greedy
// colouring (joints sharing a dynamic body differ in colour) with the joints listed colour by
// colour, and per solve one row record per joint at its position in that list.
enum { JOINT_BALL = 0, JOINT_FIXED = 1, JOINT_REVOLUTE = 2, JOINT_PRISMATIC = 3, JOINT_CYLINDRICAL = 4 };
// linear rows of a joint kind (they come first): the prismatic and cylindrical joints' two (the
// directions orthogonal to the axis), the others' three
__host__ __device__ inline int joint_linear_rows(int kind) {
  return kind == JOINT_PRISMATIC || kind == JOINT_CYLINDRICAL ? 2 : 3;
}
constexpr int kJointRec = 10;  // float4 per joint row record
struct JointBufs {
  int nj;                // joints
  int ncol;              // joint colours
  const int4* info;      // [nj] (kind, body1, body2, colour)
  const float4* a1;      // [nj] local anchors
  const float4* a2;
  const float4* f1;      // [nj] fixed: reference frames (i, j, k, w); revolute: local axes (xyz)
  const float4* f2;
  const float2* brk;     // [nj] squared break force, squared break torque
  float4* imp_lin;       // [nj] impulses cached by the last solve
  float4* imp_ang;
  int* state;            // [nj] bit 0 broken, bit 1 removed
  const int* order;      // [nj] joints colour by colour (stable in joint order)
  const int2* crange;    // [ncol] (first position in order, joints)
  // per solve, at the joint's position in `order`: (p1 | rows), (p2 | active), 6 x (dir | r),
  // (rhs 0-3), (rhs 4-5, -, -); impulses of rows 0-3 and 4-5
  float4* rec;
  float4* lam;
};
void launch_joint_assemble(const DevBodies& b, const SimParams& p, JointBufs& j, hipStream_t s);
// one workgroup over every joint colour in order (small joint sets), or one launch per colour
void launch_joint_warm(const DevBodies& b, JointBufs& j, hipStream_t s);
void launch_joint_sweep(const DevBodies& b, JointBufs& j, hipStream_t s);
void launch_joint_position(const DevBodies& b, const SimParams& p, JointBufs& j, hipStream_t s);
void launch_joint_cache(const DevBodies& b, const SimParams& p, JointBufs& j, hipStream_t s);
// residual of the joint rows (phi = w): (max, sum of squares, rows) into out[0..2]
void launch_joint_residual(const DevBodies& b, JointBufs& j, double* out, hipStream_t s);
// union of the bodies of every joint neither broken nor removed (activation islands)
void launch_joint_union(const DevBodies& b, const JointBufs& j, int* parent, uint32_t* touch, hipStream_t s);

// k_ccd.hip --------------------------------------------------------------------------------
// A queued time of impact: key = time bits << 32 | pair index (times are >= 0, so keys order by
// time, then pair) and the pair's CCD slot.
struct ToiEntry {
  unsigned long long key;
  int slot;
};
struct CcdBufs {
  float* body_time;        // [n] time each body's motion starts from (SubstepState::body_times)
  float* frozen;           // [n] per body key: freeze time, < 0 when not frozen
  int* slot_pair;          // [P] pair index of each CCD slot (pairs with a CCD-enabled side)
  int4* slot_bodies;       // [P] the slot's two bodies and their keys; a key | kSlotStatic when its
                           // body is not dynamic
  int* pos;                // [P] node of each slot's entry in the queue, -1 when it has none
  ToiEntry* heap;          // [cap] first-pass entries, then the queue's nodes below its LDS levels
  ToiEntry* stage;         // [P] entries of one re-sweep
  unsigned long long* impacts;  // [cap] keys of the confirmed impacts, in order
  uint32_t* ctl;           // control words kCcd*
  int cap;
  uint32_t* in_t;          // [n] per key: its colliders are traversed
  uint32_t* new_t;         // [n] per key: traversed from the next round on
  uint32_t* prep;          // [n] to be moved to its impact time
  uint32_t* prepared;      // [n] moved: the substep's island
  uint32_t* evaluated;     // [P]
  uint32_t* kept;          // [P] manifold not empty at the moved poses
  uint32_t* cand;          // [P] to evaluate in this round
  uint32_t* cpos;          // [P] exclusive scan of a flag array
  int2* list;              // [P] compacted pairs, pair order
  int* list_pair;          // [P] their pair indices
  uint32_t* tmp;           // scan scratch
  uint32_t* deg;           // [n] CCD slots of each dynamic body
  uint32_t* off;           // [n] exclusive scan of deg
  uint32_t* fill;          // [n]
  int* adj;                // [2P] the slots of each dynamic body, at off
};
enum : int {
  kCcdSlots = 0, kCcdEntries = 1, kCcdAnyIn = 2, kCcdMinAfter = 3, kCcdKind = 4, kCcdImpacts = 5,
  kCcdLastToi = 6, kCcdOverflow = 7, kCcdMinOver = 8, kCcdList = 9, kCcdWords = 16
};
enum : int { PREDICT_NONE = 0, PREDICT_AFTER_END = 1, PREDICT_IMPACTS = 2 };
constexpr int kSlotStatic = 1 << 30;  // body counts stay below 2^30
// predict_next_impacts up to end_time: control words reset, freeze times cleared, every CCD pair's
// time of impact, then the queue in one workgroup (kind, impacts, last impact time in ctl)
void launch_ccd_predict(const DevBodies& b, const int2* pairs, int npairs, CcdBufs& c, float end_time, float dt0,
                        float target, int on_penetration, hipStream_t s);
void launch_ccd_seed(const DevBodies& b, const int2* pairs, CcdBufs& c, int nimp, hipStream_t s);
void launch_ccd_touch(const DevBodies& b, const int2* pairs, int npairs, CcdBufs& c, hipStream_t s);
void launch_ccd_prepare(const DevBodies& b, CcdBufs& c, float dt_sub, float last_toi, hipStream_t s);
// the pairs with flag[p] != 0 into c.list / c.list_pair in pair order, their count to ctl[kCcdList]
void launch_ccd_compact(const int2* pairs, int npairs, const uint32_t* flag, CcdBufs& c, hipStream_t s);
void launch_ccd_grow(const DevBodies& b, CcdBufs& c, int nlist, const uint32_t* count, hipStream_t s);
void launch_ccd_merge(int n, CcdBufs& c, hipStream_t s);
void launch_ccd_validate(const DevBodies& b, const int2* pairs, CcdBufs& c, int nimp, hipStream_t s);
void launch_ccd_clamp(const DevBodies& b, CcdBufs& c, hipStream_t s);

// k_position.hip ---------------------------------------------------------------------------
void launch_position_colour(const DevBodies& b, const DevContacts& c, const SimParams& p, int first, int n,
                            int blocks, SolverBufs& sb, hipStream_t s);
void launch_position_run(const DevBodies& b, const SimParams& p, SolverBufs& sb, int kv0, int nk, hipStream_t s);

}  // namespace nph
