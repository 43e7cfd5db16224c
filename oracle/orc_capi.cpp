// oracle/orc_capi.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// C entry points of the CPU restatement (liboracle.so), loaded with ctypes by tests/ and by
// bench.py's cpu_baseline leg only. They mirror the product ABI in include/nphys.h (same
// parameter/body structs, same call order) with an `orc_` prefix, a precision switch (FP32 or
// FP64, nphysics being generic over RealField) and a row-order switch (reference order or the
// GPU's colour order). Read-back is always in double.
#include <cstring>
#include <string>
#include <type_traits>

#include "../include/nphys.h"
#include "orc_world.hpp"

namespace {

struct AnyWorld {
  virtual ~AnyWorld() {}
  virtual int precision() const = 0;
};

template <class T>
struct TWorld : AnyWorld {
  orc::World<T> w;
  int precision() const override { return sizeof(T) == 4 ? 32 : 64; }
};

thread_local std::string g_err;

int fail(int code, const char* msg) {
  g_err = msg;
  return code;
}

template <class T>
void apply_params(orc::World<T>& w, const nphys_params* p) {
  w.prm.t = T(p->t);
  w.prm.set_dt(T(p->dt));
  w.prm.erp = T(p->erp);
  w.prm.warmstart_coeff = T(p->warmstart_coeff);
  w.prm.restitution_velocity_threshold = T(p->restitution_velocity_threshold);
  w.prm.allowed_linear_error = T(p->allowed_linear_error);
  w.prm.allowed_angular_error = T(p->allowed_angular_error);
  w.prm.max_linear_correction = T(p->max_linear_correction);
  w.prm.max_angular_correction = T(p->max_angular_correction);
  w.prm.max_stabilization_multiplier = T(p->max_stabilization_multiplier);
  w.prm.max_velocity_iterations = p->max_velocity_iterations;
  w.prm.max_position_iterations = p->max_position_iterations;
  w.prm.max_ccd_position_iterations = p->max_ccd_position_iterations;
  w.prm.max_ccd_substeps = p->max_ccd_substeps;
  w.prm.ccd_on_penetration = p->ccd_on_penetration_enabled;
  w.gravity = orc::V3<T>{T(p->gravity[0]), T(p->gravity[1]), T(p->gravity[2])};
}

template <class F>
int with_world(void* h, F&& f) {
  if (!h) return fail(NPHYS_ERR_INVALID, "null world");
  AnyWorld* a = static_cast<AnyWorld*>(h);
  if (a->precision() == 32) return f(static_cast<TWorld<float>*>(a)->w);
  return f(static_cast<TWorld<double>*>(a)->w);
}

template <class T, class S>
orc::V3<T> v3_of(const S* p) { return orc::V3<T>{T(p[0]), T(p[1]), T(p[2])}; }
template <class T, class S>
orc::Q<T> q_of(const S* q) { return orc::Q<T>{T(q[0]), T(q[1]), T(q[2]), T(q[3])}; }
template <class T>
void put3(double* o, orc::V3<T> v) {
  o[0] = double(v.x); o[1] = double(v.y); o[2] = double(v.z);
}

template <class T>
int add_bodies_t(orc::World<T>& w, const nphys_body_desc* d, int32_t* first) {
  if (!d->status || !d->shape || !d->dims || !d->pos || !d->rot || !d->density || !d->margin ||
      !d->linear_prediction || !d->friction || !d->restitution)
    return fail(NPHYS_ERR_INVALID, "missing body array");
  if (first) *first = int32_t(w.bodies.size());
  for (int i = 0; i < d->n; ++i) {
    if (d->status[i] != orc::STATUS_DYNAMIC && d->status[i] != orc::STATUS_STATIC)
      return fail(NPHYS_ERR_INVALID, "bad status");
    if (d->shape[i] != orc::SHAPE_BALL && d->shape[i] != orc::SHAPE_CUBOID)
      return fail(NPHYS_ERR_INVALID, "bad shape");
    if (d->scene && (d->scene[i] < 0 || d->scene[i] >= NPHYS_MAX_SCENES)) return fail(NPHYS_ERR_INVALID, "bad scene");
    const orc::V3<T> zero{T(0), T(0), T(0)};
    const orc::V3<T> lv = d->linvel ? v3_of<T>(d->linvel + 3 * i) : zero;
    const orc::V3<T> av = d->angvel ? v3_of<T>(d->angvel + 3 * i) : zero;
    orc::BodyExtras<T> ex;
    if (d->linear_damping) ex.lin_damping = T(d->linear_damping[i]);
    if (d->angular_damping) ex.ang_damping = T(d->angular_damping[i]);
    if (d->max_linear_velocity) ex.max_linvel = T(d->max_linear_velocity[i]);
    if (d->max_angular_velocity) ex.max_angvel = T(d->max_angular_velocity[i]);
    if (d->gravity_enabled) ex.gravity = d->gravity_enabled[i] != 0;
    if (d->kinematic_dofs) ex.kinematic_dofs = d->kinematic_dofs[i] & 63;
    if (d->friction_combine) ex.fric_mode = d->friction_combine[i];
    if (d->restitution_combine) ex.rest_mode = d->restitution_combine[i];
    if (d->material_id) ex.mat_id = d->material_id[i] < 0 ? -1 : d->material_id[i];
    if (d->surface_velocity) ex.surf_vel = v3_of<T>(d->surface_velocity + 3 * i);
    if (d->ccd_enabled) ex.ccd = d->ccd_enabled[i] != 0;
    if (d->linear_motion_interpolation) ex.lin_interp = d->linear_motion_interpolation[i] != 0;
    if (d->static_body) ex.static_body = d->static_body[i] < 0 ? -1 : d->static_body[i];
    if (d->collision_groups) {
      const uint32_t all = (1u << 30) - 1u;  // groups 0-29
      ex.g_member = d->collision_groups[3 * i] & all;
      ex.g_white = d->collision_groups[3 * i + 1] & all;
      ex.g_black = d->collision_groups[3 * i + 2] & all;
    }
    if (ex.fric_mode < 0 || ex.fric_mode > 3 || ex.rest_mode < 0 || ex.rest_mode > 3)
      return fail(NPHYS_ERR_INVALID, "bad combine mode");
    w.add_body(d->status[i], d->shape[i], v3_of<T>(d->dims + 3 * i),
               orc::Iso<T>{v3_of<T>(d->pos + 3 * i), q_of<T>(d->rot + 4 * i)}, lv, av,
               T(d->density[i]), T(d->margin[i]), T(d->linear_prediction[i]), T(d->friction[i]),
               T(d->restitution[i]), d->scene ? int(d->scene[i]) : 0,
               d->sleep_threshold ? T(d->sleep_threshold[i]) : T(-1), ex);
  }
  return NPHYS_OK;
}

template <class T>
int add_joints_t(orc::World<T>& w, const nphys_joint_desc* d, int32_t* first) {
  if (!d || d->n < 0 || (d->n > 0 && (!d->kind || !d->body1 || !d->body2 || !d->anchor1 || !d->anchor2)))
    return fail(NPHYS_ERR_INVALID, "bad joint descriptor");
  const int nb = int(w.bodies.size());
  for (int i = 0; i < d->n; ++i) {
    const int b1 = d->body1[i], b2 = d->body2[i];
    if (b1 < 0 || b1 >= nb || b2 < 0 || b2 >= nb || b1 == b2) return fail(NPHYS_ERR_INVALID, "bad joint body");
    if (d->kind[i] < 0 || d->kind[i] > 4) return fail(NPHYS_ERR_INVALID, "bad joint kind");
    if (w.bodies[size_t(b1)].status != orc::STATUS_DYNAMIC && w.bodies[size_t(b2)].status != orc::STATUS_DYNAMIC)
      return fail(NPHYS_ERR_INVALID, "a joint needs a dynamic body");
  }
  if (first) *first = int32_t(w.joints.size());
  auto v3 = [](const float* p, int i, T x, T y, T z) {
    return p ? orc::V3<T>{T(p[3 * i]), T(p[3 * i + 1]), T(p[3 * i + 2])} : orc::V3<T>{x, y, z};
  };
  auto q4 = [](const float* p, int i) {
    return p ? orc::Q<T>{T(p[4 * i]), T(p[4 * i + 1]), T(p[4 * i + 2]), T(p[4 * i + 3])}
             : orc::Q<T>{T(0), T(0), T(0), T(1)};
  };
  for (int i = 0; i < d->n; ++i)
    w.add_joint(d->kind[i], d->body1[i], d->body2[i], v3(d->anchor1, i, 0, 0, 0), v3(d->anchor2, i, 0, 0, 0),
                q4(d->frame1, i), q4(d->frame2, i), v3(d->axis1, i, 0, 0, 1), v3(d->axis2, i, 0, 0, 1),
                d->break_force ? T(d->break_force[i]) : T(-1), d->break_torque ? T(d->break_torque[i]) : T(-1));
  return NPHYS_OK;
}

template <class T, class S>
int set_state_t(orc::World<T>& w, const S* pos, const S* rot, const S* lv, const S* av) {
  for (size_t i = 0; i < w.bodies.size(); ++i) {
    orc::Body<T>& b = w.bodies[i];
    orc::Iso<T> p = b.pos;
    if (pos) p.t = v3_of<T>(pos + 3 * i);
    if (rot) p.q = q_of<T>(rot + 4 * i);
    w.set_state(int(i), p, lv ? v3_of<T>(lv + 3 * i) : b.linvel, av ? v3_of<T>(av + 3 * i) : b.angvel);
  }
  return NPHYS_OK;
}

template <class T>
int get_state_t(orc::World<T>& w, double* pos, double* rot, double* lv, double* av) {
  for (size_t i = 0; i < w.bodies.size(); ++i) {
    const orc::Body<T>& b = w.bodies[i];
    if (pos) put3(pos + 3 * i, b.pos.t);
    if (rot) {
      rot[4 * i + 0] = double(b.pos.q.i);
      rot[4 * i + 1] = double(b.pos.q.j);
      rot[4 * i + 2] = double(b.pos.q.k);
      rot[4 * i + 3] = double(b.pos.q.w);
    }
    if (lv) put3(lv + 3 * i, b.linvel);
    if (av) put3(av + 3 * i, b.angvel);
  }
  return NPHYS_OK;
}

template <class T>
int read_contacts_t(orc::World<T>& w, int32_t* bodies, uint32_t* feature, double* w1, double* w2,
                    double* n, double* depth, int32_t* kind, double* lp1, double* lp2, double* ln,
                    int64_t cap) {
  if (int64_t(w.contacts.size()) > cap) return fail(NPHYS_ERR_CAPACITY, "contact buffer too small");
  int64_t c = 0;
  for (const auto& m : w.manifolds) {
    for (int k = 0; k < m.count; ++k, ++c) {
      const orc::Contact<T>& ct = w.contacts[m.first + k];
      if (bodies) { bodies[2 * c] = m.b1; bodies[2 * c + 1] = m.b2; }
      if (feature) feature[c] = ct.feature;
      if (w1) put3(w1 + 3 * c, ct.world1);
      if (w2) put3(w2 + 3 * c, ct.world2);
      if (n) put3(n + 3 * c, ct.normal);
      if (depth) depth[c] = double(ct.depth);
      if (kind) kind[c] = ct.kind;
      if (lp1) put3(lp1 + 3 * c, ct.lp1);
      if (lp2) put3(lp2 + 3 * c, ct.lp2);
      if (ln) put3(ln + 3 * c, ct.ln);
    }
  }
  return NPHYS_OK;
}

template <class T>
int read_impulses_t(orc::World<T>& w, uint64_t* keys, double* imp, int64_t cap) {
  if (int64_t(w.crow.size()) > cap) return fail(NPHYS_ERR_CAPACITY, "impulse buffer too small");
  for (size_t i = 0; i < w.crow.size(); ++i) {
    const orc::ContactRows& cr = w.crow[i];
    const auto& nl = cr.ground ? w.uni_g : w.uni;
    const auto& bl = cr.ground ? w.bil_g : w.bil;
    if (keys) keys[i] = nl[cr.normal].key;
    if (imp) {
      imp[3 * i + 0] = double(nl[cr.normal].impulse);
      imp[3 * i + 1] = double(bl[cr.fric[0]].impulse);
      imp[3 * i + 2] = double(bl[cr.fric[1]].impulse);
    }
  }
  return NPHYS_OK;
}

template <class T>
int read_body_dynamics_t(orc::World<T>& w, double* inv_mass, double* inv_aug, double* aabb) {
  for (size_t i = 0; i < w.bodies.size(); ++i) {
    const orc::Body<T>& b = w.bodies[i];
    if (inv_mass) inv_mass[i] = double(b.inv_mass);
    if (inv_aug)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) inv_aug[9 * i + 3 * r + c] = double(b.inv_aug.m[r][c]);
    if (aabb) {
      put3(aabb + 6 * i, b.loose.mn);
      put3(aabb + 6 * i + 3, b.loose.mx);
    }
  }
  return NPHYS_OK;
}

template <class T>
int collide_shapes_t(int k1, const double* d1, const double* p1, int k2, const double* d2,
                     const double* p2, double prediction, uint32_t* feature, double* w1,
                     double* w2, double* n, double* depth, int32_t* kind) {
  orc::ShapeDesc<T> s1{k1, k1 == orc::SHAPE_BALL ? orc::V3<T>{T(d1[0]), T(d1[0]), T(d1[0])} : v3_of<T>(d1)};
  orc::ShapeDesc<T> s2{k2, k2 == orc::SHAPE_BALL ? orc::V3<T>{T(d2[0]), T(d2[0]), T(d2[0])} : v3_of<T>(d2)};
  orc::Iso<T> i1{v3_of<T>(p1), q_of<T>(p1 + 3)}, i2{v3_of<T>(p2), q_of<T>(p2 + 3)};
  orc::Contact<T> buf[8];
  const int nc = orc::collide_pair(s1, i1, s2, i2, T(prediction), buf);
  for (int c = 0; c < nc; ++c) {
    if (feature) feature[c] = buf[c].feature;
    if (w1) put3(w1 + 3 * c, buf[c].world1);
    if (w2) put3(w2 + 3 * c, buf[c].world2);
    if (n) put3(n + 3 * c, buf[c].normal);
    if (depth) depth[c] = double(buf[c].depth);
    if (kind) kind[c] = buf[c].kind;
  }
  return nc;
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_world_create(const nphys_params* p, int32_t precision, void** out) {
  if (!p || !out) return fail(NPHYS_ERR_INVALID, "null argument");
  if (p->dt < 0) return fail(NPHYS_ERR_INVALID, "negative dt");
  if (precision == 32) {
    auto* w = new TWorld<float>();
    apply_params(w->w, p);
    *out = static_cast<AnyWorld*>(w);
  } else if (precision == 64) {
    auto* w = new TWorld<double>();
    apply_params(w->w, p);
    *out = static_cast<AnyWorld*>(w);
  } else {
    return fail(NPHYS_ERR_INVALID, "precision must be 32 or 64");
  }
  return NPHYS_OK;
}

int orc_world_destroy(void* h) {
  delete static_cast<AnyWorld*>(h);
  return NPHYS_OK;
}

int orc_set_params(void* h, const nphys_params* p) {
  if (!p) return fail(NPHYS_ERR_INVALID, "null params");
  if (p->dt < 0) return fail(NPHYS_ERR_INVALID, "negative dt");
  return with_world(h, [&](auto& w) { apply_params(w, p); return NPHYS_OK; });
}

int orc_set_order_mode(void* h, int32_t mode) {
  if (mode != orc::ORDER_REFERENCE && mode != orc::ORDER_COLOURED)
    return fail(NPHYS_ERR_INVALID, "bad order mode");
  return with_world(h, [&](auto& w) { w.order_mode = mode; return NPHYS_OK; });
}

int orc_set_residual_target(void* h, double target, int32_t max_sweeps, int32_t check_every, int32_t norm) {
  if (!(target >= 0.0) || max_sweeps < 0 || check_every < 1 || (norm != 0 && norm != 1))
    return fail(NPHYS_ERR_INVALID, "bad residual target");
  return with_world(h, [&](auto& w) {
    w.res_norm = norm;
    w.res_target = target;
    w.res_max_sweeps = max_sweeps;
    w.res_check_every = check_every;
    return NPHYS_OK;
  });
}

int orc_set_brute_force_limit(void* h, int32_t n) {
  return with_world(h, [&](auto& w) { w.brute_force_limit = n; return NPHYS_OK; });
}

int orc_add_bodies(void* h, const nphys_body_desc* d, int32_t* first) {
  if (!d || d->n < 0) return fail(NPHYS_ERR_INVALID, "bad body desc");
  return with_world(h, [&](auto& w) { return add_bodies_t(w, d, first); });
}

int orc_add_joints(void* h, const nphys_joint_desc* d, int32_t* first) {
  return with_world(h, [&](auto& w) { return add_joints_t(w, d, first); });
}

int orc_remove_joint(void* h, int32_t j) {
  return with_world(h, [&](auto& w) {
    if (j < 0 || j >= int(w.joints.size())) return fail(NPHYS_ERR_INVALID, "bad joint handle");
    w.joints[size_t(j)].removed = true;
    return NPHYS_OK;
  });
}

// lin [n][3], ang [n][3] (revolute: 2 used), broken [n]
int orc_read_joints(void* h, double* lin, double* ang, int32_t* broken, int64_t cap) {
  return with_world(h, [&](auto& w) {
    if (cap < int64_t(w.joints.size())) return fail(NPHYS_ERR_CAPACITY, "buffer too small");
    for (size_t i = 0; i < w.joints.size(); ++i) {
      const auto& j = w.joints[i];
      if (lin) put3(lin + 3 * i, j.lin_imp);
      if (ang) put3(ang + 3 * i, j.ang_imp);
      if (broken) broken[i] = j.broken ? 1 : 0;
    }
    return NPHYS_OK;
  });
}

int orc_num_bodies(void* h, int32_t* n) {
  return with_world(h, [&](auto& w) { *n = int32_t(w.bodies.size()); return NPHYS_OK; });
}

int orc_set_collision_groups(void* h, int32_t count, const int32_t* bodies, const uint32_t* groups) {
  return with_world(h, [&](auto& w) {
    const uint32_t all = (1u << 30) - 1u;  // groups 0-29
    for (int k = 0; k < count; ++k) {
      if (bodies[k] < 0 || size_t(bodies[k]) >= w.bodies.size()) return fail(NPHYS_ERR_INVALID, "bad body handle");
      w.set_collision_groups(bodies[k], groups[3 * k] & all, groups[3 * k + 1] & all, groups[3 * k + 2] & all);
    }
    return int(NPHYS_OK);
  });
}

int orc_set_state(void* h, const float* pos, const float* rot, const float* lv, const float* av) {
  return with_world(h, [&](auto& w) { return set_state_t(w, pos, rot, lv, av); });
}

int orc_set_state_f64(void* h, const double* pos, const double* rot, const double* lv,
                      const double* av) {
  return with_world(h, [&](auto& w) { return set_state_t(w, pos, rot, lv, av); });
}

int orc_get_state(void* h, double* pos, double* rot, double* lv, double* av) {
  return with_world(h, [&](auto& w) { return get_state_t(w, pos, rot, lv, av); });
}

int orc_apply_force(void* h, int32_t count, const int32_t* bodies, const float* forces, const float* points,
                    int32_t type, int32_t wake) {
  if (count < 0 || (count > 0 && (!bodies || !forces)) || type < 0 || type > 3)
    return fail(NPHYS_ERR_INVALID, "bad force arguments");
  return with_world(h, [&](auto& w) {
    using T = typename std::remove_reference<decltype(w)>::type::Scalar;
    for (int k = 0; k < count; ++k)
      if (bodies[k] < 0 || bodies[k] >= int(w.bodies.size())) return fail(NPHYS_ERR_INVALID, "bad body handle");
    for (int k = 0; k < count; ++k) {
      if (points) {
        w.apply_force_at_point(bodies[k], v3_of<T>(forces + 3 * k), v3_of<T>(points + 3 * k), type, wake != 0);
      } else {
        w.apply_force(bodies[k], v3_of<T>(forces + 6 * k), v3_of<T>(forces + 6 * k + 3), type, wake != 0);
      }
    }
    return NPHYS_OK;
  });
}

int orc_set_material_coefficient(void* h, int32_t which, int32_t id1, int32_t id2, float coeff) {
  if ((which != 0 && which != 1) || id1 < 0 || id2 < 0) return fail(NPHYS_ERR_INVALID, "bad material pair");
  return with_world(h, [&](auto& w) {
    using T = typename std::remove_reference<decltype(w)>::type::Scalar;
    const std::pair<uint32_t, uint32_t> key{uint32_t(std::min(id1, id2)), uint32_t(std::max(id1, id2))};
    (which == 0 ? w.fric_table : w.rest_table)[key] = T(coeff);
    return NPHYS_OK;
  });
}

int orc_step(void* h, int32_t nsteps) {
  if (nsteps < 0) return fail(NPHYS_ERR_INVALID, "negative step count");
  return with_world(h, [&](auto& w) {
    if (w.any_ccd && (w.any_sleeper || w.max_scene() > 0))
      return fail(NPHYS_ERR_INVALID, "CCD needs one scene and no sleep thresholds");
    for (int i = 0; i < nsteps; ++i) {
      w.step();
      if (!w.step_error.empty()) return fail(NPHYS_ERR_STATE, w.step_error.c_str());
    }
    return NPHYS_OK;
  });
}

int orc_collide(void* h) {
  return with_world(h, [&](auto& w) {
    w.collide();
    return NPHYS_OK;
  });
}

int orc_num_pairs(void* h, int64_t* n) {
  return with_world(h, [&](auto& w) { *n = int64_t(w.pairs.size()); return NPHYS_OK; });
}

int orc_read_pairs(void* h, int32_t* pairs, int64_t cap) {
  return with_world(h, [&](auto& w) {
    if (int64_t(w.pairs.size()) > cap) return fail(NPHYS_ERR_CAPACITY, "pair buffer too small");
    for (size_t i = 0; i < w.pairs.size(); ++i) {
      pairs[2 * i] = w.pairs[i].first;
      pairs[2 * i + 1] = w.pairs[i].second;
    }
    return NPHYS_OK;
  });
}

int orc_num_contacts(void* h, int64_t* n) {
  return with_world(h, [&](auto& w) { *n = int64_t(w.contacts.size()); return NPHYS_OK; });
}

int orc_read_contacts(void* h, int32_t* bodies, uint32_t* feature, double* w1, double* w2,
                      double* n, double* depth, int32_t* kind, double* lp1, double* lp2,
                      double* ln, int64_t cap) {
  return with_world(h, [&](auto& w) {
    return read_contacts_t(w, bodies, feature, w1, w2, n, depth, kind, lp1, lp2, ln, cap);
  });
}

int orc_num_manifolds(void* h, int64_t* n) {
  return with_world(h, [&](auto& w) { *n = int64_t(w.manifolds.size()); return NPHYS_OK; });
}

int orc_read_manifolds(void* h, int32_t* bodies, int32_t* first, int32_t* count, int64_t cap) {
  return with_world(h, [&](auto& w) {
    if (int64_t(w.manifolds.size()) > cap) return fail(NPHYS_ERR_CAPACITY, "manifold buffer too small");
    for (size_t m = 0; m < w.manifolds.size(); ++m) {
      const orc::Manifold& mf = w.manifolds[m];
      if (bodies) {
        bodies[2 * m] = mf.b1;
        bodies[2 * m + 1] = mf.b2;
      }
      if (first) first[m] = mf.first;
      if (count) count[m] = mf.count;
    }
    return NPHYS_OK;
  });
}

int orc_num_solved_contacts(void* h, int64_t* n) {
  return with_world(h, [&](auto& w) { *n = int64_t(w.crow.size()); return NPHYS_OK; });
}

int orc_read_impulses(void* h, uint64_t* keys, double* imp, int64_t cap) {
  return with_world(h, [&](auto& w) { return read_impulses_t(w, keys, imp, cap); });
}

int orc_residual(void* h, double* rinf, double* rrms) {
  return with_world(h, [&](auto& w) {
    *rinf = double(w.res_inf);
    *rrms = double(w.res_rms);
    return NPHYS_OK;
  });
}

int orc_read_counters(void* h, nphys_counters* c) {
  return with_world(h, [&](auto& w) {
    std::memset(c, 0, sizeof(*c));
    c->step = w.counters.step;
    c->collision_detection = w.counters.collision_detection;
    c->broad_phase = w.counters.broad_phase;
    c->narrow_phase = w.counters.narrow_phase;
    c->assembly = w.counters.assembly;
    c->velocity_resolution = w.counters.velocity_resolution;
    c->velocity_update = w.counters.velocity_update;
    c->position_resolution = w.counters.position_resolution;
    c->nsteps = w.counters.nsteps;
    c->npairs = int64_t(w.pairs.size());
    c->ncontacts = int64_t(w.crow.size());
    c->nmanifolds = int64_t(w.solved.size());
    c->nconstraints = w.n_rows;
    c->ncolours = w.ncolours;
    c->velocity_sweeps = w.velocity_sweeps;
    int64_t active = 0;
    for (const auto& b : w.bodies) active += (b.status == orc::STATUS_DYNAMIC && b.energy != 0) ? 1 : 0;
    c->active_bodies = active;
    c->ccd_time = w.counters.ccd;
    c->ccd_substeps = w.ccd_substeps;
    c->ccd_impacts = w.ccd_impacts;
    return NPHYS_OK;
  });
}

int orc_reset_counters(void* h) {
  return with_world(h, [&](auto& w) {
    w.counters = orc::Counters();
    w.velocity_sweeps = 0;
    return NPHYS_OK;
  });
}

int orc_read_activation(void* h, double* energy, int64_t cap) {
  return with_world(h, [&](auto& w) {
    if (int64_t(w.bodies.size()) > cap) return fail(NPHYS_ERR_CAPACITY, "energy buffer too small");
    for (size_t i = 0; i < w.bodies.size(); ++i) energy[i] = double(w.bodies[i].energy);
    return NPHYS_OK;
  });
}

int orc_read_colours(void* h, int32_t* colours, int64_t cap, int32_t* ncol) {
  return with_world(h, [&](auto& w) {
    if (int64_t(w.colours.size()) > cap) return fail(NPHYS_ERR_CAPACITY, "colour buffer too small");
    for (size_t i = 0; i < w.colours.size(); ++i) colours[i] = w.colours[i];
    *ncol = w.ncolours;
    return NPHYS_OK;
  });
}

int orc_read_body_dynamics(void* h, double* inv_mass, double* inv_aug, double* loose_aabb) {
  return with_world(h, [&](auto& w) { return read_body_dynamics_t(w, inv_mass, inv_aug, loose_aabb); });
}

// Stand-alone narrow phase on one shape pair; pose = (x, y, z, i, j, k, w). Returns the number of
// contacts (<= 4) written, or a negative error.
int orc_collide_shapes(int32_t precision, int32_t k1, const double* d1, const double* p1,
                       int32_t k2, const double* d2, const double* p2, double prediction,
                       uint32_t* feature, double* w1, double* w2, double* n, double* depth,
                       int32_t* kind) {
  if (precision == 32)
    return collide_shapes_t<float>(k1, d1, p1, k2, d2, p2, prediction, feature, w1, w2, n, depth, kind);
  return collide_shapes_t<double>(k1, d1, p1, k2, d2, p2, prediction, feature, w1, w2, n, depth, kind);
}

// Stand-alone time of impact of two shapes (the CCD's conservative advancement). Motion m =
// (kind, t0, x, y, z, i, j, k, w, vx, vy, vz, wx, wy, wz): kind 0 static, 1 linear, 2 rotating;
// pose at t = p0 moved for t - t0. Returns 1 with *toi and *status when an impact is found by
// max_toi, else 0.
int orc_time_of_impact(int32_t precision, int32_t k1, const double* d1, const double* m1, int32_t k2,
                       const double* d2, const double* m2, double max_toi, double target, double* toi,
                       int32_t* status) {
  auto run = [&](auto zero) -> int {
    using T = decltype(zero);
    auto shape = [](int32_t k, const double* d) {
      orc::ShapeDesc<T> sd;
      sd.kind = k;
      sd.he = k == orc::SHAPE_BALL ? orc::V3<T>{T(d[0]), T(d[0]), T(d[0])} : orc::V3<T>{T(d[0]), T(d[1]), T(d[2])};
      return sd;
    };
    auto motion = [](const double* m) {
      orc::CcdMotion<T> c;
      c.kind = int(m[0]);
      c.t0 = T(m[1]);
      c.p0 = orc::Iso<T>{orc::V3<T>{T(m[2]), T(m[3]), T(m[4])}, orc::Q<T>{T(m[5]), T(m[6]), T(m[7]), T(m[8])}};
      c.v = orc::V3<T>{T(m[9]), T(m[10]), T(m[11])};
      c.w = orc::V3<T>{T(m[12]), T(m[13]), T(m[14])};
      return c;
    };
    const orc::ToiResult<T> r =
        orc::time_of_impact(shape(k1, d1), motion(m1), shape(k2, d2), motion(m2), T(max_toi), T(target));
    *toi = double(r.toi);
    *status = r.status;
    return r.found ? 1 : 0;
  };
  return precision == 32 ? run(0.0f) : run(0.0);
}

// The deterministic manifold colouring on its own (colours[m] for manifold m).
int orc_colour_manifolds(int32_t nman, const int32_t* b1, const int32_t* b2, int32_t nbodies,
                         const int32_t* is_dynamic, int32_t* colours) {
  std::vector<orc::Manifold> ms(nman);
  for (int m = 0; m < nman; ++m) ms[m] = orc::Manifold{b1[m], b2[m], 0, 1};
  std::vector<int> dyn(is_dynamic, is_dynamic + nbodies);
  std::vector<int> col(nman);
  for (int m = 0; m < nman; ++m) col[m] = colours[m];  // entries >= 0 are kept
  const int nc = orc::colour_manifolds(ms, dyn, nbodies, col);
  for (int m = 0; m < nman; ++m) colours[m] = col[m];
  return nc;
}

}  // extern "C"
