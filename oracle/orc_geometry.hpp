// oracle/orc_geometry.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Shape AABBs and contact-manifold generation for balls and cuboids.
//
// The reference delegates all of this to ncollide3d 0.31 (unvendored; call sites
// src/world/geometrical_world.rs:57-63,301-311,319 and src/solver/nonlinear_sor_prox.rs:221-229).
// This file restates it from published methods, and each rule is "parity unpinned" against
// ncollide itself:
//   * ball AABB = centre -/+ r; cuboid AABB = centre -/+ |R| h (absolute rotation times half
//     extents).
//   * ball-ball: the closed form (normal along the centre line, strict test
//     |c2-c1|^2 < (r1+r2+prediction)^2).
//   * ball-cuboid: closest point on the box in its local frame (Ericson, Real-Time Collision
//     Detection, 5.1.3); a centre inside the box is pushed out through the face of least
//     penetration.
//   * cuboid-cuboid: separating-axis test over the 15 axes, then reference-face clipping of the
//     incident face (Sutherland-Hodgman), as described by Ericson (ch. 4/5) and Gregorius (GDC
//     2013 "The Separating Axis Test"); manifolds with more than 4 clipped points are reduced to 4
//     (deepest, farthest, largest triangle, largest added area).
// Normals point from collider 1 to collider 2; depth is positive when penetrating (the
// TrackedContact convention used by src/solver/signorini_model.rs:65-101).
#pragma once
#include <cstdint>

#include "orc_math.hpp"

namespace orc {

enum ShapeKind : int { SHAPE_BALL = 0, SHAPE_CUBOID = 1 };

// Contact kinematic kinds used by the position solver's re-evaluation.
enum KinKind : int {
  KIN_BALL_BALL = 0,
  KIN_BALL_BOX = 1,   // collider 1 is the ball
  KIN_BOX_BALL = 2,   // collider 2 is the ball
  KIN_PLANE1 = 3,     // plane attached to body 1, point attached to body 2
  KIN_PLANE2 = 4,     // plane attached to body 2, point attached to body 1
};

template <class T>
struct Contact {
  V3<T> world1, world2, normal;
  T depth;
  uint32_t feature;
  int kind;
  V3<T> lp1, lp2, ln;  // local point on body 1 / body 2; local normal (frame of the plane body)
};

template <class T>
struct ShapeDesc {
  int kind;
  V3<T> he;  // half extents (cuboid) or (r, r, r) (ball)
};

template <class T>
AABB<T> shape_aabb(const ShapeDesc<T>& s, const Iso<T>& p) {
  if (s.kind == SHAPE_BALL) {
    const V3<T> r{s.he.x, s.he.x, s.he.x};
    return AABB<T>{p.t - r, p.t + r};
  }
  const M3<T> R = qmat(p.q);
  M3<T> A;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A.m[i][j] = std::fabs(R.m[i][j]);
  const V3<T> w = mulv(A, s.he);
  return AABB<T>{p.t - w, p.t + w};
}

// ---------------------------------------------------------------- ball / ball
template <class T>
int collide_ball_ball(V3<T> c1, T r1, V3<T> c2, T r2, T prediction, bool always,
                      Contact<T>* out) {
  const V3<T> d = c2 - c1;
  const T dist2 = dot(d, d);
  const T sumr = r1 + r2;
  const T lim = sumr + prediction;
  if (!always && !(dist2 < lim * lim)) return 0;
  V3<T> n;
  T dist;
  if (dist2 != T(0)) {
    dist = std::sqrt(dist2);
    n = V3<T>{d.x / dist, d.y / dist, d.z / dist};
  } else {
    dist = T(0);
    n = V3<T>{T(1), T(0), T(0)};
  }
  Contact<T>& c = out[0];
  c.normal = n;
  c.world1 = c1 + n * r1;
  c.world2 = c2 + n * (-r2);
  c.depth = sumr - dist;
  c.feature = 0;
  c.kind = KIN_BALL_BALL;
  c.lp1 = V3<T>{T(0), T(0), T(0)};
  c.lp2 = V3<T>{T(0), T(0), T(0)};
  c.ln = V3<T>{T(0), T(0), T(0)};
  return 1;
}

// ---------------------------------------------------------------- ball / cuboid
// Box pose (bt, bq), half extents h; ball centre c, radius r. ball_first: collider 1 is the ball.
template <class T>
int collide_ball_box(V3<T> c, T r, const Iso<T>& box, V3<T> h, bool ball_first, T prediction,
                     bool always, Contact<T>* out) {
  const M3<T> R = qmat(box.q);
  const V3<T> dc = c - box.t;
  const V3<T> p = mulv_t(R, dc);  // ball centre in the box frame
  V3<T> q{tmax(-h.x, tmin(p.x, h.x)), tmax(-h.y, tmin(p.y, h.y)), tmax(-h.z, tmin(p.z, h.z))};
  const bool inside = (p.x == q.x) && (p.y == q.y) && (p.z == q.z);
  V3<T> nl;
  T depth;
  bool use_outside = false;
  if (!inside) {
    const V3<T> dl = p - q;
    const T dist2 = dot(dl, dl);
    if (dist2 > T(0)) {
      const T lim = r + prediction;
      if (!always && !(dist2 < lim * lim)) return 0;
      const T dist = std::sqrt(dist2);
      nl = V3<T>{dl.x / dist, dl.y / dist, dl.z / dist};
      depth = r - dist;
      use_outside = true;
    }
  }
  if (!use_outside) {
    // Face of least penetration (first minimum in x, y, z order).
    const T px = h.x - std::fabs(p.x), py = h.y - std::fabs(p.y), pz = h.z - std::fabs(p.z);
    int ax = 0;
    T best = px;
    if (py < best) { best = py; ax = 1; }
    if (pz < best) { best = pz; ax = 2; }
    const T s = comp(p, ax) >= T(0) ? T(1) : T(-1);
    nl = V3<T>{T(0), T(0), T(0)};
    set_comp(nl, ax, s);
    depth = r + best;
    q = p;
    set_comp(q, ax, s * comp(h, ax));
  }
  const V3<T> nw = mulv(R, nl);        // box -> ball, world frame
  const V3<T> boxw = mulv(R, q) + box.t;
  Contact<T>& ct = out[0];
  ct.feature = 0;
  ct.depth = depth;
  ct.lp1 = V3<T>{T(0), T(0), T(0)};
  ct.lp2 = V3<T>{T(0), T(0), T(0)};
  ct.ln = V3<T>{T(0), T(0), T(0)};
  if (ball_first) {
    const V3<T> n = -nw;  // ball (1) -> box (2)
    ct.normal = n;
    ct.world1 = c + n * r;
    ct.world2 = boxw;
    ct.kind = KIN_BALL_BOX;
  } else {
    const V3<T> n = nw;  // box (1) -> ball (2)
    ct.normal = n;
    ct.world1 = boxw;
    ct.world2 = c + n * (-r);
    ct.kind = KIN_BOX_BALL;
  }
  return 1;
}

// ---------------------------------------------------------------- cuboid / cuboid
template <class T>
struct ClipVert {
  V3<T> p;
  int code;   // 0..3 incident vertex; 4..19 incident edge x side plane; 20..35 plane x plane
  int label;  // line the outgoing segment lies on: 0..3 incident edge, 4+p side plane p
};

// Clip a convex polygon against { x : dot(n, x) <= off }.
template <class T>
int clip_poly(const ClipVert<T>* in, int nin, V3<T> n, T off, int plane, ClipVert<T>* out) {
  int no = 0;
  if (nin == 0) return 0;
  for (int k = 0; k < nin; ++k) {
    const ClipVert<T>& prev = in[(k + nin - 1) % nin];
    const ClipVert<T>& cur = in[k];
    const T d0 = dot(n, prev.p) - off;
    const T d1 = dot(n, cur.p) - off;
    const bool in0 = d0 <= T(0), in1 = d1 <= T(0);
    if (in0 && in1) {
      out[no++] = cur;
    } else if (in0 != in1) {
      const T tt = d0 / (d0 - d1);
      ClipVert<T> x;
      x.p = prev.p + (cur.p - prev.p) * tt;
      x.code = prev.label < 4 ? 4 + prev.label * 4 + plane : 20 + (prev.label - 4) * 4 + plane;
      if (in0) {
        // leaving: the intersection starts a segment along the clip plane
        x.label = 4 + plane;
        out[no++] = x;
      } else {
        // entering: the intersection continues the segment towards cur
        x.label = prev.label;
        out[no++] = x;
        out[no++] = cur;
      }
    }
  }
  return no;
}

template <class T>
struct BoxFrame {
  V3<T> t;
  M3<T> R;
  V3<T> h;
  V3<T> axis(int i) const { return col(R, i); }
};

template <class T>
T box_tolerance_rel() { return T(0.95); }
template <class T>
T box_tolerance_abs() { return T(0.001); }

// Returns the number of contacts (0..4) written to out.
template <class T>
int collide_box_box(const Iso<T>& pa, V3<T> ha, const Iso<T>& pb, V3<T> hb, T prediction,
                     Contact<T>* out) {
  BoxFrame<T> A{pa.t, qmat(pa.q), ha};
  BoxFrame<T> B{pb.t, qmat(pb.q), hb};
  const V3<T> d = B.t - A.t;
  V3<T> a[3], b[3];
  for (int i = 0; i < 3; ++i) { a[i] = A.axis(i); b[i] = B.axis(i); }
  T C[3][3], AC[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      C[i][j] = dot(a[i], b[j]);
      AC[i][j] = std::fabs(C[i][j]);
    }
  // Face axes of A.
  T sepA = -std::numeric_limits<T>::infinity();
  int iA = 0;
  for (int i = 0; i < 3; ++i) {
    const T rb = (comp(hb, 0) * AC[i][0] + comp(hb, 1) * AC[i][1]) + comp(hb, 2) * AC[i][2];
    const T s = std::fabs(dot(d, a[i])) - (comp(ha, i) + rb);
    if (s > prediction) return 0;
    if (s > sepA) { sepA = s; iA = i; }
  }
  // Face axes of B.
  T sepB = -std::numeric_limits<T>::infinity();
  int iB = 0;
  for (int j = 0; j < 3; ++j) {
    const T ra = (comp(ha, 0) * AC[0][j] + comp(ha, 1) * AC[1][j]) + comp(ha, 2) * AC[2][j];
    const T s = std::fabs(dot(d, b[j])) - (ra + comp(hb, j));
    if (s > prediction) return 0;
    if (s > sepB) { sepB = s; iB = j; }
  }
  // Edge axes.
  T sepE = -std::numeric_limits<T>::infinity();
  int eI = -1, eJ = -1;
  V3<T> eAxis{T(0), T(0), T(0)};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const V3<T> L = cross(a[i], b[j]);
      const T l2 = dot(L, L);
      if (l2 < T(1e-6)) continue;
      const T l = std::sqrt(l2);
      const V3<T> Ln{L.x / l, L.y / l, L.z / l};
      const T ra = (comp(ha, 0) * std::fabs(dot(a[0], Ln)) + comp(ha, 1) * std::fabs(dot(a[1], Ln))) +
                   comp(ha, 2) * std::fabs(dot(a[2], Ln));
      const T rb = (comp(hb, 0) * std::fabs(dot(b[0], Ln)) + comp(hb, 1) * std::fabs(dot(b[1], Ln))) +
                   comp(hb, 2) * std::fabs(dot(b[2], Ln));
      const T s = std::fabs(dot(d, Ln)) - (ra + rb);
      if (s > prediction) return 0;
      if (s > sepE) { sepE = s; eI = i; eJ = j; eAxis = Ln; }
    }

  const T krel = box_tolerance_rel<T>(), kabs = box_tolerance_abs<T>();
  bool refIsB = false;
  T faceSep = sepA;
  if (sepB > krel * sepA + kabs) { refIsB = true; faceSep = sepB; }

  if (eI >= 0 && sepE > krel * faceSep + kabs) {
    // ---- edge / edge
    V3<T> Ln = eAxis;
    if (dot(d, Ln) < T(0)) Ln = -Ln;
    V3<T> pA = A.t;
    for (int k = 0; k < 3; ++k) {
      if (k == eI) continue;
      const T sgn = dot(a[k], Ln) > T(0) ? T(1) : T(-1);
      pA = pA + a[k] * (sgn * comp(ha, k));
    }
    V3<T> pB = B.t;
    for (int k = 0; k < 3; ++k) {
      if (k == eJ) continue;
      const T sgn = dot(b[k], Ln) > T(0) ? T(-1) : T(1);
      pB = pB + b[k] * (sgn * comp(hb, k));
    }
    const V3<T> da = a[eI], db = b[eJ];
    const V3<T> r = pA - pB;
    const T bb = dot(da, db), c = dot(da, r), f = dot(db, r);
    const T denom = T(1) - bb * bb;
    const T hAi = comp(ha, eI), hBj = comp(hb, eJ);
    T s = T(0);
    if (denom > T(1e-6)) s = tmax(-hAi, tmin((bb * f - c) / denom, hAi));
    T t = bb * s + f;
    if (t < -hBj) {
      t = -hBj;
      s = tmax(-hAi, tmin(t * bb - c, hAi));
    } else if (t > hBj) {
      t = hBj;
      s = tmax(-hAi, tmin(t * bb - c, hAi));
    }
    const V3<T> x1 = pA + da * s;
    const V3<T> x2 = pB + db * t;
    Contact<T>& ct = out[0];
    ct.normal = Ln;
    ct.world1 = x1;
    ct.world2 = x2;
    ct.depth = dot(x1 - x2, Ln);
    ct.feature = 0x8000u | uint32_t(eI * 3 + eJ);
    ct.kind = KIN_PLANE1;
    ct.lp1 = mulv_t(A.R, x1 - A.t);
    ct.lp2 = mulv_t(B.R, x2 - B.t);
    ct.ln = mulv_t(A.R, Ln);
    return 1;
  }

  // ---- face contact: reference box Rf, incident box In
  const BoxFrame<T>& Rf = refIsB ? B : A;
  const BoxFrame<T>& In = refIsB ? A : B;
  const int ir = refIsB ? iB : iA;
  const V3<T> dRI = In.t - Rf.t;
  const V3<T> axR = Rf.axis(ir);
  const T sR = dot(dRI, axR) >= T(0) ? T(1) : T(-1);
  const V3<T> nR = axR * sR;  // reference face normal, pointing towards the incident box
  // Incident face: most anti-parallel to nR.
  int ji = 0;
  T bestc = T(-1);
  T cj = T(0);
  for (int j = 0; j < 3; ++j) {
    const T cc = dot(nR, In.axis(j));
    if (std::fabs(cc) > bestc) { bestc = std::fabs(cc); ji = j; cj = cc; }
  }
  const T sI = cj > T(0) ? T(-1) : T(1);
  const V3<T> nI = In.axis(ji) * sI;
  const V3<T> fc = In.t + nI * comp(In.h, ji);
  const int j1 = (ji + 1) % 3, j2 = (ji + 2) % 3;
  const V3<T> u = In.axis(j1) * comp(In.h, j1);
  const V3<T> v = In.axis(j2) * comp(In.h, j2);
  ClipVert<T> poly0[8], poly1[8];
  poly0[0] = ClipVert<T>{(fc + u) + v, 0, 0};
  poly0[1] = ClipVert<T>{(fc - u) + v, 1, 1};
  poly0[2] = ClipVert<T>{(fc - u) - v, 2, 2};
  poly0[3] = ClipVert<T>{(fc + u) - v, 3, 3};
  int np = 4;
  const int r1 = (ir + 1) % 3, r2 = (ir + 2) % 3;
  const V3<T> u1 = Rf.axis(r1), u2 = Rf.axis(r2);
  const T o1 = dot(u1, Rf.t), o2 = dot(u2, Rf.t);
  const T e1 = comp(Rf.h, r1), e2 = comp(Rf.h, r2);
  np = clip_poly(poly0, np, u1, o1 + e1, 0, poly1);
  np = clip_poly(poly1, np, -u1, -o1 + e1, 1, poly0);
  np = clip_poly(poly0, np, u2, o2 + e2, 2, poly1);
  np = clip_poly(poly1, np, -u2, -o2 + e2, 3, poly0);
  // Separation of each clipped point below the reference face.
  const T offR = dot(nR, Rf.t) + comp(Rf.h, ir);
  ClipVert<T> cand[8];
  T csep[8];
  int nc = 0;
  for (int k = 0; k < np; ++k) {
    const T s = dot(nR, poly0[k].p) - offR;
    if (s <= prediction) {
      cand[nc] = poly0[k];
      csep[nc] = s;
      ++nc;
    }
  }
  if (nc == 0) return 0;
  int keep[8];
  int nk = 0;
  if (nc <= 4) {
    for (int k = 0; k < nc; ++k) keep[nk++] = k;
  } else {
    int i0 = 0;
    for (int k = 1; k < nc; ++k)
      if (csep[k] < csep[i0]) i0 = k;
    int i1 = -1;
    T bd = T(-1);
    for (int k = 0; k < nc; ++k) {
      if (k == i0) continue;
      const T dd = norm2(cand[k].p - cand[i0].p);
      if (dd > bd) { bd = dd; i1 = k; }
    }
    int i2 = -1;
    T ba = T(-1);
    for (int k = 0; k < nc; ++k) {
      if (k == i0 || k == i1) continue;
      const T ar = std::fabs(dot(cross(cand[i1].p - cand[i0].p, cand[k].p - cand[i0].p), nR));
      if (ar > ba) { ba = ar; i2 = k; }
    }
    int i3 = -1;
    T bs = T(-1);
    for (int k = 0; k < nc; ++k) {
      if (k == i0 || k == i1 || k == i2) continue;
      const V3<T> pk = cand[k].p;
      const T a01 = std::fabs(dot(cross(cand[i0].p - pk, cand[i1].p - pk), nR));
      const T a12 = std::fabs(dot(cross(cand[i1].p - pk, cand[i2].p - pk), nR));
      const T a20 = std::fabs(dot(cross(cand[i2].p - pk, cand[i0].p - pk), nR));
      const T sum = (a01 + a12) + a20;
      if (sum > bs) { bs = sum; i3 = k; }
    }
    for (int k = 0; k < nc; ++k)
      if (k == i0 || k == i1 || k == i2 || k == i3) keep[nk++] = k;
  }
  const uint32_t faceR = uint32_t(ir * 2 + (sR < T(0) ? 1 : 0));
  const uint32_t faceI = uint32_t(ji * 2 + (sI < T(0) ? 1 : 0));
  const uint32_t fbase = 0x4000u | (refIsB ? 0x2000u : 0u) | (faceR << 10) | (faceI << 7);
  for (int m = 0; m < nk; ++m) {
    const ClipVert<T>& cv = cand[keep[m]];
    const T s = csep[keep[m]];
    const V3<T> onRef = cv.p - nR * s;
    Contact<T>& ct = out[m];
    ct.depth = -s;
    ct.feature = fbase | uint32_t(cv.code);
    if (!refIsB) {
      // reference face on body 1 (A), incident point on body 2 (B)
      ct.normal = nR;
      ct.world1 = onRef;
      ct.world2 = cv.p;
      ct.kind = KIN_PLANE1;
      ct.lp1 = mulv_t(A.R, onRef - A.t);
      ct.lp2 = mulv_t(B.R, cv.p - B.t);
      ct.ln = mulv_t(A.R, nR);
    } else {
      // reference face on body 2 (B), incident point on body 1 (A); normal points A -> B
      const V3<T> n = -nR;
      ct.normal = n;
      ct.world1 = cv.p;
      ct.world2 = onRef;
      ct.kind = KIN_PLANE2;
      ct.lp1 = mulv_t(A.R, cv.p - A.t);
      ct.lp2 = mulv_t(B.R, onRef - B.t);
      ct.ln = mulv_t(B.R, n);
    }
  }
  return nk;
}

// Re-evaluates a contact at the current poses (the position solver's ContactKinematic::contact,
// src/solver/nonlinear_sor_prox.rs:221-229). Returns the geometric contact without margins.
template <class T>
Contact<T> reevaluate(const Contact<T>& k, const ShapeDesc<T>& s1, const Iso<T>& p1,
                      const ShapeDesc<T>& s2, const Iso<T>& p2) {
  Contact<T> c = k;
  switch (k.kind) {
    case KIN_BALL_BALL:
      collide_ball_ball(p1.t, s1.he.x, p2.t, s2.he.x, T(0), true, &c);
      break;
    case KIN_BALL_BOX:
      collide_ball_box(p1.t, s1.he.x, p2, s2.he, true, T(0), true, &c);
      break;
    case KIN_BOX_BALL:
      collide_ball_box(p2.t, s2.he.x, p1, s1.he, false, T(0), true, &c);
      break;
    case KIN_PLANE1: {
      const V3<T> n = qrot(p1.q, k.ln);
      const V3<T> a = iso_apply(p1, k.lp1);
      const V3<T> b = iso_apply(p2, k.lp2);
      const T depth = dot(a - b, n);
      c.normal = n;
      c.depth = depth;
      c.world1 = b + n * depth;
      c.world2 = b;
      break;
    }
    case KIN_PLANE2: {
      const V3<T> n = qrot(p2.q, k.ln);
      const V3<T> a = iso_apply(p1, k.lp1);
      const V3<T> b = iso_apply(p2, k.lp2);
      const T depth = dot(a - b, n);
      c.normal = n;
      c.depth = depth;
      c.world1 = a;
      c.world2 = a - n * depth;
      break;
    }
    default:
      break;
  }
  return c;
}

// Dispatch on the shape pair; collider 1 / collider 2 in pair order.
template <class T>
int collide_pair(const ShapeDesc<T>& s1, const Iso<T>& p1, const ShapeDesc<T>& s2,
                 const Iso<T>& p2, T prediction, Contact<T>* out) {
  if (s1.kind == SHAPE_BALL && s2.kind == SHAPE_BALL)
    return collide_ball_ball(p1.t, s1.he.x, p2.t, s2.he.x, prediction, false, out);
  if (s1.kind == SHAPE_BALL)
    return collide_ball_box(p1.t, s1.he.x, p2, s2.he, true, prediction, false, out);
  if (s2.kind == SHAPE_BALL)
    return collide_ball_box(p2.t, s2.he.x, p1, s1.he, false, prediction, false, out);
  return collide_box_box(p1, s1.he, p2, s2.he, prediction, out);
}

// ---------------------------------------------------------------- time of impact
// The reference takes times of impact from ncollide 0.31's query::time_of_impact (static or
// linear motions) and query::nonlinear_time_of_impact (rigid motions with rotation), called at
// src/world/mechanical_world.rs:975-1002; both are unvendored. Restated here by one
// conservative-advancement routine (Mirtich, "Impulse-based dynamic simulation of rigid body
// systems", 1996, ch. 2): at the current time the separation d along a direction n is measured;
// no point of either shape approaches the other along n faster than
//   a = -(v2 - v1) . n + |w1| rho1 + |w2| rho2
// (rho: the largest distance from the rotation centre to a point of the shape; 0 for a ball,
// whose shape a rotation about its centre leaves unchanged), so the separation stays above the
// target for a further (d - target) / a. Parity with ncollide's iterates is unpinned.

// Separation of two shapes at the given poses, and the unit direction `n` from shape 1 to shape 2
// it is measured along. Exact distance for ball/ball and ball/cuboid (negative when they overlap);
// for cuboid/cuboid the largest separation over the 15 separating axes of collide_box_box (the
// same expressions), which equals the distance when the closest features are a face or two edges
// and is below it otherwise, so the advancement stays conservative.
template <class T>
T toi_separation(const ShapeDesc<T>& s1, const Iso<T>& p1, const ShapeDesc<T>& s2, const Iso<T>& p2, V3<T>& n) {
  if (s1.kind == SHAPE_BALL && s2.kind == SHAPE_BALL) {
    const V3<T> d = p2.t - p1.t;
    const T dist2 = dot(d, d);
    T dist = T(0);
    n = V3<T>{T(1), T(0), T(0)};
    if (dist2 > T(0)) {
      dist = std::sqrt(dist2);
      n = V3<T>{d.x / dist, d.y / dist, d.z / dist};
    }
    return dist - (s1.he.x + s2.he.x);
  }
  if (s1.kind == SHAPE_BALL || s2.kind == SHAPE_BALL) {
    const bool ball_first = s1.kind == SHAPE_BALL;
    const V3<T> c = ball_first ? p1.t : p2.t;
    const T r = ball_first ? s1.he.x : s2.he.x;
    const Iso<T>& box = ball_first ? p2 : p1;
    const V3<T> h = ball_first ? s2.he : s1.he;
    const M3<T> R = qmat(box.q);
    const V3<T> p = mulv_t(R, c - box.t);
    const V3<T> q{tmax(-h.x, tmin(p.x, h.x)), tmax(-h.y, tmin(p.y, h.y)), tmax(-h.z, tmin(p.z, h.z))};
    V3<T> nl{T(0), T(0), T(0)};
    T sep = T(0);
    bool outside = false;
    if (!((p.x == q.x) && (p.y == q.y) && (p.z == q.z))) {
      const V3<T> dl = p - q;
      const T dist2 = dot(dl, dl);
      if (dist2 > T(0)) {
        const T dist = std::sqrt(dist2);
        nl = V3<T>{dl.x / dist, dl.y / dist, dl.z / dist};
        sep = dist - r;
        outside = true;
      }
    }
    if (!outside) {  // centre inside the box: the face of least penetration
      const T px = h.x - std::fabs(p.x), py = h.y - std::fabs(p.y), pz = h.z - std::fabs(p.z);
      int ax = 0;
      T best = px;
      if (py < best) { best = py; ax = 1; }
      if (pz < best) { best = pz; ax = 2; }
      set_comp(nl, ax, comp(p, ax) >= T(0) ? T(1) : T(-1));
      sep = -(r + best);
    }
    const V3<T> nw = mulv(R, nl);  // box -> ball
    n = ball_first ? -nw : nw;
    return sep;
  }
  const M3<T> RA = qmat(p1.q), RB = qmat(p2.q);
  const V3<T> d = p2.t - p1.t;
  const V3<T> ha = s1.he, hb = s2.he;
  V3<T> a[3], b[3];
  for (int i = 0; i < 3; ++i) {
    a[i] = col(RA, i);
    b[i] = col(RB, i);
  }
  T AC[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) AC[i][j] = std::fabs(dot(a[i], b[j]));
  T best = -std::numeric_limits<T>::infinity();
  V3<T> axis = a[0];
  T scale = T(1);  // the best axis is axis / scale (edge axes are compared unnormalised)
  for (int i = 0; i < 3; ++i) {
    const T rb = (comp(hb, 0) * AC[i][0] + comp(hb, 1) * AC[i][1]) + comp(hb, 2) * AC[i][2];
    const T s = std::fabs(dot(d, a[i])) - (comp(ha, i) + rb);
    if (s > best) { best = s; axis = a[i]; }
  }
  for (int j = 0; j < 3; ++j) {
    const T ra = (comp(ha, 0) * AC[0][j] + comp(ha, 1) * AC[1][j]) + comp(ha, 2) * AC[2][j];
    const T s = std::fabs(dot(d, b[j])) - (ra + comp(hb, j));
    if (s > best) { best = s; axis = b[j]; }
  }
  // edge axes L = a_i x b_j: the separation along L / |L| is (|d.L| - (ra + rb)) / |L| with ra, rb
  // the projected half extents on the unnormalised L (one division per axis)
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const V3<T> L = cross(a[i], b[j]);
      const T l2 = dot(L, L);
      if (l2 < T(1e-6)) continue;
      const T l = std::sqrt(l2);
      const T ra = (comp(ha, 0) * std::fabs(dot(a[0], L)) + comp(ha, 1) * std::fabs(dot(a[1], L))) +
                   comp(ha, 2) * std::fabs(dot(a[2], L));
      const T rb = (comp(hb, 0) * std::fabs(dot(b[0], L)) + comp(hb, 1) * std::fabs(dot(b[1], L))) +
                   comp(hb, 2) * std::fabs(dot(b[2], L));
      const T s = (std::fabs(dot(d, L)) - (ra + rb)) / l;
      if (s > best) { best = s; axis = L; scale = l; }
    }
  if (scale != T(1)) axis = V3<T>{axis.x / scale, axis.y / scale, axis.z / scale};
  n = dot(d, axis) < T(0) ? -axis : axis;
  return best;
}

// Rigid motion of one body over a CCD interval (ncollide ConstantVelocityRigidMotion /
// ConstantLinearVelocityRigidMotion / BodyPartMotion::Static, src/object/body.rs:15-46): the pose
// at time t is p0 moved for t - t0 at linear velocity v and, for MOTION_ROTATING, angular velocity
// w about the centre (the displacement composition of RigidBody::apply_displacement).
enum MotionKind : int { MOTION_STATIC = 0, MOTION_LINEAR = 1, MOTION_ROTATING = 2 };

template <class T>
Iso<T> iso_displaced(const Iso<T>& p, V3<T> lin, V3<T> ang) {  // T(c) * Iso(lin, exp(ang)) * T(-c) * p, c = p.t
  const Q<T> qd = qexp_scaled_axis(ang);
  const V3<T> com = p.t;
  const V3<T> dt_ = com + (lin + qrot(qd, -com));
  Iso<T> np;
  np.t = dt_ + qrot(qd, p.t);
  np.q = qmul(qd, p.q);
  return np;
}

template <class T>
struct CcdMotion {
  int kind;
  Iso<T> p0;
  V3<T> v, w;
  T t0;
  Iso<T> at(T t) const {
    if (kind == MOTION_STATIC) return p0;
    const T dt = t - t0;
    if (kind == MOTION_LINEAR) return Iso<T>{p0.t + v * dt, p0.q};
    return iso_displaced(p0, v * dt, w * dt);
  }
};

enum ToiStatus : int { TOI_CONVERGED = 0, TOI_OUT_OF_ITERATIONS = 1, TOI_PENETRATING = 2 };
constexpr int kToiMaxIterations = 32;
template <class T>
T toi_tolerance() { return T(1e-4); }

template <class T>
struct ToiResult {
  bool found;
  T toi;
  int status;
};

// rho of a shape: 0 for a ball, |h| for a cuboid
template <class T>
T rotation_reach(const ShapeDesc<T>& s) {
  return s.kind == SHAPE_BALL ? T(0) : std::sqrt(dot(s.he, s.he));
}

// First time in [0, max_toi] at which the separation of the two moving shapes falls to `target`
// (+ the tolerance), by conservative advancement; every iterate is a lower bound of the first time
// the distance reaches the target. Penetrating: the shapes overlap at time 0.
template <class T>
ToiResult<T> time_of_impact(const ShapeDesc<T>& s1, const CcdMotion<T>& m1, const ShapeDesc<T>& s2,
                            const CcdMotion<T>& m2, T max_toi, T target) {
  const V3<T> z{T(0), T(0), T(0)};
  const V3<T> v1 = m1.kind == MOTION_STATIC ? z : m1.v, v2 = m2.kind == MOTION_STATIC ? z : m2.v;
  const T spin = norm(m1.kind == MOTION_ROTATING ? m1.w : z) * rotation_reach(s1) +
                 norm(m2.kind == MOTION_ROTATING ? m2.w : z) * rotation_reach(s2);
  const V3<T> dv = v2 - v1;
  T tau = T(0);
  for (int it = 0; it < kToiMaxIterations; ++it) {
    V3<T> n;
    const T d = toi_separation(s1, m1.at(tau), s2, m2.at(tau), n);
    if (it == 0 && d < T(0)) return ToiResult<T>{true, T(0), TOI_PENETRATING};
    if (d <= target + toi_tolerance<T>()) return ToiResult<T>{true, tau, TOI_CONVERGED};
    const T a = spin - dot(dv, n);
    if (!(a > T(0))) return ToiResult<T>{false, T(0), TOI_CONVERGED};
    tau = tau + (d - target) / a;
    if (tau > max_toi) return ToiResult<T>{false, T(0), TOI_CONVERGED};
  }
  return ToiResult<T>{true, tau, TOI_OUT_OF_ITERATIONS};
}

}  // namespace orc
