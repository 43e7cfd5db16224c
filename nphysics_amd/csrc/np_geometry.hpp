// nphysics_amd/csrc/np_geometry.hpp -- contact manifolds for balls and cuboids (device + host).
//
// Replaces the ncollide3d narrow phase the reference dispatches from
// src/world/geometrical_world.rs:58-61,315-320 (ball/ball, ball/cuboid, cuboid/cuboid contact
// manifold generators) and the contact re-evaluation of the position solver
// (src/solver/nonlinear_sor_prox.rs:221-229, ContactKinematic::contact). The methods:
//   * ball/ball: normal along the centre line; kept when |c2-c1|^2 < (r1+r2+prediction)^2;
//   * ball/cuboid: closest point on the box in the box frame (Ericson, Real-Time Collision
//     Detection, 5.1.3); a centre inside the box leaves through the face of least penetration;
//   * cuboid/cuboid: separating-axis test over 3+3 face axes and 9 edge axes, a face axis of B
//     or an edge axis only when its separation beats the best face of A by the relative/absolute
//     tolerance (0.95, 0.001); face contacts clip the incident face against the reference face's
//     four side planes (Sutherland-Hodgman) and keep points within the prediction distance,
//     reduced to 4 (deepest, farthest, largest triangle, largest added area); edge contacts use
//     the closest points of the two edges.
// Normals point from collider 1 to collider 2; depth > 0 means penetration. Feature ids:
//   0 for ball contacts; face contacts 0x4000 | refIsB<<13 | refFace<<10 | incFace<<7 | code,
//   edge contacts 0x8000 | (3 i + j).
#pragma once
#include "devmath.hpp"

namespace nph {

enum : int {
  KIN_BALL_BALL = 0,
  KIN_BALL_BOX = 1,
  KIN_BOX_BALL = 2,
  KIN_PLANE1 = 3,
  KIN_PLANE2 = 4,
};

struct Contact {
  f3 w1, w2, n;
  float depth;
  uint32_t feature;
  int kind;
  f3 lp1, lp2, ln;
};

NPH_HD int collide_ball_ball(f3 c1, float r1, f3 c2, float r2, float prediction, bool always,
                             Contact* out) {
  const f3 d = sub(c2, c1);
  const float dist2 = dot(d, d);
  const float sumr = r1 + r2;
  const float lim = sumr + prediction;
  if (!always && !(dist2 < lim * lim)) return 0;
  f3 n;
  float dist;
  if (dist2 != 0.0f) {
    dist = sqrtf(dist2);
    n = f3{d.x / dist, d.y / dist, d.z / dist};
  } else {
    dist = 0.0f;
    n = f3{1.0f, 0.0f, 0.0f};
  }
  Contact& c = out[0];
  c.n = n;
  c.w1 = add(c1, scale(n, r1));
  c.w2 = add(c2, scale(n, -r2));
  c.depth = sumr - dist;
  c.feature = 0;
  c.kind = KIN_BALL_BALL;
  c.lp1 = f3{0.0f, 0.0f, 0.0f};
  c.lp2 = f3{0.0f, 0.0f, 0.0f};
  c.ln = f3{0.0f, 0.0f, 0.0f};
  return 1;
}

NPH_HD int collide_ball_box(f3 c, float r, f3 bt, q4 bq, f3 h, bool ball_first, float prediction,
                            bool always, Contact* out) {
  const m33 R = qmat(bq);
  const f3 dc = sub(c, bt);
  const f3 p = mulv_t(R, dc);
  f3 q{fmaxf_(-h.x, fminf_(p.x, h.x)), fmaxf_(-h.y, fminf_(p.y, h.y)), fmaxf_(-h.z, fminf_(p.z, h.z))};
  const bool inside = (p.x == q.x) && (p.y == q.y) && (p.z == q.z);
  f3 nl{0.0f, 0.0f, 0.0f};
  float depth = 0.0f;
  bool use_outside = false;
  if (!inside) {
    const f3 dl = sub(p, q);
    const float dist2 = dot(dl, dl);
    if (dist2 > 0.0f) {
      const float lim = r + prediction;
      if (!always && !(dist2 < lim * lim)) return 0;
      const float dist = sqrtf(dist2);
      nl = f3{dl.x / dist, dl.y / dist, dl.z / dist};
      depth = r - dist;
      use_outside = true;
    }
  }
  if (!use_outside) {
    const float px = h.x - fabsf(p.x), py = h.y - fabsf(p.y), pz = h.z - fabsf(p.z);
    int ax = 0;
    float best = px;
    if (py < best) { best = py; ax = 1; }
    if (pz < best) { best = pz; ax = 2; }
    const float s = comp(p, ax) >= 0.0f ? 1.0f : -1.0f;
    nl = f3{0.0f, 0.0f, 0.0f};
    set_comp(nl, ax, s);
    depth = r + best;
    q = p;
    set_comp(q, ax, s * comp(h, ax));
  }
  const f3 nw = mulv(R, nl);
  const f3 boxw = add(mulv(R, q), bt);
  Contact& ct = out[0];
  ct.feature = 0;
  ct.depth = depth;
  ct.lp1 = f3{0.0f, 0.0f, 0.0f};
  ct.lp2 = f3{0.0f, 0.0f, 0.0f};
  ct.ln = f3{0.0f, 0.0f, 0.0f};
  if (ball_first) {
    const f3 n = neg(nw);
    ct.n = n;
    ct.w1 = add(c, scale(n, r));
    ct.w2 = boxw;
    ct.kind = KIN_BALL_BOX;
  } else {
    ct.n = nw;
    ct.w1 = boxw;
    ct.w2 = add(c, scale(nw, -r));
    ct.kind = KIN_BOX_BALL;
  }
  return 1;
}

// A clip-polygon vertex: position and tag = code | label << 8 (16 bytes).
struct ClipVert {
  f3 p;
  int tag;
};
NPH_HD int vcode(const ClipVert& v) { return v.tag & 0xff; }
NPH_HD int vlabel(const ClipVert& v) { return v.tag >> 8; }

// Strided views of per-lane scratch arrays (element k of lane t at base[k * stride]), so that a
// workgroup's lanes can keep their clip polygons interleaved in LDS.
struct PolyBuf {
  ClipVert* base;
  int stride;
  NPH_HD ClipVert& operator[](int k) const { return base[long(k) * stride]; }
};
struct SepBuf {
  float* base;
  int stride;
  NPH_HD float& operator[](int k) const { return base[long(k) * stride]; }
};

NPH_HD int clip_poly(PolyBuf in, int nin, f3 n, float off, int plane, PolyBuf out) {
  int no = 0;
  if (nin == 0) return 0;
  for (int k = 0; k < nin; ++k) {
    const ClipVert prev = in[(k + nin - 1) % nin];
    const ClipVert cur = in[k];
    const float d0 = dot(n, prev.p) - off;
    const float d1 = dot(n, cur.p) - off;
    const bool in0 = d0 <= 0.0f, in1 = d1 <= 0.0f;
    if (in0 && in1) {
      out[no++] = cur;
    } else if (in0 != in1) {
      const float tt = d0 / (d0 - d1);
      ClipVert x;
      x.p = add(prev.p, scale(sub(cur.p, prev.p), tt));
      const int pl = vlabel(prev);
      const int code = pl < 4 ? 4 + pl * 4 + plane : 20 + (pl - 4) * 4 + plane;
      if (in0) {
        x.tag = code | ((4 + plane) << 8);
        out[no++] = x;
      } else {
        x.tag = code | (pl << 8);
        out[no++] = x;
        out[no++] = cur;
      }
    }
  }
  return no;
}

// Register-friendly element access (selects instead of a dynamically indexed local array).
NPH_HD f3 sel3(f3 v0, f3 v1, f3 v2, int i) { return i == 0 ? v0 : (i == 1 ? v1 : v2); }
NPH_HD f3 mcol_sel(const m33& a, int j) {
  return sel3(f3{a.m[0][0], a.m[1][0], a.m[2][0]}, f3{a.m[0][1], a.m[1][1], a.m[2][1]},
              f3{a.m[0][2], a.m[1][2], a.m[2][2]}, j);
}

// poly0 and poly1 hold 8 vertices each and csep 8 floats.
template <class Emit>
NPH_HD int collide_box_box(f3 ta, q4 qa, f3 ha, f3 tb, q4 qb, f3 hb, float prediction, PolyBuf poly0,
                           PolyBuf poly1, SepBuf csep, Emit&& emit) {
  const m33 RA = qmat(qa);
  const m33 RB = qmat(qb);
  const f3 d = sub(tb, ta);
  f3 a[3], b[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    a[i] = mcol(RA, i);
    b[i] = mcol(RB, i);
  }
  float AC[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) AC[i][j] = fabsf(dot(a[i], b[j]));
  float sepA = -INFINITY;
  int iA = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float rb = (hb.x * AC[i][0] + hb.y * AC[i][1]) + hb.z * AC[i][2];
    const float s = fabsf(dot(d, a[i])) - (comp(ha, i) + rb);
    if (s > prediction) return 0;
    if (s > sepA) { sepA = s; iA = i; }
  }
  float sepB = -INFINITY;
  int iB = 0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float ra = (ha.x * AC[0][j] + ha.y * AC[1][j]) + ha.z * AC[2][j];
    const float s = fabsf(dot(d, b[j])) - (ra + comp(hb, j));
    if (s > prediction) return 0;
    if (s > sepB) { sepB = s; iB = j; }
  }
  float sepE = -INFINITY;
  int eI = -1, eJ = -1;
  f3 eAxis{0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const f3 L = cross(a[i], b[j]);
      const float l2 = dot(L, L);
      if (l2 < 1e-6f) continue;
      const float l = sqrtf(l2);
      const f3 Ln{L.x / l, L.y / l, L.z / l};
      const float ra = (ha.x * fabsf(dot(a[0], Ln)) + ha.y * fabsf(dot(a[1], Ln))) + ha.z * fabsf(dot(a[2], Ln));
      const float rb = (hb.x * fabsf(dot(b[0], Ln)) + hb.y * fabsf(dot(b[1], Ln))) + hb.z * fabsf(dot(b[2], Ln));
      const float s = fabsf(dot(d, Ln)) - (ra + rb);
      if (s > prediction) return 0;
      if (s > sepE) { sepE = s; eI = i; eJ = j; eAxis = Ln; }
    }

  const float krel = 0.95f, kabs = 0.001f;
  bool refIsB = false;
  float faceSep = sepA;
  if (sepB > krel * sepA + kabs) { refIsB = true; faceSep = sepB; }

  if (eI >= 0 && sepE > krel * faceSep + kabs) {
    f3 Ln = eAxis;
    if (dot(d, Ln) < 0.0f) Ln = neg(Ln);
    f3 pA = ta;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (k == eI) continue;
      const float sgn = dot(a[k], Ln) > 0.0f ? 1.0f : -1.0f;
      pA = add(pA, scale(a[k], sgn * comp(ha, k)));
    }
    f3 pB = tb;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (k == eJ) continue;
      const float sgn = dot(b[k], Ln) > 0.0f ? -1.0f : 1.0f;
      pB = add(pB, scale(b[k], sgn * comp(hb, k)));
    }
    const f3 da = sel3(a[0], a[1], a[2], eI), db = sel3(b[0], b[1], b[2], eJ);
    const f3 r = sub(pA, pB);
    const float bb = dot(da, db), c = dot(da, r), f = dot(db, r);
    const float denom = 1.0f - bb * bb;
    const float hAi = comp(ha, eI), hBj = comp(hb, eJ);
    float s = 0.0f;
    if (denom > 1e-6f) s = fmaxf_(-hAi, fminf_((bb * f - c) / denom, hAi));
    float t = bb * s + f;
    if (t < -hBj) {
      t = -hBj;
      s = fmaxf_(-hAi, fminf_(t * bb - c, hAi));
    } else if (t > hBj) {
      t = hBj;
      s = fmaxf_(-hAi, fminf_(t * bb - c, hAi));
    }
    const f3 x1 = add(pA, scale(da, s));
    const f3 x2 = add(pB, scale(db, t));
    Contact ct;
    ct.n = Ln;
    ct.w1 = x1;
    ct.w2 = x2;
    ct.depth = dot(sub(x1, x2), Ln);
    ct.feature = 0x8000u | uint32_t(eI * 3 + eJ);
    ct.kind = KIN_PLANE1;
    ct.lp1 = mulv_t(RA, sub(x1, ta));
    ct.lp2 = mulv_t(RB, sub(x2, tb));
    ct.ln = mulv_t(RA, Ln);
    emit(0, ct);
    return 1;
  }

  const m33& RR = refIsB ? RB : RA;
  const m33& RI = refIsB ? RA : RB;
  const f3 tR = refIsB ? tb : ta;
  const f3 tI = refIsB ? ta : tb;
  const f3 hR = refIsB ? hb : ha;
  const f3 hI = refIsB ? ha : hb;
  const int ir = refIsB ? iB : iA;
  const f3 dRI = sub(tI, tR);
  const f3 axR = mcol_sel(RR, ir);
  const float sR = dot(dRI, axR) >= 0.0f ? 1.0f : -1.0f;
  const f3 nR = scale(axR, sR);
  int ji = 0;
  float bestc = -1.0f;
  float cj = 0.0f;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float cc = dot(nR, mcol(RI, j));
    if (fabsf(cc) > bestc) { bestc = fabsf(cc); ji = j; cj = cc; }
  }
  const float sI = cj > 0.0f ? -1.0f : 1.0f;
  const f3 nI = scale(mcol_sel(RI, ji), sI);
  const f3 fc = add(tI, scale(nI, comp(hI, ji)));
  const int j1 = (ji + 1) % 3, j2 = (ji + 2) % 3;
  const f3 u = scale(mcol_sel(RI, j1), comp(hI, j1));
  const f3 v = scale(mcol_sel(RI, j2), comp(hI, j2));
  poly0[0] = ClipVert{add(add(fc, u), v), 0 | (0 << 8)};
  poly0[1] = ClipVert{add(sub(fc, u), v), 1 | (1 << 8)};
  poly0[2] = ClipVert{sub(sub(fc, u), v), 2 | (2 << 8)};
  poly0[3] = ClipVert{sub(add(fc, u), v), 3 | (3 << 8)};
  int np = 4;
  const int r1 = (ir + 1) % 3, r2 = (ir + 2) % 3;
  const f3 u1 = mcol_sel(RR, r1), u2 = mcol_sel(RR, r2);
  const float o1 = dot(u1, tR), o2 = dot(u2, tR);
  const float e1 = comp(hR, r1), e2 = comp(hR, r2);
  np = clip_poly(poly0, np, u1, o1 + e1, 0, poly1);
  np = clip_poly(poly1, np, neg(u1), -o1 + e1, 1, poly0);
  np = clip_poly(poly0, np, u2, o2 + e2, 2, poly1);
  np = clip_poly(poly1, np, neg(u2), -o2 + e2, 3, poly0);
  const float offR = dot(nR, tR) + comp(hR, ir);
  // candidates: the clipped polygon's vertices within the prediction distance, compacted in place
  // in poly0 (in order) with their separations in csep
  int nc = 0;
  for (int k = 0; k < np; ++k) {
    const float s = dot(nR, poly0[k].p) - offR;
    if (s <= prediction) {
      poly0[nc] = poly0[k];
      csep[nc] = s;
      ++nc;
    }
  }
  const PolyBuf cand = poly0;
  if (nc == 0) return 0;
  // at most 4 kept (in candidate order): all of them, or the deepest one and three spreading the
  // manifold's area
  int i0 = -1, i1 = -1, i2 = -1, i3 = -1;
  if (nc > 4) {
    i0 = 0;
    for (int k = 1; k < nc; ++k)
      if (csep[k] < csep[i0]) i0 = k;
    float bd = -1.0f;
    for (int k = 0; k < nc; ++k) {
      if (k == i0) continue;
      const f3 dd3 = sub(cand[k].p, cand[i0].p);
      const float dd = dot(dd3, dd3);
      if (dd > bd) { bd = dd; i1 = k; }
    }
    float ba = -1.0f;
    for (int k = 0; k < nc; ++k) {
      if (k == i0 || k == i1) continue;
      const float ar = fabsf(dot(cross(sub(cand[i1].p, cand[i0].p), sub(cand[k].p, cand[i0].p)), nR));
      if (ar > ba) { ba = ar; i2 = k; }
    }
    float bs = -1.0f;
    for (int k = 0; k < nc; ++k) {
      if (k == i0 || k == i1 || k == i2) continue;
      const f3 pk = cand[k].p;
      const float a01 = fabsf(dot(cross(sub(cand[i0].p, pk), sub(cand[i1].p, pk)), nR));
      const float a12 = fabsf(dot(cross(sub(cand[i1].p, pk), sub(cand[i2].p, pk)), nR));
      const float a20 = fabsf(dot(cross(sub(cand[i2].p, pk), sub(cand[i0].p, pk)), nR));
      const float sum = (a01 + a12) + a20;
      if (sum > bs) { bs = sum; i3 = k; }
    }
  }
  const uint32_t faceR = uint32_t(ir * 2 + (sR < 0.0f ? 1 : 0));
  const uint32_t faceI = uint32_t(ji * 2 + (sI < 0.0f ? 1 : 0));
  const uint32_t fbase = 0x4000u | (refIsB ? 0x2000u : 0u) | (faceR << 10) | (faceI << 7);
  int m = 0;
  for (int k = 0; k < nc; ++k) {
    if (nc > 4 && k != i0 && k != i1 && k != i2 && k != i3) continue;
    const ClipVert cv = cand[k];
    const float s = csep[k];
    const f3 onRef = sub(cv.p, scale(nR, s));
    Contact ct;
    ct.depth = -s;
    ct.feature = fbase | uint32_t(vcode(cv));
    if (!refIsB) {
      ct.n = nR;
      ct.w1 = onRef;
      ct.w2 = cv.p;
      ct.kind = KIN_PLANE1;
      ct.lp1 = mulv_t(RA, sub(onRef, ta));
      ct.lp2 = mulv_t(RB, sub(cv.p, tb));
      ct.ln = mulv_t(RA, nR);
    } else {
      const f3 n = neg(nR);
      ct.n = n;
      ct.w1 = cv.p;
      ct.w2 = onRef;
      ct.kind = KIN_PLANE2;
      ct.lp1 = mulv_t(RA, sub(cv.p, ta));
      ct.lp2 = mulv_t(RB, sub(onRef, tb));
      ct.ln = mulv_t(RB, n);
    }
    emit(m, ct);
    ++m;
  }
  return m;
}

// Contact generation of a pair with at least one ball (at most one contact, emitted as emit(0, c)).
template <class Emit>
NPH_HD int collide_ball_pair(int k1, f3 h1, f3 t1, q4 q1, int k2, f3 h2, f3 t2, q4 q2, float prediction,
                             Emit&& emit) {
  Contact c;
  int n;
  if (k1 == SHAPE_BALL && k2 == SHAPE_BALL) n = collide_ball_ball(t1, h1.x, t2, h2.x, prediction, false, &c);
  else if (k1 == SHAPE_BALL) n = collide_ball_box(t1, h1.x, t2, q2, h2, true, prediction, false, &c);
  else n = collide_ball_box(t2, h2.x, t1, q1, h1, false, prediction, false, &c);
  if (n) emit(0, c);
  return n;
}

// Contact re-evaluation at the current poses (geometric, without margins).
NPH_HD Contact reevaluate(int kind, f3 lp1, f3 lp2, f3 ln, int k1, f3 h1, f3 t1, q4 q1, int k2,
                          f3 h2, f3 t2, q4 q2) {
  Contact c;
  c.kind = kind;
  c.lp1 = lp1;
  c.lp2 = lp2;
  c.ln = ln;
  c.feature = 0;
  switch (kind) {
    case KIN_BALL_BALL:
      collide_ball_ball(t1, h1.x, t2, h2.x, 0.0f, true, &c);
      break;
    case KIN_BALL_BOX:
      collide_ball_box(t1, h1.x, t2, q2, h2, true, 0.0f, true, &c);
      break;
    case KIN_BOX_BALL:
      collide_ball_box(t2, h2.x, t1, q1, h1, false, 0.0f, true, &c);
      break;
    case KIN_PLANE1: {
      const f3 n = qrot(q1, ln);
      const f3 a = add(qrot(q1, lp1), t1);
      const f3 b = add(qrot(q2, lp2), t2);
      const float depth = dot(sub(a, b), n);
      c.n = n;
      c.depth = depth;
      c.w1 = add(b, scale(n, depth));
      c.w2 = b;
      break;
    }
    default: {  // KIN_PLANE2
      const f3 n = qrot(q2, ln);
      const f3 a = add(qrot(q1, lp1), t1);
      const f3 b = add(qrot(q2, lp2), t2);
      const float depth = dot(sub(a, b), n);
      c.n = n;
      c.depth = depth;
      c.w1 = a;
      c.w2 = sub(a, scale(n, depth));
      break;
    }
  }
  return c;
}

// ---------------------------------------------------------------- time of impact
// Conservative advancement (Mirtich, "Impulse-based dynamic simulation of rigid body systems",
// 1996, ch. 2) standing in for ncollide's query::time_of_impact / nonlinear_time_of_impact, which
// the reference's CCD calls (src/world/mechanical_world.rs:975-1002): at time t the separation d of
// the two shapes along a direction n is measured; no point approaches the other shape along n
// faster than a = |w1| rho1 + |w2| rho2 - (v2 - v1) . n (rho: the largest distance from the
// rotation centre to a point of the shape, 0 for a ball), so the separation stays above the target
// for a further (d - target) / a. Every iterate is a lower bound of the first time the distance
// reaches the target.

// Separation of two shapes (kinds k1, k2, half extents h1, h2; balls (r, r, r)) at the given poses
// and the unit direction n from shape 1 to shape 2 it is measured along: the distance for ball/ball
// and ball/cuboid (negative when overlapping); for cuboid/cuboid the largest separation over the 15
// axes of collide_box_box, a lower bound of the distance, exact for face and edge-edge closest
// features.
NPH_HD float toi_separation(int k1, f3 h1, f3 t1, q4 q1, int k2, f3 h2, f3 t2, q4 q2, f3& n) {
  if (k1 == SHAPE_BALL && k2 == SHAPE_BALL) {
    const f3 d = sub(t2, t1);
    const float dist2 = dot(d, d);
    float dist = 0.0f;
    n = f3{1.0f, 0.0f, 0.0f};
    if (dist2 > 0.0f) {
      dist = sqrtf(dist2);
      n = f3{d.x / dist, d.y / dist, d.z / dist};
    }
    return dist - (h1.x + h2.x);
  }
  if (k1 == SHAPE_BALL || k2 == SHAPE_BALL) {
    const bool ball_first = k1 == SHAPE_BALL;
    const f3 c = ball_first ? t1 : t2;
    const float r = ball_first ? h1.x : h2.x;
    const f3 bt = ball_first ? t2 : t1;
    const q4 bq = ball_first ? q2 : q1;
    const f3 h = ball_first ? h2 : h1;
    const m33 R = qmat(bq);
    const f3 p = mulv_t(R, sub(c, bt));
    const f3 q{fmaxf_(-h.x, fminf_(p.x, h.x)), fmaxf_(-h.y, fminf_(p.y, h.y)), fmaxf_(-h.z, fminf_(p.z, h.z))};
    f3 nl{0.0f, 0.0f, 0.0f};
    float sep = 0.0f;
    bool outside = false;
    if (!((p.x == q.x) && (p.y == q.y) && (p.z == q.z))) {
      const f3 dl = sub(p, q);
      const float dist2 = dot(dl, dl);
      if (dist2 > 0.0f) {
        const float dist = sqrtf(dist2);
        nl = f3{dl.x / dist, dl.y / dist, dl.z / dist};
        sep = dist - r;
        outside = true;
      }
    }
    if (!outside) {
      const float px = h.x - fabsf(p.x), py = h.y - fabsf(p.y), pz = h.z - fabsf(p.z);
      int ax = 0;
      float best = px;
      if (py < best) { best = py; ax = 1; }
      if (pz < best) { best = pz; ax = 2; }
      set_comp(nl, ax, comp(p, ax) >= 0.0f ? 1.0f : -1.0f);
      sep = -(r + best);
    }
    const f3 nw = mulv(R, nl);
    n = ball_first ? neg(nw) : nw;
    return sep;
  }
  const m33 RA = qmat(q1), RB = qmat(q2);
  const f3 d = sub(t2, t1);
  f3 a[3], b[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    a[i] = mcol(RA, i);
    b[i] = mcol(RB, i);
  }
  float AC[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) AC[i][j] = fabsf(dot(a[i], b[j]));
  float best = -INFINITY;
  f3 axis = a[0];
  float scale = 1.0f;  // the best axis is axis / scale (edge axes are compared unnormalised)
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float rb = (h2.x * AC[i][0] + h2.y * AC[i][1]) + h2.z * AC[i][2];
    const float s = fabsf(dot(d, a[i])) - (comp(h1, i) + rb);
    if (s > best) { best = s; axis = a[i]; }
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float ra = (h1.x * AC[0][j] + h1.y * AC[1][j]) + h1.z * AC[2][j];
    const float s = fabsf(dot(d, b[j])) - (ra + comp(h2, j));
    if (s > best) { best = s; axis = b[j]; }
  }
  // edge axes L = a_i x b_j: (|d.L| - (ra + rb)) / |L| with ra, rb projected on the unnormalised L
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const f3 L = cross(a[i], b[j]);
      const float l2 = dot(L, L);
      if (l2 < 1e-6f) continue;
      const float l = sqrtf(l2);
      const float ra = (h1.x * fabsf(dot(a[0], L)) + h1.y * fabsf(dot(a[1], L))) + h1.z * fabsf(dot(a[2], L));
      const float rb = (h2.x * fabsf(dot(b[0], L)) + h2.y * fabsf(dot(b[1], L))) + h2.z * fabsf(dot(b[2], L));
      const float s = (fabsf(dot(d, L)) - (ra + rb)) / l;
      if (s > best) { best = s; axis = L; scale = l; }
    }
  if (scale != 1.0f) axis = f3{axis.x / scale, axis.y / scale, axis.z / scale};
  n = dot(d, axis) < 0.0f ? neg(axis) : axis;
  return best;
}

// Rigid motion over a CCD interval (BodyPartMotion, src/object/body.rs:15-46): the pose at time t
// is (t0p, q0) moved for t - t0 at velocity v and, when rotating, angular velocity w about the
// centre (apply_displacement's composition); linear motions keep the rotation; static ones the pose.
enum : int { MOTION_STATIC = 0, MOTION_LINEAR = 1, MOTION_ROTATING = 2 };
struct CcdMotion {
  int kind;
  f3 t0p;
  q4 q0;
  f3 v, w;
  float t0;
  NPH_HD void at(float t, f3& tp, q4& q) const {
    tp = t0p;
    q = q0;
    if (kind == MOTION_STATIC) return;
    const float dt = t - t0;
    if (kind == MOTION_LINEAR) {
      tp = add(t0p, scale(v, dt));
      return;
    }
    apply_displacement(tp, q, scale(v, dt), scale(w, dt));
  }
};

enum : int { TOI_CONVERGED = 0, TOI_OUT_OF_ITERATIONS = 1, TOI_PENETRATING = 2 };
constexpr int kToiMaxIterations = 32;
constexpr float kToiTolerance = 1e-4f;

NPH_HD float rotation_reach(int kind, f3 h) { return kind == SHAPE_BALL ? 0.0f : sqrtf(dot(h, h)); }

// First time in [0, max_toi] at which the separation falls to target (+ kToiTolerance); returns
// false when the shapes never get that close before max_toi. *status: TOI_*.
NPH_HD bool time_of_impact(int k1, f3 h1, const CcdMotion& m1, int k2, f3 h2, const CcdMotion& m2, float max_toi,
                           float target, float* toi, int* status) {
  const f3 z{0.0f, 0.0f, 0.0f};
  const f3 v1 = m1.kind == MOTION_STATIC ? z : m1.v, v2 = m2.kind == MOTION_STATIC ? z : m2.v;
  const f3 w1 = m1.kind == MOTION_ROTATING ? m1.w : z, w2 = m2.kind == MOTION_ROTATING ? m2.w : z;
  const float spin = sqrtf(dot(w1, w1)) * rotation_reach(k1, h1) + sqrtf(dot(w2, w2)) * rotation_reach(k2, h2);
  const f3 dv = sub(v2, v1);
  float tau = 0.0f;
  for (int it = 0; it < kToiMaxIterations; ++it) {
    f3 p1, p2, n;
    q4 r1, r2;
    m1.at(tau, p1, r1);
    m2.at(tau, p2, r2);
    const float d = toi_separation(k1, h1, p1, r1, k2, h2, p2, r2, n);
    if (it == 0 && d < 0.0f) {
      *toi = 0.0f;
      *status = TOI_PENETRATING;
      return true;
    }
    if (d <= target + kToiTolerance) {
      *toi = tau;
      *status = TOI_CONVERGED;
      return true;
    }
    const float a = spin - dot(dv, n);
    if (!(a > 0.0f)) return false;
    tau = tau + (d - target) / a;
    if (tau > max_toi) return false;
  }
  *toi = tau;
  *status = TOI_OUT_OF_ITERATIONS;
  return true;
}

}  // namespace nph
