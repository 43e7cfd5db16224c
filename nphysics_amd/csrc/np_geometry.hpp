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

}  // namespace nph
