// nphysics_amd/csrc/devmath.hpp -- FP32 vector, quaternion and 3x3 helpers for the kernels.
//
// Every helper fixes its evaluation order (sums left to right, products as written) and the
// library is compiled with -ffp-contract=off, so each result is the correctly rounded FP32 value
// of the expression as written: the same value on the GPU and in any IEEE FP32 implementation
// that evaluates the same expression. Formulas:
//   dot (a.x b.x + a.y b.y) + a.z b.z; cross as usual;
//   quaternion (i, j, k, w), Hamilton product, rotation of a vector v + w t + u x t with
//   t = 2 u x v (the nalgebra forms behind src/object/rigid_body.rs:371-381);
//   rotation matrix in homogeneous form; 3x3 inverse by cofactors, zero if singular
//   (src/algebra/inertia3.rs:77-86); exp of a pure quaternion with a Maclaurin polynomial for
//   half angles below 1/4 rad (src/algebra/velocity3.rs:57-59 -> Isometry3::new).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define NPH_HD __host__ __device__ __forceinline__

namespace nph {

struct f3 {
  float x, y, z;
};

NPH_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
NPH_HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
NPH_HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
NPH_HD f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
NPH_HD f3 scale(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
NPH_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
NPH_HD f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
NPH_HD float fminf_(float a, float b) { return a < b ? a : b; }
NPH_HD float fmaxf_(float a, float b) { return a > b ? a : b; }
NPH_HD f3 vmin(f3 a, f3 b) { return f3{fminf_(a.x, b.x), fminf_(a.y, b.y), fminf_(a.z, b.z)}; }
NPH_HD f3 vmax(f3 a, f3 b) { return f3{fmaxf_(a.x, b.x), fmaxf_(a.y, b.y), fmaxf_(a.z, b.z)}; }
NPH_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
NPH_HD void set_comp(f3& a, int i, float v) {
  if (i == 0) a.x = v;
  else if (i == 1) a.y = v;
  else a.z = v;
}
NPH_HD f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
NPH_HD float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

struct q4 {
  float i, j, k, w;
};
NPH_HD q4 toq(float4 v) { return q4{v.x, v.y, v.z, v.w}; }
NPH_HD float4 fromq(q4 q) { return make_float4(q.i, q.j, q.k, q.w); }

NPH_HD q4 qmul(q4 a, q4 b) {
  q4 r;
  r.w = ((a.w * b.w - a.i * b.i) - a.j * b.j) - a.k * b.k;
  r.i = ((a.w * b.i + a.i * b.w) + a.j * b.k) - a.k * b.j;
  r.j = ((a.w * b.j - a.i * b.k) + a.j * b.w) + a.k * b.i;
  r.k = ((a.w * b.k + a.i * b.j) - a.j * b.i) + a.k * b.w;
  return r;
}

NPH_HD f3 qrot(q4 q, f3 v) {
  const f3 u{q.i, q.j, q.k};
  const f3 t = scale(cross(u, v), 2.0f);
  const f3 c = cross(u, t);
  return add(add(scale(t, q.w), c), v);
}

// Row-major 3x3.
struct m33 {
  float m[3][3];
};

NPH_HD m33 qmat(q4 q) {
  const float ww = q.w * q.w, ii = q.i * q.i, jj = q.j * q.j, kk = q.k * q.k;
  const float ij = q.i * q.j * 2.0f, wk = q.w * q.k * 2.0f, wj = q.w * q.j * 2.0f;
  const float ik = q.i * q.k * 2.0f, jk = q.j * q.k * 2.0f, wi = q.w * q.i * 2.0f;
  m33 r;
  r.m[0][0] = ((ww + ii) - jj) - kk; r.m[0][1] = ij - wk;               r.m[0][2] = wj + ik;
  r.m[1][0] = wk + ij;               r.m[1][1] = ((ww - ii) + jj) - kk; r.m[1][2] = jk - wi;
  r.m[2][0] = ik - wj;               r.m[2][1] = wi + jk;               r.m[2][2] = ((ww - ii) - jj) + kk;
  return r;
}

NPH_HD m33 mmul(const m33& a, const m33& b) {
  m33 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
  return r;
}
NPH_HD m33 mtrans(const m33& a) {
  m33 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
  return r;
}
NPH_HD f3 mulv(const m33& a, f3 v) {
  return f3{(a.m[0][0] * v.x + a.m[0][1] * v.y) + a.m[0][2] * v.z,
            (a.m[1][0] * v.x + a.m[1][1] * v.y) + a.m[1][2] * v.z,
            (a.m[2][0] * v.x + a.m[2][1] * v.y) + a.m[2][2] * v.z};
}
NPH_HD f3 mulv_t(const m33& a, f3 v) {
  return f3{(a.m[0][0] * v.x + a.m[1][0] * v.y) + a.m[2][0] * v.z,
            (a.m[0][1] * v.x + a.m[1][1] * v.y) + a.m[2][1] * v.z,
            (a.m[0][2] * v.x + a.m[1][2] * v.y) + a.m[2][2] * v.z};
}
NPH_HD f3 mcol(const m33& a, int j) { return f3{a.m[0][j], a.m[1][j], a.m[2][j]}; }

NPH_HD m33 cross_matrix(f3 v) {
  m33 r;
  r.m[0][0] = 0.0f; r.m[0][1] = -v.z; r.m[0][2] = v.y;
  r.m[1][0] = v.z;  r.m[1][1] = 0.0f; r.m[1][2] = -v.x;
  r.m[2][0] = -v.y; r.m[2][1] = v.x;  r.m[2][2] = 0.0f;
  return r;
}

NPH_HD bool try_inverse(const m33& a, m33& out) {
  const float c00 = a.m[1][1] * a.m[2][2] - a.m[2][1] * a.m[1][2];
  const float c01 = a.m[1][0] * a.m[2][2] - a.m[2][0] * a.m[1][2];
  const float c02 = a.m[1][0] * a.m[2][1] - a.m[2][0] * a.m[1][1];
  const float det = (a.m[0][0] * c00 - a.m[0][1] * c01) + a.m[0][2] * c02;
  if (det == 0.0f) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) out.m[i][j] = 0.0f;
    return false;
  }
  out.m[0][0] = c00 / det;
  out.m[0][1] = (a.m[2][1] * a.m[0][2] - a.m[0][1] * a.m[2][2]) / det;
  out.m[0][2] = (a.m[0][1] * a.m[1][2] - a.m[1][1] * a.m[0][2]) / det;
  out.m[1][0] = -c01 / det;
  out.m[1][1] = (a.m[0][0] * a.m[2][2] - a.m[2][0] * a.m[0][2]) / det;
  out.m[1][2] = (a.m[1][0] * a.m[0][2] - a.m[0][0] * a.m[1][2]) / det;
  out.m[2][0] = c02 / det;
  out.m[2][1] = (a.m[2][0] * a.m[0][1] - a.m[0][0] * a.m[2][1]) / det;
  out.m[2][2] = (a.m[0][0] * a.m[1][1] - a.m[1][0] * a.m[0][1]) / det;
  return true;
}

// exp of the pure quaternion (0, axisangle / 2).
NPH_HD q4 qexp_scaled_axis(f3 axisangle) {
  const f3 h = scale(axisangle, 0.5f);
  const float nn = dot(h, h);
  const float e = 1.1920928955078125e-07f;  // FP32 machine epsilon
  if (nn <= e * e) return q4{0.0f, 0.0f, 0.0f, 1.0f};
  const float n = sqrtf(nn);
  // sin(n)/n and cos(n): Maclaurin polynomials below 0.25; larger angles halved exactly until
  // below 0.25 and rebuilt with the double-angle formulas (the checker's arithmetic, bit for bit)
  float x = n;
  int k = 0;
  while (x >= 0.25f && k < 64) {
    x = x * 0.5f;
    ++k;
  }
  const float x2 = x * x;
  float s = 1.0f + x2 * ((float)(-1.0 / 6.0) + x2 * ((float)(1.0 / 120.0) + x2 * ((float)(-1.0 / 5040.0) + x2 * (float)(1.0 / 362880.0))));
  float c = 1.0f + x2 * (-0.5f + x2 * ((float)(1.0 / 24.0) + x2 * ((float)(-1.0 / 720.0) + x2 * (float)(1.0 / 40320.0))));
  if (k > 0) {
    float sn = x * s;
    for (int i = 0; i < k; ++i) {
      const float s2 = 2.0f * sn * c;
      const float c2 = c * c - sn * sn;
      sn = s2;
      c = c2;
    }
    s = sn / n;
  }
  return q4{h.x * s, h.y * s, h.z * s, c};
}

// atan(x) for 0 <= x <= 1 by plain arithmetic (the CPU restatement evaluates the same
// expression tree): three half-angle reductions x -> x / (1 + sqrt(1 + x^2)), then the Maclaurin
// series of atan through x^19 in Horner form, times 8.
NPH_HD float atan_unit(float x) {
  for (int i = 0; i < 3; ++i) x = x / (1.0f + sqrtf(1.0f + x * x));
  const float x2 = x * x;
  float p = float(-1.0 / 19.0);
  p = float(1.0 / 17.0) + x2 * p;
  p = float(-1.0 / 15.0) + x2 * p;
  p = float(1.0 / 13.0) + x2 * p;
  p = float(-1.0 / 11.0) + x2 * p;
  p = float(1.0 / 9.0) + x2 * p;
  p = float(-1.0 / 7.0) + x2 * p;
  p = float(1.0 / 5.0) + x2 * p;
  p = float(-1.0 / 3.0) + x2 * p;
  p = 1.0f + x2 * p;
  return 8.0f * (x * p);
}

// atan2(y, x) for y >= 0, in [0, pi].
NPH_HD float angle_of(float y, float x) {
  const float pi = float(3.14159265358979323846);
  const float ax = fabsf(x);
  if (y == 0.0f && ax == 0.0f) return 0.0f;
  const float base = ax >= y ? atan_unit(y / ax) : pi * 0.5f - atan_unit(ax / y);
  return x >= 0.0f ? base : pi - base;
}

NPH_HD q4 qconj(q4 q) { return q4{-q.i, -q.j, -q.k, q.w}; }

// Scaled axis (axis times angle) of a unit quaternion: the imaginary part, negated when w < 0,
// normalised, times 2 atan2(|imag|, |w|); zero when |imag| is at most FLT_EPSILON.
NPH_HD f3 q_scaled_axis(q4 q) {
  f3 v{q.i, q.j, q.k};
  if (q.w < 0.0f) v = neg(v);
  const float s = sqrtf(dot(v, v));
  if (s <= 1.1920928955078125e-07f) return f3{0.0f, 0.0f, 0.0f};
  const float ang = 2.0f * angle_of(s, fabsf(q.w));
  return f3{v.x / s * ang, v.y / s * ang, v.z / s * ang};
}

// Scaled axis of the shortest rotation taking unit axis a1 to unit axis a2: (a1 x a2) normalised
// times atan2(|a1 x a2|, a1 . a2); for (anti)parallel axes zero, or for opposite axes e_imin x a1
// normalised times pi, imin the index of a1's smallest |component|.
NPH_HD f3 axis_alignment_error(f3 a1, f3 a2) {
  const f3 c = cross(a1, a2);
  const float sn = sqrtf(dot(c, c));
  const float cs = dot(a1, a2);
  if (sn > 1.1920928955078125e-07f) {
    const float ang = angle_of(sn, cs);
    return f3{c.x / sn * ang, c.y / sn * ang, c.z / sn * ang};
  }
  if (cs >= 0.0f) return f3{0.0f, 0.0f, 0.0f};
  int imin = 0;
  if (fabsf(a1.y) < fabsf(comp(a1, imin))) imin = 1;
  if (fabsf(a1.z) < fabsf(comp(a1, imin))) imin = 2;
  f3 e{0.0f, 0.0f, 0.0f};
  set_comp(e, imin, 1.0f);
  const f3 d = cross(e, a1);
  const float l = sqrtf(dot(d, d));
  const float pi = float(3.14159265358979323846);
  return f3{d.x / l * pi, d.y / l * pi, d.z / l * pi};
}

// Pose update of apply_displacement: T(com) * Iso(lin, exp(ang)) * T(-com) * pose, com = t.
NPH_HD void apply_displacement(f3& t, q4& q, f3 lin, f3 ang) {
  const q4 qd = qexp_scaled_axis(ang);
  const f3 com = t;
  const f3 dt_ = add(com, add(lin, qrot(qd, neg(com))));
  t = add(dt_, qrot(qd, t));
  q = qmul(qd, q);
}

// Axis-aligned boxes.
struct aabb {
  f3 mn, mx;
};
NPH_HD aabb loosened(aabb a, float amount) {
  const f3 d{amount, amount, amount};
  return aabb{sub(a.mn, d), add(a.mx, d)};
}
NPH_HD aabb merged(aabb a, aabb b) { return aabb{vmin(a.mn, b.mn), vmax(a.mx, b.mx)}; }
NPH_HD bool contains(aabb a, aabb b) {
  return a.mn.x <= b.mn.x && a.mn.y <= b.mn.y && a.mn.z <= b.mn.z && a.mx.x >= b.mx.x &&
         a.mx.y >= b.mx.y && a.mx.z >= b.mx.z;
}
NPH_HD bool intersects(aabb a, aabb b) {
  return a.mn.x <= b.mx.x && a.mn.y <= b.mx.y && a.mn.z <= b.mx.z && a.mx.x >= b.mn.x &&
         a.mx.y >= b.mn.y && a.mx.z >= b.mn.z;
}

enum : int { SHAPE_BALL = 0, SHAPE_CUBOID = 1 };
enum : int { STATUS_DYNAMIC = 0, STATUS_STATIC = 1 };

// Shape AABB at a pose: ball centre -/+ r, cuboid centre -/+ |R| h.
NPH_HD aabb shape_aabb(int kind, f3 he, f3 t, q4 q) {
  if (kind == SHAPE_BALL) {
    const f3 r{he.x, he.x, he.x};
    return aabb{sub(t, r), add(t, r)};
  }
  const m33 R = qmat(q);
  m33 A;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A.m[i][j] = fabsf(R.m[i][j]);
  const f3 w = mulv(A, he);
  return aabb{sub(t, w), add(t, w)};
}

}  // namespace nph
