// oracle/orc_math.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Small fixed-size linear algebra for the CPU restatement of nphysics' contact pipeline.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the oracle; the
// product library (nphysics_amd/csrc) never includes or links anything under oracle/.
//
// Every operation below is written with an explicit evaluation order so that the FP32 build
// (compiled with -ffp-contract=off) reproduces the GPU kernels' arithmetic term by term.
//
// Reference semantics followed (nalgebra 0.28 is unvendored, so each formula here is the textbook
// one and its rounding is "parity unpinned" against nalgebra itself):
//   * quaternion layout (i, j, k, w) and Hamilton product: nalgebra UnitQuaternion, used by
//     src/object/rigid_body.rs:371-381 (apply_displacement) and algebra/velocity3.rs:57-59
//     (Velocity3::to_transform -> Isometry3::new(linear, angular)).
//   * quaternion * vector: v + w*t + u x t with t = 2 u x v.
//   * rotation matrix of a quaternion in homogeneous form (ww+ii-jj-kk ...).
//   * exp of a pure quaternion (from_scaled_axis): (cos|h|, h sin|h|/|h|), h = axisangle/2.
//   * 3x3 inverse by cofactors, zero when the determinant is zero (inertia3.rs:77-86).
//   * cross matrix [v]x (utils/generalized_cross.rs:70-99 -> nalgebra cross_matrix).
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>

namespace orc {

template <class T>
struct V3 {
  T x, y, z;
};

template <class T>
inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T>
inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T>
inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T>
inline T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T>
inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
inline T norm2(V3<T> a) { return dot(a, a); }
template <class T>
inline T norm(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T>
inline V3<T> vabs(V3<T> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
template <class T>
inline T comp(V3<T> a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
template <class T>
inline void set_comp(V3<T>& a, int i, T v) {
  if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v;
}
template <class T>
inline T tmin(T a, T b) { return a < b ? a : b; }
template <class T>
inline T tmax(T a, T b) { return a > b ? a : b; }
template <class T>
inline V3<T> vmin(V3<T> a, V3<T> b) { return {tmin(a.x, b.x), tmin(a.y, b.y), tmin(a.z, b.z)}; }
template <class T>
inline V3<T> vmax(V3<T> a, V3<T> b) { return {tmax(a.x, b.x), tmax(a.y, b.y), tmax(a.z, b.z)}; }

// Row-major 3x3 matrix.
template <class T>
struct M3 {
  T m[3][3];
};

template <class T>
inline M3<T> m3_zero() {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = T(0);
  return r;
}
template <class T>
inline M3<T> m3_diag(T a, T b, T c) {
  M3<T> r = m3_zero<T>();
  r.m[0][0] = a; r.m[1][1] = b; r.m[2][2] = c;
  return r;
}
// C = A * B, each entry summed in k order 0,1,2.
template <class T>
inline M3<T> mul(const M3<T>& a, const M3<T>& b) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
  return r;
}
template <class T>
inline M3<T> transpose(const M3<T>& a) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
  return r;
}
template <class T>
inline M3<T> madd(const M3<T>& a, const M3<T>& b) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j];
  return r;
}
template <class T>
inline M3<T> msub(const M3<T>& a, const M3<T>& b) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] - b.m[i][j];
  return r;
}
template <class T>
inline M3<T> mscale(const M3<T>& a, T s) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] * s;
  return r;
}
// y = A v, row sums in column order.
template <class T>
inline V3<T> mulv(const M3<T>& a, V3<T> v) {
  return {(a.m[0][0] * v.x + a.m[0][1] * v.y) + a.m[0][2] * v.z,
          (a.m[1][0] * v.x + a.m[1][1] * v.y) + a.m[1][2] * v.z,
          (a.m[2][0] * v.x + a.m[2][1] * v.y) + a.m[2][2] * v.z};
}
// y = A^T v.
template <class T>
inline V3<T> mulv_t(const M3<T>& a, V3<T> v) {
  return {(a.m[0][0] * v.x + a.m[1][0] * v.y) + a.m[2][0] * v.z,
          (a.m[0][1] * v.x + a.m[1][1] * v.y) + a.m[2][1] * v.z,
          (a.m[0][2] * v.x + a.m[1][2] * v.y) + a.m[2][2] * v.z};
}
template <class T>
inline V3<T> col(const M3<T>& a, int j) { return {a.m[0][j], a.m[1][j], a.m[2][j]}; }

// [v]x such that [v]x w = v x w.
template <class T>
inline M3<T> cross_matrix(V3<T> v) {
  M3<T> r;
  r.m[0][0] = T(0);  r.m[0][1] = -v.z; r.m[0][2] = v.y;
  r.m[1][0] = v.z;   r.m[1][1] = T(0); r.m[1][2] = -v.x;
  r.m[2][0] = -v.y;  r.m[2][1] = v.x;  r.m[2][2] = T(0);
  return r;
}

// Cofactor inverse; returns false (and zero) when the determinant is exactly zero.
template <class T>
inline bool try_inverse(const M3<T>& a, M3<T>& out) {
  const T c00 = a.m[1][1] * a.m[2][2] - a.m[2][1] * a.m[1][2];
  const T c01 = a.m[1][0] * a.m[2][2] - a.m[2][0] * a.m[1][2];
  const T c02 = a.m[1][0] * a.m[2][1] - a.m[2][0] * a.m[1][1];
  const T det = (a.m[0][0] * c00 - a.m[0][1] * c01) + a.m[0][2] * c02;
  if (det == T(0)) {
    out = m3_zero<T>();
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

// Unit quaternion stored as (i, j, k, w).
template <class T>
struct Q {
  T i, j, k, w;
};

template <class T>
inline Q<T> qmul(Q<T> a, Q<T> b) {
  Q<T> r;
  r.w = ((a.w * b.w - a.i * b.i) - a.j * b.j) - a.k * b.k;
  r.i = ((a.w * b.i + a.i * b.w) + a.j * b.k) - a.k * b.j;
  r.j = ((a.w * b.j - a.i * b.k) + a.j * b.w) + a.k * b.i;
  r.k = ((a.w * b.k + a.i * b.j) - a.j * b.i) + a.k * b.w;
  return r;
}

// q * v = v + w t + u x t, t = 2 (u x v).
template <class T>
inline V3<T> qrot(Q<T> q, V3<T> v) {
  const V3<T> u{q.i, q.j, q.k};
  const V3<T> t = cross(u, v) * T(2);
  const V3<T> c = cross(u, t);
  return (t * q.w + c) + v;
}

template <class T>
inline M3<T> qmat(Q<T> q) {
  const T ww = q.w * q.w, ii = q.i * q.i, jj = q.j * q.j, kk = q.k * q.k;
  const T ij = q.i * q.j * T(2), wk = q.w * q.k * T(2), wj = q.w * q.j * T(2);
  const T ik = q.i * q.k * T(2), jk = q.j * q.k * T(2), wi = q.w * q.i * T(2);
  M3<T> r;
  r.m[0][0] = ((ww + ii) - jj) - kk; r.m[0][1] = ij - wk;               r.m[0][2] = wj + ik;
  r.m[1][0] = wk + ij;               r.m[1][1] = ((ww - ii) + jj) - kk; r.m[1][2] = jk - wi;
  r.m[2][0] = ik - wj;               r.m[2][1] = wi + jk;               r.m[2][2] = ((ww - ii) - jj) + kk;
  return r;
}

template <class T>
inline T eps_of() { return std::numeric_limits<T>::epsilon(); }

// exp of the pure quaternion (0, axisangle/2): the rotation by |axisangle| about its axis.
// For half angles below 1/4 rad, sin(n)/n and cos(n) are evaluated by their Maclaurin series
// through n^8 (truncation < 4e-10, below one FP32 ulp of 1), in Horner form, so that the FP32
// build and the GPU kernels round identically; larger angles use the C library.
template <class T>
inline void sinc_cos(T n, T& s, T& c) {
  // sin(x)/x and cos(x) by their Maclaurin polynomials on [0, 0.25); a larger angle is halved
  // (exactly) until below 0.25 and rebuilt with sin 2x = 2 sin x cos x, cos 2x = cos^2 x - sin^2 x,
  // so every angle is plain arithmetic that the GPU reproduces bit for bit
  T x = n;
  int k = 0;
  while (x >= T(0.25) && k < 64) {
    x = x * T(0.5);
    ++k;
  }
  const T x2 = x * x;
  const T sx = T(1) + x2 * (T(-1.0 / 6.0) + x2 * (T(1.0 / 120.0) + x2 * (T(-1.0 / 5040.0) + x2 * T(1.0 / 362880.0))));
  T cx = T(1) + x2 * (T(-0.5) + x2 * (T(1.0 / 24.0) + x2 * (T(-1.0 / 720.0) + x2 * T(1.0 / 40320.0))));
  if (k == 0) {
    s = sx;
    c = cx;
    return;
  }
  T sn = x * sx;
  for (int i = 0; i < k; ++i) {
    const T s2 = T(2) * sn * cx;
    const T c2 = cx * cx - sn * sn;
    sn = s2;
    cx = c2;
  }
  s = sn / n;
  c = cx;
}

template <class T>
inline Q<T> qexp_scaled_axis(V3<T> axisangle) {
  const V3<T> h = axisangle * T(0.5);
  const T nn = norm2(h);
  const T e = eps_of<T>();
  if (nn <= e * e) return Q<T>{T(0), T(0), T(0), T(1)};
  const T n = std::sqrt(nn);
  T s, c;
  sinc_cos(n, s, c);
  return Q<T>{h.x * s, h.y * s, h.z * s, c};
}

// atan(x) for 0 <= x <= 1 by plain arithmetic (the GPU evaluates the same expression tree): three
// half-angle reductions x -> x / (1 + sqrt(1 + x^2)) (atan x = 2 atan x') bring x below
// tan(pi/32) ~ 0.0985, then the Maclaurin series of atan through x^19 in Horner form (truncation
// below 1e-20 relative there).
template <class T>
inline T atan_unit(T x) {
  for (int i = 0; i < 3; ++i) x = x / (T(1) + std::sqrt(T(1) + x * x));
  const T x2 = x * x;
  T p = T(-1.0 / 19.0);
  p = T(1.0 / 17.0) + x2 * p;
  p = T(-1.0 / 15.0) + x2 * p;
  p = T(1.0 / 13.0) + x2 * p;
  p = T(-1.0 / 11.0) + x2 * p;
  p = T(1.0 / 9.0) + x2 * p;
  p = T(-1.0 / 7.0) + x2 * p;
  p = T(1.0 / 5.0) + x2 * p;
  p = T(-1.0 / 3.0) + x2 * p;
  p = T(1) + x2 * p;
  return T(8) * (x * p);
}

// atan2(y, x) for y >= 0, in [0, pi].
template <class T>
inline T angle_of(T y, T x) {
  const T pi = T(3.14159265358979323846);
  const T ax = std::fabs(x);
  if (y == T(0) && ax == T(0)) return T(0);
  const T base = ax >= y ? atan_unit(y / ax) : pi * T(0.5) - atan_unit(ax / y);
  return x >= T(0) ? base : pi - base;
}

template <class T>
inline Q<T> qconj(Q<T> q) { return Q<T>{-q.i, -q.j, -q.k, q.w}; }

// UnitQuaternion::scaled_axis (nalgebra, unvendored; restated): axis times angle, the axis the
// imaginary part (negated when w < 0) normalised, the angle 2 atan2(|imag|, |w|) in [0, pi]; zero
// when |imag| is at most the scalar's epsilon.
template <class T>
inline V3<T> q_scaled_axis(Q<T> q) {
  V3<T> v{q.i, q.j, q.k};
  if (q.w < T(0)) v = -v;
  const T s = norm(v);
  if (s <= std::numeric_limits<T>::epsilon()) return V3<T>{T(0), T(0), T(0)};
  const T ang = T(2) * angle_of(s, std::fabs(q.w));
  return V3<T>{v.x / s * ang, v.y / s * ang, v.z / s * ang};
}

// Rigid transform.
template <class T>
struct Iso {
  V3<T> t;
  Q<T> q;
};

template <class T>
inline V3<T> iso_apply(const Iso<T>& p, V3<T> v) { return qrot(p.q, v) + p.t; }

// Axis-aligned box.
template <class T>
struct AABB {
  V3<T> mn, mx;
};
template <class T>
inline AABB<T> loosened(const AABB<T>& a, T amount) {
  const V3<T> d{amount, amount, amount};
  return AABB<T>{a.mn - d, a.mx + d};
}
template <class T>
inline AABB<T> merged(const AABB<T>& a, const AABB<T>& b) {
  return AABB<T>{vmin(a.mn, b.mn), vmax(a.mx, b.mx)};
}
// a contains b (inclusive on every face).
template <class T>
inline bool contains(const AABB<T>& a, const AABB<T>& b) {
  return a.mn.x <= b.mn.x && a.mn.y <= b.mn.y && a.mn.z <= b.mn.z && a.mx.x >= b.mx.x &&
         a.mx.y >= b.mx.y && a.mx.z >= b.mx.z;
}
// Closed-interval overlap on every axis.
template <class T>
inline bool intersects(const AABB<T>& a, const AABB<T>& b) {
  return a.mn.x <= b.mx.x && a.mn.y <= b.mx.y && a.mn.z <= b.mx.z && a.mx.x >= b.mn.x &&
         a.mx.y >= b.mn.y && a.mx.z >= b.mn.z;
}

}  // namespace orc
