// ORACLE — test infrastructure only. Nothing under oracle/ is part of the shipped library: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as the checker.
//
// Vector, quaternion and 3x3 helpers, templated on the real type R (double for the FP64 oracle, float for the
// FP32 mirror used for bit-exact discrete decisions). Every function evaluates its expression in the same order as
// the reference's maths (reference/three-avbd/src/avbd3d/ref/math.ts, and its WGSL twins in
// reference/three-avbd/src/avbd3d/gpu/layout.ts:209-242), left to right, with no fused multiply-add: the oracle is
// compiled with -ffp-contract=off.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace oracle {

template <class R> struct V3 {
  R x = 0, y = 0, z = 0;
  V3() = default;
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  R &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

template <class R> struct Q4 {
  R x = 0, y = 0, z = 0, w = 1;
  Q4() = default;
  Q4(R a, R b, R c, R d) : x(a), y(b), z(c), w(d) {}
  V3<R> xyz() const { return V3<R>(x, y, z); }
};

// Row-major 3x3 (math.ts: Float64Array(9)).
template <class R> struct M3 {
  R m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  R &operator()(int r, int c) { return m[r * 3 + c]; }
  R operator()(int r, int c) const { return m[r * 3 + c]; }
};

template <class R> inline R rmin(R a, R b) { return a < b ? a : b; }  // math.ts:16
template <class R> inline R rmax(R a, R b) { return a > b ? a : b; }  // math.ts:17
template <class R> inline R rclamp(R x, R a, R b) { return rmax(a, rmin(b, x)); }  // math.ts:18
template <class R> inline R rsign(R x) { return x < 0 ? R(-1) : (x > 0 ? R(1) : R(0)); }

template <class R> inline V3<R> add(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> sub(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> scale(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> divs(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline V3<R> neg(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> vabs(V3<R> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
template <class R> inline V3<R> mulv3(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
/** a + b * s (math.ts addScaled3). */
template <class R> inline V3<R> addScaled(V3<R> a, V3<R> b, R s) { return {a.x + b.x * s, a.y + b.y * s, a.z + b.z * s}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R lengthSq(V3<R> a) { return dot(a, a); }
template <class R> inline R length(V3<R> a) { return std::sqrt(lengthSq(a)); }
template <class R> inline V3<R> normalize(V3<R> a) { return divs(a, length(a)); }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

/** math.ts:57-67 (qmul). */
template <class R> inline Q4<R> qmul(Q4<R> a, Q4<R> b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <class R> inline R qlengthSq(Q4<R> q) { return (q.x * q.x + q.y * q.y + q.z * q.z) + q.w * q.w; }  // math.ts:70
template <class R> inline Q4<R> qinverse(Q4<R> q) {  // math.ts:72-75
  R l = qlengthSq(q);
  return {-q.x / l, -q.y / l, -q.z / l, q.w / l};
}
template <class R> inline Q4<R> qnormalize(Q4<R> q) {  // math.ts:77-80
  R l = std::sqrt(qlengthSq(q));
  return {q.x / l, q.y / l, q.z / l, q.w / l};
}
/** The demo's quat - quat: 2·vec(a·b⁻¹) (math.ts:86-89). */
template <class R> inline V3<R> qsub(Q4<R> a, Q4<R> b) {
  Q4<R> s = qmul(a, qinverse(b));
  return {s.x * 2, s.y * 2, s.z * 2};
}
/** The demo's quat + float3: normalize(a + (v, 0)·a·½) (math.ts:92-96). */
template <class R> inline Q4<R> qadd(Q4<R> a, V3<R> v) {
  Q4<R> s = qmul(Q4<R>(v.x, v.y, v.z, 0), a);
  return qnormalize(Q4<R>(a.x + s.x * R(0.5), a.y + s.y * R(0.5), a.z + s.z * R(0.5), a.w + s.w * R(0.5)));
}
/** math.ts:98-109 (rotate). */
template <class R> inline V3<R> rotate(Q4<R> q, V3<R> v) {
  R tx = (q.y * v.z - q.z * v.y) * 2;
  R ty = (q.z * v.x - q.x * v.z) * 2;
  R tz = (q.x * v.y - q.y * v.x) * 2;
  return {v.x + tx * q.w + (q.y * tz - q.z * ty), v.y + ty * q.w + (q.z * tx - q.x * tz),
          v.z + tz * q.w + (q.x * ty - q.y * tx)};
}
template <class R> inline Q4<R> conj(Q4<R> q) { return {-q.x, -q.y, -q.z, q.w}; }
template <class R> inline V3<R> rotateInv(Q4<R> q, V3<R> v) { return rotate(conj(q), v); }
/** rotate then translate (math.ts transform). */
template <class R> inline V3<R> transform(V3<R> p, Q4<R> q, V3<R> v) { return add(rotate(q, v), p); }

// 3x3 (row-major) -------------------------------------------------------------------------------------------------
template <class R> inline M3<R> diag(R a, R b, R c) {
  M3<R> o;
  o.m[0] = a;
  o.m[4] = b;
  o.m[8] = c;
  return o;
}
template <class R> inline M3<R> skew(V3<R> r) {
  M3<R> o;
  o.m[0] = 0; o.m[1] = -r.z; o.m[2] = r.y;
  o.m[3] = r.z; o.m[4] = 0; o.m[5] = -r.x;
  o.m[6] = -r.y; o.m[7] = r.x; o.m[8] = 0;
  return o;
}
template <class R> inline M3<R> transpose(const M3<R> &a) {
  M3<R> o;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) o.m[i * 3 + j] = a.m[j * 3 + i];
  return o;
}
/** Each entry dot(a.row(i), b.col(j)) (math.ts:156-171). */
template <class R> inline M3<R> mul(const M3<R> &a, const M3<R> &b) {
  M3<R> o;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      o.m[i * 3 + j] = a.m[i * 3] * b.m[j] + a.m[i * 3 + 1] * b.m[3 + j] + a.m[i * 3 + 2] * b.m[6 + j];
  return o;
}
template <class R> inline V3<R> mulv(const M3<R> &a, V3<R> v) {
  return {a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
          a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
template <class R> inline void addm(M3<R> &o, const M3<R> &a) {
  for (int i = 0; i < 9; i++) o.m[i] += a.m[i];
}
template <class R> inline M3<R> scalem(const M3<R> &a, R s) {
  M3<R> o;
  for (int i = 0; i < 9; i++) o.m[i] = a.m[i] * s;
  return o;
}
template <class R> inline M3<R> negm(const M3<R> &a) {
  M3<R> o;
  for (int i = 0; i < 9; i++) o.m[i] = -a.m[i];
  return o;
}
/** outer(a, b): entry (i, j) = b[j]·a[i] (math.ts:193-196). */
template <class R> inline M3<R> outer(V3<R> a, V3<R> b) {
  M3<R> o;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) o.m[i * 3 + j] = b[j] * a[i];
  return o;
}
/** Diagonal of the column norms (math.ts:199-204). */
template <class R> inline M3<R> diagonalize(const M3<R> &m) {
  R c0 = std::sqrt(m.m[0] * m.m[0] + m.m[3] * m.m[3] + m.m[6] * m.m[6]);
  R c1 = std::sqrt(m.m[1] * m.m[1] + m.m[4] * m.m[4] + m.m[7] * m.m[7]);
  R c2 = std::sqrt(m.m[2] * m.m[2] + m.m[5] * m.m[5] + m.m[8] * m.m[8]);
  return diag(c0, c1, c2);
}
/** Rows (n, t1, t2) (math.ts:207-219; layout.ts:236-242). */
template <class R> inline M3<R> orthonormal(V3<R> n) {
  V3<R> t1 = std::fabs(n.x) > std::fabs(n.z) ? V3<R>(-n.y, n.x, 0) : V3<R>(0, -n.z, n.y);
  t1 = normalize(t1);
  V3<R> t2 = cross(n, t1);
  M3<R> o;
  o.m[0] = n.x; o.m[1] = n.y; o.m[2] = n.z;
  o.m[3] = t1.x; o.m[4] = t1.y; o.m[5] = t1.z;
  o.m[6] = t2.x; o.m[7] = t2.y; o.m[8] = t2.z;
  return o;
}
template <class R> inline V3<R> row(const M3<R> &a, int k) { return {a.m[k * 3], a.m[k * 3 + 1], a.m[k * 3 + 2]}; }

/**
 * LDLᵀ solve of the 6x6 SPD system [aLin aCrossᵀ; aCross aAng]·x = b reading only the lower triangle
 * (math.ts:227-285). With pivotFloor > 0 every pivot D_k is floored at pivotFloor·A_kk, the GPU solver's
 * finishBody (wgsl-solve.ts:416-470, PIVOT_FLOOR layout.ts:106); pivotFloor == 0 is the upstream solve6.
 * `floored` reports a pivot that took its floor; `reciprocal` divides by each pivot through its reciprocal, formed
 * once (the HIP finishBody's form; the upstream restatement divides).
 */
template <class R>
inline void solve6(const M3<R> &aLin, const M3<R> &aAng, const M3<R> &aCross, V3<R> bLin, V3<R> bAng, V3<R> &xLin,
                   V3<R> &xAng, R pivotFloor = 0, bool *floored = nullptr, bool reciprocal = false) {
  const R A11 = aLin.m[0];
  const R A21 = aLin.m[3], A22 = aLin.m[4];
  const R A31 = aLin.m[6], A32 = aLin.m[7], A33 = aLin.m[8];
  const R A41 = aCross.m[0], A42 = aCross.m[1], A43 = aCross.m[2], A44 = aAng.m[0];
  const R A51 = aCross.m[3], A52 = aCross.m[4], A53 = aCross.m[5], A54 = aAng.m[3], A55 = aAng.m[4];
  const R A61 = aCross.m[6], A62 = aCross.m[7], A63 = aCross.m[8], A64 = aAng.m[6], A65 = aAng.m[7], A66 = aAng.m[8];
  auto floorp = [&](R d, R a) {
    if (!(pivotFloor > 0)) return d;
    if (floored && !(d > a * pivotFloor)) *floored = true;
    return rmax(d, a * pivotFloor);
  };
  // x / D, or x * (1 / D) with the reciprocal formed once per pivot (the GPU primal's form)
  auto over = [&](R x, R d, R r) { return reciprocal ? x * r : x / d; };

  const R D1 = A11;
  const R r1 = R(1) / D1;
  const R L21 = over(A21, A11, r1), L31 = over(A31, A11, r1), L41 = over(A41, A11, r1), L51 = over(A51, A11, r1),
          L61 = over(A61, A11, r1);
  const R D2 = floorp(A22 - L21 * L21 * D1, A22);
  const R r2 = R(1) / D2;
  const R L32 = over(A32 - L21 * L31 * D1, D2, r2);
  const R L42 = over(A42 - L21 * L41 * D1, D2, r2);
  const R L52 = over(A52 - L21 * L51 * D1, D2, r2);
  const R L62 = over(A62 - L21 * L61 * D1, D2, r2);
  const R D3 = floorp(A33 - (L31 * L31 * D1 + L32 * L32 * D2), A33);
  const R r3 = R(1) / D3;
  const R L43 = over(A43 - L31 * L41 * D1 - L32 * L42 * D2, D3, r3);
  const R L53 = over(A53 - L31 * L51 * D1 - L32 * L52 * D2, D3, r3);
  const R L63 = over(A63 - L31 * L61 * D1 - L32 * L62 * D2, D3, r3);
  const R D4 = floorp(A44 - (L41 * L41 * D1 + L42 * L42 * D2 + L43 * L43 * D3), A44);
  const R r4 = R(1) / D4;
  const R L54 = over(A54 - L41 * L51 * D1 - L42 * L52 * D2 - L43 * L53 * D3, D4, r4);
  const R L64 = over(A64 - L41 * L61 * D1 - L42 * L62 * D2 - L43 * L63 * D3, D4, r4);
  const R D5 = floorp(A55 - (L51 * L51 * D1 + L52 * L52 * D2 + L53 * L53 * D3 + L54 * L54 * D4), A55);
  const R r5 = R(1) / D5;
  const R L65 = over(A65 - L51 * L61 * D1 - L52 * L62 * D2 - L53 * L63 * D3 - L54 * L64 * D4, D5, r5);
  const R D6 = floorp(A66 - (L61 * L61 * D1 + L62 * L62 * D2 + L63 * L63 * D3 + L64 * L64 * D4 + L65 * L65 * D5), A66);
  const R r6 = R(1) / D6;

  const R y1 = bLin.x;
  const R y2 = bLin.y - L21 * y1;
  const R y3 = bLin.z - L31 * y1 - L32 * y2;
  const R y4 = bAng.x - L41 * y1 - L42 * y2 - L43 * y3;
  const R y5 = bAng.y - L51 * y1 - L52 * y2 - L53 * y3 - L54 * y4;
  const R y6 = bAng.z - L61 * y1 - L62 * y2 - L63 * y3 - L64 * y4 - L65 * y5;

  const R z1 = over(y1, D1, r1), z2 = over(y2, D2, r2), z3 = over(y3, D3, r3), z4 = over(y4, D4, r4),
          z5 = over(y5, D5, r5), z6 = over(y6, D6, r6);

  xAng.z = z6;
  xAng.y = z5 - L65 * xAng.z;
  xAng.x = z4 - L54 * xAng.y - L64 * xAng.z;
  xLin.z = z3 - L43 * xAng.x - L53 * xAng.y - L63 * xAng.z;
  xLin.y = z2 - L32 * xLin.z - L42 * xAng.x - L52 * xAng.y - L62 * xAng.z;
  xLin.x = z1 - L21 * xLin.y - L31 * xLin.z - L41 * xAng.x - L51 * xAng.y - L61 * xAng.z;
}

/**
 * atan2 and the wrap of an angle into [−π, π) for the hinge rows (angularConstraints.js:21-28 uses WGSL atan2, sin and
 * cos). The FP32 forms use FP32 arithmetic only, as the GPU step does: atan(t) on [0, 1] by the reduction
 * atan t = 30° + atan((√3 t − 1)/(√3 + t)) above tan 15° and the odd Taylor series to u¹¹ (error below 3e-9), octants
 * by symmetry; the wrap as x − 2π ⌊(x + π) / 2π⌋ (the value atan2(sin x, cos x) takes). The FP64 forms are the library's.
 */
inline float atanUnitF(float t) {
  const float TAN15 = 0.267949192f, SQRT3 = 1.73205081f, PI6 = 0.523598776f;
  float offset = 0.f, u = t;
  if (t > TAN15) {
    u = (t * SQRT3 - 1.f) / (SQRT3 + t);
    offset = PI6;
  }
  const float u2 = u * u;
  float p = -1.f / 11.f;
  p = p * u2 + 1.f / 9.f;
  p = p * u2 - 1.f / 7.f;
  p = p * u2 + 1.f / 5.f;
  p = p * u2 - 1.f / 3.f;
  p = p * u2 + 1.f;
  return offset + u * p;
}
inline float atan2R(float y, float x) {
  const float ax = std::fabs(x), ay = std::fabs(y);
  if (ax == 0.f && ay == 0.f) return 0.f;
  const bool steep = ay > ax;
  float a = atanUnitF(steep ? ax / ay : ay / ax);
  if (steep) a = 1.57079633f - a;
  if (x < 0.f) a = 3.14159265f - a;
  return y < 0.f ? -a : a;
}
inline double atan2R(double y, double x) { return std::atan2(y, x); }
inline float wrapAngleR(float x) { return x - 6.28318531f * std::floor((x + 3.14159265f) / 6.28318531f); }
inline double wrapAngleR(double x) { return x - 2 * M_PI * std::floor((x + M_PI) / (2 * M_PI)); }

/** lowbias32 integer hash (reference/three-avbd/src/avbd2d/gpu/layout.ts:106-115). */
inline uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

inline float bitsToFloat(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
inline uint32_t floatToBits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

}  // namespace oracle
