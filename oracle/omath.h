// CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// Small fixed-size double-precision linear algebra used by the oracle's
// restatement of mind-fcl's narrowphase.  The reference uses Eigen for
// these (include/fcl/common/types.h:69-91, include/fcl/cvx_collide/
// vec_types.h:14-22); the oracle writes the textbook operations directly.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// use anything under oracle/.
#pragma once

#include <cmath>
#include <cstdint>

namespace oracle {

struct V3 {
  double x{0}, y{0}, z{0};
  V3() = default;
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(const V3& a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(const V3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, const V3& a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(const V3& a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, const V3& b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline V3& operator-=(V3& a, const V3& b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline V3& operator*=(V3& a, double s) { a.x *= s; a.y *= s; a.z *= s; return a; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double sqnorm(const V3& a) { return dot(a, a); }
inline double norm(const V3& a) { return std::sqrt(sqnorm(a)); }
inline V3 normalized(const V3& a) {
  const double n = norm(a);
  return n > 0 ? a / n : a;
}
inline void normalize(V3& a) {
  const double n = norm(a);
  if (n > 0) a = a / n;
}
inline double absnorm(const V3& a) { return std::fabs(a.x) + std::fabs(a.y) + std::fabs(a.z); }

// Row-major 3x3 matrix.
struct M3 {
  double m[3][3]{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  V3 row(int i) const { return {m[i][0], m[i][1], m[i][2]}; }
  V3 col(int j) const { return {m[0][j], m[1][j], m[2][j]}; }
  void setCol(int j, const V3& v) { m[0][j] = v.x; m[1][j] = v.y; m[2][j] = v.z; }
  void setRow(int i, const V3& v) { m[i][0] = v.x; m[i][1] = v.y; m[i][2] = v.z; }
};

inline V3 operator*(const M3& A, const V3& v) {
  return {A.m[0][0] * v.x + A.m[0][1] * v.y + A.m[0][2] * v.z,
          A.m[1][0] * v.x + A.m[1][1] * v.y + A.m[1][2] * v.z,
          A.m[2][0] * v.x + A.m[2][1] * v.y + A.m[2][2] * v.z};
}
inline M3 operator*(const M3& A, const M3& B) {
  M3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j];
  return C;
}
inline M3 transpose(const M3& A) {
  M3 T;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T.m[i][j] = A.m[j][i];
  return T;
}
// A^T v without forming A^T.
inline V3 tmul(const M3& A, const V3& v) {
  return {A.m[0][0] * v.x + A.m[1][0] * v.y + A.m[2][0] * v.z,
          A.m[0][1] * v.x + A.m[1][1] * v.y + A.m[2][1] * v.z,
          A.m[0][2] * v.x + A.m[1][2] * v.y + A.m[2][2] * v.z};
}

// Rigid transform x -> R x + t (the reference's Transform3 = Eigen Isometry).
struct Tf {
  M3 R;
  V3 t;
};
inline V3 operator*(const Tf& T, const V3& p) { return T.R * p + T.t; }
inline Tf operator*(const Tf& A, const Tf& B) { return {A.R * B.R, A.R * B.t + A.t}; }
inline Tf inverse(const Tf& A) {
  Tf I;
  I.R = transpose(A.R);
  I.t = -(I.R * A.t);
  return I;
}

// Pose given as translation + unit quaternion (w, x, y, z), the batched
// input format; the quaternion is renormalised in double before use so the
// oracle and the device path see the same rotation.
inline Tf poseFromQuat(double tx, double ty, double tz, double qw, double qx, double qy,
                       double qz) {
  const double n = std::sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
  qw /= n; qx /= n; qy /= n; qz /= n;
  Tf T;
  T.R.m[0][0] = 1 - 2 * (qy * qy + qz * qz);
  T.R.m[0][1] = 2 * (qx * qy - qz * qw);
  T.R.m[0][2] = 2 * (qx * qz + qy * qw);
  T.R.m[1][0] = 2 * (qx * qy + qz * qw);
  T.R.m[1][1] = 1 - 2 * (qx * qx + qz * qz);
  T.R.m[1][2] = 2 * (qy * qz - qx * qw);
  T.R.m[2][0] = 2 * (qx * qz - qy * qw);
  T.R.m[2][1] = 2 * (qy * qz + qx * qw);
  T.R.m[2][2] = 1 - 2 * (qx * qx + qy * qy);
  T.t = {tx, ty, tz};
  return T;
}

}  // namespace oracle
