// Device-side fixed-size FP64/FP32 vector math for the gfx950 kernels.
#pragma once

// FP64 results must follow the oracle's operation order exactly, so the
// compiler may not fuse multiply-adds in the narrowphase code; FP32 loops
// that only rank candidates opt back in locally.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MFCL_D __device__ __forceinline__

namespace mfcl {

struct d3 {
  double x, y, z;
};

MFCL_D d3 mk(double x, double y, double z) { return d3{x, y, z}; }
MFCL_D d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
MFCL_D d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
MFCL_D d3 operator-(d3 a) { return {-a.x, -a.y, -a.z}; }
MFCL_D d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
MFCL_D d3 operator*(double s, d3 a) { return {a.x * s, a.y * s, a.z * s}; }
MFCL_D d3 operator/(d3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
MFCL_D double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MFCL_D d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
MFCL_D double sqnorm(d3 a) { return dot(a, a); }
MFCL_D double norm(d3 a) { return sqrt(dot(a, a)); }
MFCL_D double absnorm(d3 a) { return fabs(a.x) + fabs(a.y) + fabs(a.z); }
MFCL_D d3 normalized(d3 a) {
  const double n = norm(a);
  return n > 0 ? a / n : a;
}
MFCL_D double comp(d3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
MFCL_D void setComp(d3& a, int i, double v) {
  if (i == 0) a.x = v;
  else if (i == 1) a.y = v;
  else a.z = v;
}

// Row-major 3x3.
struct m33 {
  double m00, m01, m02, m10, m11, m12, m20, m21, m22;
};
MFCL_D d3 mul(const m33& A, d3 v) {
  return {A.m00 * v.x + A.m01 * v.y + A.m02 * v.z, A.m10 * v.x + A.m11 * v.y + A.m12 * v.z,
          A.m20 * v.x + A.m21 * v.y + A.m22 * v.z};
}
MFCL_D d3 tmul(const m33& A, d3 v) {
  return {A.m00 * v.x + A.m10 * v.y + A.m20 * v.z, A.m01 * v.x + A.m11 * v.y + A.m21 * v.z,
          A.m02 * v.x + A.m12 * v.y + A.m22 * v.z};
}
MFCL_D d3 col(const m33& A, int j) {
  return j == 0 ? d3{A.m00, A.m10, A.m20} : (j == 1 ? d3{A.m01, A.m11, A.m21} : d3{A.m02, A.m12, A.m22});
}
MFCL_D d3 row(const m33& A, int i) {
  return i == 0 ? d3{A.m00, A.m01, A.m02} : (i == 1 ? d3{A.m10, A.m11, A.m12} : d3{A.m20, A.m21, A.m22});
}
// A^T B
MFCL_D m33 tmulm(const m33& A, const m33& B) {
  m33 C;
  C.m00 = A.m00 * B.m00 + A.m10 * B.m10 + A.m20 * B.m20;
  C.m01 = A.m00 * B.m01 + A.m10 * B.m11 + A.m20 * B.m21;
  C.m02 = A.m00 * B.m02 + A.m10 * B.m12 + A.m20 * B.m22;
  C.m10 = A.m01 * B.m00 + A.m11 * B.m10 + A.m21 * B.m20;
  C.m11 = A.m01 * B.m01 + A.m11 * B.m11 + A.m21 * B.m21;
  C.m12 = A.m01 * B.m02 + A.m11 * B.m12 + A.m21 * B.m22;
  C.m20 = A.m02 * B.m00 + A.m12 * B.m10 + A.m22 * B.m20;
  C.m21 = A.m02 * B.m01 + A.m12 * B.m11 + A.m22 * B.m21;
  C.m22 = A.m02 * B.m02 + A.m12 * B.m12 + A.m22 * B.m22;
  return C;
}
MFCL_D m33 mulm(const m33& A, const m33& B) {
  m33 C;
  C.m00 = A.m00 * B.m00 + A.m01 * B.m10 + A.m02 * B.m20;
  C.m01 = A.m00 * B.m01 + A.m01 * B.m11 + A.m02 * B.m21;
  C.m02 = A.m00 * B.m02 + A.m01 * B.m12 + A.m02 * B.m22;
  C.m10 = A.m10 * B.m00 + A.m11 * B.m10 + A.m12 * B.m20;
  C.m11 = A.m10 * B.m01 + A.m11 * B.m11 + A.m12 * B.m21;
  C.m12 = A.m10 * B.m02 + A.m11 * B.m12 + A.m12 * B.m22;
  C.m20 = A.m20 * B.m00 + A.m21 * B.m10 + A.m22 * B.m20;
  C.m21 = A.m20 * B.m01 + A.m21 * B.m11 + A.m22 * B.m21;
  C.m22 = A.m20 * B.m02 + A.m21 * B.m12 + A.m22 * B.m22;
  return C;
}
MFCL_D m33 transpose(const m33& A) {
  return {A.m00, A.m10, A.m20, A.m01, A.m11, A.m21, A.m02, A.m12, A.m22};
}

struct tf3 {
  m33 R;
  d3 t;
};
MFCL_D d3 apply(const tf3& T, d3 p) { return mul(T.R, p) + T.t; }

// Translation + quaternion (w, x, y, z) in FP32 -> rigid transform in FP64;
// the quaternion is renormalised in double.
MFCL_D tf3 poseFromQuat(const float* p) {
  double qw = p[3], qx = p[4], qy = p[5], qz = p[6];
  const double n = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
  qw /= n; qx /= n; qy /= n; qz /= n;
  tf3 T;
  T.R.m00 = 1 - 2 * (qy * qy + qz * qz);
  T.R.m01 = 2 * (qx * qy - qz * qw);
  T.R.m02 = 2 * (qx * qz + qy * qw);
  T.R.m10 = 2 * (qx * qy + qz * qw);
  T.R.m11 = 1 - 2 * (qx * qx + qz * qz);
  T.R.m12 = 2 * (qy * qz - qx * qw);
  T.R.m20 = 2 * (qx * qz - qy * qw);
  T.R.m21 = 2 * (qy * qz + qx * qw);
  T.R.m22 = 1 - 2 * (qx * qx + qy * qy);
  T.t = {p[0], p[1], p[2]};
  return T;
}

// inverse(A) * B
MFCL_D tf3 relTf(const tf3& A, const tf3& B) {
  tf3 r;
  r.R = tmulm(A.R, B.R);
  const d3 ti = -tmul(A.R, A.t);
  r.t = tmul(A.R, B.t) + ti;
  return r;
}

}  // namespace mfcl
