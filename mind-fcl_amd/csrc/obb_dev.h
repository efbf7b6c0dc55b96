// Oriented-box nodes of the mesh trees and their FP32 overlap test, shared
// by the mesh-mesh traversal (bvh.hip) and the mesh-shape traversal
// (meshshape.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dmath.h"

namespace mfcl {

struct alignas(16) ObbNode {
  float ax[9];  // box axes, column-major: ax[0..2] = axis 0
  float c[3];   // centre
  float e[3];   // half extents
  int32_t child;  // >= 0: first child (sibling at child + 1); < 0: leaf -(tri + 1)
};
static_assert(sizeof(ObbNode) == 64, "node must be 64 bytes");

// Relative FP32 rounding scale of the overlap test's margin.
constexpr float kObbScaleErr = 16.0f * 5.9604645e-08f;

#pragma clang fp contract(fast)
// FP32 OBB overlap with a margin: returns false only when the separating
// axis clears the boxes by more than the accumulated FP32 rounding error,
// so the FP64 test (OBB-inl.h:306-436, |B| padded by 1e-6) would also
// report disjoint.  Rq (3x3 row-major) and Tq: pose of tree-2 frame in
// tree-1 frame, FP32.
MFCL_D bool obbOverlapF(const float* Rq, const float* Tq, const ObbNode& a, const ObbNode& b, float scale_err) {
  // B = a.axes^T * Rq * b.axes ; T = a.axes^T * (Rq * b.c + Tq - a.c)
  float RB[9];  // Rq * b.axes (column-major columns)
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float bx = b.ax[3 * j], by = b.ax[3 * j + 1], bz = b.ax[3 * j + 2];
    RB[3 * j] = Rq[0] * bx + Rq[1] * by + Rq[2] * bz;
    RB[3 * j + 1] = Rq[3] * bx + Rq[4] * by + Rq[5] * bz;
    RB[3 * j + 2] = Rq[6] * bx + Rq[7] * by + Rq[8] * bz;
  }
  float B[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      B[i][j] = a.ax[3 * i] * RB[3 * j] + a.ax[3 * i + 1] * RB[3 * j + 1] + a.ax[3 * i + 2] * RB[3 * j + 2];
  const float tx = Rq[0] * b.c[0] + Rq[1] * b.c[1] + Rq[2] * b.c[2] + Tq[0] - a.c[0];
  const float ty = Rq[3] * b.c[0] + Rq[4] * b.c[1] + Rq[5] * b.c[2] + Tq[1] - a.c[1];
  const float tz = Rq[6] * b.c[0] + Rq[7] * b.c[1] + Rq[8] * b.c[2] + Tq[2] - a.c[2];
  float T[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) T[i] = a.ax[3 * i] * tx + a.ax[3 * i + 1] * ty + a.ax[3 * i + 2] * tz;
  float Bf[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Bf[i][j] = fabsf(B[i][j]) + 1e-6f;
  const float* ea = a.e;
  const float* eb = b.e;
  // translational rounding error bound (absolute, in length units)
  const float m = scale_err * (fabsf(a.c[0]) + fabsf(a.c[1]) + fabsf(a.c[2]) + fabsf(b.c[0]) + fabsf(b.c[1]) +
                               fabsf(b.c[2]) + ea[0] + ea[1] + ea[2] + eb[0] + eb[1] + eb[2]) +
                  scale_err * (fabsf(Tq[0]) + fabsf(Tq[1]) + fabsf(Tq[2]));
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (fabsf(T[i]) > ea[i] + Bf[i][0] * eb[0] + Bf[i][1] * eb[1] + Bf[i][2] * eb[2] + m) return false;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float s = B[0][j] * T[0] + B[1][j] * T[1] + B[2][j] * T[2];
    if (fabsf(s) > eb[j] + Bf[0][j] * ea[0] + Bf[1][j] * ea[1] + Bf[2][j] * ea[2] + m) return false;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      const float s = T[i2] * B[i1][j] - T[i1] * B[i2][j];
      const float r = ea[i1] * Bf[i2][j] + ea[i2] * Bf[i1][j] + eb[j1] * Bf[i][j2] + eb[j2] * Bf[i][j1];
      if (fabsf(s) > r + 2.0f * m) return false;
    }
  }
  return true;
}

#pragma clang fp contract(off)

}  // namespace mfcl
