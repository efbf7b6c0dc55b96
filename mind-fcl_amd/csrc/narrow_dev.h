// Device narrowphase building blocks: support mapping, GJK (boolean and
// minimum distance), MPR, and the closed-form primitive pair tests.
//
// Semantics follow mind-fcl's cvx_collide and GJKSolver (file:line citations
// on each function); the representation is the GPU's own: one lane per pair,
// every simplex slot addressed with compile-time indices so the simplex lives
// in VGPRs, hull vertices read as 16-byte float4 records.
//
// Precision: hull support mapping scans vertices in FP32 and tracks the best
// and runner-up scores; when the FP32 gap is inside a proven rounding bound
// the scan is redone in FP64, so the selected vertex equals the FP64 argmax.
// Everything else (simplex sub-distance, portal updates, EPA polytope) is FP64.
#pragma once

#include "dmath.h"

namespace mfcl {

enum ShapeType : int { kBox = 0, kSphere = 1, kCapsule = 3, kCylinder = 5, kConvex = 100, kTriangle = 101 };

#define MFCL_LDS __attribute__((address_space(3)))

struct DevShape {
  int type;
  double p0, p1, p2;    // box: full sides; sphere: r; capsule/cylinder: r, full length
  const float4* verts;  // convex: vertices (x, y, z, unused)
  int nv;
  float rbound;         // convex: max |coordinate| over the vertices
  // convex: optional copy of the vertices in LDS (x at lv[i], y at lv[i + ls],
  // z at lv[i + 2 ls]); nullptr when the hull is read from device memory
  const MFCL_LDS float* lv;
  int ls;
  // convex with faces and more than 32 vertices: adjacency table in the
  // reference's encoding and the six axis seed vertices (support by the
  // vertex walk); nullptr: support by the naive scan
  const int* nbr;
  const int* seed;
  // walking hull: per vertex a 128-byte block, the walk record followed by
  // the coordinates of its first eight neighbours; nullptr: not available
  const float4* wblk;
};

// Vertex i of a hull, from the LDS copy when there is one (same values).
MFCL_D float3 hullVert(const DevShape& s, int i) {
  if (s.lv) return make_float3(s.lv[i], s.lv[i + s.ls], s.lv[i + 2 * s.ls]);
  const float4 q = s.verts[i];
  return make_float3(q.x, q.y, q.z);
}

// Vertex i of a walking hull and the offset of its neighbour record: the
// record offset rides in the vertex's fourth (otherwise unused) word in
// device memory, written when the faces are uploaded, so one 16-byte load
// serves both; the LDS copy has no fourth row and reads the table instead.
MFCL_D float3 hullVertRec(const DevShape& s, int i, int& rec) {
  if (s.lv) {
    rec = s.nbr[i];
    return make_float3(s.lv[i], s.lv[i + s.ls], s.lv[i + 2 * s.ls]);
  }
  const float4 q = s.verts[i];
  rec = __float_as_int(q.w);
  return make_float3(q.x, q.y, q.z);
}

// Copies a hull's vertices into LDS (SoA rows of stride ls), lanes lg, lg+G, ...
template <int G>
MFCL_D void stageHull(DevShape& s, MFCL_LDS float* dst, int ls, int lg) {
  for (int i = lg; i < s.nv; i += G) {
    const float4 q = s.verts[i];
    dst[i] = q.x;
    dst[i + ls] = q.y;
    dst[i + 2 * ls] = q.z;
  }
  s.lv = dst;
  s.ls = ls;
}

struct ShapeTables {
  const int* type;
  const float4* param;  // x, y, z as above; w = rbound for convex
  const uint32_t* vert_off;
  const uint32_t* vert_cnt;
  const float4* verts;
  const int* nbr;         // adjacency tables (mfcl_upload_convex_faces)
  const int64_t* nbr_off; // per shape, -1: naive scans; nullptr: no faces uploaded
  const int* seed;        // per shape: 8 ints, the walk's axis seed vertices
  const float4* wblk;     // walk blocks, eight float4 per vertex (indexed like verts)
};

MFCL_D DevShape loadShape(const ShapeTables& T, uint32_t i) {
  DevShape s;
  s.type = T.type[i];
  const float4 p = T.param[i];
  s.p0 = p.x;
  s.p1 = p.y;
  s.p2 = p.z;
  s.rbound = p.w;
  s.nbr = nullptr;
  s.seed = nullptr;
  s.wblk = nullptr;
  if (s.type == kConvex) {
    s.verts = T.verts + T.vert_off[i];
    s.nv = static_cast<int>(T.vert_cnt[i]);
    if (T.nbr_off) {
      const int64_t o = T.nbr_off[i];
      if (o >= 0) {
        s.nbr = T.nbr + o;
        s.seed = T.seed + 8 * static_cast<uint64_t>(i);
        if (T.wblk && s.nv <= 65535) s.wblk = T.wblk + 8 * static_cast<uint64_t>(T.vert_off[i]);
      }
    }
  } else {
    s.verts = nullptr;
    s.nv = 0;
  }
  s.lv = nullptr;
  s.ls = 0;
  return s;
}

// Support-point search over a hull, shared by the G lanes that own a pair
// (G = 1: one lane per pair).  Lane lg scans vertices lg, lg + G, ... in
// FP32 keeping the best and runner-up scores; a butterfly over the group
// merges them (ties resolve to the lower index).  When the FP32 gap is
// inside the rounding bound the group rescans in FP64, so the result is the
// first vertex attaining the FP64 maximum of dot(d, v) -- the reference's
// naive scan result (convex-inl.h:133-149), which the reference's adjacency
// walk also returns whenever that maximum is unique.
// Out of line: the scan is the largest piece of code on the support path
// and is reached from every support call site; one copy keeps the kernels'
// instruction footprint small.  Returns the index, with bit 30 set when the
// FP64 rescan ran.
template <int G>
__device__ __attribute__((noinline)) int hullArgmaxOOL(const float4* verts, const MFCL_LDS float* lv, int ls, int nv,
                                                      float rbound, double dx, double dy, double dz, int lg) {
  int bi = 0x7fffffff;
  float best = -INFINITY, second = -INFINITY;
  {
#pragma clang fp contract(fast)
    const float fx = static_cast<float>(dx), fy = static_cast<float>(dy), fz = static_cast<float>(dz);
    auto visit = [&](int i, float x, float y, float z) {
      const float val = fx * x + fy * y + fz * z;
      if (val > best) {
        second = best;
        best = val;
        bi = i;
      } else if (val > second) {
        second = val;
      }
    };
    if (lv) {
      for (int i = lg; i < nv; i += G) visit(i, lv[i], lv[i + ls], lv[i + 2 * ls]);
    } else {
      for (int i = lg; i < nv; i += G) {
        const float4 q = verts[i];
        visit(i, q.x, q.y, q.z);
      }
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    const float ob = __shfl_xor(best, off, G);
    const float os = __shfl_xor(second, off, G);
    const int oi = __shfl_xor(bi, off, G);
    if (ob > best || (ob == best && oi < bi)) {
      second = fmaxf(fmaxf(second, os), best);
      best = ob;
      bi = oi;
    } else {
      second = fmaxf(fmaxf(second, os), ob);
    }
  }
  // |fl32(d.v) - d.v| <= (rounding of d + 3 roundings of the dot) * |d|_1 max|v|
  // <= 8 * 2^-24 * |d|_1 * rbound; the gap test uses twice that.
  const float bound = 2.0f * 8.0f * 5.9604645e-08f * rbound *
                      (fabsf(static_cast<float>(dx)) + fabsf(static_cast<float>(dy)) + fabsf(static_cast<float>(dz)));
  // a NaN direction or vertex leaves no candidate: vertex 0 keeps every
  // read in bounds
  if (best - second > bound) return bi < nv ? bi : 0;
  double bd = -INFINITY;
  bi = 0x7fffffff;
  for (int i = lg; i < nv; i += G) {
    float x, y, z;
    if (lv) {
      x = lv[i]; y = lv[i + ls]; z = lv[i + 2 * ls];
    } else {
      const float4 q = verts[i];
      x = q.x; y = q.y; z = q.z;
    }
    const double val = dx * static_cast<double>(x) + dy * static_cast<double>(y) + dz * static_cast<double>(z);
    if (val > bd) {
      bd = val;
      bi = i;
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    const double ob = __shfl_xor(bd, off, G);
    const int oi = __shfl_xor(bi, off, G);
    if (ob > bd || (ob == bd && oi < bi)) {
      bd = ob;
      bi = oi;
    }
  }
  return (bi < nv ? bi : 0) | (1 << 30);
}

// convex-inl.h:202-290: the walk starts from the seed vertex of the axis
// direction (+x, -x, +y, -y, +z, -z) with the largest dot product with d
// (first on ties), then moves to the best strictly improving neighbour of the
// current vertex, skipping the two most recently left vertices, until no
// neighbour improves.  FP64 dot products as the reference's
// v_C.dot(vertices[n]).  Every step strictly increases the value, so the walk
// ends (after at most nv steps; at once for a NaN direction).
//
// Latency: a step reads the current vertex's neighbour record (its count and
// up to kWalkChunk ids in one batch; the tables are followed by 32 ints of
// padding, so the fixed-size read stays in bounds), then the neighbours'
// vertices -- each 16-byte vertex also carrying its own record offset -- in a
// second batch, and only then runs the reference's sequential comparisons in
// registers: two dependent memory round trips per step.  The two walks of one support query (one per
// shape) advance in lock step so their round trips overlap.
constexpr int kWalkChunk = 8;

struct Walk {
  d3 d;
  double ext_v;
  int ext, start, parent0, parent1, count, visited;
  bool keep;
  int id[kWalkChunk], st[kWalkChunk];
  float3 w[kWalkChunk];
};

// Walk starts: the reference seeds every walk from the best of six axis-
// direction caches (convex-inl.h:151-199); a start >= 0 replaces that seed
// with a given vertex (the caller's earlier support point for a nearby
// direction).  On a convex hull a walk ends where no neighbour improves the
// support value, which is the unique maximiser whatever the start; only
// exact ties between support values (a face orthogonal to the direction)
// can make the end depend on the start.
MFCL_D int walkSeed(const DevShape& s, d3 d) {
  int init = -1;
  double max_dot = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    // the seed directions of convex-inl.h:151-199: +-UnitX, +-UnitY, +-UnitZ
    const double sg = (i & 1) ? -1.0 : 1.0;
    const double ax = (i >> 1) == 0 ? sg : sg * 0.0;
    const double ay = (i >> 1) == 1 ? sg : sg * 0.0;
    const double az = (i >> 1) == 2 ? sg : sg * 0.0;
    const double dv = ax * d.x + ay * d.y + az * d.z;
    if (init < 0 || dv > max_dot) {
      init = s.seed[i];
      max_dot = dv;
    }
  }
  return init;
}

MFCL_D void walkInit(Walk& W, const DevShape& s, d3 d, bool active, int start = -1) {
  W.d = d;
  W.keep = active;
  W.visited = 0;
  W.ext = 0;
  W.ext_v = 0;
  W.start = 0;
  W.count = 0;
  if (!active) return;
  int init = start;
  double max_dot = 0;
  if (init < 0)
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    // the seed directions of convex-inl.h:151-199: +-UnitX, +-UnitY, +-UnitZ
    const double sg = (i & 1) ? -1.0 : 1.0;
    const double ax = (i >> 1) == 0 ? sg : sg * 0.0;
    const double ay = (i >> 1) == 1 ? sg : sg * 0.0;
    const double az = (i >> 1) == 2 ? sg : sg * 0.0;
    const double dv = ax * d.x + ay * d.y + az * d.z;
    if (init < 0 || dv > max_dot) {
      init = s.seed[i];
      max_dot = dv;
    }
  }
  W.ext = init;
  const float3 q = hullVertRec(s, init, W.start);
  W.ext_v = d.x * static_cast<double>(q.x) + d.y * static_cast<double>(q.y) + d.z * static_cast<double>(q.z);
  W.parent0 = W.parent1 = init;
  W.visited = 1;
}

// first batch: the neighbour record of the current vertex
MFCL_D void walkFetchRecord(Walk& W, const DevShape& s) {
  if (!W.keep) return;
  W.count = s.nbr[W.start];
#pragma unroll
  for (int j = 0; j < kWalkChunk; ++j) W.id[j] = s.nbr[W.start + 1 + j];
}
// second batch: the neighbours' vertices with their record offsets
MFCL_D void walkFetchNeighbours(Walk& W, const DevShape& s) {
  if (!W.keep) return;
#pragma unroll
  for (int j = 0; j < kWalkChunk; ++j) {
    const int v = j < W.count ? W.id[j] : W.ext;  // in range for the loads
    W.id[j] = v;
    W.w[j] = hullVertRec(s, v, W.st[j]);
  }
}
// the reference's comparisons over the record (neighbours beyond the first
// kWalkChunk, rare on hulls, are read here)
MFCL_D void walkCompare(Walk& W, const DevShape& s) {
  if (!W.keep) return;
  W.keep = false;
  const int old = W.ext;
  int next_start = W.start;
  for (int c0 = 0; c0 < W.count; c0 += kWalkChunk) {
    if (c0 > 0) {
#pragma unroll
      for (int j = 0; j < kWalkChunk; ++j) {
        const int nb = s.nbr[W.start + 1 + c0 + j];
        const int v = c0 + j < W.count ? nb : old;
        W.id[j] = v;
        W.w[j] = hullVertRec(s, v, W.st[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < kWalkChunk; ++j) {
      if (c0 + j >= W.count || W.id[j] == W.parent0 || W.id[j] == W.parent1) continue;
      const double v = W.d.x * static_cast<double>(W.w[j].x) + W.d.y * static_cast<double>(W.w[j].y) +
                       W.d.z * static_cast<double>(W.w[j].z);
      ++W.visited;
      if (v > W.ext_v) {
        W.parent1 = W.ext;
        W.keep = true;
        W.ext = W.id[j];
        W.ext_v = v;
        next_start = W.st[j];
      }
    }
  }
  W.parent0 = old;
  W.start = next_start;
}

// The walk over 128-byte blocks (s.wblk): a vertex's block holds its
// coordinates, its neighbour count, its first eight neighbour ids and those
// neighbours' coordinates, so a step reads one cache line -- the chosen
// neighbour's block -- instead of a record and up to eight vertices spread
// over the hull.  The comparisons, their order and the parent rule are
// walkCompare's; neighbours beyond the eighth come from the adjacency table.
struct BlkWalk {
  int cnt;
  uint32_t ids[4];
  float nx[8][3];
};
MFCL_D float3 blkLoad(const DevShape& s, int v, BlkWalk& B) {
  const float4* p = s.wblk + 8 * static_cast<uint64_t>(v);
  const float4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3], q4 = p[4], q5 = p[5], q6 = p[6], q7 = p[7];
  B.cnt = __float_as_int(q0.w);
  B.ids[0] = __float_as_uint(q1.x); B.ids[1] = __float_as_uint(q1.y);
  B.ids[2] = __float_as_uint(q1.z); B.ids[3] = __float_as_uint(q1.w);
  const float f[24] = {q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w, q4.x, q4.y, q4.z, q4.w,
                       q5.x, q5.y, q5.z, q5.w, q6.x, q6.y, q6.z, q6.w, q7.x, q7.y, q7.z, q7.w};
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    B.nx[j][0] = f[3 * j];
    B.nx[j][1] = f[3 * j + 1];
    B.nx[j][2] = f[3 * j + 2];
  }
  return make_float3(q0.x, q0.y, q0.z);
}

struct BlkState {
  BlkWalk B;
  d3 d;
  double ext_v;
  int ext, parent0, parent1, visited;
  bool keep;
};

MFCL_D void blkInit(BlkState& W, const DevShape& s, d3 d, bool active, int start = -1) {
  W.d = d;
  W.keep = active;
  W.visited = 0;
  W.ext = 0;
  W.ext_v = 0;
  W.B.cnt = 0;
  if (!active) return;
  int init = start;
  double max_dot = 0;
  if (init < 0)
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    // the seed directions of convex-inl.h:151-199: +-UnitX, +-UnitY, +-UnitZ
    const double sg = (i & 1) ? -1.0 : 1.0;
    const double ax = (i >> 1) == 0 ? sg : sg * 0.0;
    const double ay = (i >> 1) == 1 ? sg : sg * 0.0;
    const double az = (i >> 1) == 2 ? sg : sg * 0.0;
    const double dv = ax * d.x + ay * d.y + az * d.z;
    if (init < 0 || dv > max_dot) {
      init = s.seed[i];
      max_dot = dv;
    }
  }
  const float3 q = blkLoad(s, init, W.B);
  W.ext_v = d.x * static_cast<double>(q.x) + d.y * static_cast<double>(q.y) + d.z * static_cast<double>(q.z);
  W.ext = W.parent0 = W.parent1 = init;
  W.visited = 1;
}

MFCL_D void blkStep(BlkState& W, const DevShape& s) {
  if (!W.keep) return;
  W.keep = false;
  const int old = W.ext;
  const int count = W.B.cnt;
  const d3 d = W.d;
#pragma unroll
  for (int j = 0; j < kWalkChunk; ++j) {
    const int id = static_cast<int>((W.B.ids[j >> 1] >> (16 * (j & 1))) & 0xFFFFu);
    if (j >= count || id == W.parent0 || id == W.parent1) continue;
    const double v = d.x * static_cast<double>(W.B.nx[j][0]) + d.y * static_cast<double>(W.B.nx[j][1]) +
                     d.z * static_cast<double>(W.B.nx[j][2]);
    ++W.visited;
    if (v > W.ext_v) {
      W.parent1 = W.ext;
      W.keep = true;
      W.ext = id;
      W.ext_v = v;
    }
  }
  if (count > kWalkChunk) {
    const int start = __float_as_int(s.verts[old].w);
    for (int c0 = kWalkChunk; c0 < count; c0 += kWalkChunk) {
      int id[kWalkChunk];
      float3 w[kWalkChunk];
#pragma unroll
      for (int j = 0; j < kWalkChunk; ++j) {
        const int nb = s.nbr[start + 1 + c0 + j];
        id[j] = c0 + j < count ? nb : old;
        const float4 t = s.verts[id[j]];
        w[j] = make_float3(t.x, t.y, t.z);
      }
#pragma unroll
      for (int j = 0; j < kWalkChunk; ++j) {
        if (c0 + j >= count || id[j] == W.parent0 || id[j] == W.parent1) continue;
        const double v = d.x * static_cast<double>(w[j].x) + d.y * static_cast<double>(w[j].y) +
                         d.z * static_cast<double>(w[j].z);
        ++W.visited;
        if (v > W.ext_v) {
          W.parent1 = W.ext;
          W.keep = true;
          W.ext = id[j];
          W.ext_v = v;
        }
      }
    }
  }
  W.parent0 = old;
  if (W.keep) blkLoad(s, W.ext, W.B);
}

MFCL_D int hullWalkBlk(const DevShape& s, d3 d, int* scanned, int start = -1) {
  BlkState W;
  blkInit(W, s, d, true, start);
  while (W.keep) blkStep(W, s);
  *scanned += W.visited;
  return W.ext;
}

// The walk over blocks when the hull has them and is not staged in LDS,
// else over the adjacency table.
MFCL_D int hullWalk(const DevShape& s, d3 d, int* scanned, int start = -1) {
  if (s.wblk && !s.lv) return hullWalkBlk(s, d, scanned, start);
  Walk W;
  walkInit(W, s, d, true, start);
  while (W.keep) {
    walkFetchRecord(W, s);
    walkFetchNeighbours(W, s);
    walkCompare(W, s);
  }
  *scanned += W.visited;
  return W.ext;
}

// Two walks (either may be inactive) advanced together; ia / ib receive
// the extreme vertices of the active ones.
MFCL_D void hullWalk2(const DevShape& sa, d3 da, bool act_a, int& ia, const DevShape& sb, d3 db, bool act_b,
                      int& ib, int* scanned, int start_a = -1, int start_b = -1) {
  if (sa.wblk && !sa.lv && sb.wblk && !sb.lv) {
    BlkState A, B;
    blkInit(A, sa, da, act_a, start_a);
    blkInit(B, sb, db, act_b, start_b);
    while (A.keep || B.keep) {
      blkStep(A, sa);
      blkStep(B, sb);
    }
    if (act_a) ia = A.ext;
    if (act_b) ib = B.ext;
    *scanned += A.visited + B.visited;
    return;
  }
  Walk A, B;
  walkInit(A, sa, da, act_a, start_a);
  walkInit(B, sb, db, act_b, start_b);
  while (A.keep || B.keep) {
    walkFetchRecord(A, sa);
    walkFetchRecord(B, sb);
    walkFetchNeighbours(A, sa);
    walkFetchNeighbours(B, sb);
    walkCompare(A, sa);
    walkCompare(B, sb);
  }
  if (act_a) ia = A.ext;
  if (act_b) ib = B.ext;
  *scanned += A.visited + B.visited;
}

template <int G>
MFCL_D int hullArgmax(const DevShape& s, d3 d, int lg, int* scanned, int start = -1) {
  if (s.nbr) return hullWalk(s, d, scanned, start);
  const int r = hullArgmaxOOL<G>(s.verts, s.lv, s.ls, s.nv, s.rbound, d.x, d.y, d.z, lg);
  *scanned += (r >> 30) ? 2 * s.nv : s.nv;
  return r & ((1 << 30) - 1);
}

// Closed-form support points of the primitives
// (include/fcl/cvx_collide/gjk_shape.hpp:41-180); zero for a hull.
MFCL_D d3 primSupport(const DevShape& s, d3 dir) {
  switch (s.type) {
    case kBox:
      return {dir.x > 0 ? s.p0 / 2 : -s.p0 / 2, dir.y > 0 ? s.p1 / 2 : -s.p1 / 2, dir.z > 0 ? s.p2 / 2 : -s.p2 / 2};
    case kSphere:
      return dir * s.p0;
    case kCapsule: {
      const double r = s.p0, hh = s.p1 * 0.5;
      const d3 v = dir * r;
      const d3 a = d3{0, 0, hh} + v, b = d3{0, 0, -hh} + v;
      return dot(dir, a) > dot(dir, b) ? a : b;
    }
    case kCylinder: {
      const double r = s.p0, hh = s.p1 * 0.5;
      const double zd = sqrt(dir.x * dir.x + dir.y * dir.y);
      if (zd == 0.0) return {0, 0, dir.z > 0 ? hh : -hh};
      const double k = r / zd;
      return {k * dir.x, k * dir.y, dir.z > 0 ? hh : -hh};
    }
    case kTriangle: {
      // a mesh triangle (its three gathered vertices at verts), with the
      // tie rule of triangleSupport (shape_gjk_interface-inl.h:17-38)
      const float4 fa = s.verts[0], fb = s.verts[1], fc = s.verts[2];
      const d3 a = {fa.x, fa.y, fa.z}, b = {fb.x, fb.y, fb.z}, c = {fc.x, fc.y, fc.z};
      const double da = dot(dir, a), db = dot(dir, b), dc = dot(dir, c);
      if (da > db) return dc > da ? c : a;
      return dc > db ? c : b;
    }
    default:
      return {0, 0, 0};
  }
}

// include/fcl/cvx_collide/gjk_shape.hpp:41-180; shape_gjk_interface-inl.h:87-92.
template <int G>
MFCL_D d3 shapeSupport(const DevShape& s, d3 dir, int lg, int* scanned) {
  if (s.type == kConvex) {
    const float3 q = hullVert(s, hullArgmax<G>(s, dir, lg, scanned));
    return {q.x, q.y, q.z};
  }
  return primSupport(s, dir);
}

// shapeSupport plus the index of the hull vertex returned (-1 for primitives).
template <int G>
MFCL_D d3 shapeSupportI(const DevShape& s, d3 dir, int lg, int* scanned, int& idx, int start = -1) {
  if (s.type == kConvex) {
    idx = hullArgmax<G>(s, dir, lg, scanned, start);
    const float3 q = hullVert(s, idx);
    return {q.x, q.y, q.z};
  }
  idx = -1;
  return shapeSupport<G>(s, dir, lg, scanned);
}

MFCL_D d3 shapeInterior(const DevShape& s, const float4* interior_tab, uint32_t id) {
  if (s.type == kConvex) {
    const float4 c = interior_tab[id];
    return {c.x, c.y, c.z};
  }
  return {0, 0, 0};
}

// ---------------------------------------------------------------------------
// The same Minkowski difference with its configuration (both shapes, the two
// relative transforms, interior points) held in LDS instead of VGPRs.  The
// lanes that own a pair all write the same values to their slot and re-read
// them at every support call (a compiler-only memory fence stops the loads
// being hoisted into registers), which keeps ~80 FP64 registers out of the
// long-lived state of the GJK and EPA loops.
#define MFCL_LDS __attribute__((address_space(3)))

// Direction that produced a simplex vertex, and the hull vertex indices its
// support points came from (-1 for a primitive shape: recomputed from the
// direction when needed).
struct DirSlot {
  double d[3];
  int i0, i1;
};
constexpr int kDirSlots = 8;

struct MinkCfg {
  DirSlot slot[kDirSlots];
  double r1[9];  // frame 0 -> frame 1 rotation (toshape1)
  double r0[9];  // frame 1 -> frame 0 rotation (toshape0)
  double t0[3];  // frame 1 -> frame 0 translation
  double q0[3], q1[3];  // primitive parameters of shape 0 / 1
  double int0[3], int1[3];
  const float4* v0;
  const float4* v1;
  const MFCL_LDS float* lv0;
  const MFCL_LDS float* lv1;
  const int* nbr0;
  const int* nbr1;
  const int* seed0;
  const int* seed1;
  const float4* wblk0;
  const float4* wblk1;
  int ls0, ls1;
  int type0, type1, nv0, nv1;
  float rb0, rb1;
};

MFCL_D void reloadFence() { asm volatile("" ::: "memory"); }

// Wavefront-uniform copies of values that every lane holds identically.
template <class T>
MFCL_D T* uptr(T* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
  return reinterpret_cast<T*>((static_cast<uint64_t>(hi) << 32) | lo);
}
MFCL_D void uniformShape(DevShape& s) {
  s.type = __builtin_amdgcn_readfirstlane(s.type);
  s.nv = __builtin_amdgcn_readfirstlane(s.nv);
  s.verts = uptr(s.verts);
  s.nbr = uptr(s.nbr);
  s.seed = uptr(s.seed);
  s.wblk = uptr(s.wblk);
}

MFCL_D void storeM33(MFCL_LDS double* o, const m33& A) {
  o[0] = A.m00; o[1] = A.m01; o[2] = A.m02;
  o[3] = A.m10; o[4] = A.m11; o[5] = A.m12;
  o[6] = A.m20; o[7] = A.m21; o[8] = A.m22;
}
MFCL_D m33 loadM33(const MFCL_LDS double* o) {
  m33 A;
  A.m00 = o[0]; A.m01 = o[1]; A.m02 = o[2];
  A.m10 = o[3]; A.m11 = o[4]; A.m12 = o[5];
  A.m20 = o[6]; A.m21 = o[7]; A.m22 = o[8];
  return A;
}

MFCL_D void writeCfg(MFCL_LDS MinkCfg* c, const DevShape& a, const tf3& ta, const DevShape& b, const tf3& tb,
                     d3 int0, d3 int1) {
  const m33 r1 = tmulm(tb.R, ta.R);
  const tf3 t0 = relTf(ta, tb);
  storeM33(c->r1, r1);
  storeM33(c->r0, t0.R);
  c->t0[0] = t0.t.x; c->t0[1] = t0.t.y; c->t0[2] = t0.t.z;
  c->q0[0] = a.p0; c->q0[1] = a.p1; c->q0[2] = a.p2;
  c->q1[0] = b.p0; c->q1[1] = b.p1; c->q1[2] = b.p2;
  c->int0[0] = int0.x; c->int0[1] = int0.y; c->int0[2] = int0.z;
  c->int1[0] = int1.x; c->int1[1] = int1.y; c->int1[2] = int1.z;
  c->v0 = a.verts; c->v1 = b.verts;
  c->lv0 = a.lv; c->lv1 = b.lv;
  c->nbr0 = a.nbr; c->nbr1 = b.nbr;
  c->seed0 = a.seed; c->seed1 = b.seed;
  c->wblk0 = a.wblk; c->wblk1 = b.wblk;
  c->ls0 = a.ls; c->ls1 = b.ls;
  c->type0 = a.type; c->type1 = b.type;
  c->nv0 = a.nv; c->nv1 = b.nv;
  c->rb0 = a.rbound; c->rb1 = b.rbound;
}

MFCL_D d3 slotDir(const MFCL_LDS MinkCfg* c, int id) {
  reloadFence();
  const MFCL_LDS DirSlot& sl = c->slot[id & (kDirSlots - 1)];
  return d3{sl.d[0], sl.d[1], sl.d[2]};
}

template <int G>
struct MinkL {
  MFCL_LDS MinkCfg* c;
  int lg;
  int scanned;
  MFCL_D DevShape sh0() const {
    DevShape s;
    s.type = c->type0; s.p0 = c->q0[0]; s.p1 = c->q0[1]; s.p2 = c->q0[2];
    s.verts = c->v0; s.nv = c->nv0; s.rbound = c->rb0;
    s.lv = c->lv0; s.ls = c->ls0;
    s.nbr = c->nbr0; s.seed = c->seed0; s.wblk = c->wblk0;
    if constexpr (G == 64) uniformShape(s);  // one pair per wavefront
    return s;
  }
  MFCL_D DevShape sh1() const {
    DevShape s;
    s.type = c->type1; s.p0 = c->q1[0]; s.p1 = c->q1[1]; s.p2 = c->q1[2];
    s.verts = c->v1; s.nv = c->nv1; s.rbound = c->rb1;
    s.lv = c->lv1; s.ls = c->ls1;
    s.nbr = c->nbr1; s.seed = c->seed1; s.wblk = c->wblk1;
    if constexpr (G == 64) uniformShape(s);  // one pair per wavefront
    return s;
  }
  MFCL_D m33 rot1() const { return loadM33(c->r1); }
  MFCL_D tf3 tf0() const {
    tf3 t;
    t.R = loadM33(c->r0);
    t.t = d3{c->t0[0], c->t0[1], c->t0[2]};
    return t;
  }
  MFCL_D d3 support0(d3 d) {
    reloadFence();
    return shapeSupport<G>(sh0(), d, lg, &scanned);
  }
  MFCL_D d3 support1(d3 d) {
    reloadFence();
    const d3 x = shapeSupport<G>(sh1(), mul(rot1(), d), lg, &scanned);
    reloadFence();
    return apply(tf0(), x);
  }
  // support points of both shapes (vertex indices, -1 for a primitive);
  // two walking hulls walk together
  // st0 / st1: walk starts (-1: the reference's seed)
  MFCL_D void supportPair(d3 d, d3& a, d3& x, int& i0, int& i1, int st0 = -1, int st1 = -1) {
    reloadFence();
    const DevShape s0 = sh0(), s1 = sh1();
    const d3 d1 = mul(rot1(), -d);
    if (s0.nbr && s1.nbr) {
      hullWalk2(s0, d, true, i0, s1, d1, true, i1, &scanned, st0, st1);
      const float3 q0 = hullVert(s0, i0), q1 = hullVert(s1, i1);
      a = d3{q0.x, q0.y, q0.z};
      x = d3{q1.x, q1.y, q1.z};
      return;
    }
    a = shapeSupportI<G>(s0, d, lg, &scanned, i0, st0);
    reloadFence();
    x = shapeSupportI<G>(sh1(), d1, lg, &scanned, i1, st1);
  }
  // support(d) with walk starts; i0 / i1 receive the support vertices
  // (-1 for a primitive)
  MFCL_D d3 supportFrom(d3 d, int st0, int st1, int& i0, int& i1) {
    d3 a, x;
    supportPair(d, a, x, i0, i1, st0, st1);
    reloadFence();
    return a - apply(tf0(), x);
  }
  // support0(d) / support1(-d) from a known support vertex when there is one
  MFCL_D d3 point0At(int i0, d3 d) {
    if (i0 >= 0) {
      const float3 q = hullVert(sh0(), i0);
      return d3{q.x, q.y, q.z};
    }
    return support0(d);
  }
  MFCL_D d3 point1At(int i1, d3 d) {
    if (i1 >= 0) {
      const float3 q = hullVert(sh1(), i1);
      return apply(tf0(), d3{q.x, q.y, q.z});
    }
    return support1(-d);
  }
  MFCL_D d3 support(d3 d) {
    d3 a, x;
    int i0, i1;
    supportPair(d, a, x, i0, i1);
    reloadFence();
    return a - apply(tf0(), x);
  }
  // support(d), recording d and the support vertex indices in slot id.
  MFCL_D d3 supportNew(d3 d, int id) {
    int i0, i1;
    d3 a, x;
    supportPair(d, a, x, i0, i1);
    reloadFence();
    const d3 b = apply(tf0(), x);
    MFCL_LDS DirSlot& sl = c->slot[id & (kDirSlots - 1)];
    sl.d[0] = d.x; sl.d[1] = d.y; sl.d[2] = d.z;
    sl.i0 = i0; sl.i1 = i1;
    return a - b;
  }
  MFCL_D d3 dirOf(int id) {
    reloadFence();
    const MFCL_LDS DirSlot& sl = c->slot[id & (kDirSlots - 1)];
    return d3{sl.d[0], sl.d[1], sl.d[2]};
  }
  // support0(dir of slot id), from the recorded hull vertex when there is one
  MFCL_D d3 point0(int id) {
    reloadFence();
    const int i0 = c->slot[id & (kDirSlots - 1)].i0;
    if (i0 >= 0) {
      const float3 q = hullVert(sh0(), i0);
      return d3{q.x, q.y, q.z};
    }
    return support0(dirOf(id));
  }
  // support1(-dir of slot id), likewise
  MFCL_D d3 point1(int id) {
    reloadFence();
    const int i1 = c->slot[id & (kDirSlots - 1)].i1;
    if (i1 >= 0) {
      const float3 q = hullVert(sh1(), i1);
      return apply(tf0(), d3{q.x, q.y, q.z});
    }
    return support1(-dirOf(id));
  }
  MFCL_D d3 interior() {
    reloadFence();
    const d3 i1 = d3{c->int1[0], c->int1[1], c->int1[2]};
    return d3{c->int0[0], c->int0[1], c->int0[2]} - apply(tf0(), i1);
  }
};

// A simplex vertex during GJK: the Minkowski-difference point plus the id
// of the LDS slot holding the direction that produced it and the support
// vertex indices (the direction is only stored, never used by GJK's own
// arithmetic, so it stays out of registers).
struct MV {
  d3 v;
  int64_t id;  // 64-bit so the record has no padding: copies stay whole-field
               // (a 32-bit id plus padding is copied as one 64-bit word, which
               // SROA cannot split, leaving the simplex in private memory)
};

// Four fixed simplex slots; every access below uses a constant slot index.
struct Smp {
  MV a, b, c, e;  // slots 0..3
  int rank;
};

MFCL_D void mvZero(MV& m) {
  m.v.x = 0; m.v.y = 0; m.v.z = 0;
  m.id = 0;
}
MFCL_D void smpInit(Smp& s) {
  mvZero(s.a); mvZero(s.b); mvZero(s.c); mvZero(s.e);
  s.rank = -1;
}

// The same simplex with explicit directions: the EPA stage's input.
struct MVD {
  d3 v, d;
};
struct SmpD {
  MVD a, b, c, e;
  int rank;
};
MFCL_D void storeMVD(MVD& o, d3 v, d3 d) {
  o.v.x = v.x; o.v.y = v.y; o.v.z = v.z;
  o.d.x = d.x; o.d.y = d.y; o.d.z = d.z;
}

MFCL_D MV mvSel(bool c, const MV& x, const MV& y) {
  MV r;
  r.v = d3{c ? x.v.x : y.v.x, c ? x.v.y : y.v.y, c ? x.v.z : y.v.z};
  r.id = c ? x.id : y.id;
  return r;
}

// First direction slot not referenced by the simplex (at most 4 are).
MFCL_D int freeSlot(const Smp& s) {
  int used = 0;
  if (s.rank > 0) used |= 1 << static_cast<int>(s.a.id);
  if (s.rank > 1) used |= 1 << static_cast<int>(s.b.id);
  if (s.rank > 2) used |= 1 << static_cast<int>(s.c.id);
  if (s.rank > 3) used |= 1 << static_cast<int>(s.e.id);
  return __builtin_ctz(~used) & 7;
}
// Unconditional stores of selected values: conditional stores to different
// slots would be merged into one store through a pointer select, which
// forces the simplex out of registers into private memory.
MFCL_D void smpAdd(Smp& s, const MV& m) {
  const int r = s.rank < 0 ? 0 : s.rank;
  s.a = mvSel(r == 0, m, s.a);
  s.b = mvSel(r == 1, m, s.b);
  s.c = mvSel(r == 2, m, s.c);
  s.e = mvSel(r >= 3, m, s.e);
  s.rank = r + 1;
}
MFCL_D bool smpDup(const Smp& s, d3 v, double tol2) {
  if (s.rank > 0 && sqnorm(s.a.v - v) < tol2) return true;
  if (s.rank > 1 && sqnorm(s.b.v - v) < tol2) return true;
  if (s.rank > 2 && sqnorm(s.c.v - v) < tol2) return true;
  if (s.rank > 3 && sqnorm(s.e.v - v) < tol2) return true;
  return false;
}

enum GjkStatus : int { kGjkIntersect = 0, kGjkSeparated = 1, kGjkNoProgress = 2, kGjkIterLimit = 3, kGjkFailed = 4 };
enum ProjStatus : int { kProjFailed, kProjContinue, kProjIntersect, kProjZeroVolume };

// gjk.hpp:221-294 on slots (C=a, B=b, A=c).
MFCL_D int gjkProj3(Smp& s, d3& dir, double tol) {
  const MV C = s.a, B = s.b, A = s.c;
  const d3 a = A.v;
  const d3 ab = B.v - a, ac = C.v - a;
  const bool sb = dot(a, ab) >= 0, sc = dot(a, ac) >= 0;
  if (sb && sc) {
    s.rank = 1;
    s.a = A;
    dir = normalized(-a);
    return kProjContinue;
  }
  const d3 n = cross(ab, ac);
  if (dot(a, cross(n, ac)) <= 0) {
    s.b = A;
    s.rank = 2;
    dir = normalized(cross(cross(a, ac), ac));
    return kProjContinue;
  }
  if (dot(a, cross(ab, n)) <= 0) {
    s.a = B;
    s.b = A;
    s.rank = 2;
    dir = normalized(cross(cross(a, ab), ab));
    return kProjContinue;
  }
  const double area = norm(n);
  if (area < tol * tol) return kProjFailed;
  const d3 nu = n / area;
  const double nd = dot(nu, a);
  if (fabs(nd) < tol) return kProjIntersect;
  dir = nd <= 0 ? nu : -nu;
  return kProjContinue;
}

// gjk.hpp:181-363
MFCL_D int gjkProject(Smp& s, d3& dir, double tol) {
  if (s.rank == 2) {
    const MV B = s.a, A = s.b;
    const d3 a = A.v;
    const d3 a_to_b = B.v - a;
    const double ao_ab = -dot(a_to_b, a);
    const d3 abx = cross(a, a_to_b);
    if (ao_ab > 0 && sqnorm(abx) <= 0.0) return kProjIntersect;
    if (ao_ab <= 0) {
      s.rank = 1;
      s.a = A;
      dir = -normalized(a);
      return kProjContinue;
    }
    dir = normalized(cross(abx, a_to_b));
    return kProjContinue;
  }
  if (s.rank == 3) return gjkProj3(s, dir, tol);
  // rank 4: D=a, C=b, B=c, A=e
  const MV D = s.a, C = s.b, B = s.c, A = s.e;
  const d3 a = A.v;
  const d3 ab = B.v - a, ac = C.v - a, ad = D.v - a;
  d3 abc = cross(ab, ac), acd = cross(ac, ad), abd = cross(ab, ad);
  const double abc_ad = dot(abc, ad), acd_ab = dot(acd, ab), abd_ac = dot(abd, ac);
  if (fabs(abc_ad) <= 0.0) return kProjZeroVolume;
  if (abc_ad > 0) abc = -abc;
  if (acd_ab > 0) acd = -acd;
  if (abd_ac > 0) abd = -abd;
  const bool ds = dot(a, abc) > 0, cs = dot(a, abd) > 0, bs = dot(a, acd) > 0;
  if (ds && cs && bs) return kProjIntersect;
  if (!bs) {
    s.a = D; s.b = C; s.c = A;
  } else if (!cs) {
    s.a = D; s.b = B; s.c = A;
  } else {
    s.a = C; s.b = B; s.c = A;
  }
  s.rank = 3;
  return gjkProj3(s, dir, tol);
}

// ---------------------------------------------------------------- distance
// Sub-simplex outcomes are encoded as the ordered list of surviving input
// slots: bits [0,2) rank, then 2 bits per kept slot index, oldest first.
MFCL_D int code1(int a) { return 1 | (a << 2); }
MFCL_D int code2(int a, int b) { return 2 | (a << 2) | (b << 4); }
MFCL_D int code3(int a, int b, int c) { return 3 | (a << 2) | (b << 4) | (c << 6); }
MFCL_D int codeRank(int c) { return c & 3; }
MFCL_D int codeSlot(int c, int k) { return (c >> (2 + 2 * k)) & 3; }
// Re-express a code over local slots through a map local -> outer slot.
MFCL_D int codeMap(int c, int m0, int m1, int m2) {
  const int r = codeRank(c);
  int o = r;
  for (int k = 0; k < r; ++k) {
    const int l = codeSlot(c, k);
    const int g = l == 0 ? m0 : (l == 1 ? m1 : m2);
    o |= g << (2 + 2 * k);
  }
  return o;
}

// gjk_distance.hpp:128-154: segment (old x = slot 0, new y = slot 1).
MFCL_D int dist2(d3 x, d3 y, d3& out, double tol) {
  const d3 s1 = y, s2 = x;
  const d3 d = s2 - s1;
  const double l2 = sqnorm(d);
  const double t = -dot(s1, d);
  if (t <= 0 || l2 <= tol * tol) {
    out = s1;
    return code1(1);
  }
  if (t >= l2) {
    out = s2;
    return code1(0);
  }
  const double w2 = t / l2;
  out = w2 * s2 + (1.0 - w2) * s1;
  return code2(0, 1);
}

// gjk_distance.hpp:156-289: triangle (s3 = slot 0, s2 = slot 1, s1 new = slot 2).
MFCL_D int dist3(d3 x, d3 y, d3 z, d3& out, double tol) {
  const d3 s1 = z, s2 = y, s3 = x;
  const d3 s12 = s2 - s1, s13 = s3 - s1;
  const bool sep2 = dot(s1, s12) >= 0, sep3 = dot(s1, s13) >= 0;
  if (sep2 && sep3) {
    out = s1;
    return code1(2);
  }
  const d3 n = cross(s12, s13);
  const double area2 = sqnorm(n);
  const bool zero = area2 <= 0.0;
  const bool s12sep = dot(s1, cross(n, s12)) > 0;
  if (!sep2 && s12sep) return codeMap(dist2(y, z, out, tol), 1, 2, 0);
  const bool s13sep = dot(s1, cross(n, s13)) < 0;
  if (!sep3 && s13sep) return codeMap(dist2(x, z, out, tol), 0, 2, 0);
  double best = -1;
  d3 bpt = {0, 0, 0};
  int bc = 0;
  if (zero || s12sep) {
    d3 pt;
    const int c = codeMap(dist2(y, z, pt, tol), 1, 2, 0);
    const double dd = sqnorm(pt);
    if (best < 0 || dd < best) { best = dd; bpt = pt; bc = c; }
  }
  if (zero || s13sep) {
    d3 pt;
    const int c = codeMap(dist2(x, z, pt, tol), 0, 2, 0);
    const double dd = sqnorm(pt);
    if (best < 0 || dd < best) { best = dd; bpt = pt; bc = c; }
  }
  const bool s23sep = dot(s2, cross(n, s3 - s2)) > 0;
  if (zero || s23sep) {
    d3 pt;
    const int c = dist2(x, y, pt, tol);
    const double dd = sqnorm(pt);
    if (best < 0 || dd < best) { best = dd; bpt = pt; bc = c; }
  }
  if (best < 0) {
    out = n * (dot(s1, n) / area2);
    return code3(0, 1, 2);
  }
  out = bpt;
  return bc;
}

// Slot k of the simplex by value through selects (no indexed private array).
// Passes a value through an empty asm so the optimiser cannot turn a select
// between two simplex slots into a load from a computed slot address (which
// would pin the simplex in private memory).
MFCL_D void opaque(MV& m) {
  asm("" : "+v"(m.v.x), "+v"(m.v.y), "+v"(m.v.z), "+v"(m.id));
}
MFCL_D MV pick4(const MV& A, const MV& B, const MV& C, const MV& E, int k) {
  MV r = mvSel(k == 1, B, A);
  r = mvSel(k == 2, C, r);
  return mvSel(k == 3, E, r);
}

// The simplex becomes the vertices that `code` keeps of (A, B, C, E), the
// opaque copies of its slots.  Every slot is rewritten (slots beyond the
// rank hold don't-care values), so the original slots are dead once copied
// and the copies take their registers.
MFCL_D void smpApplyFrom(Smp& s, const MV& A, const MV& B, const MV& C, const MV& E, int code) {
  s.a = pick4(A, B, C, E, codeSlot(code, 0));
  s.b = pick4(A, B, C, E, codeSlot(code, 1));
  s.c = pick4(A, B, C, E, codeSlot(code, 2));
  s.e = E;
  s.rank = codeRank(code);
}

enum DistUpd : int { kNoImprovement, kUpdOK, kUpdFailed };

MFCL_D int distUpdate(Smp& s, d3& out, double tol) {
  if (s.rank == 1) {
    out = s.a.v;
    return kUpdOK;
  }
  if (s.rank != 2 && s.rank != 3 && s.rank != 4) return kUpdFailed;
  // rank 3: the triangle itself; rank 4 (gjk_distance.hpp:291-373): faces
  // (3,2,1), (3,2,0), (3,1,0) = slots (b,c,e), (a,c,e), (a,b,e).  One loop
  // body (a single inlined copy of the triangle case) with the inputs picked
  // by value selects.
  MV A = s.a, B = s.b, Cc = s.c, E = s.e;
  opaque(A); opaque(B); opaque(Cc); opaque(E);
  if (s.rank == 2) {
    smpApplyFrom(s, A, B, Cc, E, dist2(A.v, B.v, out, tol));
    return kUpdOK;
  }
  const bool r4 = s.rank == 4;
  const int nf = r4 ? 3 : 1;
  double best = -1;
  d3 bpt = {0, 0, 0};
  int bc = 0;
#pragma nounroll
  for (int f = 0; f < nf; ++f) {
    const MV x = r4 ? mvSel(f == 0, B, A) : A;
    const MV y = r4 ? mvSel(f == 2, B, Cc) : B;
    const MV z = r4 ? E : Cc;
    const int m0 = r4 ? (f == 0 ? 1 : 0) : 0;
    const int m1 = r4 ? (f == 2 ? 1 : 2) : 1;
    const int m2 = r4 ? 3 : 2;
    d3 pt;
    const int c = codeMap(dist3(x.v, y.v, z.v, pt, tol), m0, m1, m2);
    const double dd = sqnorm(pt);
    if (best < 0 || dd < best) { best = dd; bpt = pt; bc = c; }
  }
  if (best < 0) {
    s.a = A; s.b = B; s.c = Cc; s.e = E;  // unchanged, from the copies
    return kNoImprovement;
  }
  smpApplyFrom(s, A, B, Cc, E, bc);
  out = bpt;
  return kUpdOK;
}

constexpr double kBaryTol = 1e-3;  // gjk.h:125

// gjk_distance.hpp:375-486 (witness points re-evaluated from the stored
// directions, as the reference does), on the simplex vertices v1..v_rank.
template <class MK>
MFCL_D bool extractNoSub(MK& md, int rank, const MV& v1, const MV& v2, const MV& v3, d3& p0, d3& p1) {
  if (rank == 4 || rank <= 0) return false;
  if (rank == 1) {
    p0 = md.point0(v1.id);
    p1 = md.point1(v1.id);
    return true;
  }
  if (rank == 2) {
    const d3 d = v2.v - v1.v;
    const double l2 = sqnorm(d);
    if (l2 <= 0.0) {
      p0 = md.point0(v1.id);
      p1 = md.point1(v1.id);
      return true;
    }
    const double w2 = -dot(v1.v, d) / l2;
    if (w2 > 1 + kBaryTol) {
      p0 = md.point0(v2.id);
      p1 = md.point1(v2.id);
      return false;
    }
    if (w2 < -kBaryTol) {
      p0 = md.point0(v1.id);
      p1 = md.point1(v1.id);
      return false;
    }
    const double w1 = 1.0 - w2;
    p0 = md.point0(v1.id) * w1 + md.point0(v2.id) * w2;
    p1 = md.point1(v1.id) * w1 + md.point1(v2.id) * w2;
    return true;
  }
  const d3 s12 = v2.v - v1.v, s13 = v3.v - v1.v;
  const d3 n = cross(s12, s13);
  const double area2 = sqnorm(n);
  if (area2 <= 0.0) return false;
  const d3 proj = n * (dot(v1.v, n) / area2);
  const double area = sqrt(area2);
  const double w2 = norm(cross(s13, v1.v - proj)) / area;
  const double w3 = norm(cross(s12, v1.v - proj)) / area;
  const double w1 = 1.0 - w2 - w3;
  if (w1 < -kBaryTol || w2 < -kBaryTol || w3 < -kBaryTol) return false;
  p0 = md.point0(v1.id) * w1 + md.point0(v2.id) * w2 + md.point0(v3.id) * w3;
  p1 = md.point1(v1.id) * w1 + md.point1(v2.id) * w2 + md.point1(v3.id) * w3;
  return true;
}
template <class MK>
MFCL_D bool extractNoSub(MK& md, const Smp& s, d3& p0, d3& p1) {
  return extractNoSub(md, s.rank, s.a, s.b, s.c, p0, p1);
}

// gjk_distance.hpp:488-559
template <class MK>
MFCL_D bool extractTrySub(MK& md, const Smp& s, d3& p0, d3& p1) {
  if (s.rank >= 4) return false;
  if (s.rank <= 2) {
    extractNoSub(md, s, p0, p1);
    return true;
  }
  // candidates (gjk_distance.hpp:488-559): triangle (a,b,c) when its
  // extraction succeeds, then segments (a,b), (a,c), (b,c); one loop body
  double best = -1;
  d3 bp0 = {0, 0, 0}, bp1 = {0, 0, 0};
  MV A = s.a, B = s.b, Cc = s.c;
  opaque(A); opaque(B); opaque(Cc);
#pragma nounroll
  for (int k = 0; k < 4; ++k) {
    const MV v1 = mvSel(k == 3, B, A);
    const MV v2 = mvSel(k <= 1, B, Cc);
    d3 q0, q1;
    const bool ok = extractNoSub(md, k == 0 ? 3 : 2, v1, v2, Cc, q0, q1);
    if (k == 0 && !ok) continue;
    const double dd = sqnorm(q0 - q1);
    if (best < 0 || dd < best) { best = dd; bp0 = q0; bp1 = q1; }
  }
  p0 = bp0;
  p1 = bp1;
  return true;
}

// gjk_distance.hpp:11-105
template <class MK>
MFCL_D bool gjkMinDistance(MK& md, Smp& s, double tol, int max_iter, d3& p0, d3& p1, int* iters) {
  if (s.rank != 1) return false;
  d3 cur;
  if (distUpdate(s, cur, tol) != kUpdOK) return false;
  double book = norm(cur);
  // every successful exit extracts the witnesses from the current simplex:
  // one exit point, so one inlined copy of the extraction
  bool extract = book <= tol;
  if (!extract) {
    d3 nd = -cur / book;
    const double tol2 = tol * tol;
    for (int it = 0; it < max_iter; ++it) {
      (*iters)++;
      MV nv;
      nv.id = freeSlot(s);
      nv.v = md.supportNew(nd, nv.id);
      if (dot(nd, nv.v - cur) < tol) { extract = true; break; }
      if (smpDup(s, nv.v, tol2)) { extract = true; break; }
      smpAdd(s, nv);
      const int st = distUpdate(s, cur, tol);
      if (st == kNoImprovement) { extract = true; break; }
      if (st != kUpdOK) return false;
      const double nm = norm(cur);
      if (book - nm < tol) { extract = true; break; }
      if (nm < tol) { extract = true; break; }
      book = nm;
      nd = -cur / book;
    }
    if (!extract) return false;
  }
  return extractNoSub(md, s, p0, p1);
}

struct GjkOut {
  int status;
  bool sep_valid;
  d3 p0, p1;
  int iters;       // boolean iterations
  int dist_iters;  // distance iterations
};

// gjk.hpp:11-140.  want_distance selects the MinSeparationDistanceOutput path.
template <class MK>
MFCL_D void gjkEvaluate(MK& md, Smp& s, d3 guess, double tol, int max_iter, bool want_distance,
                        GjkOut& o) {
  o.sep_valid = false;
  o.iters = 0;
  o.dist_iters = 0;
  o.p0 = d3{0, 0, 0};
  o.p1 = d3{0, 0, 0};
  d3 dir = guess;
  if (sqnorm(dir) <= 0.0) dir = d3{1, 0, 0};
  dir = normalized(dir);
  MV v;
  v.id = 0;
  v.v = md.supportNew(dir, v.id);
  s.rank = -1;
  smpAdd(s, v);
  const double tol2 = tol * tol;
  if (sqnorm(v.v) <= tol2) {
    o.status = kGjkIntersect;
    return;
  }
  bool separated = dot(v.v, dir) < 0;
  int status = -1;
  if (!separated) {
    dir = -dir;
    status = kGjkIterLimit;
    for (int it = 0; it < max_iter; ++it) {
      o.iters = it + 1;
      v.id = freeSlot(s);
      v.v = md.supportNew(dir, v.id);
      if (dot(v.v, dir) < 0) {
        separated = true;
        break;
      }
      if (smpDup(s, v.v, tol2)) {
        status = kGjkNoProgress;
        break;
      }
      if (sqnorm(v.v) <= tol2) {
        smpAdd(s, v);
        o.status = kGjkIntersect;
        return;
      }
      smpAdd(s, v);
      const int p = gjkProject(s, dir, tol);
      if (p == kProjFailed) {
        o.status = kGjkFailed;
        return;
      }
      if (p == kProjIntersect) {
        o.status = kGjkIntersect;
        return;
      }
      if (p == kProjZeroVolume) {
        status = kGjkNoProgress;
        break;
      }
    }
  }
  if (separated) {
    o.status = kGjkSeparated;
    if (want_distance) {
      s.rank = -1;
      smpAdd(s, v);
      o.sep_valid = gjkMinDistance(md, s, tol, max_iter, o.p0, o.p1, &o.dist_iters);
    }
    return;
  }
  o.status = status;
  if (want_distance) o.sep_valid = extractTrySub(md, s, o.p0, o.p1);
}

// ------------------------------------------------------------------- MPR
// include/fcl/cvx_collide/mpr.hpp:34-185, findPortal :188-328, updatePortal :436-495.
enum MprStatus : int { kMprIntersect = 0, kMprSeparated = 1, kMprFailed = 2 };

template <class MK>
MFCL_D d3 mprSupport(MK& md, d3& d) {
  d = normalized(d);
  return md.support(d);
}

MFCL_D void swapd3(d3& a, d3& b) {
  const d3 t = a;
  a = b;
  b = t;
}

template <class MK>
MFCL_D int mprIntersect(MK& md, int max_iter, double tol) {
  constexpr double kDotEps = 2.220446049250313e-16;
  const d3 v0 = md.interior();
  if (sqnorm(v0) <= tol * tol) return kMprIntersect;
  d3 d1 = -v0;
  d3 v1 = mprSupport(md, d1);
  if (dot(d1, v1) < 0) return kMprSeparated;
  d3 d2 = cross(v0, v1);
  if (absnorm(d2) <= absnorm(v0) * absnorm(v1) * tol) return kMprIntersect;
  d3 v2 = mprSupport(md, d2);
  if (dot(d2, v2) < 0) return kMprSeparated;
  d3 d3v = cross(v1 - v0, v2 - v0);
  if (dot(d3v, v0) > 0) {
    swapd3(v1, v2);
    swapd3(d1, d2);
    d3v = -d3v;
  }
  d3 v3 = mprSupport(md, d3v);
  if (dot(d3v, v3) < 0) return kMprSeparated;
  const double v0n = absnorm(v0);
  int it = 0;
  while (true) {
    if (it >= max_iter) return kMprFailed;
    ++it;
    d3 a1 = v1 - v0, a2 = v2 - v0, a3 = v3 - v0;
    d3 n031 = cross(a3, a1), n012 = cross(a1, a2);
    if (dot(a2, n031) < 0) {
      swapd3(v2, v3);
      swapd3(d2, d3v);
      swapd3(a2, a3);
      swapd3(n012, n031);
      n031 = -n031;
      n012 = -n012;
    }
    if (dot(v0, n031) > kDotEps * v0n * absnorm(n031)) {
      d3 sd = -n031;
      v2 = mprSupport(md, sd);
      d2 = sd;
      if (dot(v2, sd) < 0) return kMprSeparated;
      continue;
    }
    if (dot(v0, n012) > kDotEps * v0n * absnorm(n012)) {
      d3 sd = -n012;
      v3 = mprSupport(md, sd);
      d3v = sd;
      if (dot(v3, sd) < 0) return kMprSeparated;
      continue;
    }
    const d3 n023 = cross(a2, a3);
    if (dot(v0, n023) > kDotEps * v0n * absnorm(n023)) {
      d3 sd = -n023;
      v1 = mprSupport(md, sd);
      d1 = sd;
      if (dot(v1, sd) < 0) return kMprSeparated;
      continue;
    }
    break;
  }
  for (it = 0; it < max_iter; ++it) {
    d3 n = cross(v2 - v1, v3 - v1);
    if (dot(n, v0) > 0) {
      swapd3(v2, v3);
      swapd3(d2, d3v);
      n = -n;
    }
    if (!(dot(v1, n) < 0)) return kMprIntersect;
    const d3 v4 = mprSupport(md, n);
    if (dot(v4, n) < 0) return kMprSeparated;
    if (fabs(dot(v4 - v1, n)) < tol * absnorm(n)) return kMprSeparated;
    const d3 nn = cross(v4, v0);
    if (dot(v1, nn) > 0) {
      if (dot(v2, nn) > 0) { v1 = v4; d1 = n; }
      else { v3 = v4; d3v = n; }
    } else {
      if (dot(v3, nn) > 0) { v2 = v4; d2 = n; }
      else { v1 = v4; d1 = n; }
    }
  }
  return kMprFailed;
}

// MPR penetration along a unit direction d (shape-0 frame): the distance
// along d from the origin to where the ray leaves the Minkowski difference,
// by a portal search from a mock start -d (MPR::RunDirectedPenetration,
// mpr.hpp:496-632), and the witness points as the barycentric blend of the
// final portal's support points (finalizeDirectedPenetrationResult,
// mpr.hpp:642-704).  Status codes follow MPR::DirectedPenetrationStatus.
enum DirStatus : int { kDirOK = 0, kDirFailed = 1, kDirNoIntersect = 2, kDirIterLimit = 3 };
struct DirOut {
  double distance;
  d3 p0, p1;
  int iters;
};

template <class MK>
MFCL_D void mprDirectedResult(MK& md, d3 d, d3 n, d3 v1, d3 v2, d3 v3, d3 d1, d3 d2, d3 d3v, const d3* v4,
                              DirOut& o) {
  const double nd = dot(d, n);
  if (fabs(nd) <= 0) {
    const double e1 = dot(v1, d), e2 = dot(v2, d), e3 = dot(v3, d);
    double best = -INFINITY;
    d3 bd = d1;
    if (e1 > best) { best = e1; bd = d1; }
    if (e2 > best) { best = e2; bd = d2; }
    if (e3 > best) { best = e3; bd = d3v; }
    o.distance = best;
    o.p0 = md.support0(bd);
    o.p1 = md.support1(-bd);
    return;
  }
  const double to_portal = dot(v1, n) / nd;
  o.distance = v4 ? dot(*v4, n) / nd : to_portal;
  const d3 q = to_portal * d;
  const d3 e12 = v2 - v1, e13 = v3 - v1;
  const double area = norm(cross(e12, e13));
  const double w2 = norm(cross(e13, v1 - q)) / area;
  const double w3 = norm(cross(e12, v1 - q)) / area;
  const double w1 = 1.0 - w2 - w3;
  o.p0 = md.support0(d1) * w1 + md.support0(d2) * w2 + md.support0(d3v) * w3;
  o.p1 = md.support1(-d1) * w1 + md.support1(-d2) * w2 + md.support1(-d3v) * w3;
}

template <class MK>
MFCL_D int mprDirected(MK& md, d3 d, int max_iter, double tol, DirOut& o) {
  constexpr double kDotEps = 2.220446049250313e-16;
  o.iters = 0;
  const d3 v0 = -d;
  d3 d1 = d;
  d3 v1 = mprSupport(md, d1);
  if (dot(d1, v1) < 0) return kDirNoIntersect;
  d3 d2 = cross(v0, v1);
  if (absnorm(d2) <= absnorm(v1) * tol) {
    o.distance = dot(v1, d);
    o.p0 = md.support0(d1);
    o.p1 = md.support1(-d1);
    return kDirOK;
  }
  d3 v2 = mprSupport(md, d2);
  if (dot(d2, v2) < 0) return kDirNoIntersect;
  d3 d3v = cross(v1, v2);
  if (dot(d3v, v0) > 0) {
    swapd3(v1, v2);
    swapd3(d1, d2);
    d3v = d3v * -1.0;
  }
  d3 v3 = mprSupport(md, d3v);
  if (dot(d3v, v3) < 0) return kDirNoIntersect;
  const double m2 = fmax(sqnorm(v1), fmax(sqnorm(v2), sqnorm(v3)));
  const d3 v0s = v0 * sqrt(m2);
  // the portal search of mprIntersect, from v0s
  const double v0n = absnorm(v0s);
  int it = 0;
  while (true) {
    if (it >= max_iter) return kDirFailed;
    ++it;
    d3 a1 = v1 - v0s, a2 = v2 - v0s, a3 = v3 - v0s;
    d3 n031 = cross(a3, a1), n012 = cross(a1, a2);
    if (dot(a2, n031) < 0) {
      swapd3(v2, v3);
      swapd3(d2, d3v);
      swapd3(a2, a3);
      swapd3(n012, n031);
      n031 = -n031;
      n012 = -n012;
    }
    if (dot(v0s, n031) > kDotEps * v0n * absnorm(n031)) {
      d3 sd = -n031;
      v2 = mprSupport(md, sd);
      d2 = sd;
      if (dot(v2, sd) < 0) return kDirNoIntersect;
      continue;
    }
    if (dot(v0s, n012) > kDotEps * v0n * absnorm(n012)) {
      d3 sd = -n012;
      v3 = mprSupport(md, sd);
      d3v = sd;
      if (dot(v3, sd) < 0) return kDirNoIntersect;
      continue;
    }
    const d3 n023 = cross(a2, a3);
    if (dot(v0s, n023) > kDotEps * v0n * absnorm(n023)) {
      d3 sd = -n023;
      v1 = mprSupport(md, sd);
      d1 = sd;
      if (dot(v1, sd) < 0) return kDirNoIntersect;
      continue;
    }
    break;
  }
  for (it = 1; it <= max_iter; ++it) {
    o.iters = it;
    d3 n = cross(v2 - v1, v3 - v1);
    if (dot(n, d) < 0) {
      swapd3(v2, v3);
      swapd3(d2, d3v);
      n = n * -1.0;
    }
    const d3 v4 = mprSupport(md, n);
    if (dot(v4, n) < 0) return kDirNoIntersect;
    if (fabs(dot(v4 - v1, n)) < tol * absnorm(n)) {
      mprDirectedResult(md, d, n, v1, v2, v3, d1, d2, d3v, &v4, o);
      return kDirOK;
    }
    const d3 nn = cross(v4, v0s);
    if (dot(v1, nn) > 0) {
      if (dot(v2, nn) > 0) { v1 = v4; d1 = n; }
      else { v3 = v4; d3v = n; }
    } else {
      if (dot(v3, nn) > 0) { v2 = v4; d2 = n; }
      else { v1 = v4; d1 = n; }
    }
  }
  return kDirIterLimit;
}

// MPR minimum penetration refined from a unit hint direction (shape-0
// frame): directed portal searches that switch to the current portal's
// normal whenever the portal proves that direction shorter, until a search
// converges (MPR::RunIncrementalMinimumPenetrationDistance with
// return_on_subroutine_converge, mpr_incremental_penetration.hpp:10-337).
// Status codes follow MPR::IncrementalPenetrationStatus.
enum IncStatus : int { kIncOK = 0, kIncFailed = 1, kIncNoIntersect = 2, kIncIterLimit = 4 };
enum IncExplore : int { kExNewDir = 0, kExConverged = 1, kExDegenerated = 2, kExFailed = 3, kExNoIntersect = 4 };
struct IncOut {
  double depth;
  d3 dir, p0, p1;
  int iters;
};

template <class MK>
MFCL_D void mprIncResult(MK& md, d3 d, d3 v1, d3 v2, d3 v3, d3 d1, d3 d2, d3 d3v, double lb, double ub, IncOut& o) {
  o.depth = ub;
  o.dir = d;
  const d3 q = lb * d;
  const d3 e12 = v2 - v1, e13 = v3 - v1;
  const double area = norm(cross(e12, e13));
  const double w2 = norm(cross(e13, v1 - q)) / area;
  const double w3 = norm(cross(e12, v1 - q)) / area;
  const double w1 = 1.0 - w2 - w3;
  o.p0 = md.support0(d1) * w1 + md.support0(d2) * w2 + md.support0(d3v) * w3;
  o.p1 = md.support1(-d1) * w1 + md.support1(-d2) * w2 + md.support1(-d3v) * w3;
}

template <class MK>
MFCL_D int mprExplore(MK& md, d3 d, d3& v1, d3& v2, d3& v3, d3& d1, d3& d2, d3& d3v, double& lb, double& ub,
                      d3& new_d, bool portal_valid, int max_iter, double tol) {
  constexpr double kDotEps = 2.220446049250313e-16;
  const d3 v0 = -d;
  d3 sd0 = d;
  const d3 s0 = mprSupport(md, sd0);
  if (dot(d, s0) < 0) return kExNoIntersect;
  if (!portal_valid) {
    d1 = d;
    v1 = s0;
    d2 = cross(v0, v1);
    if (absnorm(d2) <= absnorm(v1) * tol) {
      lb = dot(v1, d);
      ub = lb;
      new_d = d;
      return kExDegenerated;
    }
    v2 = mprSupport(md, d2);
    if (dot(d2, v2) < 0) return kExNoIntersect;
    d3v = cross(v1, v2);
    if (dot(d3v, v0) > 0) {
      swapd3(v1, v2);
      swapd3(d1, d2);
      d3v = d3v * -1.0;
    }
    v3 = mprSupport(md, d3v);
    if (dot(d3v, v3) < 0) return kExNoIntersect;
  }
  const double m2 = fmax(sqnorm(v1), fmax(sqnorm(v2), sqnorm(v3)));
  const d3 v0s = v0 * sqrt(m2);
  const double v0n = absnorm(v0s);
  int it = 0;
  while (true) {
    if (it >= max_iter) return kExFailed;
    ++it;
    d3 a1 = v1 - v0s, a2 = v2 - v0s, a3 = v3 - v0s;
    d3 n031 = cross(a3, a1), n012 = cross(a1, a2);
    if (dot(a2, n031) < 0) {
      swapd3(v2, v3);
      swapd3(d2, d3v);
      swapd3(a2, a3);
      swapd3(n012, n031);
      n031 = -n031;
      n012 = -n012;
    }
    if (dot(v0s, n031) > kDotEps * v0n * absnorm(n031)) {
      d3 sd = -n031;
      v2 = mprSupport(md, sd);
      d2 = sd;
      if (dot(v2, sd) < 0) return kExNoIntersect;
      continue;
    }
    if (dot(v0s, n012) > kDotEps * v0n * absnorm(n012)) {
      d3 sd = -n012;
      v3 = mprSupport(md, sd);
      d3v = sd;
      if (dot(v3, sd) < 0) return kExNoIntersect;
      continue;
    }
    const d3 n023 = cross(a2, a3);
    if (dot(v0s, n023) > kDotEps * v0n * absnorm(n023)) {
      d3 sd = -n023;
      v1 = mprSupport(md, sd);
      d1 = sd;
      if (dot(v1, sd) < 0) return kExNoIntersect;
      continue;
    }
    break;
  }
  for (it = 0; it < max_iter; ++it) {
    d3 n = cross(v2 - v1, v3 - v1);
    if (dot(n, d) < 0) {
      swapd3(v2, v3);
      swapd3(d2, d3v);
      n = n * -1.0;
    }
    const d3 v4 = mprSupport(md, n);
    const double ub_plane = dot(v4, n);
    if (ub_plane < 0) return kExNoIntersect;
    const double v1n = dot(v1, n), v4n = dot(v4, n);
    const double gap = v4n - v1n;
    const double nd = dot(d, n);
    if (fabs(nd) <= 0) {
      lb = fmax(dot(v1, d), fmax(dot(v2, d), dot(v3, d)));
      ub = lb;
    } else {
      lb = v1n / nd;
      ub = v4n / nd;
    }
    if (lb > ub_plane + tol) {
      new_d = n;
      return kExNewDir;
    }
    if (fabs(gap) <= tol) {
      new_d = n;
      return kExConverged;
    }
    const d3 nn = cross(v4, v0s);
    if (dot(v1, nn) > 0) {
      if (dot(v2, nn) > 0) { v1 = v4; d1 = n; }
      else { v3 = v4; d3v = n; }
    } else {
      if (dot(v3, nn) > 0) { v2 = v4; d2 = n; }
      else { v1 = v4; d1 = n; }
    }
  }
  return kExFailed;
}

template <class MK>
MFCL_D int mprIncremental(MK& md, d3 init, int max_iter, double tol, IncOut& o) {
  d3 d = init;
  d3 v1 = {0, 0, 0}, v2 = v1, v3 = v1, d1 = v1, d2 = v1, d3v = v1;
  double prev_lb = -1, prev_ub = -1;
  o.iters = 0;
  for (int outer = 0; outer < max_iter; ++outer) {
    o.iters = outer + 1;
    d3 new_d = {0, 0, 0};
    double lb = 0, ub = 0;
    const int e = mprExplore(md, d, v1, v2, v3, d1, d2, d3v, lb, ub, new_d, outer >= 1, max_iter, tol);
    if (e == kExNewDir) {
      d = new_d;
    } else if (e == kExConverged) {
      mprIncResult(md, d, v1, v2, v3, d1, d2, d3v, lb, ub, o);
      return kIncOK;
    } else if (e == kExDegenerated) {
      o.depth = ub;
      o.dir = new_d;
      o.p0 = md.support0(new_d);
      o.p1 = md.support1(-new_d);
      return kIncOK;
    } else if (e == kExNoIntersect) {
      return kIncNoIntersect;
    } else {
      return kIncFailed;
    }
    prev_lb = lb;
    prev_ub = ub;
  }
  mprIncResult(md, d, v1, v2, v3, d1, d2, d3v, prev_lb, prev_ub, o);
  return kIncIterLimit;
}

}  // namespace mfcl
