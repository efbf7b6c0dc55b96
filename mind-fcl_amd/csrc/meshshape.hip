// Mesh-shape collision (fcl::collide(BVHModel<OBBRSS>, shape) with
// penetration disabled): OrientedNodeBVHSolver::MeshShapeIntersect
// (include/fcl/narrowphase/detail/traversal/collision/bvh_solver-inl.h:6-72)
// tests the mesh's tree against the shape's volume and every reached leaf
// triangle against the shape (ShapeSimplexIntersect, shape_pair_intersect-
// inl.h:133-172 -> GJKSolver::shapeTriangleIntersect, gjk_solver-inl.h:
// 545-563: sphereTriangleIntersect, boxTriangleIntersect, else MPR on the
// shape and the triangle).
//
// Three launches: (1) one wavefront per query walks the mesh's OBB tree
// against the shape's box (its frame, FP32 test with the conservative
// margin of the mesh-mesh traversal), 64 nodes per step off an LDS stack,
// and appends every reached leaf as a (query, triangle) candidate; (2) 16
// lanes per candidate run the leaf test in FP64 exactly as the reference
// orders it, counting and recording colliding triangles; (3) one lane per
// query writes the result and sorts its kept triangle ids.  The set of
// colliding triangles does not depend on the tree, so booleans and complete
// sets equal the reference's; which ones are kept when more than
// max_contacts collide follows the traversal here, not the reference's
// depth-first order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "ctx.h"
#include "epa_dev.h"
#include "narrow_dev.h"
#include "obb_dev.h"

namespace mfcl {

constexpr int kMsStack = 2048;  // node ids per wavefront (LDS)
constexpr int kMsGroup = 16;    // lanes per leaf test

struct MeshShapeArgs {
  ShapeTables T;
  const double4* interior;
  const ObbNode* nodes;
  const float4* tris;
  const uint32_t* node_base;
  const uint32_t* tri_base;
  const uint32_t* mesh_ids;
  const uint32_t* shape_ids;
  const float* pose_m;
  const float* pose_s;
  uint64_t n;
  uint64_t* cand;
  uint64_t cand_cap;
  unsigned long long* cand_count;
  uint32_t* counts;    // per query: colliding triangles
  uint32_t* leaves;    // per query: candidate leaves
  uint32_t* bv_tests;  // per query: OBB tests
  uint32_t* overflow;  // per query: 1 when the LDS stack would overflow
  int max_contacts;
  int32_t* tri_out;
  int gjk_max_it;
  double gjk_tol;
};

// The shape's box in its own frame: half extents that contain it.
MFCL_D ObbNode shapeBox(int type, float4 p) {
  ObbNode b;
  for (int k = 0; k < 9; ++k) b.ax[k] = (k % 4 == 0) ? 1.f : 0.f;
  b.c[0] = b.c[1] = b.c[2] = 0.f;
  float e0, e1, e2;
  if (type == kBox) { e0 = 0.5f * p.x; e1 = 0.5f * p.y; e2 = 0.5f * p.z; }
  else if (type == kSphere) { e0 = e1 = e2 = p.x; }
  else if (type == kCapsule) { e0 = e1 = p.x; e2 = 0.5f * p.y + p.x; }
  else if (type == kCylinder) { e0 = e1 = p.x; e2 = 0.5f * p.y; }
  else { e0 = e1 = e2 = p.w; }  // convex: max |coordinate|
  b.e[0] = e0; b.e[1] = e1; b.e[2] = e2;
  b.child = 0;
  return b;
}

__global__ void __launch_bounds__(64) msCandKernel(MeshShapeArgs A) {
  __shared__ int32_t stk[kMsStack];
  const uint64_t q = blockIdx.x;
  if (q >= A.n) return;
  const int l = threadIdx.x;
  const uint32_t mesh = A.mesh_ids[q], shp = A.shape_ids[q];
  const ObbNode sb = shapeBox(A.T.type[shp], A.T.param[shp]);
  const tf3 tm = poseFromQuat(A.pose_m + 7 * q), ts = poseFromQuat(A.pose_s + 7 * q);
  const tf3 rel = relTf(tm, ts);  // shape frame in the mesh frame
  const float Rq[9] = {static_cast<float>(rel.R.m00), static_cast<float>(rel.R.m01), static_cast<float>(rel.R.m02),
                       static_cast<float>(rel.R.m10), static_cast<float>(rel.R.m11), static_cast<float>(rel.R.m12),
                       static_cast<float>(rel.R.m20), static_cast<float>(rel.R.m21), static_cast<float>(rel.R.m22)};
  const float Tq[3] = {static_cast<float>(rel.t.x), static_cast<float>(rel.t.y), static_cast<float>(rel.t.z)};
  const ObbNode* nodes = A.nodes + A.node_base[mesh];
  const unsigned long long below = (1ull << l) - 1;
  if (l == 0) stk[0] = 0;
  int top = 1;
  uint32_t tests = 0, leaves = 0;
  bool ovf = false;
  __builtin_amdgcn_wave_barrier();
  while (top > 0) {
    const int take = top < 64 ? top : 64;
    const int base = top - take;
    const int node = l < take ? stk[base + l] : -1;
    __builtin_amdgcn_wave_barrier();
    top = base;
    bool leaf = false, inner = false;
    int child = 0;
    if (node >= 0) {
      const ObbNode a = nodes[node];
      ++tests;
      if (obbOverlapF(Rq, Tq, a, sb, kObbScaleErr)) {
        child = a.child;
        leaf = child < 0;
        inner = child >= 0;
      }
    }
    const unsigned long long lm = __ballot(leaf);
    if (lm) {
      unsigned long long b0 = 0;
      if (l == 0) b0 = atomicAdd(A.cand_count, static_cast<unsigned long long>(__popcll(lm)));
      b0 = __shfl(b0, 0);
      if (leaf) {
        const uint64_t slot = b0 + __popcll(lm & below);
        if (slot < A.cand_cap) A.cand[slot] = (q << 32) | static_cast<uint32_t>(-(child + 1));
        ++leaves;
      }
    }
    const unsigned long long pm = __ballot(inner);
    const int total = 2 * __popcll(pm);
    if (top + total > kMsStack) {  // uniform: every lane sees the same counts
      ovf = true;
      break;
    }
    if (inner) {
      const int r = 2 * __popcll(pm & below);
      stk[top + r] = child;
      stk[top + r + 1] = child + 1;
    }
    top += total;
    __builtin_amdgcn_wave_barrier();
  }
  for (int off = 32; off > 0; off >>= 1) {
    tests += __shfl_xor(tests, off);
    leaves += __shfl_xor(leaves, off);
  }
  if (l == 0) {
    A.bv_tests[q] = tests;
    A.leaves[q] = leaves;
    A.overflow[q] = ovf ? 1u : 0u;
  }
}

#pragma clang fp contract(off)
// sphere_triangle-inl.h:52-74
MFCL_D double segSqrDist(d3 from, d3 to, d3 p, d3& nearest) {
  d3 diff = p - from;
  const d3 v = to - from;
  double t = dot(v, diff);
  if (t > 0) {
    const double vv = dot(v, v);
    if (t < vv) {
      t /= vv;
      diff = diff - v * t;
    } else {
      t = 1;
      diff = diff - v;
    }
  } else {
    t = 0;
  }
  nearest = from + v * t;
  return dot(diff, diff);
}

// sphere_triangle-inl.h:107-187: sphere of radius r centred at c against
// the world triangle (P1, P2, P3).  c_out (optional) receives the contact as
// the reference writes it: pos = the nearest point found, normal = that
// point minus the centre, normalised (the flipped plane normal when they
// coincide), depth = -(r - distance) (-r when they coincide).
struct SphereTriContact {
  d3 pos, normal;
  double depth;
};
MFCL_D bool sphereTriContact(double r, d3 c, d3 P1, d3 P2, d3 P3, SphereTriContact* c_out) {
  d3 n = cross(P2 - P1, P3 - P1);
  const double nn = norm(n);
  if (nn > 0) n = n / nn;
  const double rt = r + 2.220446049250313e-16;
  double dist = dot(c - P1, n);
  if (dist < 0) {
    dist *= -1;
    n = n * -1.0;
  }
  if (!(dist < rt)) return false;
  bool has = false;
  d3 cp = {0, 0, 0};
  const double r1 = dot(cross(P2 - P1, n), c - P1);
  const double r2 = dot(cross(P3 - P2, n), c - P2);
  const double r3 = dot(cross(P1 - P3, n), c - P3);
  if ((r1 > 0 && r2 > 0 && r3 > 0) || (r1 <= 0 && r2 <= 0 && r3 <= 0)) {
    has = true;
    cp = c - n * dist;
  } else {
    const double cr2 = rt * rt;
    d3 q;
    if (segSqrDist(P1, P2, c, q) < cr2) { has = true; cp = q; }
    if (segSqrDist(P2, P3, c, q) < cr2) { has = true; cp = q; }
    if (segSqrDist(P3, P1, c, q) < cr2) { has = true; cp = q; }
  }
  if (!has) return false;
  const d3 d = cp - c;
  const double d2 = dot(d, d);
  if (!(d2 < rt * rt)) return false;
  if (c_out) {
    c_out->pos = cp;
    if (d2 > 0) {
      c_out->normal = normalized(d);
      c_out->depth = -(r - sqrt(d2));
    } else {
      c_out->normal = n * -1.0;
      c_out->depth = -r;
    }
  }
  return true;
}
MFCL_D bool sphereTriHit(double r, d3 c, d3 P1, d3 P2, d3 P3) { return sphereTriContact(r, c, P1, P2, P3, nullptr); }

// box_triangle-inl.h:9-69
MFCL_D bool sepEdgeAxes(const double h[3], d3 e0, d3 v0, d3 v1) {
  const double ax = fabs(e0.x), ay = fabs(e0.y), az = fabs(e0.z);
  auto sep = [](double p0, double p1, double rad) {
    if (p0 < p1) return p0 > rad || p1 < -rad;
    return p1 > rad || p0 < -rad;
  };
  if (sep(v0.y * e0.z - v0.z * e0.y, v1.y * e0.z - v1.z * e0.y, h[1] * az + h[2] * ay)) return true;
  if (sep(v0.x * e0.z - v0.z * e0.x, v1.x * e0.z - v1.z * e0.x, h[0] * az + h[2] * ax)) return true;
  if (sep(v0.x * e0.y - v0.y * e0.x, v1.x * e0.y - v1.y * e0.x, h[0] * ay + h[1] * ax)) return true;
  return false;
}

MFCL_D void minMax3(double a, double b, double c, double& mn, double& mx) {
  mn = mx = a;
  if (b < mn) mn = b;
  if (b > mx) mx = b;
  if (c < mn) mn = c;
  if (c > mx) mx = c;
}

// box_triangle-inl.h:72-137: box of full sides s posed ts against the
// triangle (P1, P2, P3) posed tm.
MFCL_D bool boxTriHit(double s0, double s1, double s2, const tf3& ts, const tf3& tm, d3 P1, d3 P2, d3 P3) {
  const tf3 t = relTf(ts, tm);
  const d3 v0 = apply(t, P1), v1 = apply(t, P2), v2 = apply(t, P3);
  const double h[3] = {0.5 * s0, 0.5 * s1, 0.5 * s2};
  double mn, mx;
  minMax3(v0.x, v1.x, v2.x, mn, mx);
  if (mx < -h[0] || mn > h[0]) return false;
  minMax3(v0.y, v1.y, v2.y, mn, mx);
  if (mx < -h[1] || mn > h[1]) return false;
  minMax3(v0.z, v1.z, v2.z, mn, mx);
  if (mx < -h[2] || mn > h[2]) return false;
  const d3 e0 = v1 - v0;
  if (sepEdgeAxes(h, e0, v0, v2)) return false;
  const d3 e1 = v2 - v1;
  if (sepEdgeAxes(h, e1, v1, v0)) return false;
  const d3 e2 = v0 - v2;
  if (sepEdgeAxes(h, e2, v0, v1)) return false;
  const d3 n = cross(e0, e1);
  const double nv[3] = {n.x, n.y, n.z}, pv[3] = {v0.x, v0.y, v0.z};
  double pmin[3], pmax[3];
  for (int i = 0; i < 3; ++i) {
    if (nv[i] > 0.0) {
      pmin[i] = -h[i] - pv[i];
      pmax[i] = h[i] - pv[i];
    } else {
      pmin[i] = h[i] - pv[i];
      pmax[i] = -h[i] - pv[i];
    }
  }
  return dot(n, d3{pmin[0], pmin[1], pmin[2]}) <= 0 && dot(n, d3{pmax[0], pmax[1], pmax[2]}) >= 0;
}

__global__ void __launch_bounds__(256) msLeafKernel(MeshShapeArgs A, uint64_t ncand) {
  __shared__ MinkCfg cfgs[256 / kMsGroup];
  const uint64_t k = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kMsGroup;
  const int lg = static_cast<int>(threadIdx.x % kMsGroup);
  if (k >= ncand) return;  // uniform across a group
  const uint64_t c = A.cand[k];
  const uint64_t q = c >> 32;
  const uint32_t t = static_cast<uint32_t>(c & 0xFFFFFFFFull);
  const uint32_t mesh = A.mesh_ids[q], shp = A.shape_ids[q];
  const tf3 tm = poseFromQuat(A.pose_m + 7 * q), ts = poseFromQuat(A.pose_s + 7 * q);
  const float4* tri = A.tris + 3 * (static_cast<uint64_t>(A.tri_base[mesh]) + t);
  const d3 P1 = {tri[0].x, tri[0].y, tri[0].z}, P2 = {tri[1].x, tri[1].y, tri[1].z},
           P3 = {tri[2].x, tri[2].y, tri[2].z};
  const DevShape S = loadShape(A.T, shp);
  bool hit;
  if (S.type == kSphere) {
    hit = sphereTriHit(S.p0, ts.t, apply(tm, P1), apply(tm, P2), apply(tm, P3));
  } else if (S.type == kBox) {
    hit = boxTriHit(S.p0, S.p1, S.p2, ts, tm, P1, P2, P3);
  } else {
    DevShape B;
    B.type = kTriangle;
    B.p0 = B.p1 = B.p2 = 0;
    B.verts = tri;
    B.nv = 3;
    B.rbound = 0.f;
    B.lv = nullptr;
    B.ls = 0;
    B.nbr = nullptr;
    B.seed = nullptr;
    B.wblk = nullptr;
    d3 int0 = {0, 0, 0};
    if (S.type == kConvex) {
      const double4 ci = A.interior[shp];
      int0 = d3{ci.x, ci.y, ci.z};
    }
    const d3 int1 = (P1 + P2 + P3) / 3.0;
    MFCL_LDS MinkCfg* cfg = (MFCL_LDS MinkCfg*)(&cfgs[threadIdx.x / kMsGroup]);
    writeCfg(cfg, S, ts, B, tm, int0, int1);
    MinkL<kMsGroup> md{cfg, lg, 0};
    hit = mprIntersect(md, A.gjk_max_it, A.gjk_tol) == kMprIntersect;
  }
  if (lg != 0 || !hit) return;
  const uint32_t slot = atomicAdd(&A.counts[q], 1u);
  if (static_cast<int>(slot) < A.max_contacts)
    A.tri_out[q * static_cast<uint64_t>(A.max_contacts) + slot] = static_cast<int32_t>(t);
}

// Penetration enabled (shapeTriangleIntersect with contact outputs,
// gjk_solver-inl.h:545-563): one wavefront per candidate, taken from a
// shared counter.  Spheres take sphereTriContact; every other shape runs
// GJK from -UnitX and EPA on (shape, triangle) (gjk_solver-inl.h:452-527),
// the EPA polytope in the wavefront's LDS slice (cap_faces faces) and,
// when it might outgrow it, in the wavefront's device-memory polytope of the
// request's capacity.  A hit records the contact with the traversal's
// normal reversal (bvh_solver-inl.h:55-65, shape_pair_intersect-inl.h:
// 19-26): b1 = the triangle, b2 = -1.
constexpr size_t kMsCfgBytes = (sizeof(MinkCfg) + 15) & ~static_cast<size_t>(15);
constexpr size_t kMsRawBytes = (sizeof(RawF) + 15) & ~static_cast<size_t>(15);
constexpr size_t kMsPolyBytes = (sizeof(Poly) + 15) & ~static_cast<size_t>(15);
inline size_t msSliceBytes(int cap_faces) { return kMsCfgBytes + kMsRawBytes + kMsPolyBytes + arenaBytes(cap_faces); }

struct MsPen {
  mfcl_mesh_contact* contacts;  // n * max_contacts
  int cap_faces;                // LDS polytope capacity
  size_t slice_bytes;           // LDS bytes per wavefront
  char* arena;                  // device-memory polytopes, one per wavefront of the grid
  size_t arena_slice;
  int epa_max_faces, epa_max_it;
  double epa_tol;
  uint32_t* next;               // candidate cursor
};

__global__ void __launch_bounds__(256, 2) msPenKernel(MeshShapeArgs A, MsPen E, uint64_t ncand) {
  extern __shared__ __align__(16) char ms_lds[];
  const int lane = threadIdx.x & 63;
  const ArenaDims dims = arenaDims(E.cap_faces);
  MFCL_LDS char* slice = (MFCL_LDS char*)(ms_lds) + static_cast<size_t>(threadIdx.x >> 6) * E.slice_bytes;
  MFCL_LDS MinkCfg* cfg = (MFCL_LDS MinkCfg*)(slice);
  MFCL_LDS RawF* raw = (MFCL_LDS RawF*)(slice + kMsCfgBytes);
  MFCL_LDS char* base = slice + kMsCfgBytes + kMsRawBytes + kMsPolyBytes;
  Poly& poly = *reinterpret_cast<Poly*>(ms_lds + static_cast<size_t>(threadIdx.x >> 6) * E.slice_bytes +
                                        kMsCfgBytes + kMsRawBytes);
  poly.V = (MFCL_LDS EVert*)(base);
  poly.E = (MFCL_LDS EEdge*)(base + sizeof(EVert) * dims.nv);
  poly.F = (MFCL_LDS EFace*)(base + sizeof(EVert) * dims.nv + sizeof(EEdge) * dims.nv);
  poly.BL = (MFCL_LDS BorderRec*)(base + sizeof(EVert) * dims.nv + sizeof(EEdge) * dims.nv +
                                  sizeof(EFace) * dims.nf);
  poly.nv = dims.nv;
  poly.nf = dims.nf;
  const uint64_t wave = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const SpillTarget st{E.cap_faces < E.epa_max_faces ? E.arena + wave * E.arena_slice : nullptr, E.epa_max_faces};
  for (;;) {
    uint32_t k = 0;
    if (lane == 0) k = atomicAdd(E.next, 1u);
    k = __shfl(k, 0);
    if (k >= ncand) break;
    const uint64_t c = A.cand[k];
    const uint64_t q = c >> 32;
    const uint32_t t = static_cast<uint32_t>(c & 0xFFFFFFFFull);
    const uint32_t mesh = A.mesh_ids[q], shp = A.shape_ids[q];
    const tf3 tm = poseFromQuat(A.pose_m + 7 * q), ts = poseFromQuat(A.pose_s + 7 * q);
    const float4* tri = A.tris + 3 * (static_cast<uint64_t>(A.tri_base[mesh]) + t);
    const d3 P1 = {tri[0].x, tri[0].y, tri[0].z}, P2 = {tri[1].x, tri[1].y, tri[1].z},
             P3 = {tri[2].x, tri[2].y, tri[2].z};
    const DevShape S = loadShape(A.T, shp);
    bool hit = false;
    d3 pos = {0, 0, 0}, nrm = {0, 0, 0};
    double depth = 0;
    if (S.type == kSphere) {
      SphereTriContact sc;
      hit = sphereTriContact(S.p0, ts.t, apply(tm, P1), apply(tm, P2), apply(tm, P3), &sc);
      pos = sc.pos;
      nrm = sc.normal;
      depth = sc.depth;
    } else {
      DevShape B;
      B.type = kTriangle;
      B.p0 = B.p1 = B.p2 = 0;
      B.verts = tri;
      B.nv = 3;
      B.rbound = 0.f;
      B.lv = nullptr;
      B.ls = 0;
      B.nbr = nullptr;
      B.seed = nullptr;
      B.wblk = nullptr;
      writeCfg(cfg, S, ts, B, tm, d3{0, 0, 0}, d3{0, 0, 0});
      MinkL<64> md{cfg, lane, 0};
      Smp smp;
      GjkOut g;
      gjkEvaluate(md, smp, d3{-1, 0, 0}, A.gjk_tol, A.gjk_max_it, false, g);
      if (g.status == kGjkIntersect) {
        SmpD sd;
        sd.rank = smp.rank;
        storeMVD(sd.a, smp.a.v, slotDir(cfg, smp.a.id));
        storeMVD(sd.b, smp.b.v, slotDir(cfg, smp.b.id));
        storeMVD(sd.c, smp.c.v, slotDir(cfg, smp.c.id));
        storeMVD(sd.e, smp.e.v, slotDir(cfg, smp.e.id));
        EpaOut eo;
        if (st.base) epaEvaluate<true>(poly, md, sd, E.epa_tol, E.epa_max_it, eo, *raw, st);
        else epaEvaluate<false>(poly, md, sd, E.epa_tol, E.epa_max_it, eo, *raw, st);
        if (eo.status != kEpaFailed) {
          hit = true;
          d3 n1 = eo.p0 - eo.p1;
          n1 = sqnorm(n1) <= 0 ? d3{0, 0, 1} : normalized(n1);
          pos = apply(ts, 0.5 * (eo.p0 + eo.p1));
          nrm = mul(ts.R, n1);
          depth = eo.depth;
        }
      }
    }
    if (lane != 0 || !hit) continue;
    const uint32_t slot = atomicAdd(&A.counts[q], 1u);
    if (static_cast<int>(slot) < A.max_contacts) {
      mfcl_mesh_contact& o = E.contacts[q * static_cast<uint64_t>(A.max_contacts) + slot];
      o.b1 = static_cast<int32_t>(t);
      o.b2 = -1;
      o.pos[0] = pos.x; o.pos[1] = pos.y; o.pos[2] = pos.z;
      o.normal[0] = -nrm.x; o.normal[1] = -nrm.y; o.normal[2] = -nrm.z;
      o.depth = depth;
    }
  }
}

// Result headers of a penetration-enabled query; the kept contacts sorted
// by triangle id.
__global__ void msFinishPenKernel(MeshShapeArgs A, mfcl_mesh_contact* contacts, mfcl_mesh_result* out) {
  const uint64_t q = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= A.n) return;
  const uint32_t cnt = A.counts[q];
  const int kept = static_cast<int>(cnt) < A.max_contacts ? static_cast<int>(cnt) : A.max_contacts;
  mfcl_mesh_result r;
  r.hit = A.overflow[q] ? -1 : (cnt > 0 ? 1 : 0);
  r.n_contacts = kept;
  r.bv_tests = static_cast<int32_t>(A.bv_tests[q]);
  r.leaf_tests = static_cast<int32_t>(A.leaves[q]);
  out[q] = r;
  mfcl_mesh_contact* o = contacts + q * static_cast<uint64_t>(A.max_contacts);
  for (int i = 1; i < kept; ++i) {
    const mfcl_mesh_contact key = o[i];
    int j = i - 1;
    while (j >= 0 && o[j].b1 > key.b1) {
      o[j + 1] = o[j];
      --j;
    }
    o[j + 1] = key;
  }
  for (int i = kept; i < A.max_contacts; ++i) {
    mfcl_mesh_contact z;
    z.b1 = z.b2 = -1;
    z.pos[0] = z.pos[1] = z.pos[2] = 0;
    z.normal[0] = z.normal[1] = z.normal[2] = 0;
    z.depth = 0;
    o[i] = z;
  }
}

__global__ void msFinishKernel(MeshShapeArgs A, mfcl_mesh_result* out) {
  const uint64_t q = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= A.n) return;
  const uint32_t cnt = A.counts[q];
  const int kept = static_cast<int>(cnt) < A.max_contacts ? static_cast<int>(cnt) : A.max_contacts;
  mfcl_mesh_result r;
  r.hit = A.overflow[q] ? -1 : (cnt > 0 ? 1 : 0);
  r.n_contacts = kept;
  r.bv_tests = static_cast<int32_t>(A.bv_tests[q]);
  r.leaf_tests = static_cast<int32_t>(A.leaves[q]);
  out[q] = r;
  int32_t* o = A.tri_out + q * static_cast<uint64_t>(A.max_contacts);
  for (int i = 1; i < kept; ++i) {
    const int32_t key = o[i];
    int j = i - 1;
    while (j >= 0 && o[j] > key) {
      o[j + 1] = o[j];
      --j;
    }
    o[j + 1] = key;
  }
  for (int i = kept; i < A.max_contacts; ++i) o[i] = -1;
}

__global__ void msZeroKernel(uint32_t* p, uint64_t n) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

namespace {
struct DevAllocs {
  std::vector<void*> p;
  ~DevAllocs() {
    for (void* q : p) (void)hipFree(q);
  }
  void* get(size_t bytes) {
    void* q = nullptr;
    if (hipMalloc(&q, bytes ? bytes : 8) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    p.push_back(q);
    return q;
  }
};
}  // namespace

int meshShapeCollide(mfcl_ctx* ctx, const uint32_t* mesh_ids, const uint32_t* shape_ids, const float* pose_m7,
                     const float* pose_s7, uint64_t n, const mfcl_request* req, mfcl_mesh_result* out,
                     int32_t* tri_ids, mfcl_mesh_contact* contacts) {
  if (ctx->n_meshes == 0 || ctx->n_shapes == 0) {
    ctx->err = "mesh-shape collision needs uploaded meshes and shapes";
    return MFCL_E_STATE;
  }
  const bool pen = req->penetration != 0;
  if (pen != (contacts != nullptr)) {
    ctx->err = pen ? "a penetration-enabled mesh-shape query writes contacts: pass a contact array"
                   : "a contact array needs a request with penetration enabled";
    return MFCL_E_INVALID;
  }
  const int maxc = req->max_contacts;
  if (maxc < 1 || maxc > 1024) {
    ctx->err = "max_contacts must be in [1, 1024]";
    return MFCL_E_INVALID;
  }
  if (pen && (req->epa_max_faces < 4 || req->epa_max_faces > 8192)) {
    ctx->err = "epa_max_faces out of range [4, 8192]";
    return MFCL_E_INVALID;
  }
  for (uint64_t i = 0; i < n; ++i) {
    if (mesh_ids[i] >= ctx->n_meshes || shape_ids[i] >= ctx->n_shapes) {
      ctx->err = "query " + std::to_string(i) + " references a mesh or shape id out of range";
      return MFCL_E_INVALID;
    }
  }
  const hipStream_t s = ctx->stream;
  DevAllocs m;
  auto* dm = static_cast<uint32_t*>(m.get(4 * n));
  auto* ds = static_cast<uint32_t*>(m.get(4 * n));
  auto* dpm = static_cast<float*>(m.get(28 * n));
  auto* dps = static_cast<float*>(m.get(28 * n));
  auto* dcnt = static_cast<uint32_t*>(m.get(4 * 4 * n));  // counts, leaves, bv tests, overflow
  auto* dcc = static_cast<unsigned long long*>(m.get(16));  // candidate count, EPA cursor
  auto* dout = static_cast<mfcl_mesh_result*>(m.get(sizeof(mfcl_mesh_result) * n));
  auto* dtri = static_cast<int32_t*>(m.get(4 * n * static_cast<uint64_t>(maxc)));
  if (!dm || !ds || !dpm || !dps || !dcnt || !dcc || !dout || !dtri) {
    ctx->err = "out of device memory for the mesh-shape query";
    return MFCL_E_NOMEM;
  }
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dm, mesh_ids, 4 * n, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(ds, shape_ids, 4 * n, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dpm, pose_m7, 28 * n, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dps, pose_s7, 28 * n, hipMemcpyHostToDevice, s));
  MeshShapeArgs A;
  A.T = ShapeTables{ctx->d_type, ctx->d_param, ctx->d_voff, ctx->d_vcnt, ctx->d_verts, ctx->d_nbr, ctx->d_nbr_off,
                    ctx->d_seed, ctx->d_wblk};
  A.interior = ctx->d_interior;
  A.nodes = static_cast<const ObbNode*>(ctx->d_nodes);
  A.tris = static_cast<const float4*>(ctx->d_tris);
  A.node_base = ctx->d_mesh_node_base;
  A.tri_base = ctx->d_mesh_tri_base;
  A.mesh_ids = dm;
  A.shape_ids = ds;
  A.pose_m = dpm;
  A.pose_s = dps;
  A.n = n;
  A.counts = dcnt;
  A.leaves = dcnt + n;
  A.bv_tests = dcnt + 2 * n;
  A.overflow = dcnt + 3 * n;
  A.cand_count = dcc;
  A.max_contacts = maxc;
  A.tri_out = dtri;
  A.gjk_max_it = req->gjk_max_iterations;
  A.gjk_tol = req->gjk_tolerance;
  // candidates: sized for 64 per query, grown to the exact count and rerun
  uint64_t cap = std::max<uint64_t>(64 * n, 1 << 16);
  unsigned long long ncand = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    A.cand = static_cast<uint64_t*>(m.get(8 * cap));
    if (!A.cand) {
      ctx->err = "out of device memory for the mesh-shape candidates";
      return MFCL_E_NOMEM;
    }
    A.cand_cap = cap;
    msZeroKernel<<<static_cast<unsigned>((4 * n + 255) / 256), 256, 0, s>>>(dcnt, 4 * n);
    msZeroKernel<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t*>(dcc), 2);
    msCandKernel<<<static_cast<unsigned>(n), 64, 0, s>>>(A);
    MFCL_HIP_TRY(ctx, hipGetLastError());
    MFCL_HIP_TRY(ctx, hipMemcpyAsync(&ncand, dcc, 8, hipMemcpyDeviceToHost, s));
    MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (ncand <= cap) break;
    cap = ncand;
  }
  if (ncand > cap) {
    ctx->err = "mesh-shape candidate count changed between runs";
    return MFCL_E_STATE;
  }
  if (!pen) {
    if (ncand)
      msLeafKernel<<<static_cast<unsigned>((ncand * kMsGroup + 255) / 256), 256, 0, s>>>(A, ncand);
    msFinishKernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(A, dout);
    MFCL_HIP_TRY(ctx, hipGetLastError());
    MFCL_HIP_TRY(ctx, hipMemcpyAsync(out, dout, sizeof(mfcl_mesh_result) * n, hipMemcpyDeviceToHost, s));
    if (tri_ids)
      MFCL_HIP_TRY(ctx, hipMemcpyAsync(tri_ids, dtri, 4 * n * static_cast<uint64_t>(maxc), hipMemcpyDeviceToHost, s));
    MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
    return MFCL_OK;
  }
  // penetration enabled: GJK + EPA per candidate, 4 wavefronts per block
  if (ncand > 0xF0000000ull) {
    ctx->err = "too many mesh-shape candidates for one penetration-enabled call";
    return MFCL_E_INVALID;
  }
  MsPen E;
  E.epa_max_faces = req->epa_max_faces;
  E.epa_max_it = req->epa_max_iterations;
  E.epa_tol = req->epa_tolerance;
  int lds_faces = std::min(E.epa_max_faces, 64);
  while (lds_faces > 4 && 4 * msSliceBytes(lds_faces) > static_cast<size_t>(ctx->max_lds_per_block)) lds_faces /= 2;
  E.cap_faces = lds_faces;
  E.slice_bytes = msSliceBytes(lds_faces);
  const uint64_t blocks64 = std::min<uint64_t>((ncand + 3) / 4, 2ull * static_cast<uint64_t>(ctx->n_cu));
  const unsigned blocks = static_cast<unsigned>(std::max<uint64_t>(blocks64, 1));
  E.arena = nullptr;
  E.arena_slice = arenaBytes(E.epa_max_faces);
  if (lds_faces < E.epa_max_faces) {
    E.arena = static_cast<char*>(m.get(E.arena_slice * 4 * blocks));
    if (!E.arena) {
      ctx->err = "out of device memory for the mesh-shape EPA polytopes";
      return MFCL_E_NOMEM;
    }
  }
  E.contacts = static_cast<mfcl_mesh_contact*>(m.get(sizeof(mfcl_mesh_contact) * n * static_cast<uint64_t>(maxc)));
  E.next = reinterpret_cast<uint32_t*>(dcc) + 2;  // the candidate counter's third word
  if (!E.contacts) {
    ctx->err = "out of device memory for the mesh-shape contacts";
    return MFCL_E_NOMEM;
  }
  msZeroKernel<<<1, 64, 0, s>>>(E.next, 1);
  if (ncand) msPenKernel<<<blocks, 256, 4 * E.slice_bytes, s>>>(A, E, ncand);
  msFinishPenKernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(A, E.contacts, dout);
  MFCL_HIP_TRY(ctx, hipGetLastError());
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(out, dout, sizeof(mfcl_mesh_result) * n, hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(contacts, E.contacts, sizeof(mfcl_mesh_contact) * n * static_cast<uint64_t>(maxc),
                                   hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
  return MFCL_OK;
}

}  // namespace mfcl
