// Robot-arm configuration queries: forward kinematics of a serial chain of
// convex links, broadphase of every link against the AABB tree of obstacle
// boxes, signed distance of each candidate (link, box) pair, and the
// per-configuration any-collision flag and minimum distance.
//
// The reference has no arm model; a caller would run the same chain on the
// CPU and call BinaryAABB_Tree::SingleObjectCollision (binary_AABB_tree-
// inl.h:400-440) per link, then GJKSolver::shapeSignedDistance (gjk_solver-
// inl.h:820-874) per candidate pair.  This file only adds the kinematics and
// the two reductions around the library's broadphase and narrowphase.
//
// Chain: joint i turns about z (even i) or y (odd i) by q_i; link i extends
// along its local x over length L_i, its shape centred at L_i / 2.  Poses are
// composed in FP64 (quaternion product) and rounded to FP32 (t, q) for the
// narrowphase, like any caller's input.  A link's broadphase box is its
// bounding sphere (radius r_i about the shape origin) grown by the distance
// margin, so every pair closer than the margin becomes a candidate.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ctx.h"
#include "dmath.h"

namespace mfcl {

struct ArmDev {
  int n_links;
  uint32_t shape[MFCL_ARM_MAX_LINKS];
  float len[MFCL_ARM_MAX_LINKS];
  float radius[MFCL_ARM_MAX_LINKS];
};

__global__ void __launch_bounds__(256) armFkKernel(ArmDev arm, const float* q, uint64_t n_cfg, float margin,
                                                   float* pose7, float* mn3, float* mx3) {
  const uint64_t c = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n_cfg) return;
  double qw = 1, qx = 0, qy = 0, qz = 0;  // chain orientation
  double px = 0, py = 0, pz = 0;          // joint position
  for (int i = 0; i < arm.n_links; ++i) {
    double s, co;
    sincos(0.5 * static_cast<double>(q[c * arm.n_links + i]), &s, &co);
    // q <- q * (rotation about z or y)
    double nw, nx, ny, nz;
    if ((i & 1) == 0) {
      nw = qw * co - qz * s; nx = qx * co + qy * s; ny = qy * co - qx * s; nz = qz * co + qw * s;
    } else {
      nw = qw * co - qy * s; nx = qx * co - qz * s; ny = qy * co + qw * s; nz = qz * co + qx * s;
    }
    qw = nw; qx = nx; qy = ny; qz = nz;
    // local x axis in the world: first column of R(q)
    const double ax = 1 - 2 * (qy * qy + qz * qz), ay = 2 * (qx * qy + qz * qw), az = 2 * (qx * qz - qy * qw);
    const double h = 0.5 * static_cast<double>(arm.len[i]);
    const double cx = px + h * ax, cy = py + h * ay, cz = pz + h * az;
    const uint64_t k = c * arm.n_links + i;
    float* o = pose7 + 7 * k;
    o[0] = static_cast<float>(cx); o[1] = static_cast<float>(cy); o[2] = static_cast<float>(cz);
    o[3] = static_cast<float>(qw); o[4] = static_cast<float>(qx); o[5] = static_cast<float>(qy);
    o[6] = static_cast<float>(qz);
    // broadphase box around the FP32 centre the narrowphase sees
    const float rr = arm.radius[i] + margin;
    mn3[3 * k] = o[0] - rr; mn3[3 * k + 1] = o[1] - rr; mn3[3 * k + 2] = o[2] - rr;
    mx3[3 * k] = o[0] + rr; mx3[3 * k + 1] = o[1] + rr; mx3[3 * k + 2] = o[2] + rr;
    px += 2 * h * ax; py += 2 * h * ay; pz += 2 * h * az;
  }
}

// Broadphase hits -> narrowphase pairs (appended in any order) with their
// configuration ids.  Query k = c * n_links + link.
__global__ void __launch_bounds__(256) armPairsKernel(const int32_t* counts, const int64_t* ids, int cap, uint64_t nq,
                                                      ArmDev arm, uint32_t box_shape0,
                                                      uint32_t n_boxes, const float* link_pose7, const float* box_pose7,
                                                      uint32_t* pa, uint32_t* pb, float* qa, float* qb,
                                                      uint32_t* pcfg, uint64_t max_pairs, unsigned long long* n_pairs,
                                                      int32_t* any_hit) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nq) return;
  const int cnt = counts[k];
  const uint64_t cfg = k / arm.n_links;
  if (cnt < 0 || cnt > cap) {
    // broadphase overflow (more candidates than cap) or a corrupt tree: the
    // configuration is reported as failed rather than silently truncated
    atomicOr(&any_hit[cfg], 2);
    if (cnt < 0) return;
  }
  const int m = cnt < cap ? cnt : cap;
  if (m == 0) return;
  const unsigned long long base = atomicAdd(n_pairs, static_cast<unsigned long long>(m));
  const int link = static_cast<int>(k % arm.n_links);
  for (int j = 0; j < m; ++j) {
    const uint64_t p = base + j;
    if (p >= max_pairs) {
      atomicOr(&any_hit[cfg], 2);
      return;
    }
    const int64_t b = ids[k * cap + j];
    pcfg[p] = static_cast<uint32_t>(cfg);
    if (b < 0 || static_cast<uint64_t>(b) >= n_boxes) {
      // an id the tree was not built with: the slot becomes a pair with
      // invalid shape ids, which the narrowphase reports as ok = -1
      pa[p] = pb[p] = 0xFFFFFFFFu;
      continue;
    }
    pa[p] = arm.shape[link];
    pb[p] = box_shape0 + static_cast<uint32_t>(b);
    for (int t = 0; t < 7; ++t) {
      qa[7 * p + t] = link_pose7[7 * k + t];
      qb[7 * p + t] = box_pose7[7 * b + t];
    }
  }
}

// Order-preserving map of a float onto int32 (for atomicMin over signed values).
MFCL_D int32_t floatKey(float f) {
  const int32_t b = __float_as_int(f);
  return b >= 0 ? b : b ^ 0x7fffffff;
}
MFCL_D float keyFloat(int32_t k) { return __int_as_float(k >= 0 ? k : k ^ 0x7fffffff); }

__global__ void armInitKernel(uint64_t n_cfg, int32_t* any_hit, int32_t* min_key, unsigned long long* n_pairs) {
  const uint64_t c = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c == 0) *n_pairs = 0;
  if (c >= n_cfg) return;
  any_hit[c] = 0;
  min_key[c] = floatKey(INFINITY);
}

// Per pair: collision when the signed distance is negative (EPA depth);
// the configuration's minimum over its pairs with a valid result; a pair
// whose query failed (ok != 1) marks its configuration with bit 1.
__global__ void armReduceKernel(const mfcl_convex_result* res, const uint32_t* pcfg,
                                const unsigned long long* n_pairs, int32_t* any_hit, int32_t* min_key) {
  const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= *n_pairs) return;
  const mfcl_convex_result& r = res[p];
  const uint32_t c = pcfg[p];
  if (r.ok != 1) {
    atomicOr(&any_hit[c], 2);
    return;
  }
  if (r.distance < 0) atomicOr(&any_hit[c], 1);
  atomicMin(&min_key[c], floatKey(static_cast<float>(r.distance)));
}

__global__ void armFinishKernel(uint64_t n_cfg, const int32_t* min_key, float* min_dist) {
  const uint64_t c = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n_cfg) return;
  min_dist[c] = keyFloat(min_key[c]);
}

int aabbQueryLaunch(mfcl_ctx* ctx, const float* qmn, const float* qmx, uint64_t nq, int32_t cap, int32_t* counts,
                    int64_t* ids, hipStream_t s);

int armQueryLaunch(mfcl_ctx* ctx, const mfcl_arm* arm, const float* q, uint64_t n_cfg, float margin,
                   const float* box_pose7, int32_t* any_hit, float* min_dist, uint64_t* n_pairs_out, hipStream_t s) {
  if (arm->n_links < 1 || arm->n_links > MFCL_ARM_MAX_LINKS) {
    ctx->err = "n_links out of range [1, MFCL_ARM_MAX_LINKS]";
    return MFCL_E_INVALID;
  }
  if (arm->broadphase_cap < 1 || arm->broadphase_cap > 512) {
    ctx->err = "broadphase_cap out of range [1, 512]";
    return MFCL_E_INVALID;
  }
  if (!(margin >= 0)) {
    ctx->err = "margin must be >= 0";
    return MFCL_E_INVALID;
  }
  if (ctx->bp_n_nodes == 0) {
    ctx->err = "no AABB tree built";
    return MFCL_E_STATE;
  }
  for (int i = 0; i < arm->n_links; ++i)
    if (arm->link_shape[i] >= ctx->n_shapes) {
      ctx->err = "link shape id out of range";
      return MFCL_E_INVALID;
    }
  if (static_cast<uint64_t>(arm->box_shape0) + arm->n_boxes > ctx->n_shapes) {
    ctx->err = "box shape ids out of range";
    return MFCL_E_INVALID;
  }
  if (n_cfg == 0) return MFCL_OK;
  const uint64_t nq = n_cfg * static_cast<uint64_t>(arm->n_links);
  const int cap = arm->broadphase_cap;
  const uint64_t max_pairs = arm->max_pairs ? arm->max_pairs : nq * 8;
  // workspace: link poses, boxes, hit counts and ids, pairs, results, keys
  const size_t b_pose = nq * 28, b_box = nq * 24, b_cnt = nq * 4, b_ids = nq * cap * 8;
  const size_t b_pairs = max_pairs * (4 + 4 + 28 + 28 + 4), b_res = max_pairs * sizeof(mfcl_convex_result);
  // every piece below is rounded up to 256 bytes: 16 pieces of slack
  const size_t total = b_pose + b_box + b_cnt + b_ids + b_pairs + b_res + n_cfg * 4 + 16 * 8 + 16 * 256;
  if (int rc = ensure(ctx, ctx->ws_arm, total)) return rc;
  char* w = static_cast<char*>(ctx->ws_arm.p);
  auto take = [&](size_t bytes) {
    char* r = w;
    w += (bytes + 255) & ~static_cast<size_t>(255);
    return r;
  };
  float* pose7 = reinterpret_cast<float*>(take(b_pose));
  float* mn3 = reinterpret_cast<float*>(take(nq * 12));
  float* mx3 = reinterpret_cast<float*>(take(nq * 12));
  int32_t* counts = reinterpret_cast<int32_t*>(take(b_cnt));
  int64_t* ids = reinterpret_cast<int64_t*>(take(b_ids));
  uint32_t* pa = reinterpret_cast<uint32_t*>(take(max_pairs * 4));
  uint32_t* pb = reinterpret_cast<uint32_t*>(take(max_pairs * 4));
  float* qa = reinterpret_cast<float*>(take(max_pairs * 28));
  float* qb = reinterpret_cast<float*>(take(max_pairs * 28));
  uint32_t* pcfg = reinterpret_cast<uint32_t*>(take(max_pairs * 4));
  mfcl_convex_result* res = reinterpret_cast<mfcl_convex_result*>(take(b_res));
  int32_t* min_key = reinterpret_cast<int32_t*>(take(n_cfg * 4));
  unsigned long long* npairs = reinterpret_cast<unsigned long long*>(take(16 * 8));
  if (w > static_cast<char*>(ctx->ws_arm.p) + ctx->ws_arm.bytes) {
    ctx->err = "arm workspace layout exceeds its allocation";
    return MFCL_E_INVALID;
  }
  ArmDev a;
  a.n_links = arm->n_links;
  for (int i = 0; i < MFCL_ARM_MAX_LINKS; ++i) {
    a.shape[i] = i < arm->n_links ? arm->link_shape[i] : 0u;
    a.len[i] = i < arm->n_links ? arm->link_len[i] : 0.f;
    a.radius[i] = i < arm->n_links ? arm->link_radius[i] : 0.f;
  }
  const unsigned gc = static_cast<unsigned>((n_cfg + 255) / 256);
  const unsigned gq = static_cast<unsigned>((nq + 255) / 256);
  armInitKernel<<<gc, 256, 0, s>>>(n_cfg, any_hit, min_key, npairs);
  armFkKernel<<<gc, 256, 0, s>>>(a, q, n_cfg, margin, pose7, mn3, mx3);
  if (int rc = aabbQueryLaunch(ctx, mn3, mx3, nq, cap, counts, ids, s)) return rc;
  armPairsKernel<<<gq, 256, 0, s>>>(counts, ids, cap, nq, a, arm->box_shape0, arm->n_boxes, pose7, box_pose7, pa,
                                    pb, qa, qb, pcfg, max_pairs, npairs, any_hit);
  unsigned long long np = 0;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(&np, npairs, sizeof(np), hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
  if (np > max_pairs) np = max_pairs;
  if (n_pairs_out) *n_pairs_out = np;
  if (np) {
    mfcl_request req;
    mfcl_request_defaults(&req, MFCL_MODE_SIGNED_DISTANCE);
    req.gjk_max_iterations = arm->gjk_max_iterations ? arm->gjk_max_iterations : req.gjk_max_iterations;
    if (int rc = convexQueryLaunch(ctx, pa, pb, qa, qb, nullptr, np, &req, res, s)) return rc;
    armReduceKernel<<<static_cast<unsigned>((np + 255) / 256), 256, 0, s>>>(res, pcfg, npairs, any_hit, min_key);
  }
  armFinishKernel<<<gc, 256, 0, s>>>(n_cfg, min_key, min_dist);
  MFCL_HIP_TRY(ctx, hipGetLastError());
  return MFCL_OK;
}

}  // namespace mfcl
