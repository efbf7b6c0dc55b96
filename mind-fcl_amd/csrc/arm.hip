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
  if (c == 0) n_pairs[0] = n_pairs[1] = 0;
  if (c >= n_cfg) return;
  any_hit[c] = 0;
  min_key[c] = floatKey(INFINITY);
}

// Per pair: collision when the signed distance is negative (EPA depth);
// the configuration's minimum over its pairs with a valid result; a pair
// whose query failed (ok != 1) marks its configuration with bit 1.
__global__ void armReduceKernel(const mfcl_convex_result* res, const uint32_t* pcfg, uint64_t n_pairs,
                                int32_t* any_hit, int32_t* min_key) {
  const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
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

// ---- heightmap -------------------------------------------------------------
// The layers of the map (top 0 .. bottom n_layers - 1) and its pose.
struct HmDev {
  const uint16_t* h;
  int n_layers;
  uint64_t off[kMaxHmLayers];
  int fx[kMaxHmLayers], fy[kMaxHmLayers];
  float rx[kMaxHmLayers], ry[kMaxHmLayers];
  double upper_m;
  float pose7[7];
};
// The link hulls (vertex ranges in the shape table).
struct ArmHulls {
  uint64_t voff[MFCL_ARM_MAX_LINKS];
  uint32_t vcnt[MFCL_ARM_MAX_LINKS];
};

// aabbToPixelROI's corner projection (flat_heightmap-inl.h:58-87): pixel
// floor(p / res) + half clamped to [0, full - 1]; true when it was clamped.
MFCL_D bool hmProject(double p, double res, int half, int full, int& out) {
  const double v = floor(p / res) + half;
  if (v <= 0) {
    out = 0;
    return v != 0;
  }
  if (v >= full - 1) {
    out = full - 1;
    return v != full - 1;
  }
  out = static_cast<int>(v);
  return false;
}

// p strictly inside the box of its bottom-layer pixel by `slack` on every
// side (pixel x spans [(x - half) res, (x - half + 1) res], z spans [0, h]).
MFCL_D bool hmInside(const HmDev& H, d3 p, double slack) {
  const int b = H.n_layers - 1;
  const double rx = H.rx[b], ry = H.ry[b];
  const int hx = H.fx[b] / 2, hy = H.fy[b] / 2;
  const double vx = floor(p.x / rx) + hx, vy = floor(p.y / ry) + hy;
  if (!(vx >= 0 && vx < H.fx[b] && vy >= 0 && vy < H.fy[b])) return false;
  const int px = static_cast<int>(vx), py = static_cast<int>(vy);
  const int hmm = H.h[H.off[b] + static_cast<uint64_t>(py) * H.fx[b] + px];
  if (hmm == 0) return false;
  const double top = static_cast<double>(hmm) * 0.001;
  const double x0 = (px - hx) * rx, y0 = (py - hy) * ry;
  return p.z > slack && p.z < top - slack && p.x > x0 + slack && p.x < x0 + rx - slack && p.y > y0 + slack &&
         p.y < y0 + ry - slack;
}

// Heightmap candidates of one link (query k = configuration * n_links +
// link), the flat algorithm of HeightMapShapeIntersect (heightmap_solver_
// traverse-inl.h:17-98): the hull's AABB in the map frame from its
// transformed vertices (computeBV<AABB, Convex>, utility-inl.h:160-173), no
// candidate when it lies below the map's base or above its highest pixel,
// else the pixels of the AABB's region (aabbToPixelROI) with non-zero height
// at or above the AABB's bottom.  Tiles of 2^s x 2^s pixels whose upper-layer
// maximum is below the AABB's bottom are skipped whole (the same set).
//
// Before any narrowphase work a link is decided colliding when its interior
// point (the vertex mean, inside the hull) or, by 0.1 mm, one of its
// vertices lies inside the box of the pixel under it: the point is then
// common to the link and that box.  Links of configurations that already collide (with
// an obstacle box) are not searched.  Candidates beyond `cap` mark the
// configuration failed (bit 1).
__global__ void __launch_bounds__(256) armHmKernel(ArmDev arm, ArmHulls hulls, HmDev H, const float4* verts,
                                                   const double4* interior, const float* link_pose7, uint64_t nq,
                                                   int cap, uint32_t hm_shape0, uint32_t* pa, uint32_t* pb,
                                                   float* qa, float* qb, uint32_t* pcfg, uint64_t max_pairs,
                                                   unsigned long long* n_pairs, int32_t* any_hit) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nq) return;
  const uint64_t cfg = k / arm.n_links;
  const int link = static_cast<int>(k % arm.n_links);
  if (any_hit[cfg] & 1) return;
  const tf3 thm = poseFromQuat(H.pose7);
  const tf3 m = relTf(thm, poseFromQuat(link_pose7 + 7 * k));  // link -> map frame
  const float4* v = verts + hulls.voff[link];
  const int nv = static_cast<int>(hulls.vcnt[link]);
  d3 mn = {INFINITY, INFINITY, INFINITY}, mx = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < nv; ++i) {
    const float4 q = v[i];
    const d3 p = apply(m, d3{q.x, q.y, q.z});
    mn.x = fmin(mn.x, p.x); mn.y = fmin(mn.y, p.y); mn.z = fmin(mn.z, p.z);
    mx.x = fmax(mx.x, p.x); mx.y = fmax(mx.y, p.y); mx.z = fmax(mx.z, p.z);
  }
  // no vertex, a non-finite pose, the map's base or its top
  if (!(mn.x <= mx.x && mn.y <= mx.y && mn.z <= mx.z)) return;
  if (mx.z < 0 || mn.z > H.upper_m) return;
  const int b = H.n_layers - 1;
  const int fx = H.fx[b], fy = H.fy[b];
  int x0, y0, x1, y1;
  const bool c0 = hmProject(mn.x, H.rx[b], fx / 2, fx, x0) | hmProject(mn.y, H.ry[b], fy / 2, fy, y0);
  const bool c1 = hmProject(mx.x, H.rx[b], fx / 2, fx, x1) | hmProject(mx.y, H.ry[b], fy / 2, fy, y1);
  if (c0 && c1 && (x0 == x1 || y0 == y1)) return;
  x0 = min(x0, x1);
  y0 = min(y0, y1);
  {
    const double4 c = interior[arm.shape[link]];
    bool in = hmInside(H, apply(m, d3{c.x, c.y, c.z}), 0.0);
    for (int i = 0; i < nv && !in; ++i) {
      const float4 q = v[i];
      in = hmInside(H, apply(m, d3{q.x, q.y, q.z}), 1e-4);
    }
    if (in) {
      atomicOr(&any_hit[cfg], 1);
      return;
    }
  }
  const int s = b >= 2 ? 2 : b;  // tile side 2^s, read from layer b - s
  const int cl = b - s, fxc = fx >> s;
  const uint16_t* hb = H.h + H.off[b];
  const uint16_t* hc = H.h + H.off[cl];
  const double zmin = mn.z;
  int total = 0;
  unsigned long long base = 0;
  for (int pass = 0; pass < 2; ++pass) {
    int j = 0;
    for (int ty = y0 >> s; ty <= y1 >> s; ++ty)
      for (int tx = x0 >> s; tx <= x1 >> s; ++tx) {
        const int tmax = hc[static_cast<uint64_t>(ty) * fxc + tx];
        if (tmax == 0 || zmin > static_cast<double>(tmax) * 0.001) continue;
        const int ya = max(y0, ty << s), yb = min(y1, ((ty + 1) << s) - 1);
        const int xa = max(x0, tx << s), xb = min(x1, ((tx + 1) << s) - 1);
        for (int y = ya; y <= yb; ++y)
          for (int x = xa; x <= xb; ++x) {
            const int hmm = hb[static_cast<uint64_t>(y) * fx + x];
            if (hmm == 0) continue;
            const double top = static_cast<double>(hmm) * 0.001;
            if (zmin > top) continue;
            if (pass == 0 || j >= total) {
              ++j;
              continue;
            }
            const uint64_t p = base + j++;
            if (p >= max_pairs) continue;
            // the pixel's box first, as boxToShapeProcessLeafPair orders
            // them (constructBox: about the pixel's centre, the map's
            // rotation), then the link
            pcfg[p] = static_cast<uint32_t>(cfg);
            pa[p] = hm_shape0 + static_cast<uint32_t>(y * fx + x);
            pb[p] = arm.shape[link];
            const d3 c = apply(thm, d3{(x - fx / 2 + 0.5) * H.rx[b], (y - fy / 2 + 0.5) * H.ry[b], 0.5 * top});
            qa[7 * p + 0] = static_cast<float>(c.x);
            qa[7 * p + 1] = static_cast<float>(c.y);
            qa[7 * p + 2] = static_cast<float>(c.z);
            for (int t = 3; t < 7; ++t) qa[7 * p + t] = H.pose7[t];
            for (int t = 0; t < 7; ++t) qb[7 * p + t] = link_pose7[7 * k + t];
          }
      }
    if (pass == 0) {
      if (j > cap) atomicOr(&any_hit[cfg], 2);
      total = j < cap ? j : cap;
      if (total == 0) return;
      base = atomicAdd(n_pairs, static_cast<unsigned long long>(total));
      if (base + total > max_pairs) atomicOr(&any_hit[cfg], 2);
    }
  }
}

// Collide-mode results of heightmap candidates: a hit marks the
// configuration colliding, an invalid pair failed.
__global__ void armHmReduceKernel(const mfcl_convex_result* res, const uint32_t* pcfg, uint64_t n_pairs,
                                  int32_t* any_hit) {
  const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  const mfcl_convex_result& r = res[p];
  if (r.ok != 1) atomicOr(&any_hit[pcfg[p]], 2);
  else if (r.hit) atomicOr(&any_hit[pcfg[p]], 1);
}

int aabbQueryLaunch(mfcl_ctx* ctx, const float* qmn, const float* qmx, uint64_t nq, int32_t cap, int32_t* counts,
                    int64_t* ids, hipStream_t s);

// Bytes of the arm query's own workspace for n_cfg configurations (link
// poses, boxes, hit counts and ids, candidate pairs, their results, the
// minimum keys and counters; every piece rounded up to 256 bytes) and the
// candidate pair capacity.
size_t armWsBytes(const mfcl_arm* arm, uint64_t n_cfg, uint64_t* max_pairs_out) {
  const uint64_t nq = n_cfg * static_cast<uint64_t>(arm->n_links);
  const uint64_t max_pairs = arm->max_pairs ? arm->max_pairs : nq * 8;
  const size_t b_pose = nq * 28, b_box = nq * 24, b_cnt = nq * 4, b_ids = nq * arm->broadphase_cap * 8;
  const size_t b_pairs = max_pairs * (4 + 4 + 28 + 28 + 4), b_res = max_pairs * sizeof(mfcl_convex_result);
  if (max_pairs_out) *max_pairs_out = max_pairs;
  return b_pose + b_box + b_cnt + b_ids + b_pairs + b_res + n_cfg * 4 + 16 * 8 + 16 * 256;
}

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
  ctx->arm_hm_pairs = 0;
  if (arm->use_heightmap) {
    if (ctx->hm.n_layers == 0 || !ctx->d_hm) {
      ctx->err = "no heightmap uploaded";
      return MFCL_E_STATE;
    }
    if (arm->hm_cap < 1 || arm->hm_cap > 65536) {
      ctx->err = "hm_cap out of range [1, 65536]";
      return MFCL_E_INVALID;
    }
    const int b = ctx->hm.n_layers - 1;
    const uint64_t npix = static_cast<uint64_t>(ctx->hm.fx[b]) * ctx->hm.fy[b];
    if (static_cast<uint64_t>(arm->hm_box_shape0) + npix > ctx->n_shapes) {
      ctx->err = "heightmap pixel box shape ids out of range";
      return MFCL_E_INVALID;
    }
    for (uint64_t i = 0; i < npix; ++i)
      if (ctx->h_type[arm->hm_box_shape0 + i] != MFCL_SHAPE_BOX) {
        ctx->err = "heightmap pixel shapes must be boxes";
        return MFCL_E_INVALID;
      }
    for (int i = 0; i < arm->n_links; ++i)
      if (ctx->h_type[arm->link_shape[i]] != MFCL_SHAPE_CONVEX) {
        ctx->err = "a heightmap query needs convex-hull links";
        return MFCL_E_INVALID;
      }
    float qn = 0;
    for (int t = 3; t < 7; ++t) qn += arm->hm_pose7[t] * arm->hm_pose7[t];
    for (int t = 0; t < 7; ++t)
      if (!std::isfinite(arm->hm_pose7[t])) qn = 0;
    if (!(qn > 0)) {
      ctx->err = "hm_pose7 must be finite with a non-zero quaternion";
      return MFCL_E_INVALID;
    }
  }
  if (n_cfg == 0) return MFCL_OK;
  const uint64_t nq = n_cfg * static_cast<uint64_t>(arm->n_links);
  const int cap = arm->broadphase_cap;
  uint64_t max_pairs = 0;
  const size_t total = armWsBytes(arm, n_cfg, &max_pairs);
  if (total > ctx->ws_arm.bytes || max_pairs > ctx->ws_convex_pairs) {
    ctx->err = "arm workspace too small for " + std::to_string(n_cfg) +
               " configurations of this arm: call mfcl_reserve_arm first";
    return MFCL_E_STATE;
  }
  ConvexWs cws;
  if (!carveConvexWs(ctx->ws_convex.p, ctx->ws_convex.bytes, max_pairs, epaArenaBytes(kReservedArenaSlices, 256),
                     cws)) {
    ctx->err = "convex workspace layout exceeds its allocation";
    return MFCL_E_STATE;
  }
  const size_t b_pose = nq * 28, b_cnt = nq * 4, b_ids = nq * cap * 8, b_res = max_pairs * sizeof(mfcl_convex_result);
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
    if (int rc = convexQueryLaunch(ctx, pa, pb, qa, qb, nullptr, np, &req, res, nullptr, cws, s)) return rc;
    armReduceKernel<<<static_cast<unsigned>((np + 255) / 256), 256, 0, s>>>(res, pcfg, np, any_hit, min_key);
  }
  if (arm->use_heightmap) {
    // the heightmap's candidates reuse the pair buffers (stream order)
    HmDev H;
    H.h = ctx->d_hm;
    H.n_layers = ctx->hm.n_layers;
    for (int l = 0; l < kMaxHmLayers; ++l) {
      H.off[l] = ctx->hm.off[l];
      H.fx[l] = ctx->hm.fx[l];
      H.fy[l] = ctx->hm.fy[l];
      H.rx[l] = ctx->hm.rx[l];
      H.ry[l] = ctx->hm.ry[l];
    }
    H.upper_m = static_cast<double>(ctx->hm.upper_mm) * 0.001;
    for (int t = 0; t < 7; ++t) H.pose7[t] = arm->hm_pose7[t];
    ArmHulls hulls;
    for (int i = 0; i < MFCL_ARM_MAX_LINKS; ++i) {
      hulls.voff[i] = i < arm->n_links ? ctx->h_voff[arm->link_shape[i]] : 0;
      hulls.vcnt[i] = i < arm->n_links ? ctx->h_vcnt[arm->link_shape[i]] : 0;
    }
    armHmKernel<<<gq, 256, 0, s>>>(a, hulls, H, ctx->d_verts, ctx->d_interior, pose7, nq, arm->hm_cap,
                                   arm->hm_box_shape0, pa, pb, qa, qb, pcfg, max_pairs, npairs + 1, any_hit);
    unsigned long long nh = 0;
    MFCL_HIP_TRY(ctx, hipMemcpyAsync(&nh, npairs + 1, sizeof(nh), hipMemcpyDeviceToHost, s));
    MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (nh > max_pairs) nh = max_pairs;
    ctx->arm_hm_pairs = nh;
    if (nh) {
      mfcl_request req;
      mfcl_request_defaults(&req, MFCL_MODE_COLLIDE);
      req.max_contacts = 1;
      req.penetration = 0;
      req.gjk_max_iterations = arm->gjk_max_iterations ? arm->gjk_max_iterations : req.gjk_max_iterations;
      if (int rc = convexQueryLaunch(ctx, pa, pb, qa, qb, nullptr, nh, &req, res, nullptr, cws, s)) return rc;
      armHmReduceKernel<<<static_cast<unsigned>((nh + 255) / 256), 256, 0, s>>>(res, pcfg, nh, any_hit);
    }
  }
  armFinishKernel<<<gc, 256, 0, s>>>(n_cfg, min_key, min_dist);
  MFCL_HIP_TRY(ctx, hipGetLastError());
  return MFCL_OK;
}

}  // namespace mfcl
