// Convex hulls of point sets on the device, and the convex shape table built
// from them without a host round trip.
//
// The reference takes a hull's vertices and faces from its caller
// (fcl::Convex's constructor, include/fcl/geometry/shape/convex-inl.h:50-73)
// and derives the vertex adjacency of its support walk from the faces
// (convex-inl.h:293-409).  Here the hull itself is computed on the GPU:
// one wavefront per point set runs an incremental hull in its LDS slice
// (points, triangles with their three edge neighbours, conflict faces),
// inserting outside points in index order: the faces a point sees form a
// connected cap found by a flood fill from its conflict face, the cap's
// border becomes a fan of new triangles to the point, and the points that
// saw the removed faces look for a new face they see among the fan.
// Orientation tests run in FP64 on the FP32 coordinates; a point within a
// relative 1e-12 of a face's plane does not see it, so (near-)coplanar
// points are left out of the hull as qhull leaves out coplanar points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.h"
#include "dmath.h"

namespace mfcl {

constexpr int kHullMaxPoints = 768;
constexpr int kHullMaxHorizon = 192;

enum HullStatus : int32_t { kHullOk = 0, kHullDegenerate = 1, kHullTooLarge = 2, kHullFailed = 3 };

// LDS slice of one wavefront for point sets of up to cap points (faces: 2 cap).
struct HullLds {
  float *x, *y, *z;
  int16_t* conf;      // per point: conflict face, -1 inside, -2 hull vertex
  int16_t* fv;        // [3][F] vertices of face f: fv[k * F + f]
  int16_t* fnb;       // [3][F] neighbour across edge (v_k, v_k+1)
  int32_t* fvis;      // per face: 0 unseen this step, 1 visible, 2 not visible, 3 being tested; -1 free
  int32_t* start;     // per point: horizon edge starting at it (-1 none); reused as new vertex ids at the end
  int32_t* endv;      // per point: horizon edge ending at it
  int16_t* flist;     // faces touched by this step's flood fill (reset list)
  int16_t* vis;       // visible faces of this step
  int16_t* freel;     // free face slots (stack)
  int16_t* hor;       // [4][kHullMaxHorizon] horizon edges: a, b, outer face, edge index in it
  int32_t* cnt;       // counters: [0] visible, [1] touched, [2] horizon, [3] frontier end, [4] error
  int F;
};

__host__ __device__ inline size_t hullLdsBytes(int cap) {
  const size_t F = 2 * static_cast<size_t>(cap);
  size_t b = 12 * static_cast<size_t>(cap) + 2 * cap + 6 * F + 6 * F + 4 * F + 8 * static_cast<size_t>(cap) +
             2 * F + 2 * F + 2 * F + 8 * kHullMaxHorizon + 8 * 4 + 128;  // + alignment of 13 pieces
  return (b + 15) & ~static_cast<size_t>(15);
}

MFCL_D HullLds hullCarve(char* base, int cap) {
  HullLds L;
  const int F = 2 * cap;
  char* p = base;
  auto take = [&](size_t bytes) {
    char* r = p;
    p += (bytes + 7) & ~static_cast<size_t>(7);
    return r;
  };
  L.x = reinterpret_cast<float*>(take(4 * cap));
  L.y = reinterpret_cast<float*>(take(4 * cap));
  L.z = reinterpret_cast<float*>(take(4 * cap));
  L.fvis = reinterpret_cast<int32_t*>(take(4 * F));
  L.start = reinterpret_cast<int32_t*>(take(4 * cap));
  L.endv = reinterpret_cast<int32_t*>(take(4 * cap));
  L.cnt = reinterpret_cast<int32_t*>(take(4 * 8));
  L.conf = reinterpret_cast<int16_t*>(take(2 * cap));
  L.fv = reinterpret_cast<int16_t*>(take(2 * 3 * F));
  L.fnb = reinterpret_cast<int16_t*>(take(2 * 3 * F));
  L.flist = reinterpret_cast<int16_t*>(take(2 * F));
  L.vis = reinterpret_cast<int16_t*>(take(2 * F));
  L.freel = reinterpret_cast<int16_t*>(take(2 * F));
  L.hor = reinterpret_cast<int16_t*>(take(2 * 4 * kHullMaxHorizon));
  L.F = F;
  return L;
}

MFCL_D d3 hp(const HullLds& L, int i) { return d3{L.x[i], L.y[i], L.z[i]}; }

// Signed volume test of point p against face (a, b, c), normal (b-a)x(c-a):
// +1 outside (beyond the relative threshold), 0 otherwise.
MFCL_D bool sees(const HullLds& L, int a, int b, int c, int p) {
  const d3 A = hp(L, a), B = hp(L, b), C = hp(L, c), P = hp(L, p);
  const d3 u = B - A, v = C - A, w = P - A;
  const d3 n = cross(u, v);
  const double o = dot(n, w);
  if (!(o > 0)) return false;
  const double scale = sqrt(sqnorm(u) * sqnorm(v) * sqnorm(w));
  return o > 1e-12 * scale;
}
MFCL_D bool faceSees(const HullLds& L, int f, int p) {
  const int F = L.F;
  return sees(L, L.fv[f], L.fv[F + f], L.fv[2 * F + f], p);
}

MFCL_D int laneId64() { return threadIdx.x & 63; }
MFCL_D void wsync() { __builtin_amdgcn_wave_barrier(); }

MFCL_D void setFace(HullLds& L, int f, int a, int b, int c) {
  const int F = L.F;
  L.fv[f] = static_cast<int16_t>(a);
  L.fv[F + f] = static_cast<int16_t>(b);
  L.fv[2 * F + f] = static_cast<int16_t>(c);
}

// Wave argmax of val (first index on ties); all lanes get (index, value).
MFCL_D void waveArgmax(double& val, int& idx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(val, off);
    const int oi = __shfl_xor(idx, off);
    if (ov > val || (ov == val && oi < idx)) {
      val = ov;
      idx = oi;
    }
  }
}

// One wavefront per set (set order[b] for block b when order is given, else
// set b), in an LDS slice for sets of up to cap points; the slots a set's
// faces take do not depend on cap, so neither does its result.  Outputs
// (per set i, points at pt_off[i]):
// hull_idx[pt_off + k] = index (within the set) of the k-th hull vertex in
// ascending order, hull_cnt[i]; tris[6 pt_off + 3 t + j] = triangle t over
// those k-numbered vertices (outward orientation), tri_cnt[i]; status[i].
__global__ void __launch_bounds__(64) hullKernel(const float* pts, const uint64_t* pt_off, const uint32_t* pt_cnt,
                                                 const uint32_t* order, uint32_t n_sets, int cap, uint32_t* hull_cnt,
                                                 uint32_t* hull_idx, uint32_t* tri_cnt, int32_t* tris,
                                                 int32_t* status) {
  extern __shared__ __align__(16) char hull_lds[];
  if (blockIdx.x >= n_sets) return;
  const uint32_t set = order ? order[blockIdx.x] : blockIdx.x;
  const int l = laneId64();
  const int n = static_cast<int>(pt_cnt[set]);
  const uint64_t o = pt_off[set];
  auto finish = [&](int st) {
    if (l == 0) {
      status[set] = st;
      hull_cnt[set] = 0;
      tri_cnt[set] = 0;
    }
  };
  if (n > cap || n > kHullMaxPoints) { finish(kHullTooLarge); return; }
  if (n < 4) { finish(kHullDegenerate); return; }
  HullLds L = hullCarve(hull_lds, cap);
  const int F = L.F;
  for (int i = l; i < n; i += 64) {
    L.x[i] = pts[3 * (o + i)];
    L.y[i] = pts[3 * (o + i) + 1];
    L.z[i] = pts[3 * (o + i) + 2];
    L.conf[i] = -1;
    L.start[i] = -1;
    L.endv[i] = -1;
  }
  for (int f = l; f < F; f += 64) {
    L.fvis[f] = -1;
    L.fnb[f] = L.fnb[F + f] = L.fnb[2 * F + f] = 0;
  }
  wsync();
  // ---- initial tetrahedron: extreme x, farthest point, farthest from the
  // line, farthest from the plane (relative tolerances against the extent)
  int i0 = 0;
  {
    double v = -INFINITY;
    int idx = 0x7fffffff;
    for (int i = l; i < n; i += 64)
      if (-static_cast<double>(L.x[i]) > v) { v = -static_cast<double>(L.x[i]); idx = i; }
    waveArgmax(v, idx);
    i0 = idx;
  }
  const d3 P0 = hp(L, i0);
  int i1 = 0;
  double ext2 = 0;
  {
    double v = -1;
    int idx = 0x7fffffff;
    for (int i = l; i < n; i += 64) {
      const double d = sqnorm(hp(L, i) - P0);
      if (d > v) { v = d; idx = i; }
    }
    waveArgmax(v, idx);
    i1 = idx;
    ext2 = v;
  }
  if (!(ext2 > 0)) { finish(kHullDegenerate); return; }
  const d3 dir = hp(L, i1) - P0;
  int i2 = 0;
  {
    double v = -1;
    int idx = 0x7fffffff;
    for (int i = l; i < n; i += 64) {
      const double d = sqnorm(cross(dir, hp(L, i) - P0));
      if (d > v) { v = d; idx = i; }
    }
    waveArgmax(v, idx);
    i2 = idx;
    if (!(v > 1e-24 * ext2 * ext2)) { finish(kHullDegenerate); return; }
  }
  const d3 nrm = cross(dir, hp(L, i2) - P0);
  int i3 = 0;
  double vol = 0;
  {
    double v = -1;
    int idx = 0x7fffffff;
    for (int i = l; i < n; i += 64) {
      const double d = fabs(dot(nrm, hp(L, i) - P0));
      if (d > v) { v = d; idx = i; }
    }
    waveArgmax(v, idx);
    i3 = idx;
    vol = v;
    if (!(vol > 1e-12 * sqrt(sqnorm(nrm)) * sqrt(ext2))) { finish(kHullDegenerate); return; }
  }
  // faces oriented outward: the fourth vertex lies below each face
  int a = i0, b = i1, c = i2;
  const int d = i3;
  if (dot(cross(hp(L, b) - hp(L, a), hp(L, c) - hp(L, a)), hp(L, d) - hp(L, a)) > 0) {
    const int t = b;
    b = c;
    c = t;
  }
  if (l == 0) {
    setFace(L, 0, a, b, c);
    setFace(L, 1, a, d, b);
    setFace(L, 2, b, d, c);
    setFace(L, 3, c, d, a);
    for (int f = 0; f < 4; ++f) L.fvis[f] = 0;
    L.conf[a] = L.conf[b] = L.conf[c] = L.conf[d] = -2;
    for (int k = 0; k < 8; ++k) L.cnt[k] = 0;
  }
  wsync();
  // neighbours of the four faces by edge matching
  if (l < 12) {
    const int f = l / 3, k = l % 3;
    const int ea = L.fv[k * F + f], eb = L.fv[((k + 1) % 3) * F + f];
    for (int g = 0; g < 4; ++g)
      for (int kk = 0; kk < 3; ++kk)
        if (L.fv[kk * F + g] == eb && L.fv[((kk + 1) % 3) * F + g] == ea) L.fnb[k * F + f] = static_cast<int16_t>(g);
  }
  int nf_used = 4, free_top = 0;
  wsync();
  // conflict faces of the other points
  for (int i = l; i < n; i += 64) {
    if (L.conf[i] == -2) continue;
    int cf = -1;
    for (int f = 0; f < 4 && cf < 0; ++f)
      if (faceSees(L, f, i)) cf = f;
    L.conf[i] = static_cast<int16_t>(cf);
  }
  wsync();
  int cursor = 0;
  bool failed = false;
  while (!failed) {
    // next outside point in index order
    int p = -1;
    while (cursor < n) {
      const int i = cursor + l;
      const bool out = i < n && L.conf[i] >= 0;
      const unsigned long long m = __ballot(out);
      if (m) {
        p = cursor + __ffsll(m) - 1;
        break;
      }
      cursor += 64;
    }
    if (p < 0) break;
    cursor = p;
    const int f0 = L.conf[p];
    // flood fill of the faces p sees, from its conflict face
    if (l == 0) {
      L.cnt[0] = 1;  // visible
      L.cnt[1] = 1;  // touched
      L.vis[0] = static_cast<int16_t>(f0);
      L.flist[0] = static_cast<int16_t>(f0);
      L.fvis[f0] = 1;
    }
    wsync();
    int fr0 = 0, fr1 = 1;
    while (fr0 < fr1) {
      for (int e = fr0 * 3 + l; e < fr1 * 3; e += 64) {
        const int f = L.vis[e / 3], k = e % 3;
        const int g = L.fnb[k * F + f];
        if (atomicCAS(&L.fvis[g], 0, 3) == 0) {
          const bool s = faceSees(L, g, p);
          L.fvis[g] = s ? 1 : 2;
          L.flist[atomicAdd(&L.cnt[1], 1)] = static_cast<int16_t>(g);
          if (s) L.vis[atomicAdd(&L.cnt[0], 1)] = static_cast<int16_t>(g);
        }
      }
      wsync();
      fr0 = fr1;
      fr1 = L.cnt[0];
      wsync();
    }
    const int nvis = fr1;
    // horizon: edges of visible faces whose neighbour is not visible
    if (l == 0) L.cnt[2] = 0;
    wsync();
    for (int e = l; e < nvis * 3; e += 64) {
      const int f = L.vis[e / 3], k = e % 3;
      const int g = L.fnb[k * F + f];
      if (L.fvis[g] == 1) continue;
      const int ea = L.fv[k * F + f], eb = L.fv[((k + 1) % 3) * F + f];
      int kk = -1;
      for (int j = 0; j < 3; ++j)
        if (L.fv[j * F + g] == eb && L.fv[((j + 1) % 3) * F + g] == ea) kk = j;
      const int h = atomicAdd(&L.cnt[2], 1);
      if (h < kHullMaxHorizon && kk >= 0) {
        L.hor[h] = static_cast<int16_t>(ea);
        L.hor[kHullMaxHorizon + h] = static_cast<int16_t>(eb);
        L.hor[2 * kHullMaxHorizon + h] = static_cast<int16_t>(g);
        L.hor[3 * kHullMaxHorizon + h] = static_cast<int16_t>(kk);
      } else {
        atomicOr(&L.cnt[4], 1);
      }
    }
    wsync();
    const int nh = L.cnt[2];
    if (L.cnt[4] || nh < 3) {
      failed = true;
      break;
    }
    // the border must be one loop: each vertex starts and ends one edge
    for (int h = l; h < nh; h += 64) {
      if (atomicCAS(&L.start[L.hor[h]], -1, h) != -1) atomicOr(&L.cnt[4], 2);
      if (atomicCAS(&L.endv[L.hor[kHullMaxHorizon + h]], -1, h) != -1) atomicOr(&L.cnt[4], 2);
    }
    wsync();
    bool bad = L.cnt[4] != 0;
    for (int h = l; h < nh && !bad; h += 64)
      if (L.endv[L.hor[h]] < 0) bad = true;
    if (__ballot(bad)) {
      failed = true;
      break;
    }
    // slots for the fan: freed visible faces first, then fresh ones (the
    // fan never needs more than the slots left: checked before any write)
    if (nh > free_top + nvis + (F - nf_used)) {
      failed = true;
      break;
    }
    for (int j = l; j < nvis; j += 64) L.freel[free_top + j] = L.vis[j];
    free_top += nvis;
    wsync();
    // new face for horizon edge h: (a, b, p), slot slotOf(h)
    auto slotOf = [&](int h) -> int { return h < free_top ? L.freel[free_top - 1 - h] : nf_used + (h - free_top); };
    for (int h = l; h < nh; h += 64) {
      const int s = slotOf(h);
      const int ea = L.hor[h], eb = L.hor[kHullMaxHorizon + h];
      setFace(L, s, ea, eb, p);
    }
    wsync();
    for (int h = l; h < nh; h += 64) {
      const int s = slotOf(h);
      const int ea = L.hor[h], eb = L.hor[kHullMaxHorizon + h];
      const int g = L.hor[2 * kHullMaxHorizon + h], kk = L.hor[3 * kHullMaxHorizon + h];
      L.fnb[s] = static_cast<int16_t>(g);                       // across (a, b)
      L.fnb[F + s] = static_cast<int16_t>(slotOf(L.start[eb]));  // across (b, p)
      L.fnb[2 * F + s] = static_cast<int16_t>(slotOf(L.endv[ea]));  // across (p, a)
      L.fnb[kk * F + g] = static_cast<int16_t>(s);
    }
    wsync();
    // points that saw a removed face look for a new face they see
    for (int i = l; i < n; i += 64) {
      const int cf = L.conf[i];
      if (cf < 0 || L.fvis[cf] != 1 || i == p) continue;
      int nc = -1;
      for (int h = 0; h < nh && nc < 0; ++h)
        if (faceSees(L, slotOf(h), i)) nc = slotOf(h);
      L.conf[i] = static_cast<int16_t>(nc);
    }
    wsync();
    // bookkeeping for the next step
    const int taken = nh < free_top ? nh : free_top;
    const int fresh = nh - taken;
    for (int t2 = l; t2 < L.cnt[1]; t2 += 64) {
      const int f = L.flist[t2];
      L.fvis[f] = L.fvis[f] == 1 ? -1 : 0;
    }
    wsync();
    for (int h = l; h < nh; h += 64) {
      L.fvis[slotOf(h)] = 0;
      L.start[L.hor[h]] = -1;
      L.endv[L.hor[kHullMaxHorizon + h]] = -1;
    }
    if (l == 0) L.conf[p] = -2;
    free_top -= taken;
    nf_used += fresh;
    if (nf_used > F) failed = true;
    wsync();
  }
  if (failed) { finish(kHullFailed); return; }
  // hull vertices are the points the remaining faces use (a point inserted
  // earlier is no longer one once every face around it was removed)
  for (int f = l; f < nf_used; f += 64)
    if (L.fvis[f] >= 0) L.start[L.fv[f]] = L.start[L.fv[F + f]] = L.start[L.fv[2 * F + f]] = -3;
  wsync();
  // renumber the hull vertices in ascending point order
  int nv = 0;
  for (int base = 0; base < n; base += 64) {
    const int i = base + l;
    const bool v = i < n && L.start[i] == -3;
    const unsigned long long m = __ballot(v);
    if (v) {
      const int id = nv + __popcll(m & ((1ull << l) - 1));
      L.start[i] = id;
      hull_idx[o + id] = static_cast<uint32_t>(i);
    }
    nv += __popcll(m);
  }
  wsync();
  int nt = 0;
  for (int base = 0; base < nf_used; base += 64) {
    const int f = base + l;
    const bool alive = f < nf_used && L.fvis[f] >= 0;
    const unsigned long long m = __ballot(alive);
    if (alive) {
      const int t = nt + __popcll(m & ((1ull << l) - 1));
      int32_t* tp = tris + 6 * o + 3 * static_cast<uint64_t>(t);
      tp[0] = L.start[L.fv[f]];
      tp[1] = L.start[L.fv[F + f]];
      tp[2] = L.start[L.fv[2 * F + f]];
    }
    nt += __popcll(m);
  }
  if (l == 0) {
    status[set] = kHullOk;
    hull_cnt[set] = static_cast<uint32_t>(nv);
    tri_cnt[set] = static_cast<uint32_t>(nt);
  }
}

int hullsLaunch(mfcl_ctx* ctx, const float* pts, const uint64_t* pt_off, const uint32_t* pt_cnt, uint32_t n_sets,
                int max_points, uint32_t* hull_cnt, uint32_t* hull_idx, uint32_t* tri_cnt, int32_t* tris,
                int32_t* status, hipStream_t s) {
  if (n_sets == 0) return MFCL_OK;
  if (max_points < 4 || max_points > kHullMaxPoints) {
    ctx->err = "max_points out of range [4, " + std::to_string(kHullMaxPoints) + "]";
    return MFCL_E_INVALID;
  }
  const size_t lds = hullLdsBytes(max_points);
  if (lds > static_cast<size_t>(ctx->max_lds_per_block)) {
    ctx->err = "hull workspace exceeds the LDS of a workgroup";
    return MFCL_E_UNSUPPORTED;
  }
  hullKernel<<<n_sets, 64, lds, s>>>(pts, pt_off, pt_cnt, nullptr, n_sets, max_points, hull_cnt, hull_idx, tri_cnt,
                                     tris, status);
  MFCL_HIP_TRY(ctx, hipGetLastError());
  return MFCL_OK;
}

// The same with the point counts also on the host: sets are grouped by size
// class (up to 32, 64, 128, 256, 512, 768 points) and each class runs in LDS
// slices of its own size, so small sets are not held to the largest set's
// slice (more wavefronts per CU).  order_dev: n_sets device words for the
// grouped set list (filled here, read by the kernels).
int hullsLaunchClassed(mfcl_ctx* ctx, const float* pts, const uint64_t* pt_off, const uint32_t* pt_cnt,
                       const uint32_t* pt_cnt_host, uint32_t n_sets, uint32_t* order_dev, uint32_t* hull_cnt,
                       uint32_t* hull_idx, uint32_t* tri_cnt, int32_t* tris, int32_t* status, hipStream_t s) {
  if (n_sets == 0) return MFCL_OK;
  constexpr int kClasses = 6;
  const int caps[kClasses] = {32, 64, 128, 256, 512, kHullMaxPoints};
  std::vector<uint32_t> order(n_sets);
  uint32_t count[kClasses] = {}, start[kClasses] = {};
  auto cls = [&](uint32_t n) {
    int c = 0;
    while (c < kClasses - 1 && static_cast<uint32_t>(caps[c]) < n) ++c;
    return c;
  };
  for (uint32_t i = 0; i < n_sets; ++i) ++count[cls(pt_cnt_host[i])];
  for (int c = 1; c < kClasses; ++c) start[c] = start[c - 1] + count[c - 1];
  uint32_t fill[kClasses];
  for (int c = 0; c < kClasses; ++c) fill[c] = start[c];
  // largest classes first, so the longest sets do not start last
  for (uint32_t i = 0; i < n_sets; ++i) order[fill[cls(pt_cnt_host[i])]++] = i;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(order_dev, order.data(), 4 * static_cast<size_t>(n_sets), hipMemcpyHostToDevice, s));
  for (int c = kClasses - 1; c >= 0; --c) {
    if (count[c] == 0) continue;
    const size_t lds = hullLdsBytes(caps[c]);
    if (lds > static_cast<size_t>(ctx->max_lds_per_block)) {
      ctx->err = "hull workspace exceeds the LDS of a workgroup";
      return MFCL_E_UNSUPPORTED;
    }
    hullKernel<<<count[c], 64, lds, s>>>(pts, pt_off, pt_cnt, order_dev + start[c], count[c], caps[c], hull_cnt,
                                         hull_idx, tri_cnt, tris, status);
  }
  MFCL_HIP_TRY(ctx, hipGetLastError());
  // the order list is read by the kernels: keep it until they have run
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
  return MFCL_OK;
}

// ---- shape table of the hulls (mfcl_upload_convex_points) ----------------
//
// One wavefront per hull, after hullKernel: gathers the hull vertices in
// ascending point order into the vertex table, and computes what
// mfcl_upload_shapes computes on the host for a convex shape (interior point
// = vertex mean, convex-inl.h:64-70; the largest |coordinate| for the FP32
// support-scan bound; the 16-float descriptor of the warm-start network) and
// what mfcl_upload_convex_faces builds from the faces for hulls of more than
// 32 vertices (the sorted neighbour list of every vertex in the reference's
// encoding, Convex::FindVertexNeighbors, convex-inl.h:377-409).  Sums run
// sequentially in vertex order, one lane per sum, so they round as the
// host's loops do.
__global__ void __launch_bounds__(64) convexTableKernel(const float* pts, const uint64_t* pt_off,
                                                        const uint32_t* hull_cnt, const uint32_t* hull_idx,
                                                        const uint32_t* tri_cnt, const int32_t* tris,
                                                        const uint32_t* voff, const int64_t* nbr_off, int cap,
                                                        float4* verts, float4* param, double4* interior, float* desc,
                                                        int32_t* nbr) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) char tab_lds[];
  const uint32_t set = blockIdx.x;
  const int l = laneId64();
  const int nv = static_cast<int>(hull_cnt[set]);
  const int nt = static_cast<int>(tri_cnt[set]);
  const uint64_t o = pt_off[set];
  double* sums = reinterpret_cast<double*>(tab_lds);  // 16 doubles
  float* x = reinterpret_cast<float*>(sums + 16);
  float* y = x + cap;
  float* z = y + cap;
  int32_t* rec = reinterpret_cast<int32_t*>(z + cap);
  int32_t* cur = rec + cap;
  float4* vo = verts + voff[set];
  float rb = 0.f;
  for (int k = l; k < nv; k += 64) {
    const uint64_t q = o + hull_idx[o + k];
    const float px = pts[3 * q], py = pts[3 * q + 1], pz = pts[3 * q + 2];
    x[k] = px;
    y[k] = py;
    z[k] = pz;
    vo[k] = make_float4(px, py, pz, 0.f);
    rb = fmaxf(rb, fmaxf(fabsf(px), fmaxf(fabsf(py), fabsf(pz))));
  }
  float mx[3] = {0.f, 0.f, 0.f};
  for (int k = l; k < nv; k += 64) {
    mx[0] = fmaxf(mx[0], fabsf(x[k]));
    mx[1] = fmaxf(mx[1], fabsf(y[k]));
    mx[2] = fmaxf(mx[2], fabsf(z[k]));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    rb = fmaxf(rb, __shfl_xor(rb, off));
    mx[0] = fmaxf(mx[0], __shfl_xor(mx[0], off));
    mx[1] = fmaxf(mx[1], __shfl_xor(mx[1], off));
    mx[2] = fmaxf(mx[2], __shfl_xor(mx[2], off));
  }
  __syncthreads();
  // coordinate sums (lanes 0-2), radius sum and maximum (lane 3)
  if (l < 3) {
    const float* c = l == 0 ? x : (l == 1 ? y : z);
    double sacc = 0;
    for (int k = 0; k < nv; ++k) sacc += c[k];
    sums[l] = sacc;
  } else if (l == 3) {
    double rmax = 0, rsum = 0;
    for (int k = 0; k < nv; ++k) {
      const double px = x[k], py = y[k], pz = z[k];
      const double r = sqrt(px * px + py * py + pz * pz);
      rmax = fmax(rmax, r);
      rsum += r;
    }
    sums[3] = rmax;
    sums[4] = rsum;
  }
  __syncthreads();
  const double mean[3] = {sums[0] / nv, sums[1] / nv, sums[2] / nv};
  // second moments about the mean (lanes 0-5)
  if (l < 6) {
    const int i0 = l < 3 ? l : (l == 5 ? 1 : 0);
    const int i1 = l < 3 ? l : (l == 3 ? 1 : 2);
    const float* c0 = i0 == 0 ? x : (i0 == 1 ? y : z);
    const float* c1 = i1 == 0 ? x : (i1 == 1 ? y : z);
    double m = 0;
    for (int k = 0; k < nv; ++k) {
      const double u = c0[k] - mean[i0], w = c1[k] - mean[i1];
      m += u * w;
    }
    sums[8 + l] = m;
  }
  __syncthreads();
  if (l == 0) {
    const double inv = 1.0 / nv;
    interior[set] = double4{sums[0] * inv, sums[1] * inv, sums[2] * inv, 0};
    param[set] = make_float4(0.f, 0.f, 0.f, rb);
  }
  if (l < 16) {
    float d;
    if (l < 3) d = l == 0 ? mx[0] : (l == 1 ? mx[1] : mx[2]);
    else if (l < 6) d = static_cast<float>(mean[l - 3]);
    else if (l < 12) d = static_cast<float>(sums[8 + l - 6] / nv);
    else if (l == 12) d = static_cast<float>(log2(static_cast<double>(nv)) / 9.0);
    else if (l == 13) d = static_cast<float>(sums[3]);
    else if (l == 14) d = static_cast<float>(sums[4] / nv);
    else d = 1.f;
    desc[16 * static_cast<uint64_t>(set) + l] = d;
  }
  const int64_t no = nbr_off[set];
  if (no < 0) return;
  // neighbour records: on a closed, consistently oriented triangulation the
  // neighbours of a vertex are the successors of it in its triangles, each
  // once, so its degree is its triangle count
  int32_t* tab = nbr + no;
  const int32_t* tr = tris + 6 * o;
  for (int v = l; v < nv; v += 64) cur[v] = 0;
  __syncthreads();
  for (int t = l; t < nt; t += 64) {
    atomicAdd(&cur[tr[3 * t]], 1);
    atomicAdd(&cur[tr[3 * t + 1]], 1);
    atomicAdd(&cur[tr[3 * t + 2]], 1);
  }
  __syncthreads();
  int run = nv;
  for (int base = 0; base < nv; base += 64) {
    const int v = base + l;
    const int deg = v < nv ? cur[v] : 0;
    const int len = v < nv ? 1 + deg : 0;
    int incl = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(incl, off);
      if (l >= off) incl += t;
    }
    if (v < nv) {
      const int r = run + incl - len;
      rec[v] = r;
      tab[v] = r;
      tab[r] = deg;
    }
    run += __shfl(incl, 63);
  }
  __syncthreads();
  for (int v = l; v < nv; v += 64) cur[v] = 0;
  __syncthreads();
  for (int t = l; t < nt; t += 64) {
    const int a = tr[3 * t], b = tr[3 * t + 1], c = tr[3 * t + 2];
    tab[rec[a] + 1 + atomicAdd(&cur[a], 1)] = b;
    tab[rec[b] + 1 + atomicAdd(&cur[b], 1)] = c;
    tab[rec[c] + 1 + atomicAdd(&cur[c], 1)] = a;
  }
  __syncthreads();
  for (int v = l; v < nv; v += 64) {
    int32_t* r = tab + rec[v] + 1;
    const int deg = cur[v];
    for (int i = 1; i < deg; ++i) {
      const int32_t key = r[i];
      int j = i - 1;
      while (j >= 0 && r[j] > key) {
        r[j + 1] = r[j];
        --j;
      }
      r[j + 1] = key;
    }
  }
}

namespace {
// Device allocations of one call, released on every return.
struct TempDev {
  std::vector<void*> p;
  ~TempDev() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  T* get(mfcl_ctx* ctx, size_t count, int& rc) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      ctx->err = "out of device memory building the convex shape table";
      rc = MFCL_E_NOMEM;
      return nullptr;
    }
    p.push_back(q);
    return static_cast<T*>(q);
  }
};

template <class T>
int allocTable(mfcl_ctx* ctx, T** dst, size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    ctx->err = "out of device memory building the convex shape table";
    return MFCL_E_NOMEM;
  }
  *dst = static_cast<T*>(q);
  return MFCL_OK;
}
}  // namespace

int uploadConvexPoints(mfcl_ctx* ctx, const float* points, const uint64_t* pt_off, const uint32_t* pt_cnt,
                       uint32_t n_sets, uint32_t* hull_cnt_out) {
  uint64_t total = 0;
  uint32_t maxp = 4;
  for (uint32_t i = 0; i < n_sets; ++i) {
    if (pt_cnt[i] > static_cast<uint32_t>(kHullMaxPoints)) {
      ctx->err = "point set " + std::to_string(i) + " has more than " + std::to_string(kHullMaxPoints) + " points";
      return MFCL_E_INVALID;
    }
    total = std::max<uint64_t>(total, pt_off[i] + pt_cnt[i]);
    maxp = std::max(maxp, pt_cnt[i]);
  }
  const hipStream_t s = ctx->stream;
  int rc = MFCL_OK;
  TempDev tmp;
  float* dp = tmp.get<float>(ctx, 3 * total, rc);
  uint64_t* doff = rc ? nullptr : tmp.get<uint64_t>(ctx, n_sets, rc);
  uint32_t* dcnt = rc ? nullptr : tmp.get<uint32_t>(ctx, n_sets, rc);
  uint32_t* dhc = rc ? nullptr : tmp.get<uint32_t>(ctx, n_sets, rc);
  uint32_t* dtc = rc ? nullptr : tmp.get<uint32_t>(ctx, n_sets, rc);
  int32_t* dst = rc ? nullptr : tmp.get<int32_t>(ctx, n_sets, rc);
  uint32_t* dhi = rc ? nullptr : tmp.get<uint32_t>(ctx, total, rc);
  int32_t* dtr = rc ? nullptr : tmp.get<int32_t>(ctx, 6 * total, rc);
  uint32_t* dord = rc ? nullptr : tmp.get<uint32_t>(ctx, n_sets, rc);
  if (rc) return rc;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dp, points, 12 * total, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(doff, pt_off, 8 * static_cast<size_t>(n_sets), hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dcnt, pt_cnt, 4 * static_cast<size_t>(n_sets), hipMemcpyHostToDevice, s));
  if ((rc = hullsLaunchClassed(ctx, dp, doff, dcnt, pt_cnt, n_sets, dord, dhc, dhi, dtc, dtr, dst, s))) return rc;
  std::vector<uint32_t> hc(n_sets), tc(n_sets);
  std::vector<int32_t> st(n_sets);
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(hc.data(), dhc, 4 * static_cast<size_t>(n_sets), hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(tc.data(), dtc, 4 * static_cast<size_t>(n_sets), hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(st.data(), dst, 4 * static_cast<size_t>(n_sets), hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
  for (uint32_t i = 0; i < n_sets; ++i) {
    if (st[i] == kHullOk) continue;
    static const char* why[] = {"", "is degenerate (fewer than 4 points, or all on one plane)",
                                "has too many points", "has no hull within the tolerance (near-degenerate input)"};
    ctx->err = "point set " + std::to_string(i) + " " + why[st[i] >= 1 && st[i] <= 3 ? st[i] : 3];
    return MFCL_E_INVALID;
  }
  // vertex and adjacency offsets; a hull's record area holds nv offsets and
  // nv records of 1 + degree entries, and the degrees sum to 3 nt
  std::vector<uint32_t> voff(n_sets);
  std::vector<int64_t> noff(n_sets, -1);
  uint64_t nverts = 0, nnbr = 0;
  for (uint32_t i = 0; i < n_sets; ++i) {
    voff[i] = static_cast<uint32_t>(nverts);
    nverts += hc[i];
    if (hc[i] > 32) {
      noff[i] = static_cast<int64_t>(nnbr);
      nnbr += 2 * static_cast<uint64_t>(hc[i]) + 3 * static_cast<uint64_t>(tc[i]);
    }
  }
  if (nverts > 0xFFFFFFFFull || nnbr > 0x7FFFFFFFull) {
    ctx->err = "convex tables exceed 2^32 vertices or 2^31 adjacency entries";
    return MFCL_E_INVALID;
  }
  uint32_t* dvoff = tmp.get<uint32_t>(ctx, n_sets, rc);
  if (rc) return rc;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dvoff, voff.data(), 4 * static_cast<size_t>(n_sets), hipMemcpyHostToDevice, s));
  freeShapes(ctx);
  const std::vector<int32_t> type(n_sets, MFCL_SHAPE_CONVEX);
  auto fail = [&](int r) {
    freeShapes(ctx);
    return r;
  };
  if ((rc = allocTable(ctx, &ctx->d_type, n_sets)) || (rc = allocTable(ctx, &ctx->d_param, n_sets)) ||
      (rc = allocTable(ctx, &ctx->d_voff, n_sets)) || (rc = allocTable(ctx, &ctx->d_vcnt, n_sets)) ||
      (rc = allocTable(ctx, &ctx->d_verts, nverts)) || (rc = allocTable(ctx, &ctx->d_interior, n_sets)) ||
      (rc = allocTable(ctx, &ctx->d_desc, 16 * static_cast<size_t>(n_sets))) ||
      (rc = allocTable(ctx, &ctx->d_nbr, nnbr + 32)) || (rc = allocTable(ctx, &ctx->d_nbr_off, n_sets)) ||
      (rc = allocTable(ctx, &ctx->d_seed, 8 * static_cast<size_t>(n_sets))))
    return fail(rc);
  // walk blocks: 128 B per vertex; when the device cannot hold them the
  // walks read the adjacency table instead (same results, slower)
  if (hipMalloc(reinterpret_cast<void**>(&ctx->d_wblk), 8 * sizeof(float4) * std::max<uint64_t>(nverts, 1)) !=
      hipSuccess) {
    ctx->d_wblk = nullptr;
    (void)hipGetLastError();
  }
  auto chk = [&](hipError_t e) {
    if (e == hipSuccess) return MFCL_OK;
    ctx->err = std::string("HIP error: ") + hipGetErrorString(e);
    return MFCL_E_HIP;
  };
  if ((rc = chk(hipMemcpyAsync(ctx->d_type, type.data(), 4 * static_cast<size_t>(n_sets), hipMemcpyHostToDevice, s))) ||
      (rc = chk(hipMemcpyAsync(ctx->d_voff, voff.data(), 4 * static_cast<size_t>(n_sets), hipMemcpyHostToDevice, s))) ||
      (rc = chk(hipMemcpyAsync(ctx->d_vcnt, dhc, 4 * static_cast<size_t>(n_sets), hipMemcpyDeviceToDevice, s))) ||
      (rc = chk(hipMemcpyAsync(ctx->d_nbr_off, noff.data(), 8 * static_cast<size_t>(n_sets), hipMemcpyHostToDevice,
                               s))) ||
      (rc = chk(hipMemsetAsync(ctx->d_nbr + nnbr, 0, 32 * sizeof(int32_t), s))))
    return fail(rc);
  const size_t lds = 20 * static_cast<size_t>(maxp) + 16 * sizeof(double) + 64;
  convexTableKernel<<<n_sets, 64, lds, s>>>(dp, doff, dhc, dhi, dtc, dtr, dvoff, ctx->d_nbr_off,
                                             static_cast<int>(maxp), ctx->d_verts, ctx->d_param, ctx->d_interior,
                                             ctx->d_desc, ctx->d_nbr);
  if ((rc = chk(hipGetLastError()))) return fail(rc);
  ctx->n_shapes = n_sets;
  ctx->n_verts_alloc = std::max<uint64_t>(nverts, 1);
  ctx->h_vcnt = hc;
  ctx->h_voff = voff;
  ctx->h_type = type;
  if ((rc = computeWalkSeeds(ctx))) return fail(rc);
  if (hull_cnt_out) std::memcpy(hull_cnt_out, hc.data(), 4 * static_cast<size_t>(n_sets));
  return MFCL_OK;
}

}  // namespace mfcl
