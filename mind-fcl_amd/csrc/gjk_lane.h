// GJK with one pair per lane and wavefront-cooperative hull scans
// (GJKSolver::shapeSignedDistance / shapeIntersect semantics,
// gjk_solver-inl.h:93-155 and :820-874; GJK gjk.hpp:11-140, distance
// gjk_distance.hpp:11-105).
//
// The FP64 simplex logic of GJK is scalar per pair, so it runs with one pair
// per lane: the 64 pairs of a wavefront share every FP64 instruction it
// issues.  The support searches over convex hulls are the data-parallel
// part; they are served cooperatively -- whenever lanes need a support
// point, the wavefront scans each requesting lane's hull in turn with all
// 64 lanes (coalesced loads, a butterfly argmax, the same FP32 scan with a
// rounding bound and FP64 rescan as the group kernel), so every support
// point is the same vertex.  The GJK loop is a state machine so that all
// lanes meet at the one support stage:
//   support stage (converged) -> advance every lane by one step -> ...
// Primitive supports are closed forms evaluated by the requesting lane.
#pragma once

#include "narrow_dev.h"

namespace mfcl {

// A simplex holds at most 4 vertices, so freeSlot() returns at most 4.
constexpr int kLaneSlots = 5;

// Per-wavefront LDS block ([field][lane], 8.5 KB): the relative transform
// of each lane's pair and the hull
// vertex indices of the support points recorded in each direction slot
// (-1 for a primitive).
struct LaneWs {
  double tf[12][64];  // R row-major (9), t (3)
  int i0[kLaneSlots][64], i1[kLaneSlots][64];
};
// The directions themselves live in device memory, [slot][xyz][lane] per
// wavefront: they are written once per support point and read only for
// primitive witnesses and for the EPA queue.
constexpr int kLaneDirDoubles = kLaneSlots * 3 * 64;

// The Minkowski difference of one lane's pair.  Only the two shape ids and
// types stay in registers; shape parameters are re-read from the shape
// tables and the transform from LDS where needed, which keeps the long-lived
// register state of the lane machine small (occupancy hides the latency of
// the hull scans).  The rotation into frame 1 (toshape1 = R_b^T R_a) is the
// exact transpose of t0.R = R_a^T R_b (the same products summed in the same
// order), so it is applied as tmul(t0.R, .).
struct MinkLane {
  ShapeTables T;
  uint32_t ia, ib;
  int type0, type1;
  MFCL_LDS LaneWs* w;
  double* dirs;
  int lane;
  // reloadFence(): re-read at every use; hoisted out of the lane machine's
  // loop these would occupy registers for its whole duration
  MFCL_D tf3 tf0() const {
    reloadFence();
    tf3 t;
    t.R.m00 = w->tf[0][lane]; t.R.m01 = w->tf[1][lane]; t.R.m02 = w->tf[2][lane];
    t.R.m10 = w->tf[3][lane]; t.R.m11 = w->tf[4][lane]; t.R.m12 = w->tf[5][lane];
    t.R.m20 = w->tf[6][lane]; t.R.m21 = w->tf[7][lane]; t.R.m22 = w->tf[8][lane];
    t.t = d3{w->tf[9][lane], w->tf[10][lane], w->tf[11][lane]};
    return t;
  }
  MFCL_D void setTf0(const tf3& t) const {
    w->tf[0][lane] = t.R.m00; w->tf[1][lane] = t.R.m01; w->tf[2][lane] = t.R.m02;
    w->tf[3][lane] = t.R.m10; w->tf[4][lane] = t.R.m11; w->tf[5][lane] = t.R.m12;
    w->tf[6][lane] = t.R.m20; w->tf[7][lane] = t.R.m21; w->tf[8][lane] = t.R.m22;
    w->tf[9][lane] = t.t.x; w->tf[10][lane] = t.t.y; w->tf[11][lane] = t.t.z;
  }
  MFCL_D DevShape prim(int k) const {
    reloadFence();
    const float4 p = T.param[k ? ib : ia];
    DevShape s;
    s.type = k ? type1 : type0;
    s.p0 = p.x; s.p1 = p.y; s.p2 = p.z;
    s.verts = nullptr; s.nv = 0; s.rbound = p.w;
    s.lv = nullptr; s.ls = 0;
    s.nbr = nullptr; s.seed = nullptr; s.wblk = nullptr;
    return s;
  }
  MFCL_D const float4* verts(int k) const {
    reloadFence();
    return T.verts + T.vert_off[k ? ib : ia]; }
  MFCL_D void setDir(int id, d3 d) const {
    dirs[(3 * id + 0) * 64 + lane] = d.x;
    dirs[(3 * id + 1) * 64 + lane] = d.y;
    dirs[(3 * id + 2) * 64 + lane] = d.z;
  }
  MFCL_D d3 dirOf(int id) const {
    return d3{dirs[(3 * id + 0) * 64 + lane], dirs[(3 * id + 1) * 64 + lane], dirs[(3 * id + 2) * 64 + lane]};
  }
  // support0(dir of slot id), from the recorded hull vertex when there is one
  MFCL_D d3 point0(int id) const {
    const int i0 = w->i0[id][lane];
    if (i0 >= 0) {
      const float4 q = verts(0)[i0];
      return d3{q.x, q.y, q.z};
    }
    return primSupport(prim(0), dirOf(id));
  }
  // support1(-dir of slot id), likewise
  MFCL_D d3 point1(int id) const {
    const int i1 = w->i1[id][lane];
    const tf3 t0 = tf0();
    if (i1 >= 0) {
      const float4 q = verts(1)[i1];
      return apply(t0, d3{q.x, q.y, q.z});
    }
    return apply(t0, primSupport(prim(1), tmul(t0.R, -dirOf(id))));
  }
};

MFCL_D int readLaneI(int v, int src) { return __builtin_amdgcn_readlane(v, src); }
MFCL_D uint64_t readLaneU64(uint64_t v, int src) {
  const unsigned lo = __builtin_amdgcn_readlane(static_cast<unsigned>(v), src);
  const unsigned hi = __builtin_amdgcn_readlane(static_cast<unsigned>(v >> 32), src);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
MFCL_D double readLaneD(double v, int src) {
  return __longlong_as_double(static_cast<long long>(readLaneU64(static_cast<uint64_t>(__double_as_longlong(v)), src)));
}

// FP64 pass of hullArgmaxOOL<64> (first index attaining the FP64 maximum of
// dot(d, v)), inline: a call inside the lane machine's loop would make every
// value live across it occupy callee-saved registers.
MFCL_D int waveRescan64(const float4* verts, int nv, double dx, double dy, double dz, int lane) {
  double bd = -INFINITY;
  int bi = 0x7fffffff;
  for (int i = lane; i < nv; i += 64) {
    const float4 q = verts[i];
    const double val = dx * static_cast<double>(q.x) + dy * static_cast<double>(q.y) + dz * static_cast<double>(q.z);
    if (val > bd) {
      bd = val;
      bi = i;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ob = __shfl_xor(bd, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    if (ob > bd || (ob == bd && oi < bi)) {
      bd = ob;
      bi = oi;
    }
  }
  return bi < nv ? bi : 0;
}

// Argmax of dot(d_j, v) over the hulls of up to K requesting lanes at once
// (wave-uniform parameters): lane l visits vertices l, l + 64, ... of every
// hull, with the K hulls' loads issued together so their latencies overlap;
// then one butterfly per hull.  The visiting order, the merge and the FP32
// bound / FP64 rescan rule are those of hullArgmaxOOL<64>, so the indices
// are the same.  Takes requests from `m` (lowest lanes first); the lane that
// made request j receives its index in `res`.
template <int K>
MFCL_D void coopScanBatch(uint64_t& m, uint64_t vp, int nv, float rb, d3 d, int lane, int& res,
                          uint64_t& scanned) {
  int p[K], pn[K];
  const float4* pv[K];
  float prb[K];
  double dx[K], dy[K], dz[K];
  int nmax = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    p[j] = m ? __builtin_ctzll(m) : -1;
    m &= m - 1;
    const int src = p[j] < 0 ? 0 : p[j];
    pv[j] = reinterpret_cast<const float4*>(readLaneU64(vp, src));
    pn[j] = p[j] < 0 ? 0 : readLaneI(nv, src);
    prb[j] = __int_as_float(readLaneI(__float_as_int(rb), src));
    dx[j] = readLaneD(d.x, src);
    dy[j] = readLaneD(d.y, src);
    dz[j] = readLaneD(d.z, src);
    nmax = max(nmax, pn[j]);
  }
  float best[K], second[K];
  int bi[K];
  {
#pragma clang fp contract(fast)
    float fx[K], fy[K], fz[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
      best[j] = -INFINITY;
      second[j] = -INFINITY;
      bi[j] = 0x7fffffff;
      fx[j] = static_cast<float>(dx[j]);
      fy[j] = static_cast<float>(dy[j]);
      fz[j] = static_cast<float>(dz[j]);
    }
    for (int i = lane; i < nmax; i += 64) {
      float4 q[K];
#pragma unroll
      for (int j = 0; j < K; ++j) q[j] = i < pn[j] ? pv[j][i] : make_float4(0, 0, 0, 0);
#pragma unroll
      for (int j = 0; j < K; ++j) {
        if (i < pn[j]) {
          const float val = fx[j] * q[j].x + fy[j] * q[j].y + fz[j] * q[j].z;
          if (val > best[j]) {
            second[j] = best[j];
            best[j] = val;
            bi[j] = i;
          } else if (val > second[j]) {
            second[j] = val;
          }
        }
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const float ob = __shfl_xor(best[j], off, 64);
      const float os = __shfl_xor(second[j], off, 64);
      const int oi = __shfl_xor(bi[j], off, 64);
      if (ob > best[j] || (ob == best[j] && oi < bi[j])) {
        second[j] = fmaxf(fmaxf(second[j], os), best[j]);
        best[j] = ob;
        bi[j] = oi;
      } else {
        second[j] = fmaxf(fmaxf(second[j], os), ob);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (p[j] < 0) continue;
    const float bound = 2.0f * 8.0f * 5.9604645e-08f * prb[j] *
                        (fabsf(static_cast<float>(dx[j])) + fabsf(static_cast<float>(dy[j])) +
                         fabsf(static_cast<float>(dz[j])));
    int idx;
    if (best[j] - second[j] > bound) {
      idx = bi[j] < pn[j] ? bi[j] : 0;
      scanned += static_cast<uint64_t>(pn[j]);
    } else {
      idx = waveRescan64(pv[j], pn[j], dx[j], dy[j], dz[j], lane);
      scanned += static_cast<uint64_t>(2 * pn[j]);
    }
    if (lane == p[j]) res = idx;
  }
}

// Support stage: every lane with `need` gets support(d) of its Minkowski
// difference, recorded in its slot `id` (direction and vertex indices).
// Wavefront-cooperative: all 64 lanes must call this together.  `scanned`
// (wave-uniform) counts the hull vertices read by the cooperative scans,
// `walked` (per lane) those read by the lane's vertex walks.
constexpr int kScanBatch = 8;
MFCL_D void laneSupport(bool need, const MinkLane& md, d3 d, int id, MV& out, uint64_t& scanned, int& walked) {
  const d3 d1 = tmul(md.tf0().R, -d);
  int idx0 = -1, idx1 = -1;
#pragma nounroll
  for (int k = 0; k < 2; ++k) {
    const bool hull = need && (k ? md.type1 : md.type0) == kConvex;
    const uint32_t si = k ? md.ib : md.ia;
    const bool want = hull && !(md.T.nbr_off && md.T.nbr_off[si] >= 0);
    uint64_t vp = 0;
    int nv = 0;
    float rb = 0;
    if (want) {
      reloadFence();
      vp = reinterpret_cast<uint64_t>(md.T.verts + md.T.vert_off[si]);
      nv = static_cast<int>(md.T.vert_cnt[si]);
      rb = md.T.param[si].w;
    }
    const d3 dd = k ? d1 : d;
    // naive hulls: cooperative scans; walking hulls: each lane walks its
    // own (below)
    uint64_t m = __ballot(want);
    int res = -1;
    while (m) coopScanBatch<kScanBatch>(m, vp, nv, rb, dd, md.lane, res, scanned);
    if (k) idx1 = res;
    else idx0 = res;
  }
  {
    const bool w0 = need && md.type0 == kConvex && md.T.nbr_off && md.T.nbr_off[md.ia] >= 0;
    const bool w1 = need && md.type1 == kConvex && md.T.nbr_off && md.T.nbr_off[md.ib] >= 0;
    if (w0 || w1) {
      const DevShape s0 = loadShape(md.T, md.ia), s1 = loadShape(md.T, md.ib);
      // one walk after the other: the lanes' independent walks already keep
      // enough memory requests in flight, and lock-step pairs cost registers
      if (w0) idx0 = hullWalk(s0, d, &walked);
      if (w1) idx1 = hullWalk(s1, d1, &walked);
    }
  }
  if (need) {
    d3 a, x;
    if (idx0 >= 0) {
      const float4 q = md.verts(0)[idx0];
      a = d3{q.x, q.y, q.z};
    } else {
      a = primSupport(md.prim(0), d);
    }
    if (idx1 >= 0) {
      const float4 q = md.verts(1)[idx1];
      x = d3{q.x, q.y, q.z};
    } else {
      x = primSupport(md.prim(1), d1);
    }
    out.v = a - apply(md.tf0(), x);
    out.id = id;
    md.setDir(id, d);
    md.w->i0[id][md.lane] = idx0;
    md.w->i1[id][md.lane] = idx1;
  }
}

// States of the lane machine; the ones marked * consume a support point.
enum LaneState : int {
  kLsB0 = 0,       // * first support point of GJK (gjk.hpp:26-38)
  kLsB1 = 1,       // * boolean iteration (gjk.hpp:40-93)
  kLsSep = 2,      //   separated: start the distance loop (gjk_distance.hpp:11-40)
  kLsD1 = 3,       // * distance iteration (gjk_distance.hpp:42-100)
  kLsExtract = 4,  //   witnesses without sub-simplex search (gjk_distance.hpp:375-486)
  kLsNoSep = 5,    //   not separated: witnesses with sub-simplex search (gjk_distance.hpp:488-559)
  kLsDone = 6
};

// gjkEvaluate + gjkMinDistance of narrow_dev.h, one lane per pair, every
// branch and operation in the same order (the results are bit-identical),
// as a resumable machine: laneInit starts a lane's pair, laneRun advances
// the lanes of the wavefront together, laneFinish extracts a finished
// lane's witnesses.  All 64 lanes of the wavefront must call laneRun; lanes
// without a pair (or done) only take part in the support stages.
struct LaneM {
  int st;
  bool req;
  int rqid, it, ext;
  double book;
  d3 cur, dir;
  MV v;
  Smp s;
  GjkOut o;
};

MFCL_D void laneInit(LaneM& M, bool active, d3 guess) {
  smpInit(M.s);
  M.o.status = -1;
  M.o.sep_valid = false;
  M.o.iters = 0;
  M.o.dist_iters = 0;
  M.o.p0 = d3{0, 0, 0};
  M.o.p1 = d3{0, 0, 0};
  M.dir = guess;
  if (sqnorm(M.dir) <= 0.0) M.dir = d3{1, 0, 0};
  M.dir = normalized(M.dir);
  M.st = active ? kLsB0 : kLsDone;
  M.req = active;
  M.rqid = 0;
  M.it = 0;
  M.ext = 0;
  M.book = 0;
  M.cur = d3{0, 0, 0};
  mvZero(M.v);
}

// Advances the machine until at most min_running lanes of the wavefront are
// still running (0: until every lane is done).  A lane's pending support
// request carries over to the next call.
MFCL_D void laneRun(LaneM& M, const MinkLane& md, double tol, int max_iter, bool want_distance, uint64_t& scanned,
                    int& walked, int min_running) {
  const double tol2 = tol * tol;
  while (true) {
    const unsigned long long running = __ballot(M.st != kLsDone);
    if (running == 0ull || __popcll(running) <= min_running) break;
    laneSupport(M.req, md, M.dir, M.rqid, M.v, scanned, walked);
    M.req = false;
    if (M.st == kLsB0) {
      M.s.rank = -1;
      smpAdd(M.s, M.v);
      if (sqnorm(M.v.v) <= tol2) {
        M.o.status = kGjkIntersect;
        M.st = kLsDone;
      } else if (dot(M.v.v, M.dir) < 0) {
        M.st = kLsSep;
      } else {
        M.dir = -M.dir;
        M.o.status = kGjkIterLimit;
        M.it = 0;
        if (M.it < max_iter) {
          M.o.iters = 1;
          M.rqid = freeSlot(M.s);
          M.req = true;
          M.st = kLsB1;
        } else {
          M.st = kLsNoSep;
        }
      }
    } else if (M.st == kLsB1) {
      if (dot(M.v.v, M.dir) < 0) {
        M.st = kLsSep;
      } else if (smpDup(M.s, M.v.v, tol2)) {
        M.o.status = kGjkNoProgress;
        M.st = kLsNoSep;
      } else if (sqnorm(M.v.v) <= tol2) {
        smpAdd(M.s, M.v);
        M.o.status = kGjkIntersect;
        M.st = kLsDone;
      } else {
        smpAdd(M.s, M.v);
        const int p = gjkProject(M.s, M.dir, tol);
        if (p == kProjFailed) {
          M.o.status = kGjkFailed;
          M.st = kLsDone;
        } else if (p == kProjIntersect) {
          M.o.status = kGjkIntersect;
          M.st = kLsDone;
        } else if (p == kProjZeroVolume) {
          M.o.status = kGjkNoProgress;
          M.st = kLsNoSep;
        } else if (++M.it < max_iter) {
          M.o.iters = M.it + 1;
          M.rqid = freeSlot(M.s);
          M.req = true;
        } else {
          M.st = kLsNoSep;  // status stays IterLimit
        }
      }
    } else if (M.st == kLsD1) {
      if (dot(M.dir, M.v.v - M.cur) < tol || smpDup(M.s, M.v.v, tol2)) {
        M.st = kLsExtract;
      } else {
        smpAdd(M.s, M.v);
        const int u = distUpdate(M.s, M.cur, tol);
        if (u == kNoImprovement) {
          M.st = kLsExtract;
        } else if (u != kUpdOK) {
          M.st = kLsDone;  // sep_valid stays false
        } else {
          const double nm = norm(M.cur);
          if (M.book - nm < tol || nm < tol) {
            M.st = kLsExtract;
          } else {
            M.book = nm;
            M.dir = -M.cur / M.book;
            if (++M.it < max_iter) {
              M.o.dist_iters++;
              M.rqid = freeSlot(M.s);
              M.req = true;
            } else {
              M.st = kLsDone;
            }
          }
        }
      }
    }
    if (M.st == kLsSep) {
      M.o.status = kGjkSeparated;
      if (!want_distance) {
        M.st = kLsDone;
      } else {
        // gjkMinDistance from the one-vertex simplex {v}: its closest point is v
        M.s.rank = -1;
        smpAdd(M.s, M.v);
        M.cur = M.v.v;
        M.book = norm(M.cur);
        if (M.book <= tol) {
          M.st = kLsExtract;
        } else {
          M.dir = -M.cur / M.book;
          M.it = 0;
          if (M.it < max_iter) {
            M.o.dist_iters++;
            M.rqid = freeSlot(M.s);
            M.req = true;
            M.st = kLsD1;
          } else {
            M.st = kLsDone;
          }
        }
      }
    }
    if (M.st == kLsExtract) M.ext = 1;
    if (M.st == kLsNoSep && want_distance) M.ext = 2;
    if (M.st == kLsExtract || M.st == kLsNoSep) M.st = kLsDone;
  }
}

// Witnesses of a finished lane, where the machine's state is no longer live.
MFCL_D void laneFinish(LaneM& M, const MinkLane& md) {
  if (M.ext == 1) M.o.sep_valid = extractNoSub(md, M.s, M.o.p0, M.o.p1);
  if (M.ext == 2) M.o.sep_valid = extractTrySub(md, M.s, M.o.p0, M.o.p1);
}

}  // namespace mfcl
