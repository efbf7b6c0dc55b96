// Device EPA (include/fcl/cvx_collide/epa.hpp, epa_polytope*.hpp,
// epa_simplex2polytope.hpp, epa_polytope_utils.h).
//
// GPU representation: one wavefront per pair; the polytope lives in that
// wavefront's LDS slice as slot arrays of vertices, edges and faces (16-bit
// indices, FP64 geometry) that the 64 lanes scan in parallel; creation
// stamps stand in for the reference's newest-first element lists.
#pragma once

#include <type_traits>

#include "narrow_dev.h"

namespace mfcl {

constexpr int16_t kNil = -1;

// A condition that every lane of the wavefront evaluates on identical
// values (the polytope code runs one pair per wavefront in lock step), taken
// as a scalar so that the branch on it is a scalar branch.
MFCL_D bool ub(bool c) { return __builtin_amdgcn_readfirstlane(c ? 1 : 0) != 0; }

struct EVert {
  double vx, vy, vz;  // vertex on the Minkowski difference
  double dx, dy, dz;  // direction that produced it
  int32_t stamp;      // creation order; < 0: free slot
  int16_t new_edge;
  uint8_t remove, pad;
  int32_t h0, h1;     // the support vertices of the two hulls (-1: primitive or unknown)
};
struct EEdge {
  double d2, wx, wy, wz;  // squared distance to origin, witness
  int32_t stamp;
  int16_t v0, v1, f0, f1;
  uint8_t in, vis;
  uint16_t pad;
};
struct EFace {
  double d2, wx, wy, wz;
  int32_t stamp;
  int16_t e0, e1, e2;
  uint8_t in, vis;
};

// Arena sizing for a given face capacity: vertices = edges = floor(1.51 F)
// (epa_polytope.hpp:90-95).
struct ArenaDims {
  int nv, nf;
};
__host__ __device__ inline ArenaDims arenaDims(int max_faces) {
  ArenaDims d;
  d.nv = static_cast<int>(1.51 * static_cast<double>(max_faces));
  d.nf = max_faces;
  return d;
}
// Scratch record per border edge of one expansion (see Poly::expand).
struct BorderRec {
  int32_t pre;     // elements (edges and faces) created before this edge's face
  int16_t e, rank; // edge slot, rank in descending stamp order
  int16_t fe, ff;  // the list of free edge / face slots (entry k: the k-th free slot)
  int8_t t0, t1;   // lower-ranked border edges sharing endpoint v0 / v1
};
__host__ __device__ inline size_t arenaBytes(int max_faces) {
  const ArenaDims d = arenaDims(max_faces);
  size_t b = sizeof(EVert) * d.nv + sizeof(EEdge) * d.nv + sizeof(EFace) * d.nf + sizeof(BorderRec) * d.nv;
  return (b + 255) & ~static_cast<size_t>(255);
}

enum EVis : uint8_t { kUnknown = 0, kBorder = 1, kInternal = 2 };
enum FVis : uint8_t { kFUnknown = 0, kVisible = 1, kHidden = 2 };
enum EpaStatus : int { kEpaFailed = 0, kEpaOK = 1, kEpaTouching = 2, kEpaIterLimit = 3, kEpaMallocFailed = 4 };

struct MinD {
  bool in;
  double d2;
  d3 w;
};

// epa_polytope_utils.h:19-52
MFCL_D void segDist(d3 p, d3 x0, d3 x1, MinD& o) {
  const d3 d = x1 - x0;
  const d3 a = x0 - p;
  const double l2 = sqnorm(d);
  const double t = -1.0 * dot(a, d) / l2;
  if (t <= 0.0) {
    o.w = x0;
    o.in = false;
    o.d2 = sqnorm(x0 - p);
  } else if (t >= 1.0) {
    o.w = x1;
    o.in = false;
    o.d2 = sqnorm(x1 - p);
  } else {
    o.w = x0 + t * d;
    o.in = true;
    o.d2 = sqnorm(o.w);
  }
}
MFCL_D double lineDist(d3 p, d3 a, d3 b) {
  const d3 ab = b - a;
  const double l = norm(ab);
  if (l <= 0.0) return norm(p - a);
  return norm(cross(p - a, ab)) / l;
}
MFCL_D double planeDist(d3 p, d3 a, d3 b, d3 c) {
  const d3 n = cross(a - b, b - c);
  const double l = norm(n);
  if (l <= 0.0) return lineDist(p, a, b);
  return fabs(dot(n / l, a - p));
}
// epa_polytope_utils.h:92-166 with p = origin.
MFCL_D void triDist0(d3 a, d3 b, d3 c, MinD& o) {
  const d3 dl0 = a - b, dl1 = b - c, dl2 = c - a;
  const d3 n = cross(dl0, dl1);
  const double n2 = sqnorm(n);
  const double area = sqrt(n2);
  bool sub = fabs(area) <= 1e-16;
  if (!sub) {
    const d3 proj = n * (dot(a, n) / n2);  // p + n * ((a - p).n / n2) with p = 0
    const double w0 = norm(cross(dl1, b - proj)) / area;
    const double w1 = norm(cross(dl2, c - proj)) / area;
    const double w2 = 1.0 - w0 - w1;
    bool in = !(w0 < 0.0 || w0 > 1.0) && !(w1 < 0.0 || w1 > 1.0) && !(w2 < 0.0 || w2 > 1.0);
    const double chk = norm(cross(dl0, a - proj)) / area;
    if (fabs(chk - w2) > 1e-3) in = false;
    if (in) {
      o.in = true;
      o.d2 = sqnorm(proj);
      o.w = a * w0 + b * w1 + c * w2;
      return;
    }
    sub = true;
  }
  const d3 z = {0, 0, 0};
  MinD s0, s1, s2;
  segDist(z, a, b, s0);
  segDist(z, b, c, s1);
  segDist(z, c, a, s2);
  MinD best = s0;
  if (s1.d2 < best.d2) best = s1;
  if (s2.d2 < best.d2) best = s2;
  o.in = false;
  o.d2 = best.d2;
  o.w = best.w;
}

// Where a polytope's slot arrays live: the wavefront's LDS slice (the
// normal case) or a device-memory arena (a polytope that outgrew the LDS
// slice and continues at full capacity).  sync() orders one lane's writes
// before other lanes' reads of the same wavefront: LDS executes a
// wavefront's operations in order, so a compiler barrier suffices there;
// device memory needs a workgroup-scope fence.
struct LdsMem {
  template <class T>
  using P = MFCL_LDS T*;
  static MFCL_D void sync() { __builtin_amdgcn_wave_barrier(); }
};
struct GlbMem {
  template <class T>
  using P = T*;
  static MFCL_D void sync() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup"); }
};

// The polytope, processed by the 64 lanes of one wavefront together.
// Elements live in fixed slot arrays (slot s is visited by lane s % 64);
// a slot is free when its stamp is negative.  Every element gets a creation
// stamp from one counter, and the reference's element lists are
// newest-first (epa_polytope.hpp:127-290), so "earlier in the list" is
// "larger stamp": scans that the reference runs sequentially over its lists
// run here as parallel passes plus wavefront reductions ordered by stamp,
// which reproduces its tie-breaking exactly.  Slot numbers themselves never
// influence a result.  Every member function must be called by all 64 lanes
// with identical arguments.
template <class M>
struct PolyT {
  typename M::template P<EVert> V;
  typename M::template P<EEdge> E;
  typename M::template P<EFace> F;
  typename M::template P<BorderRec> BL;  // nv records
  int nv, nf;   // slot capacities: vertices and edges nv, faces nf
  int stamp;    // creation counter

  MFCL_D d3 vloc(int16_t i) const { return {V[i].vx, V[i].vy, V[i].vz}; }
  MFCL_D d3 vdir(int16_t i) const { return {V[i].dx, V[i].dy, V[i].dz}; }

  // epa_polytope.hpp:77-125: everything free
  MFCL_D void reset() {
    const int l = __lane_id();
    for (int s = l; s < nv; s += 64) {
      V[s].stamp = -1;
      E[s].stamp = -1;
    }
    for (int s = l; s < nf; s += 64) F[s].stamp = -1;
    stamp = 0;
    M::sync();
  }
  // lowest free slot (the reference pops its free list; which slot is used
  // does not matter, only the capacity does)
  template <class A>
  MFCL_D int16_t allocSlot(A arr, int cap) const {
    const int l = __lane_id();
    for (int base = 0; base < cap; base += 64) {
      const int s = base + l;
      const unsigned long long m = __ballot(s < cap && arr[s].stamp < 0);
      if (m) return static_cast<int16_t>(base + __ffsll(m) - 1);
    }
    return kNil;
  }
  MFCL_D int16_t addVertex(d3 v, d3 d, int h0 = -1, int h1 = -1) {
    const int16_t i = allocSlot(V, nv);
    if (i == kNil) return kNil;
    auto& x = V[i];
    x.vx = v.x; x.vy = v.y; x.vz = v.z;
    x.dx = d.x; x.dy = d.y; x.dz = d.z;
    x.h0 = h0; x.h1 = h1;
    x.stamp = stamp++;
    M::sync();
    return i;
  }
  MFCL_D int16_t addEdge(int16_t a, int16_t b) {
    if (a == kNil || b == kNil) return kNil;
    const int16_t i = allocSlot(E, nv);
    if (i == kNil) return kNil;
    auto& e = E[i];
    e.v0 = a;
    e.v1 = b;
    e.f0 = e.f1 = kNil;
    e.vis = kUnknown;
    MinD m;
    segDist(d3{0, 0, 0}, vloc(a), vloc(b), m);
    e.in = m.in; e.d2 = m.d2; e.wx = m.w.x; e.wy = m.w.y; e.wz = m.w.z;
    e.stamp = stamp++;
    M::sync();
    return i;
  }
  MFCL_D int16_t addFace(int16_t e1, int16_t e2, int16_t e3) {
    if (e1 == kNil || e2 == kNil || e3 == kNil) return kNil;
    const int16_t i = allocSlot(F, nf);
    if (i == kNil) return kNil;
    auto& f = F[i];
    f.e0 = e1; f.e1 = e2; f.e2 = e3;
    f.vis = kFUnknown;
    const int16_t a = E[e1].v0, b = E[e1].v1;
    const int16_t c = (E[e2].v0 != a && E[e2].v0 != b) ? E[e2].v0 : E[e2].v1;
    MinD m;
    triDist0(vloc(a), vloc(b), vloc(c), m);
    f.in = m.in; f.d2 = m.d2; f.wx = m.w.x; f.wy = m.w.y; f.wz = m.w.z;
    f.stamp = stamp++;
    M::sync();
    // the face stays in the polytope even when an edge already has two faces
    const int16_t es[3] = {e1, e2, e3};
    for (int k = 0; k < 3; ++k) {
      auto& ek = E[es[k]];
      if (ub(ek.f0 == kNil)) {
        ek.f0 = i;
      } else {
        if (ub(ek.f1 != kNil)) return kNil;
        ek.f1 = i;
      }
      M::sync();
    }
    return i;
  }
  MFCL_D void faceVerts(int16_t fi, int16_t& a, int16_t& b, int16_t& c) const {
    const auto& f = F[fi];
    a = E[f.e0].v0;
    b = E[f.e0].v1;
    c = (E[f.e1].v0 != a && E[f.e1].v0 != b) ? E[f.e1].v0 : E[f.e1].v1;
  }

  // Wavefront argmin of (d2, kind, newest first) -- the first element of
  // the reference's vertex, edge, face scans with strict '<'.
  MFCL_D static bool better(double d2, int kind, int st, double bd, int bk, int bs) {
    if (d2 < bd) return true;
    if (!(d2 == bd)) return false;
    if (kind != bk) return kind < bk;
    return st > bs;
  }
  // epa_polytope.hpp:292-345; kind: 0 none, 1 vertex, 2 edge, 3 face
  MFCL_D void minToOrigin(bool excl_vertex, int& kind, int16_t& idx) const {
    const int l = __lane_id();
    double bd = INFINITY;
    int bk = 4, bs = -1, bi = kNil;
    if (!excl_vertex)
      for (int s = l; s < nv; s += 64) {
        const int st = V[s].stamp;
        if (st < 0) continue;
        const double d2 = sqnorm(vloc(static_cast<int16_t>(s)));
        if (better(d2, 1, st, bd, bk, bs)) { bd = d2; bk = 1; bs = st; bi = s; }
      }
    for (int s = l; s < nv; s += 64) {
      const auto& e = E[s];
      const int st = e.stamp;
      if (st < 0) continue;
      if (!excl_vertex && !e.in) continue;
      if (better(e.d2, 2, st, bd, bk, bs)) { bd = e.d2; bk = 2; bs = st; bi = s; }
    }
    for (int s = l; s < nf; s += 64) {
      const auto& f = F[s];
      const int st = f.stamp;
      if (st < 0 || !f.in) continue;
      if (better(f.d2, 3, st, bd, bk, bs)) { bd = f.d2; bk = 3; bs = st; bi = s; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double od = __shfl_xor(bd, off);
      const int ok = __shfl_xor(bk, off), os = __shfl_xor(bs, off), oi = __shfl_xor(bi, off);
      if (better(od, ok, os, bd, bk, bs)) { bd = od; bk = ok; bs = os; bi = oi; }
    }
    // identical in every lane after the butterfly: taken as scalars, so
    // the callers branch on them with scalar branches
    const int ubk = __builtin_amdgcn_readfirstlane(bk);
    const int ubi = __builtin_amdgcn_readfirstlane(bi);
    const int fin = __builtin_amdgcn_readfirstlane(bd < INFINITY ? 1 : 0);
    // only finite candidates beat the reference's initial INFINITY
    if (ubk == 4 || !fin) {
      kind = 0;
      idx = kNil;
    } else {
      kind = ubk;
      idx = static_cast<int16_t>(ubi);
    }
  }
  // epa_polytope.hpp:347-405
  MFCL_D bool faceNormalOut(int16_t fi, d3& normal, double* area) const {
    int16_t ia, ib, ic;
    faceVerts(fi, ia, ib, ic);
    const d3 a = vloc(ia), b = vloc(ib), c = vloc(ic);
    const d3 n = cross(a - b, b - c);
    const double nn = norm(n);
    if (area) *area = 0.5 * nn;
    if (ub(nn <= 0.0)) return false;
    const d3 dir = n / nn;
    const double od = dot(a, dir);
    if (ub(fabs(od) >= 1e-4)) {
      normal = od > 0 ? dir : -dir;
      return true;
    }
    // max / min over the vertices: order-free, so a parallel pass
    double maxp = 0.0, minn = 0.0;
    for (int s = __lane_id(); s < nv; s += 64) {
      if (V[s].stamp < 0) continue;
      const double dv = dot(vloc(static_cast<int16_t>(s)), dir);
      if (dv > maxp) maxp = dv;
      if (dv < minn) minn = dv;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double om = __shfl_xor(maxp, off), on = __shfl_xor(minn, off);
      if (om > maxp) maxp = om;
      if (on < minn) minn = on;
    }
    normal = fabs(maxp) > fabs(minn) ? -dir : dir;
    return true;
  }
  MFCL_D d3 faceAnchor(int16_t fi) const { return vloc(E[F[fi].e0].v0); }
  // epa_polytope_expand.hpp:12-30
  MFCL_D bool outside(int16_t fi, d3 p) const {
    d3 n;
    double area = 0;
    if (!faceNormalOut(fi, n, &area)) return ub(area <= 0.0);
    return ub(dot(n, p - faceAnchor(fi)) >= 0.0);
  }
  MFCL_D bool anyLane(bool b) const { return __ballot(b) != 0ull; }

  // epa_polytope_expand.hpp:32-281. 0 ok, 1 failed, 2 malloc failed,
  // 3 (only when may_spill) this polytope's capacity might not hold the
  // expansion -- nothing has been changed then, and the caller continues
  // in a polytope of the full capacity.
  // The reference's depth-first flood fill marks the faces reachable from
  // `start` through faces that see nv; that set, and the border / internal
  // edge marks, do not depend on the visiting order, so it runs here as
  // rounds of a parallel breadth-first fill.
  MFCL_D int expand(d3 nv_, d3 nd, int nh0, int nh1, int16_t start, bool may_spill) {
    const int l = __lane_id();
    for (int s = l; s < nv; s += 64) {
      V[s].remove = 1;
      V[s].new_edge = kNil;
      E[s].vis = kUnknown;
    }
    for (int s = l; s < nf; s += 64) F[s].vis = kFUnknown;
    M::sync();
    F[start].vis = kVisible;
    M::sync();
    while (true) {
      bool changed = false;
      for (int s = l; s < nf; s += 64) {
        const auto& f = F[s];
        if (f.stamp < 0 || f.vis != kFUnknown) continue;
        const int16_t es[3] = {f.e0, f.e1, f.e2};
        bool near_visible = false;
        for (int k = 0; k < 3; ++k) {
          if (es[k] == kNil) continue;
          const auto& e = E[es[k]];
          const int16_t g = e.f0 == s ? e.f1 : e.f0;
          if (g != kNil && F[g].vis == kVisible) near_visible = true;
        }
        if (!near_visible) continue;
        const bool out = outsideLane(static_cast<int16_t>(s), nv_);
        F[s].vis = out ? kVisible : kHidden;
        changed = true;
      }
      M::sync();
      if (!anyLane(changed)) break;
    }
    // an open edge on a visible face is a malformed polytope
    bool bad = false;
    for (int s = l; s < nf; s += 64) {
      const auto& f = F[s];
      if (f.stamp < 0 || f.vis != kVisible) continue;
      const int16_t es[3] = {f.e0, f.e1, f.e2};
      for (int k = 0; k < 3; ++k) {
        if (es[k] == kNil) { bad = true; continue; }
        const auto& e = E[es[k]];
        const int16_t g = e.f0 == s ? e.f1 : e.f0;
        if (g == kNil) bad = true;
      }
    }
    if (anyLane(bad)) return 1;
    if (may_spill && !fits()) return 3;
    // edge marks from the final face classes
    for (int s = l; s < nv; s += 64) {
      auto& e = E[s];
      if (e.stamp < 0) continue;
      const bool v0 = e.f0 != kNil && F[e.f0].vis == kVisible;
      const bool v1 = e.f1 != kNil && F[e.f1].vis == kVisible;
      if (v0 && v1) e.vis = kInternal;
      else if (v0 || v1) e.vis = kBorder;
    }
    M::sync();
    for (int s = l; s < nv; s += 64) {
      const auto& e = E[s];
      if (e.stamp < 0 || e.vis == kInternal) continue;
      V[e.v0].remove = 0;
      V[e.v1].remove = 0;
    }
    M::sync();
    // drop the visible faces (detaching them from their border edges), the
    // internal edges and the vertices left without a non-internal edge
    for (int s = l; s < nf; s += 64) {
      auto& f = F[s];
      if (f.stamp < 0 || f.vis != kVisible) continue;
      const int16_t es[3] = {f.e0, f.e1, f.e2};
      for (int k = 0; k < 3; ++k) {
        if (es[k] == kNil) continue;
        auto& ek = E[es[k]];
        if (ek.vis == kInternal) continue;  // removed below
        if (ek.f0 == s) ek.f0 = ek.f1;
        ek.f1 = kNil;
      }
      f.stamp = -1;
    }
    for (int s = l; s < nv; s += 64) {
      auto& e = E[s];
      if (e.stamp >= 0 && e.vis == kInternal) e.stamp = -1;
    }
    for (int s = l; s < nv; s += 64) {
      auto& v = V[s];
      if (v.stamp >= 0 && v.remove) v.stamp = -1;
    }
    M::sync();
    const int16_t nvi = addVertex(nv_, nd, nh0, nh1);
    if (nvi == kNil) return 2;
    return coverBorder(nvi);
  }

  // New faces over the border edges (epa_polytope_expand.hpp:233-281).  The
  // reference walks the border edges newest first; for each it creates the
  // edge (new vertex, endpoint) unless an earlier border edge already did,
  // then the face, each taking the next creation stamp; a new edge's first
  // face becomes its f0, the second its f1, a third is an error.  All of
  // that follows from each border edge's rank in that order, so the lanes
  // build the faces together: rank and sharing counts by comparison with the
  // other border edges, creation positions by a prefix over the ranks, slots
  // from the free lists (which free slot is used does not matter).  A slot
  // shortage fails the expansion as the reference's allocation failure does.
  MFCL_D int coverBorder(int16_t nvi) {
    const int l = __lane_id();
    const unsigned long long below = (1ull << l) - 1;
    int nb = 0;
    for (int base = 0; base < nv; base += 64) {
      const int s = base + l;
      const bool isb = s < nv && E[s].stamp >= 0 && E[s].vis == kBorder;
      const unsigned long long m = __ballot(isb);
      if (isb) BL[nb + __popcll(m & below)].e = static_cast<int16_t>(s);
      nb += __popcll(m);
    }
    M::sync();
    for (int j = l; j < nb; j += 64) {
      const int st = E[BL[j].e].stamp;
      int rank = 0;
      for (int k = 0; k < nb; ++k) rank += E[BL[k].e].stamp > st;
      BL[j].rank = static_cast<int16_t>(rank);
    }
    M::sync();
    bool bad = false;
    int new_edges = 0;
    for (int j = l; j < nb; j += 64) {
      const auto& ej = E[BL[j].e];
      const int r = BL[j].rank, a = ej.v0, b = ej.v1;
      int t0 = 0, t1 = 0;
      for (int k = 0; k < nb; ++k) {
        if (BL[k].rank >= r) continue;
        const auto& ek = E[BL[k].e];
        t0 += ek.v0 == a || ek.v1 == a;
        t1 += ek.v0 == b || ek.v1 == b;
      }
      if (t0 > 1 || t1 > 1) bad = true;  // a third face on a new edge
      BL[j].t0 = static_cast<int8_t>(t0 > 1 ? 2 : t0);
      BL[j].t1 = static_cast<int8_t>(t1 > 1 ? 2 : t1);
      new_edges += (t0 == 0) + (t1 == 0);
    }
    if (anyLane(bad)) return 2;
    M::sync();
    for (int j = l; j < nb; j += 64) {
      const int r = BL[j].rank;
      int pre = 0;
      for (int k = 0; k < nb; ++k)
        if (BL[k].rank < r) pre += 1 + (BL[k].t0 == 0) + (BL[k].t1 == 0);
      BL[j].pre = pre;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) new_edges += __shfl_xor(new_edges, off);
    // capacity
    int free_e = 0, free_f = 0;
    for (int base = 0; base < nv; base += 64) free_e += __popcll(__ballot(base + l < nv && E[base + l].stamp < 0));
    for (int base = 0; base < nf; base += 64) free_f += __popcll(__ballot(base + l < nf && F[base + l].stamp < 0));
    new_edges = __builtin_amdgcn_readfirstlane(new_edges);
    if (new_edges > free_e || nb > free_f) return 2;
    // the free slot lists (before any slot is taken)
    {
      int ce = 0, cf = 0;
      for (int base = 0; base < nv; base += 64) {
        const int s = base + l;
        const bool fr = s < nv && E[s].stamp < 0;
        const unsigned long long m = __ballot(fr);
        const int pos = ce + __popcll(m & below);
        if (fr && pos < new_edges) BL[pos].fe = static_cast<int16_t>(s);
        ce += __popcll(m);
      }
      for (int base = 0; base < nf; base += 64) {
        const int s = base + l;
        const bool fr = s < nf && F[s].stamp < 0;
        const unsigned long long m = __ballot(fr);
        const int pos = cf + __popcll(m & below);
        if (fr && pos < nb) BL[pos].ff = static_cast<int16_t>(s);
        cf += __popcll(m);
      }
    }
    M::sync();
    // new edges: the k-th created edge takes the k-th free edge slot
    for (int j = l; j < nb; j += 64) {
      const auto& bj = BL[j];
      const auto& ej = E[bj.e];
      const int edges_before = bj.pre - bj.rank;
      int q = 0;
      for (int side = 0; side < 2; ++side) {
        if ((side ? bj.t1 : bj.t0) != 0) continue;
        const int16_t vq = side ? ej.v1 : ej.v0;
        const int16_t slot = BL[edges_before + q].fe;
        auto& x = E[slot];
        x.v0 = nvi;
        x.v1 = vq;
        x.f0 = x.f1 = kNil;
        x.vis = kUnknown;
        MinD m;
        segDist(d3{0, 0, 0}, vloc(nvi), vloc(vq), m);
        x.in = m.in; x.d2 = m.d2; x.wx = m.w.x; x.wy = m.w.y; x.wz = m.w.z;
        x.stamp = stamp + bj.pre + q;
        V[vq].new_edge = slot;
        ++q;
      }
    }
    M::sync();
    // faces: the face of rank r takes the r-th free face slot
    for (int j = l; j < nb; j += 64) {
      const auto& bj = BL[j];
      const int16_t ei = bj.e;
      auto& ej = E[ei];
      const int16_t a = ej.v0, b = ej.v1;
      const int16_t ne0 = V[a].new_edge, ne1 = V[b].new_edge;
      const int16_t slot = BL[bj.rank].ff;
      auto& f = F[slot];
      f.e0 = ei; f.e1 = ne0; f.e2 = ne1;
      f.vis = kFUnknown;
      MinD m;
      triDist0(vloc(a), vloc(b), vloc(nvi), m);
      f.in = m.in; f.d2 = m.d2; f.wx = m.w.x; f.wy = m.w.y; f.wz = m.w.z;
      f.stamp = stamp + bj.pre + (bj.t0 == 0) + (bj.t1 == 0);
    }
    M::sync();
    // links: the border edge keeps its hidden face as f0; each new edge gets
    // its lower-ranked face as f0 and the other as f1
    bool clash = false;
    for (int j = l; j < nb; j += 64) {
      const auto& bj = BL[j];
      auto& ej = E[bj.e];
      const int16_t slot = BL[bj.rank].ff;
      if (ej.f0 == kNil) ej.f0 = slot;
      else if (ej.f1 != kNil) clash = true;
      else ej.f1 = slot;
      auto& x0 = E[V[ej.v0].new_edge];
      auto& x1 = E[V[ej.v1].new_edge];
      if (bj.t0 == 0) x0.f0 = slot; else x0.f1 = slot;
      if (bj.t1 == 0) x1.f0 = slot; else x1.f1 = slot;
    }
    stamp += new_edges + nb;
    M::sync();
    return anyLane(clash) ? 2 : 0;
  }

  // Whether the expansion whose visible faces are marked certainly fits:
  // after the visible faces, the internal edges and the vertices left
  // without a non-internal edge are freed, one vertex, at most two new
  // edges per border edge and one face per border edge must find slots.
  // Marks only; nothing is changed.
  MFCL_D bool fits() const {
    const int l = __lane_id();
    int free_v = 0, free_e = 0, free_f = 0, nb = 0;
    for (int base = 0; base < nf; base += 64) {
      const int s = base + l;
      bool fr = false;
      if (s < nf) fr = F[s].stamp < 0 || F[s].vis == kVisible;
      free_f += __popcll(__ballot(fr));
    }
    for (int base = 0; base < nv; base += 64) {
      const int s = base + l;
      bool fe = false, bd = false, fv = false;
      if (s < nv) {
        const auto& e = E[s];
        if (e.stamp < 0) {
          fe = true;
        } else {
          const bool v0 = e.f0 != kNil && F[e.f0].vis == kVisible;
          const bool v1 = e.f1 != kNil && F[e.f1].vis == kVisible;
          fe = v0 && v1;
          bd = v0 != v1;
        }
        // a vertex stays when some edge of the polytope that is not internal
        // touches it; counted below as free only when it has no such edge
        fv = V[s].stamp < 0;
      }
      free_e += __popcll(__ballot(fe));
      nb += __popcll(__ballot(bd));
      free_v += __popcll(__ballot(fv));
    }
    // the vertices removed by the expansion are not counted as free: the
    // test only needs one vertex slot, so a free one must exist already
    return free_v >= 1 && free_e >= 2 * nb && free_f >= nb;
  }

 private:
  // outside() evaluated by one lane for its own face (no cross-lane work:
  // the vertex fallback of faceNormalOut is scanned by that lane alone)
  MFCL_D bool outsideLane(int16_t fi, d3 p) const {
    int16_t ia, ib, ic;
    faceVerts(fi, ia, ib, ic);
    const d3 a = vloc(ia), b = vloc(ib), c = vloc(ic);
    const d3 n = cross(a - b, b - c);
    const double nn = norm(n);
    if (nn <= 0.0) return 0.5 * nn <= 0.0;
    const d3 dir = n / nn;
    const double od = dot(a, dir);
    d3 normal;
    if (fabs(od) >= 1e-4) {
      normal = od > 0 ? dir : -dir;
    } else {
      double maxp = 0.0, minn = 0.0;
      for (int s = 0; s < nv; ++s) {
        if (V[s].stamp < 0) continue;
        const double dv = dot(vloc(static_cast<int16_t>(s)), dir);
        if (dv > maxp) maxp = dv;
        if (dv < minn) minn = dv;
      }
      normal = fabs(maxp) > fabs(minn) ? -dir : dir;
    }
    return dot(normal, p - faceAnchor(fi)) >= 0.0;
  }
};

using Poly = PolyT<LdsMem>;
using PolyG = PolyT<GlbMem>;

// Result of the EPA stage in shape-0 frame.
struct EpaOut {
  int status;
  double depth;
  d3 p0, p1;
  int iters;
  int scanned_extra;  // hull vertices scanned by the out-of-line helpers
  int spilled;        // 1: continued in the device-memory polytope
};

// The out-of-line helpers take the Minkowski difference by value (a
// reference would keep the caller's copy in private memory for the whole
// EPA loop); this adds the vertices they scan to the pair's tally on exit.
template <class MK>
struct ScanTally {
  MK& md;
  EpaOut& o;
  int start;
  MFCL_D ScanTally(MK& m, EpaOut& out) : md(m), o(out), start(m.scanned) {}
  MFCL_D ~ScanTally() { o.scanned_extra += md.scanned - start; }
};

template <class MK>
MFCL_D void touchingPoint(MK& md, d3 d, EpaOut& o) {
  const d3 mid = (md.support0(d) + md.support1(-d)) / 2;
  o.p0 = mid;
  o.p1 = mid;
}

// epa_simplex2polytope.hpp:184-217; returns 0 ok, 2 failed
__device__ __attribute__((noinline)) int tetra(Poly& P, const MVD& a, const MVD& b, const MVD& c, const MVD& d) {
  P.reset();
  int16_t v[4], e[6], f[4];
  v[0] = P.addVertex(a.v, a.d);
  v[1] = P.addVertex(b.v, b.d);
  v[2] = P.addVertex(c.v, c.d);
  v[3] = P.addVertex(d.v, d.d);
  e[0] = P.addEdge(v[0], v[1]);
  e[1] = P.addEdge(v[1], v[2]);
  e[2] = P.addEdge(v[2], v[0]);
  e[3] = P.addEdge(v[3], v[0]);
  e[4] = P.addEdge(v[3], v[1]);
  e[5] = P.addEdge(v[3], v[2]);
  f[0] = P.addFace(e[0], e[1], e[2]);
  f[1] = P.addFace(e[3], e[4], e[0]);
  f[2] = P.addFace(e[4], e[5], e[1]);
  f[3] = P.addFace(e[5], e[3], e[2]);
  for (int i = 0; i < 4; ++i)
    if (f[i] == kNil) return 2;
  return 0;
}

// s2p status: 0 OK, 1 touching, 2 failed
template <class MK>
__device__ __attribute__((noinline)) int s2p3(Poly& P, MK md, const MVD& a, const MVD& b, const MVD& c, double thr, EpaOut& o) {
  ScanTally tally(md, o);
  d3 n = cross(b.v - a.v, c.v - a.v);
  if (sqnorm(n) <= 0.0) return 2;
  n = normalized(n);
  const d3 d0 = md.support(n);
  const double t0 = planeDist(d0, a.v, b.v, c.v);
  const d3 dir1 = -n;
  const d3 d1 = md.support(dir1);
  const double t1 = planeDist(d1, a.v, b.v, c.v);
  if (t0 <= thr) { touchingPoint(md, n, o); return 1; }
  if (t1 <= thr) { touchingPoint(md, dir1, o); return 1; }
  if (t0 > t1) return tetra(P, a, b, c, MVD{d0, n});
  return tetra(P, a, b, c, MVD{d1, dir1});
}

template <class MK>
__device__ __attribute__((noinline)) int s2p2(Poly& P, MK md, const MVD& a, const MVD& b, double thr, EpaOut& o) {
  ScanTally tally(md, o);
  const double thr2 = thr * thr;
  const d3 ab = b.v - a.v;
  if (sqnorm(ab) < thr2) return 2;
  const d3 abu = normalized(ab);
  d3 dinit = {abu.y, -abu.x, 0};
  bool valid = sqnorm(dinit) > thr2;
  if (!valid) { dinit = d3{abu.z, 0, -abu.x}; valid = sqnorm(dinit) > thr2; }
  if (!valid) { dinit = d3{0, abu.z, -abu.y}; valid = sqnorm(dinit) > thr2; }
  if (!valid) return 2;
  dinit = normalized(dinit);
  d3 dcur = dinit, dfound = {0, 0, 0}, vfound = {0, 0, 0};
  bool found = false;
  for (int i = 0; i < 72; ++i) {
    const d3 vcur = md.support(dcur);
    if (lineDist(vcur, a.v, b.v) > thr) {
      vfound = vcur;
      dfound = dcur;
      found = true;
      break;
    }
    const double step = 2.0 * 3.145926 / 72;  // the reference's pi constant (sic)
    dcur = normalized(dcur + step * cross(abu, vcur));
  }
  if (!found) { touchingPoint(md, dinit, o); return 1; }
  const MVD v0 = {vfound, dfound};
  MVD v1;
  v1.d = -v0.d;
  v1.v = md.support(v1.d);
  if (lineDist(v1.v, a.v, b.v) < thr) { touchingPoint(md, v1.d, o); return 1; }
  const d3 n = cross(v0.v - a.v, v1.v - a.v);
  if (sqnorm(n) < thr2) return 2;
  MVD v2;
  v2.d = normalized(n);
  v2.v = md.support(v2.d);
  if (lineDist(v2.v, a.v, b.v) < thr) { touchingPoint(md, v2.d, o); return 1; }
  MVD v3;
  v3.d = -v2.d;
  v3.v = md.support(v3.d);
  if (lineDist(v3.v, a.v, b.v) < thr) { touchingPoint(md, v3.d, o); return 1; }
  P.reset();
  int16_t v[6], e[12], f[8];
  v[0] = P.addVertex(a.v, a.d);
  v[1] = P.addVertex(v0.v, v0.d);
  v[2] = P.addVertex(b.v, b.d);
  v[3] = P.addVertex(v1.v, v1.d);
  v[4] = P.addVertex(v2.v, v2.d);
  v[5] = P.addVertex(v3.v, v3.d);
  e[0] = P.addEdge(v[0], v[1]);
  e[1] = P.addEdge(v[1], v[2]);
  e[2] = P.addEdge(v[2], v[3]);
  e[3] = P.addEdge(v[3], v[0]);
  e[4] = P.addEdge(v[4], v[0]);
  e[5] = P.addEdge(v[4], v[1]);
  e[6] = P.addEdge(v[4], v[2]);
  e[7] = P.addEdge(v[4], v[3]);
  e[8] = P.addEdge(v[5], v[0]);
  e[9] = P.addEdge(v[5], v[1]);
  e[10] = P.addEdge(v[5], v[2]);
  e[11] = P.addEdge(v[5], v[3]);
  f[0] = P.addFace(e[4], e[5], e[0]);
  f[1] = P.addFace(e[5], e[6], e[1]);
  f[2] = P.addFace(e[6], e[7], e[2]);
  f[3] = P.addFace(e[7], e[4], e[3]);
  f[4] = P.addFace(e[8], e[9], e[0]);
  f[5] = P.addFace(e[9], e[10], e[1]);
  f[6] = P.addFace(e[10], e[11], e[2]);
  f[7] = P.addFace(e[11], e[8], e[3]);
  for (int i = 0; i < 8; ++i)
    if (f[i] == kNil) return 2;
  return 0;
}

// epa_simplex2polytope.hpp:46-110
template <class MK>
MFCL_D int simplexToPolytope(Poly& P, MK& md, const SmpD& s, double thr, EpaOut& o) {
  const double thr2 = thr * thr;
  if (s.rank > 0 && sqnorm(s.a.v) <= thr2) { touchingPoint(md, s.a.d, o); return 1; }
  if (s.rank > 1 && sqnorm(s.b.v) <= thr2) { touchingPoint(md, s.b.d, o); return 1; }
  if (s.rank > 2 && sqnorm(s.c.v) <= thr2) { touchingPoint(md, s.c.d, o); return 1; }
  if (s.rank > 3 && sqnorm(s.e.v) <= thr2) { touchingPoint(md, s.e.d, o); return 1; }
  if (s.rank == 4) {
    const d3 z = {0, 0, 0};
    if (planeDist(z, s.a.v, s.b.v, s.c.v) < thr) return s2p3(P, md, s.a, s.b, s.c, thr, o);
    if (planeDist(z, s.a.v, s.c.v, s.e.v) < thr) return s2p3(P, md, s.a, s.c, s.e, thr, o);
    if (planeDist(z, s.a.v, s.b.v, s.e.v) < thr) return s2p3(P, md, s.a, s.b, s.e, thr, o);
    if (planeDist(z, s.b.v, s.c.v, s.e.v) < thr) return s2p3(P, md, s.b, s.c, s.e, thr, o);
    return tetra(P, s.a, s.b, s.c, s.e);
  }
  if (s.rank == 3) return s2p3(P, md, s.a, s.b, s.c, thr, o);
  if (s.rank == 2) return s2p2(P, md, s.a, s.b, thr, o);
  o.p0 = md.support0(s.a.d);
  o.p1 = md.support1(-s.a.d);
  return 1;
}

struct RawF {
  int kind;
  MinD m;
  MVD a, b, c;
  int h[3][2];  // support vertices of a, b, c (-1: primitive or unknown)
};

MFCL_D void ldsSet(MFCL_LDS d3& o, d3 v) {
  o.x = v.x;
  o.y = v.y;
  o.z = v.z;
}
MFCL_D d3 ldsGet(const MFCL_LDS d3& o) { return d3{o.x, o.y, o.z}; }

template <class PT>
MFCL_D void toRaw(const PT& P, int kind, int16_t i, MFCL_LDS RawF& r) {
  r.kind = kind;
  auto rec = [&](int k, int16_t v) {
    r.h[k][0] = P.V[v].h0;
    r.h[k][1] = P.V[v].h1;
  };
  if (kind == 1) {
    const d3 v = P.vloc(i);
    r.m.in = true; r.m.d2 = sqnorm(v); ldsSet(r.m.w, v);
    ldsSet(r.a.v, v); ldsSet(r.a.d, P.vdir(i));
    rec(0, i);
  } else if (kind == 2) {
    const auto& e = P.E[i];
    r.m.in = e.in; r.m.d2 = e.d2; ldsSet(r.m.w, d3{e.wx, e.wy, e.wz});
    ldsSet(r.a.v, P.vloc(e.v0)); ldsSet(r.a.d, P.vdir(e.v0));
    ldsSet(r.b.v, P.vloc(e.v1)); ldsSet(r.b.d, P.vdir(e.v1));
    rec(0, e.v0);
    rec(1, e.v1);
  } else {
    const auto& f = P.F[i];
    r.m.in = f.in; r.m.d2 = f.d2; ldsSet(r.m.w, d3{f.wx, f.wy, f.wz});
    int16_t ia, ib, ic;
    P.faceVerts(i, ia, ib, ic);
    ldsSet(r.a.v, P.vloc(ia)); ldsSet(r.a.d, P.vdir(ia));
    ldsSet(r.b.v, P.vloc(ib)); ldsSet(r.b.d, P.vdir(ib));
    ldsSet(r.c.v, P.vloc(ic)); ldsSet(r.c.d, P.vdir(ic));
    rec(0, ia);
    rec(1, ib);
    rec(2, ic);
  }
}

// epa.hpp:11-113; returns 0 ok, 1 failed, 2 converge
// nh0 / nh1 receive the support vertices of the last support point.
template <class PT, class MK>
MFCL_D int nextDir(const PT& P, MK& md, int kind, int16_t fi, double tol, d3& nd, d3& nv, int16_t& start, int& nh0,
                   int& nh1) {
  constexpr double kOuter = 1e-3 * 1e-3;
  auto face_cand = [&](int16_t f, bool try_w) -> int {
    d3 fn;
    if (!P.faceNormalOut(f, fn, nullptr)) return 1;
    const d3 on = P.faceAnchor(f);
    const auto& F = P.F[f];
    if (try_w && ub(F.d2 > kOuter)) {
      nd = normalized(d3{F.wx, F.wy, F.wz});
      nv = md.supportFrom(nd, -1, -1, nh0, nh1);
      if (ub(dot(fn, nv - on) >= 0.0)) { start = f; return 0; }
    }
    nd = fn;
    nv = md.supportFrom(nd, -1, -1, nh0, nh1);
    if (ub(dot(fn, nv - on) > tol)) { start = f; return 0; }
    return 2;
  };
  if (kind == 3) return face_cand(fi, true);
  if (kind == 2) {
    const auto& e = P.E[fi];
    const int16_t f0 = e.f0, f1 = e.f1;
    if (ub(f0 == kNil || f1 == kNil)) return 1;  // open edge: malformed polytope
    if (ub(e.d2 > kOuter)) {
      nd = normalized(d3{e.wx, e.wy, e.wz});
      nv = md.supportFrom(nd, -1, -1, nh0, nh1);
      if (P.outside(f0, nv)) { start = f0; return 0; }
      if (P.outside(f1, nv)) { start = f1; return 0; }
    }
    const int s0 = face_cand(f0, false);
    if (s0 == 0) return 0;
    return face_cand(f1, false);
  }
  return 1;
}

// epa.hpp:269-299 (edge case: squared norm of the segment projection, as in the reference)
MFCL_D bool epaTerminate(const MFCL_LDS RawF& f, d3 nv, double tol) {
  double d2;
  if (f.kind == 2) {
    MinD m;
    segDist(nv, ldsGet(f.a.v), ldsGet(f.b.v), m);
    d2 = m.d2;
  } else {
    const double t = planeDist(nv, ldsGet(f.a.v), ldsGet(f.b.v), ldsGet(f.c.v));
    d2 = t * t;
  }
  return d2 < tol * tol;
}

// epa.hpp:341-498
// support0 / support1 of a recorded direction come from the recorded
// support vertices when there are (the same points the support searches
// returned when the polytope vertex was created).
template <class MK>
__device__ __attribute__((noinline)) void assignPair(MK md, d3 cand_d, int ch0, int ch1, const MFCL_LDS RawF& f,
                                                     EpaOut& o) {
  ScanTally tally(md, o);
  auto p0 = [&](int k, const MFCL_LDS MVD& m) { return md.point0At(f.h[k][0], ldsGet(m.d)); };
  auto p1 = [&](int k, const MFCL_LDS MVD& m) { return md.point1At(f.h[k][1], ldsGet(m.d)); };
  if (f.kind == 1) {
    o.p0 = p0(0, f.a);
    o.p1 = p1(0, f.a);
    o.depth = norm(ldsGet(f.a.v));
    return;
  }
  if (f.kind == 2) {
    o.depth = sqrt(f.m.d2);
    const d3 p = ldsGet(f.m.w), a = ldsGet(f.a.v), b = ldsGet(f.b.v);
    const d3 ab = b - a;
    int ax = 0;
    double mx = fabs(ab.x);
    if (fabs(ab.y) > mx) { ax = 1; mx = fabs(ab.y); }
    if (fabs(ab.z) > mx) { ax = 2; mx = fabs(ab.z); }
    if (mx <= 1e-10) {
      o.p0 = p0(0, f.a);
      o.p1 = p1(0, f.a);
      return;
    }
    const double s = (comp(p, ax) - comp(a, ax)) / (comp(b, ax) - comp(a, ax));
    if (!(s < 0 || s > 1.0)) {
      const double wa = 1.0 - s, wb = s;
      o.p0 = wa * p0(0, f.a) + wb * p0(1, f.b);
      o.p1 = wa * p1(0, f.a) + wb * p1(1, f.b);
      return;
    }
  } else if (f.kind == 3) {
    o.depth = sqrt(f.m.d2);
    const d3 a = ldsGet(f.a.v), b = ldsGet(f.b.v), c = ldsGet(f.c.v), p = ldsGet(f.m.w);
    const d3 dl0 = a - b, dl1 = b - c, dl2 = c - a;
    const double area = sqrt(sqnorm(cross(dl0, dl1)));
    const double w0 = norm(cross(dl1, b - p)) / area;
    const double w1 = norm(cross(dl2, c - p)) / area;
    const double chk = norm(cross(dl0, a - p)) / area;
    const double w2 = 1.0 - w0 - w1;
    if (!(fabs(chk - w2) > 0.01)) {
      o.p0 = w0 * p0(0, f.a) + w1 * p0(1, f.b) + w2 * p0(2, f.c);
      o.p1 = w0 * p1(0, f.a) + w1 * p1(1, f.b) + w2 * p1(2, f.c);
      return;
    }
  }
  const d3 q0 = md.point0At(ch0, cand_d);
  const d3 q1 = md.point1At(ch1, cand_d);
  o.p0 = q0;
  o.p1 = q1;
  o.depth = norm(q0 - q1);
}

// Copies an LDS polytope into a device-memory one of larger capacity: the
// same slot numbers, the extra slots free.  Slot numbers never influence a
// result, so continuing there equals a run at the larger capacity from the
// start.
template <class R>
MFCL_D void copyRecord(R* dst, const MFCL_LDS R* src) {
  static_assert(sizeof(R) % 4 == 0, "records copy as words");
  const MFCL_LDS uint32_t* a = reinterpret_cast<const MFCL_LDS uint32_t*>(src);
  uint32_t* b = reinterpret_cast<uint32_t*>(dst);
#pragma unroll
  for (int k = 0; k < static_cast<int>(sizeof(R) / 4); ++k) b[k] = a[k];
}
MFCL_D void spillPoly(const Poly& L, PolyG& G) {
  const int l = __lane_id();
  for (int s = l; s < G.nv; s += 64) {
    if (s < L.nv) {
      copyRecord(&G.V[s], &L.V[s]);
      copyRecord(&G.E[s], &L.E[s]);
    } else {
      G.V[s].stamp = -1;
      G.E[s].stamp = -1;
    }
  }
  for (int s = l; s < G.nf; s += 64) {
    if (s < L.nf) copyRecord(&G.F[s], &L.F[s]);
    else G.F[s].stamp = -1;
  }
  G.stamp = L.stamp;
  GlbMem::sync();
}

// Where a pair continues when its LDS polytope might not hold an
// expansion: a device-memory polytope arena (null: nowhere) of capacity
// `faces` faces.
struct SpillTarget {
  char* base;
  int faces;
};

template <class MK>
__device__ __attribute__((noinline)) EpaOut epaContinue(Poly& P, MK md, double tol, int max_iter, EpaOut o,
                                                        MFCL_LDS RawF& raw, int it, d3 nv, d3 nd, int nh0, int nh1,
                                                        int16_t start, SpillTarget st);

// The EPA iterations (epa.hpp:137-237) from iteration `it`; with `pending`
// the iteration's expansion (nv, nd, start) is still to run.  A polytope in
// LDS whose expansion may not fit (expand() == 3, only with a spill target)
// continues in a device-memory polytope of full capacity (epaContinue, out
// of line: it runs for few pairs and keeps its registers out of this loop).
template <bool SPILL, class PT, class MK>
MFCL_D void epaLoop(PT& P, MK& md, double tol, int max_iter, EpaOut& o, MFCL_LDS RawF& raw, int it, bool pending,
                    d3 nv, d3 nd, int nh0, int nh1, int16_t start, SpillTarget st) {
  while (true) {
    if (!pending) {
      int kind;
      int16_t fi;
      P.minToOrigin(false, kind, fi);
      if (kind == 0) { o.status = kEpaFailed; return; }
      if (kind == 1) {
        const double vd2 = sqnorm(P.vloc(fi));
        constexpr double kRatio = 1.0 - 1e-3;
        int k2;
        int16_t i2;
        P.minToOrigin(true, k2, i2);
        if (k2 == 0) { o.status = kEpaFailed; return; }
        if (k2 == 2) {
          if (ub(kRatio * P.E[i2].d2 < vd2)) { kind = 2; fi = i2; }
        } else if (k2 == 3) {
          if (ub(kRatio * P.F[i2].d2 < vd2)) { kind = 3; fi = i2; }
        }
      }
      if (kind != 2 && kind != 3) { o.status = kEpaFailed; return; }
      toRaw(P, kind, fi, raw);
      nd = d3{0, 0, 0};
      nv = d3{0, 0, 0};
      start = kNil;
      nh0 = nh1 = -1;
      // statuses and the start face are the same in every lane: scalars
      const int ds = __builtin_amdgcn_readfirstlane(nextDir(P, md, kind, fi, tol, nd, nv, start, nh0, nh1));
      start = static_cast<int16_t>(__builtin_amdgcn_readfirstlane(start));
      if (ds == 1) { o.status = kEpaFailed; return; }
      if (ds == 2) {
        assignPair(md, nd, nh0, nh1, raw, o);
        o.iters = it;
        o.status = kEpaOK;
        return;
      }
      if (start == kNil) { o.status = kEpaFailed; return; }
      if (ub(epaTerminate(raw, nv, tol))) {
        assignPair(md, nd, nh0, nh1, raw, o);
        o.iters = it;
        o.status = kEpaOK;
        return;
      }
    }
    pending = false;
    const int ex = __builtin_amdgcn_readfirstlane(P.expand(nv, nd, nh0, nh1, start, SPILL && st.base != nullptr));
    if constexpr (SPILL && std::is_same<PT, Poly>::value) {
      if (ex == 3) {
        o = epaContinue(P, md, tol, max_iter, o, raw, it, nv, nd, nh0, nh1, start, st);
        return;
      }
    }
    if (ex == 1) { o.status = kEpaFailed; return; }
    if (ex == 2) {
      assignPair(md, nd, nh0, nh1, raw, o);
      o.iters = it;
      o.status = kEpaMallocFailed;
      return;
    }
    ++it;
    if (it >= max_iter) {
      assignPair(md, nd, nh0, nh1, raw, o);
      o.iters = it;
      o.status = kEpaIterLimit;
      return;
    }
  }
}

template <class MK>
__device__ __attribute__((noinline)) EpaOut epaContinue(Poly& P, MK md, double tol, int max_iter, EpaOut o,
                                                        MFCL_LDS RawF& raw, int it, d3 nv, d3 nd, int nh0, int nh1,
                                                        int16_t start, SpillTarget st) {
  const ArenaDims fd = arenaDims(st.faces);
  PolyG G;
  G.V = reinterpret_cast<EVert*>(st.base);
  G.E = reinterpret_cast<EEdge*>(st.base + sizeof(EVert) * fd.nv);
  G.F = reinterpret_cast<EFace*>(st.base + sizeof(EVert) * fd.nv + sizeof(EEdge) * fd.nv);
  G.BL = reinterpret_cast<BorderRec*>(st.base + sizeof(EVert) * fd.nv + sizeof(EEdge) * fd.nv +
                                      sizeof(EFace) * fd.nf);
  G.nv = fd.nv;
  G.nf = fd.nf;
  spillPoly(P, G);
  o.spilled = 1;
  const int scanned0 = md.scanned;
  epaLoop<false>(G, md, tol, max_iter, o, raw, it, true, nv, nd, nh0, nh1, start, SpillTarget{nullptr, 0});
  o.scanned_extra += md.scanned - scanned0;
  return o;
}

// epa.hpp:137-257.  st: where the pair continues if P (of smaller capacity
// than the request's) might not hold an expansion.
template <bool SPILL, class MK>
MFCL_D void epaEvaluate(Poly& P, MK& md, const SmpD& s, double tol, int max_iter, EpaOut& o,
                        MFCL_LDS RawF& raw, SpillTarget st) {
  o.iters = 0;
  o.scanned_extra = 0;
  o.spilled = 0;
  o.depth = 0;
  o.p0 = d3{0, 0, 0};
  o.p1 = d3{0, 0, 0};
  const int s2p = simplexToPolytope(P, md, s, tol, o);
  if (s2p == 2) { o.status = kEpaFailed; return; }
  if (s2p == 1) { o.depth = 0.0; o.status = kEpaTouching; return; }
  epaLoop<SPILL>(P, md, tol, max_iter, o, raw, 0, false, d3{0, 0, 0}, d3{0, 0, 0}, -1, -1, kNil, st);
}

}  // namespace mfcl
