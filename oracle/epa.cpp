// CPU oracle -- TEST INFRASTRUCTURE ONLY (see omath.h).
// EPA: include/fcl/cvx_collide/epa.hpp (loop :137-237, direction :11-113,
// termination :269-299, witnesses :341-498), polytope bookkeeping
// epa_polytope.hpp:77-411 and epa_polytope_expand.hpp:12-281, initial polytope
// epa_simplex2polytope.hpp:46-380, point/simplex distances
// epa_polytope_utils.h:19-174.
//
// The polytope keeps the reference's three intrusive element lists (newest
// element first) and free lists, addressed by index instead of pointer, so
// every scan visits elements in the same order as the reference.
#include <vector>

#include "narrowphase.h"

namespace oracle {
namespace {

constexpr int kNull = -1;

struct MinDist {
  bool in{false};
  double d2{0};
  V3 w;
};

// epa_polytope_utils.h:19-52
void pointToSegment(const V3& p, const V3& x0, const V3& x1, bool& in, double& d2, V3* w) {
  const V3 d = x1 - x0;
  const V3 a = x0 - p;
  const double l2 = sqnorm(d);
  const double t = -1.0 * dot(a, d) / l2;
  if (t <= 0.0) {
    if (w) *w = x0;
    in = false;
    d2 = sqnorm(x0 - p);
  } else if (t >= 1.0) {
    if (w) *w = x1;
    in = false;
    d2 = sqnorm(x1 - p);
  } else {
    const V3 lw = x0 + t * d;
    if (w) *w = lw;
    in = true;
    d2 = sqnorm(lw);
  }
}
void pointToSegment(const V3& p, const V3& x0, const V3& x1, MinDist& o) {
  pointToSegment(p, x0, x1, o.in, o.d2, &o.w);
}

// epa_polytope_utils.h:61-90
double pointToLine(const V3& p, const V3& a, const V3& b) {
  const V3 ab = b - a;
  const double l = norm(ab);
  if (l <= 0.0) return norm(p - a);
  return norm(cross(p - a, ab)) / l;
}
double pointToPlane(const V3& p, const V3& a, const V3& b, const V3& c) {
  const V3 n = cross(a - b, b - c);
  const double l = norm(n);
  if (l <= 0.0) return pointToLine(p, a, b);
  return std::fabs(dot(n / l, a - p));
}

// epa_polytope_utils.h:92-166
void pointToTriangle(const V3& p, const V3& a, const V3& b, const V3& c, MinDist& o) {
  const V3 dl[3] = {a - b, b - c, c - a};
  const V3 n = cross(dl[0], dl[1]);
  const double n2 = sqnorm(n);
  const double area = std::sqrt(n2);
  auto sub = [&]() {
    MinDist s[3];
    pointToSegment(p, a, b, s[0]);
    pointToSegment(p, b, c, s[1]);
    pointToSegment(p, c, a, s[2]);
    int k = 0;
    for (int i = 1; i < 3; ++i)
      if (s[i].d2 < s[k].d2) k = i;
    o.in = false;
    o.d2 = s[k].d2;
    o.w = s[k].w;
  };
  if (std::fabs(area) <= 1e-16) {
    sub();
    return;
  }
  const double d = dot(a - p, n);
  const V3 p2proj = n * (d / n2);
  const V3 proj = p + p2proj;
  double w[3];
  w[0] = norm(cross(dl[1], b - proj)) / area;
  w[1] = norm(cross(dl[2], c - proj)) / area;
  w[2] = 1.0 - w[0] - w[1];
  o.in = true;
  for (int i = 0; i < 3; ++i)
    if (w[i] < 0.0 || w[i] > 1.0) { o.in = false; break; }
  const double chk = norm(cross(dl[0], a - proj)) / area;
  if (std::fabs(chk - w[2]) > 1e-3) o.in = false;
  if (o.in) {
    o.d2 = sqnorm(p2proj);
    o.w = a * w[0] + b * w[1] + c * w[2];
    return;
  }
  sub();
}

enum EdgeVis { EUnknown, EBorder, EInternal };
enum FaceVis { FUnknown, FVisible, FHidden };

struct PVert {
  int next{kNull}, mnext{kNull};
  MVert mv;
  MinDist md;
  bool remove{false};
  int new_edge{kNull};
};
struct PEdge {
  int next{kNull}, mnext{kNull};
  int v[2]{kNull, kNull};
  int f[2]{kNull, kNull};
  MinDist md;
  EdgeVis vis{EUnknown};
};
struct PFace {
  int next{kNull}, mnext{kNull};
  int e[3]{kNull, kNull, kNull};
  MinDist md;
  FaceVis vis{FUnknown};
};

enum class Kind { Vertex, Edge, Face, None };
struct Feat {
  Kind k{Kind::None};
  int i{kNull};
};

struct Polytope {
  std::vector<PVert> V;
  std::vector<PEdge> E;
  std::vector<PFace> F;
  int vlist{kNull}, elist{kNull}, flist{kNull};
  int vfree{kNull}, efree{kNull}, ffree{kNull};

  // epa_polytope.hpp:77-125
  void reset(int max_faces) {
    const int nv = static_cast<int>(1.51 * double(max_faces));
    V.assign(nv, PVert());
    E.assign(nv, PEdge());
    F.assign(max_faces, PFace());
    vlist = elist = flist = kNull;
    vfree = efree = ffree = kNull;
    for (int i = nv - 1; i >= 0; --i) { V[i].mnext = vfree; vfree = i; }
    for (int i = nv - 1; i >= 0; --i) { E[i].mnext = efree; efree = i; }
    for (int i = max_faces - 1; i >= 0; --i) { F[i].mnext = ffree; ffree = i; }
  }
  int faceCapacity() const { return static_cast<int>(F.size()); }

  // epa_polytope.hpp:180-206
  int addVertex(const MVert& m) {
    if (vfree == kNull) return kNull;
    const int i = vfree;
    vfree = V[i].mnext;
    PVert& v = V[i];
    v.mv = m;
    v.md.in = true;
    v.md.w = m.v;
    v.md.d2 = sqnorm(m.v);
    v.next = vlist;
    vlist = i;
    return i;
  }
  // epa_polytope.hpp:208-235
  int addEdge(int v1, int v2) {
    if (v1 == kNull || v2 == kNull) return kNull;
    if (efree == kNull) return kNull;
    const int i = efree;
    efree = E[i].mnext;
    PEdge& e = E[i];
    e.v[0] = v1;
    e.v[1] = v2;
    e.f[0] = e.f[1] = kNull;
    pointToSegment(V3(0, 0, 0), V[v1].mv.v, V[v2].mv.v, e.md);
    e.next = elist;
    elist = i;
    return i;
  }
  // epa_polytope.hpp:237-290
  int addFace(int e1, int e2, int e3) {
    if (e1 == kNull || e2 == kNull || e3 == kNull) return kNull;
    if (ffree == kNull) return kNull;
    const int i = ffree;
    ffree = F[i].mnext;
    PFace& f = F[i];
    f.e[0] = e1;
    f.e[1] = e2;
    f.e[2] = e3;
    const V3 a = V[E[e1].v[0]].mv.v;
    const V3 b = V[E[e1].v[1]].mv.v;
    const PEdge& ee = E[e2];
    const V3 c = (ee.v[0] != E[e1].v[0] && ee.v[0] != E[e1].v[1]) ? V[ee.v[0]].mv.v : V[ee.v[1]].mv.v;
    pointToTriangle(V3(0, 0, 0), a, b, c, f.md);
    f.next = flist;
    flist = i;
    for (int k = 0; k < 3; ++k) {
      PEdge& ek = E[f.e[k]];
      if (ek.f[0] == kNull) {
        ek.f[0] = i;
      } else {
        if (ek.f[1] != kNull) return kNull;
        ek.f[1] = i;
      }
    }
    return i;
  }
  // epa_polytope.hpp:12-40
  bool faceVerts(int fi, int& a, int& b, int& c) const {
    const PFace& f = F[fi];
    if (f.e[0] == kNull || f.e[1] == kNull || f.e[2] == kNull) return false;
    const PEdge& e0 = E[f.e[0]];
    a = e0.v[0];
    b = e0.v[1];
    const PEdge& e1 = E[f.e[1]];
    c = (e1.v[0] != a && e1.v[0] != b) ? e1.v[0] : e1.v[1];
    return true;
  }
  // epa_polytope.hpp:292-345
  Feat minToOrigin(bool exclude_vertex) const {
    if (vlist == kNull) return Feat{};
    double best = INFINITY;
    Feat r;
    if (!exclude_vertex)
      for (int i = vlist; i != kNull; i = V[i].next)
        if (V[i].md.d2 < best) { r = {Kind::Vertex, i}; best = V[i].md.d2; }
    for (int i = elist; i != kNull; i = E[i].next) {
      if (!exclude_vertex && !E[i].md.in) continue;
      if (E[i].md.d2 < best) { r = {Kind::Edge, i}; best = E[i].md.d2; }
    }
    for (int i = flist; i != kNull; i = F[i].next) {
      if (!F[i].md.in) continue;
      if (F[i].md.d2 < best) { r = {Kind::Face, i}; best = F[i].md.d2; }
    }
    return r;
  }
  // epa_polytope.hpp:347-405
  bool faceNormalOut(int fi, V3& normal, double* area) const {
    int ia = 0, ib = 0, ic = 0;
    faceVerts(fi, ia, ib, ic);
    const V3& a = V[ia].mv.v;
    const V3& b = V[ib].mv.v;
    const V3& c = V[ic].mv.v;
    const V3 n = cross(a - b, b - c);
    const double nn = norm(n);
    if (area) *area = 0.5 * nn;
    if (nn <= 0.0) return false;
    const V3 dir = n / nn;
    const double od = dot(a, dir);
    if (std::fabs(od) >= 1e-4) {
      normal = od > 0 ? dir : -dir;
      return true;
    }
    double maxp = 0.0, minn = 0.0;
    for (int i = vlist; i != kNull; i = V[i].next) {
      const double dv = dot(V[i].mv.v, dir);
      if (dv > maxp) maxp = dv;
      if (dv < minn) minn = dv;
    }
    normal = std::fabs(maxp) > std::fabs(minn) ? -dir : dir;
    return true;
  }
  // epa_polytope_expand.hpp:12-30
  bool outside(int fi, const V3& p, double thr = 0.0) const {
    V3 n;
    double area = 0;
    if (!faceNormalOut(fi, n, &area)) return area <= 0.0;
    const V3& on = V[E[F[fi].e[0]].v[0]].mv.v;
    return dot(n, p - on) >= thr;
  }

  enum class Expand { OK, Failed, MallocFailed };
  // epa_polytope_expand.hpp:32-281
  Expand expand(const MVert& nv, int start) {
    for (int i = vlist; i != kNull; i = V[i].next) { V[i].remove = true; V[i].new_edge = kNull; }
    for (int i = elist; i != kNull; i = E[i].next) E[i].vis = EUnknown;
    for (int i = flist; i != kNull; i = F[i].next) F[i].vis = FUnknown;
    // visible patch flood fill with a LIFO stack
    {
      std::vector<std::pair<int, int>> stk;
      F[start].vis = FVisible;
      for (int k = 0; k < 3; ++k) stk.push_back({start, k});
      while (!stk.empty()) {
        const auto t = stk.back();
        stk.pop_back();
        const int fi = t.first;
        const int ei = F[fi].e[t.second];
        if (ei == kNull) return Expand::Failed;
        PEdge& e = E[ei];
        const int g = (e.f[0] == fi) ? e.f[1] : e.f[0];
        if (g == kNull) return Expand::Failed;
        if (F[g].vis == FVisible) { e.vis = EInternal; continue; }
        if (F[g].vis == FHidden) { e.vis = EBorder; continue; }
        if (outside(g, nv.v)) {
          F[g].vis = FVisible;
          e.vis = EInternal;
          for (int k = 0; k < 3; ++k) stk.push_back({g, k});
        } else {
          F[g].vis = FHidden;
          e.vis = EBorder;
        }
      }
    }
    for (int i = elist; i != kNull; i = E[i].next)
      if (E[i].vis != EInternal) { V[E[i].v[0]].remove = false; V[E[i].v[1]].remove = false; }
    // remove visible faces, internal edges, orphan vertices (list order,
    // each reclaimed element pushed on the front of its free list)
    {
      int prev = kNull;
      int cur = flist;
      while (cur != kNull) {
        PFace& f = F[cur];
        const int nx = f.next;
        if (f.vis == FVisible) {
          for (int k = 0; k < 3; ++k) {
            PEdge& ek = E[f.e[k]];
            if (ek.f[0] == cur) ek.f[0] = ek.f[1];
            ek.f[1] = kNull;
          }
          if (prev == kNull) flist = nx; else F[prev].next = nx;
          f.next = kNull;
          f.mnext = ffree;
          ffree = cur;
        } else {
          prev = cur;
        }
        cur = nx;
      }
    }
    {
      int prev = kNull;
      int cur = elist;
      while (cur != kNull) {
        PEdge& e = E[cur];
        const int nx = e.next;
        if (e.vis == EInternal) {
          if (prev == kNull) elist = nx; else E[prev].next = nx;
          e.next = kNull;
          e.mnext = efree;
          efree = cur;
        } else {
          prev = cur;
        }
        cur = nx;
      }
    }
    {
      int prev = kNull;
      int cur = vlist;
      while (cur != kNull) {
        PVert& v = V[cur];
        const int nx = v.next;
        if (v.remove) {
          if (prev == kNull) vlist = nx; else V[prev].next = nx;
          v.next = kNull;
          v.mnext = vfree;
          vfree = cur;
        } else {
          prev = cur;
        }
        cur = nx;
      }
    }
    const int nvi = addVertex(nv);
    if (nvi == kNull) return Expand::MallocFailed;
    bool ok = true;
    for (int i = elist; i != kNull; i = E[i].next) {
      if (E[i].vis != EBorder) continue;
      int ne[2];
      for (int k = 0; k < 2; ++k) {
        PVert& vk = V[E[i].v[k]];
        if (vk.new_edge == kNull) vk.new_edge = addEdge(nvi, E[i].v[k]);
        ne[k] = vk.new_edge;
      }
      if (addFace(i, ne[0], ne[1]) == kNull) ok = false;
    }
    return ok ? Expand::OK : Expand::MallocFailed;
  }
};

enum class S2P { OK, Touching, Failed };

// epa_simplex2polytope.hpp:11-44
void touchingPoint(const MinkDiff& md, const V3& d, V3* p0, V3* p1) {
  if (!p0 && !p1) return;
  const V3 mid = (md.support0(d) + md.support1(-d)) / 2;
  if (p0) *p0 = mid;
  if (p1) *p1 = mid;
}

// epa_simplex2polytope.hpp:184-217
S2P tetra(const MVert& a, const MVert& b, const MVert& c, const MVert& d, Polytope& P) {
  P.reset(P.faceCapacity());
  int v[4];
  v[0] = P.addVertex(a);
  v[1] = P.addVertex(b);
  v[2] = P.addVertex(c);
  v[3] = P.addVertex(d);
  int e[6];
  e[0] = P.addEdge(v[0], v[1]);
  e[1] = P.addEdge(v[1], v[2]);
  e[2] = P.addEdge(v[2], v[0]);
  e[3] = P.addEdge(v[3], v[0]);
  e[4] = P.addEdge(v[3], v[1]);
  e[5] = P.addEdge(v[3], v[2]);
  int f[4];
  f[0] = P.addFace(e[0], e[1], e[2]);
  f[1] = P.addFace(e[3], e[4], e[0]);
  f[2] = P.addFace(e[4], e[5], e[1]);
  f[3] = P.addFace(e[5], e[3], e[2]);
  for (int i = 0; i < 4; ++i)
    if (f[i] == kNull) return S2P::Failed;
  return S2P::OK;
}

// epa_simplex2polytope.hpp:145-182
S2P s2p3(const MVert& a, const MVert& b, const MVert& c, const MinkDiff& md, Polytope& P, V3* p0,
         V3* p1, double thr) {
  V3 n = cross(b.v - a.v, c.v - a.v);
  if (sqnorm(n) <= 0.0) return S2P::Failed;
  normalize(n);
  const V3 d0 = md.support(n);
  const double t0 = pointToPlane(d0, a.v, b.v, c.v);
  const V3 dir1 = -n;
  const V3 d1 = md.support(dir1);
  const double t1 = pointToPlane(d1, a.v, b.v, c.v);
  if (t0 <= thr) { touchingPoint(md, n, p0, p1); return S2P::Touching; }
  if (t1 <= thr) { touchingPoint(md, dir1, p0, p1); return S2P::Touching; }
  if (t0 > t1) return tetra(a, b, c, MVert{d0, n}, P);
  return tetra(a, b, c, MVert{d1, dir1}, P);
}

// epa_simplex2polytope.hpp:112-143
S2P s2p4(const MVert& a, const MVert& b, const MVert& c, const MVert& d, const MinkDiff& md,
         Polytope& P, V3* p0, V3* p1, double thr) {
  const V3 o(0, 0, 0);
  if (pointToPlane(o, a.v, b.v, c.v) < thr) return s2p3(a, b, c, md, P, p0, p1, thr);
  if (pointToPlane(o, a.v, c.v, d.v) < thr) return s2p3(a, c, d, md, P, p0, p1, thr);
  if (pointToPlane(o, a.v, b.v, d.v) < thr) return s2p3(a, b, d, md, P, p0, p1, thr);
  if (pointToPlane(o, b.v, c.v, d.v) < thr) return s2p3(b, c, d, md, P, p0, p1, thr);
  return tetra(a, b, c, d, P);
}

// epa_simplex2polytope.hpp:219-380
S2P s2p2(const MVert& a, const MVert& b, const MinkDiff& md, Polytope& P, V3* p0, V3* p1,
         double thr) {
  const double thr2 = thr * thr;
  const V3 ab = b.v - a.v;
  if (sqnorm(ab) < thr2) return S2P::Failed;
  const V3 abu = normalized(ab);
  V3 dinit(abu.y, -abu.x, 0);
  bool valid = sqnorm(dinit) > thr2;
  if (!valid) { dinit = V3(abu.z, 0, -abu.x); valid = sqnorm(dinit) > thr2; }
  if (!valid) { dinit = V3(0, abu.z, -abu.y); valid = sqnorm(dinit) > thr2; }
  if (!valid) return S2P::Failed;
  normalize(dinit);
  constexpr int kDiv = 72;
  V3 dcur = dinit, dfound, vfound;
  bool found = false;
  for (int i = 0; i < kDiv; ++i) {
    const V3 vcur = md.support(dcur);
    if (pointToLine(vcur, a.v, b.v) > thr) {
      vfound = vcur;
      dfound = dcur;
      found = true;
      break;
    }
    constexpr double kPi = 3.145926;  // value used by the reference (sic)
    constexpr double kStep = 2.0 * kPi / kDiv;
    dcur = dcur + kStep * cross(abu, vcur);
    normalize(dcur);
  }
  if (!found) { touchingPoint(md, dinit, p0, p1); return S2P::Touching; }
  const MVert v0{vfound, dfound};
  MVert v1;
  v1.d = -v0.d;
  v1.v = md.support(v1.d);
  if (pointToLine(v1.v, a.v, b.v) < thr) { touchingPoint(md, v1.d, p0, p1); return S2P::Touching; }
  const V3 n = cross(v0.v - a.v, v1.v - a.v);
  if (sqnorm(n) < thr2) return S2P::Failed;
  MVert v2;
  v2.d = normalized(n);
  v2.v = md.support(v2.d);
  if (pointToLine(v2.v, a.v, b.v) < thr) { touchingPoint(md, v2.d, p0, p1); return S2P::Touching; }
  MVert v3;
  v3.d = -v2.d;
  v3.v = md.support(v3.d);
  if (pointToLine(v3.v, a.v, b.v) < thr) { touchingPoint(md, v3.d, p0, p1); return S2P::Touching; }
  P.reset(P.faceCapacity());
  int v[6];
  v[0] = P.addVertex(a);
  v[1] = P.addVertex(v0);
  v[2] = P.addVertex(b);
  v[3] = P.addVertex(v1);
  v[4] = P.addVertex(v2);
  v[5] = P.addVertex(v3);
  int e[12];
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
  int f[8];
  f[0] = P.addFace(e[4], e[5], e[0]);
  f[1] = P.addFace(e[5], e[6], e[1]);
  f[2] = P.addFace(e[6], e[7], e[2]);
  f[3] = P.addFace(e[7], e[4], e[3]);
  f[4] = P.addFace(e[8], e[9], e[0]);
  f[5] = P.addFace(e[9], e[10], e[1]);
  f[6] = P.addFace(e[10], e[11], e[2]);
  f[7] = P.addFace(e[11], e[8], e[3]);
  for (int i = 0; i < 8; ++i)
    if (f[i] == kNull) return S2P::Failed;
  return S2P::OK;
}

// epa_simplex2polytope.hpp:46-110
S2P simplexToPolytope(const Simplex& s, const MinkDiff& md, Polytope& P, V3* p0, V3* p1,
                      double thr) {
  const double thr2 = thr * thr;
  for (int i = 0; i < s.rank; ++i)
    if (sqnorm(s.v[i].v) <= thr2) { touchingPoint(md, s.v[i].d, p0, p1); return S2P::Touching; }
  if (s.rank == 4) return s2p4(s.v[0], s.v[1], s.v[2], s.v[3], md, P, p0, p1, thr);
  if (s.rank == 3) return s2p3(s.v[0], s.v[1], s.v[2], md, P, p0, p1, thr);
  if (s.rank == 2) return s2p2(s.v[0], s.v[1], md, P, p0, p1, thr);
  const V3& d = s.v[0].d;
  if (p0) *p0 = md.support0(d);
  if (p1) *p1 = md.support1(-d);
  return S2P::Touching;
}

struct RawFeat {
  Kind k{Kind::None};
  MinDist md;
  MVert a, b, c;
};

bool toRaw(const Polytope& P, Feat f, RawFeat& r) {
  r.k = f.k;
  if (f.k == Kind::Vertex) {
    r.md = P.V[f.i].md;
    r.a = P.V[f.i].mv;
    return true;
  }
  if (f.k == Kind::Edge) {
    const PEdge& e = P.E[f.i];
    r.md = e.md;
    r.a = P.V[e.v[0]].mv;
    r.b = P.V[e.v[1]].mv;
    return true;
  }
  if (f.k == Kind::Face) {
    r.md = P.F[f.i].md;
    int ia, ib, ic;
    if (!P.faceVerts(f.i, ia, ib, ic)) return false;
    r.a = P.V[ia].mv;
    r.b = P.V[ib].mv;
    r.c = P.V[ic].mv;
    return true;
  }
  return false;
}

enum class NextDir { OK, Failed, Converge };

// epa.hpp:11-113
NextDir nextDirection(const Polytope& P, const MinkDiff& md, Feat nf, double tol, V3& nd, V3& nv,
                      int& start) {
  constexpr double kOuter = 1e-3 * 1e-3;
  auto face_cand = [&](int fi, bool try_witness) {
    V3 fn;
    if (!P.faceNormalOut(fi, fn, nullptr)) return NextDir::Failed;
    const V3& on = P.V[P.E[P.F[fi].e[0]].v[0]].mv.v;
    const MinDist& m = P.F[fi].md;
    if (try_witness && m.d2 > kOuter) {
      nd = normalized(m.w);
      nv = md.support(nd);
      if (dot(fn, nv - on) >= 0.0) { start = fi; return NextDir::OK; }
    }
    nd = fn;
    nv = md.support(nd);
    if (dot(fn, nv - on) > tol) { start = fi; return NextDir::OK; }
    return NextDir::Converge;
  };
  if (nf.k == Kind::Face) return face_cand(nf.i, true);
  if (nf.k == Kind::Edge) {
    const PEdge& e = P.E[nf.i];
    const int f0 = e.f[0], f1 = e.f[1];
    // An open edge only arises from a malformed polytope, where the reference
    // dereferences a null face; report failure instead.
    if (f0 == kNull || f1 == kNull) return NextDir::Failed;
    if (e.md.d2 > kOuter) {
      nd = normalized(e.md.w);
      nv = md.support(nd);
      if (P.outside(f0, nv)) { start = f0; return NextDir::OK; }
      if (P.outside(f1, nv)) { start = f1; return NextDir::OK; }
    }
    const NextDir s0 = face_cand(f0, false);
    if (s0 == NextDir::OK) return s0;
    return face_cand(f1, false);
  }
  return NextDir::Failed;
}

// epa.hpp:269-299
bool terminate(const RawFeat& f, const V3& nv, double tol) {
  double d2 = 0.0;
  if (f.k == Kind::Edge) {
    bool in = false;
    pointToSegment(nv, f.a.v, f.b.v, in, d2, nullptr);
  } else {
    const double t = pointToPlane(nv, f.a.v, f.b.v, f.c.v);
    d2 = t * t;
  }
  return d2 < tol * tol;
}

// epa.hpp:376-432
bool pairFromSegment(const MinkDiff& md, const RawFeat& f, double* depth, V3* p0, V3* p1) {
  if (depth) *depth = std::sqrt(f.md.d2);
  const V3& p = f.md.w;
  const V3& a = f.a.v;
  const V3& b = f.b.v;
  const V3 ab = b - a;
  int ax = 0;
  double mx = std::fabs(ab.x);
  if (std::fabs(ab.y) > mx) { ax = 1; mx = std::fabs(ab.y); }
  if (std::fabs(ab.z) > mx) { ax = 2; mx = std::fabs(ab.z); }
  if (mx <= 1e-10) {
    if (p0) *p0 = md.support0(f.a.d);
    if (p1) *p1 = md.support1(-f.a.d);
    return true;
  }
  const double s = (p[ax] - a[ax]) / (b[ax] - a[ax]);
  if (s < 0 || s > 1.0) return false;
  const double wa = 1.0 - s, wb = s;
  if (p0) *p0 = wa * md.support0(f.a.d) + wb * md.support0(f.b.d);
  if (p1) *p1 = wa * md.support1(-f.a.d) + wb * md.support1(-f.b.d);
  return true;
}

// epa.hpp:434-498
bool pairFromFace(const MinkDiff& md, const RawFeat& f, double* depth, V3* p0, V3* p1) {
  if (depth) *depth = std::sqrt(f.md.d2);
  const V3& a = f.a.v;
  const V3& b = f.b.v;
  const V3& c = f.c.v;
  const V3& p = f.md.w;
  const V3 dl[3] = {a - b, b - c, c - a};
  const V3 n = cross(dl[0], dl[1]);
  const double area = std::sqrt(sqnorm(n));
  double w[3];
  w[0] = norm(cross(dl[1], b - p)) / area;
  w[1] = norm(cross(dl[2], c - p)) / area;
  const double chk = norm(cross(dl[0], a - p)) / area;
  w[2] = 1.0 - w[0] - w[1];
  if (std::fabs(chk - w[2]) > 0.01) return false;
  if (p0) *p0 = w[0] * md.support0(f.a.d) + w[1] * md.support0(f.b.d) + w[2] * md.support0(f.c.d);
  if (p1) *p1 = w[0] * md.support1(-f.a.d) + w[1] * md.support1(-f.b.d) + w[2] * md.support1(-f.c.d);
  return true;
}

// epa.hpp:341-374
void assignPair(const MinkDiff& md, const MVert& cand, const RawFeat& f, double* depth, V3* p0,
                V3* p1) {
  if (f.k == Kind::Vertex) {
    const V3& d = f.a.d;
    if (p0) *p0 = md.support0(d);
    if (p1) *p1 = md.support1(-d);
    if (depth) *depth = norm(f.a.v);
    return;
  }
  if (f.k == Kind::Edge) {
    if (pairFromSegment(md, f, depth, p0, p1)) return;
  } else if (f.k == Kind::Face) {
    if (pairFromFace(md, f, depth, p0, p1)) return;
  }
  const V3 q0 = md.support0(cand.d);
  const V3 q1 = md.support1(-cand.d);
  if (p0) *p0 = q0;
  if (p1) *p1 = q1;
  if (depth) *depth = norm(q0 - q1);
}

}  // namespace

// epa.hpp:137-257
EPAStatus epaEvaluate(const Simplex& simplex, const MinkDiff& md, double tol, int max_faces,
                      int max_iter, double* depth, V3* p0, V3* p1, EPAStats* stats) {
  Polytope P;
  P.reset(max_faces);
  const S2P s2p = simplexToPolytope(simplex, md, P, p0, p1, tol);
  if (s2p == S2P::Failed) return EPAStatus::Failed;
  if (s2p == S2P::Touching) {
    if (depth) *depth = 0.0;
    return EPAStatus::Touching;
  }
  int it = 0;
  RawFeat raw;
  while (true) {
    Feat f = P.minToOrigin(false);
    if (f.k == Kind::None) return EPAStatus::Failed;
    if (f.k == Kind::Vertex) {
      const double vd2 = P.V[f.i].md.d2;
      constexpr double kRatio = 1.0 - 1e-3;
      const Feat o = P.minToOrigin(true);
      if (o.k == Kind::None) return EPAStatus::Failed;
      if (o.k == Kind::Edge) {
        if (kRatio * P.E[o.i].md.d2 < vd2) f = o;
      } else if (o.k == Kind::Face) {
        if (kRatio * P.F[o.i].md.d2 < vd2) f = o;
      }
    }
    if (f.k != Kind::Edge && f.k != Kind::Face) return EPAStatus::Failed;
    toRaw(P, f, raw);
    V3 nd, nv;
    int start = kNull;
    const NextDir ds = nextDirection(P, md, f, tol, nd, nv, start);
    if (ds == NextDir::Failed) return EPAStatus::Failed;
    if (ds == NextDir::Converge) {
      assignPair(md, MVert{nv, nd}, raw, depth, p0, p1);
      if (stats) stats->iterations = it;
      return EPAStatus::OK;
    }
    if (start == kNull) return EPAStatus::Failed;
    const MVert next{nv, nd};
    if (terminate(raw, nv, tol)) {
      assignPair(md, next, raw, depth, p0, p1);
      if (stats) stats->iterations = it;
      return EPAStatus::OK;
    }
    const auto ex = P.expand(next, start);
    if (ex == Polytope::Expand::Failed) return EPAStatus::Failed;
    if (ex == Polytope::Expand::MallocFailed) {
      assignPair(md, next, raw, depth, p0, p1);
      if (stats) stats->iterations = it;
      return EPAStatus::MallocFailed;
    }
    ++it;
    if (it >= max_iter) {
      assignPair(md, next, raw, depth, p0, p1);
      if (stats) stats->iterations = it;
      return EPAStatus::IterationLimit;
    }
  }
}

}  // namespace oracle
