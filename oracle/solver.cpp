// CPU oracle -- TEST INFRASTRUCTURE ONLY (see omath.h).
// fcl::collide on primitive / convex geometry (collision_interface-inl.h:13-32,
// collision-inl.h:68-116, shape_pair_intersect-inl.h:49-114,
// gjk_solver-inl.h:71-148 and the closed-form registrations :209-217) and
// GJKSolver::shapeSignedDistance (gjk_solver-inl.h:820-874).
#include <algorithm>

#include "narrowphase.h"

namespace oracle {
namespace {

// gjk_solver-inl.h:71-137, the generic support-mapping path.
bool genericIntersect(const Shape& a, const Tf& tfa, const Shape& b, const Tf& tfb,
                      const SolverParams& sp, std::vector<ContactPt>* contacts, CollideOut& o) {
  const V3 guess = sp.guess_valid ? sp.guess : V3(1, 0, 0);
  const MinkDiff md = makeMinkDiff(a, tfa, b, tfb);
  if (contacts == nullptr) {
    o.method = 2;
    const MPRStatus r = mprIntersect(md, sp.gjk_max_iterations, sp.gjk_tolerance);
    if (r == MPRStatus::Intersect) return true;
    if (r == MPRStatus::Separated) return false;
  }
  o.method = 1;
  Simplex s;
  GJKStats gs;
  const GJKStatus g = gjkEvaluate(md, s, -guess, sp.gjk_tolerance, sp.gjk_max_iterations, nullptr, &gs);
  o.gjk_status = static_cast<int>(g);
  o.gjk_iters = gs.iterations;
  if (contacts == nullptr) return g == GJKStatus::Intersect;
  if (g != GJKStatus::Intersect) return false;
  double depth = 0;
  V3 p1, p2;
  EPAStats es;
  const EPAStatus e = epaEvaluate(s, md, sp.epa_tolerance, sp.epa_max_face_num,
                                  sp.epa_max_iterations, &depth, &p1, &p2, &es);
  o.epa_status = static_cast<int>(e);
  o.epa_iters = es.iterations;
  if (e == EPAStatus::Failed) return false;
  V3 n1 = p1 - p2;
  if (sqnorm(n1) <= 0) n1 = V3(0, 0, 1);
  else normalize(n1);
  const V3 pt1 = 0.5 * (p1 + p2);
  contacts->push_back({tfa.R * n1, tfa * pt1, depth});
  return true;
}

void flip(std::vector<ContactPt>* c) {
  if (c)
    for (auto& x : *c) x.normal = -x.normal;
}

bool shapeIntersect(const Shape& a, const Tf& tfa, const Shape& b, const Tf& tfb,
                    const SolverParams& sp, std::vector<ContactPt>* c, CollideOut& o) {
  const int ta = a.type, tb = b.type;
  o.method = 0;
  if (ta == kSphere && tb == kSphere) return sphereSphereIntersect(a.p[0], tfa, b.p[0], tfb, c);
  if (ta == kBox && tb == kBox) return boxBoxIntersect(V3(a.p[0], a.p[1], a.p[2]), tfa, V3(b.p[0], b.p[1], b.p[2]), tfb, c);
  if (ta == kSphere && tb == kCapsule) return sphereCapsuleIntersect(a.p[0], tfa, b.p[0], b.p[1], tfb, c);
  if (ta == kCapsule && tb == kSphere) {
    const bool r = sphereCapsuleIntersect(b.p[0], tfb, a.p[0], a.p[1], tfa, c);
    flip(c);
    return r;
  }
  if (ta == kSphere && tb == kBox) return sphereBoxIntersect(a.p[0], tfa, V3(b.p[0], b.p[1], b.p[2]), tfb, c);
  if (ta == kBox && tb == kSphere) {
    const bool r = sphereBoxIntersect(b.p[0], tfb, V3(a.p[0], a.p[1], a.p[2]), tfa, c);
    flip(c);
    return r;
  }
  if (ta == kSphere && tb == kCylinder) return sphereCylinderIntersect(a.p[0], tfa, b.p[0], b.p[1], tfb, c);
  if (ta == kCylinder && tb == kSphere) {
    const bool r = sphereCylinderIntersect(b.p[0], tfb, a.p[0], a.p[1], tfa, c);
    flip(c);
    return r;
  }
  return genericIntersect(a, tfa, b, tfb, sp, c, o);
}

}  // namespace

void collide(const Shape& a, const Tf& tfa, const Shape& b, const Tf& tfb, int max_contacts,
             bool penetration, const SolverParams& sp, CollideOut& out) {
  out = CollideOut();
  if (max_contacts <= 0) return;
  if (!penetration) {
    out.hit = shapeIntersect(a, tfa, b, tfb, sp, nullptr, out);
    out.n = out.hit ? 1 : 0;
    return;
  }
  std::vector<ContactPt> cs;
  out.hit = shapeIntersect(a, tfa, b, tfb, sp, &cs, out);
  if (!out.hit) return;
  // shape_pair_intersect-inl.h:91-113: keep the deepest contacts when they
  // do not all fit (std::partial_sort with depth-descending order).
  const int free_space = max_contacts;
  int n = static_cast<int>(cs.size());
  if (free_space < n) {
    std::partial_sort(cs.begin(), cs.begin() + free_space, cs.end(),
                      [](const ContactPt& x, const ContactPt& y) { return x.depth > y.depth; });
    n = free_space;
  }
  n = std::min(n, 4);
  for (int i = 0; i < n; ++i) out.c[i] = cs[i];
  out.n = n;
}

void signedDistance(const Shape& a, const Tf& tfa, const Shape& b, const Tf& tfb,
                    const SolverParams& sp, SignedDistOut& out) {
  out = SignedDistOut();
  const V3 guess = sp.guess_valid ? sp.guess : V3(1, 0, 0);
  const MinkDiff md = makeMinkDiff(a, tfa, b, tfb);
  Simplex s;
  GJKDistOut d;
  GJKStats gs;
  const GJKStatus g = gjkEvaluate(md, s, -guess, sp.gjk_tolerance, sp.gjk_max_iterations, &d, &gs);
  out.gjk_status = static_cast<int>(g);
  out.gjk_iters = gs.iterations + gs.dist_iterations;
  out.support_calls = gs.support_calls;
  if (g == GJKStatus::Separated && d.sep_valid) {
    out.pa = tfa * d.p0;
    out.pb = tfa * d.p1;
    out.distance = norm(d.p0 - d.p1);
    out.ok = true;
    return;
  }
  if (g == GJKStatus::Intersect) {
    V3 p1, p2;
    double depth = 0;
    EPAStats es;
    const EPAStatus e = epaEvaluate(s, md, sp.epa_tolerance, sp.epa_max_face_num,
                                    sp.epa_max_iterations, &depth, &p1, &p2, &es);
    out.epa_status = static_cast<int>(e);
    out.epa_iters = es.iterations;
    if (e != EPAStatus::Failed) {
      out.pa = tfa * p1;
      out.pb = tfa * p2;
      out.distance = -depth;
      out.ok = true;
      return;
    }
    return;
  }
  out.distance = -1;
}

}  // namespace oracle
