// CPU oracle -- TEST INFRASTRUCTURE ONLY (see omath.h).
// MPR boolean: include/fcl/cvx_collide/mpr.hpp:34-185 (RunIntersect),
// :188-328 (findPortal), :436-495 (updatePortal); directed penetration
// :496-632 (RunDirectedPenetration), :642-704 (its result); incremental
// minimum penetration: mpr_incremental_penetration.hpp:10-337.
#include <algorithm>
#include <cmath>
#include <limits>
#include <utility>

#include "narrowphase.h"

namespace oracle {
namespace {

constexpr double kDotEps = 2.220446049250313e-16;  // numeric_limits<double>::epsilon(), mpr.h:103

// mpr.hpp:11-16: the direction is normalised in place.
V3 support(const MinkDiff& md, V3& d) {
  normalize(d);
  return md.support(d);
}

enum class Portal { IterationLimit, Separated, Found };

Portal findPortal(const MinkDiff& md, const V3& v0, V3& v1, V3& v2, V3& v3, V3& d1, V3& d2,
                  V3& d3, int max_iter) {
  const double v0n = absnorm(v0);
  int it = 0;
  while (true) {
    if (it >= max_iter) return Portal::IterationLimit;
    ++it;
    V3 a1 = v1 - v0, a2 = v2 - v0, a3 = v3 - v0;
    V3 n031 = cross(a3, a1);
    V3 n012 = cross(a1, a2);
    const double vol = dot(a2, n031);
    if (vol < 0) {
      std::swap(v2, v3);
      std::swap(d2, d3);
      std::swap(a2, a3);
      std::swap(n012, n031);
      n031 = -n031;
      n012 = -n012;
    }
    if (dot(v0, n031) > kDotEps * v0n * absnorm(n031)) {
      V3 s = -n031;
      v2 = support(md, s);
      d2 = s;
      if (dot(v2, s) < 0) return Portal::Separated;
      continue;
    }
    if (dot(v0, n012) > kDotEps * v0n * absnorm(n012)) {
      V3 s = -n012;
      v3 = support(md, s);
      d3 = s;
      if (dot(v3, s) < 0) return Portal::Separated;
      continue;
    }
    const V3 n023 = cross(a2, a3);
    if (dot(v0, n023) > kDotEps * v0n * absnorm(n023)) {
      V3 s = -n023;
      v1 = support(md, s);
      d1 = s;
      if (dot(v1, s) < 0) return Portal::Separated;
      continue;
    }
    return Portal::Found;
  }
}

void updatePortal(const V3& v0, const V3& v4, const V3& n, V3& v1, V3& v2, V3& v3, V3& d1,
                  V3& d2, V3& d3) {
  const V3 nn = cross(v4, v0);
  if (dot(v1, nn) > 0) {
    if (dot(v2, nn) > 0) { v1 = v4; d1 = n; }
    else { v3 = v4; d3 = n; }
  } else {
    if (dot(v3, nn) > 0) { v2 = v4; d2 = n; }
    else { v1 = v4; d1 = n; }
  }
}

}  // namespace

MPRStatus mprIntersect(const MinkDiff& md, int max_iter, double tol) {
  const V3 v0 = md.interior();
  if (sqnorm(v0) <= tol * tol) return MPRStatus::Intersect;
  V3 d1 = -v0;
  V3 v1 = support(md, d1);
  if (dot(d1, v1) < 0) return MPRStatus::Separated;
  V3 d2 = cross(v0, v1);
  if (absnorm(d2) <= absnorm(v0) * absnorm(v1) * tol) return MPRStatus::Intersect;
  V3 v2 = support(md, d2);
  if (dot(d2, v2) < 0) return MPRStatus::Separated;
  V3 d3 = cross(v1 - v0, v2 - v0);
  if (dot(d3, v0) > 0) {
    std::swap(v1, v2);
    std::swap(d1, d2);
    d3 = -d3;
  }
  V3 v3 = support(md, d3);
  if (dot(d3, v3) < 0) return MPRStatus::Separated;
  const Portal ps = findPortal(md, v0, v1, v2, v3, d1, d2, d3, max_iter);
  if (ps == Portal::IterationLimit) return MPRStatus::Failed;
  if (ps == Portal::Separated) return MPRStatus::Separated;
  int it = 0;
  while (it < max_iter) {
    ++it;
    V3 n = cross(v2 - v1, v3 - v1);
    if (dot(n, v0) > 0) {
      std::swap(v2, v3);
      std::swap(d2, d3);
      n = -n;
    }
    if (!(dot(v1, n) < 0)) return MPRStatus::Intersect;
    const V3 v4 = support(md, n);  // n is normalised in place
    if (dot(v4, n) < 0) return MPRStatus::Separated;
    if (std::fabs(dot(v4 - v1, n)) < tol * absnorm(n)) return MPRStatus::Separated;
    updatePortal(v0, v4, n, v1, v2, v3, d1, d2, d3);
  }
  return MPRStatus::Failed;
}

namespace {

// mpr.hpp:642-704: the penetration distance along d through the portal
// (v1, v2, v3) with normal n, and the witness points as the barycentric
// blend of the portal vertices' support points.
void directedResult(const MinkDiff& md, const V3& d, const V3& n, const V3& v1, const V3& v2, const V3& v3,
                    const V3& d1, const V3& d2, const V3& d3, const V3* v4, DirectedOut& o) {
  o.portal_normal = n;
  const double nd = dot(d, n);
  if (std::fabs(nd) <= 0) {
    const double e[3] = {dot(v1, d), dot(v2, d), dot(v3, d)};
    const V3* dirs[3] = {&d1, &d2, &d3};
    double best = -std::numeric_limits<double>::infinity();
    int bi = 0;
    for (int i = 0; i < 3; ++i)
      if (e[i] > best) {
        best = e[i];
        bi = i;
      }
    o.distance = best;
    o.p0 = md.support0(*dirs[bi]);
    o.p1 = md.support1(-*dirs[bi]);
    return;
  }
  const double to_portal = dot(v1, n) / nd;
  o.distance = v4 ? dot(*v4, n) / nd : to_portal;
  const V3 q = to_portal * d;
  const V3 e12 = v2 - v1, e13 = v3 - v1;
  const double area = norm(cross(e12, e13));
  const double w2 = norm(cross(e13, v1 - q)) / area;
  const double w3 = norm(cross(e12, v1 - q)) / area;
  const double w1 = 1.0 - w2 - w3;
  o.p0 = md.support0(d1) * w1 + md.support0(d2) * w2 + md.support0(d3) * w3;
  o.p1 = md.support1(-d1) * w1 + md.support1(-d2) * w2 + md.support1(-d3) * w3;
}

}  // namespace

DirectedStatus mprDirected(const MinkDiff& md, const V3& d, int max_iter, double tol, DirectedOut& o) {
  // the ray starts at a mock point -d (not necessarily inside the difference)
  const V3 v0 = -d;
  V3 d1 = d;
  V3 v1 = support(md, d1);
  if (dot(d1, v1) < 0) return DirectedStatus::FailedNoIntersect;
  V3 d2 = cross(v0, v1);
  if (absnorm(d2) <= absnorm(v1) * tol) {  // v1 on the ray
    o.distance = dot(v1, d);
    o.portal_normal = d1;
    o.p0 = md.support0(d1);
    o.p1 = md.support1(-d1);
    return DirectedStatus::OK;
  }
  V3 v2 = support(md, d2);
  if (dot(d2, v2) < 0) return DirectedStatus::FailedNoIntersect;
  V3 d3 = cross(v1, v2);
  if (dot(d3, v0) > 0) {
    std::swap(v1, v2);
    std::swap(d1, d2);
    d3 = d3 * -1.0;
  }
  V3 v3 = support(md, d3);
  if (dot(d3, v3) < 0) return DirectedStatus::FailedNoIntersect;
  const double m2 = std::max(sqnorm(v1), std::max(sqnorm(v2), sqnorm(v3)));
  const V3 v0s = v0 * std::sqrt(m2);
  const Portal ps = findPortal(md, v0s, v1, v2, v3, d1, d2, d3, max_iter);
  if (ps == Portal::IterationLimit) return DirectedStatus::Failed;
  if (ps == Portal::Separated) return DirectedStatus::FailedNoIntersect;
  for (int it = 1; it <= max_iter; ++it) {
    o.iterations = it;
    V3 n = cross(v2 - v1, v3 - v1);
    if (dot(n, d) < 0) {
      std::swap(v2, v3);
      std::swap(d2, d3);
      n = n * -1.0;
    }
    const V3 v4 = support(md, n);  // n is normalised in place
    if (dot(v4, n) < 0) return DirectedStatus::FailedNoIntersect;
    if (std::fabs(dot(v4 - v1, n)) < tol * absnorm(n)) {
      directedResult(md, d, n, v1, v2, v3, d1, d2, d3, &v4, o);
      return DirectedStatus::OK;
    }
    updatePortal(v0s, v4, n, v1, v2, v3, d1, d2, d3);
  }
  return DirectedStatus::FailedRefinementIterationLimit;
}

namespace {

enum class Explore { NewDirection, Converged, Degenerated, Failed, FailedNoIntersect };

// mpr_incremental_penetration.hpp:10-34: the distance along d to the
// portal plane (lb) and to the parallel plane through v4 (ub).
void exploredBounds(const V3& d, const V3& v1, const V3& v2, const V3& v3, const V3& n, double v1n, double v4n,
                    double& lb, double& ub) {
  const double nd = dot(d, n);
  if (std::fabs(nd) <= 0) {
    lb = std::max(dot(v1, d), std::max(dot(v2, d), dot(v3, d)));
    ub = lb;
    return;
  }
  lb = v1n / nd;
  ub = v4n / nd;
}

// mpr_incremental_penetration.hpp:36-186: a directed portal search along d
// that stops early when the portal's normal is a better direction.
Explore exploreDirection(const MinkDiff& md, const V3& d, V3& v1, V3& v2, V3& v3, V3& d1, V3& d2, V3& d3,
                         double& lb, double& ub, V3& new_d, bool portal_valid, int max_iter, double tol) {
  const V3 v0 = -d;
  V3 sd = d;
  const V3 s0 = support(md, sd);
  if (dot(d, s0) < 0) return Explore::FailedNoIntersect;
  if (!portal_valid) {
    d1 = d;
    v1 = s0;
    d2 = cross(v0, v1);
    if (absnorm(d2) <= absnorm(v1) * tol) {
      lb = dot(v1, d);
      ub = lb;
      new_d = d;
      return Explore::Degenerated;
    }
    v2 = support(md, d2);
    if (dot(d2, v2) < 0) return Explore::FailedNoIntersect;
    d3 = cross(v1, v2);
    if (dot(d3, v0) > 0) {
      std::swap(v1, v2);
      std::swap(d1, d2);
      d3 = d3 * -1.0;
    }
    v3 = support(md, d3);
    if (dot(d3, v3) < 0) return Explore::FailedNoIntersect;
  }
  const double m2 = std::max(sqnorm(v1), std::max(sqnorm(v2), sqnorm(v3)));
  const V3 v0s = v0 * std::sqrt(m2);
  const Portal ps = findPortal(md, v0s, v1, v2, v3, d1, d2, d3, max_iter);
  if (ps == Portal::IterationLimit) return Explore::Failed;
  if (ps == Portal::Separated) return Explore::FailedNoIntersect;
  for (int it = 0; it < max_iter; ++it) {
    V3 n = cross(v2 - v1, v3 - v1);
    if (dot(n, d) < 0) {
      std::swap(v2, v3);
      std::swap(d2, d3);
      n = n * -1.0;
    }
    const V3 v4 = support(md, n);  // n is normalised in place
    const double ub_plane = dot(v4, n);
    if (ub_plane < 0) return Explore::FailedNoIntersect;
    const double v1n = dot(v1, n), v4n = dot(v4, n);
    const double gap = v4n - v1n;
    exploredBounds(d, v1, v2, v3, n, v1n, v4n, lb, ub);
    if (lb > ub_plane + tol) {
      new_d = n;
      return Explore::NewDirection;
    }
    if (std::fabs(gap) <= tol) {
      new_d = n;
      return Explore::Converged;
    }
    updatePortal(v0s, v4, n, v1, v2, v3, d1, d2, d3);
  }
  return Explore::Failed;
}

// mpr_incremental_penetration.hpp:298-334
void incrementalResult(const MinkDiff& md, const V3& d, const V3& v1, const V3& v2, const V3& v3, const V3& d1,
                       const V3& d2, const V3& d3, double lb, double ub, IncrementalOut& o) {
  o.depth = ub;
  o.direction = d;
  const V3 q = lb * d;
  const V3 e12 = v2 - v1, e13 = v3 - v1;
  const double area = norm(cross(e12, e13));
  const double w2 = norm(cross(e13, v1 - q)) / area;
  const double w3 = norm(cross(e12, v1 - q)) / area;
  const double w1 = 1.0 - w2 - w3;
  o.p0 = md.support0(d1) * w1 + md.support0(d2) * w2 + md.support0(d3) * w3;
  o.p1 = md.support1(-d1) * w1 + md.support1(-d2) * w2 + md.support1(-d3) * w3;
}

}  // namespace

IncrementalStatus mprIncremental(const MinkDiff& md, const V3& init, int max_iter, double tol, IncrementalOut& o) {
  V3 d = init;
  V3 v1, v2, v3, d1, d2, d3;
  V3 prev_dir = init;
  double prev_lb = -1, prev_ub = -1;
  for (int outer = 0; outer < max_iter; ++outer) {
    o.iterations = outer + 1;
    V3 new_d;
    double lb = 0, ub = 0;
    const Explore e = exploreDirection(md, d, v1, v2, v3, d1, d2, d3, lb, ub, new_d, outer >= 1, max_iter, tol);
    switch (e) {
      case Explore::NewDirection:
        d = new_d;
        break;
      case Explore::Converged:
        incrementalResult(md, d, v1, v2, v3, d1, d2, d3, lb, ub, o);
        return IncrementalStatus::OK;
      case Explore::Degenerated:
        o.depth = ub;
        o.direction = new_d;
        o.p0 = md.support0(new_d);
        o.p1 = md.support1(-new_d);
        return IncrementalStatus::OK;
      case Explore::FailedNoIntersect:
        return IncrementalStatus::FailedNoIntersect;
      default:
        return IncrementalStatus::Failed;
    }
    prev_dir = new_d;
    prev_lb = lb;
    prev_ub = ub;
  }
  (void)prev_dir;
  incrementalResult(md, d, v1, v2, v3, d1, d2, d3, prev_lb, prev_ub, o);
  return IncrementalStatus::IterationLimit;
}

}  // namespace oracle
