// CPU oracle -- TEST INFRASTRUCTURE ONLY (see omath.h).
// Support mapping of the primitive shapes and of convex hulls.
#include <map>
#include <set>
#include <utility>

#include "narrowphase.h"

namespace oracle {

// include/fcl/geometry/shape/convex-inl.h:377-409 (adjacency from faces, sorted
// neighbour ids), :301-360 (watertight check: every edge used by exactly two
// faces), :50-72 (interior point = vertex mean; walking only above 32 vertices).
void buildHull(Hull& h, const float* xyz, int n_verts, const int* tris, int n_tris) {
  h.verts.resize(n_verts);
  V3 sum(0, 0, 0);
  for (int i = 0; i < n_verts; ++i) {
    h.verts[i] = V3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    sum += h.verts[i];
  }
  h.interior = sum * (1.0 / n_verts);
  std::vector<std::set<int>> nb(n_verts);
  std::map<std::pair<int, int>, int> edge_faces;
  for (int f = 0; f < n_tris; ++f) {
    int prev = tris[3 * f + 2];
    for (int k = 0; k < 3; ++k) {
      const int v = tris[3 * f + k];
      nb[v].insert(prev);
      nb[prev].insert(v);
      ++edge_faces[{std::min(v, prev), std::max(v, prev)}];
      prev = v;
    }
  }
  h.nbr.assign(n_verts, 0);
  bool connected = true;
  for (int v = 0; v < n_verts; ++v) {
    h.nbr[v] = static_cast<int>(h.nbr.size());
    h.nbr.push_back(static_cast<int>(nb[v].size()));
    h.nbr.insert(h.nbr.end(), nb[v].begin(), nb[v].end());
    if (nb[v].empty()) connected = false;
  }
  bool watertight = n_tris > 0;
  for (const auto& kv : edge_faces)
    if (kv.second != 2) watertight = false;
  h.walk = (n_verts > 32) && watertight && connected;
  const V3 axes[6] = {V3(1, 0, 0), V3(-1, 0, 0), V3(0, 1, 0), V3(0, -1, 0), V3(0, 0, 1), V3(0, 0, -1)};
  for (int i = 0; i < 6; ++i) {
    h.cache_dir[i] = axes[i];
    h.cache_idx[i] = hullExtremeNaive(h, axes[i]);
  }
}

// convex-inl.h:133-149: first index attaining the strict maximum.
int hullExtremeNaive(const Hull& h, const V3& d) {
  int best = 0;
  double best_v = dot(d, h.verts[0]);
  for (int i = 1; i < static_cast<int>(h.verts.size()); ++i) {
    const double v = dot(d, h.verts[i]);
    if (v > best_v) {
      best = i;
      best_v = v;
    }
  }
  return best;
}

// convex-inl.h:202-290: seed from the best of six axis-direction caches, then
// greedy walk over neighbours, skipping the two most recent parents.
int hullExtremeWalk(const Hull& h, const V3& d) {
  int init = -1;
  double max_dot = 0;
  for (int i = 0; i < 6; ++i) {
    const double dv = dot(h.cache_dir[i], d);
    if (init < 0 || dv > max_dot) {
      init = h.cache_idx[i];
      max_dot = dv;
    }
  }
  int ext = init;
  double ext_v = dot(d, h.verts[ext]);
  int parent0 = init, parent1 = init;
  bool keep = true;
  while (keep) {
    keep = false;
    const int start = h.nbr[ext];
    const int count = h.nbr[start];
    const int old = ext;
    for (int k = start + 1; k <= start + count; ++k) {
      const int n = h.nbr[k];
      if (n == parent0 || n == parent1) continue;
      const double nv = dot(d, h.verts[n]);
      if (nv > ext_v) {
        parent1 = ext;
        keep = true;
        ext = n;
        ext_v = nv;
      }
    }
    parent0 = old;
  }
  return ext;
}

// include/fcl/cvx_collide/gjk_shape.hpp:41-180 and
// include/fcl/geometry/shape/shape_gjk_interface-inl.h:87-92.
V3 shapeSupport(const Shape& s, const V3& dir) {
  switch (s.type) {
    case kBox:
      return V3(dir.x > 0 ? s.p[0] / 2 : -s.p[0] / 2, dir.y > 0 ? s.p[1] / 2 : -s.p[1] / 2,
                dir.z > 0 ? s.p[2] / 2 : -s.p[2] / 2);
    case kSphere:
      return dir * s.p[0];
    case kCapsule: {
      const double r = s.p[0];
      const double half_h = s.p[1] * 0.5;
      const V3 v = dir * r;
      const V3 pos1 = V3(0, 0, half_h) + v;
      const V3 pos2 = V3(0, 0, -half_h) + v;
      return dot(dir, pos1) > dot(dir, pos2) ? pos1 : pos2;
    }
    case kCylinder: {
      const double r = s.p[0];
      const double half_h = s.p[1] * 0.5;
      const double zdist = std::sqrt(dir.x * dir.x + dir.y * dir.y);
      if (zdist == 0.0) return V3(0, 0, dir.z > 0 ? half_h : -half_h);
      const double k = r / zdist;
      return V3(k * dir.x, k * dir.y, dir.z > 0 ? half_h : -half_h);
    }
    case kConvex: {
      const Hull& h = *s.hull;
      return h.verts[h.walk ? hullExtremeWalk(h, dir) : hullExtremeNaive(h, dir)];
    }
    case kTriangle: {
      // triangleSupport, shape_gjk_interface-inl.h:17-38 (its tie rule)
      const double da = dot(dir, s.tri[0]), db = dot(dir, s.tri[1]), dc = dot(dir, s.tri[2]);
      if (da > db) return dc > da ? s.tri[2] : s.tri[0];
      return dc > db ? s.tri[2] : s.tri[1];
    }
    default:
      return V3(0, 0, 0);
  }
}

// gjk_shape.hpp:28-32 (origin), gjk_solver_cvx-inl.h:117-125 (hull mean),
// shape_gjk_interface-inl.h:8-15 (triangle centroid).
V3 shapeInterior(const Shape& s) {
  if (s.type == kConvex) return s.hull->interior;
  if (s.type == kTriangle) return (s.tri[0] + s.tri[1] + s.tri[2]) / 3.0;
  return V3(0, 0, 0);
}

// gjk_solver-inl.h:79-85: toshape1 = R2^T R1, toshape0 = tf1^-1 * tf2.
MinkDiff makeMinkDiff(const Shape& a, const Tf& tfa, const Shape& b, const Tf& tfb) {
  MinkDiff md;
  md.s0 = a;
  md.s1 = b;
  md.toshape1 = transpose(tfb.R) * tfa.R;
  md.toshape0 = inverse(tfa) * tfb;
  return md;
}

}  // namespace oracle
