// CPU oracle -- TEST INFRASTRUCTURE ONLY (see omath.h).
// GJK boolean (include/fcl/cvx_collide/gjk.hpp:11-363) and GJK minimum
// distance (include/fcl/cvx_collide/gjk_distance.hpp:11-559).
#include "narrowphase.h"

namespace oracle {
namespace {

enum class Proj { Failed, Continue, Intersect, FailedZeroVolume };
enum class DistUpd { NoImprovement, OK, Failed };
constexpr double kBaryTol = 1e-3;  // gjk.h:125

struct Ctx {
  const MinkDiff& md;
  double tol;
  int max_iter;
  GJKStats* st;
  MVert sv(const V3& d) const {
    if (st) st->support_calls++;
    return MVert{md.support(d), d};
  }
};

// gjk.hpp:181-219
Proj proj2(Simplex& s, V3& dir) {
  const MVert B = s.v[0];
  const MVert A = s.v[1];
  const V3& a = A.v;
  const V3 a_to_b = B.v - a;
  const double ao_dot_ab = -dot(a_to_b, a);
  const V3 ab_cross_ao = cross(a, a_to_b);
  if (ao_dot_ab > 0 && sqnorm(ab_cross_ao) <= 0.0) return Proj::Intersect;
  if (ao_dot_ab <= 0) {
    s.reset();
    s.add(A);
    dir = -normalized(a);
    return Proj::Continue;
  }
  dir = cross(ab_cross_ao, a_to_b);
  normalize(dir);
  return Proj::Continue;
}

// gjk.hpp:221-294
Proj proj3(Simplex& s, V3& dir, double tol) {
  const MVert C = s.v[0];
  const MVert B = s.v[1];
  const MVert A = s.v[2];
  const V3& a = A.v;
  const V3 a_to_b = B.v - a;
  const V3 a_to_c = C.v - a;
  const bool sep_b = dot(a, a_to_b) >= 0;
  const bool sep_c = dot(a, a_to_c) >= 0;
  if (sep_b && sep_c) {
    s.reset();
    s.add(A);
    dir = -a;
    normalize(dir);
    return Proj::Continue;
  }
  const V3 n = cross(a_to_b, a_to_c);
  const V3 ac_n = cross(n, a_to_c);
  if (dot(a, ac_n) <= 0) {
    s.v[1] = A;
    s.rank = 2;
    dir = cross(cross(a, a_to_c), a_to_c);
    normalize(dir);
    return Proj::Continue;
  }
  const V3 ab_n = cross(a_to_b, n);
  if (dot(a, ab_n) <= 0) {
    s.v[0] = B;
    s.v[1] = A;
    s.rank = 2;
    dir = cross(cross(a, a_to_b), a_to_b);
    normalize(dir);
    return Proj::Continue;
  }
  const double area = norm(n);
  if (area < tol * tol) return Proj::Failed;
  const V3 nu = n / area;
  const double nd = dot(nu, a);
  if (std::fabs(nd) < tol) return Proj::Intersect;
  dir = nd <= 0 ? nu : -nu;
  return Proj::Continue;
}

// gjk.hpp:296-363
Proj proj4(Simplex& s, V3& dir, double tol) {
  const MVert D = s.v[0];
  const MVert C = s.v[1];
  const MVert B = s.v[2];
  const MVert A = s.v[3];
  const V3& a = A.v;
  const V3 ab = B.v - a, ac = C.v - a, ad = D.v - a;
  V3 abc = cross(ab, ac), acd = cross(ac, ad), abd = cross(ab, ad);
  const double abc_ad = dot(abc, ad);
  const double acd_ab = dot(acd, ab);
  const double abd_ac = dot(abd, ac);
  if (std::fabs(abc_ad) <= 0.0) return Proj::FailedZeroVolume;
  if (abc_ad > 0) abc = -abc;
  if (acd_ab > 0) acd = -acd;
  if (abd_ac > 0) abd = -abd;
  const bool d_same = dot(a, abc) > 0;
  const bool c_same = dot(a, abd) > 0;
  const bool b_same = dot(a, acd) > 0;
  if (d_same && c_same && b_same) return Proj::Intersect;
  if (!b_same) {
    s.v[2] = A;
    s.rank = 3;
  } else if (!c_same) {
    s.v[1] = B;
    s.v[2] = A;
    s.rank = 3;
  } else {
    s.v[0] = C;
    s.v[1] = B;
    s.v[2] = A;
    s.rank = 3;
  }
  return proj3(s, dir, tol);
}

Proj project(Simplex& s, V3& dir, double tol) {
  if (s.rank == 2) return proj2(s, dir);
  if (s.rank == 3) return proj3(s, dir, tol);
  return proj4(s, dir, tol);
}

// gjk_distance.hpp:128-154
void dist2(Simplex& s, V3& out, double tol) {
  const V3& s1 = s.v[1].v;
  const V3& s2 = s.v[0].v;
  const V3 d = s2 - s1;
  const double l2 = sqnorm(d);
  const double t = -dot(s1, d);
  if (t <= 0 || l2 <= tol * tol) {
    out = s1;
    s.v[0] = s.v[1];
    s.rank = 1;
  } else if (t >= l2) {
    out = s2;
    s.rank = 1;
  } else {
    const double w2 = t / l2;
    out = w2 * s2 + (1.0 - w2) * s1;
  }
}

// gjk_distance.hpp:156-289
void dist3(Simplex& s, V3& out, double tol) {
  const V3& s1 = s.v[2].v;
  const V3& s2 = s.v[1].v;
  const V3& s3 = s.v[0].v;
  const V3 s12 = s2 - s1, s13 = s3 - s1;
  const bool sep2 = dot(s1, s12) >= 0;
  const bool sep3 = dot(s1, s13) >= 0;
  if (sep2 && sep3) {
    out = s1;
    s.v[0] = s.v[2];
    s.rank = 1;
    return;
  }
  const V3 n = cross(s12, s13);
  const double area2 = sqnorm(n);
  const bool zero_area = area2 <= 0.0;
  const V3 n12 = cross(n, s12);
  const bool s12_sep = dot(s1, n12) > 0;
  if (!sep2 && s12_sep) {
    s.v[0] = s.v[1];
    s.v[1] = s.v[2];
    s.rank = 2;
    dist2(s, out, tol);
    return;
  }
  const V3 n13 = cross(n, s13);
  const bool s13_sep = dot(s1, n13) < 0;
  if (!sep3 && s13_sep) {
    s.v[1] = s.v[2];
    s.rank = 2;
    dist2(s, out, tol);
    return;
  }
  double best = -1;
  V3 best_pt, pt;
  Simplex best_s, c;
  if (zero_area || s12_sep) {
    c = s;
    c.v[0] = c.v[1];
    c.v[1] = c.v[2];
    c.rank = 2;
    dist2(c, pt, tol);
    const double d = sqnorm(pt);
    if (best < 0 || d < best) { best = d; best_s = c; best_pt = pt; }
  }
  if (zero_area || s13_sep) {
    c = s;
    c.v[1] = c.v[2];
    c.rank = 2;
    dist2(c, pt, tol);
    const double d = sqnorm(pt);
    if (best < 0 || d < best) { best = d; best_s = c; best_pt = pt; }
  }
  const V3 n23 = cross(n, s3 - s2);
  const bool s23_sep = dot(s2, n23) > 0;
  if (zero_area || s23_sep) {
    c = s;
    c.rank = 2;
    dist2(c, pt, tol);
    const double d = sqnorm(pt);
    if (best < 0 || d < best) { best = d; best_s = c; best_pt = pt; }
  }
  if (best < 0) {
    const double dd = dot(s1, n);
    out = n * (dd / area2);
  } else {
    s = best_s;
    out = best_pt;
  }
}

// gjk_distance.hpp:291-373
DistUpd dist4(Simplex& s, V3& out, double tol) {
  double best = -1;
  V3 best_pt, pt;
  Simplex best_s, c;
  c = s;  // keep [3,2,1]
  c.v[0] = c.v[1]; c.v[1] = c.v[2]; c.v[2] = c.v[3]; c.rank = 3;
  dist3(c, pt, tol);
  { const double d = sqnorm(pt); if (best < 0 || d < best) { best = d; best_pt = pt; best_s = c; } }
  c = s;  // keep [3,2,0]
  c.v[1] = c.v[2]; c.v[2] = c.v[3]; c.rank = 3;
  dist3(c, pt, tol);
  { const double d = sqnorm(pt); if (best < 0 || d < best) { best = d; best_pt = pt; best_s = c; } }
  c = s;  // keep [3,1,0]
  c.v[2] = c.v[3]; c.rank = 3;
  dist3(c, pt, tol);
  { const double d = sqnorm(pt); if (best < 0 || d < best) { best = d; best_pt = pt; best_s = c; } }
  if (best < 0) return DistUpd::NoImprovement;
  s = best_s;
  out = best_pt;
  return DistUpd::OK;
}

DistUpd distUpdate(Simplex& s, V3& out, double tol) {
  if (s.rank == 1) { out = s.v[0].v; return DistUpd::OK; }
  if (s.rank == 2) { dist2(s, out, tol); return DistUpd::OK; }
  if (s.rank == 3) { dist3(s, out, tol); return DistUpd::OK; }
  if (s.rank == 4) return dist4(s, out, tol);
  return DistUpd::Failed;
}

// gjk_distance.hpp:375-486
bool extractNoSub(const MinkDiff& md, const Simplex& s, V3& p0, V3& p1) {
  if (s.rank == 4 || s.rank <= 0) return false;
  if (s.rank == 1) {
    p0 = md.support0(s.v[0].d);
    p1 = md.support1(-s.v[0].d);
    return true;
  }
  if (s.rank == 2) {
    const MVert& v1 = s.v[0];
    const MVert& v2 = s.v[1];
    const V3 d = v2.v - v1.v;
    const double l2 = sqnorm(d);
    if (l2 <= 0.0) {
      p0 = md.support0(v1.d);
      p1 = md.support1(-v1.d);
      return true;
    }
    const double w2 = -dot(v1.v, d) / l2;
    if (w2 > 1 + kBaryTol) {
      p0 = md.support0(v2.d);
      p1 = md.support1(-v2.d);
      return false;
    } else if (w2 < -kBaryTol) {
      p0 = md.support0(v1.d);
      p1 = md.support1(-v1.d);
      return false;
    }
    const double w1 = 1.0 - w2;
    p0 = md.support0(v1.d) * w1 + md.support0(v2.d) * w2;
    p1 = md.support1(-v1.d) * w1 + md.support1(-v2.d) * w2;
    return true;
  }
  const MVert& v1 = s.v[0];
  const MVert& v2 = s.v[1];
  const MVert& v3 = s.v[2];
  const V3 s12 = v2.v - v1.v, s13 = v3.v - v1.v;
  const V3 n = cross(s12, s13);
  const double area2 = sqnorm(n);
  if (area2 <= 0.0) return false;
  const double dd = dot(v1.v, n);
  const V3 proj = n * (dd / area2);
  const double area = std::sqrt(area2);
  const double w2 = norm(cross(s13, v1.v - proj)) / area;
  const double w3 = norm(cross(s12, v1.v - proj)) / area;
  const double w1 = 1.0 - w2 - w3;
  if (w1 < -kBaryTol || w2 < -kBaryTol || w3 < -kBaryTol) return false;
  p0 = md.support0(v1.d) * w1 + md.support0(v2.d) * w2 + md.support0(v3.d) * w3;
  p1 = md.support1(-v1.d) * w1 + md.support1(-v2.d) * w2 + md.support1(-v3.d) * w3;
  return true;
}

// gjk_distance.hpp:488-559
bool extractTrySub(const MinkDiff& md, const Simplex& s, V3& p0, V3& p1) {
  if (s.rank >= 4) return false;
  if (s.rank <= 2) {
    extractNoSub(md, s, p0, p1);
    return true;
  }
  double best = -1;
  V3 bp0, bp1, q0, q1;
  Simplex c = s;
  if (extractNoSub(md, c, q0, q1)) {
    const double d = sqnorm(q0 - q1);
    if (best < 0 || d < best) { best = d; bp0 = q0; bp1 = q1; }
  }
  c = s; c.rank = 2;
  extractNoSub(md, c, q0, q1);
  { const double d = sqnorm(q0 - q1); if (best < 0 || d < best) { best = d; bp0 = q0; bp1 = q1; } }
  c = s; c.v[1] = c.v[2]; c.rank = 2;
  extractNoSub(md, c, q0, q1);
  { const double d = sqnorm(q0 - q1); if (best < 0 || d < best) { best = d; bp0 = q0; bp1 = q1; } }
  c = s; c.v[0] = c.v[1]; c.v[1] = c.v[2]; c.rank = 2;
  extractNoSub(md, c, q0, q1);
  { const double d = sqnorm(q0 - q1); if (best < 0 || d < best) { best = d; bp0 = q0; bp1 = q1; } }
  p0 = bp0;
  p1 = bp1;
  return true;
}

// gjk_distance.hpp:11-105
bool minDistance(const Ctx& cx, Simplex& s, V3& p0, V3& p1) {
  if (s.rank != 1) return false;
  V3 cur;
  if (distUpdate(s, cur, cx.tol) != DistUpd::OK) return false;
  double book = norm(cur);
  if (book <= cx.tol) return extractNoSub(cx.md, s, p0, p1);
  V3 nd = -cur / book;
  const double tol2 = cx.tol * cx.tol;
  int it = 0;
  while (it < cx.max_iter) {
    ++it;
    if (cx.st) cx.st->dist_iterations++;
    const MVert nv = cx.sv(nd);
    if (dot(nd, nv.v - cur) < cx.tol) return extractNoSub(cx.md, s, p0, p1);
    for (int j = 0; j < s.rank; ++j)
      if (sqnorm(s.v[j].v - nv.v) < tol2) return extractNoSub(cx.md, s, p0, p1);
    s.add(nv);
    const DistUpd st = distUpdate(s, cur, cx.tol);
    if (st == DistUpd::NoImprovement) return extractNoSub(cx.md, s, p0, p1);
    if (st != DistUpd::OK) return false;
    const double nm = norm(cur);
    if (book - nm < cx.tol) return extractNoSub(cx.md, s, p0, p1);
    if (nm < cx.tol) return extractNoSub(cx.md, s, p0, p1);
    book = nm;
    nd = -cur / book;
  }
  return false;
}

}  // namespace

// gjk.hpp:11-140
GJKStatus gjkEvaluate(const MinkDiff& md, Simplex& simplex, const V3& guess, double tol,
                      int max_iter, GJKDistOut* out, GJKStats* st) {
  Ctx cx{md, tol, max_iter, st};
  V3 dir = guess;
  if (sqnorm(dir) <= 0.0) dir = V3(1, 0, 0);
  normalize(dir);

  auto separated = [&](const MVert& sv) {
    if (out == nullptr) return GJKStatus::Separated;
    out->cert_valid = true;
    simplex.reset();
    simplex.add(sv);
    V3 p0, p1;
    const bool ok = minDistance(cx, simplex, p0, p1);
    out->sep_valid = ok;
    if (ok) { out->p0 = p0; out->p1 = p1; }
    return GJKStatus::Separated;
  };
  auto undetermined = [&](GJKStatus r) {
    if (out != nullptr) {
      out->cert_valid = false;
      V3 p0, p1;
      const bool ok = extractTrySub(md, simplex, p0, p1);
      out->sep_valid = ok;
      if (ok) { out->p0 = p0; out->p1 = p1; }
    }
    return r;
  };

  MVert v = cx.sv(dir);
  simplex.reset();
  simplex.add(v);
  const double tol2 = tol * tol;
  if (sqnorm(v.v) <= tol2) return GJKStatus::Intersect;
  if (dot(v.v, dir) < 0) return separated(v);
  dir = -dir;
  int it = 0;
  while (it < max_iter) {
    ++it;
    if (st) st->iterations = it;
    v = cx.sv(dir);
    if (dot(v.v, dir) < 0) return separated(v);
    for (int j = 0; j < simplex.rank; ++j)
      if (sqnorm(simplex.v[j].v - v.v) < tol2) return undetermined(GJKStatus::ConvergeNoProgress);
    if (sqnorm(v.v) <= tol2) {
      simplex.add(v);
      return GJKStatus::Intersect;
    }
    simplex.add(v);
    const Proj p = project(simplex, dir, tol);
    if (p == Proj::Failed) return GJKStatus::Failed;
    if (p == Proj::Intersect) return GJKStatus::Intersect;
    if (p == Proj::FailedZeroVolume) return undetermined(GJKStatus::ConvergeNoProgress);
  }
  return undetermined(GJKStatus::IterationLimit);
}

}  // namespace oracle
