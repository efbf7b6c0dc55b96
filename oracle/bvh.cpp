// CPU oracle -- TEST INFRASTRUCTURE ONLY (see omath.h).
// OBBRSS mesh-mesh collision and broadphase AABB tree (file map in bvh.h).
#include "bvh.h"

#include <algorithm>
#include <limits>
#include <utility>

namespace oracle {
namespace {

// include/fcl/math/geometry-inl.h:237-337 (cyclic Jacobi, at most 50 sweeps).
void eigenJacobi(const M3& m, V3& dout, M3& vout) {
  M3 R = m;
  const int n = 3;
  double b[3], z[3], d[3];
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int ip = 0; ip < n; ++ip) {
    b[ip] = d[ip] = R.m[ip][ip];
    z[ip] = 0;
  }
  for (int i = 0; i < 50; ++i) {
    double sm = 0;
    for (int ip = 0; ip < n; ++ip)
      for (int iq = ip + 1; iq < n; ++iq) sm += std::fabs(R.m[ip][iq]);
    if (sm == 0.0) {
      for (int c = 0; c < 3; ++c) vout.setCol(c, V3(v[c][0], v[c][1], v[c][2]));
      dout = V3(d[0], d[1], d[2]);
      return;
    }
    const double tresh = i < 3 ? 0.2 * sm / (n * n) : 0.0;
    for (int ip = 0; ip < n; ++ip) {
      for (int iq = ip + 1; iq < n; ++iq) {
        const double g = 100.0 * std::fabs(R.m[ip][iq]);
        if (i > 3 && std::fabs(d[ip]) + g == std::fabs(d[ip]) && std::fabs(d[iq]) + g == std::fabs(d[iq])) {
          R.m[ip][iq] = 0.0;
        } else if (std::fabs(R.m[ip][iq]) > tresh) {
          double h = d[iq] - d[ip];
          double t;
          if (std::fabs(h) + g == std::fabs(h)) {
            t = R.m[ip][iq] / h;
          } else {
            const double theta = 0.5 * h / R.m[ip][iq];
            t = 1.0 / (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
            if (theta < 0.0) t = -t;
          }
          const double c = 1.0 / std::sqrt(1 + t * t);
          const double s = t * c;
          const double tau = s / (1.0 + c);
          h = t * R.m[ip][iq];
          z[ip] -= h;
          z[iq] += h;
          d[ip] -= h;
          d[iq] += h;
          R.m[ip][iq] = 0.0;
          for (int j = 0; j < ip; ++j) {
            const double gg = R.m[j][ip], hh = R.m[j][iq];
            R.m[j][ip] = gg - s * (hh + gg * tau);
            R.m[j][iq] = hh + s * (gg - hh * tau);
          }
          for (int j = ip + 1; j < iq; ++j) {
            const double gg = R.m[ip][j], hh = R.m[j][iq];
            R.m[ip][j] = gg - s * (hh + gg * tau);
            R.m[j][iq] = hh + s * (gg - hh * tau);
          }
          for (int j = iq + 1; j < n; ++j) {
            const double gg = R.m[ip][j], hh = R.m[iq][j];
            R.m[ip][j] = gg - s * (hh + gg * tau);
            R.m[iq][j] = hh + s * (gg - hh * tau);
          }
          for (int j = 0; j < n; ++j) {
            const double gg = v[j][ip], hh = v[j][iq];
            v[j][ip] = gg - s * (hh + gg * tau);
            v[j][iq] = hh + s * (gg - hh * tau);
          }
        }
      }
    }
    for (int ip = 0; ip < n; ++ip) {
      b[ip] += z[ip];
      d[ip] = b[ip];
      z[ip] = 0.0;
    }
  }
}

// geometry-inl.h:340-366: columns = eigenvectors of the largest and middle
// eigenvalue (taken as rows of the Jacobi output), third = their cross product.
M3 axisFromEigen(const M3& V, const V3& s) {
  int mn, md, mx;
  if (s[0] > s[1]) { mx = 0; mn = 1; } else { mn = 0; mx = 1; }
  if (s[2] < s[mn]) { md = mn; mn = 2; }
  else if (s[2] > s[mx]) { md = mx; mx = 2; }
  else { md = 2; }
  M3 a;
  a.setCol(0, V.row(mx));
  a.setCol(1, V.row(md));
  a.setCol(2, cross(a.col(0), a.col(1)));
  return a;
}

// BV_fitter-inl.h:324-346 with geometry-inl.h:713-780 (covariance) and
// :148-180 (extent in the fitted frame).  Only the OBB part of OBBRSS is kept:
// collision overlap uses the OBB alone (OBBRSS-inl.h:131-137).
OBB fitOBB(const Mesh& m, const unsigned* idx, int n) {
  V3 S1(0, 0, 0);
  double S2[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  int np = 0;
  for (int i = 0; i < n; ++i) {
    const int* t = &m.tris[3 * idx[i]];
    const V3& p1 = m.verts[t[0]];
    const V3& p2 = m.verts[t[1]];
    const V3& p3 = m.verts[t[2]];
    S1 += p1 + p2 + p3;
    S2[0][0] += p1.x * p1.x + p2.x * p2.x + p3.x * p3.x;
    S2[1][1] += p1.y * p1.y + p2.y * p2.y + p3.y * p3.y;
    S2[2][2] += p1.z * p1.z + p2.z * p2.z + p3.z * p3.z;
    S2[0][1] += p1.x * p1.y + p2.x * p2.y + p3.x * p3.y;
    S2[0][2] += p1.x * p1.z + p2.x * p2.z + p3.x * p3.z;
    S2[1][2] += p1.y * p1.z + p2.y * p2.z + p3.y * p3.z;
    np += 3;
  }
  M3 M;
  M.m[0][0] = S2[0][0] - S1.x * S1.x / np;
  M.m[1][1] = S2[1][1] - S1.y * S1.y / np;
  M.m[2][2] = S2[2][2] - S1.z * S1.z / np;
  M.m[0][1] = S2[0][1] - S1.x * S1.y / np;
  M.m[1][2] = S2[1][2] - S1.y * S1.z / np;
  M.m[0][2] = S2[0][2] - S1.x * S1.z / np;
  M.m[1][0] = M.m[0][1];
  M.m[2][0] = M.m[0][2];
  M.m[2][1] = M.m[1][2];
  V3 s;
  M3 E;
  eigenJacobi(M, s, E);
  OBB bv;
  bv.axis = axisFromEigen(E, s);
  const double big = std::numeric_limits<double>::max();
  V3 mnc(big, big, big), mxc(-big, -big, -big);
  for (int i = 0; i < n; ++i) {
    const int* t = &m.tris[3 * idx[i]];
    for (int j = 0; j < 3; ++j) {
      const V3& p = m.verts[t[j]];
      const V3 pr(dot(bv.axis.col(0), p), dot(bv.axis.col(1), p), dot(bv.axis.col(2), p));
      for (int k = 0; k < 3; ++k) {
        if (pr[k] > mxc[k]) mxc[k] = pr[k];
        if (pr[k] < mnc[k]) mnc[k] = pr[k];
      }
    }
  }
  const V3 o = (mxc + mnc) / 2;
  bv.To = bv.axis * o;
  bv.extent = (mxc - mnc) / 2;
  return bv;
}

}  // namespace

// BVH_model-inl.h:449-568: explicit stack, children allocated adjacently,
// left pushed last so it is built first; mean split along the major axis
// (BV_splitter-inl.h:411-500).
void buildMeshBVH(Mesh& m, const float* xyz, int n_verts, const int* tris, int n_tris) {
  m.verts.resize(n_verts);
  for (int i = 0; i < n_verts; ++i) m.verts[i] = V3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  m.tris.assign(tris, tris + 3 * n_tris);
  std::vector<unsigned> prim(n_tris);
  for (int i = 0; i < n_tris; ++i) prim[i] = i;
  m.nodes.clear();
  m.nodes.reserve(2 * n_tris - 1);
  struct Task { int node, first, num; };
  std::vector<Task> stk;
  m.nodes.emplace_back();
  stk.push_back({0, 0, n_tris});
  while (!stk.empty()) {
    const Task t = stk.back();
    stk.pop_back();
    m.nodes[t.node].first_primitive = t.first;
    m.nodes[t.node].num_primitives = t.num;
    unsigned* cur = prim.data() + t.first;
    m.nodes[t.node].obb = fitOBB(m, cur, t.num);
    const OBB& bv = m.nodes[t.node].obb;
    const V3 split_vec = bv.axis.col(0);
    double c[3] = {0, 0, 0};
    for (int i = 0; i < t.num; ++i) {
      const int* tr = &m.tris[3 * cur[i]];
      const V3& p1 = m.verts[tr[0]];
      const V3& p2 = m.verts[tr[1]];
      const V3& p3 = m.verts[tr[2]];
      c[0] += (p1.x + p2.x + p3.x) / 3;
      c[1] += (p1.y + p2.y + p3.y) / 3;
      c[2] += (p1.z + p2.z + p3.z) / 3;
    }
    const double split_val = (c[0] * split_vec.x + c[1] * split_vec.y + c[2] * split_vec.z) / t.num;
    if (t.num == 1) {
      m.nodes[t.node].first_child = -(static_cast<int>(cur[0]) + 1);
      continue;
    }
    int c1 = 0;
    for (int i = 0; i < t.num; ++i) {
      const int* tr = &m.tris[3 * cur[i]];
      const V3 p = (m.verts[tr[0]] + m.verts[tr[1]] + m.verts[tr[2]]) / 3.0;
      if (!(dot(split_vec, p) > split_val)) {
        std::swap(cur[i], cur[c1]);
        c1++;
      }
    }
    if (c1 == 0 || c1 == t.num) c1 = t.num / 2;
    const int left = static_cast<int>(m.nodes.size());
    m.nodes.emplace_back();
    m.nodes[t.node].first_child = left;
    const int right = static_cast<int>(m.nodes.size());
    m.nodes.emplace_back();
    stk.push_back({right, t.first + c1, t.num - c1});
    stk.push_back({left, t.first, c1});
  }
}

// OBB-inl.h:320-436: 15-axis SAT, |B| padded by 1e-6.
bool obbDisjoint(const M3& B, const V3& T, const V3& a, const V3& b) {
  const double reps = 1e-6;
  M3 Bf;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Bf.m[i][j] = std::fabs(B.m[i][j]) + reps;
  const auto& Bm = B.m;
  const auto& F = Bf.m;
  double t, s;
  t = std::fabs(T[0]);
  if (t > (a[0] + dot(Bf.row(0), b))) return true;
  s = dot(B.col(0), T);
  t = std::fabs(s);
  if (t > (b[0] + dot(Bf.col(0), a))) return true;
  t = std::fabs(T[1]);
  if (t > (a[1] + dot(Bf.row(1), b))) return true;
  t = std::fabs(T[2]);
  if (t > (a[2] + dot(Bf.row(2), b))) return true;
  s = dot(B.col(1), T);
  t = std::fabs(s);
  if (t > (b[1] + dot(Bf.col(1), a))) return true;
  s = dot(B.col(2), T);
  t = std::fabs(s);
  if (t > (b[2] + dot(Bf.col(2), a))) return true;
  s = T[2] * Bm[1][0] - T[1] * Bm[2][0]; t = std::fabs(s);
  if (t > (a[1] * F[2][0] + a[2] * F[1][0] + b[1] * F[0][2] + b[2] * F[0][1])) return true;
  s = T[2] * Bm[1][1] - T[1] * Bm[2][1]; t = std::fabs(s);
  if (t > (a[1] * F[2][1] + a[2] * F[1][1] + b[0] * F[0][2] + b[2] * F[0][0])) return true;
  s = T[2] * Bm[1][2] - T[1] * Bm[2][2]; t = std::fabs(s);
  if (t > (a[1] * F[2][2] + a[2] * F[1][2] + b[0] * F[0][1] + b[1] * F[0][0])) return true;
  s = T[0] * Bm[2][0] - T[2] * Bm[0][0]; t = std::fabs(s);
  if (t > (a[0] * F[2][0] + a[2] * F[0][0] + b[1] * F[1][2] + b[2] * F[1][1])) return true;
  s = T[0] * Bm[2][1] - T[2] * Bm[0][1]; t = std::fabs(s);
  if (t > (a[0] * F[2][1] + a[2] * F[0][1] + b[0] * F[1][2] + b[2] * F[1][0])) return true;
  s = T[0] * Bm[2][2] - T[2] * Bm[0][2]; t = std::fabs(s);
  if (t > (a[0] * F[2][2] + a[2] * F[0][2] + b[0] * F[1][1] + b[1] * F[1][0])) return true;
  s = T[1] * Bm[0][0] - T[0] * Bm[1][0]; t = std::fabs(s);
  if (t > (a[0] * F[1][0] + a[1] * F[0][0] + b[1] * F[2][2] + b[2] * F[2][1])) return true;
  s = T[1] * Bm[0][1] - T[0] * Bm[1][1]; t = std::fabs(s);
  if (t > (a[0] * F[1][1] + a[1] * F[0][1] + b[0] * F[2][2] + b[2] * F[2][0])) return true;
  s = T[1] * Bm[0][2] - T[0] * Bm[1][2]; t = std::fabs(s);
  if (t > (a[0] * F[1][2] + a[1] * F[0][2] + b[0] * F[2][1] + b[1] * F[2][0])) return true;
  return false;
}

// OBB-inl.h:306-316
bool obbOverlap(const M3& R0, const V3& T0, const OBB& b1, const OBB& b2) {
  const M3 R0b2 = R0 * b2.axis;
  const M3 R = transpose(b1.axis) * R0b2;
  const V3 Tt = R0 * b2.To + T0 - b1.To;
  const V3 T = tmul(b1.axis, Tt);
  return !obbDisjoint(R, T, b1.extent, b2.extent);
}

namespace {
// intersect-inl.h:1082-1103
bool project6(const V3& ax, const V3& p1, const V3& p2, const V3& p3, const V3& q1, const V3& q2,
              const V3& q3) {
  const double P1 = dot(ax, p1), P2 = dot(ax, p2), P3 = dot(ax, p3);
  const double Q1 = dot(ax, q1), Q2 = dot(ax, q2), Q3 = dot(ax, q3);
  const double mn1 = std::min(P1, std::min(P2, P3));
  const double mx2 = std::max(Q1, std::max(Q2, Q3));
  if (mn1 > mx2) return false;
  const double mx1 = std::max(P1, std::max(P2, P3));
  const double mn2 = std::min(Q1, std::min(Q2, Q3));
  if (mn2 > mx1) return false;
  return true;
}
}  // namespace

// intersect-inl.h:594-613 (Q mapped into P's frame) and :724-790 (17 axes).
bool triTriIntersect(const V3& P1, const V3& P2, const V3& P3, const V3& Q1in, const V3& Q2in,
                     const V3& Q3in, const M3& R, const V3& T) {
  const V3 Q1 = R * Q1in + T, Q2 = R * Q2in + T, Q3 = R * Q3in + T;
  const V3 p1 = P1 - P1, p2 = P2 - P1, p3 = P3 - P1;
  const V3 q1 = Q1 - P1, q2 = Q2 - P1, q3 = Q3 - P1;
  const V3 e1 = p2 - p1, e2 = p3 - p2, e3 = p1 - p3;
  const V3 f1 = q2 - q1, f2 = q3 - q2, f3 = q1 - q3;
  const V3 n1 = cross(e1, e2);
  if (!project6(n1, p1, p2, p3, q1, q2, q3)) return false;
  const V3 m1 = cross(f1, f2);
  if (!project6(m1, p1, p2, p3, q1, q2, q3)) return false;
  const V3 ax[15] = {cross(e1, f1), cross(e1, f2), cross(e1, f3), cross(e2, f1), cross(e2, f2),
                     cross(e2, f3), cross(e3, f1), cross(e3, f2), cross(e3, f3), cross(e1, n1),
                     cross(e2, n1), cross(e3, n1), cross(f1, m1), cross(f2, m1), cross(f3, m1)};
  for (int i = 0; i < 15; ++i)
    if (!project6(ax[i], p1, p2, p3, q1, q2, q3)) return false;
  return true;
}

namespace {
// intersect-inl.h:1036-1051 with normalize() of geometry-inl.h:188-197:
// unit normal and offset of the triangle's plane; false (and n, t left as
// given) for a degenerate triangle.
bool buildTrianglePlane(const V3& v1, const V3& v2, const V3& v3, V3* n, double* t) {
  V3 n_ = cross(v2 - v1, v3 - v1);
  const double sq = sqnorm(n_);
  if (!(sq > 0)) return false;
  const double l = std::sqrt(sq);
  n_ = V3(n_.x / l, n_.y / l, n_.z / l);
  *n = n_;
  *t = dot(n_, v1);
  return true;
}

// intersect-inl.h:845-882 (getEpsilon() = 1e-5, distanceToPlane = n.v - t)
void computeDeepestPoints(const V3* pts, int np, const V3& n, double t, double* depth, V3* deepest, int* nd) {
  constexpr double kEps = 1e-5;
  double max_depth = -std::numeric_limits<double>::max();
  int k = 0, num_neg = 0, num_pos = 0, num_zero = 0;
  for (int i = 0; i < np; ++i) {
    const double dist = -(dot(n, pts[i]) - t);
    if (dist > kEps) num_pos++;
    else if (dist < -kEps) num_neg++;
    else num_zero++;
    if (dist > max_depth) {
      max_depth = dist;
      k = 1;
      deepest[0] = pts[i];
    } else if (dist + 1e-6 >= max_depth) {
      k++;
      deepest[k - 1] = pts[i];
    }
  }
  if (max_depth < -kEps) k = 0;
  if (num_zero == 0 && (num_neg == 0 || num_pos == 0)) k = 0;
  *depth = max_depth;
  *nd = k;
}
}  // namespace

bool triTriContacts(const V3& P1, const V3& P2, const V3& P3, const V3& Q1in, const V3& Q2in, const V3& Q3in,
                    const M3& R, const V3& T, V3 pts[2], int* npts, double* depth, V3* normal) {
  if (!triTriIntersect(P1, P2, P3, Q1in, Q2in, Q3in, R, T)) return false;
  const V3 Q1 = R * Q1in + T, Q2 = R * Q2in + T, Q3 = R * Q3in + T;
  // the reference leaves n / t unset for a degenerate triangle; zero here
  V3 n1(0, 0, 0), n2(0, 0, 0);
  double t1 = 0, t2 = 0;
  buildTrianglePlane(P1, P2, P3, &n1, &t1);
  buildTrianglePlane(Q1, Q2, Q3, &n2, &t2);
  V3 d1[3], d2[3];
  int k1 = 0, k2 = 0;
  double pd1 = 0, pd2 = 0;
  const V3 P[3] = {P1, P2, P3}, Q[3] = {Q1, Q2, Q3};
  computeDeepestPoints(Q, 3, n1, t1, &pd2, d2, &k2);
  computeDeepestPoints(P, 3, n2, t2, &pd1, d1, &k1);
  if (pd1 > pd2) {
    *npts = std::min(k2, 2);
    for (int i = 0; i < *npts; ++i) pts[i] = d2[i];
    *normal = n1;
    *depth = pd2;
  } else {
    *npts = std::min(k1, 2);
    for (int i = 0; i < *npts; ++i) pts[i] = d1[i];
    *normal = -n2;
    *depth = pd1;
  }
  return true;
}

// bvh_solver-inl.h:75-164: explicit stack seeded with (0, 0); descend into
// node 1 when node 2 is a leaf or node 1 is the larger inner node; the right
// child is pushed last and so visited first.
void meshCollide(const Mesh& m1, const Tf& tf1, const Mesh& m2, const Tf& tf2, int max_contacts,
                 MeshCollideOut& out, bool penetration) {
  out = MeshCollideOut();
  const M3 R = transpose(tf1.R) * tf2.R;
  const V3 T = tmul(tf1.R, tf2.t - tf1.t);
  std::vector<std::pair<int, int>> stk;
  stk.emplace_back(0, 0);
  while (!stk.empty()) {
    const auto t = stk.back();
    stk.pop_back();
    const BVNode& a = m1.nodes[t.first];
    const BVNode& b = m2.nodes[t.second];
    out.bv_tests++;
    if (!obbOverlap(R, T, a.obb, b.obb)) continue;
    const bool la = a.first_child < 0, lb = b.first_child < 0;
    if (la && lb) {
      const int p1 = -(a.first_child + 1), p2 = -(b.first_child + 1);
      const int* ta = &m1.tris[3 * p1];
      const int* tb = &m2.tris[3 * p2];
      out.leaf_tests++;
      if (!penetration) {
        if (triTriIntersect(m1.verts[ta[0]], m1.verts[ta[1]], m1.verts[ta[2]], m2.verts[tb[0]],
                            m2.verts[tb[1]], m2.verts[tb[2]], R, T)) {
          out.contacts.emplace_back(p1, p2);
          if (max_contacts > 0 && static_cast<int>(out.contacts.size()) >= max_contacts) return;
        }
        continue;
      }
      // shape_pair_intersect-inl.h:226-254: as many contacts as deepest
      // points, truncated to the free slots; positions and normals mapped
      // to the world by the transform the reference passes there, which is
      // the second mesh's (SimplexIntersect calls trianglePairIntersect
      // with tf2 in both transform slots, :269)
      V3 pts[2];
      int np = 0;
      double depth = 0;
      V3 nrm;
      if (triTriContacts(m1.verts[ta[0]], m1.verts[ta[1]], m1.verts[ta[2]], m2.verts[tb[0]], m2.verts[tb[1]],
                         m2.verts[tb[2]], R, T, pts, &np, &depth, &nrm)) {
        const int have = static_cast<int>(out.pen.size());
        if (max_contacts > 0 && max_contacts < have + np) np = max_contacts > have ? max_contacts - have : 0;
        for (int i = 0; i < np; ++i) {
          out.pen.push_back({p1, p2, tf2 * pts[i], tf2.R * nrm, depth});
          out.contacts.emplace_back(p1, p2);
        }
        if (max_contacts > 0 && static_cast<int>(out.pen.size()) >= max_contacts) return;
      }
      continue;
    }
    const bool go1 = lb || (!la && sqnorm(a.obb.extent) > sqnorm(b.obb.extent));
    if (go1) {
      stk.emplace_back(a.first_child, t.second);
      stk.emplace_back(a.first_child + 1, t.second);
    } else {
      stk.emplace_back(t.first, b.first_child);
      stk.emplace_back(t.first, b.first_child + 1);
    }
  }
}

bool aabbOverlap(const AABBd& a, const AABBd& b) {
  if (a.mn.x > b.mx.x || a.mn.y > b.mx.y || a.mn.z > b.mx.z) return false;
  if (a.mx.x < b.mn.x || a.mx.y < b.mn.y || a.mx.z < b.mn.z) return false;
  return true;
}

// binary_AABB_tree-inl.h:70-195: union box, longest axis, nth_element at n/2
// on box centres; left child built first.
void buildAABBTree(AABBTree& t, const float* mn, const float* mx, const int64_t* ids, int n) {
  t.nodes.clear();
  t.root = -1;
  if (n == 0) return;
  struct Obj { AABBd bv; int64_t id; };
  std::vector<Obj> objs(n);
  for (int i = 0; i < n; ++i) {
    objs[i].bv.mn = V3(mn[3 * i], mn[3 * i + 1], mn[3 * i + 2]);
    objs[i].bv.mx = V3(mx[3 * i], mx[3 * i + 1], mx[3 * i + 2]);
    objs[i].id = ids[i];
  }
  struct Task { int begin, n, parent; bool left; };
  std::vector<Task> stk;
  stk.push_back({0, n, -1, false});
  while (!stk.empty()) {
    const Task task = stk.back();
    stk.pop_back();
    const int idx = static_cast<int>(t.nodes.size());
    t.nodes.emplace_back();
    if (task.parent >= 0) t.nodes[task.parent].child[task.left ? 0 : 1] = idx;
    else t.root = idx;
    if (task.n == 1) {
      t.nodes[idx].bv = objs[task.begin].bv;
      t.nodes[idx].user_id = objs[task.begin].id;
      t.nodes[idx].leaf = true;
      continue;
    }
    AABBd u = objs[task.begin].bv;
    for (int i = 1; i < task.n; ++i) {
      const AABBd& b = objs[task.begin + i].bv;
      u.mn = V3(std::min(u.mn.x, b.mn.x), std::min(u.mn.y, b.mn.y), std::min(u.mn.z, b.mn.z));
      u.mx = V3(std::max(u.mx.x, b.mx.x), std::max(u.mx.y, b.mx.y), std::max(u.mx.z, b.mx.z));
    }
    const int nc = task.n / 2;
    int axis = 0;
    const double ext[3] = {u.mx.x - u.mn.x, u.mx.y - u.mn.y, u.mx.z - u.mn.z};
    if (ext[1] > ext[0]) axis = 1;
    if (ext[2] > ext[axis]) axis = 2;
    std::nth_element(objs.begin() + task.begin, objs.begin() + task.begin + nc,
                     objs.begin() + task.begin + task.n, [axis](const Obj& a, const Obj& b) {
                       return ((a.bv.mn + a.bv.mx) * 0.5)[axis] < ((b.bv.mn + b.bv.mx) * 0.5)[axis];
                     });
    t.nodes[idx].bv = u;
    t.nodes[idx].leaf = false;
    stk.push_back({task.begin + nc, task.n - nc, idx, false});
    stk.push_back({task.begin, nc, idx, true});
  }
}

// binary_AABB_tree-inl.h:400-440
void aabbTreeQuery(const AABBTree& t, const AABBd& q, std::vector<int64_t>& hits,
                   int64_t* nodes_visited) {
  if (t.root < 0) return;
  std::vector<int> stk{t.root};
  while (!stk.empty()) {
    const int i = stk.back();
    stk.pop_back();
    if (nodes_visited) (*nodes_visited)++;
    const auto& nd = t.nodes[i];
    if (!aabbOverlap(q, nd.bv)) continue;
    if (nd.leaf) {
      hits.push_back(nd.user_id);
    } else {
      stk.push_back(nd.child[1]);
      stk.push_back(nd.child[0]);
    }
  }
}

}  // namespace oracle
