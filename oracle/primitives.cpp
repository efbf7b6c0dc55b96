// CPU oracle -- TEST INFRASTRUCTURE ONLY (see omath.h).
// Closed-form pair tests that fcl::collide uses instead of GJK/EPA
// (include/fcl/narrowphase/detail/gjk_solver-inl.h:209-217).
#include <algorithm>
#include <cstring>
#include <limits>

#include "narrowphase.h"

namespace oracle {

namespace {
constexpr double kEps = std::numeric_limits<double>::epsilon();
}

// primitive_shape_algorithm/sphere_sphere-inl.h:45-68
bool sphereSphereIntersect(double r1, const Tf& tf1, double r2, const Tf& tf2,
                           std::vector<ContactPt>* c) {
  const V3 diff = tf2.t - tf1.t;
  const double len = norm(diff);
  if (len > r1 + r2) return false;
  if (c) {
    const V3 n = len > 0 ? diff / len : diff;
    const V3 p = tf1.t + diff * r1 / (r1 + r2);
    c->push_back({n, p, r1 + r2 - len});
  }
  return true;
}

// primitive_shape_algorithm/sphere_capsule-inl.h:50-101
bool sphereCapsuleIntersect(double r1, const Tf& tf1, double r2, double lz2, const Tf& tf2,
                            std::vector<ContactPt>* c) {
  const V3 pos1(0., 0., 0.5 * lz2);
  const V3 pos2(0., 0., -0.5 * lz2);
  const V3 sc = inverse(tf2) * tf1.t;
  // lineSegmentPointClosestToPoint
  V3 sp;
  {
    const V3 v = pos2 - pos1;
    const V3 w = sc - pos1;
    const double c1 = dot(w, v);
    const double c2 = dot(v, v);
    if (c1 <= 0) sp = pos1;
    else if (c2 <= c1) sp = pos2;
    else sp = pos1 + v * (c1 / c2);
  }
  const V3 diff = sc - sp;
  const double distance = norm(diff) - r1 - r2;
  if (distance > 0) return false;
  const V3 ln = -normalized(diff);
  if (c) c->push_back({tf2.R * ln, tf2 * (sp + ln * distance), -distance});
  return true;
}

// primitive_shape_algorithm/sphere_box-inl.h:48-136
bool sphereBoxIntersect(double r, const Tf& X_FS, const V3& side, const Tf& X_FB,
                        std::vector<ContactPt>* c) {
  const Tf X_BS = inverse(X_FB) * X_FS;
  const V3 p_BC = X_BS.t;
  V3 p_BN;
  bool clamped = false;
  const V3 half = side / 2;
  for (int i = 0; i < 3; ++i) {
    p_BN[i] = p_BC[i];
    if (p_BC[i] < -half[i]) { clamped = true; p_BN[i] = -half[i]; }
    if (p_BC[i] > half[i]) { clamped = true; p_BN[i] = half[i]; }
  }
  const V3 p_CN = p_BN - p_BC;
  const double d2 = sqnorm(p_CN);
  if (d2 > r * r) return false;
  if (c) {
    double depth = 0;
    V3 n, p_BP;
    const double eps = 16 * kEps;
    if (clamped && d2 > eps * eps) {
      const double dist = std::sqrt(d2);
      n = p_CN / dist;
      depth = r - dist;
      p_BP = p_BN + n * (depth * 0.5);
    } else {
      double min_d = std::numeric_limits<double>::infinity();
      int ax = -1;
      for (int i = 0; i < 3; ++i) {
        const double d = p_BC[i] >= 0 ? half[i] - p_BC[i] : p_BC[i] + half[i];
        if (d + eps < min_d) { min_d = d; ax = i; }
      }
      n = V3(0, 0, 0);
      n[ax] = p_BC[ax] >= 0 ? -1 : 1;
      depth = min_d + r;
      p_BP = p_BC + n * ((r - min_d) / 2);
    }
    c->push_back({X_FB.R * n, X_FB * p_BP, depth});
  }
  return true;
}

// primitive_shape_algorithm/sphere_cylinder-inl.h:48-150
bool sphereCylinderIntersect(double r_s, const Tf& X_FS, double cr, double clz, const Tf& X_FC,
                             std::vector<ContactPt>* c) {
  const Tf X_CS = inverse(X_FC) * X_FS;
  const V3 p_CS = X_CS.t;
  V3 p_CN = p_CS;
  bool outside = false;
  const double hh = clz / 2;
  if (p_CS.z > hh) { outside = true; p_CN.z = hh; }
  else if (p_CS.z < -hh) { outside = true; p_CN.z = -hh; }
  {
    double rx = p_CS.x, ry = p_CS.y;
    const double sq = rx * rx + ry * ry;
    if (sq > cr * cr) {
      outside = true;
      const double k = cr / std::sqrt(sq);
      rx *= k;
      ry *= k;
      p_CN.x = rx;
      p_CN.y = ry;
    }
  }
  const V3 p_SN = p_CN - p_CS;
  const double d2 = sqnorm(p_SN);
  if (d2 > r_s * r_s) return false;
  if (c) {
    double depth = 0;
    V3 n, p_CP;
    const double eps = 16 * kEps;
    if (outside && d2 > eps * eps) {
      const double d = std::sqrt(d2);
      n = p_SN / d;
      depth = r_s - d;
      p_CP = p_CN + n * (depth * 0.5);
    } else {
      const double face = p_CS.z >= 0 ? clz / 2 - p_CS.z : p_CS.z + clz / 2;
      const double dxy = std::sqrt(p_CS.x * p_CS.x + p_CS.y * p_CS.y);
      const double barrel = cr - dxy;
      if (barrel < face - eps) {
        if (dxy > eps) {
          n = V3(-p_CS.x / dxy, -p_CS.y / dxy, 0);
          depth = r_s + barrel;
          p_CP = p_CS + n * ((r_s - barrel) / 2);
        } else {
          n = V3(-1, 0, 0);
          depth = r_s + cr;
          p_CP = p_CS + n * ((r_s - barrel) / 2);
        }
      } else {
        n = V3(0, 0, p_CS.z >= 0 ? -1 : 1);
        depth = face + r_s;
        p_CP = p_CS + n * ((r_s - face) / 2);
      }
    }
    c->push_back({X_FC.R * n, X_FC * p_CP, depth});
  }
  return true;
}

namespace {

// box_box-inl.h:52-73
void lineClosestApproach(const V3& pa, const V3& ua, const V3& pb, const V3& ub, double* alpha,
                         double* beta) {
  const V3 p = pb - pa;
  const double uaub = dot(ua, ub);
  const double q1 = dot(ua, p);
  const double q2 = -dot(ub, p);
  double d = 1 - uaub * uaub;
  if (d <= static_cast<double>(0.0001f)) {
    *alpha = 0;
    *beta = 0;
  } else {
    d = 1 / d;
    *alpha = (q1 + uaub * q2) * d;
    *beta = (uaub * q1 + q2) * d;
  }
}

// box_box-inl.h:76-136: clip the incident quad by the reference rectangle
// [-h0,h0] x [-h1,h1]; at most 8 output vertices.
int intersectRectQuad(const double h[2], double p[8], double ret[16]) {
  int nq = 4, nr = 0;
  double buffer[16];
  double* q = p;
  double* r = ret;
  for (int dir = 0; dir <= 1; ++dir) {
    for (int sign = -1; sign <= 1; sign += 2) {
      double* pq = q;
      double* pr = r;
      nr = 0;
      for (int i = nq; i > 0; --i) {
        if (sign * pq[dir] < h[dir]) {
          pr[0] = pq[0];
          pr[1] = pq[1];
          pr += 2;
          nr++;
          if (nr & 8) { q = r; goto done; }
        }
        double* nextq = (i > 1) ? pq + 2 : q;
        if ((sign * pq[dir] < h[dir]) ^ (sign * nextq[dir] < h[dir])) {
          pr[1 - dir] = pq[1 - dir] + (nextq[1 - dir] - pq[1 - dir]) / (nextq[dir] - pq[dir]) *
                                          (sign * h[dir] - pq[dir]);
          pr[dir] = sign * h[dir];
          pr += 2;
          nr++;
          if (nr & 8) { q = r; goto done; }
        }
        pq += 2;
      }
      q = r;
      r = (q == ret) ? buffer : ret;
      nq = nr;
    }
  }
done:
  if (q != ret) std::memcpy(ret, q, nr * 2 * sizeof(double));
  return nr;
}

// box_box-inl.h:139-211: keep m of n polygon points, spread in angle around
// the centroid starting from point i0.
void cullPoints(int n, double p[], int m, int i0, int iret[]) {
  double a, cx, cy, q;
  switch (n) {
    case 1:
      cx = p[0];
      cy = p[1];
      break;
    case 2:
      cx = 0.5 * (p[0] + p[2]);
      cy = 0.5 * (p[1] + p[3]);
      break;
    default:
      a = 0;
      cx = 0;
      cy = 0;
      for (int i = 0; i < n - 1; ++i) {
        q = p[i * 2] * p[i * 2 + 3] - p[i * 2 + 2] * p[i * 2 + 1];
        a += q;
        cx += q * (p[i * 2] + p[i * 2 + 2]);
        cy += q * (p[i * 2 + 1] + p[i * 2 + 3]);
      }
      q = p[n * 2 - 2] * p[1] - p[0] * p[n * 2 - 1];
      if (std::fabs(a + q) > kEps) a = 1 / (3 * (a + q));
      else a = static_cast<double>(1e18f);
      cx = a * (cx + q * (p[n * 2 - 2] + p[0]));
      cy = a * (cy + q * (p[n * 2 - 1] + p[1]));
  }
  double A[8];
  for (int i = 0; i < n; ++i) A[i] = std::atan2(p[i * 2 + 1] - cy, p[i * 2] - cx);
  int avail[8];
  for (int i = 0; i < n; ++i) avail[i] = 1;
  avail[i0] = 0;
  iret[0] = i0;
  iret++;
  const double pi = 3.141592653589793238462643383279502884197169399375105820974944592;
  for (int j = 1; j < m; ++j) {
    a = j * (2 * pi / m) + A[i0];
    if (a > pi) a -= 2 * pi;
    double maxdiff = 1e9, diff;
    *iret = i0;
    for (int i = 0; i < n; ++i) {
      if (avail[i]) {
        diff = std::fabs(A[i] - a);
        if (diff > pi) diff = 2 * pi - diff;
        if (diff < maxdiff) { maxdiff = diff; *iret = i; }
      }
    }
    avail[*iret] = 0;
    iret++;
  }
}

}  // namespace

// box_box-inl.h:214-803 (boxBox2 with maxc = 4) and :823-841.
// Returns the ODE return code (0 = separated); contacts appended to c.
static int boxBox2(const V3& side1, const M3& R1, const V3& T1, const V3& side2, const M3& R2,
                   const V3& T2, std::vector<ContactPt>& contacts) {
  const int maxc_in = 4;
  const double fudge = 1.05;
  V3 normalC;
  double s, s2, l;
  int invert_normal, code;
  const V3 p = T2 - T1;
  const V3 pp = tmul(R1, p);
  const V3 A = side1 * 0.5;
  const V3 B = side2 * 0.5;
  const M3 R = transpose(R1) * R2;
  M3 Q;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Q.m[i][j] = std::fabs(R.m[i][j]);
  int best_col = -1;
  const M3* normalR = nullptr;
  s = -std::numeric_limits<double>::max();
  invert_normal = 0;
  code = 0;

  auto face_axis = [&](double t, double proj_sum, const M3* Rsrc, int col, int cd) -> bool {
    s2 = std::fabs(t) - proj_sum;
    if (s2 > 0) return true;
    if (s2 > s) { s = s2; best_col = col; normalR = Rsrc; invert_normal = (t < 0); code = cd; }
    return false;
  };
  if (face_axis(pp[0], dot(Q.row(0), B) + A[0], &R1, 0, 1)) return 0;
  if (face_axis(pp[1], dot(Q.row(1), B) + A[1], &R1, 1, 2)) return 0;
  if (face_axis(pp[2], dot(Q.row(2), B) + A[2], &R1, 2, 3)) return 0;
  if (face_axis(dot(R2.col(0), p), dot(Q.col(0), A) + B[0], &R2, 0, 4)) return 0;
  if (face_axis(dot(R2.col(1), p), dot(Q.col(1), A) + B[1], &R2, 1, 5)) return 0;
  if (face_axis(dot(R2.col(2), p), dot(Q.col(2), A) + B[2], &R2, 2, 6)) return 0;

  {
    const double amax = std::max(std::max(A.x, std::max(A.y, A.z)), std::max(B.x, std::max(B.y, B.z)));
    const double scale = std::max(amax, 1.0) * 10 * kEps;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Q.m[i][j] += scale;
  }
  auto edge_axis = [&](double t, double proj_sum, const V3& n, int cd) -> bool {
    s2 = std::fabs(t) - proj_sum;
    if (s2 > 0) return true;
    l = norm(n);
    if (l > kEps) {
      s2 /= l;
      if (s2 * fudge > s) { s = s2; best_col = -1; normalC = n / l; invert_normal = (t < 0); code = cd; }
    }
    return false;
  };
  const auto& Rm = R.m;
  const auto& Qm = Q.m;
  if (edge_axis(pp[2] * Rm[1][0] - pp[1] * Rm[2][0], A[1] * Qm[2][0] + A[2] * Qm[1][0] + B[1] * Qm[0][2] + B[2] * Qm[0][1], V3(0, -Rm[2][0], Rm[1][0]), 7)) return 0;
  if (edge_axis(pp[2] * Rm[1][1] - pp[1] * Rm[2][1], A[1] * Qm[2][1] + A[2] * Qm[1][1] + B[0] * Qm[0][2] + B[2] * Qm[0][0], V3(0, -Rm[2][1], Rm[1][1]), 8)) return 0;
  if (edge_axis(pp[2] * Rm[1][2] - pp[1] * Rm[2][2], A[1] * Qm[2][2] + A[2] * Qm[1][2] + B[0] * Qm[0][1] + B[1] * Qm[0][0], V3(0, -Rm[2][2], Rm[1][2]), 9)) return 0;
  if (edge_axis(pp[0] * Rm[2][0] - pp[2] * Rm[0][0], A[0] * Qm[2][0] + A[2] * Qm[0][0] + B[1] * Qm[1][2] + B[2] * Qm[1][1], V3(Rm[2][0], 0, -Rm[0][0]), 10)) return 0;
  if (edge_axis(pp[0] * Rm[2][1] - pp[2] * Rm[0][1], A[0] * Qm[2][1] + A[2] * Qm[0][1] + B[0] * Qm[1][2] + B[2] * Qm[1][0], V3(Rm[2][1], 0, -Rm[0][1]), 11)) return 0;
  if (edge_axis(pp[0] * Rm[2][2] - pp[2] * Rm[0][2], A[0] * Qm[2][2] + A[2] * Qm[0][2] + B[0] * Qm[1][1] + B[1] * Qm[1][0], V3(Rm[2][2], 0, -Rm[0][2]), 12)) return 0;
  if (edge_axis(pp[1] * Rm[0][0] - pp[0] * Rm[1][0], A[0] * Qm[1][0] + A[1] * Qm[0][0] + B[1] * Qm[2][2] + B[2] * Qm[2][1], V3(-Rm[1][0], Rm[0][0], 0), 13)) return 0;
  if (edge_axis(pp[1] * Rm[0][1] - pp[0] * Rm[1][1], A[0] * Qm[1][1] + A[1] * Qm[0][1] + B[0] * Qm[2][2] + B[2] * Qm[2][0], V3(-Rm[1][1], Rm[0][1], 0), 14)) return 0;
  if (edge_axis(pp[1] * Rm[0][2] - pp[0] * Rm[1][2], A[0] * Qm[1][2] + A[1] * Qm[0][2] + B[0] * Qm[2][1] + B[1] * Qm[2][0], V3(-Rm[1][2], Rm[0][2], 0), 15)) return 0;

  if (!code) return 0;
  V3 normal = best_col != -1 ? normalR->col(best_col) : R1 * normalC;
  if (invert_normal) normal = -normal;
  const double depth = -s;

  if (code > 6) {
    V3 pa = T1;
    for (int j = 0; j < 3; ++j) {
      const double sign = dot(R1.col(j), normal) > 0 ? 1 : -1;
      pa += R1.col(j) * (A[j] * sign);
    }
    V3 pb = T2;
    for (int j = 0; j < 3; ++j) {
      const double sign = dot(R2.col(j), normal) > 0 ? -1 : 1;
      pb += R2.col(j) * (B[j] * sign);
    }
    double alpha, beta;
    const V3 ua = R1.col((code - 7) / 3);
    const V3 ub = R2.col((code - 7) % 3);
    lineClosestApproach(pa, ua, pb, ub, &alpha, &beta);
    pa += ua * alpha;
    pb += ub * beta;
    contacts.push_back({normal, (pa + pb) * 0.5, depth});
    return code;
  }

  const M3 *Ra, *Rb;
  const V3 *pa, *pb, *Sa, *Sb;
  if (code <= 3) { Ra = &R1; Rb = &R2; pa = &T1; pb = &T2; Sa = &A; Sb = &B; }
  else { Ra = &R2; Rb = &R1; pa = &T2; pb = &T1; Sa = &B; Sb = &A; }
  const V3 normal2 = code <= 3 ? normal : -normal;
  const V3 nr = tmul(*Rb, normal2);
  const V3 anr(std::fabs(nr.x), std::fabs(nr.y), std::fabs(nr.z));
  int lanr, a1, a2;
  if (anr[1] > anr[0]) {
    if (anr[1] > anr[2]) { a1 = 0; lanr = 1; a2 = 2; }
    else { a1 = 0; a2 = 1; lanr = 2; }
  } else {
    if (anr[0] > anr[2]) { lanr = 0; a1 = 1; a2 = 2; }
    else { a1 = 0; a2 = 1; lanr = 2; }
  }
  V3 center;
  if (nr[lanr] < 0) center = (*pb) - (*pa) + Rb->col(lanr) * ((*Sb)[lanr]);
  else center = (*pb) - (*pa) - Rb->col(lanr) * ((*Sb)[lanr]);
  const int codeN = code <= 3 ? code - 1 : code - 4;
  int code1, code2;
  if (codeN == 0) { code1 = 1; code2 = 2; }
  else if (codeN == 1) { code1 = 0; code2 = 2; }
  else { code1 = 0; code2 = 1; }
  double quad[8];
  const double c1 = dot(Ra->col(code1), center);
  const double c2 = dot(Ra->col(code2), center);
  double m11 = dot(Rb->col(a1), Ra->col(code1));
  double m12 = dot(Rb->col(a2), Ra->col(code1));
  double m21 = dot(Rb->col(a1), Ra->col(code2));
  double m22 = dot(Rb->col(a2), Ra->col(code2));
  const double k1 = m11 * (*Sb)[a1];
  const double k2 = m21 * (*Sb)[a1];
  const double k3 = m12 * (*Sb)[a2];
  const double k4 = m22 * (*Sb)[a2];
  quad[0] = c1 - k1 - k3;
  quad[1] = c2 - k2 - k4;
  quad[2] = c1 - k1 + k3;
  quad[3] = c2 - k2 + k4;
  quad[4] = c1 + k1 + k3;
  quad[5] = c2 + k2 + k4;
  quad[6] = c1 + k1 - k3;
  quad[7] = c2 + k2 - k4;
  const double rect[2] = {(*Sa)[code1], (*Sa)[code2]};
  double ret[16];
  const int n_int = intersectRectQuad(rect, quad, ret);
  if (n_int < 1) return code;
  V3 points[8];
  double dep[8];
  const double det1 = 1.f / (m11 * m22 - m12 * m21);
  m11 *= det1;
  m12 *= det1;
  m21 *= det1;
  m22 *= det1;
  int cnum = 0;
  for (int j = 0; j < n_int; ++j) {
    const double kk1 = m22 * (ret[j * 2] - c1) - m12 * (ret[j * 2 + 1] - c2);
    const double kk2 = -m21 * (ret[j * 2] - c1) + m11 * (ret[j * 2 + 1] - c2);
    points[cnum] = center + Rb->col(a1) * kk1 + Rb->col(a2) * kk2;
    dep[cnum] = (*Sa)[codeN] - dot(normal2, points[cnum]);
    if (dep[cnum] >= 0) {
      ret[cnum * 2] = ret[j * 2];
      ret[cnum * 2 + 1] = ret[j * 2 + 1];
      cnum++;
    }
  }
  if (cnum < 1) return code;
  int maxc = maxc_in;
  if (maxc > cnum) maxc = cnum;
  if (maxc < 1) maxc = 1;
  int iret[] = {0, 1, 2, 3, 4, 5, 6, 7};
  if (cnum > maxc) {
    int i1 = 0;
    double maxdepth = dep[0];
    for (int i = 1; i < cnum; ++i)
      if (dep[i] > maxdepth) { maxdepth = dep[i]; i1 = i; }
    cullPoints(cnum, ret, maxc, i1, iret);
    cnum = maxc;
  }
  for (int j = 0; j < cnum; ++j) {
    const int i = iret[j];
    const V3 pw = code < 4 ? points[i] + (*pa) + normal * (dep[i] / 2)
                           : points[i] + (*pa) - normal * (dep[i] / 2);
    contacts.push_back({normal, pw, dep[i]});
  }
  return code;
}

bool boxBoxIntersect(const V3& side1, const Tf& tf1, const V3& side2, const Tf& tf2,
                     std::vector<ContactPt>* c) {
  std::vector<ContactPt> contacts;
  const int rc = boxBox2(side1, tf1.R, tf1.t, side2, tf2.R, tf2.t, contacts);
  if (c) *c = contacts;
  return rc != 0;
}

// Textbook closest points between segments (Ericson, Real-Time Collision
// Detection, 5.1.9), minus the radii; a closed-form check for GJK distance
// on capsules (the reference's own check uses capsule_capsule-inl.h).
double capsuleCapsuleDistance(double r1, double lz1, const Tf& tf1, double r2, double lz2,
                              const Tf& tf2) {
  const V3 p1 = tf1 * V3(0, 0, -lz1 / 2), q1 = tf1 * V3(0, 0, lz1 / 2);
  const V3 p2 = tf2 * V3(0, 0, -lz2 / 2), q2 = tf2 * V3(0, 0, lz2 / 2);
  const V3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
  const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  double s, t;
  const double c = dot(d1, r), b = dot(d1, d2), den = a * e - b * b;
  s = den != 0 ? std::clamp((b * f - c * e) / den, 0.0, 1.0) : 0.0;
  t = (b * s + f) / e;
  if (t < 0) { t = 0; s = std::clamp(-c / a, 0.0, 1.0); }
  else if (t > 1) { t = 1; s = std::clamp((b - c) / a, 0.0, 1.0); }
  const V3 c1 = p1 + d1 * s, c2 = p2 + d2 * t;
  return norm(c1 - c2) - r1 - r2;
}

// ---- mesh-shape leaf tests -------------------------------------------------

namespace {

// sphere_triangle-inl.h:52-74: squared distance from p to segment
// [from, to] and the nearest point.
double segmentSqrDistance(const V3& from, const V3& to, const V3& p, V3& nearest) {
  V3 diff = p - from;
  const V3 v = to - from;
  double t = dot(v, diff);
  if (t > 0) {
    const double vv = dot(v, v);
    if (t < vv) {
      t /= vv;
      diff -= v * t;
    } else {
      t = 1;
      diff -= v;
    }
  } else {
    t = 0;
  }
  nearest = from + v * t;
  return dot(diff, diff);
}

// sphere_triangle-inl.h:77-104
bool projectInTriangle(const V3& p1, const V3& p2, const V3& p3, const V3& n, const V3& p) {
  const V3 e1 = p2 - p1, e2 = p3 - p2, e3 = p1 - p3;
  const double r1 = dot(cross(e1, n), p - p1);
  const double r2 = dot(cross(e2, n), p - p2);
  const double r3 = dot(cross(e3, n), p - p3);
  return (r1 > 0 && r2 > 0 && r3 > 0) || (r1 <= 0 && r2 <= 0 && r3 <= 0);
}

// box_triangle-inl.h:9-69: is the box separated from the triangle along
// box axis x / y / z crossed with the edge e0 (v0 on the edge, v1 off it)?
bool separatedByEdgeAxes(const double h[3], const V3& e0, const V3& v0, const V3& v1) {
  const double ax = std::fabs(e0.x), ay = std::fabs(e0.y), az = std::fabs(e0.z);
  auto sep = [](double p0, double p1, double rad) {
    if (p0 < p1) return p0 > rad || p1 < -rad;
    return p1 > rad || p0 < -rad;
  };
  if (sep(v0.y * e0.z - v0.z * e0.y, v1.y * e0.z - v1.z * e0.y, h[1] * az + h[2] * ay)) return true;
  if (sep(v0.x * e0.z - v0.z * e0.x, v1.x * e0.z - v1.z * e0.x, h[0] * az + h[2] * ax)) return true;
  if (sep(v0.x * e0.y - v0.y * e0.x, v1.x * e0.y - v1.y * e0.x, h[0] * ay + h[1] * ax)) return true;
  return false;
}

void minMax3(double a, double b, double c, double& mn, double& mx) {
  mn = mx = a;
  if (b < mn) mn = b;
  if (b > mx) mx = b;
  if (c < mn) mn = c;
  if (c > mx) mx = c;
}

}  // namespace

// sphere_triangle-inl.h:107-187.  c (optional) receives the contact as the
// reference writes it: pos = the nearest point found, normal = that point
// minus the centre, normalised (the flipped plane normal when they
// coincide), depth = -(radius - distance) (-radius when they coincide).
bool sphereTriangleContact(double radius, const Tf& tf_sphere, const V3& P1, const V3& P2, const V3& P3,
                           ContactPt* c) {
  V3 n = cross(P2 - P1, P3 - P1);
  normalize(n);
  const V3 ctr = tf_sphere.t;
  const double rt = radius + std::numeric_limits<double>::epsilon();
  double dist = dot(ctr - P1, n);
  if (dist < 0) {
    dist *= -1;
    n *= -1;
  }
  if (!(dist < rt)) return false;
  bool has = false;
  V3 cp;
  if (projectInTriangle(P1, P2, P3, n, ctr)) {
    has = true;
    cp = ctr - n * dist;
  } else {
    const double cr2 = rt * rt;
    V3 q;
    if (segmentSqrDistance(P1, P2, ctr, q) < cr2) { has = true; cp = q; }
    if (segmentSqrDistance(P2, P3, ctr, q) < cr2) { has = true; cp = q; }
    if (segmentSqrDistance(P3, P1, ctr, q) < cr2) { has = true; cp = q; }
  }
  if (!has) return false;
  const V3 d = cp - ctr;
  const double d2 = dot(d, d);
  if (!(d2 < rt * rt)) return false;
  if (c) {
    if (d2 > 0) {
      const double l = std::sqrt(d2);
      c->normal = normalized(d);
      c->pos = cp;
      c->depth = -(radius - l);
    } else {
      c->normal = -n;
      c->pos = cp;
      c->depth = -radius;
    }
  }
  return true;
}

bool sphereTriangleHit(double radius, const Tf& tf_sphere, const V3& P1, const V3& P2, const V3& P3) {
  return sphereTriangleContact(radius, tf_sphere, P1, P2, P3, nullptr);
}

bool boxTriangleHit(const double side[3], const Tf& tf_box, const V3& P1, const V3& P2, const V3& P3,
                    const Tf& tf_tri) {
  const Tf t = inverse(tf_box) * tf_tri;
  const V3 v0 = t * P1, v1 = t * P2, v2 = t * P3;
  const double h[3] = {0.5 * side[0], 0.5 * side[1], 0.5 * side[2]};
  double mn, mx;
  minMax3(v0.x, v1.x, v2.x, mn, mx);
  if (mx < -h[0] || mn > h[0]) return false;
  minMax3(v0.y, v1.y, v2.y, mn, mx);
  if (mx < -h[1] || mn > h[1]) return false;
  minMax3(v0.z, v1.z, v2.z, mn, mx);
  if (mx < -h[2] || mn > h[2]) return false;
  const V3 e0 = v1 - v0;
  if (separatedByEdgeAxes(h, e0, v0, v2)) return false;
  const V3 e1 = v2 - v1;
  if (separatedByEdgeAxes(h, e1, v1, v0)) return false;
  const V3 e2 = v0 - v2;
  if (separatedByEdgeAxes(h, e2, v0, v1)) return false;
  const V3 n = cross(e0, e1);
  const double nv[3] = {n.x, n.y, n.z}, pv[3] = {v0.x, v0.y, v0.z};
  double pmin[3], pmax[3];
  for (int i = 0; i < 3; ++i) {
    if (nv[i] > 0.0) {
      pmin[i] = -h[i] - pv[i];
      pmax[i] = h[i] - pv[i];
    } else {
      pmin[i] = h[i] - pv[i];
      pmax[i] = -h[i] - pv[i];
    }
  }
  return dot(n, V3(pmin[0], pmin[1], pmin[2])) <= 0 && dot(n, V3(pmax[0], pmax[1], pmax[2])) >= 0;
}

bool shapeTriangleHit(const Shape& s, const Tf& tf_shape, const V3 tri[3], const Tf& tf_tri, int max_iter,
                      double tol) {
  if (s.type == kSphere) return sphereTriangleHit(s.p[0], tf_shape, tf_tri * tri[0], tf_tri * tri[1], tf_tri * tri[2]);
  if (s.type == kBox) return boxTriangleHit(s.p, tf_shape, tri[0], tri[1], tri[2], tf_tri);
  Shape t;
  t.type = kTriangle;
  t.tri[0] = tri[0];
  t.tri[1] = tri[1];
  t.tri[2] = tri[2];
  const MinkDiff md = makeMinkDiff(s, tf_shape, t, tf_tri);
  return mprIntersect(md, max_iter, tol) == MPRStatus::Intersect;
}

// The leaf test with penetration enabled (GJKSolver::shapeTriangleIntersect
// with contact outputs, gjk_solver-inl.h:545-563): the sphere's closed form
// with its contact, else GJK from -guess (guess = UnitX) and EPA on
// (shape, TriangleP) (shapeTransformedTriangleIntersectIndepImplRun,
// gjk_solver-inl.h:452-527: boxes too -- their closed form is boolean only):
// a hit when GJK intersects and EPA does not fail; pos = tf_shape * the
// witness midpoint, normal = tf_shape.R * normalize(p0 - p1) (UnitZ when
// zero), depth = the EPA depth.  The contact is before the mesh-shape
// traversal's normal reversal.
bool shapeTriangleContact(const Shape& s, const Tf& tf_shape, const V3 tri[3], const Tf& tf_tri,
                          const SolverParams& sp, ContactPt& c, int* gjk_status, int* epa_status) {
  if (gjk_status) *gjk_status = -1;
  if (epa_status) *epa_status = -1;
  if (s.type == kSphere)
    return sphereTriangleContact(s.p[0], tf_shape, tf_tri * tri[0], tf_tri * tri[1], tf_tri * tri[2], &c);
  Shape t;
  t.type = kTriangle;
  t.tri[0] = tri[0];
  t.tri[1] = tri[1];
  t.tri[2] = tri[2];
  const MinkDiff md = makeMinkDiff(s, tf_shape, t, tf_tri);
  const V3 guess = sp.guess_valid ? sp.guess : V3(1, 0, 0);
  Simplex smp;
  const GJKStatus g = gjkEvaluate(md, smp, -guess, sp.gjk_tolerance, sp.gjk_max_iterations, nullptr);
  if (gjk_status) *gjk_status = static_cast<int>(g);
  if (g != GJKStatus::Intersect) return false;
  double depth = 0;
  V3 p0, p1;
  const EPAStatus e = epaEvaluate(smp, md, sp.epa_tolerance, sp.epa_max_face_num, sp.epa_max_iterations, &depth,
                                  &p0, &p1);
  if (epa_status) *epa_status = static_cast<int>(e);
  if (e == EPAStatus::Failed) return false;
  V3 n1 = p0 - p1;
  if (sqnorm(n1) <= 0) n1 = V3(0, 0, 1);
  else normalize(n1);
  c.pos = tf_shape * (0.5 * (p0 + p1));
  c.normal = tf_shape.R * n1;
  c.depth = depth;
  return true;
}

}  // namespace oracle
