// Device closed-form pair tests that fcl::collide uses in place of GJK/EPA
// (include/fcl/narrowphase/detail/gjk_solver-inl.h:209-217); each function
// cites the primitive_shape_algorithm file it follows.  FP64 throughout.
// Contacts: up to 4 (normal from shape 1 to shape 2, world position, depth).
#pragma once

#include "dmath.h"

namespace mfcl {

struct PContact {
  d3 n, p;
  double depth;
};
struct PContacts {
  PContact c[4];
  int n;
};

constexpr double kDblEps = 2.220446049250313e-16;

// std::partial_sort(c, c + k, c + n, deeper-first) as libstdc++ performs it
// (heap select over the first k, then a heap sort of them), which is what
// the reference's truncation to the free contact slots runs
// (shape_pair_intersect-inl.h:91-113): equal depths keep the same relative
// order as there.  n <= 4.
MFCL_D bool deeperThan(const PContact& a, const PContact& b) { return a.depth > b.depth; }
MFCL_D void heapPush(PContact* f, int hole, int top, const PContact& v) {
  int parent = (hole - 1) / 2;
  while (hole > top && deeperThan(f[parent], v)) {
    f[hole] = f[parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  f[hole] = v;
}
MFCL_D void heapAdjust(PContact* f, int hole, int len, const PContact& v) {
  const int top = hole;
  int child = hole;
  while (child < (len - 1) / 2) {
    child = 2 * (child + 1);
    if (deeperThan(f[child], f[child - 1])) --child;
    f[hole] = f[child];
    hole = child;
  }
  if ((len & 1) == 0 && child == (len - 2) / 2) {
    child = 2 * (child + 1);
    f[hole] = f[child - 1];
    hole = child - 1;
  }
  heapPush(f, hole, top, v);
}
MFCL_D void heapPop(PContact* f, int len, int result) {
  const PContact v = f[result];
  f[result] = f[0];
  heapAdjust(f, 0, len, v);
}
MFCL_D void partialSortDeepest(PContact* f, int k, int n) {
  if (k >= 2) {
    for (int parent = (k - 2) / 2;; --parent) {
      const PContact v = f[parent];
      heapAdjust(f, parent, k, v);
      if (parent == 0) break;
    }
  }
  for (int i = k; i < n; ++i)
    if (deeperThan(f[i], f[0])) heapPop(f, k, i);
  for (int len = k; len > 1;) {
    --len;
    heapPop(f, len, len);
  }
}

// sphere_sphere-inl.h:45-68
MFCL_D bool primSphereSphere(double r1, const tf3& t1, double r2, const tf3& t2, PContacts* out) {
  const d3 diff = t2.t - t1.t;
  const double len = norm(diff);
  if (len > r1 + r2) return false;
  if (out) {
    const d3 n = len > 0 ? diff / len : diff;
    out->c[out->n++] = {n, t1.t + diff * r1 / (r1 + r2), r1 + r2 - len};
  }
  return true;
}

// sphere_capsule-inl.h:50-101
MFCL_D bool primSphereCapsule(double r1, const tf3& t1, double r2, double lz, const tf3& t2, PContacts* out) {
  const d3 pos1 = {0, 0, 0.5 * lz}, pos2 = {0, 0, -0.5 * lz};
  const d3 sc = tmul(t2.R, t1.t) + (-tmul(t2.R, t2.t));
  d3 sp;
  const d3 v = pos2 - pos1, w = sc - pos1;
  const double c1 = dot(w, v), c2 = dot(v, v);
  if (c1 <= 0) sp = pos1;
  else if (c2 <= c1) sp = pos2;
  else sp = pos1 + v * (c1 / c2);
  const d3 diff = sc - sp;
  const double dist = norm(diff) - r1 - r2;
  if (dist > 0) return false;
  const d3 ln = -normalized(diff);
  if (out) out->c[out->n++] = {mul(t2.R, ln), apply(t2, sp + ln * dist), -dist};
  return true;
}

// sphere_box-inl.h:48-136
MFCL_D bool primSphereBox(double r, const tf3& X_FS, d3 side, const tf3& X_FB, PContacts* out) {
  const tf3 X_BS = relTf(X_FB, X_FS);
  const d3 p_BC = X_BS.t;
  const d3 half = side / 2;
  d3 p_BN = p_BC;
  bool clamped = false;
  for (int i = 0; i < 3; ++i) {
    const double q = comp(p_BC, i), h = comp(half, i);
    if (q < -h) { clamped = true; setComp(p_BN, i, -h); }
    if (q > h) { clamped = true; setComp(p_BN, i, h); }
  }
  const d3 p_CN = p_BN - p_BC;
  const double d2 = sqnorm(p_CN);
  if (d2 > r * r) return false;
  if (out) {
    const double eps = 16 * kDblEps;
    double depth;
    d3 n, p_BP;
    if (clamped && d2 > eps * eps) {
      const double dist = sqrt(d2);
      n = p_CN / dist;
      depth = r - dist;
      p_BP = p_BN + n * (depth * 0.5);
    } else {
      double min_d = INFINITY;
      int ax = -1;
      for (int i = 0; i < 3; ++i) {
        const double q = comp(p_BC, i), h = comp(half, i);
        const double d = q >= 0 ? h - q : q + h;
        if (d + eps < min_d) { min_d = d; ax = i; }
      }
      n = d3{0, 0, 0};
      setComp(n, ax, comp(p_BC, ax) >= 0 ? -1 : 1);
      depth = min_d + r;
      p_BP = p_BC + n * ((r - min_d) / 2);
    }
    out->c[out->n++] = {mul(X_FB.R, n), apply(X_FB, p_BP), depth};
  }
  return true;
}

// sphere_cylinder-inl.h:48-150
MFCL_D bool primSphereCylinder(double rs, const tf3& X_FS, double cr, double clz, const tf3& X_FC, PContacts* out) {
  const tf3 X_CS = relTf(X_FC, X_FS);
  const d3 p_CS = X_CS.t;
  d3 p_CN = p_CS;
  bool outside = false;
  const double hh = clz / 2;
  if (p_CS.z > hh) { outside = true; p_CN.z = hh; }
  else if (p_CS.z < -hh) { outside = true; p_CN.z = -hh; }
  {
    double rx = p_CS.x, ry = p_CS.y;
    const double sq = rx * rx + ry * ry;
    if (sq > cr * cr) {
      outside = true;
      const double k = cr / sqrt(sq);
      rx *= k;
      ry *= k;
      p_CN.x = rx;
      p_CN.y = ry;
    }
  }
  const d3 p_SN = p_CN - p_CS;
  const double d2 = sqnorm(p_SN);
  if (d2 > rs * rs) return false;
  if (out) {
    const double eps = 16 * kDblEps;
    double depth;
    d3 n, p_CP;
    if (outside && d2 > eps * eps) {
      const double d = sqrt(d2);
      n = p_SN / d;
      depth = rs - d;
      p_CP = p_CN + n * (depth * 0.5);
    } else {
      const double face = p_CS.z >= 0 ? clz / 2 - p_CS.z : p_CS.z + clz / 2;
      const double dxy = sqrt(p_CS.x * p_CS.x + p_CS.y * p_CS.y);
      const double barrel = cr - dxy;
      if (barrel < face - eps) {
        if (dxy > eps) {
          n = d3{-p_CS.x / dxy, -p_CS.y / dxy, 0};
          depth = rs + barrel;
        } else {
          n = d3{-1, 0, 0};
          depth = rs + cr;
        }
        p_CP = p_CS + n * ((rs - barrel) / 2);
      } else {
        n = d3{0, 0, p_CS.z >= 0 ? -1.0 : 1.0};
        depth = face + rs;
        p_CP = p_CS + n * ((rs - face) / 2);
      }
    }
    out->c[out->n++] = {mul(X_FC.R, n), apply(X_FC, p_CP), depth};
  }
  return true;
}

// box_box-inl.h:76-136: Sutherland-Hodgman clip of the incident quad against
// the reference rectangle [-h0,h0] x [-h1,h1]; at most 8 output points.
MFCL_D int rectQuadClip(const double h[2], const double quad[8], double ret[16]) {
  double bufA[16], bufB[16];
  int nq = 4;
  for (int i = 0; i < 8; ++i) bufA[i] = quad[i];
  double* q = bufA;
  double* r = bufB;
  for (int dir = 0; dir <= 1; ++dir) {
    for (int sign = -1; sign <= 1; sign += 2) {
      int nr = 0;
      for (int i = 0; i < nq; ++i) {
        const double* pq = q + 2 * i;
        const double* nx = (i + 1 < nq) ? pq + 2 : q;
        const bool in_p = sign * pq[dir] < h[dir];
        if (in_p) {
          r[2 * nr] = pq[0];
          r[2 * nr + 1] = pq[1];
          if (++nr == 8) goto full;
        }
        if (in_p ^ (sign * nx[dir] < h[dir])) {
          r[2 * nr + 1 - dir] = pq[1 - dir] + (nx[1 - dir] - pq[1 - dir]) / (nx[dir] - pq[dir]) * (sign * h[dir] - pq[dir]);
          r[2 * nr + dir] = sign * h[dir];
          if (++nr == 8) goto full;
        }
      }
      {
        double* t = q;
        q = r;
        r = t;
        nq = nr;
      }
      continue;
    full:
      for (int i = 0; i < 16; ++i) ret[i] = r[i];
      return 8;
    }
  }
  for (int i = 0; i < 2 * nq; ++i) ret[i] = q[i];
  return nq;
}

// box_box-inl.h:139-211
MFCL_D void cullPts(int n, const double* p, int m, int i0, int* iret) {
  double a, cx, cy, q;
  if (n == 1) {
    cx = p[0];
    cy = p[1];
  } else if (n == 2) {
    cx = 0.5 * (p[0] + p[2]);
    cy = 0.5 * (p[1] + p[3]);
  } else {
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
    if (fabs(a + q) > kDblEps) a = 1 / (3 * (a + q));
    else a = 999999984306749440.0;  // 1e18f
    cx = a * (cx + q * (p[n * 2 - 2] + p[0]));
    cy = a * (cy + q * (p[n * 2 - 1] + p[1]));
  }
  double A[8];
  for (int i = 0; i < n; ++i) A[i] = atan2(p[i * 2 + 1] - cy, p[i * 2] - cx);
  int avail[8];
  for (int i = 0; i < n; ++i) avail[i] = 1;
  avail[i0] = 0;
  iret[0] = i0;
  const double pi = 3.14159265358979323846;
  for (int j = 1; j < m; ++j) {
    a = j * (2 * pi / m) + A[i0];
    if (a > pi) a -= 2 * pi;
    double maxdiff = 1e9;
    int pick = i0;
    for (int i = 0; i < n; ++i) {
      if (avail[i]) {
        double diff = fabs(A[i] - a);
        if (diff > pi) diff = 2 * pi - diff;
        if (diff < maxdiff) { maxdiff = diff; pick = i; }
      }
    }
    avail[pick] = 0;
    iret[j] = pick;
  }
}

// box_box-inl.h:214-803 (maxc = 4) and :823-841; returns the ODE code.
MFCL_D int primBoxBox(d3 side1, const tf3& t1, d3 side2, const tf3& t2, PContacts* out) {
  const m33& R1 = t1.R;
  const m33& R2 = t2.R;
  const d3 T1 = t1.t, T2 = t2.t;
  const double fudge = 1.05;
  const d3 p = T2 - T1;
  const d3 pp = tmul(R1, p);
  const d3 A = side1 * 0.5, B = side2 * 0.5;
  const m33 R = tmulm(R1, R2);
  double Q[3][3] = {{fabs(R.m00), fabs(R.m01), fabs(R.m02)},
                    {fabs(R.m10), fabs(R.m11), fabs(R.m12)},
                    {fabs(R.m20), fabs(R.m21), fabs(R.m22)}};
  double s = -1.7976931348623157e308, s2;
  int best_col = -1, which_R = 0, invert = 0, code = 0;
  d3 normalC = {0, 0, 0};
  // face axes of box 1 and box 2
  {
    const double tp[3] = {pp.x, pp.y, pp.z};
    const double a[3] = {A.x, A.y, A.z};
    for (int i = 0; i < 3; ++i) {
      const double t = tp[i];
      s2 = fabs(t) - (Q[i][0] * B.x + Q[i][1] * B.y + Q[i][2] * B.z + a[i]);
      if (s2 > 0) return 0;
      if (s2 > s) { s = s2; best_col = i; which_R = 1; invert = t < 0; code = 1 + i; }
    }
    const double b[3] = {B.x, B.y, B.z};
    for (int j = 0; j < 3; ++j) {
      const double t = dot(col(R2, j), p);
      s2 = fabs(t) - (Q[0][j] * A.x + Q[1][j] * A.y + Q[2][j] * A.z + b[j]);
      if (s2 > 0) return 0;
      if (s2 > s) { s = s2; best_col = j; which_R = 2; invert = t < 0; code = 4 + j; }
    }
  }
  {
    const double amax = fmax(fmax(A.x, fmax(A.y, A.z)), fmax(B.x, fmax(B.y, B.z)));
    const double sc = fmax(amax, 1.0) * 10 * kDblEps;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Q[i][j] += sc;
  }
  const double Rm[3][3] = {{R.m00, R.m01, R.m02}, {R.m10, R.m11, R.m12}, {R.m20, R.m21, R.m22}};
  const double a_[3] = {A.x, A.y, A.z}, b_[3] = {B.x, B.y, B.z}, p_[3] = {pp.x, pp.y, pp.z};
  // edge axes u_i x v_j, i = edge of box 1, j = edge of box 2 (box_box-inl.h:360-500)
  for (int i = 0; i < 3; ++i) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    for (int j = 0; j < 3; ++j) {
      const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      const double t = p_[i2] * Rm[i1][j] - p_[i1] * Rm[i2][j];
      const double a1 = i1 < i2 ? a_[i1] : a_[i2], a2 = i1 < i2 ? a_[i2] : a_[i1];
      const double qa1 = i1 < i2 ? Q[i2][j] : Q[i1][j], qa2 = i1 < i2 ? Q[i1][j] : Q[i2][j];
      const double b1 = j1 < j2 ? b_[j1] : b_[j2], b2 = j1 < j2 ? b_[j2] : b_[j1];
      const double qb1 = j1 < j2 ? Q[i][j2] : Q[i][j1], qb2 = j1 < j2 ? Q[i][j1] : Q[i][j2];
      s2 = fabs(t) - (a1 * qa1 + a2 * qa2 + b1 * qb1 + b2 * qb2);
      if (s2 > 0) return 0;
      double nv[3] = {0, 0, 0};
      nv[i1] = -Rm[i2][j];
      nv[i2] = Rm[i1][j];
      const d3 n = {nv[0], nv[1], nv[2]};
      const double l = norm(n);
      if (l > kDblEps) {
        s2 /= l;
        if (s2 * fudge > s) { s = s2; best_col = -1; normalC = n / l; invert = t < 0; code = 7 + 3 * i + j; }
      }
    }
  }
  if (!code) return 0;
  d3 normal = best_col != -1 ? col(which_R == 1 ? R1 : R2, best_col) : mul(R1, normalC);
  if (invert) normal = -normal;
  const double depth = -s;
  if (!out) return code;
  if (code > 6) {
    d3 pa = T1;
    for (int j = 0; j < 3; ++j) pa = pa + col(R1, j) * (comp(A, j) * (dot(col(R1, j), normal) > 0 ? 1.0 : -1.0));
    d3 pb = T2;
    for (int j = 0; j < 3; ++j) pb = pb + col(R2, j) * (comp(B, j) * (dot(col(R2, j), normal) > 0 ? -1.0 : 1.0));
    const d3 ua = col(R1, (code - 7) / 3), ub = col(R2, (code - 7) % 3);
    const d3 pd = pb - pa;
    const double uaub = dot(ua, ub), q1 = dot(ua, pd), q2 = -dot(ub, pd);
    double d = 1 - uaub * uaub, alpha = 0, beta = 0;
    if (!(d <= static_cast<double>(0.0001f))) {
      d = 1 / d;
      alpha = (q1 + uaub * q2) * d;
      beta = (uaub * q1 + q2) * d;
    }
    pa = pa + ua * alpha;
    pb = pb + ub * beta;
    out->c[out->n++] = {normal, (pa + pb) * 0.5, depth};
    return code;
  }
  const bool ref1 = code <= 3;
  const m33& Ra = ref1 ? R1 : R2;
  const m33& Rb = ref1 ? R2 : R1;
  const d3 pa = ref1 ? T1 : T2, pb = ref1 ? T2 : T1;
  const d3 Sa = ref1 ? A : B, Sb = ref1 ? B : A;
  const d3 normal2 = ref1 ? normal : -normal;
  const d3 nr = tmul(Rb, normal2);
  const double anr[3] = {fabs(nr.x), fabs(nr.y), fabs(nr.z)};
  int lanr, a1, a2;
  if (anr[1] > anr[0]) {
    if (anr[1] > anr[2]) { a1 = 0; lanr = 1; a2 = 2; }
    else { a1 = 0; a2 = 1; lanr = 2; }
  } else {
    if (anr[0] > anr[2]) { lanr = 0; a1 = 1; a2 = 2; }
    else { a1 = 0; a2 = 1; lanr = 2; }
  }
  const d3 center = comp(nr, lanr) < 0 ? pb - pa + col(Rb, lanr) * comp(Sb, lanr)
                                        : pb - pa - col(Rb, lanr) * comp(Sb, lanr);
  const int codeN = ref1 ? code - 1 : code - 4;
  const int code1 = codeN == 0 ? 1 : 0;
  const int code2 = codeN == 2 ? 1 : 2;
  const double c1 = dot(col(Ra, code1), center), c2 = dot(col(Ra, code2), center);
  double m11 = dot(col(Rb, a1), col(Ra, code1));
  double m12 = dot(col(Rb, a2), col(Ra, code1));
  double m21 = dot(col(Rb, a1), col(Ra, code2));
  double m22 = dot(col(Rb, a2), col(Ra, code2));
  const double k1 = m11 * comp(Sb, a1), k2 = m21 * comp(Sb, a1), k3 = m12 * comp(Sb, a2), k4 = m22 * comp(Sb, a2);
  const double quad[8] = {c1 - k1 - k3, c2 - k2 - k4, c1 - k1 + k3, c2 - k2 + k4,
                          c1 + k1 + k3, c2 + k2 + k4, c1 + k1 - k3, c2 + k2 - k4};
  const double rect[2] = {comp(Sa, code1), comp(Sa, code2)};
  double ret[16];
  const int n_int = rectQuadClip(rect, quad, ret);
  if (n_int < 1) return code;
  d3 points[8];
  double dep[8];
  const double det1 = static_cast<double>(1.f) / (m11 * m22 - m12 * m21);
  m11 *= det1; m12 *= det1; m21 *= det1; m22 *= det1;
  int cnum = 0;
  for (int j = 0; j < n_int; ++j) {
    const double kk1 = m22 * (ret[j * 2] - c1) - m12 * (ret[j * 2 + 1] - c2);
    const double kk2 = -m21 * (ret[j * 2] - c1) + m11 * (ret[j * 2 + 1] - c2);
    points[cnum] = center + col(Rb, a1) * kk1 + col(Rb, a2) * kk2;
    dep[cnum] = comp(Sa, codeN) - dot(normal2, points[cnum]);
    if (dep[cnum] >= 0) {
      ret[cnum * 2] = ret[j * 2];
      ret[cnum * 2 + 1] = ret[j * 2 + 1];
      cnum++;
    }
  }
  if (cnum < 1) return code;
  int maxc = 4;
  if (maxc > cnum) maxc = cnum;
  int iret[8] = {0, 1, 2, 3, 4, 5, 6, 7};
  if (cnum > maxc) {
    int i1 = 0;
    double md = dep[0];
    for (int i = 1; i < cnum; ++i)
      if (dep[i] > md) { md = dep[i]; i1 = i; }
    cullPts(cnum, ret, maxc, i1, iret);
    cnum = maxc;
  }
  for (int j = 0; j < cnum; ++j) {
    const int i = iret[j];
    const d3 pw = code < 4 ? points[i] + pa + normal * (dep[i] / 2) : points[i] + pa - normal * (dep[i] / 2);
    out->c[out->n++] = {normal, pw, dep[i]};
  }
  return code;
}

}  // namespace mfcl
