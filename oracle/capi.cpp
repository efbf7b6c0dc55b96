// CPU oracle -- TEST INFRASTRUCTURE ONLY (see omath.h).
//
// Flat C entry points so that tests/ (via ctypes) and bench.py's cpu_baseline
// leg can drive the oracle.  Batched calls split the pairs into contiguous
// static chunks over a std::thread pool, the way a caller parallelises the
// reference's re-entrant per-pair fcl::collide (README.md:13).
#include <chrono>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "bvh.h"
#include "narrowphase.h"

using namespace oracle;

namespace {

template <class F>
void parallelFor(int64_t n, int n_threads, F&& f) {
  if (n_threads <= 1 || n < 2) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + n_threads - 1) / n_threads;
  for (int t = 0; t < n_threads; ++t) {
    const int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([b, e, &f]() {
      for (int64_t i = b; i < e; ++i) f(i);
    });
  }
  for (auto& x : th) x.join();
}

Tf tfFrom12(const double* p) {
  Tf t;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t.R.m[i][j] = p[3 * i + j];
  t.t = V3(p[9], p[10], p[11]);
  return t;
}

struct HullSet {
  std::vector<Hull> hulls;
};

struct ShapeTable {
  const HullSet* hs;
  const int* type;
  const double* params;
  const int* hull_idx;
  Shape get(int i) const {
    Shape s;
    s.type = type[i];
    s.p[0] = params[3 * i];
    s.p[1] = params[3 * i + 1];
    s.p[2] = params[3 * i + 2];
    if (s.type == kConvex) s.hull = &hs->hulls[hull_idx[i]];
    return s;
  }
};

SolverParams solverFrom(const double* sp6) {
  SolverParams sp;
  if (sp6) {
    sp.gjk_max_iterations = static_cast<int>(sp6[0]);
    sp.gjk_tolerance = sp6[1];
    sp.epa_max_face_num = static_cast<int>(sp6[2]);
    sp.epa_max_iterations = static_cast<int>(sp6[3]);
    sp.epa_tolerance = sp6[4];
  }
  return sp;
}

struct MeshSet {
  std::vector<Mesh> meshes;
};

}  // namespace

extern "C" {

struct OrcCollideRec {
  int32_t hit, n, method, gjk_status, epa_status, gjk_iters, epa_iters, pad;
  double depth;
  double normal[3];
  double pos[3];
};

struct OrcSdRec {
  int32_t ok, gjk_status, epa_status, gjk_iters, epa_iters, support_calls;
  double distance;
  double pa[3];
  double pb[3];
};

int orc_hardware_threads() { return static_cast<int>(std::thread::hardware_concurrency()); }

void orc_pose_from_quat(const float* pose7, int64_t n, double* out12) {
  for (int64_t i = 0; i < n; ++i) {
    const float* p = pose7 + 7 * i;
    const Tf t = poseFromQuat(p[0], p[1], p[2], p[3], p[4], p[5], p[6]);
    double* o = out12 + 12 * i;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) o[3 * r + c] = t.R.m[r][c];
    o[9] = t.t.x;
    o[10] = t.t.y;
    o[11] = t.t.z;
  }
}

void* orc_hullset_create(const float* xyz, const int64_t* vert_off, const int* vert_cnt,
                         const int* tris, const int64_t* tri_off, const int* tri_cnt, int n_hulls,
                         int n_threads) {
  auto* hs = new HullSet();
  hs->hulls.resize(n_hulls);
  parallelFor(n_hulls, n_threads, [&](int64_t i) {
    buildHull(hs->hulls[i], xyz + 3 * vert_off[i], vert_cnt[i], tris ? tris + 3 * tri_off[i] : nullptr,
              tris ? tri_cnt[i] : 0);
  });
  return hs;
}
void orc_hullset_destroy(void* h) { delete static_cast<HullSet*>(h); }
int orc_hull_walks(void* h, int i) { return static_cast<HullSet*>(h)->hulls[i].walk ? 1 : 0; }

// Support point of shape i in its own frame along direction d (unit).
void orc_support(void* hs, const int* type, const double* params, const int* hull_idx, int i,
                 const double* d, double* out) {
  ShapeTable st{static_cast<HullSet*>(hs), type, params, hull_idx};
  const V3 s = shapeSupport(st.get(i), V3(d[0], d[1], d[2]));
  out[0] = s.x;
  out[1] = s.y;
  out[2] = s.z;
}

// fcl::collide over a batch of pairs.  Returns wall seconds of the query loop.
// contacts7 (may be NULL): every contact the call keeps, 4 slots of
// (normal xyz, pos xyz, depth) per pair in CollisionResult order.
double orc_collide_contacts_batch(void* hs, const int* type, const double* params, const int* hull_idx,
                                  const int* pa, const int* pb, const double* tfa12, const double* tfb12,
                                  int64_t n, int max_contacts, int penetration, const double* sp6,
                                  const double* guess3, int n_threads, OrcCollideRec* out, double* contacts7) {
  ShapeTable st{static_cast<HullSet*>(hs), type, params, hull_idx};
  const SolverParams base = solverFrom(sp6);
  const auto t0 = std::chrono::steady_clock::now();
  parallelFor(n, n_threads, [&](int64_t i) {
    SolverParams sp = base;
    if (guess3) {
      sp.guess_valid = true;
      sp.guess = V3(guess3[3 * i], guess3[3 * i + 1], guess3[3 * i + 2]);
    }
    CollideOut o;
    collide(st.get(pa[i]), tfFrom12(tfa12 + 12 * i), st.get(pb[i]), tfFrom12(tfb12 + 12 * i),
            max_contacts, penetration != 0, sp, o);
    OrcCollideRec& r = out[i];
    std::memset(&r, 0, sizeof(r));
    r.hit = o.hit;
    r.n = o.n;
    r.method = o.method;
    r.gjk_status = o.gjk_status;
    r.epa_status = o.epa_status;
    r.gjk_iters = o.gjk_iters;
    r.epa_iters = o.epa_iters;
    if (o.n > 0) {
      r.depth = o.c[0].depth;
      r.normal[0] = o.c[0].normal.x; r.normal[1] = o.c[0].normal.y; r.normal[2] = o.c[0].normal.z;
      r.pos[0] = o.c[0].pos.x; r.pos[1] = o.c[0].pos.y; r.pos[2] = o.c[0].pos.z;
    }
    if (contacts7) {
      double* c = contacts7 + 28 * i;
      for (int k = 0; k < 28; ++k) c[k] = 0;
      for (int k = 0; k < o.n && k < 4; ++k) {
        const ContactPt& x = o.c[k];
        double* d = c + 7 * k;
        d[0] = x.normal.x; d[1] = x.normal.y; d[2] = x.normal.z;
        d[3] = x.pos.x; d[4] = x.pos.y; d[5] = x.pos.z;
        d[6] = x.depth;
      }
    }
  });
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

double orc_collide_batch(void* hs, const int* type, const double* params, const int* hull_idx,
                         const int* pa, const int* pb, const double* tfa12, const double* tfb12,
                         int64_t n, int max_contacts, int penetration, const double* sp6,
                         const double* guess3, int n_threads, OrcCollideRec* out) {
  return orc_collide_contacts_batch(hs, type, params, hull_idx, pa, pb, tfa12, tfb12, n, max_contacts,
                                    penetration, sp6, guess3, n_threads, out, nullptr);
}

// fcl::collisionPenetration with the MPR modes on shape pairs
// (include/fcl/narrowphase/collision_penetration-inl.h:196-244): the
// boolean collide (penetration disabled), and for its contact the MPR
// search from dir3 (world unit vector, the escape direction of shape b)
// with MPR(128, tol) (computePenetrationMPR, :105-194).  Directed
// (incremental == 0): depth = distance along dir3, normal = dir3.
// Incremental: depth = the refined minimum penetration, normal = its
// direction in the world (an iteration-limit result counts as valid).
// Failures: depth -1, pos 0, normal = dir3.  pos = tf_a * the witness
// midpoint.  epa_status holds the MPR status, epa_iters its iterations.
double orc_mpr_penetration_batch(void* hs, const int* type, const double* params, const int* hull_idx,
                                 const int* pa, const int* pb, const double* tfa12, const double* tfb12,
                                 const double* dir3, int64_t n, int max_contacts, const double* sp6, double mpr_tol,
                                 int incremental, int n_threads, OrcCollideRec* out) {
  ShapeTable st{static_cast<HullSet*>(hs), type, params, hull_idx};
  const SolverParams sp = solverFrom(sp6);
  const auto t0 = std::chrono::steady_clock::now();
  parallelFor(n, n_threads, [&](int64_t i) {
    const Shape a = st.get(pa[i]), b = st.get(pb[i]);
    const Tf ta = tfFrom12(tfa12 + 12 * i), tb = tfFrom12(tfb12 + 12 * i);
    CollideOut c;
    collide(a, ta, b, tb, max_contacts, false, sp, c);
    OrcCollideRec& r = out[i];
    std::memset(&r, 0, sizeof(r));
    r.hit = c.hit;
    r.n = c.n;
    r.method = c.method;
    r.gjk_status = c.gjk_status;
    r.epa_status = -1;
    if (c.n == 0) return;
    const V3 dw(dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2]);
    const MinkDiff md = makeMinkDiff(a, ta, b, tb);
    const V3 dl = transpose(ta.R) * dw;
    bool ok;
    V3 p0, p1, nw = dw;
    double depth = 0;
    if (incremental) {
      IncrementalOut o;
      const IncrementalStatus s = mprIncremental(md, dl, 128, mpr_tol, o);
      r.epa_status = static_cast<int>(s);
      r.epa_iters = o.iterations;
      ok = s == IncrementalStatus::OK || s == IncrementalStatus::IterationLimit;
      if (ok) {
        depth = o.depth;
        p0 = o.p0;
        p1 = o.p1;
        nw = ta.R * o.direction;
      }
    } else {
      DirectedOut o;
      const DirectedStatus s = mprDirected(md, dl, 128, mpr_tol, o);
      r.epa_status = static_cast<int>(s);
      r.epa_iters = o.iterations;
      ok = s == DirectedStatus::OK;
      if (ok) {
        depth = o.distance;
        p0 = o.p0;
        p1 = o.p1;
      }
    }
    r.normal[0] = nw.x; r.normal[1] = nw.y; r.normal[2] = nw.z;
    if (ok) {
      r.depth = depth;
      const V3 pw = ta * (0.5 * (p0 + p1));
      r.pos[0] = pw.x; r.pos[1] = pw.y; r.pos[2] = pw.z;
    } else {
      r.depth = -1;
    }
  });
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// GJKSolver::shapeSignedDistance over a batch.  Returns wall seconds.
double orc_signed_distance_batch(void* hs, const int* type, const double* params,
                                 const int* hull_idx, const int* pa, const int* pb,
                                 const double* tfa12, const double* tfb12, int64_t n,
                                 const double* sp6, const double* guess3, int n_threads,
                                 OrcSdRec* out) {
  ShapeTable st{static_cast<HullSet*>(hs), type, params, hull_idx};
  const SolverParams base = solverFrom(sp6);
  const auto t0 = std::chrono::steady_clock::now();
  parallelFor(n, n_threads, [&](int64_t i) {
    SolverParams sp = base;
    if (guess3) {
      sp.guess_valid = true;
      sp.guess = V3(guess3[3 * i], guess3[3 * i + 1], guess3[3 * i + 2]);
    }
    SignedDistOut o;
    signedDistance(st.get(pa[i]), tfFrom12(tfa12 + 12 * i), st.get(pb[i]), tfFrom12(tfb12 + 12 * i),
                   sp, o);
    OrcSdRec& r = out[i];
    r.ok = o.ok;
    r.gjk_status = o.gjk_status;
    r.epa_status = o.epa_status;
    r.gjk_iters = o.gjk_iters;
    r.epa_iters = o.epa_iters;
    r.support_calls = o.support_calls;
    r.distance = o.distance;
    r.pa[0] = o.pa.x; r.pa[1] = o.pa.y; r.pa[2] = o.pa.z;
    r.pb[0] = o.pb.x; r.pb[1] = o.pb.y; r.pb[2] = o.pb.z;
  });
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Low-level single-pair algorithm calls for known-answer tests.
// mode 0: GJK boolean status; 1: MPR status; 2: GJK with distance (dist_out
// receives distance, p0 xyz, p1 xyz in shape-0 frame; returns status).
// The Minkowski difference is built as in the reference tests
// (toshape0 = tf1^-1 tf2, toshape1 = rotation of toshape0^-1).
int orc_pair_algorithm(void* hs, const int* type, const double* params, const int* hull_idx,
                       int ia, int ib, const double* tfa12, const double* tfb12, int mode,
                       int max_iter, double tol, const double* guess3, double* dist_out) {
  ShapeTable st{static_cast<HullSet*>(hs), type, params, hull_idx};
  const MinkDiff md = makeMinkDiff(st.get(ia), tfFrom12(tfa12), st.get(ib), tfFrom12(tfb12));
  const V3 guess = guess3 ? V3(guess3[0], guess3[1], guess3[2]) : V3(1, 0, 0);
  if (mode == 1) return static_cast<int>(mprIntersect(md, max_iter, tol));
  Simplex s;
  if (mode == 0) return static_cast<int>(gjkEvaluate(md, s, guess, tol, max_iter, nullptr));
  GJKDistOut d;
  const GJKStatus g = gjkEvaluate(md, s, guess, tol, max_iter, &d);
  if (dist_out) {
    dist_out[0] = d.sep_valid ? norm(d.p0 - d.p1) : -1.0;
    dist_out[1] = d.p0.x; dist_out[2] = d.p0.y; dist_out[3] = d.p0.z;
    dist_out[4] = d.p1.x; dist_out[5] = d.p1.y; dist_out[6] = d.p1.z;
  }
  return static_cast<int>(g);
}

double orc_capsule_capsule_distance(double r1, double lz1, const double* tf1, double r2, double lz2,
                                    const double* tf2) {
  return capsuleCapsuleDistance(r1, lz1, tfFrom12(tf1), r2, lz2, tfFrom12(tf2));
}

// ---------------------------------------------------------------- meshes
void* orc_meshset_create(const float* xyz, const int64_t* vert_off, const int* vert_cnt,
                         const int* tris, const int64_t* tri_off, const int* tri_cnt, int n_meshes,
                         int n_threads) {
  auto* ms = new MeshSet();
  ms->meshes.resize(n_meshes);
  parallelFor(n_meshes, n_threads, [&](int64_t i) {
    buildMeshBVH(ms->meshes[i], xyz + 3 * vert_off[i], vert_cnt[i], tris + 3 * tri_off[i], tri_cnt[i]);
  });
  return ms;
}
void orc_meshset_destroy(void* m) { delete static_cast<MeshSet*>(m); }

// Mesh-shape collision, penetration disabled (OrientedNodeBVHSolver::
// MeshShapeIntersect, include/fcl/narrowphase/detail/traversal/collision/
// bvh_solver-inl.h:6-72, leaf ShapeSimplexIntersect, shape_pair_intersect-
// inl.h:133-172): the colliding triangles of mesh `mesh` (posed tfm12)
// against shape `shape` (posed tfs12), by testing every triangle -- the set
// does not depend on the tree -- in ascending id order.  Writes at most cap
// ids, returns the count.
int64_t orc_mesh_shape_all(void* ms, void* hs, const int* type, const double* params, const int* hull_idx,
                           int mesh, int shape, const double* tfm12, const double* tfs12, const double* sp6,
                           int* out, int64_t cap) {
  const Mesh& m = static_cast<MeshSet*>(ms)->meshes[mesh];
  ShapeTable st{static_cast<HullSet*>(hs), type, params, hull_idx};
  const SolverParams sp = solverFrom(sp6);
  const Shape sh = st.get(shape);
  const Tf tm = tfFrom12(tfm12), ts = tfFrom12(tfs12);
  int64_t n = 0;
  const int nt = static_cast<int>(m.tris.size() / 3);
  for (int t = 0; t < nt; ++t) {
    const V3 tri[3] = {m.verts[m.tris[3 * t]], m.verts[m.tris[3 * t + 1]], m.verts[m.tris[3 * t + 2]]};
    if (shapeTriangleHit(sh, ts, tri, tm, sp.gjk_max_iterations, sp.gjk_tolerance)) {
      if (n < cap) out[n] = t;
      ++n;
    }
  }
  return n;
}

// The same over a batch of (mesh, shape) queries: hit flag and count.
double orc_mesh_shape_batch(void* ms, void* hs, const int* type, const double* params, const int* hull_idx,
                            const int* mesh, const int* shape, const double* tfm12, const double* tfs12, int64_t n,
                            const double* sp6, int n_threads, int* hit, int64_t* count) {
  const auto t0 = std::chrono::steady_clock::now();
  parallelFor(n, n_threads, [&](int64_t i) {
    count[i] = orc_mesh_shape_all(ms, hs, type, params, hull_idx, mesh[i], shape[i], tfm12 + 12 * i,
                                  tfs12 + 12 * i, sp6, nullptr, 0);
    hit[i] = count[i] > 0;
  });
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// One leaf test (for pins): shape `shape` posed tfs12 against the triangle
// P (9 doubles) posed tfm12.
int orc_shape_triangle_hit(void* hs, const int* type, const double* params, const int* hull_idx, int shape,
                           const double* tfs12, const double* P9, const double* tfm12, const double* sp6) {
  ShapeTable st{static_cast<HullSet*>(hs), type, params, hull_idx};
  const SolverParams sp = solverFrom(sp6);
  const V3 tri[3] = {V3(P9[0], P9[1], P9[2]), V3(P9[3], P9[4], P9[5]), V3(P9[6], P9[7], P9[8])};
  return shapeTriangleHit(st.get(shape), tfFrom12(tfs12), tri, tfFrom12(tfm12), sp.gjk_max_iterations,
                          sp.gjk_tolerance);
}
int orc_mesh_num_nodes(void* m, int i) {
  return static_cast<int>(static_cast<MeshSet*>(m)->meshes[i].nodes.size());
}

// Mesh-mesh collide (CollisionRequest(max_contacts), penetration disabled).
// contacts: n * max(1, max_contacts) * 2 int32 (b1, b2); with max_contacts
// <= 0 only the count and counters are written.  Returns wall seconds.
double orc_mesh_collide_batch(void* m, const int* pa, const int* pb, const double* tfa12,
                              const double* tfb12, int64_t n, int max_contacts, int n_threads,
                              int32_t* hit, int32_t* ncontacts, int32_t* contacts,
                              int64_t* bv_tests, int64_t* leaf_tests) {
  const auto* ms = static_cast<MeshSet*>(m);
  const auto t0 = std::chrono::steady_clock::now();
  parallelFor(n, n_threads, [&](int64_t i) {
    MeshCollideOut o;
    meshCollide(ms->meshes[pa[i]], tfFrom12(tfa12 + 12 * i), ms->meshes[pb[i]],
                tfFrom12(tfb12 + 12 * i), max_contacts, o);
    hit[i] = o.contacts.empty() ? 0 : 1;
    ncontacts[i] = static_cast<int32_t>(o.contacts.size());
    if (contacts && max_contacts > 0) {
      for (int k = 0; k < static_cast<int>(o.contacts.size()) && k < max_contacts; ++k) {
        contacts[(i * max_contacts + k) * 2] = o.contacts[k].first;
        contacts[(i * max_contacts + k) * 2 + 1] = o.contacts[k].second;
      }
    }
    if (bv_tests) bv_tests[i] = o.bv_tests;
    if (leaf_tests) leaf_tests[i] = o.leaf_tests;
  });
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct OrcMeshContact {
  int32_t b1, b2;
  double pos[3], normal[3], depth;
};

static void putContact(OrcMeshContact& c, const MeshContact& m) {
  c.b1 = m.b1;
  c.b2 = m.b2;
  c.pos[0] = m.pos.x; c.pos[1] = m.pos.y; c.pos[2] = m.pos.z;
  c.normal[0] = m.normal.x; c.normal[1] = m.normal.y; c.normal[2] = m.normal.z;
  c.depth = m.depth;
}

// Penetration-enabled mesh-mesh collide over a batch: per pair the number of
// contacts and the first max_contacts of them (CollisionResult order).
double orc_mesh_collide_pen_batch(void* m, const int* pa, const int* pb, const double* tfa12, const double* tfb12,
                                  int64_t n, int max_contacts, int n_threads, int32_t* ncontacts,
                                  OrcMeshContact* contacts) {
  const auto* ms = static_cast<MeshSet*>(m);
  const auto t0 = std::chrono::steady_clock::now();
  parallelFor(n, n_threads, [&](int64_t i) {
    MeshCollideOut o;
    meshCollide(ms->meshes[pa[i]], tfFrom12(tfa12 + 12 * i), ms->meshes[pb[i]], tfFrom12(tfb12 + 12 * i),
                max_contacts, o, true);
    ncontacts[i] = static_cast<int32_t>(o.pen.size());
    for (int k = 0; k < static_cast<int>(o.pen.size()) && k < max_contacts; ++k)
      putContact(contacts[i * max_contacts + k], o.pen[k]);
  });
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Every penetration contact of one mesh pair (full traversal), up to cap;
// returns the total count.
int64_t orc_mesh_collide_all_pen(void* m, int ia, int ib, const double* tfa12, const double* tfb12,
                                 OrcMeshContact* out, int64_t cap) {
  const auto* ms = static_cast<MeshSet*>(m);
  MeshCollideOut o;
  meshCollide(ms->meshes[ia], tfFrom12(tfa12), ms->meshes[ib], tfFrom12(tfb12), 0, o, true);
  for (int64_t k = 0; k < static_cast<int64_t>(o.pen.size()) && k < cap; ++k) putContact(out[k], o.pen[k]);
  return static_cast<int64_t>(o.pen.size());
}

// Penetration-enabled mesh-shape collide, every colliding triangle of one
// query in ascending order (the set does not depend on the tree): b1 = the
// triangle, b2 = -1 (Contact::NONE), the leaf contact with its normal
// reversed (ContactMeta::reverse_normal, bvh_solver-inl.h:55-65,
// shape_pair_intersect-inl.h:19-26).  Up to cap records; returns the count.
int64_t orc_mesh_shape_contacts_all(void* ms, void* hs, const int* type, const double* params, const int* hull_idx,
                                    int mesh, int shape, const double* tfm12, const double* tfs12,
                                    const double* sp6, OrcMeshContact* out, int64_t cap) {
  const Mesh& m = static_cast<MeshSet*>(ms)->meshes[mesh];
  ShapeTable st{static_cast<HullSet*>(hs), type, params, hull_idx};
  const SolverParams sp = solverFrom(sp6);
  const Shape sh = st.get(shape);
  const Tf tm = tfFrom12(tfm12), ts = tfFrom12(tfs12);
  int64_t n = 0;
  const int nt = static_cast<int>(m.tris.size() / 3);
  for (int t = 0; t < nt; ++t) {
    const V3 tri[3] = {m.verts[m.tris[3 * t]], m.verts[m.tris[3 * t + 1]], m.verts[m.tris[3 * t + 2]]};
    ContactPt c;
    if (!shapeTriangleContact(sh, ts, tri, tm, sp, c)) continue;
    if (n < cap) {
      MeshContact mc;
      mc.b1 = t;
      mc.b2 = -1;
      mc.pos = c.pos;
      mc.normal = -c.normal;
      mc.depth = c.depth;
      putContact(out[n], mc);
    }
    ++n;
  }
  return n;
}

// One leaf test with penetration (for pins): returns 1 on a hit and writes
// pos (3), normal (3, before the traversal's reversal), depth, GJK and EPA
// status into r9.
int orc_shape_triangle_contact(void* hs, const int* type, const double* params, const int* hull_idx, int shape,
                               const double* tfs12, const double* P9, const double* tfm12, const double* sp6,
                               double* r9) {
  ShapeTable st{static_cast<HullSet*>(hs), type, params, hull_idx};
  const SolverParams sp = solverFrom(sp6);
  const V3 tri[3] = {V3(P9[0], P9[1], P9[2]), V3(P9[3], P9[4], P9[5]), V3(P9[6], P9[7], P9[8])};
  ContactPt c;
  int gs = -1, es = -1;
  const bool hit = shapeTriangleContact(st.get(shape), tfFrom12(tfs12), tri, tfFrom12(tfm12), sp, c, &gs, &es);
  r9[0] = c.pos.x; r9[1] = c.pos.y; r9[2] = c.pos.z;
  r9[3] = c.normal.x; r9[4] = c.normal.y; r9[5] = c.normal.z;
  r9[6] = c.depth;
  r9[7] = gs;
  r9[8] = es;
  return hit ? 1 : 0;
}

// One triangle pair (Q given in its own frame, mapped by R, T):
// intersect-inl.h:594-613 with contacts.  Returns 1 when they intersect.
int orc_tri_tri_contacts(const double* P9, const double* Q9, const double* R9, const double* T3, double* pts6,
                         int* npts, double* depth, double* normal3) {
  M3 R;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R.m[i][j] = R9[3 * i + j];
  const V3 T(T3[0], T3[1], T3[2]);
  V3 pts[2], nrm;
  int np = 0;
  double d = 0;
  const bool hit = triTriContacts(V3(P9[0], P9[1], P9[2]), V3(P9[3], P9[4], P9[5]), V3(P9[6], P9[7], P9[8]),
                                  V3(Q9[0], Q9[1], Q9[2]), V3(Q9[3], Q9[4], Q9[5]), V3(Q9[6], Q9[7], Q9[8]), R, T,
                                  pts, &np, &d, &nrm);
  *npts = hit ? np : 0;
  *depth = d;
  normal3[0] = nrm.x; normal3[1] = nrm.y; normal3[2] = nrm.z;
  for (int k = 0; k < (hit ? np : 0); ++k) {
    pts6[3 * k] = pts[k].x; pts6[3 * k + 1] = pts[k].y; pts6[3 * k + 2] = pts[k].z;
  }
  return hit ? 1 : 0;
}

// All colliding triangle pairs of one mesh pair (full traversal), written as
// (b1, b2) up to cap entries; returns the total count.
int64_t orc_mesh_collide_all(void* m, int ia, int ib, const double* tfa12, const double* tfb12,
                             int32_t* out_pairs, int64_t cap) {
  const auto* ms = static_cast<MeshSet*>(m);
  MeshCollideOut o;
  meshCollide(ms->meshes[ia], tfFrom12(tfa12), ms->meshes[ib], tfFrom12(tfb12), 0, o);
  for (int64_t k = 0; k < static_cast<int64_t>(o.contacts.size()) && k < cap; ++k) {
    out_pairs[2 * k] = o.contacts[k].first;
    out_pairs[2 * k + 1] = o.contacts[k].second;
  }
  return static_cast<int64_t>(o.contacts.size());
}

// Every colliding triangle pair by exhaustive search (no BVH), with the same
// relative transform and triangle test the traversal uses; pins the BVH
// build and traversal (test/test_fcl_collision.cpp:915-1059 compares the
// BVH result against the same brute force).
int64_t orc_mesh_collide_brute(void* m, int ia, int ib, const double* tfa12, const double* tfb12,
                               int32_t* out_pairs, int64_t cap) {
  const auto* ms = static_cast<MeshSet*>(m);
  const Mesh& m1 = ms->meshes[ia];
  const Mesh& m2 = ms->meshes[ib];
  const Tf tf1 = tfFrom12(tfa12), tf2 = tfFrom12(tfb12);
  const M3 R = transpose(tf1.R) * tf2.R;
  const V3 T = tmul(tf1.R, tf2.t - tf1.t);
  const int n1 = static_cast<int>(m1.tris.size() / 3), n2 = static_cast<int>(m2.tris.size() / 3);
  int64_t k = 0;
  for (int p1 = 0; p1 < n1; ++p1) {
    const int* ta = &m1.tris[3 * p1];
    for (int p2 = 0; p2 < n2; ++p2) {
      const int* tb = &m2.tris[3 * p2];
      if (!triTriIntersect(m1.verts[ta[0]], m1.verts[ta[1]], m1.verts[ta[2]], m2.verts[tb[0]],
                           m2.verts[tb[1]], m2.verts[tb[2]], R, T))
        continue;
      if (k < cap) {
        out_pairs[2 * k] = p1;
        out_pairs[2 * k + 1] = p2;
      }
      ++k;
    }
  }
  return k;
}

// Brute-force triangle-triangle test of two triangles given in their own
// frames with the relative transform of the second into the first.
int orc_tri_tri(const double* P9, const double* Q9, const double* R9, const double* T3) {
  M3 R;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R.m[i][j] = R9[3 * i + j];
  return triTriIntersect(V3(P9[0], P9[1], P9[2]), V3(P9[3], P9[4], P9[5]), V3(P9[6], P9[7], P9[8]),
                         V3(Q9[0], Q9[1], Q9[2]), V3(Q9[3], Q9[4], Q9[5]), V3(Q9[6], Q9[7], Q9[8]),
                         R, V3(T3[0], T3[1], T3[2]))
             ? 1
             : 0;
}

int orc_obb_disjoint(const double* B9, const double* T3, const double* a3, const double* b3) {
  M3 B;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B.m[i][j] = B9[3 * i + j];
  return obbDisjoint(B, V3(T3[0], T3[1], T3[2]), V3(a3[0], a3[1], a3[2]), V3(b3[0], b3[1], b3[2])) ? 1 : 0;
}

// ------------------------------------------------------------ broadphase
void* orc_aabbtree_create(const float* mn, const float* mx, const int64_t* ids, int n) {
  auto* t = new AABBTree();
  buildAABBTree(*t, mn, mx, ids, n);
  return t;
}
void orc_aabbtree_destroy(void* t) { delete static_cast<AABBTree*>(t); }
// For each query box: up to cap hits (leaf user ids, reference visit order).
// Returns wall seconds.
double orc_aabbtree_query_batch(void* t, const float* qmn, const float* qmx, int64_t nq, int cap,
                                int n_threads, int32_t* counts, int64_t* ids, int64_t* visited) {
  const auto* tree = static_cast<AABBTree*>(t);
  const auto t0 = std::chrono::steady_clock::now();
  parallelFor(nq, n_threads, [&](int64_t i) {
    AABBd q;
    q.mn = V3(qmn[3 * i], qmn[3 * i + 1], qmn[3 * i + 2]);
    q.mx = V3(qmx[3 * i], qmx[3 * i + 1], qmx[3 * i + 2]);
    std::vector<int64_t> hits;
    int64_t v = 0;
    aabbTreeQuery(*tree, q, hits, &v);
    counts[i] = static_cast<int32_t>(hits.size());
    for (int k = 0; k < static_cast<int>(hits.size()) && k < cap; ++k) ids[i * cap + k] = hits[k];
    if (visited) visited[i] = v;
  });
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
