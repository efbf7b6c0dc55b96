// CPU oracle -- TEST INFRASTRUCTURE ONLY (see omath.h).
//
// Double-precision restatement of mind-fcl's convex narrowphase:
//   shapes + support mapping  include/fcl/cvx_collide/gjk_shape.hpp:41-180,
//                             include/fcl/geometry/shape/convex-inl.h:111-290
//   MinkowskiDiff             include/fcl/cvx_collide/minkowski_diff.hpp:14-85
//   GJK boolean               include/fcl/cvx_collide/gjk.hpp:11-363
//   GJK min distance          include/fcl/cvx_collide/gjk_distance.hpp:11-559
//   EPA                       include/fcl/cvx_collide/epa.hpp, epa_polytope*.hpp,
//                             epa_simplex2polytope.hpp, epa_polytope_utils.h
//   MPR boolean               include/fcl/cvx_collide/mpr.hpp:34-495
//   GJKSolver dispatch        include/fcl/narrowphase/detail/gjk_solver-inl.h:71-148,762-874
//   closed-form primitives    include/fcl/narrowphase/detail/primitive_shape_algorithm/
#pragma once

#include <array>
#include <vector>

#include "omath.h"

namespace oracle {

// Shape type codes follow cvx_collide::GJKShapeType (include/fcl/cvx_collide/gjk_shape.h:12-28).
enum ShapeType : int { kBox = 0, kSphere = 1, kCapsule = 3, kCylinder = 5, kConvex = 100, kTriangle = 101 };

// A convex hull: vertices plus the vertex adjacency used by the reference's
// hill-climbing support search (include/fcl/geometry/shape/convex.h:212-259).
struct Hull {
  std::vector<V3> verts;
  // Compact adjacency: first V entries are offsets into `nbr`; at each offset
  // the neighbour count followed by the neighbour indices.
  std::vector<int> nbr;
  bool walk{false};                // edge walking enabled (> 32 vertices and watertight)
  std::array<V3, 6> cache_dir;     // +x -x +y -y +z -z
  std::array<int, 6> cache_idx{};  // naive argmax for each axis direction
  V3 interior;                     // vertex mean
};

// Build a hull from vertices and triangle faces (each face: 3 vertex ids).
void buildHull(Hull& h, const float* xyz, int n_verts, const int* tris, int n_tris);

struct Shape {
  int type{kBox};
  // box: full side lengths; sphere: radius; capsule/cylinder: radius, full length.
  double p[3]{0, 0, 0};
  const Hull* hull{nullptr};
  V3 tri[3];  // triangle (TriangleP): its vertices
};

V3 shapeSupport(const Shape& s, const V3& d);
V3 shapeInterior(const Shape& s);
int hullExtremeNaive(const Hull& h, const V3& d);
int hullExtremeWalk(const Hull& h, const V3& d);

struct MinkDiff {
  Shape s0, s1;
  M3 toshape1;  // rotation from shape0 frame to shape1 frame
  Tf toshape0;  // transform from shape1 frame to shape0 frame
  V3 support0(const V3& d) const { return shapeSupport(s0, d); }
  V3 support1(const V3& d) const { return toshape0 * shapeSupport(s1, toshape1 * d); }
  V3 support(const V3& d) const { return support0(d) - support1(-d); }
  V3 interior() const { return shapeInterior(s0) - toshape0 * shapeInterior(s1); }
};
MinkDiff makeMinkDiff(const Shape& a, const Tf& tfa, const Shape& b, const Tf& tfb);

struct MVert {
  V3 v;  // vertex on the Minkowski difference
  V3 d;  // direction used to obtain it
};

struct Simplex {
  MVert v[4];
  int rank{-1};
  void reset() { rank = -1; }
  void add(const MVert& m) {
    if (rank < 0) rank = 0;
    v[rank++] = m;
  }
};

enum class GJKStatus { Intersect, Separated, ConvergeNoProgress, IterationLimit, Failed };

struct GJKDistOut {
  bool cert_valid{false};
  bool sep_valid{false};
  V3 p0, p1;  // both in shape-0 frame
};

struct GJKStats {
  int iterations{0};  // boolean loop iterations
  int dist_iterations{0};
  int support_calls{0};
};

GJKStatus gjkEvaluate(const MinkDiff& md, Simplex& simplex, const V3& guess, double tol,
                      int max_iter, GJKDistOut* dist_out, GJKStats* stats = nullptr);

enum class EPAStatus { Failed, OK, Touching, IterationLimit, MallocFailed };
struct EPAStats {
  int iterations{0};
};
EPAStatus epaEvaluate(const Simplex& simplex, const MinkDiff& md, double tol, int max_faces,
                      int max_iter, double* depth, V3* p0, V3* p1, EPAStats* stats = nullptr);

enum class MPRStatus { Intersect, Separated, Failed };
MPRStatus mprIntersect(const MinkDiff& md, int max_iter, double tol);

// MPR penetration along a given unit direction d (shape-0 frame)
// (MPR::RunDirectedPenetration, include/fcl/cvx_collide/mpr.hpp:496-632,
// finalizeDirectedPenetrationResult :642-704); status codes follow
// MPR::DirectedPenetrationStatus (mpr.h:45-50).
enum class DirectedStatus { OK = 0, Failed = 1, FailedNoIntersect = 2, FailedRefinementIterationLimit = 3 };
struct DirectedOut {
  double distance{0};
  V3 portal_normal;
  V3 p0, p1;  // shape-0 frame
  int iterations{0};  // portal refinement iterations
};
DirectedStatus mprDirected(const MinkDiff& md, const V3& d, int max_iter, double tol, DirectedOut& out);

// MPR minimum penetration refined from a unit hint direction d (shape-0
// frame) (MPR::RunIncrementalMinimumPenetrationDistance with
// return_on_subroutine_converge, include/fcl/cvx_collide/
// mpr_incremental_penetration.hpp:10-337); status codes follow
// MPR::IncrementalPenetrationStatus (mpr.h:76-82).
enum class IncrementalStatus { OK = 0, Failed = 1, FailedNoIntersect = 2, DirectedPenetrationFailed = 3,
                               IterationLimit = 4 };
struct IncrementalOut {
  double depth{0};
  V3 direction;  // penetration direction (shape-0 frame)
  V3 p0, p1;     // shape-0 frame
  int iterations{0};  // outer iterations
};
IncrementalStatus mprIncremental(const MinkDiff& md, const V3& d, int max_iter, double tol, IncrementalOut& out);

// Mesh-shape leaf tests (a triangle of a mesh posed by tf_tri against a
// shape posed by tf_shape, penetration disabled): GJKSolver::
// shapeTriangleIntersect (include/fcl/narrowphase/detail/gjk_solver-inl.h:
// 545-563) -> sphereTriangleIntersect (primitive_shape_algorithm/
// sphere_triangle-inl.h:107-175), boxTriangleIntersect (box_triangle-inl.h:
// 72-137), else MPR on (shape, TriangleP) (gjk_solver-inl.h:452-480).
bool sphereTriangleHit(double r, const Tf& tf_sphere, const V3& P1, const V3& P2, const V3& P3);
bool boxTriangleHit(const double side[3], const Tf& tf_box, const V3& P1, const V3& P2, const V3& P3,
                    const Tf& tf_tri);
bool shapeTriangleHit(const Shape& s, const Tf& tf_shape, const V3 tri[3], const Tf& tf_tri, int max_iter,
                      double tol);

// ---------------------------------------------------------------------------
// GJKSolver-level queries (include/fcl/narrowphase/detail/gjk_solver-inl.h).
struct SolverParams {
  int gjk_max_iterations{128};
  double gjk_tolerance{1e-6};
  int epa_max_face_num{256};
  int epa_max_iterations{255};
  double epa_tolerance{1e-6};
  bool guess_valid{false};
  V3 guess{1, 0, 0};
};

struct ContactPt {
  V3 normal;
  V3 pos;
  double depth{0};
};

// One collide call with max_contacts and penetration mode, as fcl::collide on
// two primitive/convex geometries.  Returns the number of contacts written
// (penetration disabled: a hit writes one contact with zero fields).
struct CollideOut {
  bool hit{false};
  int n{0};
  ContactPt c[4];
  int method{0};  // 0 = closed form, 1 = GJK/EPA, 2 = MPR
  int gjk_status{-1};
  int epa_status{-1};
  int gjk_iters{0};
  int epa_iters{0};
};
void collide(const Shape& a, const Tf& tfa, const Shape& b, const Tf& tfb, int max_contacts,
             bool penetration, const SolverParams& sp, CollideOut& out);
// Mesh-shape leaf tests with penetration enabled: sphereTriangleIntersect's
// contact (sphere_triangle-inl.h:107-187, triangle in the world) and the
// generic GJK + EPA contact on (shape, TriangleP) (gjk_solver-inl.h:452-527,
// 545-563).  The contact is before the traversal's normal reversal.
bool sphereTriangleContact(double r, const Tf& tf_sphere, const V3& P1, const V3& P2, const V3& P3,
                           ContactPt* c);
bool shapeTriangleContact(const Shape& s, const Tf& tf_shape, const V3 tri[3], const Tf& tf_tri,
                          const SolverParams& sp, ContactPt& c, int* gjk_status = nullptr,
                          int* epa_status = nullptr);

// shapeSignedDistance semantics (gjk_solver-inl.h:820-874): returns false on
// failure.  distance > 0 separated (witnesses in world), < 0 = -depth.
struct SignedDistOut {
  bool ok{false};
  double distance{-1};
  V3 pa, pb;
  int gjk_status{-1};
  int epa_status{-1};
  int gjk_iters{0};
  int epa_iters{0};
  int support_calls{0};
};
void signedDistance(const Shape& a, const Tf& tfa, const Shape& b, const Tf& tfb,
                    const SolverParams& sp, SignedDistOut& out);

// Closed forms (primitive_shape_algorithm/*).
bool sphereSphereIntersect(double r1, const Tf& tf1, double r2, const Tf& tf2,
                           std::vector<ContactPt>* c);
bool sphereCapsuleIntersect(double r1, const Tf& tf1, double r2, double lz2, const Tf& tf2,
                            std::vector<ContactPt>* c);
bool sphereBoxIntersect(double r, const Tf& tfs, const V3& side, const Tf& tfb,
                        std::vector<ContactPt>* c);
bool sphereCylinderIntersect(double r, const Tf& tfs, double cr, double clz, const Tf& tfc,
                             std::vector<ContactPt>* c);
bool boxBoxIntersect(const V3& side1, const Tf& tf1, const V3& side2, const Tf& tf2,
                     std::vector<ContactPt>* c);
// Segment-segment capsule distance (capsule_capsule-inl.h); used by tests as
// a closed-form check on GJK distance.
double capsuleCapsuleDistance(double r1, double lz1, const Tf& tf1, double r2, double lz2,
                              const Tf& tf2);

}  // namespace oracle
