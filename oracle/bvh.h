// CPU oracle -- TEST INFRASTRUCTURE ONLY (see omath.h).
//
// OBBRSS mesh-mesh collision and the broadphase AABB tree:
//   BVH build        include/fcl/geometry/bvh/BVH_model-inl.h:397-568,
//                    detail/BV_fitter-inl.h:324-346, detail/BV_splitter-inl.h:411-500,
//                    include/fcl/math/geometry-inl.h:148-180,237-365,713-780
//   OBB overlap      include/fcl/math/bv/OBB-inl.h:306-436, OBBRSS-inl.h:131-137
//   tri-tri SAT      narrowphase/detail/traversal/collision/intersect-inl.h:594-790,1082-1103
//   traversal        narrowphase/detail/traversal/collision/bvh_solver-inl.h:75-164
//   AABB tree        include/fcl/broadphase/binary_AABB_tree-inl.h:70-195,400-500
#pragma once

#include <cstdint>
#include <vector>

#include "omath.h"

namespace oracle {

struct OBB {
  M3 axis;  // columns are the box axes
  V3 To;    // centre
  V3 extent;
};

struct BVNode {
  OBB obb;
  int first_child{0};  // < 0: leaf with primitive -(first_child + 1)
  int first_primitive{0};
  int num_primitives{0};
};

struct Mesh {
  std::vector<V3> verts;
  std::vector<int> tris;  // 3 per triangle
  std::vector<BVNode> nodes;
};

void buildMeshBVH(Mesh& m, const float* xyz, int n_verts, const int* tris, int n_tris);

bool obbDisjoint(const M3& B, const V3& T, const V3& a, const V3& b);
bool obbOverlap(const M3& R0, const V3& T0, const OBB& b1, const OBB& b2);
bool triTriIntersect(const V3& P1, const V3& P2, const V3& P3, const V3& Q1, const V3& Q2,
                     const V3& Q3, const M3& R, const V3& T);

// A contact of a penetration-enabled mesh query (world frame as the
// reference computes it, see triTriContacts).
struct MeshContact {
  int b1, b2;
  V3 pos, normal;
  double depth;
};
struct MeshCollideOut {
  std::vector<std::pair<int, int>> contacts;  // (b1, b2) in discovery order (one per contact)
  std::vector<MeshContact> pen;               // penetration mode: the contacts themselves
  int64_t bv_tests{0};
  int64_t leaf_tests{0};
};
// max_contacts <= 0: collect every contact.  penetration: contacts carry
// position, normal and depth (intersect-inl.h:796-842), and a colliding
// triangle pair adds as many contacts as deepest points it yields (0-2).
void meshCollide(const Mesh& m1, const Tf& tf1, const Mesh& m2, const Tf& tf2, int max_contacts,
                 MeshCollideOut& out, bool penetration = false);
// intersect-inl.h:724-842: the 17-axis test and, when the triangles
// intersect, up to 2 contact points (P's frame), depth and normal.
bool triTriContacts(const V3& P1, const V3& P2, const V3& P3, const V3& Q1in, const V3& Q2in, const V3& Q3in,
                    const M3& R, const V3& T, V3 pts[2], int* npts, double* depth, V3* normal);

// Broadphase AABB tree.
struct AABBd {
  V3 mn, mx;
};
struct AABBTree {
  struct Node {
    AABBd bv;
    int child[2]{-1, -1};
    int64_t user_id{-1};
    bool leaf{false};
  };
  std::vector<Node> nodes;
  int root{-1};
};
void buildAABBTree(AABBTree& t, const float* mn, const float* mx, const int64_t* ids, int n);
bool aabbOverlap(const AABBd& a, const AABBd& b);
// Leaves whose boxes overlap the query box, in the reference's visit order.
void aabbTreeQuery(const AABBTree& t, const AABBd& q, std::vector<int64_t>& hits,
                   int64_t* nodes_visited = nullptr);

}  // namespace oracle
