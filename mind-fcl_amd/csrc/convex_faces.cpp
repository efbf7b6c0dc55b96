// Convex-hull faces -> the vertex-adjacency walk data of fcl::Convex
// (host side of mfcl_upload_convex_faces).
//
// The reference builds, per Convex, the sorted neighbour list of every
// vertex from its faces (Convex::FindVertexNeighbors, convex-inl.h:377-409),
// enables the support walk when the hull has more than 32 vertices and the
// mesh is watertight (every edge shared by exactly two faces) and connected
// (every vertex on a face) (ValidateTopology, convex-inl.h:293-375), and
// seeds each walk from the naive extreme vertices of the six axis directions
// (initExtremeViaNeighborCache, convex-inl.h:151-199).  The same tables are
// built here, once per distinct face list: shapes that share a face range
// (the reference's shared_ptr<const std::vector<int>> faces) share the
// adjacency in device memory; the six seed vertices depend on the vertex
// coordinates and are computed per shape on the device (computeWalkSeeds).
//
// Device layout: one int32 array `nbr`; a topology with n vertices occupies
// nbr[o .. o + n + sum(1 + degree)) in the reference's encoding (entry v is
// the local offset of vertex v's record [count, neighbours...]), followed by
// 32 ints of padding for the walk's fixed-size record reads.  Per shape:
// nbr offset o (-1: naive scans) and 8 ints (6 seed vertices, 2 unused).
#include <algorithm>
#include <map>
#include <string>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>

#include "ctx.h"

namespace mfcl {

namespace {

struct Topo {
  std::vector<int32_t> nbr;  // reference encoding, local indices
  bool walk{false};
};

// Returns false (with a message) when a face references a vertex outside
// [0, nv) or has fewer than 3 vertices.
bool buildTopo(const int32_t* faces, uint32_t num_faces, uint64_t face_len, uint32_t nv, Topo& t,
               std::string& err) {
  // neighbour pairs (both directions) and undirected face edges
  std::vector<std::pair<int32_t, int32_t>> adj, edges;
  uint64_t k = 0;
  for (uint32_t f = 0; f < num_faces; ++f) {
    if (k >= face_len) {
      err = "face list shorter than its face count";
      return false;
    }
    const int32_t cnt = faces[k];
    if (cnt < 3 || k + 1 + static_cast<uint64_t>(cnt) > face_len) {
      err = "face " + std::to_string(f) + " has an invalid vertex count";
      return false;
    }
    int32_t prev = faces[k + cnt];
    for (int32_t i = 1; i <= cnt; ++i) {
      const int32_t v = faces[k + i];
      if (v < 0 || static_cast<uint32_t>(v) >= nv || prev < 0 || static_cast<uint32_t>(prev) >= nv) {
        err = "face " + std::to_string(f) + " references a vertex outside the hull";
        return false;
      }
      adj.emplace_back(v, prev);
      adj.emplace_back(prev, v);
      edges.emplace_back(std::min(v, prev), std::max(v, prev));
      prev = v;
    }
    k += 1 + static_cast<uint64_t>(cnt);
  }
  std::sort(adj.begin(), adj.end());
  adj.erase(std::unique(adj.begin(), adj.end()), adj.end());
  std::sort(edges.begin(), edges.end());
  t.nbr.assign(nv, 0);
  t.nbr.reserve(nv + nv + adj.size());
  bool connected = true;
  size_t a = 0;
  for (uint32_t v = 0; v < nv; ++v) {
    t.nbr[v] = static_cast<int32_t>(t.nbr.size());
    const size_t a0 = a;
    while (a < adj.size() && adj[a].first == static_cast<int32_t>(v)) ++a;
    t.nbr.push_back(static_cast<int32_t>(a - a0));
    for (size_t j = a0; j < a; ++j) t.nbr.push_back(adj[j].second);
    if (a == a0) connected = false;
  }
  // every distinct edge must appear in exactly two faces
  bool watertight = true;
  for (size_t i = 0; i < edges.size();) {
    size_t j = i;
    while (j < edges.size() && edges[j] == edges[i]) ++j;
    if (j - i != 2) watertight = false;
    i = j;
  }
  t.walk = nv > 32 && watertight && connected;
  return true;
}

}  // namespace

int uploadConvexFaces(mfcl_ctx* ctx, const int32_t* faces, const uint64_t* face_off, const uint32_t* num_faces,
                      uint64_t faces_len) {
  const uint32_t n = ctx->n_shapes;
  std::vector<int64_t> off(n, -1);
  std::vector<int32_t> all;
  // distinct (face offset, face count, vertex count) -> offset in `all` (-1: no walk)
  std::map<std::tuple<uint64_t, uint32_t, uint32_t>, int64_t> seen;
  for (uint32_t i = 0; i < n; ++i) {
    if (ctx->h_type[i] != MFCL_SHAPE_CONVEX || num_faces[i] == 0) continue;
    const uint32_t nv = ctx->h_vcnt[i];
    if (face_off[i] > faces_len) {
      ctx->err = "convex shape " + std::to_string(i) + ": face offset beyond the face array";
      return MFCL_E_INVALID;
    }
    const auto key = std::make_tuple(face_off[i], num_faces[i], nv);
    auto it = seen.find(key);
    if (it == seen.end()) {
      Topo t;
      std::string e;
      if (!buildTopo(faces + face_off[i], num_faces[i], faces_len - face_off[i], nv, t, e)) {
        ctx->err = "convex shape " + std::to_string(i) + ": " + e;
        return MFCL_E_INVALID;
      }
      int64_t o = -1;
      if (t.walk) {
        o = static_cast<int64_t>(all.size());
        all.insert(all.end(), t.nbr.begin(), t.nbr.end());
      }
      it = seen.emplace(key, o).first;
    }
    off[i] = it->second;
  }
  // the walk data (record offsets, walk blocks) is stored per vertex, so
  // walking shapes whose vertex ranges overlap must be the same hull with
  // the same adjacency
  {
    std::vector<std::tuple<uint32_t, uint32_t, int64_t, uint32_t>> r;
    for (uint32_t i = 0; i < n; ++i)
      if (off[i] >= 0) r.emplace_back(ctx->h_voff[i], ctx->h_vcnt[i], off[i], i);
    std::sort(r.begin(), r.end());
    uint64_t end = 0;
    for (size_t k = 0; k < r.size(); ++k) {
      const auto [vo, vc, to, id] = r[k];
      if (k > 0) {
        const auto [po, pc, pt, pid] = r[k - 1];
        const bool same = vo == po && vc == pc && to == pt;
        if (!same && vo < end) {
          ctx->err = "convex shapes " + std::to_string(pid) + " and " + std::to_string(id) +
                     " share vertices but not their faces (the support walk needs one adjacency per vertex)";
          return MFCL_E_INVALID;
        }
      }
      end = std::max<uint64_t>(end, static_cast<uint64_t>(vo) + vc);
    }
  }
  if (all.size() > 0x7FFFFFFFull) {
    ctx->err = "adjacency tables exceed 2^31 entries";
    return MFCL_E_INVALID;
  }
  // fixed-size reads of a neighbour record may run past the last table
  all.resize(all.size() + 32, 0);
  int rc;
  if ((rc = uploadVec(ctx, reinterpret_cast<void**>(&ctx->d_nbr), all.data(), all.size() * sizeof(int32_t))))
    return rc;
  if ((rc = uploadVec(ctx, reinterpret_cast<void**>(&ctx->d_nbr_off), off.data(), off.size() * sizeof(int64_t))))
    return rc;
  if ((rc = uploadVec(ctx, reinterpret_cast<void**>(&ctx->d_seed), nullptr, 8 * sizeof(int32_t) * n))) return rc;
  // walk blocks: 128 B per vertex; when the device cannot hold them the
  // walks read the adjacency table instead (same results, slower)
  if (ctx->d_wblk) (void)hipFree(ctx->d_wblk);
  ctx->d_wblk = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&ctx->d_wblk), 8 * sizeof(float4) * std::max<uint64_t>(ctx->n_verts_alloc, 1)) !=
      hipSuccess) {
    ctx->d_wblk = nullptr;
    (void)hipGetLastError();
  }
  return computeWalkSeeds(ctx);
}

}  // namespace mfcl
