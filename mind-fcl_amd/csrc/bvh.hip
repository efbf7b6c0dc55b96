// OBB-tree mesh-mesh collision for gfx950.
//
// Host: one OBB tree per mesh, built top-down with covariance (PCA) boxes and
// a median split of triangle centroids along the box's major axis, laid out
// breadth-first (siblings adjacent) as 64-byte FP32 nodes; node boxes are
// inflated when rounded to FP32 so they still contain their triangles.
//
// Device: a persistent grid of wavefronts; each wavefront takes one query at
// a time from an atomic counter and runs the bounding-volume test tree
// (reference: narrowphase/detail/traversal/collision/bvh_solver-inl.h:75-164)
// 64 node pairs per step off a wave-private stack in device memory:
// FP32 OBB separating-axis tests with a rounding margin (never prunes a pair
// the FP64 test keeps), ballot/prefix compaction of children and contacts,
// and the 17-axis triangle test in FP64 with contraction off, so a leaf
// verdict equals the reference's double-precision verdict.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "ctx.h"
#include "dmath.h"
#include "obb_dev.h"

namespace mfcl {


// ------------------------------------------------------------ host build
namespace {

struct HV {
  double x, y, z;
};
inline HV hsub(HV a, HV b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double hdot(HV a, HV b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline HV hcross(HV a, HV b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (textbook;
// Golub & Van Loan 8.5).  Returns eigenvectors as columns of V.
void symEigen(double A[3][3], double w[3], double V[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = std::fabs(A[0][1]) + std::fabs(A[0][2]) + std::fabs(A[1][2]);
    if (off < 1e-300) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (std::fabs(A[p][q]) < 1e-300) continue;
        const double theta = (A[q][q] - A[p][p]) / (2 * A[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < 3; ++k) {
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < 3; ++i) w[i] = A[i][i];
}

struct Builder {
  const float* v;   // mesh vertices xyz
  const uint32_t* t;  // mesh triangles (local ids)
  std::vector<HV> cen;  // centroids
  std::vector<uint32_t> prim;   // triangle ids, leaf order once built
  std::vector<uint32_t> range;  // per node: first slot of prim it covers, count

  HV vert(uint32_t i) const { return {v[3 * i], v[3 * i + 1], v[3 * i + 2]}; }

  ObbNode fit(const uint32_t* idx, int n, HV axis_out[3]) const {
    HV m = {0, 0, 0};
    double S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const int np = 3 * n;
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        const HV p = vert(t[3 * idx[i] + k]);
        m.x += p.x; m.y += p.y; m.z += p.z;
        const double q[3] = {p.x, p.y, p.z};
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) S[a][b] += q[a] * q[b];
      }
    const double mm[3] = {m.x / np, m.y / np, m.z / np};
    double C[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) C[a][b] = S[a][b] / np - mm[a] * mm[b];
    double w[3], V[3][3];
    symEigen(C, w, V);
    int order[3] = {0, 1, 2};
    std::sort(order, order + 3, [&](int a, int b) { return w[a] > w[b]; });
    HV ax0 = {V[0][order[0]], V[1][order[0]], V[2][order[0]]};
    HV ax1 = {V[0][order[1]], V[1][order[1]], V[2][order[1]]};
    const double n0 = std::sqrt(hdot(ax0, ax0));
    ax0 = {ax0.x / n0, ax0.y / n0, ax0.z / n0};
    const double d01 = hdot(ax0, ax1);
    ax1 = {ax1.x - d01 * ax0.x, ax1.y - d01 * ax0.y, ax1.z - d01 * ax0.z};
    double n1 = std::sqrt(hdot(ax1, ax1));
    if (n1 < 1e-12) {  // degenerate spread: any perpendicular
      ax1 = std::fabs(ax0.x) < 0.9 ? hcross(ax0, HV{1, 0, 0}) : hcross(ax0, HV{0, 1, 0});
      n1 = std::sqrt(hdot(ax1, ax1));
    }
    ax1 = {ax1.x / n1, ax1.y / n1, ax1.z / n1};
    const HV ax2 = hcross(ax0, ax1);
    const HV axes[3] = {ax0, ax1, ax2};
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        const HV p = vert(t[3 * idx[i] + k]);
        for (int a = 0; a < 3; ++a) {
          const double pr = hdot(axes[a], p);
          lo[a] = std::min(lo[a], pr);
          hi[a] = std::max(hi[a], pr);
        }
      }
    ObbNode nd;
    HV c = {0, 0, 0};
    double emax = 0;
    for (int a = 0; a < 3; ++a) {
      const double mid = 0.5 * (lo[a] + hi[a]);
      c.x += axes[a].x * mid; c.y += axes[a].y * mid; c.z += axes[a].z * mid;
      emax = std::max(emax, 0.5 * (hi[a] - lo[a]));
    }
    // FP32 rounding of axes (<= 2^-24 each) and centre moves a contained point
    // by at most ~4 * 2^-24 * (|c| + |e|); pad by 2^-16 of that scale.
    const double pad = (std::fabs(c.x) + std::fabs(c.y) + std::fabs(c.z) + 3 * emax) * 1.5e-5 + 1e-12;
    for (int a = 0; a < 3; ++a) {
      nd.ax[3 * a] = static_cast<float>(axes[a].x);
      nd.ax[3 * a + 1] = static_cast<float>(axes[a].y);
      nd.ax[3 * a + 2] = static_cast<float>(axes[a].z);
      nd.e[a] = static_cast<float>(0.5 * (hi[a] - lo[a]) + pad);
      axis_out[a] = axes[a];
    }
    nd.c[0] = static_cast<float>(c.x);
    nd.c[1] = static_cast<float>(c.y);
    nd.c[2] = static_cast<float>(c.z);
    nd.child = 0;
    return nd;
  }

  // Breadth-first build: a FIFO of (node, first, count) ranges.
  void build(uint32_t n_tris, std::vector<ObbNode>& nodes) {
    cen.resize(n_tris);
    prim.resize(n_tris);
    for (uint32_t i = 0; i < n_tris; ++i) {
      const HV a = vert(t[3 * i]), b = vert(t[3 * i + 1]), c = vert(t[3 * i + 2]);
      cen[i] = {(a.x + b.x + c.x) / 3, (a.y + b.y + c.y) / 3, (a.z + b.z + c.z) / 3};
      prim[i] = i;
    }
    nodes.clear();
    nodes.reserve(2 * n_tris);
    struct R { uint32_t node, first, count; };
    std::vector<R> fifo;
    fifo.reserve(2 * n_tris);
    nodes.emplace_back();
    fifo.push_back({0, 0, n_tris});
    range.assign(2 * (2 * static_cast<size_t>(n_tris) - 1), 0);
    for (size_t h = 0; h < fifo.size(); ++h) {
      const R r = fifo[h];
      range[2 * r.node] = r.first;
      range[2 * r.node + 1] = r.count;
      HV axes[3];
      ObbNode nd = fit(prim.data() + r.first, static_cast<int>(r.count), axes);
      if (r.count == 1) {
        nd.child = -static_cast<int32_t>(prim[r.first]) - 1;
        nodes[r.node] = nd;
        continue;
      }
      const uint32_t half = r.count / 2;
      const HV ax = axes[0];
      std::nth_element(prim.begin() + r.first, prim.begin() + r.first + half, prim.begin() + r.first + r.count,
                       [&](uint32_t x, uint32_t y) { return hdot(ax, cen[x]) < hdot(ax, cen[y]); });
      const uint32_t left = static_cast<uint32_t>(nodes.size());
      nodes.emplace_back();
      nodes.emplace_back();
      nd.child = static_cast<int32_t>(left);
      nodes[r.node] = nd;
      fifo.push_back({left, r.first, half});
      fifo.push_back({left + 1, r.first + half, r.count - half});
    }
  }
};

}  // namespace

int buildMeshSet(std::string& err, const float* verts, const uint64_t* voff, const uint32_t* vcnt,
                 const uint32_t* tris, const uint64_t* toff, const uint32_t* tcnt, uint32_t n, MeshSetHost& m) {
  for (uint32_t i = 0; i < n; ++i) {
    if (tcnt[i] == 0 || vcnt[i] == 0) {
      err = "mesh " + std::to_string(i) + " is empty";
      return MFCL_E_INVALID;
    }
    for (uint64_t k = 0; k < 3ull * tcnt[i]; ++k)
      if (tris[3 * toff[i] + k] >= vcnt[i]) {
        err = "mesh " + std::to_string(i) + " has a triangle vertex id out of range";
        return MFCL_E_INVALID;
      }
  }
  std::vector<std::vector<ObbNode>> per(n);
  std::vector<std::vector<uint32_t>> rng(n), prm(n);
  const unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  {
    std::vector<std::thread> th;
    for (unsigned w = 0; w < nth; ++w)
      th.emplace_back([&, w]() {
        for (uint32_t i = w; i < n; i += nth) {
          Builder b{verts + 3 * voff[i], tris + 3 * toff[i], {}, {}, {}};
          b.build(tcnt[i], per[i]);
          rng[i] = std::move(b.range);
          prm[i] = std::move(b.prim);
        }
      });
    for (auto& x : th) x.join();
  }
  m = MeshSetHost();
  m.vert_base.resize(n);
  m.vert_count.assign(vcnt, vcnt + n);
  m.tri_base.resize(n);
  m.tri_count.assign(tcnt, tcnt + n);
  m.node_base.resize(n);
  m.node_count.resize(n);
  uint64_t nv = 0, nt = 0, nn = 0;
  for (uint32_t i = 0; i < n; ++i) {
    m.vert_base[i] = nv;
    m.tri_base[i] = static_cast<uint32_t>(nt);
    m.node_base[i] = static_cast<uint32_t>(nn);
    m.node_count[i] = static_cast<uint32_t>(per[i].size());
    nv += vcnt[i];
    nt += tcnt[i];
    nn += per[i].size();
  }
  if (nn > 0x7FFFFFFFull || nt > 0x7FFFFFFFull) {
    err = "mesh set too large (more than 2^31 nodes or triangles)";
    return MFCL_E_INVALID;
  }
  m.verts.resize(3 * nv);
  m.tris.resize(3 * nt);
  m.nodes.resize(nn * sizeof(ObbNode));
  m.node_range.resize(2 * nn);
  m.prim.resize(nt);
  for (uint32_t i = 0; i < n; ++i) {
    std::memcpy(m.verts.data() + 3 * m.vert_base[i], verts + 3 * voff[i], 12 * static_cast<size_t>(vcnt[i]));
    std::memcpy(m.tris.data() + 3 * static_cast<size_t>(m.tri_base[i]), tris + 3 * toff[i],
                12 * static_cast<size_t>(tcnt[i]));
    std::memcpy(m.nodes.data() + sizeof(ObbNode) * m.node_base[i], per[i].data(), per[i].size() * sizeof(ObbNode));
    std::memcpy(m.node_range.data() + 2 * static_cast<size_t>(m.node_base[i]), rng[i].data(), 8 * per[i].size());
    std::memcpy(m.prim.data() + m.tri_base[i], prm[i].data(), 4 * static_cast<size_t>(tcnt[i]));
  }
  return MFCL_OK;
}

// Nodes over more triangles than this are refitted by a whole wavefront.
constexpr uint32_t kRefitLaneMax = 32;

int uploadMeshSet(mfcl_ctx* ctx, const MeshSetHost& m) {
  const uint32_t n = static_cast<uint32_t>(m.node_base.size());
  const size_t nn = m.nodes.size() / sizeof(ObbNode), nt = m.prim.size();
  MeshHost h;
  h.node_base = m.node_base;
  h.node_count = m.node_count;
  h.tri_base = m.tri_base;
  h.tri_count = m.tri_count;
  h.vert_base = m.vert_base;
  h.vert_count = m.vert_count;
  h.large_base.resize(n);
  h.large_count.resize(n);
  std::vector<uint32_t> large;
  for (uint32_t i = 0; i < n; ++i) {
    h.large_base[i] = static_cast<uint32_t>(large.size());
    for (uint32_t k = 0; k < m.node_count[i]; ++k)
      if (m.node_range[2 * (static_cast<size_t>(m.node_base[i]) + k) + 1] > kRefitLaneMax)
        large.push_back(m.node_base[i] + k);
    h.large_count[i] = static_cast<uint32_t>(large.size()) - h.large_base[i];
  }
  // triangles with their vertices gathered, as the traversal reads them
  std::vector<float4> tri4(3 * nt);
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t k = 0; k < m.tri_count[i]; ++k)
      for (int j = 0; j < 3; ++j) {
        const size_t t = static_cast<size_t>(m.tri_base[i]) + k;
        const float* p = m.verts.data() + 3 * (m.vert_base[i] + m.tris[3 * t + j]);
        tri4[3 * t + j] = float4{p[0], p[1], p[2], 0.f};
      }
  auto up = [&](void** dst, const void* src, size_t bytes) -> int {
    MFCL_HIP_TRY(ctx, hipMalloc(dst, std::max<size_t>(bytes, 4)));
    if (bytes) MFCL_HIP_TRY(ctx, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return MFCL_OK;
  };
  int rc;
  if ((rc = up(&ctx->d_nodes, m.nodes.data(), m.nodes.size())) ||
      (rc = up(&ctx->d_tris, tri4.data(), tri4.size() * sizeof(float4))) ||
      (rc = up(reinterpret_cast<void**>(&ctx->d_mesh_node_base), m.node_base.data(), 4 * static_cast<size_t>(n))) ||
      (rc = up(reinterpret_cast<void**>(&ctx->d_mesh_node_count), m.node_count.data(), 4 * static_cast<size_t>(n))) ||
      (rc = up(reinterpret_cast<void**>(&ctx->d_mesh_tri_base), m.tri_base.data(), 4 * static_cast<size_t>(n))) ||
      (rc = up(reinterpret_cast<void**>(&ctx->d_mesh_verts), m.verts.data(), 4 * m.verts.size())) ||
      (rc = up(reinterpret_cast<void**>(&ctx->d_mesh_tri_idx), m.tris.data(), 4 * m.tris.size())) ||
      (rc = up(reinterpret_cast<void**>(&ctx->d_node_range), m.node_range.data(), 4 * m.node_range.size())) ||
      (rc = up(reinterpret_cast<void**>(&ctx->d_mesh_prim), m.prim.data(), 4 * m.prim.size())) ||
      (rc = up(reinterpret_cast<void**>(&ctx->d_refit_large), large.data(), 4 * large.size())))
    return rc;
  (void)nn;
  ctx->mesh_host = std::move(h);
  ctx->n_meshes = n;
  return MFCL_OK;
}

int buildAndUploadMeshes(mfcl_ctx* ctx, const float* verts, const uint64_t* voff, const uint32_t* vcnt,
                         const uint32_t* tris, const uint64_t* toff, const uint32_t* tcnt, uint32_t n) {
  MeshSetHost m;
  if (int rc = buildMeshSet(ctx->err, verts, voff, vcnt, tris, toff, tcnt, n, m)) return rc;
  return uploadMeshSet(ctx, m);
}

int downloadMeshSet(mfcl_ctx* ctx, MeshSetHost& m) {
  const MeshHost& h = ctx->mesh_host;
  const uint32_t n = ctx->n_meshes;
  m = MeshSetHost();
  m.vert_base = h.vert_base;
  m.vert_count = h.vert_count;
  m.tri_base = h.tri_base;
  m.tri_count = h.tri_count;
  m.node_base = h.node_base;
  m.node_count = h.node_count;
  const uint64_t nv = n ? h.vert_base[n - 1] + h.vert_count[n - 1] : 0;
  const uint64_t nt = n ? static_cast<uint64_t>(h.tri_base[n - 1]) + h.tri_count[n - 1] : 0;
  const uint64_t nn = n ? static_cast<uint64_t>(h.node_base[n - 1]) + h.node_count[n - 1] : 0;
  m.verts.resize(3 * nv);
  m.tris.resize(3 * nt);
  m.nodes.resize(nn * sizeof(ObbNode));
  m.node_range.resize(2 * nn);
  m.prim.resize(nt);
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  MFCL_HIP_TRY(ctx, hipMemcpy(m.verts.data(), ctx->d_mesh_verts, 4 * m.verts.size(), hipMemcpyDeviceToHost));
  MFCL_HIP_TRY(ctx, hipMemcpy(m.tris.data(), ctx->d_mesh_tri_idx, 4 * m.tris.size(), hipMemcpyDeviceToHost));
  MFCL_HIP_TRY(ctx, hipMemcpy(m.nodes.data(), ctx->d_nodes, m.nodes.size(), hipMemcpyDeviceToHost));
  MFCL_HIP_TRY(ctx, hipMemcpy(m.node_range.data(), ctx->d_node_range, 4 * m.node_range.size(), hipMemcpyDeviceToHost));
  MFCL_HIP_TRY(ctx, hipMemcpy(m.prim.data(), ctx->d_mesh_prim, 4 * m.prim.size(), hipMemcpyDeviceToHost));
  return MFCL_OK;
}

// ------------------------------------------------------------ refit
// BVHModel::beginReplaceModel / replaceSubModel / endReplaceModel(refit)
// (BVH_model-inl.h:306-375, refit 572-640) keep the tree and recompute
// every node's volume for the new vertices.  Here every node keeps its axes
// (FP32, as stored) and the triangles it covers, and gets the centre and
// half extents of their vertices' projections on those axes, padded by the
// same rule as the build, so it contains its triangles again; the
// traversal's pair sets stay exact because the leaf tests decide them.

// Centre and half extents from the projection ranges on the node's axes.
MFCL_D void refitWrite(ObbNode& nd, const double lo[3], const double hi[3]) {
  double cx = 0, cy = 0, cz = 0, emax = 0;
  for (int a = 0; a < 3; ++a) {
    const double mid = 0.5 * (lo[a] + hi[a]);
    cx += static_cast<double>(nd.ax[3 * a]) * mid;
    cy += static_cast<double>(nd.ax[3 * a + 1]) * mid;
    cz += static_cast<double>(nd.ax[3 * a + 2]) * mid;
    emax = fmax(emax, 0.5 * (hi[a] - lo[a]));
  }
  const double pad = (fabs(cx) + fabs(cy) + fabs(cz) + 3 * emax) * 1.5e-5 + 1e-12;
  for (int a = 0; a < 3; ++a) nd.e[a] = static_cast<float>(0.5 * (hi[a] - lo[a]) + pad);
  nd.c[0] = static_cast<float>(cx);
  nd.c[1] = static_cast<float>(cy);
  nd.c[2] = static_cast<float>(cz);
}

MFCL_D void refitAccum(const ObbNode& nd, const float4* tri, double lo[3], double hi[3]) {
  for (int j = 0; j < 3; ++j) {
    const float4 p = tri[j];
    for (int a = 0; a < 3; ++a) {
      const double pr = static_cast<double>(nd.ax[3 * a]) * p.x + static_cast<double>(nd.ax[3 * a + 1]) * p.y +
                        static_cast<double>(nd.ax[3 * a + 2]) * p.z;
      lo[a] = fmin(lo[a], pr);
      hi[a] = fmax(hi[a], pr);
    }
  }
}

// New vertices into the vertex copy and the gathered triangles.
__global__ void meshGatherKernel(const float* v, const uint32_t* idx, uint32_t n_tris, float4* tri4) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tris) return;
  for (int j = 0; j < 3; ++j) {
    const uint32_t k = idx[3 * t + j];
    tri4[3 * static_cast<uint64_t>(t) + j] = make_float4(v[3 * k], v[3 * k + 1], v[3 * k + 2], 0.f);
  }
}

// One lane per node over at most kRefitLaneMax triangles.
__global__ void refitLaneKernel(ObbNode* nodes, const uint32_t* range, const uint32_t* prim, const float4* tri4,
                                uint32_t node0, uint32_t n_nodes) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_nodes) return;
  const uint32_t g = node0 + k;
  const uint32_t first = range[2 * g], count = range[2 * g + 1];
  if (count > kRefitLaneMax) return;
  ObbNode nd = nodes[g];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < count; ++i) refitAccum(nd, tri4 + 3 * static_cast<uint64_t>(prim[first + i]), lo, hi);
  refitWrite(nd, lo, hi);
  for (int a = 0; a < 3; ++a) {
    nodes[g].c[a] = nd.c[a];
    nodes[g].e[a] = nd.e[a];
  }
}

// One wavefront per listed node.
__global__ void __launch_bounds__(64) refitWaveKernel(ObbNode* nodes, const uint32_t* range, const uint32_t* prim,
                                                      const float4* tri4, const uint32_t* list) {
  const uint32_t g = list[blockIdx.x];
  const uint32_t first = range[2 * g], count = range[2 * g + 1];
  const int l = threadIdx.x;
  ObbNode nd = nodes[g];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = l; i < count; i += 64) refitAccum(nd, tri4 + 3 * static_cast<uint64_t>(prim[first + i]), lo, hi);
  for (int off = 32; off > 0; off >>= 1)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fmin(lo[a], __shfl_xor(lo[a], off));
      hi[a] = fmax(hi[a], __shfl_xor(hi[a], off));
    }
  if (l != 0) return;
  refitWrite(nd, lo, hi);
  for (int a = 0; a < 3; ++a) {
    nodes[g].c[a] = nd.c[a];
    nodes[g].e[a] = nd.e[a];
  }
}

static int meshGather(mfcl_ctx* ctx, uint32_t mesh, const float* verts, bool device_ptr, hipStream_t s) {
  const MeshHost& h = ctx->mesh_host;
  float* vdst = ctx->d_mesh_verts + 3 * h.vert_base[mesh];
  const size_t vbytes = 12 * static_cast<size_t>(h.vert_count[mesh]);
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(vdst, verts, vbytes, device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                   s));
  const uint32_t tb = h.tri_base[mesh], nt = h.tri_count[mesh];
  float4* tri4 = static_cast<float4*>(ctx->d_tris) + 3 * static_cast<uint64_t>(tb);
  meshGatherKernel<<<(nt + 255) / 256, 256, 0, s>>>(vdst, ctx->d_mesh_tri_idx + 3 * static_cast<uint64_t>(tb), nt,
                                                    tri4);
  MFCL_HIP_TRY(ctx, hipGetLastError());
  return MFCL_OK;
}

int meshReplaceVertices(mfcl_ctx* ctx, uint32_t mesh, const float* verts, bool device_ptr, hipStream_t s) {
  if (int rc = meshGather(ctx, mesh, verts, device_ptr, s)) return rc;
  const MeshHost& h = ctx->mesh_host;
  const uint32_t tb = h.tri_base[mesh];
  float4* tri4 = static_cast<float4*>(ctx->d_tris) + 3 * static_cast<uint64_t>(tb);
  ObbNode* nodes = static_cast<ObbNode*>(ctx->d_nodes);
  const uint32_t nb = h.node_base[mesh], nn = h.node_count[mesh];
  refitLaneKernel<<<(nn + 255) / 256, 256, 0, s>>>(nodes, ctx->d_node_range, ctx->d_mesh_prim + tb, tri4, nb, nn);
  if (h.large_count[mesh])
    refitWaveKernel<<<h.large_count[mesh], 64, 0, s>>>(nodes, ctx->d_node_range, ctx->d_mesh_prim + tb, tri4,
                                                       ctx->d_refit_large + h.large_base[mesh]);
  MFCL_HIP_TRY(ctx, hipGetLastError());
  return MFCL_OK;
}

// endReplaceModel(refit = false): the tree is rebuilt on the host from the
// new vertices; it has the same node count, so it replaces the old one in
// place.
int meshRebuild(mfcl_ctx* ctx, uint32_t mesh, const float* verts_host) {
  const MeshHost& h = ctx->mesh_host;
  const uint32_t nt = h.tri_count[mesh], tb = h.tri_base[mesh], nb = h.node_base[mesh];
  std::vector<uint32_t> idx(3 * static_cast<size_t>(nt));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  MFCL_HIP_TRY(ctx, hipMemcpy(idx.data(), ctx->d_mesh_tri_idx + 3 * static_cast<uint64_t>(tb), 12 * static_cast<size_t>(nt),
                              hipMemcpyDeviceToHost));
  Builder b{verts_host, idx.data(), {}, {}, {}};
  std::vector<ObbNode> nodes;
  b.build(nt, nodes);
  if (nodes.size() != h.node_count[mesh]) {
    ctx->err = "rebuilt tree has a different node count";
    return MFCL_E_STATE;
  }
  MFCL_HIP_TRY(ctx, hipMemcpy(static_cast<ObbNode*>(ctx->d_nodes) + nb, nodes.data(), nodes.size() * sizeof(ObbNode),
                              hipMemcpyHostToDevice));
  MFCL_HIP_TRY(ctx, hipMemcpy(ctx->d_node_range + 2 * static_cast<uint64_t>(nb), b.range.data(), 4 * b.range.size(),
                              hipMemcpyHostToDevice));
  MFCL_HIP_TRY(ctx, hipMemcpy(ctx->d_mesh_prim + tb, b.prim.data(), 4 * static_cast<size_t>(nt), hipMemcpyHostToDevice));
  // the wavefront-refit list of this mesh may change with the partition
  std::vector<uint32_t> large;
  for (uint32_t k = 0; k < nodes.size(); ++k)
    if (b.range[2 * k + 1] > kRefitLaneMax) large.push_back(nb + k);
  if (large.size() != h.large_count[mesh]) {
    ctx->err = "rebuilt tree has a different refit layout";
    return MFCL_E_STATE;
  }
  if (!large.empty())
    MFCL_HIP_TRY(ctx, hipMemcpy(ctx->d_refit_large + h.large_base[mesh], large.data(), 4 * large.size(),
                                hipMemcpyHostToDevice));
  if (int rc = meshGather(ctx, mesh, verts_host, false, ctx->stream)) return rc;
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return MFCL_OK;
}

// --------------------------------------------------------------- device
struct MeshArgs {
  const ObbNode* nodes;
  const float4* tris;
  const uint32_t* node_base;
  const uint32_t* node_count;
  const uint32_t* tri_base;
  const uint32_t* pa;
  const uint32_t* pb;
  const float* posea;
  const float* poseb;
  uint64_t n;
  uint32_t n_meshes;
  int max_contacts;
  mfcl_mesh_result* out;
  int32_t* contacts;
  mfcl_mesh_contact* pen_contacts;  // penetration mode: max_contacts records per query
  uint64_t* stacks;
  uint32_t stack_cap;
  uint32_t* next_query;
  unsigned long long* totals;  // [0] bv tests, [1] leaf tests
};

struct F3 {
  float x, y, z;
};
MFCL_D F3 f3(const float* p) { return {p[0], p[1], p[2]}; }
MFCL_D float fdot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

#pragma clang fp contract(off)
// intersect-inl.h:1082-1103 and :594-790 in FP64 without FMA contraction.
MFCL_D bool project6d(d3 ax, d3 p1, d3 p2, d3 p3, d3 q1, d3 q2, d3 q3) {
  const double P1 = dot(ax, p1), P2 = dot(ax, p2), P3 = dot(ax, p3);
  const double Q1 = dot(ax, q1), Q2 = dot(ax, q2), Q3 = dot(ax, q3);
  const double mn1 = fmin(P1, fmin(P2, P3)), mx2 = fmax(Q1, fmax(Q2, Q3));
  if (mn1 > mx2) return false;
  const double mx1 = fmax(P1, fmax(P2, P3)), mn2 = fmin(Q1, fmin(Q2, Q3));
  if (mn2 > mx1) return false;
  return true;
}

MFCL_D bool triTriD(const float4* P, const float4* Qm, const m33& R, d3 T) {
  const d3 P1 = {P[0].x, P[0].y, P[0].z}, P2 = {P[1].x, P[1].y, P[1].z}, P3 = {P[2].x, P[2].y, P[2].z};
  const d3 Q1 = mul(R, d3{Qm[0].x, Qm[0].y, Qm[0].z}) + T;
  const d3 Q2 = mul(R, d3{Qm[1].x, Qm[1].y, Qm[1].z}) + T;
  const d3 Q3 = mul(R, d3{Qm[2].x, Qm[2].y, Qm[2].z}) + T;
  const d3 p1 = P1 - P1, p2 = P2 - P1, p3 = P3 - P1;
  const d3 q1 = Q1 - P1, q2 = Q2 - P1, q3 = Q3 - P1;
  const d3 e1 = p2 - p1, e2 = p3 - p2, e3 = p1 - p3;
  const d3 f1 = q2 - q1, f2 = q3 - q2, f3 = q1 - q3;
  const d3 n1 = cross(e1, e2);
  if (!project6d(n1, p1, p2, p3, q1, q2, q3)) return false;
  const d3 m1 = cross(f1, f2);
  if (!project6d(m1, p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(e1, f1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(e1, f2), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(e1, f3), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(e2, f1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(e2, f2), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(e2, f3), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(e3, f1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(e3, f2), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(e3, f3), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(e1, n1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(e2, n1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(e3, n1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(f1, m1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(f2, m1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6d(cross(f3, m1), p1, p2, p3, q1, q2, q3)) return false;
  return true;
}

MFCL_D void relPose(const float* pa, const float* pb, m33& R, d3& T) {
  const tf3 ta = poseFromQuat(pa), tb = poseFromQuat(pb);
  R = tmulm(ta.R, tb.R);
  T = tmul(ta.R, tb.t - ta.t);
}

__device__ __forceinline__ unsigned long long ballot64(bool p) { return __ballot(p); }
__device__ __forceinline__ int laneId() { return threadIdx.x & 63; }
__device__ __forceinline__ int prefixBelow(unsigned long long m) {
  return __popcll(m & ((1ull << laneId()) - 1ull));
}

// intersect-inl.h:1036-1051 with normalize() of geometry-inl.h:188-197.
MFCL_D bool buildTrianglePlaneD(d3 v1, d3 v2, d3 v3, d3& n, double& t) {
  d3 n_ = cross(v2 - v1, v3 - v1);
  const double sq = sqnorm(n_);
  if (!(sq > 0)) return false;
  const double l = sqrt(sq);
  n_ = d3{n_.x / l, n_.y / l, n_.z / l};
  n = n_;
  t = dot(n_, v1);
  return true;
}

// intersect-inl.h:845-882 (getEpsilon() = 1e-5, distanceToPlane = n.v - t).
MFCL_D void deepestPointsD(const d3* pts, d3 n, double t, double& depth, d3* deepest, int& nd) {
  constexpr double kEps = 1e-5;
  double max_depth = -1.7976931348623157e308;
  int k = 0, num_neg = 0, num_pos = 0, num_zero = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
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
  depth = max_depth;
  nd = k;
}

// Contacts of an intersecting triangle pair (intersect-inl.h:796-842): up to
// two points in the first mesh's frame, the depth and the normal.
MFCL_D void triContactsD(const float4* P, const float4* Qm, const m33& R, d3 T, d3 pts[2], int& npts, double& depth,
                         d3& normal) {
  const d3 P1 = {P[0].x, P[0].y, P[0].z}, P2 = {P[1].x, P[1].y, P[1].z}, P3 = {P[2].x, P[2].y, P[2].z};
  const d3 Q1 = mul(R, d3{Qm[0].x, Qm[0].y, Qm[0].z}) + T;
  const d3 Q2 = mul(R, d3{Qm[1].x, Qm[1].y, Qm[1].z}) + T;
  const d3 Q3 = mul(R, d3{Qm[2].x, Qm[2].y, Qm[2].z}) + T;
  d3 n1 = {0, 0, 0}, n2 = {0, 0, 0};
  double t1 = 0, t2 = 0;
  buildTrianglePlaneD(P1, P2, P3, n1, t1);
  buildTrianglePlaneD(Q1, Q2, Q3, n2, t2);
  const d3 Pv[3] = {P1, P2, P3}, Qv[3] = {Q1, Q2, Q3};
  d3 d1[3], d2[3];
  int k1 = 0, k2 = 0;
  double pd1 = 0, pd2 = 0;
  deepestPointsD(Qv, n1, t1, pd2, d2, k2);
  deepestPointsD(Pv, n2, t2, pd1, d1, k1);
  if (pd1 > pd2) {
    npts = k2 < 2 ? k2 : 2;
    pts[0] = d2[0];
    pts[1] = d2[1];
    normal = n1;
    depth = pd2;
  } else {
    npts = k1 < 2 ? k1 : 2;
    pts[0] = d1[0];
    pts[1] = d1[1];
    normal = -n2;
    depth = pd1;
  }
}

// One 256-thread block per query (queries from an atomic counter, a
// persistent grid): each step pops up to 256 node pairs of the bounding-
// volume test tree (bvh_solver-inl.h:75-164) off a block stack -- the top
// kMeshLds entries in LDS, the rest in a per-block device-memory region --
// tests them (FP32 OBB SAT with a rounding margin, FP64 triangle test), and
// pushes children and contacts with block-wide prefix sums.  A whole block
// per query cuts the number of dependent steps of the costliest queries,
// which set the launch's duration.  Descent rule as the reference (the
// larger inner node first); the pop order differs, so with a contact cap the
// kept contacts are a subset of the full set rather than the reference's
// first ones.  PEN: contacts with position, normal and depth (each
// colliding triangle pair adds 0-2 of them, shape_pair_intersect-inl.h:
// 226-254), and a query collides when it has a contact, as a
// CollisionResult with contacts does.
constexpr int kMeshBlock = 256;
constexpr int kMeshLds = 3072;  // LDS stack entries per block

// Exclusive prefix of v over the block (4 wavefronts) and the total; 64-bit
// so several 16-bit counters travel in one scan.
MFCL_D uint64_t blockScan(uint64_t v, uint64_t& total, uint64_t* wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint64_t base = 0;
  total = 0;
#pragma unroll
  for (int k = 0; k < kMeshBlock / 64; ++k) {
    const uint64_t c = wsum[k];
    if (k < w) base += c;
    total += c;
  }
  __syncthreads();
  return base + x - v;
}

template <bool PEN>
__global__ void __launch_bounds__(kMeshBlock, 3) meshKernel(MeshArgs A) {
  __shared__ uint64_t lstk[kMeshLds];
  __shared__ uint64_t wsum[kMeshBlock / 64];
  __shared__ uint32_t s_q;
  const int t = threadIdx.x;
  uint64_t* gstk = A.stacks + static_cast<uint64_t>(blockIdx.x) * A.stack_cap;
  const float kScaleErr = 16.0f * 5.9604645e-08f;
  unsigned long long tot_bv = 0, tot_leaf = 0;
  while (true) {
    if (t == 0) s_q = atomicAdd(A.next_query, 1u);
    __syncthreads();
    const uint32_t q = s_q;
    __syncthreads();
    if (q >= A.n) break;
    const uint32_t ma = A.pa[q], mb = A.pb[q];
    if (ma >= A.n_meshes || mb >= A.n_meshes) {
      if (t == 0) A.out[q] = mfcl_mesh_result{-1, -1, 0, 0};
      continue;
    }
    const uint32_t base1 = A.node_base[ma], base2 = A.node_base[mb];
    const uint32_t cnt1 = A.node_count[ma], cnt2 = A.node_count[mb];
    const float4* tri1 = A.tris + 3ull * A.tri_base[ma];
    const float4* tri2 = A.tris + 3ull * A.tri_base[mb];
    m33 R;
    d3 T;
    relPose(A.posea + 7ull * q, A.poseb + 7ull * q, R, T);
    const float Rq[9] = {(float)R.m00, (float)R.m01, (float)R.m02, (float)R.m10, (float)R.m11,
                         (float)R.m12, (float)R.m20, (float)R.m21, (float)R.m22};
    const float Tq[3] = {(float)T.x, (float)T.y, (float)T.z};
    tf3 tf2;
    if constexpr (PEN) tf2 = poseFromQuat(A.poseb + 7ull * q);
    // block-uniform traversal state, kept identical in every thread
    int top = 1, gtop = 0, ncont = 0, bv = 0, leaf = 0;
    bool overflow = false;
    if (t == 0) lstk[0] = 0ull;
    __syncthreads();
    int32_t* cout = A.contacts + 2ull * A.max_contacts * q;
    mfcl_mesh_contact* pout = A.pen_contacts + static_cast<uint64_t>(A.max_contacts) * q;
    while (top + gtop > 0 && ncont < A.max_contacts) {
      if (top == 0) {
        // refill the LDS stack from the device-memory part
        const int k = gtop < kMeshBlock ? gtop : kMeshBlock;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        if (t < k) lstk[t] = gstk[gtop - k + t];
        __syncthreads();
        gtop -= k;
        top = k;
      }
      const int k = top < kMeshBlock ? top : kMeshBlock;
      const bool act = t < k;
      const uint64_t e = act ? lstk[top - k + t] : 0ull;
      top -= k;
      const uint32_t n1 = static_cast<uint32_t>(e >> 32), n2 = static_cast<uint32_t>(e & 0xFFFFFFFFu);
      bool is_leafpair = false, push1 = false, push2 = false, bad = false;
      int32_t c1 = 0, c2 = 0;
      if (act) {
        bad = n1 >= cnt1 || n2 >= cnt2;
        if (!bad) {
          const ObbNode a = A.nodes[base1 + n1];
          const ObbNode b = A.nodes[base2 + n2];
          if (obbOverlapF(Rq, Tq, a, b, kScaleErr)) {
            const bool la = a.child < 0, lb = b.child < 0;
            c1 = a.child;
            c2 = b.child;
            if (la && lb) {
              is_leafpair = true;
            } else {
              const float sa = a.e[0] * a.e[0] + a.e[1] * a.e[1] + a.e[2] * a.e[2];
              const float sb = b.e[0] * b.e[0] + b.e[1] * b.e[1] + b.e[2] * b.e[2];
              if (lb || (!la && sa > sb)) push1 = true;
              else push2 = true;
            }
          }
        }
      }
      bv += k;
      // leaf tests (FP64); the number of contacts each lane adds
      int t1 = 0, t2 = 0, nadd = 0;
      d3 pts[2], nrm = {0, 0, 0};
      double depth = 0;
      if (is_leafpair) {
        t1 = -c1 - 1;
        t2 = -c2 - 1;
        if (triTriD(tri1 + 3 * t1, tri2 + 3 * t2, R, T)) {
          if constexpr (PEN) triContactsD(tri1 + 3 * t1, tri2 + 3 * t2, R, T, pts, nadd, depth, nrm);
          else nadd = 1;
        }
      }
      // one scan for four counters (16 bits each): contacts added, pushes,
      // leaf tests, bad node ids
      const bool pushes = push1 || push2;
      const uint64_t mine = static_cast<uint64_t>(nadd) | (static_cast<uint64_t>(pushes ? 1 : 0) << 16) |
                            (static_cast<uint64_t>(is_leafpair ? 1 : 0) << 32) |
                            (static_cast<uint64_t>(bad ? 1 : 0) << 48);
      uint64_t tot = 0;
      const uint64_t pre = blockScan(mine, tot, wsum);
      const int n_add = static_cast<int>(tot & 0xFFFF), n_push = static_cast<int>((tot >> 16) & 0xFFFF);
      leaf += static_cast<int>((tot >> 32) & 0xFFFF);
      if ((tot >> 48) & 0xFFFF) {
        overflow = true;
        break;
      }
      const int slot0 = ncont + static_cast<int>(pre & 0xFFFF);
      for (int j = 0; j < nadd; ++j) {
        const int slot = slot0 + j;
        if (slot >= A.max_contacts) break;
        if (A.contacts) {
          cout[2 * slot] = t1;
          cout[2 * slot + 1] = t2;
        }
        if constexpr (PEN) {
          // the reference maps them with the second mesh's transform
          // (shape_pair_intersect-inl.h:244-248 with tf1 = tf2 passed at :269)
          const d3 pw = apply(tf2, pts[j]), nw = mul(tf2.R, nrm);
          mfcl_mesh_contact c;
          c.b1 = t1;
          c.b2 = t2;
          c.pos[0] = pw.x; c.pos[1] = pw.y; c.pos[2] = pw.z;
          c.normal[0] = nw.x; c.normal[1] = nw.y; c.normal[2] = nw.z;
          c.depth = depth;
          pout[slot] = c;
        }
      }
      ncont += n_add;
      if (ncont > A.max_contacts) ncont = A.max_contacts;
      // pushes: two per expanding thread, LDS first, the excess to device memory
      const int ppos = 2 * static_cast<int>((pre >> 16) & 0xFFFF);
      const int total = 2 * n_push;
      if (gtop + (top + total > kMeshLds ? top + total - kMeshLds : 0) > static_cast<int>(A.stack_cap)) {
        overflow = true;
        break;
      }
      if (pushes) {
        uint64_t x, y;
        if (push1) {
          x = (static_cast<uint64_t>(c1) << 32) | n2;
          y = (static_cast<uint64_t>(c1 + 1) << 32) | n2;
        } else {
          x = (static_cast<uint64_t>(n1) << 32) | static_cast<uint32_t>(c2);
          y = (static_cast<uint64_t>(n1) << 32) | static_cast<uint32_t>(c2 + 1);
        }
        const uint64_t v2[2] = {x, y};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int pos = top + ppos + h;
          if (pos < kMeshLds) lstk[pos] = v2[h];
          else gstk[gtop + pos - kMeshLds] = v2[h];
        }
      }
      if (top + total > kMeshLds) {
        gtop += top + total - kMeshLds;
        top = kMeshLds;
      } else {
        top += total;
      }
      __syncthreads();
    }
    if (t == 0) {
      mfcl_mesh_result r;
      r.hit = overflow ? -1 : (ncont > 0 ? 1 : 0);
      r.n_contacts = overflow ? -1 : ncont;
      r.bv_tests = bv;
      r.leaf_tests = leaf;
      A.out[q] = r;
    }
    tot_bv += bv;
    tot_leaf += leaf;
    __syncthreads();
  }
  if (t == 0) {
    atomicAdd(&A.totals[0], tot_bv);
    atomicAdd(&A.totals[1], tot_leaf);
  }
}

__global__ void zeroU32(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// The traversal's workspace: 2 KB of counters, then one device-memory stack
// of mesh_stack entries per persistent block (mesh_blocks_per_cu blocks of
// 4 wavefronts per CU).
int reserveMesh(mfcl_ctx* ctx) {
  const uint32_t waves = static_cast<uint32_t>(ctx->n_cu * ctx->tune.mesh_blocks_per_cu * 4);
  const uint32_t cap = static_cast<uint32_t>(ctx->tune.mesh_stack);
  if (ctx->ws_arena.p && ctx->mesh_ws_waves == waves && ctx->mesh_ws_cap == cap) return MFCL_OK;
  ctx->mesh_ws_waves = ctx->mesh_ws_cap = 0;
  // one device-memory stack per block (4 wavefronts)
  if (int rc = ensure(ctx, ctx->ws_arena, 2048 + static_cast<size_t>(waves / 4) * cap * sizeof(uint64_t))) return rc;
  ctx->mesh_ws_waves = waves;
  ctx->mesh_ws_cap = cap;
  return MFCL_OK;
}

int meshCollideLaunch(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                      const float* poseb, uint64_t n, int32_t max_contacts, mfcl_mesh_result* out,
                      int32_t* contacts, mfcl_mesh_contact* pen_contacts, hipStream_t s) {
  if (n == 0) return MFCL_OK;
  if (ctx->n_meshes == 0) {
    ctx->err = "no meshes uploaded";
    return MFCL_E_STATE;
  }
  const uint32_t waves = ctx->mesh_ws_waves, cap = ctx->mesh_ws_cap;
  if (!ctx->ws_arena.p || waves == 0 || cap != static_cast<uint32_t>(ctx->tune.mesh_stack) ||
      waves != static_cast<uint32_t>(ctx->n_cu * ctx->tune.mesh_blocks_per_cu * 4)) {
    ctx->err = "mesh traversal stacks not sized for the current knobs";
    return MFCL_E_STATE;
  }
  const int blocks = static_cast<int>(waves / 4);  // one 256-thread block per query
  char* misc = static_cast<char*>(ctx->ws_arena.p);
  uint32_t* next = reinterpret_cast<uint32_t*>(misc + 1024);
  unsigned long long* totals = reinterpret_cast<unsigned long long*>(misc + 1536);
  MeshArgs A;
  A.nodes = static_cast<const ObbNode*>(ctx->d_nodes);
  A.tris = static_cast<const float4*>(ctx->d_tris);
  A.node_base = ctx->d_mesh_node_base;
  A.node_count = ctx->d_mesh_node_count;
  A.tri_base = ctx->d_mesh_tri_base;
  A.pa = pa;
  A.pb = pb;
  A.posea = posea;
  A.poseb = poseb;
  A.n = n;
  A.n_meshes = ctx->n_meshes;
  A.max_contacts = max_contacts;
  A.out = out;
  A.contacts = contacts;
  A.pen_contacts = pen_contacts;
  A.stacks = reinterpret_cast<uint64_t*>(misc + 2048);
  A.stack_cap = cap;
  A.next_query = next;
  A.totals = totals;
  zeroU32<<<1, 512, 0, s>>>(reinterpret_cast<uint32_t*>(misc), 512);
  if (ctx->timing) MFCL_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
  if (pen_contacts) meshKernel<true><<<blocks, kMeshBlock, 0, s>>>(A);
  else meshKernel<false><<<blocks, kMeshBlock, 0, s>>>(A);
  if (ctx->timing) MFCL_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  MFCL_HIP_TRY(ctx, hipGetLastError());
  ctx->d_counters = reinterpret_cast<uint64_t*>(totals);
  return MFCL_OK;
}

}  // namespace mfcl
