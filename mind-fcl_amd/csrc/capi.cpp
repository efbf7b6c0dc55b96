// Host side of the C ABI declared in include/mfcl.h: context lifetime,
// geometry upload, workspace management, and host-pointer convenience
// wrappers around the device-pointer launchers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "ctx.h"

namespace mfcl {

int ensure(mfcl_ctx* ctx, DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (b.bytes >= bytes) return MFCL_OK;
  if (b.p) {
    MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    MFCL_HIP_TRY(ctx, hipDeviceSynchronize());
    MFCL_HIP_TRY(ctx, hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  const size_t grown = bytes + bytes / 4;
  hipError_t e = hipMalloc(&b.p, grown);
  if (e != hipSuccess) {
    ctx->err = std::string("workspace allocation failed: ") + hipGetErrorString(e);
    b.p = nullptr;
    return MFCL_E_NOMEM;
  }
  b.bytes = grown;
  return MFCL_OK;
}

static void freePtr(void*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void freeShapes(mfcl_ctx* c) {
  freePtr(reinterpret_cast<void*&>(c->d_type));
  freePtr(reinterpret_cast<void*&>(c->d_param));
  freePtr(reinterpret_cast<void*&>(c->d_voff));
  freePtr(reinterpret_cast<void*&>(c->d_vcnt));
  freePtr(reinterpret_cast<void*&>(c->d_verts));
  freePtr(reinterpret_cast<void*&>(c->d_interior));
  freePtr(reinterpret_cast<void*&>(c->d_desc));
  freePtr(reinterpret_cast<void*&>(c->d_nbr));
  freePtr(reinterpret_cast<void*&>(c->d_nbr_off));
  freePtr(reinterpret_cast<void*&>(c->d_seed));
  freePtr(reinterpret_cast<void*&>(c->d_wblk));
  c->n_shapes = 0;
  c->n_verts_alloc = 0;
}

void freeMeshes(mfcl_ctx* c) {
  freePtr(c->d_nodes);
  freePtr(c->d_tris);
  freePtr(reinterpret_cast<void*&>(c->d_mesh_node_base));
  freePtr(reinterpret_cast<void*&>(c->d_mesh_node_count));
  freePtr(reinterpret_cast<void*&>(c->d_mesh_tri_base));
  freePtr(reinterpret_cast<void*&>(c->d_mesh_verts));
  freePtr(reinterpret_cast<void*&>(c->d_mesh_tri_idx));
  freePtr(reinterpret_cast<void*&>(c->d_node_range));
  freePtr(reinterpret_cast<void*&>(c->d_mesh_prim));
  freePtr(reinterpret_cast<void*&>(c->d_refit_large));
  c->n_meshes = 0;
  c->mesh_host = MeshHost();
}

void freeHeightmap(mfcl_ctx* c) {
  freePtr(reinterpret_cast<void*&>(c->d_hm));
  c->hm = HeightmapHost();
}

void freeAll(mfcl_ctx* c) {
  freeShapes(c);
  freeMeshes(c);
  freeHeightmap(c);
  freePtr(c->d_bp_nodes);
  freePtr(reinterpret_cast<void*&>(c->d_bp_ids));
  freePtr(c->ws_convex.p);
  freePtr(c->ws_arena.p);
  freePtr(c->ws_io.p);
  freePtr(c->ws_arm.p);
  c->ws_convex = c->ws_arena = c->ws_io = c->ws_arm = DevBuf();
  c->ws_convex_pairs = c->ws_arm_configs = 0;
  c->mesh_ws_waves = c->mesh_ws_cap = 0;
  c->d_counters = nullptr;
  c->d_qcounts = nullptr;
  mlpFree(c);
}

int reserveConvex(mfcl_ctx* ctx, uint64_t n) {
  if (n <= ctx->ws_convex_pairs && ctx->ws_convex.p) return MFCL_OK;
  // the counters of the last query live in this buffer
  ctx->d_counters = nullptr;
  ctx->d_qcounts = nullptr;
  const size_t arena = epaArenaBytes(kReservedArenaSlices, 256);
  if (int rc = ensure(ctx, ctx->ws_convex, convexWsBytes(n, arena))) return rc;
  ctx->ws_convex_pairs = convexWsMaxPairs(ctx->ws_convex.bytes, arena);
  return MFCL_OK;
}

// Host builders (bvh_build.cpp, broadphase.hip).
int buildAndUploadMeshes(mfcl_ctx* ctx, const float* verts, const uint64_t* voff, const uint32_t* vcnt,
                         const uint32_t* tris, const uint64_t* toff, const uint32_t* tcnt, uint32_t n);
int buildAndUploadAabbTree(mfcl_ctx* ctx, const float* mn, const float* mx, const int64_t* ids, uint32_t n);

}  // namespace mfcl

using namespace mfcl;

template <class T>
static int upload(mfcl_ctx* ctx, T** dst, const T* src, size_t count) {
  if (count == 0) count = 1;
  MFCL_HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(dst), sizeof(T) * count));
  if (src) MFCL_HIP_TRY(ctx, hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice));
  return MFCL_OK;
}

int mfcl::uploadVec(mfcl_ctx* ctx, void** dptr, const void* src, size_t bytes) {
  freePtr(*dptr);
  MFCL_HIP_TRY(ctx, hipMalloc(dptr, bytes ? bytes : 4));
  if (src && bytes) MFCL_HIP_TRY(ctx, hipMemcpy(*dptr, src, bytes, hipMemcpyHostToDevice));
  return MFCL_OK;
}

// Calls that synchronise, allocate or copy host memory cannot be recorded.
#define MFCL_NOT_WHILE_CAPTURING(ctx)                                              \
  do {                                                                             \
    if ((ctx) && (ctx)->capturing) {                                               \
      (ctx)->err = "not allowed while a query sequence is being recorded";         \
      return MFCL_E_STATE;                                                         \
    }                                                                              \
  } while (0)

struct mfcl_graph {
  mfcl_ctx* ctx;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" {

void mfcl_request_defaults(mfcl_request* req, int32_t mode) {
  std::memset(req, 0, sizeof(*req));
  req->mode = mode;
  req->max_contacts = 1;
  req->penetration = mode == MFCL_MODE_SIGNED_DISTANCE ? 1 : 0;
  req->gjk_max_iterations = 128;
  req->gjk_tolerance = 1e-6;
  req->epa_max_faces = 256;
  req->epa_max_iterations = 255;
  req->epa_tolerance = 1e-6;
  req->use_guess = 0;
}

int mfcl_create(mfcl_ctx** out, int device) {
  if (!out) return MFCL_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MFCL_E_HIP;
  if (device < 0 || device >= n) return MFCL_E_INVALID;
  auto* c = new (std::nothrow) mfcl_ctx();
  if (!c) return MFCL_E_NOMEM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return MFCL_E_HIP;
  }
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete c;
      return MFCL_E_HIP;
    }
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
    c->max_lds_per_block = lds;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    c->n_cu = cus;
  // the static workspace's address on this device (left null on failure:
  // mfcl_convex_query_static_dev then reports it)
  if (staticScratch(device, &c->d_static_scratch) != MFCL_OK) c->d_static_scratch = nullptr;
  *out = c;
  return MFCL_OK;
}

int mfcl_destroy(mfcl_ctx* ctx) {
  if (!ctx) return MFCL_E_INVALID;
  (void)hipSetDevice(ctx->device);
  if (ctx->capturing) {  // abandon an unfinished recording
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(ctx->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    ctx->capturing = false;
  }
  (void)hipStreamSynchronize(ctx->stream);
  freeAll(ctx);
  for (auto& e : ctx->ev) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MFCL_OK;
}

const char* mfcl_last_error(const mfcl_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mfcl_sync(mfcl_ctx* ctx) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  MFCL_HIP_TRY(ctx, hipDeviceSynchronize());
  return MFCL_OK;
}

int mfcl_set_timing(mfcl_ctx* ctx, int enabled) {
  if (!ctx) return MFCL_E_INVALID;
  ctx->timing = enabled != 0;
  return MFCL_OK;
}

int mfcl_last_timing(mfcl_ctx* ctx, float* ms6) {
  if (!ctx || !ms6) return MFCL_E_INVALID;
  std::memcpy(ms6, ctx->last_ms, sizeof(ctx->last_ms));
  return MFCL_OK;
}

static int refreshCounters(mfcl_ctx* ctx) {
  if (ctx->d_counters && ctx->d_qcounts) {
    MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    uint64_t dev[2] = {0, 0};
    uint32_t q[5] = {0, 0, 0, 0, 0};  // EPA queue count, overflow count, 2 cursors, continued in device memory
    MFCL_HIP_TRY(ctx, hipMemcpy(dev, ctx->d_counters, sizeof(dev), hipMemcpyDeviceToHost));
    MFCL_HIP_TRY(ctx, hipMemcpy(q, ctx->d_qcounts, sizeof(q), hipMemcpyDeviceToHost));
    ctx->counters[5] = q[4];
    ctx->counters[1] = dev[0] + dev[1];
    ctx->counters[2] = q[0];
    ctx->counters[3] = q[1];
    ctx->counters[4] = dev[1];
  }
  return MFCL_OK;
}

int mfcl_last_counters(mfcl_ctx* ctx, uint64_t* c4) {
  if (!ctx || !c4) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (int rc = refreshCounters(ctx)) return rc;
  std::memcpy(c4, ctx->counters, 4 * sizeof(uint64_t));
  return MFCL_OK;
}

int mfcl_last_counters_n(mfcl_ctx* ctx, uint64_t* c, int32_t n) {
  if (!ctx || !c || n < 0) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (int rc = refreshCounters(ctx)) return rc;
  for (int32_t k = 0; k < n; ++k) c[k] = k < 6 ? ctx->counters[k] : 0;
  return MFCL_OK;
}

int mfcl_tune(mfcl_ctx* ctx, int32_t knob, int32_t value) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  switch (knob) {
    case MFCL_TUNE_GJK_GROUP:
      if (value != 1 && value != 16 && value != 32 && value != 64) break;
      ctx->tune.gjk_group = value;
      return MFCL_OK;
    case MFCL_TUNE_EPA_THREADS:
      if (value < 256 || value % 256 != 0 || value > (1 << 20)) break;
      ctx->tune.epa_threads = value;
      return MFCL_OK;
    case MFCL_TUNE_MESH_BLOCKS_PER_CU:
      if (value < 1 || value > 16) break;
      ctx->tune.mesh_blocks_per_cu = value;
      return ctx->n_meshes ? reserveMesh(ctx) : MFCL_OK;
    case MFCL_TUNE_EPA_LDS_FACES:
      if (value < 16 || value > 128) break;
      ctx->tune.epa_lds_faces = value;
      return MFCL_OK;
    case MFCL_TUNE_GUESS_POLICY:
      if (value < 0 || value > 1) break;
      ctx->tune.guess_policy = value;
      return MFCL_OK;
    case MFCL_TUNE_EPA_SPILL:
      if (value < 0 || value > 1) break;
      ctx->tune.epa_spill = value;
      return MFCL_OK;
    case MFCL_TUNE_LANE_REFILL:
      if (value < 0 || value > 63) break;
      ctx->tune.lane_refill = value;
      return MFCL_OK;
    case MFCL_TUNE_MLP_WAVES:
      if (value != 4 && value != 8 && value != 16) break;
      ctx->tune.mlp_waves = value;
      return MFCL_OK;
    case MFCL_TUNE_MESH_STACK:
      if (value < 1024 || value > (1 << 18)) break;
      ctx->tune.mesh_stack = value;
      return ctx->n_meshes ? reserveMesh(ctx) : MFCL_OK;
    default:
      break;
  }
  ctx->err = "invalid tuning knob/value";
  return MFCL_E_INVALID;
}

int mfcl_graph_begin(mfcl_ctx* ctx) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MFCL_HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  ctx->timing_before_capture = ctx->timing;
  ctx->timing = false;
  return MFCL_OK;
}

int mfcl_graph_end(mfcl_ctx* ctx, mfcl_graph** out) {
  if (!ctx || !out) return MFCL_E_INVALID;
  *out = nullptr;
  if (!ctx->capturing) {
    ctx->err = "no query sequence is being recorded";
    return MFCL_E_STATE;
  }
  ctx->capturing = false;
  ctx->timing = ctx->timing_before_capture;
  hipGraph_t g = nullptr;
  MFCL_HIP_TRY(ctx, hipStreamEndCapture(ctx->stream, &g));
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    ctx->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
    return MFCL_E_HIP;
  }
  auto* r = new (std::nothrow) mfcl_graph{ctx, g, ex};
  if (!r) {
    (void)hipGraphExecDestroy(ex);
    (void)hipGraphDestroy(g);
    return MFCL_E_NOMEM;
  }
  *out = r;
  return MFCL_OK;
}

int mfcl_graph_launch(mfcl_graph* g, void* stream) {
  if (!g) return MFCL_E_INVALID;
  mfcl_ctx* ctx = g->ctx;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  MFCL_HIP_TRY(ctx, hipGraphLaunch(g->exec, s));
  return MFCL_OK;
}

int mfcl_graph_destroy(mfcl_graph* g) {
  if (!g) return MFCL_E_INVALID;
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  delete g;
  return MFCL_OK;
}

int mfcl_reserve(mfcl_ctx* ctx, int32_t kind, uint64_t count) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  switch (kind) {
    case MFCL_WS_CONVEX:
      if (count > 0xFFFFFFFFull) {
        ctx->err = "too many pairs for one query";
        return MFCL_E_INVALID;
      }
      return reserveConvex(ctx, count);
    case MFCL_WS_MLP:
      return mlpReserve(ctx, count);
    default:
      ctx->err = "unknown workspace kind";
      return MFCL_E_INVALID;
  }
}

int mfcl_reserved_pairs(mfcl_ctx* ctx, uint64_t* pairs) {
  if (!ctx || !pairs) return MFCL_E_INVALID;
  *pairs = ctx->ws_convex_pairs;
  return MFCL_OK;
}

int mfcl_dev_alloc(mfcl_ctx* ctx, uint64_t bytes, void** dptr) {
  if (!ctx || !dptr) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MFCL_HIP_TRY(ctx, hipMalloc(dptr, bytes ? bytes : 1));
  return MFCL_OK;
}
int mfcl_dev_free(mfcl_ctx* ctx, void* dptr) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipFree(dptr));
  return MFCL_OK;
}
int mfcl_h2d(mfcl_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return MFCL_OK;
}
int mfcl_d2h(mfcl_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  MFCL_HIP_TRY(ctx, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return MFCL_OK;
}

// 16-float shape descriptor for the warm-start MLP: local half-extents (3),
// vertex mean (3), vertex second moments about the mean xx yy zz xy xz yz
// (6), log2(vertex count) / 9, max |v|, mean |v|, and 1 for a hull (0 for a
// primitive, whose extents and moments are those of its solid box).
static void shapeDescriptor(int type, const float* p, const float* xyz, uint32_t n, float* d) {
  for (int k = 0; k < 16; ++k) d[k] = 0.f;
  if (type == MFCL_SHAPE_CONVEX) {
    double mx[3] = {0, 0, 0}, mean[3] = {0, 0, 0}, rmax = 0, rsum = 0;
    for (uint32_t k = 0; k < n; ++k)
      for (int a = 0; a < 3; ++a) {
        mx[a] = std::max(mx[a], static_cast<double>(std::fabs(xyz[3 * k + a])));
        mean[a] += xyz[3 * k + a];
      }
    for (int a = 0; a < 3; ++a) mean[a] /= n;
    double m2[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t k = 0; k < n; ++k) {
      const double px = xyz[3 * k], py = xyz[3 * k + 1], pz = xyz[3 * k + 2];
      const double x = px - mean[0], y = py - mean[1], z = pz - mean[2];
      m2[0] += x * x; m2[1] += y * y; m2[2] += z * z; m2[3] += x * y; m2[4] += x * z; m2[5] += y * z;
      const double r = std::sqrt(px * px + py * py + pz * pz);
      rmax = std::max(rmax, r);
      rsum += r;
    }
    for (int a = 0; a < 3; ++a) { d[a] = static_cast<float>(mx[a]); d[3 + a] = static_cast<float>(mean[a]); }
    for (int a = 0; a < 6; ++a) d[6 + a] = static_cast<float>(m2[a] / n);
    d[12] = static_cast<float>(std::log2(static_cast<double>(n)) / 9.0);
    d[13] = static_cast<float>(rmax);
    d[14] = static_cast<float>(rsum / n);
    d[15] = 1.f;
    return;
  }
  double h[3] = {0, 0, 0};
  if (type == MFCL_SHAPE_BOX) { h[0] = p[0] / 2; h[1] = p[1] / 2; h[2] = p[2] / 2; }
  else if (type == MFCL_SHAPE_SPHERE) { h[0] = h[1] = h[2] = p[0]; }
  else if (type == MFCL_SHAPE_CAPSULE) { h[0] = h[1] = p[0]; h[2] = p[0] + p[1] / 2; }
  else { h[0] = h[1] = p[0]; h[2] = p[1] / 2; }
  for (int a = 0; a < 3; ++a) { d[a] = static_cast<float>(h[a]); d[6 + a] = static_cast<float>(h[a] * h[a] / 3); }
  d[13] = static_cast<float>(std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]));
  d[14] = d[13];
}

int mfcl_upload_shapes(mfcl_ctx* ctx, const int32_t* type, const float* params4, const uint64_t* vert_off,
                       const uint32_t* vert_cnt, uint32_t n_shapes, const float* verts_xyz,
                       uint64_t n_verts) {
  if (!ctx || !type || !params4 || n_shapes == 0) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  freeShapes(ctx);
  std::vector<float4> param(n_shapes);
  std::vector<uint32_t> voff(n_shapes, 0), vcnt(n_shapes, 0);
  std::vector<double4> interior(n_shapes, double4{0, 0, 0, 0});
  std::vector<float> desc(16 * static_cast<size_t>(n_shapes), 0.f);
  std::vector<float4> verts(std::max<uint64_t>(n_verts, 1));
  for (uint64_t k = 0; k < n_verts; ++k)
    verts[k] = float4{verts_xyz[3 * k], verts_xyz[3 * k + 1], verts_xyz[3 * k + 2], 0.f};
  for (uint32_t i = 0; i < n_shapes; ++i) {
    const int t = type[i];
    param[i] = float4{params4[4 * i], params4[4 * i + 1], params4[4 * i + 2], 0.f};
    if (t == MFCL_SHAPE_CONVEX) {
      if (!vert_off || !vert_cnt || vert_cnt[i] == 0 || vert_off[i] + vert_cnt[i] > n_verts ||
          vert_off[i] > 0xFFFFFFFFull) {
        ctx->err = "convex shape " + std::to_string(i) + " has an invalid vertex range";
        return MFCL_E_INVALID;
      }
      voff[i] = static_cast<uint32_t>(vert_off[i]);
      vcnt[i] = vert_cnt[i];
      // interior point = vertex mean in double (convex-inl.h:64-70); rbound
      // = max |coordinate| for the FP32 support-scan error bound.
      double sx = 0, sy = 0, sz = 0;
      float rb = 0;
      for (uint32_t k = 0; k < vcnt[i]; ++k) {
        const float* p = verts_xyz + 3 * (vert_off[i] + k);
        sx += p[0];
        sy += p[1];
        sz += p[2];
        rb = std::max(rb, std::max(std::fabs(p[0]), std::max(std::fabs(p[1]), std::fabs(p[2]))));
      }
      const double inv = 1.0 / vcnt[i];
      interior[i] = double4{sx * inv, sy * inv, sz * inv, 0};
      param[i].w = rb;
      shapeDescriptor(t, params4 + 4 * i, verts_xyz + 3 * vert_off[i], vcnt[i], &desc[16 * static_cast<size_t>(i)]);
    } else if (t != MFCL_SHAPE_BOX && t != MFCL_SHAPE_SPHERE && t != MFCL_SHAPE_CAPSULE &&
               t != MFCL_SHAPE_CYLINDER) {
      ctx->err = "unsupported shape type " + std::to_string(t) + " for shape " + std::to_string(i);
      return MFCL_E_UNSUPPORTED;
    } else {
      shapeDescriptor(t, params4 + 4 * i, nullptr, 0, &desc[16 * static_cast<size_t>(i)]);
    }
  }
  int rc;
  if ((rc = upload(ctx, &ctx->d_type, type, n_shapes))) return rc;
  if ((rc = upload(ctx, &ctx->d_param, param.data(), n_shapes))) return rc;
  if ((rc = upload(ctx, &ctx->d_voff, voff.data(), n_shapes))) return rc;
  if ((rc = upload(ctx, &ctx->d_vcnt, vcnt.data(), n_shapes))) return rc;
  if ((rc = upload(ctx, &ctx->d_verts, verts.data(), verts.size()))) return rc;
  if ((rc = upload(ctx, &ctx->d_interior, interior.data(), n_shapes))) return rc;
  if ((rc = upload(ctx, &ctx->d_desc, desc.data(), desc.size()))) return rc;
  ctx->h_vcnt = vcnt;
  ctx->h_voff = voff;
  ctx->n_verts_alloc = verts.size();
  ctx->h_type.assign(type, type + n_shapes);
  ctx->n_shapes = n_shapes;
  return MFCL_OK;
}

int mfcl_upload_convex_faces(mfcl_ctx* ctx, const int32_t* faces, uint64_t faces_len, const uint64_t* face_off,
                             const uint32_t* num_faces) {
  if (!ctx || !face_off || !num_faces || (faces_len && !faces)) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (ctx->n_shapes == 0) {
    ctx->err = "no shapes uploaded";
    return MFCL_E_STATE;
  }
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const int rc = uploadConvexFaces(ctx, faces, face_off, num_faces, faces_len);
  if (rc) {
    freePtr(reinterpret_cast<void*&>(ctx->d_nbr));
    freePtr(reinterpret_cast<void*&>(ctx->d_nbr_off));
    freePtr(reinterpret_cast<void*&>(ctx->d_seed));
    freePtr(reinterpret_cast<void*&>(ctx->d_wblk));
  }
  return rc;
}

static int finishTiming(mfcl_ctx* ctx, int first_ev, int n_ev, int slot0) {
  if (!ctx->timing) return MFCL_OK;
  MFCL_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[first_ev + n_ev - 1]));
  for (int k = 0; k + 1 < n_ev; ++k) {
    float ms = 0;
    MFCL_HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[first_ev + k], ctx->ev[first_ev + k + 1]));
    ctx->last_ms[slot0 + k] = ms;
  }
  return MFCL_OK;
}

int mfcl_convex_collide_contacts_dev(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                                     const float* poseb, const float* guess, uint64_t n, const mfcl_request* req,
                                     mfcl_convex_result* out, mfcl_contact* contacts, void* stream) {
  if (!ctx || !req) return MFCL_E_INVALID;
  if (n && (!pa || !pb || !posea || !poseb || !out)) {
    ctx->err = "null device pointer";
    return MFCL_E_INVALID;
  }
  if (n == 0) return MFCL_OK;
  if (n > ctx->ws_convex_pairs) {
    ctx->err = "convex workspace holds " + std::to_string(ctx->ws_convex_pairs) + " pairs, query has " +
               std::to_string(n) + ": call mfcl_reserve(ctx, MFCL_WS_CONVEX, n) first";
    return MFCL_E_STATE;
  }
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  ctx->counters[0] = n;
  ConvexWs ws;
  if (!carveConvexWs(ctx->ws_convex.p, ctx->ws_convex.bytes, n, epaArenaBytes(kReservedArenaSlices, 256), ws)) {
    ctx->err = "convex workspace layout exceeds its allocation";
    return MFCL_E_STATE;
  }
  const int rc = convexQueryLaunch(ctx, pa, pb, posea, poseb, guess, n, req, out, contacts, ws, s);
  if (rc) return rc;
  if (ctx->timing) return finishTiming(ctx, 0, 4, 0);
  return MFCL_OK;
}

int mfcl_convex_query_static_dev(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                                 const float* poseb, const float* guess, uint64_t n, const mfcl_request* req,
                                 mfcl_convex_result* out, void* stream) {
  if (!ctx || !req) return MFCL_E_INVALID;
  if (n && (!pa || !pb || !posea || !poseb || !out)) {
    ctx->err = "null device pointer";
    return MFCL_E_INVALID;
  }
  if (n == 0) return MFCL_OK;
  const uint64_t max_pairs = staticScratchMaxPairs();
  if (n > max_pairs) {
    ctx->err = "mfcl_convex_query_static_dev takes at most " + std::to_string(max_pairs) + " pairs per call (" +
               std::to_string(kStaticScratchBytes >> 20) + " MiB static workspace), got " + std::to_string(n);
    return MFCL_E_INVALID;
  }
  if (!ctx->d_static_scratch) {
    ctx->err = "the static workspace's device address could not be looked up at mfcl_create";
    return MFCL_E_STATE;
  }
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  ctx->counters[0] = n;
  ConvexWs ws;
  if (!carveConvexWs(ctx->d_static_scratch, kStaticScratchBytes, n, epaArenaBytes(kStaticArenaSlices, 256), ws)) {
    ctx->err = "static workspace layout exceeds its size";
    return MFCL_E_STATE;
  }
  const int rc = convexQueryLaunch(ctx, pa, pb, posea, poseb, guess, n, req, out, nullptr, ws, s);
  if (rc) return rc;
  if (ctx->timing) return finishTiming(ctx, 0, 4, 0);
  return MFCL_OK;
}

int mfcl_convex_query_static_dev_max_pairs(uint64_t* max_pairs) {
  if (!max_pairs) return MFCL_E_INVALID;
  *max_pairs = staticScratchMaxPairs();
  return MFCL_OK;
}

int mfcl_convex_query_dev(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                          const float* poseb, const float* guess, uint64_t n, const mfcl_request* req,
                          mfcl_convex_result* out, void* stream) {
  return mfcl_convex_collide_contacts_dev(ctx, pa, pb, posea, poseb, guess, n, req, out, nullptr, stream);
}

int mfcl_convex_query(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                      const float* poseb, const float* guess, uint64_t n, const mfcl_request* req,
                      mfcl_convex_result* out) {
  MFCL_NOT_WHILE_CAPTURING(ctx);
  return mfcl_convex_collide_contacts(ctx, pa, pb, posea, poseb, guess, n, req, out, nullptr);
}

int mfcl_convex_collide_contacts(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                                 const float* poseb, const float* guess, uint64_t n, const mfcl_request* req,
                                 mfcl_convex_result* out, mfcl_contact* contacts) {
  if (!ctx || !req) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (n == 0) return MFCL_OK;
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (n > 0xFFFFFFFFull) {
    ctx->err = "too many pairs in one query";
    return MFCL_E_INVALID;
  }
  if (contacts && !(req->mode == MFCL_MODE_COLLIDE && req->penetration)) {
    ctx->err = "a contact array needs a collide request with penetration";
    return MFCL_E_INVALID;
  }
  if (int rc = reserveConvex(ctx, n)) return rc;
  for (uint64_t i = 0; i < n; ++i)
    if (pa[i] >= ctx->n_shapes || pb[i] >= ctx->n_shapes) {
      ctx->err = "pair " + std::to_string(i) + " references a shape id out of range";
      return MFCL_E_INVALID;
    }
  const int cstride = req->max_contacts < MFCL_MAX_PAIR_CONTACTS ? req->max_contacts : MFCL_MAX_PAIR_CONTACTS;
  const size_t b_ids = 4 * n, b_pose = 28 * n, b_guess = req->use_guess ? 12 * n : 0,
               b_out = sizeof(mfcl_convex_result) * n,
               b_con = contacts && cstride > 0 ? sizeof(mfcl_contact) * n * cstride : 0;
  const size_t total = 2 * b_ids + 2 * b_pose + b_guess + b_out + b_con + 7 * 256;
  if (int rc = ensure(ctx, ctx->ws_io, total)) return rc;
  char* base = static_cast<char*>(ctx->ws_io.p);
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += (bytes + 255) & ~static_cast<size_t>(255);
    return p;
  };
  uint32_t* dpa = reinterpret_cast<uint32_t*>(carve(b_ids));
  uint32_t* dpb = reinterpret_cast<uint32_t*>(carve(b_ids));
  float* dqa = reinterpret_cast<float*>(carve(b_pose));
  float* dqb = reinterpret_cast<float*>(carve(b_pose));
  float* dg = b_guess ? reinterpret_cast<float*>(carve(b_guess)) : nullptr;
  mfcl_convex_result* dout = reinterpret_cast<mfcl_convex_result*>(carve(b_out));
  mfcl_contact* dcon = b_con ? reinterpret_cast<mfcl_contact*>(carve(b_con)) : nullptr;
  hipStream_t s = ctx->stream;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dpa, pa, b_ids, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dpb, pb, b_ids, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dqa, posea, b_pose, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dqb, poseb, b_pose, hipMemcpyHostToDevice, s));
  if (dg) MFCL_HIP_TRY(ctx, hipMemcpyAsync(dg, guess, b_guess, hipMemcpyHostToDevice, s));
  // the device pass writes only the contacts each pair keeps: start from the caller's records
  if (dcon) MFCL_HIP_TRY(ctx, hipMemcpyAsync(dcon, contacts, b_con, hipMemcpyHostToDevice, s));
  int rc = mfcl_convex_collide_contacts_dev(ctx, dpa, dpb, dqa, dqb, dg, n, req, dout, dcon, s);
  if (rc) return rc;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(out, dout, b_out, hipMemcpyDeviceToHost, s));
  if (dcon) MFCL_HIP_TRY(ctx, hipMemcpyAsync(contacts, dcon, b_con, hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
  return MFCL_OK;
}

static int mprPenDev(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea, const float* poseb,
                     const float* dir3, uint64_t n, const mfcl_request* req, mfcl_convex_result* out, void* stream,
                     bool incremental) {
  if (!ctx || !req) return MFCL_E_INVALID;
  if (n && !dir3) {
    ctx->err = "null device pointer";
    return MFCL_E_INVALID;
  }
  if (req->mode != MFCL_MODE_COLLIDE) {
    ctx->err = "MPR penetration needs a collide request";
    return MFCL_E_INVALID;
  }
  mfcl_request boolean = *req;
  boolean.penetration = 0;
  boolean.use_guess = 0;
  // the boolean collide first (timed as the query's stages), then the pass
  if (int rc = mfcl_convex_collide_contacts_dev(ctx, pa, pb, posea, poseb, nullptr, n, &boolean, out, nullptr, stream))
    return rc;
  if (n == 0) return MFCL_OK;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  // MPR(128, distanceTolerance) as collisionPenetrationMPR builds it (:208-210)
  return mprPenLaunch(ctx, pa, pb, posea, poseb, dir3, n, incremental, 128, req->epa_tolerance, out, s);
}

static int mprPenHost(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                      const float* poseb, const float* dir3, uint64_t n, const mfcl_request* req,
                      mfcl_convex_result* out, bool incremental) {
  if (!ctx || !req) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (n == 0) return MFCL_OK;
  if (!pa || !pb || !posea || !poseb || !dir3 || !out) return MFCL_E_INVALID;
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (n > 0xFFFFFFFFull) {
    ctx->err = "too many pairs in one query";
    return MFCL_E_INVALID;
  }
  if (int rc = reserveConvex(ctx, n)) return rc;
  for (uint64_t i = 0; i < n; ++i)
    if (pa[i] >= ctx->n_shapes || pb[i] >= ctx->n_shapes) {
      ctx->err = "pair " + std::to_string(i) + " references a shape id out of range";
      return MFCL_E_INVALID;
    }
  const size_t b_ids = 4 * n, b_pose = 28 * n, b_dir = 12 * n, b_out = sizeof(mfcl_convex_result) * n;
  if (int rc = ensure(ctx, ctx->ws_io, 2 * b_ids + 2 * b_pose + b_dir + b_out + 6 * 256)) return rc;
  char* base = static_cast<char*>(ctx->ws_io.p);
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += (bytes + 255) & ~static_cast<size_t>(255);
    return p;
  };
  uint32_t* dpa = reinterpret_cast<uint32_t*>(carve(b_ids));
  uint32_t* dpb = reinterpret_cast<uint32_t*>(carve(b_ids));
  float* dqa = reinterpret_cast<float*>(carve(b_pose));
  float* dqb = reinterpret_cast<float*>(carve(b_pose));
  float* dd = reinterpret_cast<float*>(carve(b_dir));
  mfcl_convex_result* dout = reinterpret_cast<mfcl_convex_result*>(carve(b_out));
  hipStream_t s = ctx->stream;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dpa, pa, b_ids, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dpb, pb, b_ids, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dqa, posea, b_pose, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dqb, poseb, b_pose, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dd, dir3, b_dir, hipMemcpyHostToDevice, s));
  if (int rc = mprPenDev(ctx, dpa, dpb, dqa, dqb, dd, n, req, dout, s, incremental)) return rc;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(out, dout, b_out, hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
  return MFCL_OK;
}

int mfcl_convex_directed_penetration_dev(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb,
                                         const float* posea, const float* poseb, const float* dir3, uint64_t n,
                                         const mfcl_request* req, mfcl_convex_result* out, void* stream) {
  return mprPenDev(ctx, pa, pb, posea, poseb, dir3, n, req, out, stream, false);
}

int mfcl_convex_directed_penetration(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                                     const float* poseb, const float* dir3, uint64_t n, const mfcl_request* req,
                                     mfcl_convex_result* out) {
  return mprPenHost(ctx, pa, pb, posea, poseb, dir3, n, req, out, false);
}

int mfcl_convex_incremental_penetration_dev(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb,
                                            const float* posea, const float* poseb, const float* hint3, uint64_t n,
                                            const mfcl_request* req, mfcl_convex_result* out, void* stream) {
  return mprPenDev(ctx, pa, pb, posea, poseb, hint3, n, req, out, stream, true);
}

int mfcl_convex_incremental_penetration(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                                        const float* poseb, const float* hint3, uint64_t n, const mfcl_request* req,
                                        mfcl_convex_result* out) {
  return mprPenHost(ctx, pa, pb, posea, poseb, hint3, n, req, out, true);
}

int mfcl_convex_hulls_dev(mfcl_ctx* ctx, const float* points, const uint64_t* pt_off, const uint32_t* pt_cnt,
                          uint32_t n_sets, int32_t max_points, uint32_t* hull_cnt, uint32_t* hull_idx,
                          uint32_t* tri_cnt, int32_t* tris, int32_t* status, void* stream) {
  if (!ctx) return MFCL_E_INVALID;
  if (n_sets && (!points || !pt_off || !pt_cnt || !hull_cnt || !hull_idx || !tri_cnt || !tris || !status)) {
    ctx->err = "null device pointer";
    return MFCL_E_INVALID;
  }
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return hullsLaunch(ctx, points, pt_off, pt_cnt, n_sets, max_points, hull_cnt, hull_idx, tri_cnt, tris, status, s);
}

int mfcl_convex_hulls(mfcl_ctx* ctx, const float* points, const uint64_t* pt_off, const uint32_t* pt_cnt,
                      uint32_t n_sets, uint32_t* hull_cnt, uint32_t* hull_idx, uint32_t* tri_cnt, int32_t* tris,
                      int32_t* status) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (n_sets == 0) return MFCL_OK;
  if (!points || !pt_off || !pt_cnt || !hull_cnt || !hull_idx || !tri_cnt || !tris || !status) return MFCL_E_INVALID;
  uint64_t total = 0;
  for (uint32_t i = 0; i < n_sets; ++i) total = std::max<uint64_t>(total, pt_off[i] + pt_cnt[i]);
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t b_pts = 12 * total, b_off = 8 * static_cast<size_t>(n_sets), b_cnt = 4 * static_cast<size_t>(n_sets),
               b_idx = 4 * total, b_tri = 24 * total;
  if (int rc = ensure(ctx, ctx->ws_io, b_pts + b_off + 5 * b_cnt + b_idx + b_tri + 9 * 256)) return rc;
  char* base = static_cast<char*>(ctx->ws_io.p);
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += (bytes + 255) & ~static_cast<size_t>(255);
    return p;
  };
  float* dp = reinterpret_cast<float*>(carve(b_pts));
  uint64_t* doff = reinterpret_cast<uint64_t*>(carve(b_off));
  uint32_t* dcnt = reinterpret_cast<uint32_t*>(carve(b_cnt));
  uint32_t* dhc = reinterpret_cast<uint32_t*>(carve(b_cnt));
  uint32_t* dtc = reinterpret_cast<uint32_t*>(carve(b_cnt));
  int32_t* dst = reinterpret_cast<int32_t*>(carve(b_cnt));
  uint32_t* dhi = reinterpret_cast<uint32_t*>(carve(b_idx));
  int32_t* dtr = reinterpret_cast<int32_t*>(carve(b_tri));
  uint32_t* dord = reinterpret_cast<uint32_t*>(carve(b_cnt));
  hipStream_t s = ctx->stream;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dp, points, b_pts, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(doff, pt_off, b_off, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dcnt, pt_cnt, b_cnt, hipMemcpyHostToDevice, s));
  if (int rc = hullsLaunchClassed(ctx, dp, doff, dcnt, pt_cnt, n_sets, dord, dhc, dhi, dtc, dtr, dst, s)) return rc;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(hull_cnt, dhc, b_cnt, hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(tri_cnt, dtc, b_cnt, hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(status, dst, b_cnt, hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(hull_idx, dhi, b_idx, hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(tris, dtr, b_tri, hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
  return MFCL_OK;
}

int mfcl_upload_convex_points(mfcl_ctx* ctx, const float* points, const uint64_t* pt_off, const uint32_t* pt_cnt,
                              uint32_t n_sets, uint32_t* hull_cnt) {
  if (!ctx || !points || !pt_off || !pt_cnt || n_sets == 0) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return uploadConvexPoints(ctx, points, pt_off, pt_cnt, n_sets, hull_cnt);
}

int mfcl_upload_meshes(mfcl_ctx* ctx, const float* verts_xyz, const uint64_t* vert_off, const uint32_t* vert_cnt,
                       const uint32_t* tris, const uint64_t* tri_off, const uint32_t* tri_cnt,
                       uint32_t n_meshes) {
  if (!ctx || !verts_xyz || !vert_off || !vert_cnt || !tris || !tri_off || !tri_cnt || n_meshes == 0)
    return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  freeMeshes(ctx);
  if (int rc = buildAndUploadMeshes(ctx, verts_xyz, vert_off, vert_cnt, tris, tri_off, tri_cnt, n_meshes)) return rc;
  return reserveMesh(ctx);
}

int mfcl_mesh_num_nodes(mfcl_ctx* ctx, uint32_t mesh, uint32_t* n_nodes) {
  if (!ctx || !n_nodes || mesh >= ctx->n_meshes) return MFCL_E_INVALID;
  *n_nodes = ctx->mesh_host.node_count[mesh];
  return MFCL_OK;
}

int mfcl_mesh_get_nodes(mfcl_ctx* ctx, uint32_t mesh, void* nodes) {
  if (!ctx || !nodes || mesh >= ctx->n_meshes) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  MFCL_HIP_TRY(ctx, hipMemcpy(nodes, static_cast<char*>(ctx->d_nodes) + 64ull * ctx->mesh_host.node_base[mesh],
                              64ull * ctx->mesh_host.node_count[mesh], hipMemcpyDeviceToHost));
  return MFCL_OK;
}

int mfcl_mesh_replace_vertices(mfcl_ctx* ctx, uint32_t mesh, const float* verts_xyz, int32_t refit) {
  if (!ctx || !verts_xyz || mesh >= ctx->n_meshes) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (!refit) return meshRebuild(ctx, mesh, verts_xyz);
  if (int rc = meshReplaceVertices(ctx, mesh, verts_xyz, false, ctx->stream)) return rc;
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return MFCL_OK;
}

int mfcl_mesh_replace_vertices_dev(mfcl_ctx* ctx, uint32_t mesh, const float* verts_xyz, void* stream) {
  if (!ctx || !verts_xyz || mesh >= ctx->n_meshes) return MFCL_E_INVALID;
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  return meshReplaceVertices(ctx, mesh, verts_xyz, true, stream ? static_cast<hipStream_t>(stream) : ctx->stream);
}

int mfcl_save_meshes(mfcl_ctx* ctx, const char* path) {
  if (!ctx || !path) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (ctx->n_meshes == 0) {
    ctx->err = "no meshes uploaded";
    return MFCL_E_STATE;
  }
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MeshSetHost m;
  if (int rc = downloadMeshSet(ctx, m)) return rc;
  return writeMeshFile(m, path, ctx->err);
}

int mfcl_load_meshes(mfcl_ctx* ctx, const char* path) {
  if (!ctx || !path) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MeshSetHost m;
  if (int rc = readMeshFile(path, m, ctx->err)) return rc;
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  freeMeshes(ctx);
  if (int rc = uploadMeshSet(ctx, m)) {
    freeMeshes(ctx);
    return rc;
  }
  return reserveMesh(ctx);
}

// Host-only entry points report errors through their return code alone
// (there is no context to hold a message); the last message is kept here.
static thread_local std::string g_host_err;
const char* mfcl_host_last_error(void) { return g_host_err.c_str(); }

int mfcl_build_mesh_file(const char* path, const float* verts_xyz, const uint64_t* vert_off, const uint32_t* vert_cnt,
                         const uint32_t* tris, const uint64_t* tri_off, const uint32_t* tri_cnt, uint32_t n_meshes) {
  if (!path || !verts_xyz || !vert_off || !vert_cnt || !tris || !tri_off || !tri_cnt || n_meshes == 0)
    return MFCL_E_INVALID;
  MeshSetHost m;
  if (int rc = buildMeshSet(g_host_err, verts_xyz, vert_off, vert_cnt, tris, tri_off, tri_cnt, n_meshes, m))
    return rc;
  return writeMeshFile(m, path, g_host_err);
}

int mfcl_mesh_file_info(const char* path, uint32_t* n_meshes, uint64_t* n_tris, uint64_t* n_nodes) {
  if (!path || !n_meshes || !n_tris || !n_nodes) return MFCL_E_INVALID;
  MeshSetHost m;
  if (int rc = readMeshFile(path, m, g_host_err)) return rc;
  *n_meshes = static_cast<uint32_t>(m.node_base.size());
  *n_tris = m.prim.size();
  *n_nodes = m.nodes.size() / 64;
  return MFCL_OK;
}

int mfcl_load_obj(const char* path, float* verts_xyz, uint64_t* n_verts, uint32_t* tris, uint64_t* n_tris) {
  if (!path || !n_verts || !n_tris) return MFCL_E_INVALID;
  std::vector<float> v;
  std::vector<uint32_t> t;
  if (int rc = loadObj(path, v, t, g_host_err)) return rc;
  const uint64_t nv = v.size() / 3, nt = t.size() / 3;
  if (verts_xyz || tris) {
    if (!verts_xyz || !tris || *n_verts < nv || *n_tris < nt) return MFCL_E_INVALID;
    std::memcpy(verts_xyz, v.data(), 4 * v.size());
    std::memcpy(tris, t.data(), 4 * t.size());
  }
  *n_verts = nv;
  *n_tris = nt;
  return MFCL_OK;
}

int mfcl_mesh_shape_collide(mfcl_ctx* ctx, const uint32_t* mesh_ids, const uint32_t* shape_ids,
                            const float* pose_mesh7, const float* pose_shape7, uint64_t n, const mfcl_request* req,
                            mfcl_mesh_result* out, int32_t* tri_ids) {
  if (!ctx || !req) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (n == 0) return MFCL_OK;
  if (!mesh_ids || !shape_ids || !pose_mesh7 || !pose_shape7 || !out) return MFCL_E_INVALID;
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (n > 0xFFFFFFFFull) {
    ctx->err = "too many queries in one call";
    return MFCL_E_INVALID;
  }
  if (req->penetration) {
    ctx->err = "mfcl_mesh_shape_collide takes penetration-disabled requests; use mfcl_mesh_shape_collide_contacts";
    return MFCL_E_INVALID;
  }
  return meshShapeCollide(ctx, mesh_ids, shape_ids, pose_mesh7, pose_shape7, n, req, out, tri_ids, nullptr);
}

int mfcl_mesh_shape_collide_contacts(mfcl_ctx* ctx, const uint32_t* mesh_ids, const uint32_t* shape_ids,
                                     const float* pose_mesh7, const float* pose_shape7, uint64_t n,
                                     const mfcl_request* req, mfcl_mesh_result* out, mfcl_mesh_contact* contacts) {
  if (!ctx || !req) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (n == 0) return MFCL_OK;
  if (!mesh_ids || !shape_ids || !pose_mesh7 || !pose_shape7 || !out || !contacts) return MFCL_E_INVALID;
  if (!req->penetration) {
    ctx->err = "mfcl_mesh_shape_collide_contacts needs a request with penetration enabled";
    return MFCL_E_INVALID;
  }
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (n > 0xFFFFFFFFull) {
    ctx->err = "too many queries in one call";
    return MFCL_E_INVALID;
  }
  return meshShapeCollide(ctx, mesh_ids, shape_ids, pose_mesh7, pose_shape7, n, req, out, nullptr, contacts);
}

static int meshCollideDev(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                          const float* poseb, uint64_t n, int32_t max_contacts, mfcl_mesh_result* out,
                          int32_t* contacts, mfcl_mesh_contact* pen, void* stream) {
  if (!ctx) return MFCL_E_INVALID;
  if (max_contacts < 1 || max_contacts > 1024) {
    ctx->err = "max_contacts must be in [1, 1024]";
    return MFCL_E_INVALID;
  }
  if (n && (!pa || !pb || !posea || !poseb || !out)) {
    ctx->err = "null device pointer";
    return MFCL_E_INVALID;
  }
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  const int rc = meshCollideLaunch(ctx, pa, pb, posea, poseb, n, max_contacts, out, contacts, pen, s);
  if (rc) return rc;
  if (ctx->timing && n) return finishTiming(ctx, 4, 2, 3);
  return MFCL_OK;
}

int mfcl_mesh_collide_dev(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                          const float* poseb, uint64_t n, int32_t max_contacts, mfcl_mesh_result* out,
                          int32_t* contacts, void* stream) {
  return meshCollideDev(ctx, pa, pb, posea, poseb, n, max_contacts, out, contacts, nullptr, stream);
}

int mfcl_mesh_collide_contacts_dev(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                                   const float* poseb, uint64_t n, int32_t max_contacts, mfcl_mesh_result* out,
                                   mfcl_mesh_contact* contacts, void* stream) {
  if (ctx && n && !contacts) {
    ctx->err = "null contact array";
    return MFCL_E_INVALID;
  }
  return meshCollideDev(ctx, pa, pb, posea, poseb, n, max_contacts, out, nullptr, contacts, stream);
}

static int meshCollideHost(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                           const float* poseb, uint64_t n, int32_t max_contacts, mfcl_mesh_result* out,
                           int32_t* contacts, mfcl_mesh_contact* pen) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (n == 0) return MFCL_OK;
  if (max_contacts < 1 || max_contacts > 1024) {
    ctx->err = "max_contacts must be in [1, 1024]";
    return MFCL_E_INVALID;
  }
  for (uint64_t i = 0; i < n; ++i)
    if (pa[i] >= ctx->n_meshes || pb[i] >= ctx->n_meshes) {
      ctx->err = "pair " + std::to_string(i) + " references a mesh id out of range";
      return MFCL_E_INVALID;
    }
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t b_ids = 4 * n, b_pose = 28 * n, b_out = sizeof(mfcl_mesh_result) * n,
               b_c = sizeof(int32_t) * 2 * max_contacts * n, b_p = pen ? sizeof(mfcl_mesh_contact) * max_contacts * n : 0;
  if (int rc = ensure(ctx, ctx->ws_io, 2 * b_ids + 2 * b_pose + b_out + b_c + b_p + 7 * 256)) return rc;
  char* base = static_cast<char*>(ctx->ws_io.p);
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += (bytes + 255) & ~static_cast<size_t>(255);
    return p;
  };
  uint32_t* dpa = reinterpret_cast<uint32_t*>(carve(b_ids));
  uint32_t* dpb = reinterpret_cast<uint32_t*>(carve(b_ids));
  float* dqa = reinterpret_cast<float*>(carve(b_pose));
  float* dqb = reinterpret_cast<float*>(carve(b_pose));
  mfcl_mesh_result* dout = reinterpret_cast<mfcl_mesh_result*>(carve(b_out));
  int32_t* dc = reinterpret_cast<int32_t*>(carve(b_c));
  mfcl_mesh_contact* dp = pen ? reinterpret_cast<mfcl_mesh_contact*>(carve(b_p)) : nullptr;
  hipStream_t s = ctx->stream;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dpa, pa, b_ids, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dpb, pb, b_ids, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dqa, posea, b_pose, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dqb, poseb, b_pose, hipMemcpyHostToDevice, s));
  int rc = meshCollideDev(ctx, dpa, dpb, dqa, dqb, n, max_contacts, dout, pen ? nullptr : dc, dp, s);
  if (rc) return rc;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(out, dout, b_out, hipMemcpyDeviceToHost, s));
  if (contacts && !pen) MFCL_HIP_TRY(ctx, hipMemcpyAsync(contacts, dc, b_c, hipMemcpyDeviceToHost, s));
  if (pen) MFCL_HIP_TRY(ctx, hipMemcpyAsync(pen, dp, b_p, hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
  return MFCL_OK;
}

int mfcl_mesh_collide(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                      const float* poseb, uint64_t n, int32_t max_contacts, mfcl_mesh_result* out,
                      int32_t* contacts) {
  return meshCollideHost(ctx, pa, pb, posea, poseb, n, max_contacts, out, contacts, nullptr);
}

int mfcl_mesh_collide_contacts(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                               const float* poseb, uint64_t n, int32_t max_contacts, mfcl_mesh_result* out,
                               mfcl_mesh_contact* contacts) {
  if (ctx && n && !contacts) {
    ctx->err = "null contact array";
    return MFCL_E_INVALID;
  }
  return meshCollideHost(ctx, pa, pb, posea, poseb, n, max_contacts, out, nullptr, contacts);
}

int mfcl_aabbtree_build(mfcl_ctx* ctx, const float* mn3, const float* mx3, const int64_t* ids, uint32_t n) {
  if (!ctx || !mn3 || !mx3 || !ids || n == 0) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return buildAndUploadAabbTree(ctx, mn3, mx3, ids, n);
}

int mfcl_aabbtree_query_dev(mfcl_ctx* ctx, const float* qmn3, const float* qmx3, uint64_t nq, int32_t cap,
                            int32_t* counts, int64_t* ids, void* stream) {
  if (!ctx) return MFCL_E_INVALID;
  if (cap < 0) {
    ctx->err = "cap must be >= 0";
    return MFCL_E_INVALID;
  }
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  const int rc = aabbQueryLaunch(ctx, qmn3, qmx3, nq, cap, counts, ids, s);
  if (rc) return rc;
  if (ctx->timing && nq) return finishTiming(ctx, 6, 2, 4);
  return MFCL_OK;
}

int mfcl_aabbtree_query(mfcl_ctx* ctx, const float* qmn3, const float* qmx3, uint64_t nq, int32_t cap,
                        int32_t* counts, int64_t* ids) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (nq == 0) return MFCL_OK;
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t b_box = 12 * nq, b_cnt = 4 * nq, b_ids = sizeof(int64_t) * (cap > 0 ? cap : 1) * nq;
  if (int rc = ensure(ctx, ctx->ws_io, 2 * b_box + b_cnt + b_ids + 4 * 256)) return rc;
  char* base = static_cast<char*>(ctx->ws_io.p);
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += (bytes + 255) & ~static_cast<size_t>(255);
    return p;
  };
  float* dmn = reinterpret_cast<float*>(carve(b_box));
  float* dmx = reinterpret_cast<float*>(carve(b_box));
  int32_t* dcnt = reinterpret_cast<int32_t*>(carve(b_cnt));
  int64_t* dids = reinterpret_cast<int64_t*>(carve(b_ids));
  hipStream_t s = ctx->stream;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dmn, qmn3, b_box, hipMemcpyHostToDevice, s));
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(dmx, qmx3, b_box, hipMemcpyHostToDevice, s));
  int rc = mfcl_aabbtree_query_dev(ctx, dmn, dmx, nq, cap, dcnt, dids, s);
  if (rc) return rc;
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(counts, dcnt, b_cnt, hipMemcpyDeviceToHost, s));
  if (ids && cap > 0) MFCL_HIP_TRY(ctx, hipMemcpyAsync(ids, dids, b_ids, hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
  return MFCL_OK;
}

// ---- warm-start MLP ---------------------------------------------------------
int mfcl_mlp_init(mfcl_ctx* ctx, uint64_t seed) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  return mlpInit(ctx, seed);
}

int mfcl_mlp_num_params(void) { return kMlpParams; }

int mfcl_mlp_set_params(mfcl_ctx* ctx, const float* params, int32_t n) {
  if (!ctx || !params) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  return mlpSetParams(ctx, params, n, true);
}

int mfcl_mlp_get_params(mfcl_ctx* ctx, float* params, float* grads, int32_t n) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  return mlpGetParams(ctx, params, n, grads);
}

int mfcl_mlp_features_dev(mfcl_ctx* ctx, const uint32_t* pair_a, const uint32_t* pair_b, const float* pose_a7,
                          const float* pose_b7, uint64_t n, float* feat64, void* stream) {
  if (!ctx || (n && (!pair_a || !pair_b || !pose_a7 || !pose_b7 || !feat64))) return MFCL_E_INVALID;
  return mlpFeatures(ctx, pair_a, pair_b, pose_a7, pose_b7, n, feat64,
                     stream ? static_cast<hipStream_t>(stream) : ctx->stream);
}

int mfcl_mlp_guess_dev(mfcl_ctx* ctx, const uint32_t* pair_a, const uint32_t* pair_b, const float* pose_a7,
                       const float* pose_b7, uint64_t n, float* guess3, float* dist_estimate, void* stream) {
  if (!ctx || (n && (!pair_a || !pair_b || !pose_a7 || !pose_b7 || !guess3))) return MFCL_E_INVALID;
  const int rc = mlpGuess(ctx, pair_a, pair_b, pose_a7, pose_b7, n, guess3, dist_estimate,
                          stream ? static_cast<hipStream_t>(stream) : ctx->stream);
  if (rc) return rc;
  if (ctx->timing && n) {
    MFCL_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[9]));
    float ms = 0;
    MFCL_HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[8], ctx->ev[9]));
    ctx->last_ms[5] = ms;
  }
  return MFCL_OK;
}

int mfcl_mlp_forward_dev(mfcl_ctx* ctx, const float* feat64, uint64_t n, float* out4, void* stream) {
  if (!ctx || (n && (!feat64 || !out4))) return MFCL_E_INVALID;
  return mlpForwardFeat(ctx, feat64, n, out4, stream ? static_cast<hipStream_t>(stream) : ctx->stream);
}

int mfcl_mlp_train_step_dev(mfcl_ctx* ctx, const float* feat64, const float* target4, uint64_t n,
                            const float* hyper6, int32_t apply_update, float* loss_host, void* stream) {
  if (!ctx || !feat64 || !target4) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  const int rc = mlpTrainStep(ctx, feat64, target4, n, hyper6, apply_update, loss_host, s);
  if (rc) return rc;
  if (loss_host) MFCL_HIP_TRY(ctx, hipStreamSynchronize(s));
  return MFCL_OK;
}

// LayeredHeightMap(res_x, res_y, half_x, half_y) (layered_heightmap-inl.h:
// 8-51): halve the shape (doubling the resolution) while both halves exceed
// one, then build the layers from that top back down to the given bottom;
// each upper pixel is the maximum of its 2 x 2 children (rebuildNextLayer,
// :76-101).
int mfcl_upload_heightmap(mfcl_ctx* ctx, const uint16_t* heights_mm, uint16_t half_x, uint16_t half_y, float res_x,
                          float res_y) {
  if (!ctx) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  auto pow2 = [](uint32_t v) { return v > 0 && (v & (v - 1)) == 0; };
  if (!heights_mm || !pow2(half_x) || !pow2(half_y) || half_x > 16384 || half_y > 16384 || !(res_x > 0) ||
      !(res_y > 0) || !std::isfinite(res_x) || !std::isfinite(res_y)) {
    ctx->err = "heightmap: half shapes must be powers of two in [1, 16384] and resolutions positive";
    return MFCL_E_INVALID;
  }
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  freeHeightmap(ctx);
  int nx = half_x, ny = half_y, k = 0;
  while (nx > 1 && ny > 1) {
    nx /= 2;
    ny /= 2;
    ++k;
  }
  HeightmapHost h;
  h.n_layers = k + 1;
  std::vector<std::vector<uint16_t>> layer(h.n_layers);
  uint64_t total = 0;
  for (int l = 0; l < h.n_layers; ++l) {
    const int up = k - l;  // halvings above the bottom
    h.fx[l] = 2 * (half_x >> up);
    h.fy[l] = 2 * (half_y >> up);
    h.rx[l] = std::ldexp(res_x, up);
    h.ry[l] = std::ldexp(res_y, up);
    h.off[l] = total;
    total += static_cast<uint64_t>(h.fx[l]) * h.fy[l];
  }
  const int b = h.n_layers - 1;
  layer[b].assign(heights_mm, heights_mm + static_cast<size_t>(h.fx[b]) * h.fy[b]);
  for (int l = b - 1; l >= 0; --l) {
    layer[l].assign(static_cast<size_t>(h.fx[l]) * h.fy[l], 0);
    for (int y = 0; y < h.fy[l + 1]; ++y)
      for (int x = 0; x < h.fx[l + 1]; ++x) {
        uint16_t& u = layer[l][static_cast<size_t>(y / 2) * h.fx[l] + x / 2];
        u = std::max(u, layer[l + 1][static_cast<size_t>(y) * h.fx[l + 1] + x]);
      }
  }
  for (uint16_t v : layer[b]) h.upper_mm = std::max<int>(h.upper_mm, v);
  std::vector<uint16_t> all;
  all.reserve(total);
  for (auto& v : layer) all.insert(all.end(), v.begin(), v.end());
  if (int rc = uploadVec(ctx, reinterpret_cast<void**>(&ctx->d_hm), all.data(), all.size() * sizeof(uint16_t)))
    return rc;
  ctx->hm = h;
  return MFCL_OK;
}

int mfcl_arm_last_heightmap_pairs(mfcl_ctx* ctx, uint64_t* n_pairs) {
  if (!ctx || !n_pairs) return MFCL_E_INVALID;
  *n_pairs = ctx->arm_hm_pairs;
  return MFCL_OK;
}

int mfcl_arm_query_dev(mfcl_ctx* ctx, const mfcl_arm* arm, const float* q, uint64_t n_configs, float margin,
                       const float* box_pose7, int32_t* any_hit, float* min_dist, uint64_t* n_pairs, void* stream) {
  if (!ctx || !arm) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (n_configs && (!q || !box_pose7 || !any_hit || !min_dist)) {
    ctx->err = "null device pointer";
    return MFCL_E_INVALID;
  }
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  if (n_pairs) *n_pairs = 0;
  return armQueryLaunch(ctx, arm, q, n_configs, margin, box_pose7, any_hit, min_dist, n_pairs, s);
}

int mfcl_reserve_arm(mfcl_ctx* ctx, const mfcl_arm* arm, uint64_t n_configs) {
  if (!ctx || !arm) return MFCL_E_INVALID;
  MFCL_NOT_WHILE_CAPTURING(ctx);
  if (arm->n_links < 1 || arm->n_links > MFCL_ARM_MAX_LINKS || arm->broadphase_cap < 1 ||
      arm->broadphase_cap > 512) {
    ctx->err = "n_links or broadphase_cap out of range";
    return MFCL_E_INVALID;
  }
  MFCL_HIP_TRY(ctx, hipSetDevice(ctx->device));
  uint64_t max_pairs = 0;
  const size_t bytes = armWsBytes(arm, n_configs ? n_configs : 1, &max_pairs);
  if (max_pairs > 0xFFFFFFFFull) {
    ctx->err = "candidate pair capacity too large for one query";
    return MFCL_E_INVALID;
  }
  if (int rc = ensure(ctx, ctx->ws_arm, bytes)) return rc;
  return reserveConvex(ctx, max_pairs);
}

}  // extern "C"

