// Library-internal context: device geometry, workspace and timing state.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mfcl.h"

#define MFCL_HIP_TRY(ctx, expr)                                               \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);         \
      return MFCL_E_HIP;                                                      \
    }                                                                         \
  } while (0)

namespace mfcl {

struct DevBuf {
  void* p{nullptr};
  size_t bytes{0};
};

// Warm-start MLP: 39 -> 128 -> 128 -> 128 -> 4 (inputs padded to 64,
// outputs to 32).  FP32 master parameters, packed in this order.
constexpr int kMlpIn = 39, kMlpInPad = 64, kHid = 128, kMlpOut = 4, kMlpOutPad = 32;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kMlpInPad;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kHid;
constexpr int kOffW4 = kOffB3 + kHid;
constexpr int kOffB4 = kOffW4 + kMlpOutPad * kHid;
constexpr int kMlpParams = kOffB4 + kMlpOutPad;

struct MlpState {
  float* params{nullptr};  // FP32 master weights
  float* grad{nullptr};
  float* m{nullptr};
  float* v{nullptr};
  float* loss{nullptr};
  void* wbuf{nullptr};     // BF16 copies of W (forward) and W^T (backward)
  __bf16 *w1{nullptr}, *w2{nullptr}, *w3{nullptr}, *w4{nullptr};
  __bf16 *w1t{nullptr}, *w2t{nullptr}, *w3t{nullptr}, *w4t{nullptr};
  DevBuf ws;               // activations
  void* pack{nullptr};     // fused-inference weight image (BF16, K-permuted) + biases
  int step{0};
};

struct MeshHost {
  std::vector<uint32_t> node_base;  // first node of each mesh
  std::vector<uint32_t> node_count;
  std::vector<uint32_t> tri_base;   // first triangle (and leaf-order slot) of each mesh
  std::vector<uint32_t> tri_count;
  std::vector<uint64_t> vert_base;  // first vertex of each mesh in the vertex copy
  std::vector<uint32_t> vert_count;
  std::vector<uint32_t> large_base;  // refit: each mesh's nodes over many triangles,
  std::vector<uint32_t> large_count; // a slice of the device list d_refit_large
};

// A mesh set as stored in host memory (built trees, or read from a file):
// vertices (xyz per mesh, concatenated), triangles (local vertex ids),
// 64-byte OBB nodes breadth-first, per node the range of leaf-order
// triangle slots it covers (first relative to its mesh, count), and per
// mesh the leaf-order triangle ids.
struct MeshSetHost {
  std::vector<float> verts;
  std::vector<uint64_t> vert_base;
  std::vector<uint32_t> vert_count;
  std::vector<uint32_t> tris;
  std::vector<uint32_t> tri_base, tri_count;
  std::vector<char> nodes;            // 64 B per node
  std::vector<uint32_t> node_base, node_count;
  std::vector<uint32_t> node_range;   // 2 per node: first, count
  std::vector<uint32_t> prim;         // per triangle slot
};

// Layers of a heightmap, top (0) to bottom (n_layers - 1), as
// LayeredHeightMap builds them: layer l has full shape (fx, fy) = 2 x its half
// shape and resolution (rx, ry); its pixels start at element off of the
// device array.  upper_mm: the largest height.
constexpr int kMaxHmLayers = 17;
struct HeightmapHost {
  int n_layers{0};
  uint64_t off[kMaxHmLayers]{};
  int fx[kMaxHmLayers]{}, fy[kMaxHmLayers]{};
  float rx[kMaxHmLayers]{}, ry[kMaxHmLayers]{};
  int upper_mm{0};
};

}  // namespace mfcl

struct mfcl_ctx {
  int device{0};
  int max_lds_per_block{65536};  // device limit on dynamic + static LDS per workgroup
  int n_cu{256};                 // compute units
  // launch-shape knobs (set by mfcl_tune / the autotuner)
  struct {
    int gjk_group{1};           // lanes per pair in the GJK stage (1, 16, 32, 64)
    int epa_threads{131072};    // persistent EPA lanes (multiple of 256): 64 per wavefront
    int guess_policy{1};        // warm start: 0 MLP direction for every pair, 1 not for predicted penetration
    int epa_lds_faces{64};      // face capacity of the LDS polytope (overflow continues or reruns at full size)
    int epa_spill{1};           // overflow continues in a device-memory polytope (1) or reruns (0)
    int epa_overflow_waves{512};   // persistent wavefronts (one per block) for overflow pairs
    int mesh_blocks_per_cu{3};  // persistent mesh-traversal blocks (4 waves, one query each) per CU
    int mesh_stack{16384};      // node-pair stack entries per wavefront
    int mlp_waves{8};           // wavefronts per block of the fused warm-start inference
    int lane_refill{48};        // lane kernel: 0 one pair per lane; k > 0 refill lanes when <= k run
  } tune;
  std::string err;
  hipStream_t stream{nullptr};

  // shape table
  uint32_t n_shapes{0};
  int32_t* d_type{nullptr};
  float4* d_param{nullptr};
  uint32_t* d_voff{nullptr};
  uint32_t* d_vcnt{nullptr};
  float4* d_verts{nullptr};
  double4* d_interior{nullptr};
  float* d_desc{nullptr};  // 16-float shape descriptors (warm-start MLP inputs)
  // convex faces (mfcl_upload_convex_faces): adjacency for the support walk
  int32_t* d_nbr{nullptr};      // adjacency tables in the reference's encoding
  int64_t* d_nbr_off{nullptr};  // per shape: table offset, -1 = naive scans; null = no faces
  int32_t* d_seed{nullptr};     // per shape: 8 ints, the walk's six axis seed vertices
  float4* d_wblk{nullptr};      // per vertex: 8 float4 walk blocks (record + 8 neighbours' x, y, z)
  uint64_t n_verts_alloc{0};    // float4 entries of d_verts
  std::vector<uint32_t> h_vcnt, h_voff;
  std::vector<int32_t> h_type;

  // mesh set
  uint32_t n_meshes{0};
  void* d_nodes{nullptr};   // mfcl::ObbNode[]
  void* d_tris{nullptr};    // float4 x 3 per triangle (vertices pre-gathered)
  uint32_t* d_mesh_node_base{nullptr};
  uint32_t* d_mesh_node_count{nullptr};
  uint32_t* d_mesh_tri_base{nullptr};
  float* d_mesh_verts{nullptr};         // current vertices (xyz), for refits and saves
  uint32_t* d_mesh_tri_idx{nullptr};    // triangles: 3 local vertex ids
  uint32_t* d_node_range{nullptr};      // per node: first leaf-order slot (mesh-relative), count
  uint32_t* d_mesh_prim{nullptr};       // per leaf-order slot: mesh-local triangle id
  uint32_t* d_refit_large{nullptr};     // nodes refitted by a whole wavefront
  mfcl::MeshHost mesh_host;

  // broadphase tree
  uint32_t bp_n_nodes{0};
  void* d_bp_nodes{nullptr};
  int64_t* d_bp_ids{nullptr};

  // heightmap (mfcl_upload_heightmap): every layer's uint16 mm heights in one
  // array, top layer first
  uint16_t* d_hm{nullptr};
  mfcl::HeightmapHost hm;
  uint64_t arm_hm_pairs{0};  // heightmap candidate pairs of the last arm query

  // device workspaces, sized by mfcl_reserve / mfcl_reserve_arm / the mesh
  // upload (query entry points never grow them) and, for the host-pointer
  // wrappers, ws_io
  mfcl::DevBuf ws_convex, ws_arena, ws_io, ws_arm;
  uint64_t ws_convex_pairs{0};  // pairs ws_convex holds
  uint64_t ws_arm_configs{0};   // configurations ws_arm holds (for ws_arm_key's arm)
  uint64_t ws_arm_key[4]{};     // links, broadphase cap, max pairs per config, pair capacity
  uint32_t mesh_ws_waves{0}, mesh_ws_cap{0};  // stacks ws_arena holds
  uint32_t* d_qcounts{nullptr};  // EPA queue / overflow counts of the last convex query
  void* d_static_scratch{nullptr};  // this device's static workspace (mfcl_convex_query_static_dev)

  mfcl::MlpState mlp;

  // graph capture (mfcl_graph_begin .. mfcl_graph_end)
  bool capturing{false};
  bool timing_before_capture{false};

  // timing
  bool timing{false};
  hipEvent_t ev[16]{};
  float last_ms[6]{0, 0, 0, 0, 0, 0};
  uint64_t counters[6]{0, 0, 0, 0, 0, 0};
  uint64_t* d_counters{nullptr};  // device-side work counters
};

namespace mfcl {

int ensure(mfcl_ctx* ctx, DevBuf& b, size_t bytes);
// capi.cpp: frees the heightmap
void freeHeightmap(mfcl_ctx* ctx);
void freeAll(mfcl_ctx* ctx);
// hipMalloc + copy of `bytes` bytes (frees *dptr first).
int uploadVec(mfcl_ctx* ctx, void** dptr, const void* src, size_t bytes);
// convex_faces.cpp: adjacency tables for the support walk; narrowphase.hip:
// the walk's per-shape seed vertices (device pass over the uploaded hulls).
int uploadConvexFaces(mfcl_ctx* ctx, const int32_t* faces, const uint64_t* face_off, const uint32_t* num_faces,
                      uint64_t faces_len);
int computeWalkSeeds(mfcl_ctx* ctx);
// narrowphase.hip: the MPR penetration pass (directed or incremental) over
// a collide's results
int mprPenLaunch(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea, const float* poseb,
                 const float* dirs, uint64_t n, bool incremental, int max_it, double tol, mfcl_convex_result* out,
                 hipStream_t s);
// meshshape.hip: mesh-shape collision over host arrays
int meshShapeCollide(mfcl_ctx* ctx, const uint32_t* mesh_ids, const uint32_t* shape_ids, const float* pose_m7,
                     const float* pose_s7, uint64_t n, const mfcl_request* req, mfcl_mesh_result* out,
                     int32_t* tri_ids, mfcl_mesh_contact* contacts);
// capi.cpp: frees the shape table (convex tables included)
void freeShapes(mfcl_ctx* ctx);
// hull.hip
int uploadConvexPoints(mfcl_ctx* ctx, const float* points, const uint64_t* pt_off, const uint32_t* pt_cnt,
                       uint32_t n_sets, uint32_t* hull_cnt_out);
int hullsLaunch(mfcl_ctx* ctx, const float* pts, const uint64_t* pt_off, const uint32_t* pt_cnt, uint32_t n_sets,
                int max_points, uint32_t* hull_cnt, uint32_t* hull_idx, uint32_t* tri_cnt, int32_t* tris,
                int32_t* status, hipStream_t s);
int hullsLaunchClassed(mfcl_ctx* ctx, const float* pts, const uint64_t* pt_off, const uint32_t* pt_cnt,
                       const uint32_t* pt_cnt_host, uint32_t n_sets, uint32_t* order_dev, uint32_t* hull_cnt,
                       uint32_t* hull_idx, uint32_t* tri_cnt, int32_t* tris, int32_t* status, hipStream_t s);
// arm.hip
size_t armWsBytes(const mfcl_arm* arm, uint64_t n_cfg, uint64_t* max_pairs_out);
int armQueryLaunch(mfcl_ctx* ctx, const mfcl_arm* arm, const float* q, uint64_t n_cfg, float margin,
                   const float* box_pose7, int32_t* any_hit, float* min_dist, uint64_t* n_pairs_out, hipStream_t s);

// Convex-query workspace (narrowphase.hip): pieces carved from one buffer.
struct ConvexWs {
  uint32_t* misc;      // 2 KB of counters
  uint32_t* order;     // bucketed pair order
  uint8_t* key;        // bucket keys
  void* queue;         // EPA queue
  void* oflow;         // EPA overflow queue
  double* lane_dirs;   // lane kernel direction slots
  void* arena;         // EPA continuation polytopes (device memory), one per wavefront that may need one
  size_t arena_bytes;
};
// EPA continuation arena of the reserved workspace (2048 polytopes of 256
// faces) and of the static one (256 such polytopes).
size_t epaArenaBytes(int slices, int max_faces);
constexpr int kReservedArenaSlices = 2048, kStaticArenaSlices = 256;
size_t convexWsBytes(uint64_t n, size_t arena_bytes);
uint64_t convexWsMaxPairs(size_t bytes, size_t arena_bytes);
bool carveConvexWs(void* base, size_t bytes, uint64_t n, size_t arena_bytes, ConvexWs& w);
// The static workspace of mfcl_convex_query_static_dev (narrowphase.hip):
// 64 MiB per GPU; staticScratch() returns its device address on `device`
// (looked up once per device, with that device current).
constexpr size_t kStaticScratchBytes = size_t{64} << 20;
int staticScratch(int device, void** out);
uint64_t staticScratchMaxPairs();
// Grows ctx->ws_convex to hold n pairs (host call; may synchronise).
int reserveConvex(mfcl_ctx* ctx, uint64_t n);
// Mesh traversal stacks for the current tune knobs (bvh.hip; host call).
int reserveMesh(mfcl_ctx* ctx);
// bvh.hip: OBB trees of a mesh set on the host; upload of a built set
int buildMeshSet(std::string& err, const float* verts, const uint64_t* voff, const uint32_t* vcnt,
                 const uint32_t* tris, const uint64_t* toff, const uint32_t* tcnt, uint32_t n, MeshSetHost& out);
int uploadMeshSet(mfcl_ctx* ctx, const MeshSetHost& m);
int downloadMeshSet(mfcl_ctx* ctx, MeshSetHost& m);
int meshReplaceVertices(mfcl_ctx* ctx, uint32_t mesh, const float* verts, bool device_ptr, hipStream_t s);
int meshRebuild(mfcl_ctx* ctx, uint32_t mesh, const float* verts_host);
// meshio.cpp
int loadObj(const char* path, std::vector<float>& v, std::vector<uint32_t>& t, std::string& err);
int writeMeshFile(const MeshSetHost& m, const char* path, std::string& err);
int readMeshFile(const char* path, MeshSetHost& m, std::string& err);

// Launchers (narrowphase.hip, bvh.hip, broadphase.hip).
int convexQueryLaunch(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                      const float* poseb, const float* guess, uint64_t n, const mfcl_request* req,
                      mfcl_convex_result* out, mfcl_contact* contacts, const ConvexWs& ws, hipStream_t s);
int meshCollideLaunch(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                      const float* poseb, uint64_t n, int32_t max_contacts, mfcl_mesh_result* out,
                      int32_t* contacts, mfcl_mesh_contact* pen_contacts, hipStream_t s);
int aabbQueryLaunch(mfcl_ctx* ctx, const float* qmn, const float* qmx, uint64_t nq, int32_t cap,
                    int32_t* counts, int64_t* ids, hipStream_t s);

// Warm-start MLP (mlp.hip).
int mlpInit(mfcl_ctx* ctx, uint64_t seed);
int mlpSetParams(mfcl_ctx* ctx, const float* host, int n, bool reset_opt);
int mlpGetParams(mfcl_ctx* ctx, float* host, int n, float* grad_host);
int mlpFeatures(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea, const float* poseb,
                uint64_t n, float* xf_out, hipStream_t s);
int mlpGuess(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea, const float* poseb,
             uint64_t n, float* guess3, float* dist, hipStream_t s);
int mlpForwardFeat(mfcl_ctx* ctx, const float* xf, uint64_t n, float* y4_out, hipStream_t s);
int mlpReserve(mfcl_ctx* ctx, uint64_t rows);
int mlpTrainStep(mfcl_ctx* ctx, const float* xf, const float* target4, uint64_t n, const float* hyper,
                 int apply_update, float* loss_out, hipStream_t s);
void mlpFree(mfcl_ctx* ctx);

}  // namespace mfcl
