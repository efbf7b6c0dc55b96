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

struct MeshHost {
  std::vector<uint32_t> node_base;  // first node of each mesh
  std::vector<uint32_t> node_count;
};

}  // namespace mfcl

struct mfcl_ctx {
  int device{0};
  // launch-shape knobs (set by mfcl_tune / the autotuner)
  struct {
    int gjk_group{16};          // lanes per pair in the GJK stage (1, 8, 16)
    int epa_threads{32768};     // persistent EPA lanes (multiple of 256): 64 per wavefront
    int epa_lds_faces{64};      // face capacity of the LDS polytope (overflow reruns at full size)
    int epa_overflow_waves{4096};  // persistent wavefronts for overflow pairs
    int mesh_blocks_per_cu{4};  // persistent mesh-traversal blocks (4 waves) per CU
    int mesh_stack{16384};      // node-pair stack entries per wavefront
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
  std::vector<uint32_t> h_vcnt;

  // mesh set
  uint32_t n_meshes{0};
  void* d_nodes{nullptr};   // mfcl::ObbNode[]
  void* d_tris{nullptr};    // float4 x 3 per triangle (vertices pre-gathered)
  uint32_t* d_mesh_node_base{nullptr};
  uint32_t* d_mesh_node_count{nullptr};
  uint32_t* d_mesh_tri_base{nullptr};
  mfcl::MeshHost mesh_host;

  // broadphase tree
  uint32_t bp_n_nodes{0};
  void* d_bp_nodes{nullptr};
  int64_t* d_bp_ids{nullptr};

  // growable device workspaces, keyed by purpose
  mfcl::DevBuf ws_order, ws_queue, ws_arena, ws_misc, ws_io, ws_oflow;

  // timing
  bool timing{false};
  hipEvent_t ev[16]{};
  float last_ms[6]{0, 0, 0, 0, 0, 0};
  uint64_t counters[4]{0, 0, 0, 0};
  uint64_t* d_counters{nullptr};  // device-side work counters
};

namespace mfcl {

int ensure(mfcl_ctx* ctx, DevBuf& b, size_t bytes);
void freeAll(mfcl_ctx* ctx);

// Launchers (narrowphase.hip, bvh.hip, broadphase.hip).
int convexQueryLaunch(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                      const float* poseb, const float* guess, uint64_t n, const mfcl_request* req,
                      mfcl_convex_result* out, hipStream_t s);
int meshCollideLaunch(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                      const float* poseb, uint64_t n, int32_t max_contacts, mfcl_mesh_result* out,
                      int32_t* contacts, hipStream_t s);
int aabbQueryLaunch(mfcl_ctx* ctx, const float* qmn, const float* qmx, uint64_t nq, int32_t cap,
                    int32_t* counts, int64_t* ids, hipStream_t s);

}  // namespace mfcl
