// Broadphase AABB tree for gfx950.
//
// Host build: median split of box centres along the longest axis of the
// union box (the reference's rule, include/fcl/broadphase/
// binary_AABB_tree-inl.h:138-160), emitted breadth-first into a flat array of
// 32-byte FP32 nodes with siblings adjacent.
// Device query: one lane per query box, short stack in registers, children
// pushed right-then-left so leaves are met in the reference's order
// (binary_AABB_tree-inl.h:400-440); leaf ids are written up to `cap` per query.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ctx.h"

namespace mfcl {

struct alignas(16) AabbNode {
  float mn[3];
  int32_t child;  // >= 0: first child (sibling at child + 1); < 0: leaf -(object + 1)
  float mx[3];
  int32_t pad;
};
static_assert(sizeof(AabbNode) == 32, "node must be 32 bytes");

constexpr int kBpStack = 64;

int buildAndUploadAabbTree(mfcl_ctx* ctx, const float* mn, const float* mx, const int64_t* ids, uint32_t n) {
  struct Obj {
    float mn[3], mx[3];
    uint32_t idx;
  };
  std::vector<Obj> objs(n);
  for (uint32_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      objs[i].mn[k] = mn[3 * i + k];
      objs[i].mx[k] = mx[3 * i + k];
      if (!(objs[i].mn[k] <= objs[i].mx[k])) {
        ctx->err = "box " + std::to_string(i) + " has min > max";
        return MFCL_E_INVALID;
      }
    }
    objs[i].idx = i;
  }
  std::vector<AabbNode> nodes;
  nodes.reserve(2 * n);
  struct R { uint32_t node, first, count; };
  std::vector<R> fifo;
  fifo.reserve(2 * n);
  nodes.emplace_back();
  fifo.push_back({0, 0, n});
  int depth_guard = 0;
  for (size_t h = 0; h < fifo.size(); ++h) {
    const R r = fifo[h];
    AabbNode nd;
    for (int k = 0; k < 3; ++k) {
      nd.mn[k] = objs[r.first].mn[k];
      nd.mx[k] = objs[r.first].mx[k];
    }
    for (uint32_t i = 1; i < r.count; ++i)
      for (int k = 0; k < 3; ++k) {
        nd.mn[k] = std::min(nd.mn[k], objs[r.first + i].mn[k]);
        nd.mx[k] = std::max(nd.mx[k], objs[r.first + i].mx[k]);
      }
    nd.pad = 0;
    if (r.count == 1) {
      nd.child = -static_cast<int32_t>(objs[r.first].idx) - 1;
      nodes[r.node] = nd;
      continue;
    }
    int axis = 0;
    const float ext[3] = {nd.mx[0] - nd.mn[0], nd.mx[1] - nd.mn[1], nd.mx[2] - nd.mn[2]};
    if (ext[1] > ext[0]) axis = 1;
    if (ext[2] > ext[axis]) axis = 2;
    const uint32_t half = r.count / 2;
    std::nth_element(objs.begin() + r.first, objs.begin() + r.first + half, objs.begin() + r.first + r.count,
                     [axis](const Obj& a, const Obj& b) {
                       return 0.5f * (a.mn[axis] + a.mx[axis]) < 0.5f * (b.mn[axis] + b.mx[axis]);
                     });
    const uint32_t left = static_cast<uint32_t>(nodes.size());
    nodes.emplace_back();
    nodes.emplace_back();
    nd.child = static_cast<int32_t>(left);
    nodes[r.node] = nd;
    fifo.push_back({left, r.first, half});
    fifo.push_back({left + 1, r.first + half, r.count - half});
    (void)depth_guard;
  }
  // depth bound for the fixed device stack: balanced median split gives
  // depth <= ceil(log2 n) + 1, and the stack never holds more than depth + 1
  if (n > (1u << 30)) {
    ctx->err = "too many boxes";
    return MFCL_E_INVALID;
  }
  if (ctx->d_bp_nodes) (void)hipFree(ctx->d_bp_nodes);
  if (ctx->d_bp_ids) (void)hipFree(ctx->d_bp_ids);
  ctx->d_bp_nodes = nullptr;
  ctx->d_bp_ids = nullptr;
  MFCL_HIP_TRY(ctx, hipMalloc(&ctx->d_bp_nodes, nodes.size() * sizeof(AabbNode)));
  MFCL_HIP_TRY(ctx, hipMemcpy(ctx->d_bp_nodes, nodes.data(), nodes.size() * sizeof(AabbNode), hipMemcpyHostToDevice));
  MFCL_HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->d_bp_ids), n * sizeof(int64_t)));
  MFCL_HIP_TRY(ctx, hipMemcpy(ctx->d_bp_ids, ids, n * sizeof(int64_t), hipMemcpyHostToDevice));
  ctx->bp_n_nodes = static_cast<uint32_t>(nodes.size());
  return MFCL_OK;
}

__global__ void __launch_bounds__(256) aabbQueryKernel(const AabbNode* nodes, uint32_t n_nodes, const int64_t* ids,
                                                       const float* qmn, const float* qmx, uint64_t nq, int cap,
                                                       int32_t* counts, int64_t* out) {
  const uint64_t q = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const float a0 = qmn[3 * q], a1 = qmn[3 * q + 1], a2 = qmn[3 * q + 2];
  const float b0 = qmx[3 * q], b1 = qmx[3 * q + 1], b2 = qmx[3 * q + 2];
  uint32_t stk[kBpStack];
  int sp = 0;
  stk[sp++] = 0;
  int cnt = 0;
  while (sp > 0) {
    const uint32_t i = stk[--sp];
    if (i >= n_nodes) {
      cnt = -1;  // corrupt tree: report instead of reading out of bounds
      break;
    }
    const AabbNode nd = nodes[i];
    // AABB-inl.h:82-88
    if (a0 > nd.mx[0] || a1 > nd.mx[1] || a2 > nd.mx[2]) continue;
    if (b0 < nd.mn[0] || b1 < nd.mn[1] || b2 < nd.mn[2]) continue;
    if (nd.child < 0) {
      if (cnt < cap) out[q * static_cast<uint64_t>(cap) + cnt] = ids[-nd.child - 1];
      ++cnt;
    } else {
      if (sp + 2 > kBpStack) {
        cnt = -1;
        break;
      }
      stk[sp++] = static_cast<uint32_t>(nd.child) + 1;
      stk[sp++] = static_cast<uint32_t>(nd.child);
    }
  }
  counts[q] = cnt;
}

int aabbQueryLaunch(mfcl_ctx* ctx, const float* qmn, const float* qmx, uint64_t nq, int32_t cap, int32_t* counts,
                    int64_t* ids, hipStream_t s) {
  if (nq == 0) return MFCL_OK;
  if (ctx->bp_n_nodes == 0) {
    ctx->err = "no AABB tree built";
    return MFCL_E_STATE;
  }
  const unsigned nb = static_cast<unsigned>((nq + 255) / 256);
  if (ctx->timing) MFCL_HIP_TRY(ctx, hipEventRecord(ctx->ev[6], s));
  aabbQueryKernel<<<nb, 256, 0, s>>>(static_cast<const AabbNode*>(ctx->d_bp_nodes), ctx->bp_n_nodes,
                                     ctx->d_bp_ids, qmn, qmx, nq, cap, counts, ids);
  if (ctx->timing) MFCL_HIP_TRY(ctx, hipEventRecord(ctx->ev[7], s));
  MFCL_HIP_TRY(ctx, hipGetLastError());
  return MFCL_OK;
}

}  // namespace mfcl
