// Batched convex narrowphase for gfx950.
//
// Pipeline per query (one stream, no host round trip):
//   1. costKey / bucketScan / bucketScatter: counting sort of the pairs by
//      (closed-form vs support-mapping, log2 of hull vertex count) so the
//      lanes of a wavefront run loops of similar trip count;
//   2. narrowKernel: one lane per pair -- closed forms, MPR, GJK boolean or
//      GJK distance; pairs that need EPA append their simplex to a queue;
//   3. epaWaveKernel: persistent grid, one wavefront per queued pair, polytope
//      in LDS; pairs that outgrow the LDS arena rerun with the full capacity
//      in device memory.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ctx.h"
#include "epa_dev.h"
#include "gjk_lane.h"
#include "prim_dev.h"

namespace mfcl {

constexpr int kBuckets = 128;
constexpr int kBlock = 256;

struct PairIO {
  const uint32_t* pa;
  const uint32_t* pb;
  const float* posea;
  const float* poseb;
  const float* guess;
  mfcl_convex_result* out;
  mfcl_contact* contacts;  // optional: cstride records per pair
  int cstride;             // min(max_contacts, 4) when contacts is set
};

struct QEntry {
  SmpD s;
  uint32_t pair;
  uint32_t pad;
};

struct Params {
  int mode, max_contacts, penetration, gjk_max_it, epa_max_faces, epa_max_it, use_guess;
  uint32_t n_shapes;
  double gjk_tol, epa_tol;
};

MFCL_D int shapeCost(const ShapeTables& T, uint32_t i) {
  return T.type[i] == kConvex ? static_cast<int>(T.vert_cnt[i]) : 1;
}

MFCL_D bool closedForm(int ta, int tb) {
  if (ta == kSphere && (tb == kSphere || tb == kCapsule || tb == kBox || tb == kCylinder)) return true;
  if (tb == kSphere && (ta == kCapsule || ta == kBox || ta == kCylinder)) return true;
  return ta == kBox && tb == kBox;
}

// Closed-form pairs first (4 buckets per octave of the summed vertex
// count), then the others by how many of the two shapes walk (0, 1, 2: a
// wavefront's walks and cooperative scans then rarely mix) and, within that,
// 3 buckets per octave of the summed vertex count.
MFCL_D uint8_t pairKey(const ShapeTables& T, const PairIO& io, uint64_t i, const Params& P) {
  const uint32_t a = io.pa[i], b = io.pb[i];
  if (a >= P.n_shapes || b >= P.n_shapes) return 0;
  const bool cf = P.mode == MFCL_MODE_COLLIDE && closedForm(T.type[a], T.type[b]);
  const int cost = shapeCost(T, a) + shapeCost(T, b);
  if (cf) return static_cast<uint8_t>(min(63, static_cast<int>(4.0f * log2f(static_cast<float>(cost)))));
  auto walks = [&](uint32_t k) { return T.type[k] == kConvex && T.nbr_off && T.nbr_off[k] >= 0; };
  const int c = (walks(a) ? 1 : 0) + (walks(b) ? 1 : 0);
  const int lvl = max(0, min(20, static_cast<int>(3.0f * log2f(static_cast<float>(cost))) - 10));
  return static_cast<uint8_t>(64 + 21 * c + lvl);
}

// Block-local histogram in LDS, one global atomic per bin per block.
__global__ void __launch_bounds__(kBlock) costKeyKernel(ShapeTables T, PairIO io, uint64_t n, Params P,
                                                        uint8_t* key, uint32_t* hist) {
  __shared__ uint32_t h[kBuckets];
  if (threadIdx.x < kBuckets) h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    const uint8_t k = pairKey(T, io, i, P);
    key[i] = k;
    atomicAdd(&h[k], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kBuckets && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

__global__ void bucketScanKernel(uint32_t* hist, uint32_t* offs) {
  // single block of kBuckets threads: exclusive scan
  __shared__ uint32_t s[kBuckets];
  const int t = threadIdx.x;
  s[t] = hist[t];
  __syncthreads();
  if (t == 0) {
    uint32_t acc = 0;
    for (int k = 0; k < kBuckets; ++k) {
      const uint32_t v = s[k];
      s[k] = acc;
      acc += v;
    }
  }
  __syncthreads();
  offs[t] = s[t];
}

// Each block reserves its bins' ranges with one atomic per bin, then places
// its pairs with LDS atomics (order inside a bin is irrelevant).
__global__ void __launch_bounds__(kBlock) bucketScatterKernel(const uint8_t* key, uint64_t n, uint32_t* offs,
                                                              uint32_t* order) {
  __shared__ uint32_t cnt[kBuckets], base[kBuckets];
  if (threadIdx.x < kBuckets) cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  uint8_t k = 0;
  uint32_t local = 0;
  if (i < n) {
    k = key[i];
    local = atomicAdd(&cnt[k], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kBuckets && cnt[threadIdx.x]) base[threadIdx.x] = atomicAdd(&offs[threadIdx.x], cnt[threadIdx.x]);
  __syncthreads();
  if (i < n) order[base[k] + local] = static_cast<uint32_t>(i);
}

MFCL_D void clearResult(mfcl_convex_result& r) {
  r.hit = 0;
  r.n_contacts = 0;
  r.method = 0;
  r.gjk_status = -1;
  r.epa_status = -1;
  r.ok = 0;
  r.gjk_iters = 0;
  r.epa_iters = 0;
  r.distance = 0;
  r.depth = 0;
  r.normal[0] = r.normal[1] = r.normal[2] = 0;
  r.pos[0] = r.pos[1] = r.pos[2] = 0;
  r.pa[0] = r.pa[1] = r.pa[2] = 0;
  r.pb[0] = r.pb[1] = r.pb[2] = 0;
}

// Result fields gathered in registers and stored once.
struct ResAcc {
  d3 normal, pos, pa, pb;
};
MFCL_D void storeResult(mfcl_convex_result* dst, const mfcl_convex_result& r, const ResAcc& x) {
  mfcl_convex_result o = r;
  o.normal[0] = x.normal.x; o.normal[1] = x.normal.y; o.normal[2] = x.normal.z;
  o.pos[0] = x.pos.x; o.pos[1] = x.pos.y; o.pos[2] = x.pos.z;
  o.pa[0] = x.pa.x; o.pa[1] = x.pa.y; o.pa[2] = x.pa.z;
  o.pb[0] = x.pb.x; o.pb[1] = x.pb.y; o.pb[2] = x.pb.z;
  *dst = o;
}

MFCL_D uint16_t sat16(int v) { return static_cast<uint16_t>(v > 65535 ? 65535 : v); }

// Closed-form dispatch with the reference's operand swaps and normal flips
// (gjk_solver-inl.h:158-217).  Returns hit; fills contacts if requested.
MFCL_D bool runClosedForm(const DevShape& A, const tf3& ta, const DevShape& B, const tf3& tb, PContacts* pc) {
  const int x = A.type, y = B.type;
  bool hit = false, flip = false;
  if (x == kSphere && y == kSphere) hit = primSphereSphere(A.p0, ta, B.p0, tb, pc);
  else if (x == kBox && y == kBox) hit = primBoxBox(d3{A.p0, A.p1, A.p2}, ta, d3{B.p0, B.p1, B.p2}, tb, pc) != 0;
  else if (x == kSphere && y == kCapsule) hit = primSphereCapsule(A.p0, ta, B.p0, B.p1, tb, pc);
  else if (x == kCapsule && y == kSphere) { hit = primSphereCapsule(B.p0, tb, A.p0, A.p1, ta, pc); flip = true; }
  else if (x == kSphere && y == kBox) hit = primSphereBox(A.p0, ta, d3{B.p0, B.p1, B.p2}, tb, pc);
  else if (x == kBox && y == kSphere) { hit = primSphereBox(B.p0, tb, d3{A.p0, A.p1, A.p2}, ta, pc); flip = true; }
  else if (x == kSphere && y == kCylinder) hit = primSphereCylinder(A.p0, ta, B.p0, B.p1, tb, pc);
  else if (x == kCylinder && y == kSphere) { hit = primSphereCylinder(B.p0, tb, A.p0, A.p1, ta, pc); flip = true; }
  if (flip && pc)
    for (int k = 0; k < pc->n; ++k) pc->c[k].n = -pc->c[k].n;
  return hit;
}

MFCL_D void storeContact(mfcl_contact* dst, const d3& n, const d3& p, double depth) {
  mfcl_contact c;
  c.normal[0] = n.x; c.normal[1] = n.y; c.normal[2] = n.z;
  c.pos[0] = p.x; c.pos[1] = p.y; c.pos[2] = p.z;
  c.depth = depth;
  *dst = c;
}

// Contacts of a colliding closed-form pair into the result: the deepest
// max_contacts when they do not all fit (shape_pair_intersect-inl.h:91-113),
// contact 0 into the record, all kept ones into the contact array.
MFCL_D void finishClosedForm(PContacts& pc, const Params& P, const PairIO& io, uint32_t i, mfcl_convex_result& r,
                             ResAcc& acc, bool writer) {
  int kept = pc.n;
  if (pc.n > P.max_contacts) {
    partialSortDeepest(pc.c, P.max_contacts, pc.n);
    kept = P.max_contacts;
  }
  r.n_contacts = kept;
  if (kept > 0) {
    r.depth = pc.c[0].depth;
    r.distance = -r.depth;
    acc.normal = pc.c[0].n;
    acc.pos = pc.c[0].p;
  }
  if (io.contacts && writer)
    for (int k = 0; k < kept && k < io.cstride; ++k)
      storeContact(io.contacts + static_cast<uint64_t>(i) * io.cstride + k, pc.c[k].n, pc.c[k].p, pc.c[k].depth);
}

MFCL_D void loadPair(const ShapeTables& T, const double4* interior, const PairIO& io, uint32_t i,
                     DevShape& A, DevShape& B, tf3& ta, tf3& tb, uint32_t& ia, uint32_t& ib) {
  ia = io.pa[i];
  ib = io.pb[i];
  A = loadShape(T, ia);
  B = loadShape(T, ib);
  ta = poseFromQuat(io.posea + 7 * static_cast<uint64_t>(i));
  tb = poseFromQuat(io.poseb + 7 * static_cast<uint64_t>(i));
}

// LDS staging capacity (vertices per group, both hulls together): 48 KB of
// SoA float rows per 256-lane block whatever the group size.
template <int G>
constexpr int kStage = 48 * 1024 / 12 / (kBlock / G);

template <int G>
MFCL_D void stagePair(DevShape& A, DevShape& B, MFCL_LDS float* st, int lg) {
  int used = 0;
  if (A.type == kConvex && A.nv <= kStage<G>) {
    stageHull<G>(A, st, kStage<G>, lg);
    used = A.nv;
  }
  if (B.type == kConvex && B.nv <= kStage<G> - used) stageHull<G>(B, st + used, kStage<G>, lg);
  // the group's lanes read each other's LDS writes: LDS executes a
  // wavefront's operations in order; this keeps the compiler from moving
  // the reads above the writes
  __builtin_amdgcn_wave_barrier();
}

// One group of G lanes per pair (G divides 64); all lanes of a group run
// the same FP64 logic on identical values, share the support scans, and
// lane 0 of the group writes the result / queues the EPA work.
constexpr int kModeBool = 0, kModePen = 1, kModeSd = 2;

template <int G, int MODE>
__global__ void __launch_bounds__(kBlock) narrowKernel(ShapeTables T, const double4* interior, PairIO io,
                                                       const uint32_t* order, uint64_t n, Params P,
                                                       QEntry* queue, uint32_t* qcount,
                                                       unsigned long long* scanned_total) {
  // Per group: the Minkowski configuration and an LDS copy of the two hulls
  // (SoA rows of kStage<G> floats) when they fit, so the repeated support
  // scans of GJK read LDS instead of device memory.
  __shared__ MinkCfg cfgs[kBlock / G];
  __shared__ float stage[kBlock / G][3 * kStage<G>];
  const uint64_t slot = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
  const int lg = static_cast<int>(threadIdx.x % G);
  MFCL_LDS MinkCfg* cfg = (MFCL_LDS MinkCfg*)(&cfgs[threadIdx.x / G]);
  int scanned = 0;
  if (slot < n) {
    const uint32_t i = order[n - 1 - slot];  // costly pairs first, as laneKernel
    mfcl_convex_result r;
    clearResult(r);
    ResAcc acc;
    acc.normal = acc.pos = acc.pa = acc.pb = d3{0, 0, 0};
    const bool bad_ids = io.pa[i] >= P.n_shapes || io.pb[i] >= P.n_shapes;
    bool queued = false;
    Smp s;
    smpInit(s);
    if (bad_ids) {
      r.ok = -1;  // shape id out of range
    } else {
      DevShape A, B;
      tf3 ta, tb;
      uint32_t ia, ib;
      loadPair(T, interior, io, i, A, B, ta, tb, ia, ib);
      if constexpr (MODE != kModeSd) {
        if (P.max_contacts <= 0) {
          // the reference warns and adds nothing (collision-inl.h:81-84)
        } else if (closedForm(A.type, B.type)) {
          PContacts pc;
          pc.n = 0;
          const bool hit = runClosedForm(A, ta, B, tb, MODE == kModePen ? &pc : nullptr);
          r.hit = hit;
          r.ok = 1;
          if (hit) {
            if constexpr (MODE == kModeBool) {
              r.n_contacts = 1;
            } else {
              finishClosedForm(pc, P, io, i, r, acc, lg == 0);
            }
          }
        } else {
          d3 int0 = {0, 0, 0}, int1 = {0, 0, 0};
          if (A.type == kConvex) { const double4 c = interior[ia]; int0 = d3{c.x, c.y, c.z}; }
          if (B.type == kConvex) { const double4 c = interior[ib]; int1 = d3{c.x, c.y, c.z}; }
          stagePair<G>(A, B, (MFCL_LDS float*)(&stage[threadIdx.x / G][0]), lg);
          writeCfg(cfg, A, ta, B, tb, int0, int1);
          MinkL<G> md{cfg, lg, 0};
          r.method = 1;
          bool decided = false;
          if constexpr (MODE == kModeBool) {
            r.method = 2;
            const int m = mprIntersect(md, P.gjk_max_it, P.gjk_tol);
            if (m != kMprFailed) {
              r.hit = m == kMprIntersect;
              r.n_contacts = r.hit;
              r.ok = 1;
              decided = true;
            }
          }
          if (!decided) {
            d3 guess = {1, 0, 0};
            if (P.use_guess) guess = d3{io.guess[3 * i], io.guess[3 * i + 1], io.guess[3 * i + 2]};
            GjkOut g;
            gjkEvaluate(md, s, -guess, P.gjk_tol, P.gjk_max_it, false, g);
            r.gjk_status = static_cast<int8_t>(g.status);
            r.gjk_iters = sat16(g.iters);
            r.ok = 1;
            if constexpr (MODE == kModeBool) {
              r.hit = g.status == kGjkIntersect;
              r.n_contacts = r.hit;
            } else if (g.status == kGjkIntersect) {
              queued = true;
            }
          }
          scanned = md.scanned;
        }
      } else {
        // signed distance: no closed forms (gjk_solver-inl.h:820-874)
        stagePair<G>(A, B, (MFCL_LDS float*)(&stage[threadIdx.x / G][0]), lg);
        writeCfg(cfg, A, ta, B, tb, d3{0, 0, 0}, d3{0, 0, 0});
        MinkL<G> md{cfg, lg, 0};
        r.method = 1;
        d3 guess = {1, 0, 0};
        if (P.use_guess) guess = d3{io.guess[3 * i], io.guess[3 * i + 1], io.guess[3 * i + 2]};
        GjkOut g;
        gjkEvaluate(md, s, -guess, P.gjk_tol, P.gjk_max_it, true, g);
        r.gjk_status = static_cast<int8_t>(g.status);
        r.gjk_iters = sat16(g.iters + g.dist_iters);
        if (g.status == kGjkSeparated && g.sep_valid) {
          acc.pa = apply(ta, g.p0);
          acc.pb = apply(ta, g.p1);
          r.distance = norm(g.p0 - g.p1);
          r.ok = 1;
        } else if (g.status == kGjkIntersect) {
          queued = true;
        } else {
          r.distance = -1;
        }
        scanned = md.scanned;
      }
    }
    if (lg == 0) {
      if (queued) {
        const uint32_t q = atomicAdd(qcount, 1u);
        // element-wise stores: an aggregate copy out of the simplex would
        // become a memcpy that keeps the simplex in private memory
        QEntry* qe = &queue[q];
        qe->s.rank = s.rank;
        storeMVD(qe->s.a, s.a.v, slotDir(cfg, s.a.id));
        storeMVD(qe->s.b, s.b.v, slotDir(cfg, s.b.id));
        storeMVD(qe->s.c, s.c.v, slotDir(cfg, s.c.id));
        storeMVD(qe->s.e, s.e.v, slotDir(cfg, s.e.id));
        qe->pair = i;
      }
      // queued pairs keep their GJK status here; the EPA stage completes them
      storeResult(&io.out[i], r, acc);
    }
  }
  if (lg != 0) scanned = 0;
  // wave-level reduction of the scan counter, one atomic per wave
  unsigned long long v = static_cast<unsigned long long>(scanned);
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(scanned_total, v);
}

// Lane-kernel pieces.  laneSetup: pair i's closed form or invalid ids are
// finished at once (false); otherwise the lane's Minkowski difference is set
// up and GJK's start direction returned (true).
template <int MODE>
MFCL_D bool laneSetup(const ShapeTables& T, const PairIO& io, const Params& P, uint32_t i, MinkLane& md,
                      d3& guess) {
  guess = d3{1, 0, 0};
  mfcl_convex_result r;
  clearResult(r);
  ResAcc acc;
  acc.normal = acc.pos = acc.pa = acc.pb = d3{0, 0, 0};
  const uint32_t ia = io.pa[i], ib = io.pb[i];
  if (ia >= P.n_shapes || ib >= P.n_shapes) {
    r.ok = -1;  // shape id out of range
  } else {
    const int ta_ = T.type[ia], tb_ = T.type[ib];
    if (MODE == kModePen && P.max_contacts <= 0) {
      // the reference warns and adds nothing (collision-inl.h:81-84)
    } else if (MODE == kModePen && closedForm(ta_, tb_)) {
      const tf3 ta = poseFromQuat(io.posea + 7 * static_cast<uint64_t>(i));
      const tf3 tb = poseFromQuat(io.poseb + 7 * static_cast<uint64_t>(i));
      PContacts pc;
      pc.n = 0;
      const bool hit = runClosedForm(loadShape(T, ia), ta, loadShape(T, ib), tb, &pc);
      r.hit = hit;
      r.ok = 1;
      if (hit) finishClosedForm(pc, P, io, i, r, acc, true);
    } else {
      md.ia = ia;
      md.ib = ib;
      md.type0 = ta_;
      md.type1 = tb_;
      md.setTf0(relTf(poseFromQuat(io.posea + 7 * static_cast<uint64_t>(i)),
                      poseFromQuat(io.poseb + 7 * static_cast<uint64_t>(i))));
      if (P.use_guess) guess = d3{io.guess[3 * i], io.guess[3 * i + 1], io.guess[3 * i + 2]};
      return true;
    }
  }
  storeResult(&io.out[i], r, acc);
  return false;
}

// laneStore: pair i's result after its GJK (and its EPA queue entry when it
// intersects; the EPA stage completes those).
template <int MODE>
MFCL_D void laneStore(const PairIO& io, uint32_t i, const MinkLane& md, const LaneM& M, QEntry* queue,
                      uint32_t* qcount) {
  const GjkOut& g = M.o;
  const Smp& s = M.s;
  mfcl_convex_result r;
  clearResult(r);
  ResAcc acc;
  acc.normal = acc.pos = acc.pa = acc.pb = d3{0, 0, 0};
  r.method = 1;
  r.gjk_status = static_cast<int8_t>(g.status);
  bool queued = false;
  if constexpr (MODE == kModeSd) {
    r.gjk_iters = sat16(g.iters + g.dist_iters);
    if (g.status == kGjkSeparated && g.sep_valid) {
      const tf3 ta = poseFromQuat(io.posea + 7 * static_cast<uint64_t>(i));
      acc.pa = apply(ta, g.p0);
      acc.pb = apply(ta, g.p1);
      r.distance = norm(g.p0 - g.p1);
      r.ok = 1;
    } else if (g.status == kGjkIntersect) {
      queued = true;
    } else {
      r.distance = -1;
    }
  } else {
    r.gjk_iters = sat16(g.iters);
    r.ok = 1;
    queued = g.status == kGjkIntersect;
  }
  if (queued) {
    QEntry* qe = &queue[atomicAdd(qcount, 1u)];
    qe->s.rank = s.rank;
    storeMVD(qe->s.a, s.a.v, md.dirOf(static_cast<int>(s.a.id)));
    storeMVD(qe->s.b, s.b.v, md.dirOf(static_cast<int>(s.b.id)));
    storeMVD(qe->s.c, s.c.v, md.dirOf(static_cast<int>(s.c.id)));
    storeMVD(qe->s.e, s.e.v, md.dirOf(static_cast<int>(s.e.id)));
    qe->pair = i;
  }
  // queued pairs keep their GJK status here; the EPA stage completes them
  storeResult(&io.out[i], r, acc);
}

// One lane per pair (penetration and signed-distance modes): closed forms
// per lane, GJK through the lane machine of gjk_lane.h.  Results, EPA queue
// entries and scan counts equal those of narrowKernel.  dir_ws holds
// kLaneDirDoubles per wavefront of the grid.
template <int MODE>
__global__ void __launch_bounds__(kBlock) laneKernel(ShapeTables T, PairIO io, const uint32_t* order, uint64_t n,
                                                     Params P, QEntry* queue, uint32_t* qcount,
                                                     unsigned long long* scanned_total, double* dir_ws) {
  __shared__ LaneWs ws[kBlock / 64];
  const uint64_t slot = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  MinkLane md;
  md.T = T;
  md.w = (MFCL_LDS LaneWs*)(&ws[threadIdx.x >> 6]);
  md.dirs = dir_ws + (slot >> 6) * kLaneDirDoubles;
  md.lane = static_cast<int>(threadIdx.x & 63);
  md.ia = md.ib = 0;
  md.type0 = md.type1 = kSphere;
  uint32_t i = 0;
  bool run_gjk = false;
  d3 guess = {1, 0, 0};
  if (slot < n) {
    // the bucketed order runs from cheap to costly pairs: the grid takes it
    // from the costly end so that the largest pairs start first
    i = order[n - 1 - slot];
    run_gjk = laneSetup<MODE>(T, io, P, i, md, guess);
  }
  uint64_t scanned = 0;
  int walked = 0;
  LaneM M;
  laneInit(M, run_gjk, -guess);
  laneRun(M, md, P.gjk_tol, P.gjk_max_it, MODE == kModeSd, scanned, walked, 0);
  laneFinish(M, md);
  if (run_gjk) laneStore<MODE>(io, i, md, M, queue, qcount);
  // cooperative scans are counted once per wavefront, walks per lane
  unsigned long long v = static_cast<unsigned long long>(walked);
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  v += scanned;
  if (md.lane == 0 && v) atomicAdd(scanned_total, v);
}

// The same with refill: a persistent grid (two wavefronts per SIMD) whose
// lanes take pairs from a shared cursor (the costly end of the bucketed
// order first).  When at most refill_at lanes of a wavefront are still
// running, the finished lanes write their results and take new pairs, so a
// wavefront does not wait on its slowest pair.  Each pair's result and EPA
// queue entry equal laneKernel's (a pair's computation does not depend on
// its lane or wavefront).  The penetration-mode kernel is bounded to two
// wavefronts per SIMD (it spills a little to fit); the signed-distance one
// fits without the bound.
template <int MODE>
__device__ __attribute__((always_inline)) inline void laneRefillBody(
    ShapeTables T, PairIO io, const uint32_t* order, uint64_t n, Params P, QEntry* queue, uint32_t* qcount,
    unsigned long long* scanned_total, double* dir_ws, uint32_t* cursor, int refill_at, MFCL_LDS LaneWs* ws) {
  const uint64_t wave = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  MinkLane md;
  md.T = T;
  md.w = ws + (threadIdx.x >> 6);
  md.dirs = dir_ws + wave * kLaneDirDoubles;
  md.lane = static_cast<int>(threadIdx.x & 63);
  md.ia = md.ib = 0;
  md.type0 = md.type1 = kSphere;
  const unsigned long long below = (1ull << md.lane) - 1;
  uint64_t scanned = 0;
  int walked = 0;
  LaneM M;
  laneInit(M, false, d3{1, 0, 0});
  uint32_t i = 0;
  bool have = false;       // the lane holds a pair in GJK
  bool exhausted = false;  // the cursor has passed n (wavefront-uniform)
  while (true) {
    // lanes without a pair take the next ones; closed forms finish at once
    // and their lanes take again
    while (!exhausted) {
      const unsigned long long need = __ballot(!have);
      if (need == 0ull) break;
      const uint32_t cnt = static_cast<uint32_t>(__popcll(need));
      uint32_t base = 0;
      if (md.lane == 0) base = atomicAdd(cursor, cnt);
      base = __builtin_amdgcn_readfirstlane(base);
      if (static_cast<uint64_t>(base) + cnt >= n) exhausted = true;
      if (!have) {
        const uint64_t k = static_cast<uint64_t>(base) + static_cast<uint64_t>(__popcll(need & below));
        if (k < n) {
          i = order[n - 1 - k];
          d3 guess;
          if (laneSetup<MODE>(T, io, P, i, md, guess)) {
            laneInit(M, true, -guess);
            have = true;
          }
        }
      }
    }
    if (__ballot(have) == 0ull) break;
    laneRun(M, md, P.gjk_tol, P.gjk_max_it, MODE == kModeSd, scanned, walked, exhausted ? 0 : refill_at);
    if (have && M.st == kLsDone) {
      laneFinish(M, md);
      laneStore<MODE>(io, i, md, M, queue, qcount);
      have = false;
    }
  }
  unsigned long long v = static_cast<unsigned long long>(walked);
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  v += scanned;
  if (md.lane == 0 && v) atomicAdd(scanned_total, v);
}

// three wavefronts per SIMD: the spills this costs (168 VGPRs) are cheaper
// than the latency left exposed at two; four spill too much
__global__ void __launch_bounds__(kBlock, 3) laneKernelRSd(ShapeTables T, PairIO io, const uint32_t* order, uint64_t n,
                                                        Params P, QEntry* queue, uint32_t* qcount,
                                                        unsigned long long* scanned_total, double* dir_ws,
                                                        uint32_t* cursor, int refill_at) {
  __shared__ LaneWs ws[kBlock / 64];
  laneRefillBody<kModeSd>(T, io, order, n, P, queue, qcount, scanned_total, dir_ws, cursor, refill_at,
                          (MFCL_LDS LaneWs*)(&ws[0]));
}
__global__ void __launch_bounds__(kBlock, 2) laneKernelRPen(ShapeTables T, PairIO io, const uint32_t* order,
                                                            uint64_t n, Params P, QEntry* queue, uint32_t* qcount,
                                                            unsigned long long* scanned_total, double* dir_ws,
                                                            uint32_t* cursor, int refill_at) {
  __shared__ LaneWs ws[kBlock / 64];
  laneRefillBody<kModePen>(T, io, order, n, P, queue, qcount, scanned_total, dir_ws, cursor, refill_at,
                           (MFCL_LDS LaneWs*)(&ws[0]));
}

// EPA stage: one wavefront per queued pair; the 64 lanes share every
// support scan and run the polytope bookkeeping in lockstep on identical
// values.  Each wavefront owns an LDS slice: the pair's Minkowski
// configuration, the closest-feature record and the polytope arena with
// capacity cap_faces.  A pair that exhausts the arena (status MallocFailed
// while cap_faces < epa_max_faces) is re-queued to `overflow` and rerun by a
// second launch with the requested capacity (one wavefront per block), so
// every result equals a run at full capacity.
constexpr size_t kCfgBytes = (sizeof(MinkCfg) + 15) & ~static_cast<size_t>(15);
constexpr size_t kRawBytes = (sizeof(RawF) + 15) & ~static_cast<size_t>(15);
constexpr size_t kPolyBytes = (sizeof(Poly) + 15) & ~static_cast<size_t>(15);
inline size_t epaSliceBytes(int cap_faces) { return kCfgBytes + kRawBytes + kPolyBytes + arenaBytes(cap_faces); }

// Wavefront w of the grid may continue a pair that outgrows its LDS
// polytope in the device-memory polytope arena + w * arena_slice (full
// capacity) when w < arena_slices; the others rerun such a pair.
struct EpaArena {
  char* base;
  uint32_t slices;
  size_t slice;
  uint32_t* spilled;  // pairs continued in device memory
};

template <bool SPILL>
__global__ void __launch_bounds__(kBlock, 2) epaWaveKernel(ShapeTables T, PairIO io, Params P, const QEntry* queue,
                                                        const uint32_t* qcount, int cap_faces, size_t slice_bytes,
                                                        QEntry* overflow, uint32_t* ocount, uint32_t* next,
                                                        unsigned long long* scanned_total, EpaArena ga) {
  extern __shared__ __align__(16) char lds_arena[];
  const int lane = threadIdx.x & 63;
  const ArenaDims dims = arenaDims(cap_faces);
  MFCL_LDS char* slice = (MFCL_LDS char*)(lds_arena) + static_cast<size_t>(threadIdx.x >> 6) * slice_bytes;
  MFCL_LDS MinkCfg* cfg = (MFCL_LDS MinkCfg*)(slice);
  MFCL_LDS RawF* raw = (MFCL_LDS RawF*)(slice + kCfgBytes);
  MFCL_LDS char* base = slice + kCfgBytes + kRawBytes + kPolyBytes;
  // the polytope header (list heads, free lists) lives in LDS too, so the
  // out-of-line helpers that take it by reference do not pin it in scratch
  Poly& poly = *reinterpret_cast<Poly*>(lds_arena + static_cast<size_t>(threadIdx.x >> 6) * slice_bytes +
                                        kCfgBytes + kRawBytes);
  poly.V = (MFCL_LDS EVert*)(base);
  poly.E = (MFCL_LDS EEdge*)(base + sizeof(EVert) * dims.nv);
  poly.F = (MFCL_LDS EFace*)(base + sizeof(EVert) * dims.nv + sizeof(EEdge) * dims.nv);
  poly.BL = (MFCL_LDS BorderRec*)(base + sizeof(EVert) * dims.nv + sizeof(EEdge) * dims.nv +
                                  sizeof(EFace) * dims.nf);
  poly.nv = dims.nv;
  poly.nf = dims.nf;
  // this wavefront's device-memory polytope of the request's full capacity
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const bool can_spill = cap_faces < P.epa_max_faces && ga.base != nullptr && wave < ga.slices;
  const SpillTarget st{can_spill ? ga.base + static_cast<size_t>(wave) * ga.slice : nullptr, P.epa_max_faces};
  const uint32_t count = *qcount;
  int scanned = 0;
  // pairs are taken one at a time from a shared counter: EPA iteration
  // counts vary widely (p99 / mean ~ 3), so a static share per wavefront
  // would leave the grid waiting on its most loaded wavefronts.  The GJK
  // stage queues the pairs roughly in descending cost order (bucketed by
  // hull sizes), so the largest pairs, which tend to need the most EPA
  // iterations, start first and the tail shortens.
  for (;;) {
    uint32_t q = 0;
    if (lane == 0) q = atomicAdd(next, 1u);
    q = __shfl(q, 0);
    if (q >= count) break;
    const QEntry e = queue[q];
    const uint32_t i = e.pair;
    DevShape A, B;
    tf3 ta, tb;
    uint32_t ia, ib;
    loadPair(T, nullptr, io, i, A, B, ta, tb, ia, ib);
    writeCfg(cfg, A, ta, B, tb, d3{0, 0, 0}, d3{0, 0, 0});
    MinkL<64> md{cfg, lane, 0};
    EpaOut eo;
    epaEvaluate<SPILL>(poly, md, e.s, P.epa_tol, P.epa_max_it, eo, *raw, st);
    scanned += md.scanned + eo.scanned_extra;
    if (eo.spilled && lane == 0) atomicAdd(ga.spilled, 1u);
    if (eo.status == kEpaMallocFailed && cap_faces < P.epa_max_faces) {
      if (lane == 0) {
        const uint32_t o = atomicAdd(ocount, 1u);
        overflow[o] = e;
      }
      continue;
    }
    if (lane != 0) continue;
    mfcl_convex_result& r = io.out[i];
    r.epa_status = static_cast<int8_t>(eo.status);
    r.epa_iters = sat16(eo.iters);
    if (eo.status == kEpaFailed) {
      r.hit = 0;
      r.ok = 0;
      continue;
    }
    r.hit = 1;
    r.ok = 1;
    if (P.mode == MFCL_MODE_COLLIDE) {
      // gjk_solver-inl.h:117-129
      d3 n1 = eo.p0 - eo.p1;
      n1 = sqnorm(n1) <= 0 ? d3{0, 0, 1} : normalized(n1);
      const d3 pt = 0.5 * (eo.p0 + eo.p1);
      r.n_contacts = 1;
      r.depth = eo.depth;
      r.distance = -eo.depth;
      const d3 nw = mul(ta.R, n1), pw = apply(ta, pt);
      r.normal[0] = nw.x; r.normal[1] = nw.y; r.normal[2] = nw.z;
      r.pos[0] = pw.x; r.pos[1] = pw.y; r.pos[2] = pw.z;
      if (io.contacts) storeContact(io.contacts + static_cast<uint64_t>(i) * io.cstride, nw, pw, eo.depth);
    } else {
      const d3 wa = apply(ta, eo.p0), wb = apply(ta, eo.p1);
      r.pa[0] = wa.x; r.pa[1] = wa.y; r.pa[2] = wa.z;
      r.pb[0] = wb.x; r.pb[1] = wb.y; r.pb[2] = wb.z;
      r.distance = -eo.depth;
      r.depth = eo.depth;
    }
  }
  if (lane != 0) scanned = 0;
  unsigned long long v = static_cast<unsigned long long>(scanned);
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (lane == 0 && v) atomicAdd(scanned_total, v);
}

// MPR penetration (fcl::collisionPenetration in its DirectedPenetration or
// IncrementalMinimumPenetration mode, collision_penetration-inl.h:196-244):
// after the boolean collide, one wavefront per pair it reported colliding
// runs the MPR search from the pair's direction (world unit vector: the
// escape direction of shape b, or the hint) and replaces the pair's
// contact (computePenetrationMPR, :105-194): depth = the distance along the
// direction / the refined minimum penetration (-1 when the search fails),
// pos = the witness midpoint (zero on failure), normal = the direction /
// the refined direction in the world.  epa_status / epa_iters hold the MPR
// status and its iterations.
template <bool INC>
__global__ void __launch_bounds__(kBlock) mprPenKernel(ShapeTables T, PairIO io, const float* dirs, uint64_t n,
                                                       int max_it, double tol) {
  __shared__ MinkCfg cfgs[kBlock / 64];
  const uint64_t i = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (i >= n) return;  // uniform across the wavefront
  mfcl_convex_result& r = io.out[i];
  if (r.n_contacts <= 0) return;
  MFCL_LDS MinkCfg* cfg = (MFCL_LDS MinkCfg*)(&cfgs[threadIdx.x >> 6]);
  DevShape A, B;
  tf3 ta, tb;
  uint32_t ia, ib;
  loadPair(T, nullptr, io, static_cast<uint32_t>(i), A, B, ta, tb, ia, ib);
  writeCfg(cfg, A, ta, B, tb, d3{0, 0, 0}, d3{0, 0, 0});
  MinkL<64> md{cfg, lane, 0};
  const d3 dw = {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
  const d3 dl = tmul(ta.R, dw);
  int st, iters;
  bool ok;
  double depth = 0;
  d3 p0 = {0, 0, 0}, p1 = p0, nw = dw;
  if constexpr (INC) {
    IncOut o;
    o.depth = 0;
    o.dir = o.p0 = o.p1 = d3{0, 0, 0};
    st = mprIncremental(md, dl, max_it, tol, o);
    iters = o.iters;
    ok = st == kIncOK || st == kIncIterLimit;
    if (ok) {
      depth = o.depth;
      p0 = o.p0;
      p1 = o.p1;
      nw = mul(ta.R, o.dir);
    }
  } else {
    DirOut o;
    o.distance = 0;
    o.p0 = o.p1 = d3{0, 0, 0};
    st = mprDirected(md, dl, max_it, tol, o);
    iters = o.iters;
    ok = st == kDirOK;
    if (ok) {
      depth = o.distance;
      p0 = o.p0;
      p1 = o.p1;
    }
  }
  if (lane != 0) return;
  r.epa_status = static_cast<int8_t>(st);
  r.epa_iters = sat16(iters);
  r.normal[0] = nw.x; r.normal[1] = nw.y; r.normal[2] = nw.z;
  if (ok) {
    const d3 pw = apply(ta, 0.5 * (p0 + p1));
    r.depth = depth;
    r.pos[0] = pw.x; r.pos[1] = pw.y; r.pos[2] = pw.z;
  } else {
    r.depth = -1;
    r.pos[0] = r.pos[1] = r.pos[2] = 0;
  }
  r.distance = -r.depth;
}

int mprPenLaunch(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea, const float* poseb,
                 const float* dirs, uint64_t n, bool incremental, int max_it, double tol, mfcl_convex_result* out,
                 hipStream_t s) {
  ShapeTables T{ctx->d_type, ctx->d_param, ctx->d_voff, ctx->d_vcnt, ctx->d_verts, ctx->d_nbr, ctx->d_nbr_off,
                ctx->d_seed, ctx->d_wblk};
  PairIO io{pa, pb, posea, poseb, nullptr, out, nullptr, 0};
  const unsigned blocks = static_cast<unsigned>((64 * n + kBlock - 1) / kBlock);
  if (incremental) mprPenKernel<true><<<blocks, kBlock, 0, s>>>(T, io, dirs, n, max_it, tol);
  else mprPenKernel<false><<<blocks, kBlock, 0, s>>>(T, io, dirs, n, max_it, tol);
  MFCL_HIP_TRY(ctx, hipGetLastError());
  return MFCL_OK;
}

__global__ void zeroKernel(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}



// Seed vertices of the support walk (convex-inl.h:151-199): per walking
// hull, the first vertex attaining the maximum of +x, -x, +y, -y, +z, -z
// (convex-inl.h:133-149; the dot product with a unit axis is the coordinate).
__global__ void walkSeedKernel(ShapeTables T, uint32_t n, int* seed) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || T.nbr_off[i] < 0) return;
  const float4* v = T.verts + T.vert_off[i];
  const int nv = static_cast<int>(T.vert_cnt[i]);
  int idx[6] = {0, 0, 0, 0, 0, 0};
  double best[6];
  {
    const float4 q = v[0];
    best[0] = q.x; best[1] = -static_cast<double>(q.x);
    best[2] = q.y; best[3] = -static_cast<double>(q.y);
    best[4] = q.z; best[5] = -static_cast<double>(q.z);
  }
  for (int k = 1; k < nv; ++k) {
    const float4 q = v[k];
    const double c[6] = {q.x, -static_cast<double>(q.x), q.y, -static_cast<double>(q.y), q.z,
                         -static_cast<double>(q.z)};
#pragma unroll
    for (int a = 0; a < 6; ++a)
      if (c[a] > best[a]) {
        best[a] = c[a];
        idx[a] = k;
      }
  }
#pragma unroll
  for (int a = 0; a < 6; ++a) seed[8 * static_cast<uint64_t>(i) + a] = idx[a];
  seed[8 * static_cast<uint64_t>(i) + 6] = seed[8 * static_cast<uint64_t>(i) + 7] = 0;
}

// Per walking hull (one block each, its threads over the vertices): every
// vertex's neighbour-record offset goes into the vertex's fourth word (read
// together with the vertex by the table walk).
__global__ void walkRecKernel(ShapeTables T, float4* verts_rw) {
  const uint32_t i = blockIdx.x;
  if (T.nbr_off[i] < 0) return;
  const int nv = static_cast<int>(T.vert_cnt[i]);
  const int* tab = T.nbr + T.nbr_off[i];
  for (int k = threadIdx.x; k < nv; k += blockDim.x) {
    const uint64_t g = T.vert_off[i] + static_cast<uint64_t>(k);
    verts_rw[g].w = __int_as_float(tab[k]);
  }
}

// Walk blocks (one block of threads per walking hull of at most 65535
// vertices): per vertex 8 float4 = (x, y, z, neighbour count), the first
// eight neighbour ids as 16-bit pairs, then those neighbours' x, y, z
// (24 floats, zero past the count).  Runs after walkRecKernel has written
// the vertices' fourth words; reads only their coordinates.
__global__ void walkBlkKernel(ShapeTables T, float4* wblk) {
  const uint32_t i = blockIdx.x;
  if (T.nbr_off[i] < 0) return;
  const int nv = static_cast<int>(T.vert_cnt[i]);
  if (nv > 65535) return;
  const int* tab = T.nbr + T.nbr_off[i];
  const float4* v = T.verts + T.vert_off[i];
  for (int k = threadIdx.x; k < nv; k += blockDim.x) {
    const uint64_t g = T.vert_off[i] + static_cast<uint64_t>(k);
    const int* rec = tab + tab[k];
    const int cnt = rec[0];
    uint32_t u[4] = {0, 0, 0, 0};
    float f[24];
#pragma unroll
    for (int j = 0; j < 24; ++j) f[j] = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (j >= cnt) break;
      const int nb = rec[1 + j];
      u[j >> 1] |= static_cast<uint32_t>(nb) << (16 * (j & 1));
      const float4 q = v[nb];
      f[3 * j] = q.x;
      f[3 * j + 1] = q.y;
      f[3 * j + 2] = q.z;
    }
    const float4 self = v[k];
    float4* o = wblk + 8 * g;
    o[0] = make_float4(self.x, self.y, self.z, __int_as_float(cnt));
    o[1] = make_float4(__uint_as_float(u[0]), __uint_as_float(u[1]), __uint_as_float(u[2]), __uint_as_float(u[3]));
#pragma unroll
    for (int j = 0; j < 6; ++j) o[2 + j] = make_float4(f[4 * j], f[4 * j + 1], f[4 * j + 2], f[4 * j + 3]);
  }
}

int computeWalkSeeds(mfcl_ctx* ctx) {
  ShapeTables T{ctx->d_type, ctx->d_param, ctx->d_voff, ctx->d_vcnt, ctx->d_verts, ctx->d_nbr, ctx->d_nbr_off,
                ctx->d_seed, ctx->d_wblk};
  const uint32_t n = ctx->n_shapes;
  walkSeedKernel<<<(n + 255) / 256, 256, 0, ctx->stream>>>(T, n, ctx->d_seed);
  walkRecKernel<<<n, 64, 0, ctx->stream>>>(T, ctx->d_verts);
  if (ctx->d_wblk) walkBlkKernel<<<n, 64, 0, ctx->stream>>>(T, ctx->d_wblk);
  MFCL_HIP_TRY(ctx, hipGetLastError());
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return MFCL_OK;
}

// Workspace of one convex query over n pairs, carved from one device buffer
// (256-byte aligned pieces): the counters block (2 KB: bucket histogram and
// offsets, EPA queue and overflow counts at byte 1024, EPA work cursors at
// 1032 / 1036, 64-bit work counters at 1536), the bucket order (4n B) and
// keys (n B), the EPA queue and overflow queue (one QEntry per pair each,
// the most that can be queued) and the lane kernel's direction slots
// (kLaneDirDoubles doubles per wavefront of its grid).
static size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }
static size_t laneDirBytes(uint64_t n) {
  const uint64_t waves = (n + kBlock - 1) / kBlock * (kBlock / 64);
  return static_cast<size_t>(waves) * kLaneDirDoubles * sizeof(double);
}
size_t epaArenaBytes(int slices, int max_faces) { return static_cast<size_t>(slices) * arenaBytes(max_faces); }
size_t convexWsBytes(uint64_t n, size_t arena_bytes) {
  if (n == 0) n = 1;
  return align256(2048) + align256(4 * n) + align256(n) + 2 * align256(n * sizeof(QEntry)) +
         align256(laneDirBytes(n)) + align256(arena_bytes);
}
uint64_t convexWsMaxPairs(size_t bytes, size_t arena_bytes) {
  // convexWsBytes grows by at least sizeof(QEntry) per pair: bisect
  uint64_t lo = 0, hi = bytes / (2 * sizeof(QEntry)) + 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) / 2;
    if (convexWsBytes(mid, arena_bytes) <= bytes) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
bool carveConvexWs(void* base, size_t bytes, uint64_t n, size_t arena_bytes, ConvexWs& w) {
  if (!base || convexWsBytes(n, arena_bytes) > bytes) return false;
  char* p = static_cast<char*>(base);
  auto take = [&](size_t b) {
    char* r = p;
    p += align256(b);
    return r;
  };
  w.misc = reinterpret_cast<uint32_t*>(take(2048));
  w.order = reinterpret_cast<uint32_t*>(take(4 * (n ? n : 1)));
  w.key = reinterpret_cast<uint8_t*>(take(n ? n : 1));
  w.queue = take((n ? n : 1) * sizeof(QEntry));
  w.oflow = take((n ? n : 1) * sizeof(QEntry));
  w.lane_dirs = reinterpret_cast<double*>(take(laneDirBytes(n ? n : 1)));
  w.arena = arena_bytes ? take(arena_bytes) : nullptr;
  w.arena_bytes = arena_bytes;
  return true;
}

// Workspace of mfcl_convex_query_static_dev, for callers that cannot
// allocate: a file-scope device array, one copy per GPU (its own address on
// each), held until the process ends.  Host code never passes the array's
// name to a kernel (on the host it is only a stand-in for the symbol): it
// uses the device address that staticScratch() looked up once per device.
static __device__ __align__(256) char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

int staticScratch(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxScratchDevices) return MFCL_E_INVALID;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    // the caller has made `device` current
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return MFCL_E_HIP;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return MFCL_OK;
}

uint64_t staticScratchMaxPairs() {
  return convexWsMaxPairs(kStaticScratchBytes, epaArenaBytes(kStaticArenaSlices, 256));
}

int convexQueryLaunch(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea,
                      const float* poseb, const float* guess, uint64_t n, const mfcl_request* req,
                      mfcl_convex_result* out, mfcl_contact* contacts, const ConvexWs& ws, hipStream_t s) {
  if (n == 0) return MFCL_OK;
  if (ctx->n_shapes == 0) {
    ctx->err = "no shapes uploaded";
    return MFCL_E_STATE;
  }
  if (n > 0xFFFFFFFFull) {
    ctx->err = "too many pairs in one query";
    return MFCL_E_INVALID;
  }
  if (req->epa_max_faces < 4 || req->epa_max_faces > 8192) {
    ctx->err = "epa_max_faces out of range [4, 8192]";
    return MFCL_E_INVALID;
  }
  if (req->use_guess && guess == nullptr) {
    ctx->err = "use_guess set but guess3 is NULL";
    return MFCL_E_INVALID;
  }
  Params P;
  P.mode = req->mode;
  P.max_contacts = req->max_contacts;
  P.penetration = req->penetration;
  P.gjk_max_it = req->gjk_max_iterations;
  P.gjk_tol = req->gjk_tolerance;
  P.epa_max_faces = req->epa_max_faces;
  P.epa_max_it = req->epa_max_iterations;
  P.epa_tol = req->epa_tolerance;
  P.use_guess = req->use_guess;
  P.n_shapes = ctx->n_shapes;

  if (contacts && !(P.mode == MFCL_MODE_COLLIDE && P.penetration)) {
    ctx->err = "a contact array needs a collide request with penetration";
    return MFCL_E_INVALID;
  }
  const size_t full_slice = epaSliceBytes(req->epa_max_faces);
  if (full_slice > static_cast<size_t>(ctx->max_lds_per_block)) {
    ctx->err = "epa_max_faces = " + std::to_string(req->epa_max_faces) + " needs " + std::to_string(full_slice) +
               " bytes of LDS per wavefront; this device allows " + std::to_string(ctx->max_lds_per_block);
    return MFCL_E_UNSUPPORTED;
  }
  uint32_t* hist = ws.misc;
  uint32_t* offs = hist + kBuckets;
  uint32_t* qcount = offs + kBuckets;
  uint32_t* ocount = qcount + 1;
  QEntry* oflow = static_cast<QEntry*>(ws.oflow);
  unsigned long long* sc = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(ws.misc) + 1536);
  ctx->d_counters = reinterpret_cast<uint64_t*>(sc);
  ctx->d_qcounts = qcount;
  uint32_t* order = ws.order;
  uint8_t* key = ws.key;
  QEntry* queue = static_cast<QEntry*>(ws.queue);

  ShapeTables T{ctx->d_type, ctx->d_param, ctx->d_voff, ctx->d_vcnt, ctx->d_verts, ctx->d_nbr, ctx->d_nbr_off,
                ctx->d_seed, ctx->d_wblk};
  PairIO io{pa, pb, posea, poseb, guess, out, contacts, contacts ? (P.max_contacts < 4 ? P.max_contacts : 4) : 0};
  const unsigned nb = static_cast<unsigned>((n + kBlock - 1) / kBlock);
  if (ctx->timing) MFCL_HIP_TRY(ctx, hipEventRecord(ctx->ev[0], s));
  zeroKernel<<<1, 512, 0, s>>>(hist, 512);
  costKeyKernel<<<nb, kBlock, 0, s>>>(T, io, n, P, key, hist);
  bucketScanKernel<<<1, kBuckets, 0, s>>>(hist, offs);
  bucketScatterKernel<<<nb, kBlock, 0, s>>>(key, n, offs, order);
  if (ctx->timing) MFCL_HIP_TRY(ctx, hipEventRecord(ctx->ev[1], s));
  {
    const int mode = P.mode == MFCL_MODE_SIGNED_DISTANCE ? kModeSd : (P.penetration ? kModePen : kModeBool);
    // the lane kernel has no boolean mode: boolean queries take 16-lane groups
    const int G = ctx->tune.gjk_group == 1 && mode == kModeBool ? 16 : ctx->tune.gjk_group;
    const unsigned gb = static_cast<unsigned>((n * G + kBlock - 1) / kBlock);
#define MFCL_NARROW(GG, MM) \
  narrowKernel<GG, MM><<<gb, kBlock, 0, s>>>(T, ctx->d_interior, io, order, n, P, queue, qcount, sc)
#define MFCL_NARROW_MODES(GG)                 \
  if (mode == kModeSd) MFCL_NARROW(GG, kModeSd); \
  else if (mode == kModePen) MFCL_NARROW(GG, kModePen); \
  else MFCL_NARROW(GG, kModeBool)
    if (G == 1) {
      const unsigned lb = static_cast<unsigned>((n + kBlock - 1) / kBlock);
      double* dws = ws.lane_dirs;
      const int refill = ctx->tune.lane_refill;
      if (refill > 0) {
        // persistent: as many blocks per CU as the kernel's launch bound
        // keeps resident (3 for distance, 2 for penetration), at most one
        // lane per pair
        const uint64_t per_cu = mode == kModeSd ? 3ull : 2ull;
        const unsigned rb = static_cast<unsigned>(std::min<uint64_t>(lb, per_cu * static_cast<uint64_t>(ctx->n_cu)));
        uint32_t* cursor = qcount + 8;  // zeroed with the histogram
        if (mode == kModeSd)
          laneKernelRSd<<<rb, kBlock, 0, s>>>(T, io, order, n, P, queue, qcount, sc, dws, cursor, refill);
        else
          laneKernelRPen<<<rb, kBlock, 0, s>>>(T, io, order, n, P, queue, qcount, sc, dws, cursor, refill);
      } else if (mode == kModeSd) {
        laneKernel<kModeSd><<<lb, kBlock, 0, s>>>(T, io, order, n, P, queue, qcount, sc, dws);
      } else {
        laneKernel<kModePen><<<lb, kBlock, 0, s>>>(T, io, order, n, P, queue, qcount, sc, dws);
      }
    } else switch (G) {
      case 32: MFCL_NARROW_MODES(32); break;
      case 64: MFCL_NARROW_MODES(64); break;
      default: MFCL_NARROW_MODES(16); break;
    }
#undef MFCL_NARROW_MODES
#undef MFCL_NARROW
  }
  if (ctx->timing) MFCL_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  {
    int cap_small = P.epa_max_faces < ctx->tune.epa_lds_faces ? P.epa_max_faces : ctx->tune.epa_lds_faces;
    while (cap_small > 4 && 4 * epaSliceBytes(cap_small) > static_cast<size_t>(ctx->max_lds_per_block)) cap_small /= 2;
    const size_t slice = epaSliceBytes(cap_small);
    const unsigned eb = static_cast<unsigned>(ctx->tune.epa_threads / kBlock);
    EpaArena ga{nullptr, 0, arenaBytes(P.epa_max_faces), ocount + 3};
    if (ws.arena && cap_small < P.epa_max_faces && ctx->tune.epa_spill) {
      ga.base = static_cast<char*>(ws.arena);
      ga.slices = static_cast<uint32_t>(std::min<size_t>(ws.arena_bytes / ga.slice, 4u * eb));
    }
    if (ga.base)
      epaWaveKernel<true><<<eb, kBlock, 4 * slice, s>>>(T, io, P, queue, qcount, cap_small, slice, oflow, ocount,
                                                        ocount + 1, sc + 1, ga);
    else
      epaWaveKernel<false><<<eb, kBlock, 4 * slice, s>>>(T, io, P, queue, qcount, cap_small, slice, oflow, ocount,
                                                         ocount + 1, sc + 1, ga);
    // pairs that outgrew the LDS polytope on a wavefront without a device-
    // memory polytope: rerun at full capacity, one wavefront per block
    if (cap_small < P.epa_max_faces) {
      const unsigned ob = static_cast<unsigned>(ctx->tune.epa_overflow_waves);
      EpaArena none{nullptr, 0, 0, ocount + 3};
      epaWaveKernel<false><<<ob, 64, full_slice, s>>>(T, io, P, oflow, ocount, P.epa_max_faces, full_slice, oflow,
                                                      ocount, ocount + 2, sc + 1, none);
    }
  }
  if (ctx->timing) MFCL_HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
  MFCL_HIP_TRY(ctx, hipGetLastError());
  return MFCL_OK;
}

}  // namespace mfcl
