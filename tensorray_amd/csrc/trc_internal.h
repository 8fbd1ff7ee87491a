// Internal state behind the C ABI: context and resident scene.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/trc.h"
#include "common/tr_layout.h"
#include "host/scene_host.h"

struct trc_ctx {
  int device = 0;
  int num_cus = 0;
  std::string last_error;
  // traversal stack spill per stream ([depth][threads] of the largest grid),
  // allocated on a stream's first launch; guarded by mu
  std::map<hipStream_t, uint32_t*> spill;
  std::mutex mu;
  // this device's copy of the fixed scratch of trc_render_fwd_noalloc
  // (hipGetSymbolAddress, looked up once per device at context creation)
  void* scratch = nullptr;
  // job tables of batched launches: a ring of pinned staging + device
  // buffers, each reused once the batch that last used it has finished
  struct BatchSlot {
    void* host = nullptr;
    void* dev = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
  };
  static constexpr int kBatchSlots = 4;
  BatchSlot batch[kBatchSlots];
  unsigned batch_next = 0;
  std::mutex batch_mu;
};

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

namespace trc {
// Device-side scene update (kernels/update.hip).
// Edges per block of the order-preserving edge selection.
constexpr int kEdgeTile = 1024;
// Scalars the update writes on the device; the host reads them back once.
struct UpdateScalars {
  EdgeCount sec;          // secondary edges: count and length integral
  float light_integral;   // Distribution1D integral of the emitter table
  float pad;
  double sah;             // SAH cost of the refit tree (rebuild decision)
};
// Device arrays of the update, owned by the scene.
struct SceneUpdateDev {
  const EdgeDev* all_edges = nullptr;  // every mesh edge (global ids), edge order
  uint32_t n_edges = 0;
  const uint32_t* vs_off = nullptr;    // smooth faces per vertex (CSR, triangle order) or null
  const uint32_t* vs_list = nullptr;
  const uint32_t* refit_nodes = nullptr;  // BVH nodes by level, deepest first
  std::vector<uint32_t> level_off;        // host copy of the level offsets
  uint32_t n_nodes = 0;
  float* nbox = nullptr;                  // [nodes][6] unpadded node boxes
  uint32_t* scan = nullptr;               // [edge_scan_words(n_edges)] block counts / offsets
  float* part = nullptr;                  // [ceil(n_edges / kCdfChunk)] CDF chunk sums / offsets
  UpdateScalars* scal = nullptr;
};
}  // namespace trc

struct trc_scene {
  trc_ctx* ctx = nullptr;
  trc::SceneHost host;
  bool configured = false;
  uint64_t version = 0;           // bumped by every change of the device scene (plans check it)
  std::vector<DevBuf> bufs;       // all device allocations of this scene
  trc::SceneDev dev{};
  trc::TexDesc* dev_texs = nullptr;
  int64_t device_bytes = 0;
  // device-side update of the derived data and the BVH (trc_scene_set_vertices)
  trc::SceneUpdateDev upd;
  trc::UpdateScalars* upd_host = nullptr;  // pinned readback of upd.scal
  double sah_built = 0.0;                  // SAH cost of the tree when it was last built
  int32_t refits = 0, rebuilds = 0;
  // the host copies of positions, derived data and BVH boxes lag behind the
  // device after a device update (sync_host brings them up to date)
  bool host_stale = false;
  // primary-edge list of the last camera, selected on the device (capacity:
  // every edge)
  trc::EdgeDev* prim_e = nullptr;
  float* prim_len = nullptr;
  float* prim_cdf = nullptr;
  trc::EdgeCount* prim_cnt = nullptr;
  // guiding of the direct boundary term: per-sample values, leaf weights,
  // leaf CDF and the weight integral
  DevBuf guide_buf;
  // the guide in guide_buf was kept (trc_options.g_keep) for this key
  struct GuideKey {
    bool valid = false;
    uint64_t version = 0;
    uint32_t seed = 0;
    int32_t spp = 0, n = 0;
    float eps = 0.0f;
  } guide_key;
  // survivor queue of the primary / direct / indirect boundary kernels
  DevBuf bq_buf;
  // adjoint of the vertex normals [V][3] (kept zero between calls), present
  // when a differentiable mesh is smooth-shaded
  float* vn_adj = nullptr;
};

namespace trc {

// launch configuration shared by the megakernels
constexpr int kBlock = 128;
// Waves per SIMD the path kernels are compiled for (register budget:
// 512 / waves VGPRs).  Traversal is latency-bound, so the forward trades a
// few spilled registers for 5 resident waves (96 VGPRs); the adjoint keeps
// 3 (168 VGPRs).  Measured on the metric workload: forward 6.39 -> 5.09 ms
// going from 4 to 5 waves, 6.38 ms at 6; adjoint 2.15 -> 1.86 ms going from
// 2 to 3.
constexpr int kFwdWaves = 5;
constexpr int kAdjWaves = 3;
// The boundary (edge) kernels: 5 waves; primary / pixel / indirect edge
// terms at 512^2 x 4 spp on the metric scene: 2.35 / 2.53 / 3.09 ms at their
// natural 2 waves, 1.44 / 1.40 / 2.04 ms at 5.
constexpr int kBndWaves = 5;
constexpr int kStackLds = 16;
// Postponed leaves per lane (LDS).  Forward on the metric scene with
// lane-serial leaf tests at 8 / 12 / 16 entries: 4.69 / 4.68 / 5.90 ms at
// depth 4, 5.34 / 5.24 / 6.50 ms at depth 8 (16 entries leave room for
// fewer blocks per CU); 8 leave room for the wave-wide tests' slots.
constexpr int kLeafQueue = 8;
static_assert(kLeafQueue <= 15, "the queue length is kept in 4 bits of a query's state");
constexpr int kTraceLdsWords = kStackLds + kLeafQueue;  // LDS words per lane for traversal
// LDS words per lane of the wave-wide leaf tests' result slots (a 64-bit
// key and two barycentrics), after the traversal words
constexpr int kTraceSlotWords = 4;

struct LaunchCfg {
  int grid;             // persistent grid size (blocks)
  uint32_t* spill;      // spill buffer [depth][grid*kBlock]
  int cap;              // stack capacity
  int cus;              // compute units of the device
  uint32_t* tiles;      // the forward's tile counter (one word, zeroed per launch); null: tiles dealt statically
};

// A persistent grid-stride kernel should not launch more blocks than can be
// resident at once: blocks beyond that only start when resident ones have
// finished all their iterations, and then run at low occupancy.  Clamp the
// grid to (resident blocks per CU for this kernel) x CUs.
int resident_blocks_per_cu(const void* kernel, int block);  // cached occupancy query, 0 if unknown

template <class Kernel>
inline LaunchCfg fit_grid(Kernel kernel, LaunchCfg L, int block = kBlock) {
  const int per_cu = resident_blocks_per_cu((const void*)kernel, block);
  if (per_cu > 0 && L.cus > 0) {
    const int cap = per_cu * L.cus;
    if (L.grid > cap) L.grid = cap;
  }
  return L;
}

hipError_t launch_intersect(const SceneDev& S, const LaunchCfg& L, const float* rays, int n, int32_t* tri,
                            float* uvt, int any_hit, hipStream_t st);
// Privatized vertex-gradient accumulation: `copies` zeroed copies of the
// [V][3] gradient; block b adds into copy b % copies, and a reduction adds
// them into dpos.  Cuts float-atomic contention on small meshes, where every
// path lands on the same few hundred vertices.
struct GradCopies {
  float* base = nullptr;  // [copies][n3] or null (atomics go to dpos directly)
  uint32_t copies = 1;
  uint32_t n3 = 0;        // 3 * V
  uint32_t keep = 0;      // nonzero: the copies outlive the launch, and the reduction re-zeroes them
};
// Number of copies for a mesh of `n3` gradient floats (1 = none).
uint32_t grad_copies_for(uint64_t n3);
hipError_t launch_fill_zero(float* p, uint64_t n, hipStream_t st);
hipError_t launch_adam(float* x, const float* g, float* m, float* v, uint64_t n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, hipStream_t st);
// vmax_bits: one device word of scratch
hipError_t launch_adam_uniform(float* x, const float* g, float* m, float* v, uint64_t n, float lr, float b1, float b2,
                               float bc1, float bc2, unsigned int* vmax_bits, hipStream_t st);
// Sparse symmetric matrix (CSR) and the device scalars of a 3-column
// conjugate-gradient solve.
struct CsrDev {
  int n;
  const int* row;    // [n + 1]
  const int* col;    // [nnz]
  const float* val;  // [nnz]
};
// Blocks of the conjugate-gradient kernels (grid-stride over the rows); each
// writes its share of a dot product to a partial, and the partials are
// summed in block order, so the solve is deterministic.
constexpr int kCgBlocks = 1024;
struct CgScalars {
  float rr[3], rr_new[3], pap[3], rr0[3];  // rr0: |b|^2 per column
  unsigned int active;  // bit c: column c not converged
  unsigned int iters;
  float part_pap[kCgBlocks][3];  // p . Ap
  float part_rr[kCgBlocks][3];   // r . r
  float part_b[kCgBlocks][3];    // b . b
};
hipError_t launch_csr_spmv3(const CsrDev& M, const float* x, float* y, hipStream_t st);
// M x = b for the three columns of [n][3] b, x (x: initial guess on entry);
// r, p, ap: [n][3] work vectors; stops a column once |r| <= tol |b|.
hipError_t launch_cg(const CsrDev& M, const float* b, float* x, float* r, float* p, float* ap, CgScalars* s,
                     int max_iter, float tol, hipStream_t st);
hipError_t launch_grad_reduce(const GradCopies& g, float* dpos, hipStream_t st);

// One job of a batched launch (device copy of trc_job).
struct JobDev {
  SceneDev S;
  CameraDev C;
  uint32_t seed;
  int32_t max_bounces, spp;
  int32_t solo;            // launched on its own (single-request kernels over its range), no tiles
  uint64_t first, count;   // path range of the job
  uint64_t tile0;          // first tile of the job in the batch
  float* paths;
  void* ckpt;
  const float* dLdI;
  float* dpos;
  float* dmat;
  float* dtex;
  float* vn;
  GradCopies gc;
};
// tiles of `tile` paths; jobs sorted by tile0
hipError_t launch_render_batch(const JobDev* jobs_dev, const JobDev* jobs_host, int njobs, uint64_t ntiles,
                               uint32_t tile, int adjoint, const LaunchCfg& L, hipStream_t st);

// ckpt: null or a path checkpoint of checkpoint_bytes(spp*W*H, max_bounces)
// bytes.  first / count: the range of path indices rendered (count 0: all).
// paths_first: the path index whose radiance goes to paths[0] (0: paths holds
// the request's whole [spp*W*H][3]).
hipError_t launch_render_paths(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                               const LaunchCfg& L, float* paths, void* ckpt, hipStream_t st, uint64_t first = 0,
                               uint64_t count = 0, uint64_t paths_first = 0);
uint64_t checkpoint_bytes(uint64_t npaths, int max_bounces);
hipError_t launch_render_adj_ckpt(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                                  const LaunchCfg& L, const void* ckpt, const float* dLdI, float* dpos, float* dmat,
                                  float* dtex, float* vn_adj, const GradCopies& gc, hipStream_t st,
                                  uint64_t first = 0, uint64_t count = 0, bool fold = true);
hipError_t launch_render_stats(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                               const LaunchCfg& L, unsigned long long* counters, hipStream_t st);
hipError_t launch_film_reduce(const float* paths, int spp, int wh, float* film, hipStream_t st);
// One chunk of the same reduction: paths holds samples [s0, s0 + ns) of every
// pixel; film carries the running (unscaled) sums between chunks, so chunking
// the samples gives the same bits as launch_film_reduce.
hipError_t launch_film_accum(const float* paths, int ns, int wh, float* film, bool first_chunk, bool last_chunk,
                             int spp, hipStream_t st);
// Fixed device scratch of the allocation-free forward (one array per device,
// held until the process ends): its size, and its address on the current
// device (hipGetSymbolAddress).
size_t noalloc_scratch_bytes();
hipError_t noalloc_scratch_address(void** out);
// ---- device-side scene update (kernels/update.hip) ----
// Words of the block-count scratch of an edge selection over n_edges edges.
uint32_t edge_scan_words(uint32_t n_edges);
// Recomputes face normals and areas, vertex normals, the emitter table, the
// secondary-edge list, the triangle records and the BVH boxes from S.pos.
hipError_t launch_scene_update(const SceneDev& S, const SceneUpdateDev& U, hipStream_t st);
// Primary edges seen from `eye` (boundary or silhouette) into out_* (capacity
// U.n_edges), their count and length integral into *count.
hipError_t launch_prim_edges(const SceneDev& S, const SceneUpdateDev& U, const float eye[3], EdgeDev* out_e,
                             float* out_len, float* out_cdf, EdgeCount* count, hipStream_t st);
// Guide of the direct boundary term: one leaf per secondary edge i over
// x in [sec.cdf[i-1], sec.cdf[i]], y, z in [0, 1].
struct GuideDev {
  float* vals;  // [n * spp] per-sample |baseVal| / spp
  float* poly;  // [n] leaf weights after the eps mix
  float* cdf;   // [n] leaf CDF (weight * volume)
  float* fsum;  // [1] sum of weight * volume
  float* part;  // [ceil(n / kSumChunk) + 1] chunk sums / offsets of the blocked reductions
  int32_t n, spp;
  float eps;
  int32_t enabled;
};
// Reductions over the guide's leaves run in a fixed blocked order: chunks of
// kSumChunk consecutive values summed sequentially, then the chunk sums
// sequentially; a prefix is the sequential prefix of the preceding chunk
// sums plus the sequential prefix within the chunk.  (The reference leaves
// the order to its reduction library.)  Up to kSumChunk values it is the
// plain sequential order.
constexpr int kSumChunk = 1024;
// Builds the guide on `st`; g.enabled is ignored.
hipError_t launch_guide_build(const SceneDev& S, const LaunchCfg& L, uint32_t seed, const GuideDev& g,
                              hipStream_t st);
// Bytes of the survivor queue launch_boundary needs for `n_samples` samples
// of `kind` (0 for the pixel boundary, which runs in one pass).
uint64_t boundary_queue_bytes(int kind, uint64_t n_samples);
// kind: 0 primary edges, 1 pixel boundary, 2 secondary direct, 3 secondary
// indirect; queue: boundary_queue_bytes(kind, spp * W * H) bytes of scratch.
// Gradient mode (dimg null): dpos += d/dV of <dLdI, boundary term>.
// Tangent mode: dimg [3][W*H] += derivative image along tpos [V][3].
hipError_t launch_boundary(int kind, const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                           const LaunchCfg& L, const EdgeSet& prim, const GuideDev& guide, const float* dLdI,
                           float* dpos, const float* tpos, float* dimg, void* queue, hipStream_t st);
// Interior derivative image: dimg [3][W*H] += d film / d theta along tpos
// [V][3] and tmat [M][8] (or null); tvn [V][3] scratch for the vertex-normal
// tangents when a differentiable mesh is smooth-shaded, else null.
hipError_t launch_render_deriv(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                               const LaunchCfg& L, const float* tpos, const float* tmat, float* tvn, float* dimg,
                               hipStream_t st);
// vn_adj: [V][3] zeroed scratch for smooth-shaded differentiable meshes or
// null; it is folded into dpos and cleared again by a second kernel.
hipError_t launch_render_adj(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                             const LaunchCfg& L, const float* dLdI, float* dpos, float* dmat, float* dtex,
                             float* vn_adj, const GradCopies& gc, hipStream_t st);

}  // namespace trc
