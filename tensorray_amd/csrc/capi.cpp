// C ABI implementation (include/trc.h): context, resident scenes, launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>

#include "trc_internal.h"

using namespace trc;

namespace {

struct Fail : std::runtime_error {
  trc_status code;
  Fail(trc_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    trc_status c = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? TRC_ERR_NOMEM : TRC_ERR_HIP;
    throw Fail(c, std::string(what) + ": " + hipGetErrorString(e));
  }
}

template <class F>
trc_status guarded(trc_ctx* ctx, F&& f) {
  try {
    f();
    return TRC_OK;
  } catch (const Fail& e) {
    if (ctx) ctx->last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    if (ctx) ctx->last_error = e.what();
    return TRC_ERR_INVALID;
  }
}

constexpr int kSpillDepth = 64;
static_assert(kStackLds + kSpillDepth <= 255, "a query's stack depth is kept in 8 bits of TraceState");
constexpr int kBlocksPerCU = 2 * kFwdWaves;  // 128-lane blocks: 2 waves each, 4 SIMDs

// Address of the allocation-free forward's scratch on `device` (the current
// device), looked up once per device: each GPU holds its own copy of the
// array at its own address.
void* device_scratch(int device) {
  static std::mutex mu;
  static std::map<int, void*> addr;
  std::lock_guard<std::mutex> lk(mu);
  auto it = addr.find(device);
  if (it != addr.end()) return it->second;
  void* p = nullptr;
  hip_check(trc::noalloc_scratch_address(&p), "hipGetSymbolAddress(scratch)");
  if (!p) throw Fail(TRC_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
  addr.emplace(device, p);
  return p;
}

}  // namespace

namespace trc {
int resident_blocks_per_cu(const void* kernel, int block) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_pair(kernel, block);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess) per_cu = 0;
  cache.emplace(key, per_cu);
  return per_cu;
}
}  // namespace trc

namespace {

// Stream-ordered zeroed gradient copies for a scene of n_verts vertices
// (none when the mesh is large enough that atomics rarely collide).
GradCopies make_grad_copies(int n_verts, hipStream_t st) {
  GradCopies g;
  const uint64_t n3 = 3ull * (uint64_t)n_verts;
  g.copies = grad_copies_for(n3);
  g.n3 = (uint32_t)n3;
  if (g.copies <= 1) return g;
  const uint64_t n = n3 * g.copies;
  hip_check(hipMallocAsync((void**)&g.base, n * sizeof(float), st), "hipMallocAsync(gradient copies)");
  hip_check(launch_fill_zero(g.base, n, st), "k_fill_zero");
  return g;
}

// Words of a stream's spill buffer: the traversal stacks' spill area, then
// the forward's tile counter (kTileCtrWords words, one used).
constexpr size_t kTileCtrWords = 64;
static size_t spill_words(int max_grid) { return (size_t)kSpillDepth * (size_t)max_grid * kBlock; }

// Launch shape for `work_items` items over scene `s`, with the spill buffer of
// stream `st` (none when with_spill is false; the caller supplies one).
LaunchCfg make_cfg(trc_ctx* ctx, const trc_scene* s, hipStream_t st, long long work_items, bool with_spill = true) {
  LaunchCfg L;
  int max_grid = ctx->num_cus * kBlocksPerCU;
  long long need = (work_items + kBlock - 1) / kBlock;
  L.grid = (int)(need < max_grid ? (need < 1 ? 1 : need) : max_grid);
  int stack_need = s->dev.stack_need;
  if (stack_need > kStackLds + kSpillDepth)
    throw Fail(TRC_ERR_OVERFLOW, "BVH too deep for the traversal stack (" + std::to_string(stack_need) + ")");
  L.cap = kStackLds + kSpillDepth;
  L.cus = ctx->num_cus;
  L.spill = nullptr;
  L.tiles = nullptr;
  if (!with_spill) return L;
  // one spill buffer per (context, stream): launches on different streams
  // never share one (word 0 doubles as the chunk counter of the adjoint)
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->spill.find(st);
  if (it == ctx->spill.end()) {
    uint32_t* p = nullptr;
    hip_check(hipMalloc(&p, (spill_words(max_grid) + kTileCtrWords) * 4), "hipMalloc(spill)");
    it = ctx->spill.emplace(st, p).first;
  }
  L.spill = it->second;
  L.tiles = it->second + spill_words(max_grid);
  return L;
}

template <class T>
T* upload(trc_scene* s, const std::vector<T>& v, hipStream_t st) {
  if (v.empty()) return nullptr;
  DevBuf b;
  b.bytes = v.size() * sizeof(T);
  hip_check(hipMalloc(&b.ptr, b.bytes), "hipMalloc(scene)");
  s->bufs.push_back(b);
  s->device_bytes += (int64_t)b.bytes;
  hip_check(hipMemcpyAsync(b.ptr, v.data(), b.bytes, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  return (T*)b.ptr;
}

template <class T>
void reupload(const T* dst, const std::vector<T>& v, hipStream_t st) {
  if (v.empty()) return;
  hip_check(hipMemcpyAsync((void*)dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st),
            "hipMemcpyAsync");
}

// An uninitialised device buffer owned by the scene.
void* alloc_dev(trc_scene* s, size_t bytes) {
  DevBuf b;
  b.bytes = bytes == 0 ? 16 : bytes;
  hip_check(hipMalloc(&b.ptr, b.bytes), "hipMalloc(scene)");
  s->bufs.push_back(b);
  s->device_bytes += (int64_t)b.bytes;
  return b.ptr;
}

// `v` padded with zero records to `cap` entries, uploaded.
template <class T>
T* upload_cap(trc_scene* s, std::vector<T> v, size_t cap, hipStream_t st) {
  v.resize(std::max<size_t>(std::max(cap, v.size()), 1));
  return upload(s, v, st);
}

void free_scene_bufs(trc_scene* s) {
  for (auto& b : s->bufs) hipFree(b.ptr);
  s->bufs.clear();
  s->upd = SceneUpdateDev{};
  s->prim_e = nullptr;
  s->prim_len = s->prim_cdf = nullptr;
  s->prim_cnt = nullptr;
  if (s->guide_buf.ptr) hipFree(s->guide_buf.ptr);
  s->guide_buf = DevBuf{};
  s->guide_key.valid = false;
  if (s->bq_buf.ptr) hipFree(s->bq_buf.ptr);
  s->bq_buf = DevBuf{};
  s->vn_adj = nullptr;
  s->device_bytes = 0;
  s->dev_texs = nullptr;
}

void upload_scene(trc_scene* s, hipStream_t st) {
  free_scene_bufs(s);
  SceneHost& h = s->host;
  SceneDev d{};
  d.pos = upload(s, h.pos, st);
  d.idx = upload(s, h.idx, st);
  d.vnrm = upload(s, h.vnrm, st);
  d.uv = upload(s, h.uv, st);
  d.uv_idx = upload(s, h.uv_idx, st);
  d.shade = upload(s, h.shade, st);
  d.tris = upload(s, h.tris, st);
  d.nodes = upload(s, h.nodes, st);
  d.mats = upload(s, h.materials, st);
  std::vector<TexDesc> texs;
  for (size_t i = 0; i < h.textures.size(); ++i) {
    TexDesc t;
    t.texels = upload(s, h.textures[i], st);
    t.w = h.tex_w[i];
    t.h = h.tex_h[i];
    t.grad_off = texs.empty() ? 0 : texs.back().grad_off + 3 * (int64_t)texs.back().w * texs.back().h;
    texs.push_back(t);
  }
  d.texs = upload(s, texs, st);
  d.light_tri = upload(s, h.light_tri, st);
  d.light_cdf = upload(s, h.light_cdf, st);
  d.light_f = upload(s, h.light_f, st);
  d.emit_rad = upload(s, h.emit_rad, st);
  // secondary-edge arrays sized for every edge: a device update may select
  // any number of them
  const size_t E = h.edges.size();
  d.sec.e = upload_cap(s, h.sec_edges, E, st);
  d.sec.len = upload_cap(s, h.sec_len, E, st);
  d.sec.cdf = upload_cap(s, h.sec_cdf, E, st);
  d.sec.n = (int32_t)h.sec_edges.size();
  d.sec.integral = h.sec_integral;
  d.n_light = (int32_t)h.light_tri.size();
  d.light_integral = h.light_integral;
  d.n_tris = h.num_triangles();
  d.n_mats = (int32_t)h.materials.size();
  d.stack_need = h.bvh_stats.stack_need;
  d.n_verts = h.num_vertices();
  // smooth-shaded differentiable faces per vertex, in triangle order
  std::vector<uint32_t> vf_off(h.num_vertices() + 1, 0u), vf_list;
  const uint32_t T = (uint32_t)h.num_triangles();
  auto smooth_diff = [&](uint32_t t) { return (h.shade[t].flags & (kTriSmooth | kTriDiff)) == (kTriSmooth | kTriDiff); };
  for (uint32_t t = 0; t < T; ++t)
    if (smooth_diff(t))
      for (int j = 0; j < 3; ++j) vf_off[h.idx[3 * t + j] + 1]++;
  for (size_t v = 0; v + 1 < vf_off.size(); ++v) vf_off[v + 1] += vf_off[v];
  if (vf_off.back() > 0) {
    vf_list.resize(vf_off.back());
    std::vector<uint32_t> fill(vf_off.begin(), vf_off.end() - 1);
    for (uint32_t t = 0; t < T; ++t)
      if (smooth_diff(t))
        for (int j = 0; j < 3; ++j) vf_list[fill[h.idx[3 * t + j]]++] = t;
    d.vf_off = upload(s, vf_off, st);
    d.vf_list = upload(s, vf_list, st);
    std::vector<float> zeros(3 * (size_t)h.num_vertices(), 0.0f);
    s->vn_adj = (float*)upload(s, zeros, st);
  }
  // device-side update: every edge, smooth faces per vertex, the refit plan
  SceneUpdateDev& U = s->upd;
  {
    std::vector<EdgeDev> all(E);
    for (size_t i = 0; i < E; ++i) {
      const EdgeHost& e = h.edges[i];
      all[i] = EdgeDev{e.v0, e.v1, e.v2, e.f0, e.f1, {0, 0, 0}};
    }
    U.all_edges = upload(s, all, st);
    U.n_edges = (uint32_t)E;
  }
  {
    std::vector<uint32_t> off(h.num_vertices() + 1, 0u), list;
    auto smooth = [&](uint32_t t) { return (h.shade[t].flags & kTriSmooth) != 0u; };
    for (uint32_t t = 0; t < T; ++t)
      if (smooth(t))
        for (int j = 0; j < 3; ++j) off[h.idx[3 * t + j] + 1]++;
    for (size_t v = 0; v + 1 < off.size(); ++v) off[v + 1] += off[v];
    if (off.back() > 0) {
      list.resize(off.back());
      std::vector<uint32_t> fill(off.begin(), off.end() - 1);
      for (uint32_t t = 0; t < T; ++t)
        if (smooth(t))
          for (int j = 0; j < 3; ++j) list[fill[h.idx[3 * t + j]]++] = t;
      U.vs_off = upload(s, off, st);
      U.vs_list = upload(s, list, st);
    }
  }
  U.refit_nodes = upload(s, h.refit_plan.nodes, st);
  U.level_off = h.refit_plan.level_off;
  U.n_nodes = (uint32_t)h.nodes.size();
  U.nbox = (float*)alloc_dev(s, 6 * sizeof(float) * (size_t)U.n_nodes);
  U.scan = (uint32_t*)alloc_dev(s, 4 * (size_t)edge_scan_words((uint32_t)E));
  U.part = (float*)alloc_dev(s, 4 * std::max<size_t>(1, (E + kCdfChunk - 1) / kCdfChunk));
  {
    std::vector<UpdateScalars> sc(1);
    sc[0].sec = EdgeCount{(int32_t)h.sec_edges.size(), h.sec_integral};
    sc[0].light_integral = h.light_integral;
    sc[0].pad = 0.0f;
    sc[0].sah = h.bvh_stats.sah_cost;
    U.scal = upload(s, sc, st);
  }
  s->prim_e = (EdgeDev*)alloc_dev(s, sizeof(EdgeDev) * std::max<size_t>(E, 1));
  s->prim_len = (float*)alloc_dev(s, sizeof(float) * std::max<size_t>(E, 1));
  s->prim_cdf = (float*)alloc_dev(s, sizeof(float) * std::max<size_t>(E, 1));
  s->prim_cnt = (EdgeCount*)alloc_dev(s, sizeof(EdgeCount));
  s->sah_built = h.bvh_stats.sah_cost;
  s->host_stale = false;
  s->dev = d;
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

// Brings the host copies of the positions and of everything derived from
// them (normals, emitter table, secondary edges, BVH boxes) up to date after
// device updates.  Only explicit queries and a rebuild need them.
void sync_host(trc_scene* s) {
  if (!s->host_stale) return;
  hip_check(hipMemcpy(s->host.pos.data(), s->dev.pos, s->host.pos.size() * sizeof(float), hipMemcpyDeviceToHost),
            "hipMemcpy(positions)");
  s->host.update_positions(nullptr);
  s->host.refit_bvh();
  s->host.bvh_stats.sah_cost = s->upd_host ? s->upd_host->sah : s->host.bvh_stats.sah_cost;
  s->host_stale = false;
}

CameraDev to_cam(const trc_camera* c) {
  if (!c) throw Fail(TRC_ERR_INVALID, "camera is NULL");
  if (c->width <= 0 || c->height <= 0) throw Fail(TRC_ERR_INVALID, "camera size must be positive");
  if (!(c->fov > 0.0f && c->fov < 180.0f)) throw Fail(TRC_ERR_INVALID, "camera fov out of range");
  return make_camera(c->pos, c->target, c->up, c->width, c->height, c->fov);
}

void check_render_common(const trc_scene* s, const trc_options* o) {
  if (!s->configured) throw Fail(TRC_ERR_STATE, "scene is not configured");
  if (!o) throw Fail(TRC_ERR_INVALID, "options are NULL");
  if (o->max_bounces < 0 || o->max_bounces > kMaxBounces)
    throw Fail(TRC_ERR_INVALID, "max_bounces must be in [0, " + std::to_string(kMaxBounces) + "]");
  if (s->dev.n_light <= 0) throw Fail(TRC_ERR_INVALID, "scene has no area emitter");
}

void check_render(const trc_scene* s, const trc_options* o) {
  if (!s->configured) throw Fail(TRC_ERR_STATE, "scene is not configured");
  if (!o) throw Fail(TRC_ERR_INVALID, "options are NULL");
  if (o->spp <= 0) throw Fail(TRC_ERR_INVALID, "spp must be positive");
  if (o->max_bounces < 0 || o->max_bounces > kMaxBounces)
    throw Fail(TRC_ERR_INVALID, "max_bounces must be in [0, " + std::to_string(kMaxBounces) + "]");
  if (s->dev.n_light <= 0) throw Fail(TRC_ERR_INVALID, "scene has no area emitter");
}

}  // namespace

extern "C" {

trc_status trc_context_create_host(trc_ctx** out) {
  if (!out) return TRC_ERR_INVALID;
  trc_ctx* c = new trc_ctx();
  c->device = -1;
  *out = c;
  return TRC_OK;
}

trc_status trc_context_create(int32_t device, trc_ctx** out) {
  if (!out) return TRC_ERR_INVALID;
  *out = nullptr;
  trc_ctx* c = new trc_ctx();
  trc_status st = guarded(c, [&] {
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) throw Fail(TRC_ERR_INVALID, "no such device");
    hip_check(hipSetDevice(device), "hipSetDevice");
    c->device = device;
    hip_check(hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device),
              "hipDeviceGetAttribute");
    c->scratch = device_scratch(device);
  });
  if (st != TRC_OK) {
    // keep the context so the caller can read the message
    *out = c;
    return st;
  }
  *out = c;
  return TRC_OK;
}

trc_status trc_context_destroy(trc_ctx* ctx) {
  if (!ctx) return TRC_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (auto& kv : ctx->spill) hipFree(kv.second);
    ctx->spill.clear();
  }
  for (auto& b : ctx->batch) {
    if (b.done) {
      hipEventSynchronize(b.done);
      hipEventDestroy(b.done);
    }
    if (b.host) hipHostFree(b.host);
    if (b.dev) hipFree(b.dev);
  }
  delete ctx;
  return TRC_OK;
}

const char* trc_last_error(const trc_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int32_t trc_device_cus(const trc_ctx* ctx) { return ctx ? ctx->num_cus : 0; }

trc_status trc_scene_create(trc_ctx* ctx, trc_scene** out) {
  if (!ctx || !out) return TRC_ERR_INVALID;
  trc_scene* s = new trc_scene();
  s->ctx = ctx;
  *out = s;
  return TRC_OK;
}

trc_status trc_scene_destroy(trc_scene* s) {
  if (!s) return TRC_ERR_INVALID;
  if (s->ctx->device >= 0) {
    hipSetDevice(s->ctx->device);
    free_scene_bufs(s);
    if (s->upd_host) hipHostFree(s->upd_host);
    s->upd_host = nullptr;
  }
  delete s;
  return TRC_OK;
}

int32_t trc_scene_add_material(trc_scene* s, int32_t type, const float albedo[3], const float spec[3],
                               float roughness, int32_t texture) {
  if (!s) return -1;
  if (type != TRC_MAT_DIFFUSE && type != TRC_MAT_MICROFACET && type != TRC_MAT_ROUGHCONDUCTOR) {
    s->ctx->last_error = "unknown material type";
    return -1;
  }
  if (type == TRC_MAT_ROUGHCONDUCTOR && texture >= 0) {
    s->ctx->last_error = "roughconductor takes no texture";
    return -1;
  }
  Material m{};
  m.type = type;
  m.tex = texture;
  m.roughness = roughness;
  for (int k = 0; k < 3; ++k) {
    m.albedo[k] = albedo ? albedo[k] : 0.0f;
    m.spec[k] = spec ? spec[k] : 0.0f;
    m.eta[k] = m.albedo[k];  // rough conductor: albedo slot = eta, spec slot = k
    m.k[k] = m.spec[k];
  }
  s->host.materials.push_back(m);
  s->configured = false;
  return (int32_t)s->host.materials.size() - 1;
}

int32_t trc_scene_add_texture(trc_scene* s, const float* texels, int32_t w, int32_t h) {
  if (!s || !texels || w <= 0 || h <= 0) return -1;
  s->host.textures.emplace_back(texels, texels + (size_t)w * h * 3);
  s->host.tex_w.push_back(w);
  s->host.tex_h.push_back(h);
  s->configured = false;
  return (int32_t)s->host.textures.size() - 1;
}

int32_t trc_scene_add_mesh(trc_scene* s, const float* pos, int32_t nv, const uint32_t* idx, int32_t nt,
                           const float* uv, int32_t nuv, const uint32_t* uv_idx, int32_t material,
                           const float* emit, int32_t flags) {
  if (!s || !pos || !idx || nv <= 0 || nt <= 0) {
    if (s) s->ctx->last_error = "add_mesh: empty or NULL buffers";
    return -1;
  }
  MeshHost m;
  m.pos.assign(pos, pos + 3 * (size_t)nv);
  m.idx.assign(idx, idx + 3 * (size_t)nt);
  if (uv && uv_idx && nuv > 0) {
    m.uv.assign(uv, uv + 2 * (size_t)nuv);
    m.uv_idx.assign(uv_idx, uv_idx + 3 * (size_t)nt);
  }
  m.material = material;
  if (emit) {
    m.is_emitter = true;
    for (int k = 0; k < 3; ++k) m.radiance[k] = emit[k];
  }
  m.smooth = (flags & TRC_MESH_SMOOTH) != 0;
  m.diff = (flags & TRC_MESH_DIFF) != 0;
  s->host.meshes.push_back(std::move(m));
  s->configured = false;
  return (int32_t)s->host.meshes.size() - 1;
}

namespace {
void prepare_host(trc_scene* s) {
  for (auto& m : s->host.materials)
    if (m.tex >= (int32_t)s->host.textures.size()) throw Fail(TRC_ERR_INVALID, "material texture out of range");
  s->host.flatten();
  s->host.build_bvh();
}
}  // namespace

trc_status trc_scene_prepare(trc_scene* s) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] { prepare_host(s); });
}

trc_status trc_scene_configure(trc_scene* s, trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
    prepare_host(s);
    upload_scene(s, (hipStream_t)stream);
    s->configured = true;
    ++s->version;
  });
}

// A refit tree whose SAH cost grows past this factor of the cost it had
// when it was built is rebuilt.
constexpr double kRebuildFactor = 1.25;

trc_status trc_scene_set_vertices(trc_scene* s, const float* pos, int32_t nv, int32_t refit, trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    if (!s->configured) throw Fail(TRC_ERR_STATE, "scene is not configured");
    if (!pos || nv != s->host.num_vertices()) throw Fail(TRC_ERR_INVALID, "vertex count mismatch");
    hipStream_t st = (hipStream_t)stream;
    SceneDev& d = s->dev;
    if (!refit) {
      // rebuild: the host builds the tree over the new positions
      hip_check(hipMemcpyAsync(s->host.pos.data(), pos, s->host.pos.size() * sizeof(float), hipMemcpyDefault, st),
                "hipMemcpyAsync");
      hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
      s->host_stale = false;
      s->host.update_positions(nullptr);
      s->host.build_bvh();
      upload_scene(s, st);
      ++s->rebuilds;
      ++s->version;
      return;
    }
    // refit on the device: positions in place, then every derived array and
    // the BVH boxes; only the scalars the launches take by value (edge count,
    // integrals) and the tree's SAH cost come back to the host
    hip_check(hipMemcpyAsync((void*)d.pos, pos, s->host.pos.size() * sizeof(float), hipMemcpyDefault, st),
              "hipMemcpyAsync(positions)");
    hip_check(launch_scene_update(d, s->upd, st), "scene update kernels");
    if (!s->upd_host) hip_check(hipHostMalloc((void**)&s->upd_host, sizeof(UpdateScalars)), "hipHostMalloc");
    hip_check(hipMemcpyAsync(s->upd_host, s->upd.scal, sizeof(UpdateScalars), hipMemcpyDeviceToHost, st),
              "hipMemcpyAsync(scalars)");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    d.sec.n = s->upd_host->sec.n;
    d.sec.integral = s->upd_host->sec.integral;
    d.light_integral = s->upd_host->light_integral;
    s->host.bvh_stats.sah_cost = s->upd_host->sah;
    s->host_stale = true;
    ++s->refits;
    ++s->version;
    if (s->upd_host->sah > kRebuildFactor * s->sah_built) {
      sync_host(s);
      s->host.build_bvh();
      upload_scene(s, st);
      ++s->rebuilds;
    }
  });
}

trc_status trc_scene_set_material(trc_scene* s, int32_t mi, const float albedo[3], const float spec[3],
                                  float roughness) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    if (mi < 0 || mi >= (int32_t)s->host.materials.size()) throw Fail(TRC_ERR_INVALID, "material out of range");
    Material& m = s->host.materials[mi];
    for (int k = 0; k < 3; ++k) {
      if (albedo) m.albedo[k] = albedo[k];
      if (spec) m.spec[k] = spec[k];
      m.eta[k] = m.albedo[k];
      m.k[k] = m.spec[k];
    }
    m.roughness = roughness;
    ++s->version;
    if (s->configured) {
      hip_check(hipMemcpy((void*)(s->dev.mats + mi), &m, sizeof(Material), hipMemcpyHostToDevice), "hipMemcpy");
    }
  });
}

trc_status trc_scene_get_info(const trc_scene* s, trc_scene_info* out) {
  if (!s || !out) return TRC_ERR_INVALID;
  const SceneHost& h = s->host;
  out->num_vertices = h.num_vertices();
  out->num_triangles = h.num_triangles();
  out->num_materials = (int32_t)h.materials.size();
  out->num_emitter_triangles = (int32_t)h.light_tri.size();
  out->num_edges = (int32_t)h.edges.size();
  out->bvh_nodes = h.bvh_stats.nodes;
  out->bvh_leaves = h.bvh_stats.leaves;
  out->bvh_max_depth = h.bvh_stats.max_depth;
  out->bvh_stack_need = h.bvh_stats.stack_need;
  out->bvh_sah_cost = h.bvh_stats.sah_cost;
  out->bvh_refits = s->refits;
  out->bvh_rebuilds = s->rebuilds;
  out->device_bytes = s->device_bytes;
  int64_t tf = 0;
  for (size_t i = 0; i < h.textures.size(); ++i) tf += (int64_t)h.textures[i].size();
  out->texture_floats = tf;
  return TRC_OK;
}

trc_status trc_scene_get_vertices(const trc_scene* s, float* out) {
  if (!s || !out) return TRC_ERR_INVALID;
  trc_scene* m = const_cast<trc_scene*>(s);
  return guarded(m->ctx, [&] {
    sync_host(m);
    std::memcpy(out, s->host.pos.data(), s->host.pos.size() * 4);
  });
}

trc_status trc_scene_get_indices(const trc_scene* s, uint32_t* out) {
  if (!s || !out) return TRC_ERR_INVALID;
  std::memcpy(out, s->host.idx.data(), s->host.idx.size() * 4);
  return TRC_OK;
}

trc_status trc_scene_get_bvh(const trc_scene* s, uint32_t* nodes, uint32_t* order) {
  if (!s) return TRC_ERR_INVALID;
  trc_scene* m = const_cast<trc_scene*>(s);
  return guarded(m->ctx, [&] {
    sync_host(m);
    if (nodes) std::memcpy(nodes, s->host.nodes.data(), s->host.nodes.size() * sizeof(Bvh4Node));
    if (order) std::memcpy(order, s->host.tri_order.data(), s->host.tri_order.size() * 4);
  });
}

trc_status trc_scene_get_edges(const trc_scene* s, int32_t* out) {
  if (!s || !out) return TRC_ERR_INVALID;
  for (size_t i = 0; i < s->host.edges.size(); ++i) {
    const EdgeHost& e = s->host.edges[i];
    out[5 * i] = (int32_t)e.v0; out[5 * i + 1] = (int32_t)e.v1; out[5 * i + 2] = (int32_t)e.v2;
    out[5 * i + 3] = e.f0; out[5 * i + 4] = e.f1;
  }
  return TRC_OK;
}

trc_status trc_intersect(trc_scene* s, const float* rays, int32_t n, int32_t* tri, float* uvt, int32_t any_hit,
                         trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    if (!s->configured) throw Fail(TRC_ERR_STATE, "scene is not configured");
    if (n < 0 || (n > 0 && (!rays || !tri || !uvt))) throw Fail(TRC_ERR_INVALID, "bad ray buffers");
    if (n == 0) return;
    LaunchCfg L = make_cfg(s->ctx, s, (hipStream_t)stream, n);
    hip_check(launch_intersect(s->dev, L, rays, n, tri, uvt, any_hit, (hipStream_t)stream), "k_intersect");
  });
}

trc_status trc_render_paths(trc_scene* s, const trc_camera* cam, const trc_options* opt, float* paths,
                            trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    check_render(s, opt);
    CameraDev C = to_cam(cam);
    if (!paths) throw Fail(TRC_ERR_INVALID, "paths buffer is NULL");
    long long np = (long long)opt->spp * C.width * C.height;
    LaunchCfg L = make_cfg(s->ctx, s, (hipStream_t)stream, np);
    hip_check(launch_render_paths(s->dev, C, opt->seed, opt->max_bounces, opt->spp, L, paths, nullptr,
                                  (hipStream_t)stream),
              "k_render_paths");
  });
}

static void render_fwd_impl(trc_scene* s, const trc_camera* cam, const trc_options* opt, float* film, void* ckpt,
                     int64_t ckpt_bytes, trc_stream stream) {
  check_render(s, opt);
  CameraDev C = to_cam(cam);
  if (!film) throw Fail(TRC_ERR_INVALID, "film is NULL");
  hipStream_t st = (hipStream_t)stream;
  long long np = (long long)opt->spp * C.width * C.height;
  if (ckpt && ckpt_bytes < (int64_t)checkpoint_bytes((uint64_t)np, opt->max_bounces))
    throw Fail(TRC_ERR_INVALID, "checkpoint buffer too small (trc_checkpoint_bytes)");
  float* paths = nullptr;
  hip_check(hipMallocAsync((void**)&paths, (size_t)np * 12, st), "hipMallocAsync(paths)");
  LaunchCfg L = make_cfg(s->ctx, s, st, np);
  hipError_t e = launch_render_paths(s->dev, C, opt->seed, opt->max_bounces, opt->spp, L, paths, ckpt, st);
  if (e == hipSuccess) e = launch_film_reduce(paths, opt->spp, C.width * C.height, film, st);
  hipFreeAsync(paths, st);
  hip_check(e, "k_render_paths/k_film_reduce");
}

trc_status trc_render_fwd(trc_scene* s, const trc_camera* cam, const trc_options* opt, float* film,
                          trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] { render_fwd_impl(s, cam, opt, film, nullptr, 0, stream); });
}

trc_status trc_render_fwd_noalloc(trc_scene* s, const trc_camera* cam, const trc_options* opt, float* film,
                                  trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    check_render(s, opt);
    CameraDev C = to_cam(cam);
    if (!film) throw Fail(TRC_ERR_INVALID, "film is NULL");
    trc_ctx* ctx = s->ctx;
    if (!ctx->scratch) throw Fail(TRC_ERR_STATE, "context has no device scratch");
    hipStream_t st = (hipStream_t)stream;
    const uint64_t wh = (uint64_t)C.width * (uint64_t)C.height;
    const uint64_t np = (uint64_t)opt->spp * wh;
    // launch shape as trc_render_fwd's, but the traversal stack gets only the
    // spill depth this scene's BVH needs, and the spill slice and the chunk
    // of per-path radiance live in the fixed scratch
    LaunchCfg L = make_cfg(ctx, s, st, (long long)np, /*with_spill=*/false);
    const int cap = std::max(s->dev.stack_need, kStackLds);
    L.cap = cap;
    const uint64_t spill_bytes = ((uint64_t)(cap - kStackLds) * (uint64_t)L.grid * kBlock * 4 + 255) / 256 * 256;
    const uint64_t total = trc::noalloc_scratch_bytes();
    const uint64_t per_sample = 12 * wh;
    if (spill_bytes + per_sample > total)
      throw Fail(TRC_ERR_INVALID, "trc_render_fwd_noalloc: a " + std::to_string(C.width) + "x" +
                                      std::to_string(C.height) + " image needs " +
                                      std::to_string(spill_bytes + per_sample) + " bytes of scratch, more than its " +
                                      std::to_string(total));
    unsigned char* base = (unsigned char*)ctx->scratch;
    L.spill = (uint32_t*)base;
    float* chunk = (float*)(base + spill_bytes);
    const int chunk_spp = (int)std::min<uint64_t>((total - spill_bytes) / per_sample, (uint64_t)opt->spp);
    hipError_t e = hipSuccess;
    for (int s0 = 0; s0 < opt->spp && e == hipSuccess; s0 += chunk_spp) {
      const int ns = std::min(chunk_spp, opt->spp - s0);
      const uint64_t first = (uint64_t)s0 * wh;
      e = launch_render_paths(s->dev, C, opt->seed, opt->max_bounces, opt->spp, L, chunk, nullptr, st, first,
                              (uint64_t)ns * wh, first);
      if (e == hipSuccess)
        e = launch_film_accum(chunk, ns, (int)wh, film, s0 == 0, s0 + ns == opt->spp, opt->spp, st);
    }
    hip_check(e, "k_render_paths/k_film_accum");
  });
}

int64_t trc_checkpoint_bytes(const trc_camera* cam, const trc_options* opt) {
  if (!cam || !opt || cam->width <= 0 || cam->height <= 0 || opt->spp <= 0 || opt->max_bounces < 0 ||
      opt->max_bounces > kMaxBounces)
    return -1;
  const uint64_t np = (uint64_t)opt->spp * (uint64_t)cam->width * (uint64_t)cam->height;
  return (int64_t)checkpoint_bytes(np, opt->max_bounces);
}

trc_status trc_render_fwd_ckpt(trc_scene* s, const trc_camera* cam, const trc_options* opt, float* film,
                               void* ckpt, int64_t ckpt_bytes, trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    if (!ckpt) throw Fail(TRC_ERR_INVALID, "checkpoint is NULL");
    render_fwd_impl(s, cam, opt, film, ckpt, ckpt_bytes, stream);
  });
}

// ---- batched launches ----
// Validates the jobs and fills their device records; counts the floats of
// the privatized gradient copies the adjoint jobs need (bases unset).
static std::vector<JobDev> prepare_batch(trc_ctx* ctx, const trc_job* jobs, int32_t n, int32_t adjoint,
                                         const trc_scene** deepest, uint64_t* total, uint64_t* grad_floats) {
  if (ctx->device < 0) throw Fail(TRC_ERR_STATE, "host-only context cannot render");
  if (n < 0 || (n > 0 && !jobs)) throw Fail(TRC_ERR_INVALID, "bad job list");
  std::vector<JobDev> jd((size_t)n);
  *deepest = nullptr;
  *total = 0;
  *grad_floats = 0;
  for (int i = 0; i < n; ++i) {
    const trc_job& J = jobs[i];
    const std::string tag = "job " + std::to_string(i) + ": ";
    if (!J.scene || J.scene->ctx != ctx) throw Fail(TRC_ERR_INVALID, tag + "bad scene");
    check_render(J.scene, &J.opt);
    CameraDev C = to_cam(&J.cam);
    const uint64_t np = (uint64_t)J.opt.spp * (uint64_t)C.width * (uint64_t)C.height;
    if (J.num_paths == 0 || J.first_path + J.num_paths > np) throw Fail(TRC_ERR_INVALID, tag + "path range outside spp*W*H");
    if (!J.ckpt) throw Fail(TRC_ERR_INVALID, tag + "checkpoint is NULL");
    if (!adjoint && !J.paths) throw Fail(TRC_ERR_INVALID, tag + "paths is NULL");
    if (adjoint && (!J.dLdI || !J.dpos)) throw Fail(TRC_ERR_INVALID, tag + "dLdI / dpos is NULL");
    if (adjoint && J.scene->dev.vf_off && !J.vn_scratch)
      throw Fail(TRC_ERR_INVALID, tag + "smooth differentiable mesh needs vn_scratch");
    JobDev& d = jd[(size_t)i];
    std::memset(&d, 0, sizeof(d));
    d.S = J.scene->dev;
    d.C = C;
    d.seed = J.opt.seed;
    d.max_bounces = J.opt.max_bounces;
    d.spp = J.opt.spp;
    d.first = J.first_path;
    d.count = J.num_paths;
    d.paths = J.paths;
    d.ckpt = J.ckpt;
    d.dLdI = J.dLdI;
    d.dpos = J.dpos;
    d.dmat = J.dmat;
    d.dtex = J.dtex;
    d.vn = J.scene->dev.vf_off ? J.vn_scratch : nullptr;
    d.gc = GradCopies{};
    if (adjoint) {
      const uint64_t n3 = 3ull * (uint64_t)d.S.n_verts;
      const uint32_t k = grad_copies_for(n3);
      // jobs that accumulate into the same dpos share one set of copies
      bool shared = false;
      for (int q = 0; q < i && !shared; ++q)
        shared = jd[(size_t)q].dpos == d.dpos && jd[(size_t)q].gc.copies > 1;
      if (k > 1) {
        d.gc.copies = k;
        d.gc.n3 = (uint32_t)n3;
        if (!shared) *grad_floats += n3 * k;
      }
    }
    *total += J.num_paths;
    if (!*deepest || J.scene->dev.stack_need > (*deepest)->dev.stack_need) *deepest = J.scene;
  }
  return jd;
}

static void assign_grad_copies(std::vector<JobDev>& jd, float* base) {
  uint64_t off = 0;
  for (size_t i = 0; i < jd.size(); ++i) {
    JobDev& d = jd[i];
    if (d.gc.copies <= 1) continue;
    const JobDev* owner = nullptr;
    for (size_t q = 0; q < i && !owner; ++q)
      if (jd[q].dpos == d.dpos && jd[q].gc.base) owner = &jd[q];
    if (owner) {
      d.gc.base = owner->gc.base;
      continue;
    }
    d.gc.base = base + off;
    off += (uint64_t)d.gc.n3 * d.gc.copies;
  }
}

// Default tile: about four tiles per resident wave (waves take tiles
// round-robin, so finer tiles even out paths of unequal length), at least
// one path per lane, at most 4096 paths.
static uint64_t default_tile(const trc_ctx* ctx, uint64_t total) {
  const uint64_t waves = (uint64_t)ctx->num_cus * kBlocksPerCU * (kBlock / 64);
  uint64_t tile = total / (waves * 4) / 64 * 64;
  return tile < 64 ? 64 : (tile > 4096 ? 4096 : tile);
}

// In a batch that mixes scenes, jobs of at least this many paths (several
// times the resident lanes) fill the GPU on their own: they run through the
// single-request kernels over their path range, one scene at a time in the
// caches, and only the rest is tiled.  (Serving at saturation: 667 -> 702
// Mpaths/s.)  So does a batch of a single large job, which then launches
// exactly what the plain path launches.  A batch of several jobs of one
// scene -- the views of a multi-view plan -- is tiled whole, which keeps
// the tail of each view hidden behind the others.
constexpr uint64_t kSoloPaths = 4u << 20;

static uint64_t mark_solo(std::vector<JobDev>& jd) {
  bool mixed = jd.size() == 1;
  for (const JobDev& d : jd) mixed = mixed || d.S.nodes != jd[0].S.nodes;
  uint64_t tiled = 0;
  for (JobDev& d : jd) {
    d.solo = mixed && d.count >= kSoloPaths ? 1 : 0;
    if (!d.solo) tiled += d.count;
  }
  return tiled;
}

static uint64_t assign_tiles(std::vector<JobDev>& jd, uint64_t tile) {
  uint64_t ntiles = 0;
  for (JobDev& d : jd) {
    d.tile0 = ntiles;
    if (!d.solo) ntiles += (d.count + tile - 1) / tile;
  }
  return ntiles;
}

trc_status trc_render_batch(trc_ctx* ctx, const trc_job* jobs, int32_t n, int32_t adjoint, trc_stream stream) {
  if (!ctx) return TRC_ERR_INVALID;
  return guarded(ctx, [&] {
    hipStream_t st = (hipStream_t)stream;
    const trc_scene* deepest;
    uint64_t total, gfloats;
    std::vector<JobDev> jd = prepare_batch(ctx, jobs, n, adjoint, &deepest, &total, &gfloats);
    if (n == 0) return;
    float* gbase = nullptr;
    if (gfloats) {
      hip_check(hipMallocAsync((void**)&gbase, gfloats * sizeof(float), st), "hipMallocAsync(gradient copies)");
      hip_check(launch_fill_zero(gbase, gfloats, st), "k_fill_zero");
      assign_grad_copies(jd, gbase);
    }
    LaunchCfg L = make_cfg(ctx, deepest, st, (long long)total);
    const uint64_t tiled = mark_solo(jd);
    const uint64_t tile = default_tile(ctx, tiled);
    const uint64_t ntiles = assign_tiles(jd, tile);
    // job table: pinned staging -> device (stream-ordered), from a ring slot
    // whose previous batch has completed
    const size_t bytes = jd.size() * sizeof(JobDev);
    std::lock_guard<std::mutex> lk(ctx->batch_mu);
    trc_ctx::BatchSlot& b = ctx->batch[ctx->batch_next++ % trc_ctx::kBatchSlots];
    if (b.done) hip_check(hipEventSynchronize(b.done), "hipEventSynchronize");
    else hip_check(hipEventCreateWithFlags(&b.done, hipEventDisableTiming), "hipEventCreate");
    if (b.cap < bytes) {
      if (b.host) hipHostFree(b.host);
      if (b.dev) hipFree(b.dev);
      b.host = nullptr;
      b.dev = nullptr;
      b.cap = 0;
      hip_check(hipHostMalloc(&b.host, bytes, hipHostMallocDefault), "hipHostMalloc(jobs)");
      hip_check(hipMalloc(&b.dev, bytes), "hipMalloc(jobs)");
      b.cap = bytes;
    }
    std::memcpy(b.host, jd.data(), bytes);
    hip_check(hipMemcpyAsync(b.dev, b.host, bytes, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    hipError_t e = launch_render_batch((const JobDev*)b.dev, jd.data(), n, ntiles, (uint32_t)tile, adjoint, L, st);
    if (gbase) hipFreeAsync(gbase, st);
    hip_check(e, adjoint ? "k_render_batch_adj" : "k_render_batch_fwd");
    hip_check(hipEventRecord(b.done, st), "hipEventRecord");
  });
}

// ---- prepared batches ----
struct trc_plan {
  trc_ctx* ctx = nullptr;
  int adjoint = 0;
  std::vector<JobDev> jd;
  JobDev* dev_jobs = nullptr;
  float* gbase = nullptr;      // persistent gradient copies (re-zeroed by each launch's reduction)
  uint32_t* spill = nullptr;   // the plan's own traversal spill buffer
  LaunchCfg L{};
  uint64_t total = 0, tiled = 0, tile = 0, ntiles = 0;
  // the scenes the job table snapshots, with their versions at creation
  std::vector<std::pair<const trc_scene*, uint64_t>> scenes;
};

namespace {
void plan_free(trc_plan* p) {
  if (p->dev_jobs) hipFree(p->dev_jobs);
  if (p->gbase) hipFree(p->gbase);
  if (p->spill) hipFree(p->spill);
  p->dev_jobs = nullptr;
  p->gbase = nullptr;
  p->spill = nullptr;
}

void plan_upload_tiles(trc_plan* p, uint64_t tile) {
  p->tile = tile;
  p->ntiles = assign_tiles(p->jd, tile);
  hip_check(hipMemcpy(p->dev_jobs, p->jd.data(), p->jd.size() * sizeof(JobDev), hipMemcpyHostToDevice),
            "hipMemcpy(jobs)");
}
}  // namespace

trc_status trc_plan_create(trc_ctx* ctx, const trc_job* jobs, int32_t n, int32_t adjoint, trc_plan** out) {
  if (!ctx || !out) return TRC_ERR_INVALID;
  *out = nullptr;
  return guarded(ctx, [&] {
    if (n <= 0) throw Fail(TRC_ERR_INVALID, "a plan needs at least one job");
    std::unique_ptr<trc_plan> p(new trc_plan);
    p->ctx = ctx;
    p->adjoint = adjoint ? 1 : 0;
    const trc_scene* deepest;
    uint64_t gfloats;
    p->jd = prepare_batch(ctx, jobs, n, p->adjoint, &deepest, &p->total, &gfloats);
    for (int i = 0; i < n; ++i) {
      bool seen = false;
      for (auto& sv : p->scenes) seen = seen || sv.first == jobs[i].scene;
      if (!seen) p->scenes.emplace_back(jobs[i].scene, jobs[i].scene->version);
    }
    try {
      hip_check(hipSetDevice(ctx->device), "hipSetDevice");
      hip_check(hipMalloc((void**)&p->dev_jobs, p->jd.size() * sizeof(JobDev)), "hipMalloc(plan jobs)");
      if (gfloats) {
        hip_check(hipMalloc((void**)&p->gbase, gfloats * sizeof(float)), "hipMalloc(plan gradient copies)");
        hip_check(launch_fill_zero(p->gbase, gfloats, nullptr), "k_fill_zero");
        assign_grad_copies(p->jd, p->gbase);
        for (JobDev& d : p->jd) d.gc.keep = 1;
      }
      // launch shape of the deepest scene, with a spill buffer owned by the plan
      // so that launches never allocate
      p->L = make_cfg(ctx, deepest, nullptr, (long long)p->total, false);
      const size_t words = spill_words(ctx->num_cus * kBlocksPerCU);
      hip_check(hipMalloc((void**)&p->spill, (words + kTileCtrWords) * 4), "hipMalloc(plan spill)");
      p->L.spill = p->spill;
      p->L.tiles = p->spill + words;
      p->tiled = mark_solo(p->jd);
      plan_upload_tiles(p.get(), default_tile(ctx, p->tiled));
      hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    } catch (...) {
      plan_free(p.get());
      throw;
    }
    *out = p.release();
  });
}

trc_status trc_plan_set_tile(trc_plan* plan, uint32_t tile_paths) {
  if (!plan) return TRC_ERR_INVALID;
  return guarded(plan->ctx, [&] {
    if (tile_paths % 64 != 0) throw Fail(TRC_ERR_INVALID, "tile_paths must be a multiple of 64");
    plan_upload_tiles(plan, tile_paths ? tile_paths : default_tile(plan->ctx, plan->tiled));
  });
}

uint32_t trc_plan_tile(const trc_plan* plan) { return plan ? (uint32_t)plan->tile : 0u; }

static void plan_check_scenes(const trc_plan* p) {
  for (auto& sv : p->scenes)
    if (sv.first->version != sv.second)
      throw Fail(TRC_ERR_STATE, "a scene of the plan changed since trc_plan_create: create the plan again");
}

trc_status trc_plan_launch(trc_plan* plan, trc_stream stream) {
  if (!plan) return TRC_ERR_INVALID;
  return guarded(plan->ctx, [&] {
    plan_check_scenes(plan);
    hip_check(launch_render_batch(plan->dev_jobs, plan->jd.data(), (int)plan->jd.size(), plan->ntiles,
                                  (uint32_t)plan->tile, plan->adjoint, plan->L, (hipStream_t)stream),
              plan->adjoint ? "k_render_batch_adj" : "k_render_batch_fwd");
  });
}

trc_status trc_plan_tune(trc_plan* plan, const uint32_t* candidates, int32_t n, int32_t reps, uint32_t* best,
                         trc_stream stream) {
  if (!plan) return TRC_ERR_INVALID;
  return guarded(plan->ctx, [&] {
    if (n <= 0 || !candidates || reps <= 0) throw Fail(TRC_ERR_INVALID, "no candidates to tune over");
    plan_check_scenes(plan);
    for (int i = 0; i < n; ++i)
      if (candidates[i] == 0 || candidates[i] % 64 != 0)
        throw Fail(TRC_ERR_INVALID, "candidate tiles must be positive multiples of 64");
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    float best_ms = 0.0f;
    uint32_t best_tile = 0;
    std::string err;
    for (int i = 0; i < n && err.empty(); ++i) {
      plan_upload_tiles(plan, candidates[i]);
      std::vector<float> ms;
      for (int r = 0; r <= reps; ++r) {  // the first launch warms up
        hipError_t e = hipEventRecord(e0, st);
        if (e == hipSuccess)
          e = launch_render_batch(plan->dev_jobs, plan->jd.data(), (int)plan->jd.size(), plan->ntiles,
                                  (uint32_t)plan->tile, plan->adjoint, plan->L, st);
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e != hipSuccess) {
          err = std::string("plan tuning: ") + hipGetErrorString(e);
          break;
        }
        if (r > 0) ms.push_back(t);
      }
      if (!err.empty()) break;
      std::sort(ms.begin(), ms.end());
      const float med = ms[ms.size() / 2];
      if (best_tile == 0 || med < best_ms) {
        best_ms = med;
        best_tile = candidates[i];
      }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (!err.empty()) throw Fail(TRC_ERR_HIP, err);
    plan_upload_tiles(plan, best_tile);
    if (best) *best = best_tile;
  });
}

trc_status trc_plan_destroy(trc_plan* plan) {
  if (!plan) return TRC_OK;
  hipDeviceSynchronize();
  plan_free(plan);
  delete plan;
  return TRC_OK;
}

trc_status trc_render_adj_ckpt(trc_scene* s, const trc_camera* cam, const trc_options* opt, const void* ckpt,
                               int64_t ckpt_bytes, const float* dLdI, float* dpos, float* dmat, float* dtex,
                               trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    check_render(s, opt);
    CameraDev C = to_cam(cam);
    if (!dLdI || !dpos) throw Fail(TRC_ERR_INVALID, "dLdI / dpos is NULL");
    if (!ckpt) throw Fail(TRC_ERR_INVALID, "checkpoint is NULL");
    long long np = (long long)opt->spp * C.width * C.height;
    if (ckpt_bytes < (int64_t)checkpoint_bytes((uint64_t)np, opt->max_bounces))
      throw Fail(TRC_ERR_INVALID, "checkpoint buffer too small (trc_checkpoint_bytes)");
    LaunchCfg L = make_cfg(s->ctx, s, (hipStream_t)stream, np);
    GradCopies gc = make_grad_copies(s->dev.n_verts, (hipStream_t)stream);
    hipError_t e = launch_render_adj_ckpt(s->dev, C, opt->seed, opt->max_bounces, opt->spp, L, ckpt, dLdI, dpos, dmat,
                                          dtex, s->vn_adj, gc, (hipStream_t)stream);
    if (gc.base) hipFreeAsync(gc.base, (hipStream_t)stream);
    hip_check(e, "k_render_adj_ckpt");
  });
}

trc_status trc_adam_step(trc_ctx* ctx, float* x, const float* grad, float* m, float* v, uint64_t n, int32_t step,
                         float lr, float beta1, float beta2, float eps, trc_stream stream) {
  if (!ctx) return TRC_ERR_INVALID;
  return guarded(ctx, [&] {
    if (ctx->device < 0) throw Fail(TRC_ERR_STATE, "host-only context cannot run kernels");
    if (n && (!x || !grad || !m || !v)) throw Fail(TRC_ERR_INVALID, "NULL Adam buffer");
    if (step < 1) throw Fail(TRC_ERR_INVALID, "Adam step count starts at 1");
    // bias corrections 1 - beta^step, evaluated in double and rounded once
    const float bc1 = (float)(1.0 - std::pow((double)beta1, (double)step));
    const float bc2 = (float)(1.0 - std::pow((double)beta2, (double)step));
    hip_check(launch_adam(x, grad, m, v, n, lr, beta1, beta2, eps, bc1, bc2, (hipStream_t)stream), "k_adam");
  });
}

trc_status trc_adam_uniform_step(trc_ctx* ctx, float* x, const float* grad, float* m, float* v, uint64_t n,
                                 int32_t step, float lr, float beta1, float beta2, trc_stream stream) {
  if (!ctx) return TRC_ERR_INVALID;
  return guarded(ctx, [&] {
    if (ctx->device < 0) throw Fail(TRC_ERR_STATE, "host-only context cannot run kernels");
    if (n && (!x || !grad || !m || !v)) throw Fail(TRC_ERR_INVALID, "NULL Adam buffer");
    if (step < 1) throw Fail(TRC_ERR_INVALID, "Adam step count starts at 1");
    hipStream_t st = (hipStream_t)stream;
    const float bc1 = (float)(1.0 - std::pow((double)beta1, (double)step));
    const float bc2 = (float)(1.0 - std::pow((double)beta2, (double)step));
    unsigned int* vmax = nullptr;
    hip_check(hipMallocAsync((void**)&vmax, sizeof(unsigned int), st), "hipMallocAsync");
    hipError_t e = launch_adam_uniform(x, grad, m, v, n, lr, beta1, beta2, bc1, bc2, vmax, st);
    hipFreeAsync(vmax, st);
    hip_check(e, "k_adam_uniform");
  });
}

// ---- Laplacian solver ----
struct trc_lsolver {
  trc_ctx* ctx = nullptr;
  int n = 0;
  int* row = nullptr;
  int* col = nullptr;
  float* val = nullptr;
  float* work = nullptr;      // r, p, ap: 3 x [n][3]
  CgScalars* scal = nullptr;
  CsrDev dev() const { return CsrDev{n, row, col, val}; }
};

namespace {
void lsolver_free(trc_lsolver* s) {
  for (void* p : {(void*)s->row, (void*)s->col, (void*)s->val, (void*)s->work, (void*)s->scal})
    if (p) hipFree(p);
  s->row = s->col = nullptr;
  s->val = s->work = nullptr;
  s->scal = nullptr;
}
}  // namespace

trc_status trc_lsolver_create(trc_ctx* ctx, const uint32_t* faces, int32_t nf, int32_t nv, float lambda,
                              trc_lsolver** out) {
  if (!ctx || !out) return TRC_ERR_INVALID;
  *out = nullptr;
  return guarded(ctx, [&] {
    if (ctx->device < 0) throw Fail(TRC_ERR_STATE, "host-only context cannot run kernels");
    if (nv <= 0 || nf < 0 || (nf > 0 && !faces)) throw Fail(TRC_ERR_INVALID, "bad mesh");
    if (!(lambda >= 0.0f)) throw Fail(TRC_ERR_INVALID, "lambda must be >= 0");
    // undirected edges of the triangles, then CSR rows with the diagonal
    std::vector<std::pair<int, int>> e;
    e.reserve(6 * (size_t)nf);
    for (int32_t f = 0; f < nf; ++f)
      for (int k = 0; k < 3; ++k) {
        const uint32_t a = faces[3 * (size_t)f + k], b = faces[3 * (size_t)f + (k + 1) % 3];
        if (a >= (uint32_t)nv || b >= (uint32_t)nv) throw Fail(TRC_ERR_INVALID, "face index out of range");
        if (a == b) continue;
        e.emplace_back((int)a, (int)b);
        e.emplace_back((int)b, (int)a);
      }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    std::vector<int> row((size_t)nv + 1, 0), col;
    std::vector<float> val;
    col.reserve(e.size() + (size_t)nv);
    val.reserve(e.size() + (size_t)nv);
    size_t k = 0;
    for (int i = 0; i < nv; ++i) {
      size_t k0 = k;
      while (k < e.size() && e[k].first == i) ++k;
      const int deg = (int)(k - k0);
      bool diag_done = false;
      for (size_t q = k0; q < k; ++q) {
        if (!diag_done && e[q].second > i) {
          col.push_back(i);
          val.push_back(1.0f + lambda * (float)deg);
          diag_done = true;
        }
        col.push_back(e[q].second);
        val.push_back(-lambda);
      }
      if (!diag_done) {
        col.push_back(i);
        val.push_back(1.0f + lambda * (float)deg);
      }
      row[(size_t)i + 1] = (int)col.size();
    }
    std::unique_ptr<trc_lsolver> s(new trc_lsolver);
    s->ctx = ctx;
    s->n = nv;
    try {
      hip_check(hipSetDevice(ctx->device), "hipSetDevice");
      hip_check(hipMalloc((void**)&s->row, row.size() * sizeof(int)), "hipMalloc(csr)");
      hip_check(hipMalloc((void**)&s->col, col.size() * sizeof(int)), "hipMalloc(csr)");
      hip_check(hipMalloc((void**)&s->val, val.size() * sizeof(float)), "hipMalloc(csr)");
      hip_check(hipMalloc((void**)&s->work, 9 * (size_t)nv * sizeof(float)), "hipMalloc(cg)");
      hip_check(hipMalloc((void**)&s->scal, sizeof(CgScalars)), "hipMalloc(cg)");
      hip_check(hipMemcpy(s->row, row.data(), row.size() * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
      hip_check(hipMemcpy(s->col, col.data(), col.size() * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
      hip_check(hipMemcpy(s->val, val.data(), val.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    } catch (...) {
      lsolver_free(s.get());
      throw;
    }
    *out = s.release();
  });
}

trc_status trc_lsolver_apply(trc_lsolver* s, const float* x, float* y, trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    if (!x || !y || x == y) throw Fail(TRC_ERR_INVALID, "bad vectors");
    hip_check(launch_csr_spmv3(s->dev(), x, y, (hipStream_t)stream), "k_csr_spmv3");
  });
}

trc_status trc_lsolver_solve(trc_lsolver* s, const float* b, float* x, int32_t max_iter, float tol,
                             trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    if (!b || !x || b == x) throw Fail(TRC_ERR_INVALID, "bad vectors");
    if (max_iter < 0 || !(tol > 0.0f && tol < 1.0f)) throw Fail(TRC_ERR_INVALID, "bad solver limits");
    float* r = s->work;
    float* p = r + 3 * (size_t)s->n;
    float* ap = p + 3 * (size_t)s->n;
    hip_check(launch_cg(s->dev(), b, x, r, p, ap, s->scal, max_iter, tol, (hipStream_t)stream), "k_cg");
  });
}

trc_status trc_lsolver_status(trc_lsolver* s, int32_t* iters, float rel_residual[3]) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    CgScalars h;
    hip_check(hipMemcpy(&h, s->scal, sizeof(h), hipMemcpyDeviceToHost), "hipMemcpy");
    if (iters) *iters = (int32_t)h.iters;
    if (rel_residual)
      for (int c = 0; c < 3; ++c) rel_residual[c] = h.rr0[c] > 0.0f ? std::sqrt(h.rr[c] / h.rr0[c]) : 0.0f;  // |r| / |b|
  });
}

trc_status trc_lsolver_destroy(trc_lsolver* s) {
  if (!s) return TRC_OK;
  hipDeviceSynchronize();
  lsolver_free(s);
  delete s;
  return TRC_OK;
}

trc_status trc_film_from_paths(trc_ctx* ctx, const float* paths, int32_t spp, int32_t wh, float* film,
                               trc_stream stream) {
  if (!ctx) return TRC_ERR_INVALID;
  return guarded(ctx, [&] {
    if (!paths || !film || spp <= 0 || wh <= 0) throw Fail(TRC_ERR_INVALID, "bad film reduction arguments");
    hip_check(launch_film_reduce(paths, spp, wh, film, (hipStream_t)stream), "k_film_reduce");
  });
}

trc_status trc_render_stats(trc_scene* s, const trc_camera* cam, const trc_options* opt, uint64_t* counters,
                            trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    check_render(s, opt);
    CameraDev C = to_cam(cam);
    if (!counters) throw Fail(TRC_ERR_INVALID, "counters is NULL");
    hipStream_t st = (hipStream_t)stream;
    long long np = (long long)opt->spp * C.width * C.height;
    unsigned long long* d = nullptr;
    hip_check(hipMalloc((void**)&d, 4 * sizeof(unsigned long long)), "hipMalloc(counters)");
    hipError_t e = hipMemsetAsync(d, 0, 4 * sizeof(unsigned long long), st);
    LaunchCfg L = make_cfg(s->ctx, s, st, np);
    if (e == hipSuccess) e = launch_render_stats(s->dev, C, opt->seed, opt->max_bounces, opt->spp, L, d, st);
    if (e == hipSuccess) e = hipMemcpyAsync(counters, d, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(d);
    hip_check(e, "k_render_stats");
  });
}

// One boundary integrator: gradient mode (dLdI, dpos) or, with dimg, the
// derivative image along tpos.  Runs inside guarded().
static void boundary_call(trc_scene* s, const trc_camera* cam, const trc_options* opt, int32_t kind,
                          const float* dLdI, float* dpos, const float* tpos, float* dimg, trc_stream stream) {
  check_render_common(s, opt);
  CameraDev C = to_cam(cam);
  if (kind < 0 || kind > 3) throw Fail(TRC_ERR_INVALID, "unknown boundary kind");
  const int spp = kind == 0 ? opt->sppe : kind == 1 ? opt->sppe0 : kind == 2 ? opt->sppse0 : opt->sppse1;
  if (spp < 0) throw Fail(TRC_ERR_INVALID, "negative boundary spp");
  if (spp == 0) return;
  const long long N = (long long)spp * C.width * C.height;
  if (kind == 1 && N % 4 != 0) throw Fail(TRC_ERR_INVALID, "pixel-boundary samples (sppe0*W*H) must be a multiple of 4");
  hipStream_t st = (hipStream_t)stream;
  EdgeSet prim{};
  if (kind == 0) {
    // boundary and silhouette edges seen from the camera, selected on the
    // device; the kernels read their count and length integral from there
    if (s->upd.n_edges == 0) return;
    hip_check(launch_prim_edges(s->dev, s->upd, C.eye, s->prim_e, s->prim_len, s->prim_cdf, s->prim_cnt, st),
              "primary edge selection");
    prim.e = s->prim_e;
    prim.len = s->prim_len;
    prim.cdf = s->prim_cdf;
    prim.dyn = s->prim_cnt;
  }
  if (kind >= 2 && s->dev.sec.n == 0) return;
  GuideDev guide{};
  if (kind == 2 && opt->g_direct) {
    const int n = s->dev.sec.n;
    if (opt->g_direct_spp < 1 || opt->g_direct_spp > 4096)
      throw Fail(TRC_ERR_INVALID, "g_direct_spp must be in [1, 4096]");
    if (!(opt->g_eps >= 0.0f && opt->g_eps <= 1.0f)) throw Fail(TRC_ERR_INVALID, "g_eps must be in [0, 1]");
    if (opt->g_direct_max_size > 0 && n > opt->g_direct_max_size)
      throw Fail(TRC_ERR_INVALID, "guiding: " + std::to_string(n) + " secondary edges exceed g_direct_max_size " +
                                      std::to_string(opt->g_direct_max_size));
    const size_t nv = (size_t)n * (size_t)opt->g_direct_spp;
    const size_t nparts = ((size_t)n + kSumChunk - 1) / kSumChunk + 1;
    const size_t need = (nv + 2 * (size_t)n + 1 + nparts) * sizeof(float);
    const uint32_t gseed = opt->g_keep ? opt->g_seed : opt->seed;
    const trc_scene::GuideKey key{true, s->version, gseed, opt->g_direct_spp, n, opt->g_eps};
    const trc_scene::GuideKey& k = s->guide_key;
    const bool reuse = opt->g_keep && k.valid && k.version == key.version && k.seed == key.seed &&
                       k.spp == key.spp && k.n == key.n && k.eps == key.eps && s->guide_buf.bytes >= need;
    if (s->guide_buf.bytes < need) {
      hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
      if (s->guide_buf.ptr) hipFree(s->guide_buf.ptr);
      s->guide_buf.ptr = nullptr;
      s->guide_buf.bytes = 0;
      s->guide_key.valid = false;
      hip_check(hipMalloc(&s->guide_buf.ptr, need), "hipMalloc(guide)");
      s->guide_buf.bytes = need;
    }
    float* base = (float*)s->guide_buf.ptr;
    guide.vals = base;
    guide.poly = base + nv;
    guide.cdf = base + nv + n;
    guide.fsum = base + nv + 2 * (size_t)n;
    guide.part = base + nv + 2 * (size_t)n + 1;
    guide.n = n;
    guide.spp = opt->g_direct_spp;
    guide.eps = opt->g_eps;
    guide.enabled = 1;
    if (!reuse) {
      s->guide_key.valid = false;
      LaunchCfg Lg = make_cfg(s->ctx, s, st, (long long)nv);
      hip_check(launch_guide_build(s->dev, Lg, gseed, guide, st), "k_guide");
      if (opt->g_keep) s->guide_key = key;
    }
  }
  LaunchCfg L = make_cfg(s->ctx, s, st, N);
  const size_t qb = (size_t)boundary_queue_bytes(kind, (uint64_t)N);
  if (s->bq_buf.bytes < qb) {
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (s->bq_buf.ptr) hipFree(s->bq_buf.ptr);
    s->bq_buf.ptr = nullptr;
    s->bq_buf.bytes = 0;
    hip_check(hipMalloc(&s->bq_buf.ptr, qb), "hipMalloc(boundary queue)");
    s->bq_buf.bytes = qb;
  }
  hip_check(launch_boundary(kind, s->dev, C, opt->seed, opt->max_bounces, spp, L, prim, guide, dLdI, dpos, tpos,
                            dimg, s->bq_buf.ptr, st),
            "k_boundary");
}

trc_status trc_render_boundary(trc_scene* s, const trc_camera* cam, const trc_options* opt, int32_t kind,
                               const float* dLdI, float* dpos, trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    if (!dLdI || !dpos) throw Fail(TRC_ERR_INVALID, "dLdI / dpos is NULL");
    boundary_call(s, cam, opt, kind, dLdI, dpos, nullptr, nullptr, stream);
  });
}

trc_status trc_render_deriv(trc_scene* s, const trc_camera* cam, const trc_options* opt, int32_t kind,
                            const float* tpos, const float* tmat, float* dimg, trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    if (!tpos || !dimg) throw Fail(TRC_ERR_INVALID, "tpos / dimg is NULL");
    if (kind < -1 || kind > 3) throw Fail(TRC_ERR_INVALID, "unknown integrator kind");
    if (kind >= 0) {
      boundary_call(s, cam, opt, kind, nullptr, nullptr, tpos, dimg, stream);
      return;
    }
    check_render(s, opt);
    CameraDev C = to_cam(cam);
    hipStream_t st = (hipStream_t)stream;
    const long long np = (long long)opt->spp * C.width * C.height;
    LaunchCfg L = make_cfg(s->ctx, s, st, np);
    // vertex-normal tangents only where a differentiable mesh is smooth
    float* tvn = nullptr;
    if (s->vn_adj)
      hip_check(hipMallocAsync((void**)&tvn, 3 * (size_t)s->dev.n_verts * sizeof(float), st), "hipMallocAsync(tvn)");
    const hipError_t e = launch_render_deriv(s->dev, C, opt->seed, opt->max_bounces, opt->spp, L, tpos, tmat, tvn,
                                             dimg, st);
    if (tvn) hipFreeAsync(tvn, st);
    hip_check(e, "k_render_deriv");
  });
}

trc_status trc_render_adj(trc_scene* s, const trc_camera* cam, const trc_options* opt, const float* dLdI,
                          float* dpos, float* dmat, float* dtex, trc_stream stream) {
  if (!s) return TRC_ERR_INVALID;
  return guarded(s->ctx, [&] {
    check_render(s, opt);
    CameraDev C = to_cam(cam);
    if (!dLdI || !dpos) throw Fail(TRC_ERR_INVALID, "dLdI / dpos is NULL");
    long long np = (long long)opt->spp * C.width * C.height;
    LaunchCfg L = make_cfg(s->ctx, s, (hipStream_t)stream, np);
    GradCopies gc = make_grad_copies(s->dev.n_verts, (hipStream_t)stream);
    hipError_t e = launch_render_adj(s->dev, C, opt->seed, opt->max_bounces, opt->spp, L, dLdI, dpos, dmat, dtex,
                                     s->vn_adj, gc, (hipStream_t)stream);
    if (gc.base) hipFreeAsync(gc.base, (hipStream_t)stream);
    hip_check(e, "k_render_adj");
  });
}

}  // extern "C"
