// Host classes of the `TensorRay` module over the C ABI (see host_scene.h).
#include "host_scene.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <unordered_map>

namespace trpy {

// ---------------------------------------------------------------------------
// environment
// ---------------------------------------------------------------------------
Env& env() {
  static Env e;
  return e;
}

void env_create(int device) {
  if (env().ctx) return;
  trc_ctx* c = nullptr;
  trc_status st = trc_context_create(device, &c);
  if (st != TRC_OK) throw std::runtime_error("env_create: no usable HIP device");
  env().ctx = c;
  env().rnd_seed = 0;
}

// Scenes holding a device scene of the current context.
static std::set<Scene*>& live_scenes() {
  static std::set<Scene*> s;
  return s;
}

// Releases the device context and empties the parameter pool.  Scenes that
// outlive it lose their device state first (they must be configured again
// under a new context); destroying it afterwards would touch the freed one.
void env_release() {
  param_pool().clear();
  for (Scene* s : live_scenes())
    if (s->sc) {
      trc_scene_destroy(s->sc);
      s->sc = nullptr;
    }
  live_scenes().clear();
  if (env().ctx) trc_context_destroy(env().ctx);
  env().ctx = nullptr;
}

trc_ctx* ctx() {
  if (!env().ctx) throw std::runtime_error("TensorRay environment not created (call env_create first)");
  return env().ctx;
}

void check(trc_status st, const char* what) {
  if (st != TRC_OK) throw std::runtime_error(std::string(what) + ": " + trc_last_error(env().ctx));
}

static void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Copies between caller pointers (host or device).  Without a visible GPU
// every pointer is a host pointer, so scene assembly and parameter access
// also work on machines without one.
static bool have_device() {
  static int n = -1;
  if (n < 0) {
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  }
  return n > 0;
}
static void any_copy(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  if (!have_device()) {
    std::memcpy(dst, src, bytes);
    return;
  }
  hip_ok(hipMemcpy(dst, src, bytes, hipMemcpyDefault), "hipMemcpy");
}

// ---------------------------------------------------------------------------
// tensors
// ---------------------------------------------------------------------------
int64_t TensorShape::elements() const {
  int64_t n = 1;
  for (int d : dims) n *= d;
  return n;
}

DevArray::DevArray(int64_t count) : n(count) {
  if (n > 0) hip_ok(hipMalloc((void**)&p, (size_t)n * sizeof(float)), "hipMalloc");
  zero();
}
DevArray::~DevArray() {
  if (p) (void)hipFree(p);
}
void DevArray::zero() {
  if (p) hip_ok(hipMemset(p, 0, (size_t)n * sizeof(float)), "hipMemset");
}

void Tensorf::value_to(float* dst) const {
  if (!dst) throw std::runtime_error("value_to_torch: null pointer");
  if (dev) {
    hip_ok(hipMemcpy(dst, dev->p, (size_t)dev->n * sizeof(float), hipMemcpyDefault), "hipMemcpy");
  } else if (!host.empty()) {
    const std::vector<float> v = resolved();
    any_copy(dst, v.data(), v.size() * sizeof(float));
  }
}

Tensorf Tensorf::from_ptr(const float* src, const TensorShape& shape) {
  Tensorf t;
  t.shape = shape;
  t.dev = std::make_shared<DevArray>(shape.size());
  if (shape.size() > 0)
    hip_ok(hipMemcpy(t.dev->p, src, (size_t)shape.size() * sizeof(float), hipMemcpyDefault), "hipMemcpy");
  return t;
}

Tensorf Tensorf::scalar(float v, bool rg) {
  Tensorf t;
  t.shape = TensorShape({1}, kScalar1);
  t.host = {v};
  t.requires_grad = rg;
  return t;
}

Tensorf Tensorf::vec3(float x, float y, float z, bool rg) {
  Tensorf t;
  t.shape = TensorShape({1}, kVec3);
  t.host = {x, y, z};
  t.requires_grad = rg;
  return t;
}

static const Param* scalar_param(int i) {
  if (i < 0 || i >= (int)param_pool().size()) return nullptr;
  const Param& p = param_pool()[i];
  return p.kind == kParamScalar && p.value.size() == 1 ? &p : nullptr;
}

std::vector<float> Tensorf::resolved() const {
  if (const Param* p = scalar_param(lin_param)) {
    const float x = p->value[0];
    return {lin_dir[0] * x, lin_dir[1] * x, lin_dir[2] * x};
  }
  if (const Param* p = scalar_param(param)) return p->value;
  return host;
}

int register_scalar_param(float v) {
  Param p;
  p.kind = kParamScalar;
  p.value = {v};
  param_pool().push_back(p);
  return (int)param_pool().size() - 1;
}

// Tensor * Tensor for scene set-up (Tensor.h expression operators): two
// constants multiply elementwise (a single value broadcasts); a constant
// Vec3 times a scalar parameter stays linked to the parameter, so that
// set_var / configure move it and gradients and derivative images reach it.
Tensorf tensor_mul(const Tensorf& a, const Tensorf& b) {
  const bool pa = scalar_param(a.param) != nullptr, pb = scalar_param(b.param) != nullptr;
  if (a.lin_param >= 0 || b.lin_param >= 0 || (pa && pb))
    throw std::runtime_error("Tensorf product: only constant * constant and constant Vec3 * scalar parameter");
  if (pa || pb) {
    const Tensorf& k = pa ? b : a;
    const int prm = pa ? a.param : b.param;
    std::array<float, 3> dir;
    if (k.host.size() == 3) dir = {k.host[0], k.host[1], k.host[2]};
    else throw std::runtime_error("Tensorf product with a parameter: the constant must be a Vec3");
    Tensorf t = Tensorf::vec3(0.0f, 0.0f, 0.0f, false);
    t.lin_param = prm;
    t.lin_dir = dir;
    t.host = t.resolved();
    return t;
  }
  const std::vector<float> x = a.resolved(), y = b.resolved();
  if (x.empty() || y.empty()) throw std::runtime_error("Tensorf product of an empty tensor");
  if (x.size() != y.size() && x.size() != 1 && y.size() != 1)
    throw std::runtime_error("Tensorf product: shapes do not broadcast");
  Tensorf t = x.size() >= y.size() ? a : b;
  t.param = t.lin_param = -1;
  t.requires_grad = false;
  const size_t n = std::max(x.size(), y.size());
  t.host.resize(n);
  for (size_t i = 0; i < n; ++i) t.host[i] = x[x.size() == 1 ? 0 : i] * y[y.size() == 1 ? 0 : i];
  return t;
}

// ---------------------------------------------------------------------------
// small 4x4 algebra (row-major, double)
// ---------------------------------------------------------------------------
using M4 = std::array<double, 16>;
static M4 m4_identity() { return {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}; }
static M4 m4_mul(const M4& a, const M4& b) {
  M4 r{};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += a[4 * i + k] * b[4 * k + j];
      r[4 * i + j] = s;
    }
  return r;
}
static M4 m4_translate(double x, double y, double z) {
  M4 m = m4_identity();
  m[3] = x; m[7] = y; m[11] = z;
  return m;
}
static M4 m4_scale(double x, double y, double z) {
  M4 m = m4_identity();
  m[0] = x; m[5] = y; m[10] = z;
  return m;
}
// Rotation by `rad` radians about `axis` (Rodrigues).
static M4 m4_rotate(double rad, double ax, double ay, double az) {
  double l = std::sqrt(ax * ax + ay * ay + az * az);
  if (l == 0.0) return m4_identity();
  ax /= l; ay /= l; az /= l;
  double c = std::cos(rad), s = std::sin(rad), t = 1.0 - c;
  M4 m = m4_identity();
  m[0] = c + t * ax * ax;      m[1] = t * ax * ay - s * az; m[2] = t * ax * az + s * ay;
  m[4] = t * ax * ay + s * az; m[5] = c + t * ay * ay;      m[6] = t * ay * az - s * ax;
  m[8] = t * ax * az - s * ay; m[9] = t * ay * az + s * ax; m[10] = c + t * az * az;
  return m;
}
// d/drad of the rotation part of m4_rotate.
static std::array<double, 9> rot_deriv(double rad, double ax, double ay, double az) {
  double l = std::sqrt(ax * ax + ay * ay + az * az);
  std::array<double, 9> d{};
  if (l == 0.0) return d;
  ax /= l; ay /= l; az /= l;
  double c = std::cos(rad), s = std::sin(rad);
  d[0] = -s + s * ax * ax;      d[1] = s * ax * ay - c * az; d[2] = s * ax * az + c * ay;
  d[3] = s * ax * ay + c * az;  d[4] = -s + s * ay * ay;     d[5] = s * ay * az - c * ax;
  d[6] = s * ax * az - c * ay;  d[7] = s * ay * az + c * ax; d[8] = -s + s * az * az;
  return d;
}

// ---------------------------------------------------------------------------
// BSDF / Camera / Primitive
// ---------------------------------------------------------------------------
static int64_t texture_param_size(const BSDF& b, const std::string& slot);

// BSDF::DiffTexture (BSDF.cpp:350-356): registers the named texture (or
// constant) as a parameter, in call order like the reference's pool.
void BSDF::diff_texture(const std::string& name) {
  const std::string slot = name.empty() ? "albedo" : name;
  if (texture_param_size(*this, slot) < 0) throw std::runtime_error("Texture " + slot + " not supported by this BSDF");
  Param p;
  p.kind = kParamTexture;
  p.bsdf = shared_from_this();
  p.slot = slot;
  param_pool().push_back(p);
}

Camera::Camera(const Vector3f& p, const Vector3f& t, const Vector3f& u, int w, int h, float f) { init(p, t, u, w, h, f); }
void Camera::init(const Vector3f& p, const Vector3f& t, const Vector3f& u, int w, int h, float f) {
  pos = {p.x, p.y, p.z};
  target = {t.x, t.y, t.z};
  up = {u.x, u.y, u.z};
  width = w;
  height = h;
  fov = f;
}
void Camera::update(const Camera& c) {
  pos = c.pos;
  target = c.target;
  up = c.up;
  width = c.width;
  height = c.height;
  fov = c.fov;
}
void Camera::resize(int w, int h) {
  width = w;
  height = h;
}
trc_camera Camera::to_trc() const {
  trc_camera c;
  for (int i = 0; i < 3; ++i) {
    c.pos[i] = pos[i];
    c.target[i] = target[i];
    c.up[i] = up[i];
  }
  c.width = width;
  c.height = height;
  c.fov = fov;
  return c;
}

// ObjMesh::processEdges semantics through a host-only libtrc context.
static std::vector<int32_t> mesh_edges(const Primitive& p) {
  trc_ctx* hc = nullptr;
  check(trc_context_create_host(&hc), "trc_context_create_host");
  trc_scene* s = nullptr;
  std::vector<int32_t> out;
  try {
    check(trc_scene_create(hc, &s), "trc_scene_create");
    float alb[3] = {0.5f, 0.5f, 0.5f}, sp[3] = {0, 0, 0};
    int m = trc_scene_add_material(s, TRC_MAT_DIFFUSE, alb, sp, 0.0f, -1);
    if (trc_scene_add_mesh(s, p.raw.data(), p.vertex_count(), p.idx.data(), p.face_count(), nullptr, 0, nullptr, m,
                           nullptr, 0) < 0)
      throw std::runtime_error("invalid mesh");
    trc_status st = trc_scene_prepare(s);
    if (st != TRC_OK) throw std::runtime_error(std::string("trc_scene_prepare: ") + trc_last_error(hc));
    trc_scene_info info;
    trc_scene_get_info(s, &info);
    out.resize(5 * (size_t)info.num_edges);
    if (info.num_edges) trc_scene_get_edges(s, out.data());
  } catch (...) {
    if (s) trc_scene_destroy(s);
    trc_context_destroy(hc);
    throw;
  }
  trc_scene_destroy(s);
  trc_context_destroy(hc);
  return out;
}

int Primitive::edge_count() const { return (int)(mesh_edges(*this).size() / 5); }

void Primitive::face_indices(int* out) const {
  for (size_t i = 0; i < idx.size(); ++i) out[i] = (int)idx[i];
}

// GetEdgeData (Primitive.cpp:435-442): vertex-0 ids of all edges, then vertex-1 ids.
void Primitive::edge_data(int* out) const {
  std::vector<int32_t> e = mesh_edges(*this);
  const size_t n = e.size() / 5;
  for (size_t i = 0; i < n; ++i) {
    out[i] = e[5 * i];
    out[n + i] = e[5 * i + 1];
  }
}

// GetObjCenter (Primitive.cpp:444-450): centre of the object-space bounding box.
Tensorf Primitive::get_obj_center() const {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < raw.size(); ++i) {
    lo[i % 3] = std::min(lo[i % 3], raw[i]);
    hi[i % 3] = std::max(hi[i % 3], raw[i]);
  }
  return Tensorf::vec3(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]), false);
}

// ExportMesh (Primitive.cpp:452-470): object-space vertices and faces as OBJ.
void Primitive::export_mesh(const std::string& filename) const {
  std::ofstream f(filename);
  if (!f) throw std::runtime_error("export_mesh: cannot open " + filename);
  for (size_t v = 0; v < raw.size() / 3; ++v) f << "v " << raw[3 * v] << " " << raw[3 * v + 1] << " " << raw[3 * v + 2] << "\n";
  for (size_t t = 0; t < idx.size() / 3; ++t)
    f << "f " << idx[3 * t] + 1 << " " << idx[3 * t + 1] + 1 << " " << idx[3 * t + 2] + 1 << "\n";
}

static void register_shape_param(Primitive& s, ParamKind k) {
  Param p;
  p.kind = k;
  p.prim = s.shared_from_this();
  param_pool().push_back(p);
}
// DiffTranslation / DiffRotation / DiffAllVertexPos (Primitive.cpp:374-388).
void Primitive::diff_translation() {
  translation = Tensorf::vec3(0.0f, 0.0f, 0.0f, true);
  register_shape_param(*this, kParamTranslation);
}
void Primitive::diff_rotation(const Tensorf& axis) {
  if (axis.host.size() != 3) throw std::runtime_error("diff_rotation: the axis must be a Vec3 tensor");
  rotate_axis = axis;
  rotate_angle = Tensorf::scalar(0.0f, true);
  register_shape_param(*this, kParamRotation);
}
void Primitive::diff_all_vertex_pos() {
  if (diff_pos) return;
  diff_pos = true;
  register_shape_param(*this, kParamVertices);
}

// Primitive::Configure (Primitive.cpp:134-210): world = T(translation) or
// R(angle, axis), times to_world, times the multi-pose pre-transform
// C^-1 R C about obj_center.
M4 Primitive::world_matrix() const {
  M4 cur = to_world;
  if (rotate_matrix.host.size() == 16) {
    M4 rm;
    for (int i = 0; i < 16; ++i) rm[i] = rotate_matrix.host[i];
    M4 c = m4_identity(), ci = m4_identity();
    if (obj_center.host.size() == 3) {
      c = m4_translate(-obj_center.host[0], -obj_center.host[1], -obj_center.host[2]);
      ci = m4_translate(obj_center.host[0], obj_center.host[1], obj_center.host[2]);
    }
    cur = m4_mul(to_world, m4_mul(ci, m4_mul(rm, c)));
  }
  const std::vector<float> tr = translation.resolved();
  if (tr.size() == 3) return m4_mul(m4_translate(tr[0], tr[1], tr[2]), cur);
  if (rotate_axis.host.size() == 3 && rotate_angle.host.size() == 1)
    return m4_mul(m4_rotate(rotate_angle.host[0], rotate_axis.host[0], rotate_axis.host[1], rotate_axis.host[2]), cur);
  return cur;
}

void Primitive::world_positions(std::vector<float>& out, std::array<double, 9>& lin) const {
  const M4 m = world_matrix();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) lin[3 * i + j] = m[4 * i + j];
  const size_t V = raw.size() / 3;
  out.resize(3 * V);
  for (size_t v = 0; v < V; ++v) {
    const double x = raw[3 * v], y = raw[3 * v + 1], z = raw[3 * v + 2];
    for (int i = 0; i < 3; ++i)
      out[3 * v + i] = (float)(m[4 * i] * x + m[4 * i + 1] * y + m[4 * i + 2] * z + m[4 * i + 3]);
  }
  const std::vector<float> vt = vertex_translation.resolved();
  if (vertex_id >= 0 && (size_t)vertex_id < V && vt.size() == 3)
    for (int i = 0; i < 3; ++i) out[3 * vertex_id + i] += vt[i];
}

// ---------------------------------------------------------------------------
// render options
// ---------------------------------------------------------------------------
trc_options RenderOptions::to_trc() const {
  trc_options t;
  std::memset(&t, 0, sizeof(t));
  t.seed = (uint32_t)env().rnd_seed;
  t.max_bounces = max_bounces;
  t.spp = spp;
  t.sppe = sppe;
  t.sppse0 = sppse0;
  t.sppse1 = sppse1;
  t.sppe0 = sppe0;
  t.quiet = quiet ? 1 : 0;
  t.g_direct = g_direct ? 1 : 0;
  t.g_direct_depth = g_direct_depth;
  t.g_direct_max_size = g_direct_max_size;
  t.g_direct_spp = g_direct_spp;
  t.g_direct_thold = g_direct_thold;
  t.g_eps = g_eps;
  t.g_keep = g_keep ? 1 : 0;
  t.g_seed = g_seed;
  return t;
}

// ---------------------------------------------------------------------------
// scene
// ---------------------------------------------------------------------------
Scene::~Scene() {
  live_scenes().erase(this);
  if (sc) trc_scene_destroy(sc);
  if (active_scene() == this) active_scene() = nullptr;
}

int Scene::get_width(int i) const {
  if (i < 0 || i >= (int)cameras.size()) throw std::runtime_error("camera index out of range");
  return cameras[i]->width;
}
int Scene::get_height(int i) const {
  if (i < 0 || i >= (int)cameras.size()) throw std::runtime_error("camera index out of range");
  return cameras[i]->height;
}

static std::string topology_key(const Scene& s) {
  std::ostringstream k;
  for (const auto& p : s.shapes)
    k << p->raw.size() << ':' << p->idx.size() << ':' << p->uv.size() << ':' << p->bsdf << ':' << p->smooth << ':'
      << p->emitter << ':' << p->diff_pos << ':' << p->translation.host.size() << ':' << p->rotate_axis.host.size()
      << ':' << p->radiance[0] << ',' << p->radiance[1] << ','
      << p->radiance[2] << ';';
  k << '|';
  for (const auto& b : s.bsdfs) {
    k << b->type << ':' << b->tex_w << 'x' << b->tex_h << ':';
    // texel values are part of the uploaded scene
    uint64_t h = 1469598103934665603ull;
    for (float v : b->texels) {
      uint32_t u;
      std::memcpy(&u, &v, 4);
      h = (h ^ u) * 1099511628211ull;
    }
    k << h << ';';
  }
  return k.str();
}

void Scene::configure() {
  ++version;
  if (shapes.empty()) throw std::runtime_error("configure: scene has no shapes");
  if (bsdfs.empty()) throw std::runtime_error("configure: scene has no BSDFs");
  // world positions (Primitive::Configure for every shape)
  world.clear();
  lin.assign(shapes.size(), {});
  vert_off.assign(shapes.size(), 0);
  std::vector<float> wp;
  for (size_t i = 0; i < shapes.size(); ++i) {
    vert_off[i] = (int)(world.size() / 3);
    shapes[i]->world_positions(wp, lin[i]);
    world.insert(world.end(), wp.begin(), wp.end());
  }
  nv = (int)(world.size() / 3);
  const std::string key = topology_key(*this);
  if (sc && key == topo_key) {
    // same topology: new vertex positions (BVH refit) and material constants
    check(trc_scene_set_vertices(sc, world.data(), nv, 1, nullptr), "trc_scene_set_vertices");
    for (size_t i = 0; i < bsdfs.size(); ++i) {
      const BSDF& b = *bsdfs[i];
      check(trc_scene_set_material(sc, (int)i, b.albedo.data(), b.spec.data(), b.roughness), "trc_scene_set_material");
    }
    return;
  }
  if (sc) {
    trc_scene_destroy(sc);
    sc = nullptr;
  }
  check(trc_scene_create(ctx(), &sc), "trc_scene_create");
  live_scenes().insert(this);
  tex_index.assign(bsdfs.size(), -1);
  tex_off.assign(bsdfs.size(), -1);
  tex_floats = 0;
  for (size_t i = 0; i < bsdfs.size(); ++i) {
    const BSDF& b = *bsdfs[i];
    int ti = -1;
    if (!b.texels.empty()) {
      ti = trc_scene_add_texture(sc, b.texels.data(), b.tex_w, b.tex_h);
      if (ti < 0) check(TRC_ERR_INVALID, "trc_scene_add_texture");
      tex_index[i] = ti;
      tex_off[i] = tex_floats;
      tex_floats += (int64_t)b.texels.size();
    }
    if (trc_scene_add_material(sc, b.type, b.albedo.data(), b.spec.data(), b.roughness, ti) != (int)i)
      check(TRC_ERR_INVALID, "trc_scene_add_material");
  }
  for (size_t i = 0; i < shapes.size(); ++i) {
    const Primitive& p = *shapes[i];
    if (p.bsdf < 0 || p.bsdf >= (int)bsdfs.size()) throw std::runtime_error("shape refers to an unknown BSDF");
    const int flags = (p.smooth ? TRC_MESH_SMOOTH : 0) |
                      (p.diff_pos || !p.translation.host.empty() || !p.rotate_axis.host.empty() ||
                               p.vertex_translation.lin_param >= 0
                           ? TRC_MESH_DIFF
                           : 0);
    const float* uv = p.uv.empty() ? nullptr : p.uv.data();
    const uint32_t* uvi = p.uv_idx.empty() ? nullptr : p.uv_idx.data();
    if (trc_scene_add_mesh(sc, world.data() + 3 * (size_t)vert_off[i], p.vertex_count(), p.idx.data(), p.face_count(),
                           uv, (int)(p.uv.size() / 2), uvi, p.bsdf, p.emitter ? p.radiance.data() : nullptr,
                           flags) != (int)i)
      check(TRC_ERR_INVALID, "trc_scene_add_mesh");
  }
  check(trc_scene_configure(sc, nullptr), "trc_scene_configure");
  topo_key = key;
  dpos.reset();
  dmat.reset();
  dtex.reset();
  grads_pending = false;
}

void Scene::ensure_grad_buffers() {
  if (!sc) throw std::runtime_error("scene is not configured");
  if (!dpos || dpos->n != 3 * (int64_t)nv) dpos.reset(new DevArray(3 * (int64_t)nv));
  if (!dmat || dmat->n != 8 * (int64_t)bsdfs.size()) dmat.reset(new DevArray(8 * (int64_t)bsdfs.size()));
  if (tex_floats > 0 && (!dtex || dtex->n != tex_floats)) dtex.reset(new DevArray(tex_floats));
}

Scene*& active_scene() {
  static Scene* s = nullptr;
  return s;
}

// ---------------------------------------------------------------------------
// parameter pool
// ---------------------------------------------------------------------------
std::vector<Param>& param_pool() {
  static std::vector<Param> p;
  return p;
}

static int64_t texture_param_size(const BSDF& b, const std::string& slot) {
  const bool tex = !b.texels.empty();
  if (b.type == TRC_MAT_DIFFUSE) {
    if (slot == "albedo" || slot == "reflectance") return tex ? (int64_t)b.texels.size() : 3;
  } else if (b.type == TRC_MAT_MICROFACET) {
    if (slot == "diffuseReflectance" || slot == "albedo") return tex ? (int64_t)b.texels.size() : 3;
    if (slot == "specularReflectance") return 3;
    if (slot == "roughness") return 1;
  } else if (b.type == TRC_MAT_ROUGHCONDUCTOR) {
    if (slot == "eta" || slot == "k") return 3;
    if (slot == "roughness") return 1;
  }
  return -1;
}

int64_t Param::size() const {
  if (kind == kParamScalar) return 1;
  if (kind == kParamTexture) {
    auto b = bsdf.lock();
    return b ? texture_param_size(*b, slot) : 0;
  }
  auto p = prim.lock();
  if (!p) return 0;
  if (kind == kParamVertices) return 3 * (int64_t)p->vertex_count();
  if (kind == kParamTranslation) return 3;
  return 1;
}

int get_num_param() { return (int)param_pool().size(); }

int get_size_param(std::vector<int>& sizes) {
  sizes.clear();
  int total = 0;
  for (const Param& p : param_pool()) {
    sizes.push_back((int)p.size());
    total += (int)p.size();
  }
  return total;
}

static Param& param_at(int i) {
  if (i < 0 || i >= (int)param_pool().size()) throw std::runtime_error("parameter index out of range");
  return param_pool()[i];
}

// Values are exchanged in the reference's layouts: Vec3 tensors as [3][n]
// planes (vertex positions and texels), scalars as is.
static void copy_out(float* dst, const std::vector<float>& v) { any_copy(dst, v.data(), v.size() * sizeof(float)); }
static std::vector<float> copy_in(const float* src, size_t n) {
  std::vector<float> v(n);
  any_copy(v.data(), src, n * sizeof(float));
  return v;
}
static std::vector<float> aos_to_planes(const std::vector<float>& a) {
  const size_t n = a.size() / 3;
  std::vector<float> s(a.size());
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) s[c * n + i] = a[3 * i + c];
  return s;
}
static std::vector<float> planes_to_aos(const std::vector<float>& s) {
  const size_t n = s.size() / 3;
  std::vector<float> a(s.size());
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) a[3 * i + c] = s[c * n + i];
  return a;
}

static std::vector<float> texture_value(const BSDF& b, const std::string& slot) {
  if (slot == "roughness") return {b.roughness};
  if (slot == "specularReflectance" || slot == "k") return {b.spec[0], b.spec[1], b.spec[2]};
  if (!b.texels.empty() && slot != "eta") return aos_to_planes(b.texels);
  return {b.albedo[0], b.albedo[1], b.albedo[2]};
}

static void set_texture_value(BSDF& b, const std::string& slot, const std::vector<float>& v) {
  if (slot == "roughness") {
    b.roughness = v[0];
  } else if (slot == "specularReflectance" || slot == "k") {
    for (int c = 0; c < 3; ++c) b.spec[c] = v[c];
  } else if (!b.texels.empty() && slot != "eta") {
    b.texels = planes_to_aos(v);
  } else {
    for (int c = 0; c < 3; ++c) b.albedo[c] = v[c];
  }
}

void get_var(int index, float* dst, int offset) {
  Param& p = param_at(index);
  dst += offset;
  if (p.kind == kParamScalar) {
    copy_out(dst, p.value);
    return;
  }
  if (p.kind == kParamTexture) {
    auto b = p.bsdf.lock();
    if (!b) throw std::runtime_error("parameter's BSDF no longer exists");
    copy_out(dst, texture_value(*b, p.slot));
    return;
  }
  auto s = p.prim.lock();
  if (!s) throw std::runtime_error("parameter's shape no longer exists");
  if (p.kind == kParamVertices) copy_out(dst, aos_to_planes(s->raw));
  else if (p.kind == kParamTranslation) copy_out(dst, s->translation.host);
  else copy_out(dst, s->rotate_angle.host);
}

void set_var(int index, const float* src, int offset) {
  Param& p = param_at(index);
  src += offset;
  const size_t n = (size_t)p.size();
  std::vector<float> v = copy_in(src, n);
  if (p.kind == kParamScalar) {
    p.value = v;
    return;
  }
  if (p.kind == kParamTexture) {
    auto b = p.bsdf.lock();
    if (!b) throw std::runtime_error("parameter's BSDF no longer exists");
    set_texture_value(*b, p.slot, v);
    return;
  }
  auto s = p.prim.lock();
  if (!s) throw std::runtime_error("parameter's shape no longer exists");
  if (p.kind == kParamVertices) s->raw = planes_to_aos(v);
  else if (p.kind == kParamTranslation) s->translation.host = v;
  else s->rotate_angle.host = v;
}

// Moves the device gradients of the last render_d calls into the pool:
// world-space vertex gradients g_v become raw-position gradients L^T g_v,
// translation gradients sum_v g_v, rotation gradients
// sum_v g_v . (dR/dangle (C p_v)), material slots and texels as stored.
void Scene::fold_gradients() {
  if (!grads_pending) return;
  grads_pending = false;
  std::vector<float> gp(3 * (size_t)nv), gm(8 * bsdfs.size()), gt((size_t)tex_floats);
  hip_ok(hipMemcpy(gp.data(), dpos->p, gp.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
  hip_ok(hipMemcpy(gm.data(), dmat->p, gm.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
  if (tex_floats) hip_ok(hipMemcpy(gt.data(), dtex->p, gt.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
  dpos->zero();
  dmat->zero();
  if (dtex) dtex->zero();
  for (Param& p : param_pool()) {
    if (p.grad.size() != (size_t)p.size()) p.grad.assign((size_t)p.size(), 0.0);
    if (p.kind == kParamScalar) continue;  // below
    if (p.kind == kParamTexture) {
      auto b = p.bsdf.lock();
      if (!b) continue;
      int bi = -1;
      for (size_t i = 0; i < bsdfs.size(); ++i)
        if (bsdfs[i] == b) bi = (int)i;
      if (bi < 0) continue;
      const float* g = &gm[8 * (size_t)bi];
      if (p.slot == "roughness") {
        p.grad[0] += g[0];
      } else if (p.slot == "specularReflectance" || p.slot == "k") {
        for (int c = 0; c < 3; ++c) p.grad[c] += g[4 + c];
      } else if (!b->texels.empty() && p.slot != "eta") {
        const size_t n = b->texels.size() / 3;
        for (size_t i = 0; i < n; ++i)
          for (int c = 0; c < 3; ++c) p.grad[c * n + i] += gt[(size_t)tex_off[bi] + 3 * i + c];
      } else {
        for (int c = 0; c < 3; ++c) p.grad[c] += g[1 + c];
      }
      continue;
    }
    auto s = p.prim.lock();
    if (!s) continue;
    int si = -1;
    for (size_t i = 0; i < shapes.size(); ++i)
      if (shapes[i] == s) si = (int)i;
    if (si < 0) continue;
    const float* g = &gp[3 * (size_t)vert_off[si]];
    const size_t V = (size_t)s->vertex_count();
    const auto& L = lin[si];
    if (p.kind == kParamVertices) {
      for (size_t v = 0; v < V; ++v)
        for (int j = 0; j < 3; ++j)
          p.grad[j * V + v] += L[j] * g[3 * v] + L[3 + j] * g[3 * v + 1] + L[6 + j] * g[3 * v + 2];
    } else if (p.kind == kParamTranslation) {
      for (size_t v = 0; v < V; ++v)
        for (int j = 0; j < 3; ++j) p.grad[j] += g[3 * v + j];
    } else {
      // world = R(angle) * cur: d p / d angle = dR (cur p)
      Primitive tmp = *s;
      tmp.rotate_axis = Tensorf();
      tmp.rotate_angle = Tensorf();
      std::vector<float> q;
      std::array<double, 9> unused;
      tmp.world_positions(q, unused);
      const auto d = rot_deriv(s->rotate_angle.host[0], s->rotate_axis.host[0], s->rotate_axis.host[1],
                               s->rotate_axis.host[2]);
      double acc = 0.0;
      for (size_t v = 0; v < V; ++v)
        for (int i = 0; i < 3; ++i)
          acc += g[3 * v + i] * (d[3 * i] * q[3 * v] + d[3 * i + 1] * q[3 * v + 1] + d[3 * i + 2] * q[3 * v + 2]);
      p.grad[0] += acc;
    }
  }
  // scalar parameters: the shapes' translations linked to them
  for (size_t k = 0; k < param_pool().size(); ++k) {
    Param& p = param_pool()[k];
    if (p.kind != kParamScalar) continue;
    for (size_t si = 0; si < shapes.size(); ++si) {
      const Primitive& sh = *shapes[si];
      const float* g = &gp[3 * (size_t)vert_off[si]];
      const size_t V = (size_t)sh.vertex_count();
      if (sh.translation.lin_param == (int)k)
        for (size_t v = 0; v < V; ++v)
          for (int j = 0; j < 3; ++j) p.grad[0] += (double)g[3 * v + j] * sh.translation.lin_dir[j];
      if (sh.vertex_translation.lin_param == (int)k && sh.vertex_id >= 0 && (size_t)sh.vertex_id < V)
        for (int j = 0; j < 3; ++j)
          p.grad[0] += (double)g[3 * (size_t)sh.vertex_id + j] * sh.vertex_translation.lin_dir[j];
    }
  }
}

void get_grad(int index, float* dst, int offset, bool clear) {
  if (active_scene()) active_scene()->fold_gradients();
  Param& p = param_at(index);
  const size_t n = (size_t)p.size();
  if (p.grad.size() != n) p.grad.assign(n, 0.0);
  std::vector<float> g(n);
  for (size_t i = 0; i < n; ++i) g[i] = (float)p.grad[i];
  copy_out(dst + offset, g);
  if (clear) std::fill(p.grad.begin(), p.grad.end(), 0.0);
}

void set_rnd_seed(uint64_t seed) { env().rnd_seed = seed; }

// ---------------------------------------------------------------------------
// integrators
// ---------------------------------------------------------------------------
static const Camera& camera0(const Scene& s) {
  if (s.cameras.empty()) throw std::runtime_error("scene has no camera");
  if (!s.sc) throw std::runtime_error("scene is not configured (call configure)");
  return *s.cameras[0];
}

static const float* dldi_device(const Scene& s, const Tensorf& dLdI, std::shared_ptr<DevArray>& hold) {
  const Camera& c = camera0(s);
  const int64_t need = 3 * (int64_t)c.width * c.height;
  if (dLdI.dev) {
    if (dLdI.dev->n != need) throw std::runtime_error("dLdI has the wrong size for the camera");
    return dLdI.dev->p;
  }
  if ((int64_t)dLdI.host.size() != need) throw std::runtime_error("dLdI has the wrong size for the camera");
  hold = std::make_shared<DevArray>(need);
  hip_ok(hipMemcpy(hold->p, dLdI.host.data(), (size_t)need * 4, hipMemcpyHostToDevice), "hipMemcpy");
  return hold->p;
}

// Key of a checkpoint: the scene's process-unique id and version, and the
// exact bits of every camera float (two views that differ in the last bit
// get different keys), the seed, depth and spp.
static std::string ckpt_key_of(const Scene& s, const trc_camera& c, const trc_options& o) {
  std::ostringstream k;
  auto bits = [](float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
  };
  k << s.uid << ':' << s.version << ':' << c.width << 'x' << c.height << ':' << std::hex << bits(c.fov);
  for (int i = 0; i < 3; ++i) k << ':' << bits(c.pos[i]) << ',' << bits(c.target[i]) << ',' << bits(c.up[i]);
  k << std::dec << ':' << o.seed << ':' << o.max_bounces << ':' << o.spp;
  return k.str();
}

// PathTracer::RenderC: image [W*H] Vec3 as [3][H*W] planes on the device.
Tensorf PathTracer::render_c(Scene& scene, const RenderOptions& o) {
  const Camera& c = camera0(scene);
  set_param(o);
  Tensorf img;
  img.shape = TensorShape({c.width * c.height}, kVec3);
  img.dev = std::make_shared<DevArray>(3 * (int64_t)c.width * c.height);
  trc_camera tc = c.to_trc();
  trc_options to = o.to_trc();
  if (o.checkpoint) {
    const int64_t nb = trc_checkpoint_bytes(&tc, &to);
    if (nb < 0) check(TRC_ERR_INVALID, "renderC");
    const int64_t nf = (nb + 3) / 4;
    if (!ckpt || ckpt->n < nf) ckpt = std::make_shared<DevArray>(nf);
    check(trc_render_fwd_ckpt(scene.sc, &tc, &to, img.dev->p, ckpt->p, 4 * ckpt->n, nullptr), "renderC");
    ckpt_key = ckpt_key_of(scene, tc, to);
  } else {
    check(trc_render_fwd(scene.sc, &tc, &to, img.dev->p, nullptr), "renderC");
  }
  hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return img;
}

// PathTracer::RenderD: interior adjoint, accumulated into the scene's
// gradient buffers (read back through get_grad).
Tensorf PathTracer::render_d(Scene& scene, const RenderOptions& o, const Tensorf& dLdI) {
  const Camera& c = camera0(scene);
  set_param(o);
  if (o.export_deriv) return derivative_images(scene, o, -1);
  scene.ensure_grad_buffers();
  std::shared_ptr<DevArray> hold;
  const float* g = dldi_device(scene, dLdI, hold);
  trc_camera tc = c.to_trc();
  trc_options to = o.to_trc();
  if (o.checkpoint && ckpt && ckpt_key == ckpt_key_of(scene, tc, to)) {
    check(trc_render_adj_ckpt(scene.sc, &tc, &to, ckpt->p, 4 * ckpt->n, g, scene.dpos->p, scene.dmat->p,
                              scene.dtex ? scene.dtex->p : nullptr, nullptr),
          "renderD");
  } else {
    check(trc_render_adj(scene.sc, &tc, &to, g, scene.dpos->p, scene.dmat->p, scene.dtex ? scene.dtex->p : nullptr,
                         nullptr),
          "renderD");
  }
  hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
  scene.grads_pending = true;
  active_scene() = &scene;
  return Tensorf();
}

Tensorf BoundaryIntegrator::render_d(Scene& scene, const RenderOptions& o, const Tensorf& dLdI) {
  const Camera& c = camera0(scene);
  set_param(o);
  if (o.export_deriv) return derivative_images(scene, o, kind);
  scene.ensure_grad_buffers();
  std::shared_ptr<DevArray> hold;
  const float* g = dldi_device(scene, dLdI, hold);
  trc_camera tc = c.to_trc();
  trc_options to = o.to_trc();
  check(trc_render_boundary(scene.sc, &tc, &to, kind, g, scene.dpos->p, nullptr), "renderD");
  hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
  scene.grads_pending = true;
  active_scene() = &scene;
  return Tensorf();
}

// RenderD with export_deriv (Integrator.h:106-121, GradiantImageHandler,
// Integrator.cpp:305-379): one forward-mode derivative image per element of
// every pool parameter, in pool order, concatenated along the pixel axis
// of each colour plane ([3][nimg * W * H]).  Each element enters the
// device as a tangent of the world vertex positions (and of the material
// slots) and trc_render_deriv does the rest.
static constexpr int64_t kMaxDerivImages = 64;

Tensorf derivative_images(Scene& scene, const RenderOptions& o, int kind) {
  const Camera& c = camera0(scene);
  const int64_t WH = (int64_t)c.width * c.height;
  auto& pool = param_pool();
  int64_t nimg = 0;
  for (const Param& p : pool) nimg += p.size();
  if (nimg == 0) throw std::runtime_error("export_deriv: no differentiable parameter");
  if (nimg > kMaxDerivImages)
    throw std::runtime_error("export_deriv: " + std::to_string(nimg) + " derivative images asked for, at most " +
                             std::to_string(kMaxDerivImages) + " (one per parameter element)");
  const size_t nv3 = 3 * (size_t)scene.nv, nm8 = 8 * scene.bsdfs.size();
  auto out = std::make_shared<DevArray>(3 * WH * nimg);
  out->zero();
  DevArray img(3 * WH), tp((int64_t)nv3), tm((int64_t)nm8);
  const trc_camera tc = c.to_trc();
  const trc_options to = o.to_trc();
  int64_t k = 0;
  auto emit = [&](const std::vector<float>& tpos, const std::vector<float>* tmat) {
    img.zero();
    hip_ok(hipMemcpy(tp.p, tpos.data(), nv3 * 4, hipMemcpyHostToDevice), "hipMemcpy");
    if (tmat) hip_ok(hipMemcpy(tm.p, tmat->data(), nm8 * 4, hipMemcpyHostToDevice), "hipMemcpy");
    check(trc_render_deriv(scene.sc, &tc, &to, kind, tp.p, tmat ? tm.p : nullptr, img.p, nullptr), "renderD");
    for (int ch = 0; ch < 3; ++ch)
      hip_ok(hipMemcpy(out->p + ch * nimg * WH + k * WH, img.p + ch * WH, (size_t)WH * 4, hipMemcpyDeviceToDevice),
             "hipMemcpy");
    ++k;
  };
  auto shape_index = [&](const std::shared_ptr<Primitive>& sp) {
    for (size_t i = 0; i < scene.shapes.size(); ++i)
      if (scene.shapes[i] == sp) return (int)i;
    return -1;
  };
  for (size_t pi = 0; pi < pool.size(); ++pi) {
    const Param& p = pool[pi];
    const int64_t n = p.size();
    if (p.kind == kParamScalar) {
      std::vector<float> t(nv3, 0.0f);
      for (size_t si = 0; si < scene.shapes.size(); ++si) {
        const Primitive& sh = *scene.shapes[si];
        const size_t vo = (size_t)scene.vert_off[si], V = (size_t)sh.vertex_count();
        if (sh.translation.lin_param == (int)pi)
          for (size_t v = 0; v < V; ++v)
            for (int j = 0; j < 3; ++j) t[3 * (vo + v) + j] += sh.translation.lin_dir[j];
        if (sh.vertex_translation.lin_param == (int)pi && sh.vertex_id >= 0 && (size_t)sh.vertex_id < V)
          for (int j = 0; j < 3; ++j) t[3 * (vo + (size_t)sh.vertex_id) + j] += sh.vertex_translation.lin_dir[j];
      }
      emit(t, nullptr);
    } else if (p.kind == kParamTexture) {
      auto b = p.bsdf.lock();
      int bi = -1;
      for (size_t i = 0; i < scene.bsdfs.size(); ++i)
        if (b && scene.bsdfs[i] == b) bi = (int)i;
      if (bi < 0) throw std::runtime_error("export_deriv: parameter's BSDF is not in the scene");
      int first;
      if (p.slot == "roughness") first = 0;
      else if (p.slot == "specularReflectance" || p.slot == "k") first = 4;
      else if (!b->texels.empty()) throw std::runtime_error("export_deriv: texture texels carry no derivative image");
      else first = 1;  // albedo / reflectance / diffuseReflectance / eta constants
      const std::vector<float> zero(nv3, 0.0f);
      for (int64_t e = 0; e < n; ++e) {
        std::vector<float> m(nm8, 0.0f);
        m[8 * (size_t)bi + (size_t)first + (size_t)e] = 1.0f;
        emit(zero, &m);
      }
    } else {
      auto sp = p.prim.lock();
      const int si = sp ? shape_index(sp) : -1;
      if (si < 0) throw std::runtime_error("export_deriv: parameter's shape is not in the scene");
      const size_t vo = (size_t)scene.vert_off[si], V = (size_t)sp->vertex_count();
      const auto& L = scene.lin[si];
      if (p.kind == kParamVertices) {
        for (int j = 0; j < 3; ++j)
          for (size_t v = 0; v < V; ++v) {  // element j * V + v: raw coordinate j of vertex v
            std::vector<float> t(nv3, 0.0f);
            for (int i = 0; i < 3; ++i) t[3 * (vo + v) + i] = (float)L[3 * i + j];
            emit(t, nullptr);
          }
      } else if (p.kind == kParamTranslation) {
        for (int j = 0; j < 3; ++j) {
          std::vector<float> t(nv3, 0.0f);
          for (size_t v = 0; v < V; ++v) t[3 * (vo + v) + j] = 1.0f;
          emit(t, nullptr);
        }
      } else {
        // world = R(angle) * cur: d p / d angle = dR (cur p)
        Primitive tmp = *sp;
        tmp.rotate_axis = Tensorf();
        tmp.rotate_angle = Tensorf();
        std::vector<float> q;
        std::array<double, 9> unused;
        tmp.world_positions(q, unused);
        const auto d = rot_deriv(sp->rotate_angle.host[0], sp->rotate_axis.host[0], sp->rotate_axis.host[1],
                                 sp->rotate_axis.host[2]);
        std::vector<float> t(nv3, 0.0f);
        for (size_t v = 0; v < V; ++v)
          for (int i = 0; i < 3; ++i)
            t[3 * (vo + v) + i] =
                (float)(d[3 * i] * q[3 * v] + d[3 * i + 1] * q[3 * v + 1] + d[3 * i + 2] * q[3 * v + 2]);
        emit(t, nullptr);
      }
    }
  }
  hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
  Tensorf r;
  r.shape = TensorShape({(int)(nimg * WH)}, kVec3);
  r.dev = out;
  return r;
}

// ---------------------------------------------------------------------------
// OBJ
// ---------------------------------------------------------------------------
void load_obj(const std::string& filename, Primitive& out) {
  std::ifstream f(filename);
  if (!f) throw std::runtime_error("cannot open " + filename);
  std::vector<float> pos, uv;
  std::vector<uint32_t> idx, uv_idx;
  bool has_normals = false;
  std::string line;
  auto fix = [](long i, size_t n) -> long { return i < 0 ? (long)n + i : i - 1; };
  while (std::getline(f, line)) {
    std::istringstream ss(line);
    std::string tag;
    ss >> tag;
    if (tag == "v") {
      float x, y, z;
      ss >> x >> y >> z;
      pos.insert(pos.end(), {x, y, z});
    } else if (tag == "vt") {
      float u = 0, v = 0;
      ss >> u >> v;
      uv.insert(uv.end(), {u, v});
    } else if (tag == "vn") {
      has_normals = true;
    } else if (tag == "f") {
      std::vector<long> pv, tv;
      std::string tok;
      while (ss >> tok) {
        long a = 0, b = 0;
        size_t s1 = tok.find('/');
        a = std::stol(tok.substr(0, s1));
        if (s1 != std::string::npos) {
          size_t s2 = tok.find('/', s1 + 1);
          std::string t = tok.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
          if (!t.empty()) b = std::stol(t);
        }
        pv.push_back(fix(a, pos.size() / 3));
        tv.push_back(b ? fix(b, uv.size() / 2) : -1);
      }
      for (size_t k = 1; k + 1 < pv.size(); ++k) {
        const long tri[3] = {pv[0], pv[k], pv[k + 1]};
        const long tuv[3] = {tv[0], tv[k], tv[k + 1]};
        for (int j = 0; j < 3; ++j) {
          if (tri[j] < 0 || (size_t)tri[j] >= pos.size() / 3) throw std::runtime_error("bad face index in " + filename);
          idx.push_back((uint32_t)tri[j]);
          uv_idx.push_back(tuv[j] < 0 ? 0u : (uint32_t)tuv[j]);
        }
      }
    }
  }
  out.raw = std::move(pos);
  out.idx = std::move(idx);
  if (!uv.empty()) {
    out.uv = std::move(uv);
    out.uv_idx = std::move(uv_idx);
  } else {
    out.uv.clear();
    out.uv_idx.clear();
  }
  (void)has_normals;
}

// ---------------------------------------------------------------------------
// textures: PFM / binary PPM
// ---------------------------------------------------------------------------
void load_texture(const std::string& filename, std::vector<float>& texels, int& w, int& h) {
  std::ifstream f(filename, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + filename);
  std::string magic;
  f >> magic >> w >> h;
  if (magic == "PF") {
    double scale;
    f >> scale;
    f.get();
    std::vector<float> rows((size_t)w * h * 3);
    f.read((char*)rows.data(), rows.size() * 4);
    if (scale > 0) {  // big endian
      for (float& v : rows) {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        u = __builtin_bswap32(u);
        std::memcpy(&v, &u, 4);
      }
    }
    // PFM stores rows bottom to top
    texels.resize(rows.size());
    for (int y = 0; y < h; ++y)
      std::memcpy(&texels[(size_t)y * w * 3], &rows[(size_t)(h - 1 - y) * w * 3], (size_t)w * 12);
  } else if (magic == "P6") {
    int maxv;
    f >> maxv;
    f.get();
    std::vector<unsigned char> px((size_t)w * h * 3);
    f.read((char*)px.data(), px.size());
    texels.resize(px.size());
    for (size_t i = 0; i < px.size(); ++i) texels[i] = px[i] / (float)maxv;
  } else {
    throw std::runtime_error("unsupported texture format (PFM or binary PPM): " + filename);
  }
}

// ---------------------------------------------------------------------------
// scene XML
// ---------------------------------------------------------------------------
namespace {
struct XNode {
  std::string name;
  std::map<std::string, std::string> attr;
  std::vector<XNode> kids;
  const XNode* child(const std::string& n) const {
    for (const XNode& k : kids)
      if (k.name == n) return &k;
    return nullptr;
  }
  // child element whose `name` attribute is one of `names`
  const XNode* named(std::initializer_list<const char*> names) const {
    for (const XNode& k : kids) {
      auto it = k.attr.find("name");
      if (it == k.attr.end()) continue;
      for (const char* n : names)
        if (it->second == n) return &k;
    }
    return nullptr;
  }
  std::string get(const std::string& a, const std::string& def = "") const {
    auto it = attr.find(a);
    return it == attr.end() ? def : it->second;
  }
};

struct XParser {
  const std::string& s;
  size_t i = 0;
  explicit XParser(const std::string& t) : s(t) {}
  [[noreturn]] void fail(const std::string& m) { throw std::runtime_error("XML parsing failed: " + m); }
  void ws() {
    while (i < s.size() && isspace((unsigned char)s[i])) ++i;
  }
  void skip_misc() {
    for (;;) {
      ws();
      if (s.compare(i, 4, "<!--") == 0) {
        size_t e = s.find("-->", i);
        if (e == std::string::npos) fail("unterminated comment");
        i = e + 3;
      } else if (s.compare(i, 2, "<?") == 0) {
        size_t e = s.find("?>", i);
        if (e == std::string::npos) fail("unterminated declaration");
        i = e + 2;
      } else {
        return;
      }
    }
  }
  std::string ident() {
    size_t b = i;
    while (i < s.size() && (isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' || s[i] == ':' || s[i] == '.')) ++i;
    if (b == i) fail("expected a name at offset " + std::to_string(i));
    return s.substr(b, i - b);
  }
  XNode element() {
    skip_misc();
    if (i >= s.size() || s[i] != '<') fail("expected '<' at offset " + std::to_string(i));
    ++i;
    XNode n;
    n.name = ident();
    for (;;) {
      ws();
      if (s.compare(i, 2, "/>") == 0) {
        i += 2;
        return n;
      }
      if (i < s.size() && s[i] == '>') {
        ++i;
        break;
      }
      std::string a = ident();
      ws();
      if (i >= s.size() || s[i] != '=') fail("expected '=' after attribute " + a);
      ++i;
      ws();
      char q = s[i];
      if (q != '"' && q != '\'') fail("expected a quoted value for " + a);
      size_t e = s.find(q, i + 1);
      if (e == std::string::npos) fail("unterminated attribute value");
      n.attr[a] = s.substr(i + 1, e - i - 1);
      i = e + 1;
    }
    for (;;) {
      skip_misc();
      size_t lt = s.find('<', i);
      if (lt == std::string::npos) fail("unterminated element " + n.name);
      i = lt;  // text content is ignored
      skip_misc();
      if (s.compare(i, 2, "</") == 0) {
        i += 2;
        std::string c = ident();
        if (c != n.name) fail("mismatched </" + c + "> for <" + n.name + ">");
        ws();
        if (i >= s.size() || s[i] != '>') fail("expected '>'");
        ++i;
        return n;
      }
      n.kids.push_back(element());
    }
  }
};

std::array<float, 3> parse_vec3(const std::string& t) {
  std::array<float, 3> v{0, 0, 0};
  std::string u = t;
  std::replace(u.begin(), u.end(), ',', ' ');
  std::istringstream ss(u);
  float a;
  int k = 0;
  while (k < 3 && ss >> a) v[k++] = a;
  if (k == 1) v[1] = v[2] = v[0];
  return v;
}

// load_rgb (SceneLoader.cpp:89-101): <float> or <rgb>/<spectrum>
std::array<float, 3> load_rgb(const XNode* n) {
  if (!n) throw std::runtime_error("missing colour value");
  if (n->name == "float") {
    float v = std::stof(n->get("value", "0"));
    return {v, v, v};
  }
  if (n->name == "rgb" || n->name == "spectrum") return parse_vec3(n->get("value"));
  throw std::runtime_error("unsupported colour element <" + n->name + ">");
}

std::string resolve(const std::string& base, const std::string& f) {
  if (!f.empty() && f[0] == '/') return f;
  size_t s = base.find_last_of('/');
  return s == std::string::npos ? f : base.substr(0, s + 1) + f;
}
}  // namespace

void load_scene_xml(const std::string& filename, Scene& scene) {
  std::ifstream f(filename);
  if (!f) throw std::runtime_error("cannot open " + filename);
  std::stringstream buf;
  buf << f.rdbuf();
  const std::string text = buf.str();
  XParser p(text);
  XNode root = p.element();
  if (root.name != "scene") throw std::runtime_error("XML root must be <scene>");
  std::unordered_map<std::string, int> bsdf_ids;
  // BSDFs (SceneLoader.cpp:200-310)
  for (const XNode& n : root.kids) {
    if (n.name != "bsdf") continue;
    auto b = std::make_shared<BSDF>();
    b->id = n.get("id");
    if (b->id.empty()) throw std::runtime_error("BSDF must have an id");
    const std::string type = n.get("type");
    if (type == "diffuse") {
      b->type = TRC_MAT_DIFFUSE;
      const XNode* r = n.named({"reflectance"});
      if (r && r->name == "texture") {
        const XNode* fn = r->named({"filename"});
        if (!fn) throw std::runtime_error("Failed to retrieve bitmap filename");
        load_texture(resolve(filename, fn->get("value")), b->texels, b->tex_w, b->tex_h);
      } else {
        b->albedo = load_rgb(r);
      }
    } else if (type == "roughconductor") {
      b->type = TRC_MAT_ROUGHCONDUCTOR;
      const XNode* a = n.named({"alpha"});
      b->roughness = a ? std::stof(a->get("value")) : 0.1f;
      b->albedo = load_rgb(n.named({"eta"}));
      b->spec = load_rgb(n.named({"k"}));
    } else if (type == "microfacet") {
      b->type = TRC_MAT_MICROFACET;
      const XNode* r = n.named({"roughness"});
      b->roughness = r ? std::stof(r->get("value")) : 0.3f;
      const XNode* d = n.named({"diffuseReflectance"});
      if (d && d->name == "texture") {
        const XNode* fn = d->named({"filename"});
        if (!fn) throw std::runtime_error("Failed to retrieve bitmap filename");
        load_texture(resolve(filename, fn->get("value")), b->texels, b->tex_w, b->tex_h);
      } else if (d) {
        b->albedo = load_rgb(d);
      }
      const XNode* s = n.named({"specularReflectance"});
      if (s) b->spec = load_rgb(s);
    } else {
      throw std::runtime_error("Unsupported BSDF: " + type);
    }
    bsdf_ids[b->id] = (int)scene.bsdfs.size();
    scene.bsdfs.push_back(b);
  }
  // shapes (SceneLoader.cpp:314-383)
  for (const XNode& n : root.kids) {
    if (n.name != "shape") continue;
    const std::string type = n.get("type");
    if (type != "obj") throw std::runtime_error("Unsupported shape: " + type);
    auto s = std::make_shared<Primitive>();
    s->id = n.get("id");
    const XNode* ref = n.child("ref");
    if (!ref) throw std::runtime_error("Missing BSDF reference");
    auto it = bsdf_ids.find(ref->get("id"));
    if (it == bsdf_ids.end()) throw std::runtime_error("Unknown BSDF id: " + ref->get("id"));
    s->bsdf = it->second;
    if (const XNode* t = n.child("transform")) {
      const std::string tn = t->get("name");
      if (tn != "to_world" && tn != "toWorld") throw std::runtime_error("Invalid transformation name: " + tn);
      M4 m = m4_identity();
      for (const XNode& op : t->kids) {
        auto fl = [&](const char* a, double d) { return op.attr.count(a) ? std::stod(op.get(a)) : d; };
        if (op.name == "translate") m = m4_mul(m4_translate(fl("x", 0), fl("y", 0), fl("z", 0)), m);
        else if (op.name == "rotate")
          m = m4_mul(m4_rotate(fl("angle", 0) * M_PI / 180.0, fl("x", 0), fl("y", 0), fl("z", 0)), m);
        else if (op.name == "scale") m = m4_mul(m4_scale(fl("x", 1), fl("y", 1), fl("z", 1)), m);
        else throw std::runtime_error("Unsupported transformation: " + op.name);
      }
      s->to_world = m;
    }
    const XNode* fn = n.named({"filename"});
    if (!fn) throw std::runtime_error("shape without filename");
    load_obj(resolve(filename, fn->get("value")), *s);
    bool face_normals = true;
    if (const XNode* fnn = n.named({"face_normals", "faceNormals"})) face_normals = fnn->get("value") == "true";
    s->smooth = !face_normals;
    if (const XNode* e = n.child("emitter")) {
      if (e->get("type") != "area") throw std::runtime_error("Unsupported emitter: " + e->get("type"));
      s->emitter = true;
      s->radiance = load_rgb(e->named({"radiance"}));
      auto l = std::make_shared<Light>();
      l->shape = (int)scene.shapes.size();
      l->radiance = s->radiance;
      scene.emitters.push_back(l);
    }
    scene.shapes.push_back(s);
  }
  // sensors (SceneLoader.cpp:385-410)
  for (const XNode& n : root.kids) {
    if (n.name != "sensor") continue;
    if (n.get("type") != "perspective") throw std::runtime_error("Unsupported sensor: " + n.get("type"));
    const XNode* film = n.child("film");
    int w = 512, h = 512;
    if (film) {
      if (const XNode* x = film->named({"width"})) w = std::stoi(x->get("value"));
      if (const XNode* y = film->named({"height"})) h = std::stoi(y->get("value"));
    }
    const XNode* t = n.child("transform");
    const XNode* la = t ? t->child("lookat") : nullptr;
    if (!la) throw std::runtime_error("LookAt transformation for sensor is missing!");
    auto o = parse_vec3(la->get("origin")), tg = parse_vec3(la->get("target")), up = parse_vec3(la->get("up"));
    const XNode* fov = n.named({"fov"});
    auto c = std::make_shared<Camera>(Vector3f(o[0], o[1], o[2]), Vector3f(tg[0], tg[1], tg[2]),
                                      Vector3f(up[0], up[1], up[2]), w, h, fov ? std::stof(fov->get("value")) : 45.0f);
    scene.cameras.push_back(c);
  }
}

void Scene::load_file(const std::string& filename) { load_scene_xml(filename, *this); }

}  // namespace trpy
