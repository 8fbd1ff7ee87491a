// Host-side scene, parameter and integrator classes of the `TensorRay`
// Python module, layered on the C ABI of include/trc.h.
//
// They mirror the reference's host API (TensorRay.cpp:48-222): Scene with
// shapes / bsdfs / emitters / cameras, Shape (Primitive) with its
// differentiable parameters (diff_all_vertex_pos / diff_translation /
// diff_rotation, Primitive.cpp:134-210 and 374-388), BSDF::DiffTexture
// (BSDF.cpp:350-356), RenderOptions (Integrator.h:47-78), the integrators'
// renderC / renderD and the ParameterPool accessors get_size_param /
// get_var / set_var / get_grad (pybind_utils.h, Tensor.h:1408-1419,
// 1675-1685).  All device work goes through libtrc; images and gradient
// buffers are HIP device allocations owned by these classes.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/trc.h"

namespace trpy {

// ---- environment (EnvCreate / EnvRelease, pybind_utils.h:38-61) ----
struct Env {
  trc_ctx* ctx = nullptr;
  uint64_t rnd_seed = 0;  // set_rnd_seed
};
Env& env();
void env_create(int device);
void env_release();
trc_ctx* ctx();
void check(trc_status st, const char* what);

// ---- pointer wrappers (ptr_wrapper<T>, pybind_utils.h) ----
template <class T>
struct PtrWrapper {
  T* p;
  explicit PtrWrapper(std::size_t addr) : p(reinterpret_cast<T*>(addr)) {}
};
using FloatPtr = PtrWrapper<float>;
using IntPtr = PtrWrapper<int>;

struct Vector3f {
  float x = 0.0f, y = 0.0f, z = 0.0f;
  Vector3f() = default;
  Vector3f(float a, float b, float c) : x(a), y(b), z(c) {}
};

// ---- tensors ----
enum VecType { kScalar1 = 1, kVec2 = 2, kVec3 = 3, kVec4 = 4, kMat4x4 = 16 };

struct TensorShape {
  std::vector<int> dims;
  int vtype = kScalar1;
  TensorShape() = default;
  TensorShape(std::vector<int> d, int v) : dims(std::move(d)), vtype(v) {}
  int64_t elements() const;  // product of dims
  int64_t size() const { return elements() * vtype; }
};

struct DevArray {
  float* p = nullptr;
  int64_t n = 0;
  DevArray() = default;
  explicit DevArray(int64_t count);
  ~DevArray();
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  void zero();
};

// A Tensorf: small tensors (parameters, scalars) live on the host, images
// and adjoint images on the device ([vtype][elements] planes, as the
// reference's SoA Vec3 tensors).
struct Tensorf {
  TensorShape shape;
  std::vector<float> host;
  std::shared_ptr<DevArray> dev;
  bool requires_grad = false;
  // A standalone parameter (Tensorf(val, requiresGrad=True)) is pool entry
  // `param`; a product of a constant Vec3 and such a parameter (the
  // reference's `Tensorf(x, y, z) * param` expression) is linked to it:
  // value = lin_dir * current value of pool entry lin_param.
  int param = -1;
  int lin_param = -1;
  std::array<float, 3> lin_dir{0.0f, 0.0f, 0.0f};
  std::vector<float> resolved() const;  // current host value (links followed)
  Tensorf() = default;
  int64_t size() const { return dev ? dev->n : (int64_t)host.size(); }
  void value_to(float* dst) const;
  const float* device_ptr() const { return dev ? dev->p : nullptr; }
  static Tensorf from_ptr(const float* src, const TensorShape& shape);
  static Tensorf scalar(float v, bool rg);
  static Tensorf vec3(float x, float y, float z, bool rg);
};

// ---- scene description ----
struct BSDF : std::enable_shared_from_this<BSDF> {
  int type = TRC_MAT_DIFFUSE;
  std::array<float, 3> albedo{0.5f, 0.5f, 0.5f};  // diffuse reflectance / kd / eta
  std::array<float, 3> spec{0.04f, 0.04f, 0.04f};  // F0 / k
  float roughness = 0.3f;
  std::vector<float> texels;  // albedo / kd texture [h][w][3], empty = constant
  int tex_w = 0, tex_h = 0;
  std::string id;
  void diff_texture(const std::string& name);
};

struct Light {
  int shape = -1;
  std::array<float, 3> radiance{0.0f, 0.0f, 0.0f};
};

struct Camera {
  std::array<float, 3> pos{0, 0, 0}, target{0, 0, -1}, up{0, 1, 0};
  int width = 1, height = 1;
  float fov = 45.0f;
  Camera() = default;
  Camera(const Vector3f& p, const Vector3f& t, const Vector3f& u, int w, int h, float f);
  void init(const Vector3f& p, const Vector3f& t, const Vector3f& u, int w, int h, float f);
  // Camera::Update (Camera.h:110-113): takes over another camera's pose,
  // resolution and fov.
  void update(const Camera& c);
  void resize(int w, int h);
  trc_camera to_trc() const;
  Tensorf pos_tensor() const { return Tensorf::vec3(pos[0], pos[1], pos[2], false); }
};

struct Primitive : std::enable_shared_from_this<Primitive> {
  std::string id;
  std::vector<float> raw;      // object-space positions [V][3]
  std::vector<uint32_t> idx;   // [T][3]
  std::vector<float> uv;       // [nuv][2]
  std::vector<uint32_t> uv_idx;
  int bsdf = 0;
  bool smooth = false;
  std::array<double, 16> to_world{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};  // row-major
  // differentiable parameters
  bool diff_pos = false;
  Tensorf translation;    // empty or Vec3
  Tensorf rotate_axis;    // empty or Vec3
  Tensorf rotate_angle;   // scalar, radians
  Tensorf obj_center;     // multi-pose
  Tensorf rotate_matrix;  // multi-pose, 4x4 row-major
  int vertex_id = -1;
  Tensorf vertex_translation;
  // emitter
  bool emitter = false;
  std::array<float, 3> radiance{0, 0, 0};

  int vertex_count() const { return (int)(raw.size() / 3); }
  int face_count() const { return (int)(idx.size() / 3); }
  int edge_count() const;
  void face_indices(int* out) const;
  void edge_data(int* out) const;
  Tensorf get_obj_center() const;
  void export_mesh(const std::string& filename) const;
  void diff_translation();
  void diff_rotation(const Tensorf& axis);
  void diff_all_vertex_pos();
  // world positions and the 3x3 linear part of the world transform
  void world_positions(std::vector<float>& out, std::array<double, 9>& lin) const;
  std::array<double, 16> world_matrix() const;
};

struct RenderOptions {
  int seed = 1234, max_bounces = 1, spp = 1, sppe = 0, sppse0 = 0, sppse1 = 0, sppe0 = 0;
  int spp_batch = 1, sppe_batch = 0, sppse0_batch = 0, sppse1_batch = 0, sppe0_batch = 0;
  bool export_deriv = false, quiet = true;
  // keep the path checkpoint of renderC for the next renderD with the same
  // scene state, camera and options (adjoint without re-tracing)
  bool checkpoint = false;
  bool g_direct = false;
  int g_direct_depth = 0, g_direct_max_size = 100000, g_direct_spp = 16;
  float g_direct_thold = 0.01f, g_eps = 0.01f;
  // guide reuse across views (trc_options.g_keep / g_seed); off as in the reference
  bool g_keep = false;
  uint32_t g_seed = 0;
  RenderOptions() = default;
  RenderOptions(int s, int b, int sp, int se, int s0, int s1, int e0)
      : seed(s), max_bounces(b), spp(sp), sppe(se), sppse0(s0), sppse1(s1), sppe0(e0),
        spp_batch(sp), sppe_batch(se), sppse0_batch(s0), sppse1_batch(s1), sppe0_batch(e0) {}
  trc_options to_trc() const;
};

struct Scene {
  std::vector<std::shared_ptr<Primitive>> shapes;
  std::vector<std::shared_ptr<BSDF>> bsdfs;
  std::vector<std::shared_ptr<Light>> emitters;
  std::vector<std::shared_ptr<Camera>> cameras;
  trc_scene* sc = nullptr;
  // layout of the configured device scene
  std::vector<int> vert_off;       // per shape, into the flattened vertex array
  std::vector<int> tex_index;      // per bsdf, texture index or -1
  std::vector<int64_t> tex_off;    // per bsdf, float offset into the texture-gradient buffer
  int64_t tex_floats = 0;
  int nv = 0;
  std::string topo_key;            // topology signature of the configured scene
  std::vector<std::array<double, 9>> lin;  // per shape, linear part of its world transform
  std::vector<float> world;        // flattened world positions
  // gradient accumulators (device) of the render_d calls since the last fold
  std::unique_ptr<DevArray> dpos, dmat, dtex;
  bool grads_pending = false;
  uint64_t version = 0;  // bumped by every configure()
  const uint64_t uid = next_uid();  // unique per Scene object of the process (checkpoint keys)

  Scene() = default;
  ~Scene();
  void load_file(const std::string& filename);
  void configure();
  int get_width(int i = 0) const;
  int get_height(int i = 0) const;
  int num_bsdf() const { return (int)bsdfs.size(); }
  void ensure_grad_buffers();
  void fold_gradients();  // moves device gradients into the parameter pool

 private:
  static uint64_t next_uid() {
    static std::atomic<uint64_t> n{0};
    return ++n;
  }
};

// ---- parameter pool (Tensor.h:1408-1419, 1675-1685) ----
enum ParamKind { kParamVertices, kParamTranslation, kParamRotation, kParamTexture, kParamScalar };
struct Param {
  ParamKind kind;
  std::weak_ptr<Primitive> prim;
  std::weak_ptr<BSDF> bsdf;
  std::string slot;  // texture slot name
  std::vector<float> value;  // kParamScalar: its value
  std::vector<double> grad;
  int64_t size() const;
};
std::vector<Param>& param_pool();
// Registers a standalone scalar parameter; returns its pool index.
int register_scalar_param(float v);
// The reference's Tensorf product for the shapes it takes in scene
// set-up: constant * constant, and constant Vec3 * scalar parameter.
Tensorf tensor_mul(const Tensorf& a, const Tensorf& b);
int get_num_param();
int get_size_param(std::vector<int>& sizes);
void get_var(int index, float* dst, int offset);
void set_var(int index, const float* src, int offset);
void get_grad(int index, float* dst, int offset, bool clear);
void set_rnd_seed(uint64_t seed);
Scene*& active_scene();  // the scene whose gradients feed the pool

// ---- integrators (Integrator.h:96-133) ----
struct Integrator {
  RenderOptions opts;
  virtual ~Integrator() = default;
  void set_param(const RenderOptions& o) { opts = o; }
};
struct PathTracer : Integrator {
  std::shared_ptr<DevArray> ckpt;  // path checkpoint of the last renderC (RenderOptions.checkpoint)
  std::string ckpt_key;
  Tensorf render_c(Scene& scene, const RenderOptions& o);
  Tensorf render_d(Scene& scene, const RenderOptions& o, const Tensorf& dLdI);
};
// RenderD with RenderOptions.export_deriv (Integrator.h:106-121): one
// derivative image [W*H] Vec3 per element of every pool parameter,
// concatenated; `kind` -1 = interior, 0..3 = boundary terms.
Tensorf derivative_images(Scene& scene, const RenderOptions& o, int kind);
struct BoundaryIntegrator : Integrator {
  int kind;
  explicit BoundaryIntegrator(int k) : kind(k) {}
  Tensorf render_d(Scene& scene, const RenderOptions& o, const Tensorf& dLdI);
};
struct PrimaryEdgeIntegrator : BoundaryIntegrator { PrimaryEdgeIntegrator() : BoundaryIntegrator(0) {} };
struct PixelBoundaryIntegrator : BoundaryIntegrator { PixelBoundaryIntegrator() : BoundaryIntegrator(1) {} };
struct DirectEdgeIntegrator : BoundaryIntegrator { DirectEdgeIntegrator() : BoundaryIntegrator(2) {} };
struct IndirectEdgeIntegrator : BoundaryIntegrator { IndirectEdgeIntegrator() : BoundaryIntegrator(3) {} };

// ---- file formats ----
// Wavefront OBJ: v / vt / vn / f (polygons fan-triangulated, negative indices).
void load_obj(const std::string& filename, Primitive& out);
// Mitsuba-style scene XML subset (SceneLoader.cpp:176-410).
void load_scene_xml(const std::string& filename, Scene& scene);
// Textures: PFM (float RGB) or binary PPM (8-bit, sRGB-encoded values taken as is / 255).
void load_texture(const std::string& filename, std::vector<float>& texels, int& w, int& h);

}  // namespace trpy
