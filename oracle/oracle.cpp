// ORACLE -- test infrastructure only.  A scalar CPU restatement of the
// reference integrator (TensorRay, LibTensorRay/Renderer) used as the checker
// for the HIP path: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it.  The product never links or calls this file.
//
// What it restates (reference file:line):
//   camera rays          Camera.cpp:224-277, BaseCamera.cpp:45-88, Matrix.cpp:181-220
//   primary weight       Camera.cpp:279-303 (EvalFilter, box filter)
//   PostIntersect        Scene.cpp:668-730
//   face normal / area   Primitive.cpp:299-321; vertex normals Primitive.cpp:323-340
//   light                Scene.cpp:471-520 (ConfigureLight), Light.cpp:134-184
//   distribution         Distribution.cpp:37-58, Distribution.cu:37-48
//   BSDFs                BSDF.cpp:100-302, Diffuse.cpp:37-85, Microfacet.cpp:37-165
//   sampling             Utils.cpp:70-140
//   path tracer          PathTracer.cpp:37-123, Integrator.cpp:55-303
//   hit records          ptx.cu:92-113 (miss: tri -1, u = v = 0, t = -1)
//   reverse-mode AD      Tensor.cpp:601-635 (a scalar tape instead of tensors)
//   RNG                  Philox4x32-10 (Salmon et al., SC'11) in place of the
//                        reference's cuRAND stream (Tensor.h:2792-2802)
//
// Numerics: every float expression is written in the order of the reference
// expression it restates and the file is compiled with -ffp-contract=off, so
// the values equal the HIP kernels' bit for bit.  Elementary functions use
// their own range reduction + Taylor series (sin/cos, exp2) so that host and
// device agree; their coefficients are the Taylor coefficients 1/n!.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------------------
// Philox4x32-10
// ---------------------------------------------------------------------------
struct U4 {
  uint32_t x, y, z, w;
};
static inline U4 philox(U4 c, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    U4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}
static inline float unitf(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }
struct R4 {
  float a, b, c, d;
};
static inline R4 rnd4(uint32_t seed, uint32_t domain, uint64_t index, uint32_t block) {
  U4 v = philox(U4{(uint32_t)index, (uint32_t)(index >> 32), block, 0u}, seed, domain);
  return R4{unitf(v.x), unitf(v.y), unitf(v.z), unitf(v.w)};
}
constexpr uint32_t kDomInterior = 0x11u;

// ---------------------------------------------------------------------------
// Elementary functions (Cody-Waite reduction with a two-part pi/2 and
// Taylor series; see the header comment)
// ---------------------------------------------------------------------------
static inline void f_sincos(float x, float* so, float* co) {
  float kf = rintf(x * 0.636619772367581343f);
  float r = fmaf(-kf, 1.57079637050628662109375f, x);
  r = fmaf(-kf, -4.37113900018624283e-8f, r);
  float r2 = r * r;
  float ps = -2.50521083854417188e-8f;
  ps = fmaf(ps, r2, 2.75573192239858907e-6f);
  ps = fmaf(ps, r2, -1.98412698412698413e-4f);
  ps = fmaf(ps, r2, 8.33333333333333333e-3f);
  ps = fmaf(ps, r2, -1.66666666666666667e-1f);
  float sr = fmaf(ps * r2, r, r);
  float pc = 2.08767569878680990e-9f;
  pc = fmaf(pc, r2, -2.75573192239858907e-7f);
  pc = fmaf(pc, r2, 2.48015873015873016e-5f);
  pc = fmaf(pc, r2, -1.38888888888888889e-3f);
  pc = fmaf(pc, r2, 4.16666666666666667e-2f);
  pc = fmaf(pc, r2, -0.5f);
  float cr = fmaf(pc, r2, 1.0f);
  int k = (int)kf & 3;
  if (k == 0) { *so = sr; *co = cr; }
  else if (k == 1) { *so = cr; *co = -sr; }
  else if (k == 2) { *so = -sr; *co = -cr; }
  else { *so = -cr; *co = sr; }
}
static inline float f_exp2(float x) {
  if (x < -126.0f) return 0.0f;
  if (x > 127.0f) return INFINITY;
  float kf = floorf(x);
  float y = (x - kf) * 0.693147180559945309f;
  float p = 2.75573192239858907e-7f;
  p = fmaf(p, y, 2.75573192239858907e-6f);
  p = fmaf(p, y, 2.48015873015873016e-5f);
  p = fmaf(p, y, 1.98412698412698413e-4f);
  p = fmaf(p, y, 1.38888888888888889e-3f);
  p = fmaf(p, y, 8.33333333333333333e-3f);
  p = fmaf(p, y, 4.16666666666666667e-2f);
  p = fmaf(p, y, 1.66666666666666667e-1f);
  p = fmaf(p, y, 0.5f);
  p = fmaf(p, y, 1.0f);
  p = fmaf(p, y, 1.0f);
  return ldexpf(p, (int)kf);
}

// ---------------------------------------------------------------------------
// Scalar reverse-mode AD (tape per thread, per path)
// ---------------------------------------------------------------------------
struct Node {
  int a, b;
  double da, db;
};
struct Tape {
  std::vector<Node> nodes;
  std::vector<double> adj;
  void clear() { nodes.clear(); }
};
static thread_local Tape* g_tape = nullptr;

struct Var {
  float v;
  int i;  // tape index, -1 = constant
  Var() : v(0.0f), i(-1) {}
  Var(float x) : v(x), i(-1) {}
  Var(float x, int idx) : v(x), i(idx) {}
};
static inline int push_node(int a, double da, int b, double db) {
  if (a < 0 && b < 0) return -1;
  g_tape->nodes.push_back(Node{a, b, da, db});
  return (int)g_tape->nodes.size() - 1;
}
static inline Var leaf(float v) {
  g_tape->nodes.push_back(Node{-1, -1, 0.0, 0.0});
  return Var(v, (int)g_tape->nodes.size() - 1);
}
static inline Var operator+(Var a, Var b) { return Var(a.v + b.v, push_node(a.i, 1.0, b.i, 1.0)); }
static inline Var operator-(Var a, Var b) { return Var(a.v - b.v, push_node(a.i, 1.0, b.i, -1.0)); }
static inline Var operator-(Var a) { return Var(-a.v, push_node(a.i, -1.0, -1, 0.0)); }
static inline Var operator*(Var a, Var b) { return Var(a.v * b.v, push_node(a.i, (double)b.v, b.i, (double)a.v)); }
static inline Var operator/(Var a, Var b) {
  double ib = 1.0 / (double)b.v;
  return Var(a.v / b.v, push_node(a.i, ib, b.i, -(double)a.v * ib * ib));
}
static inline Var sqrt_(Var a) {
  float r = sqrtf(a.v);
  return Var(r, push_node(a.i, 0.5 / (double)r, -1, 0.0));
}
static inline Var abs_(Var a) {
  return Var(fabsf(a.v), push_node(a.i, a.v > 0.0f ? 1.0 : (a.v < 0.0f ? -1.0 : 0.0), -1, 0.0));
}
static inline Var exp2_(Var a) {
  float r = f_exp2(a.v);
  return Var(r, push_node(a.i, (double)r * 0.693147180559945309, -1, 0.0));
}
static inline Var detach(Var a) { return Var(a.v); }
static inline float val(Var a) { return a.v; }
static inline float sqrt_(float a) { return sqrtf(a); }
static inline float abs_(float a) { return fabsf(a); }
static inline float exp2_(float a) { return f_exp2(a); }
static inline float detach(float a) { return a; }
static inline float val(float a) { return a; }

// Twin evaluation for finite differences of the material-form estimator:
// `b` is the unperturbed float value that drives every comparison, branch
// and detach(); `s` is the perturbed evaluation in float64.  With a vertex
// coordinate shifted by +-delta in `s` only, (F(+d) - F(-d)) / 2d is a
// central difference of exactly the function the reverse-mode AD
// differentiates (same hit triangles, barycentrics and detached factors).
struct Pair {
  float b;
  double s;
  Pair() : b(0.0f), s(0.0) {}
  Pair(float x) : b(x), s((double)x) {}
  Pair(float x, double y) : b(x), s(y) {}
};
static inline Pair operator+(Pair a, Pair c) { return Pair(a.b + c.b, a.s + c.s); }
static inline Pair operator-(Pair a, Pair c) { return Pair(a.b - c.b, a.s - c.s); }
static inline Pair operator-(Pair a) { return Pair(-a.b, -a.s); }
static inline Pair operator*(Pair a, Pair c) { return Pair(a.b * c.b, a.s * c.s); }
static inline Pair operator/(Pair a, Pair c) { return Pair(a.b / c.b, a.s / c.s); }
static inline Pair sqrt_(Pair a) { return Pair(sqrtf(a.b), std::sqrt(a.s)); }
static inline Pair abs_(Pair a) { return Pair(fabsf(a.b), a.b >= 0.0f ? a.s : -a.s); }
static inline Pair exp2_(Pair a) { return Pair(f_exp2(a.b), std::exp2(a.s)); }
static inline Pair detach(Pair a) { return Pair(a.b, (double)a.b); }
static inline float val(Pair a) { return a.b; }

template <class S>
struct V3 {
  S x, y, z;
};
template <class S> static inline V3<S> mk(S x, S y, S z) { return V3<S>{x, y, z}; }
template <class S> static inline V3<S> operator+(V3<S> a, V3<S> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class S> static inline V3<S> operator-(V3<S> a, V3<S> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class S> static inline V3<S> operator-(V3<S> a) { return {-a.x, -a.y, -a.z}; }
template <class S> static inline V3<S> operator*(V3<S> a, V3<S> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class S> static inline V3<S> operator*(V3<S> a, S s) { return {a.x * s, a.y * s, a.z * s}; }
template <class S> static inline V3<S> operator*(S s, V3<S> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class S> static inline V3<S> operator/(V3<S> a, S s) { return {a.x / s, a.y / s, a.z / s}; }
template <class S> static inline S dot(V3<S> a, V3<S> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class S> static inline V3<S> cross(V3<S> a, V3<S> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class S> static inline S len2(V3<S> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <class S> static inline S length(V3<S> a) { return sqrt_(len2(a)); }
template <class S> static inline V3<S> normalize(V3<S> a) { return a / sqrt_(len2(a)); }
template <class S> static inline V3<S> detach3(V3<S> a) { return {detach(a.x), detach(a.y), detach(a.z)}; }
template <class S> static inline V3<float> val3(V3<S> a) { return {val(a.x), val(a.y), val(a.z)}; }
template <class S> static inline V3<S> cast3(V3<float> a) { return {S(a.x), S(a.y), S(a.z)}; }
template <class S> static inline bool finite3(V3<S> a) {
  return std::isfinite(val(a.x)) && std::isfinite(val(a.y)) && std::isfinite(val(a.z));
}
template <class S> static inline S sel(bool c, S a, S b) { return c ? a : b; }
using F3 = V3<float>;

// ---------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------
constexpr float EPSILON_ = 1e-6f;
constexpr float EDGE_EPSILON_ = 1e-8f;
constexpr float SHADOW_EPSILON_ = 1e-2f;
constexpr float INV_PI = 0.31830988618379067154f;
constexpr float PI = 3.14159265358979323846f;

struct Mat {
  int type, tex;
  float roughness;
  F3 albedo, spec;
};

struct BNode {
  float lo[3], hi[3];
  int left, right;  // inner children or -1
  int first, count; // leaf range
};

struct Scene {
  int V = 0, T = 0;
  std::vector<float> pos;        // [V][3]
  std::vector<uint32_t> idx;     // [T][3]
  std::vector<int> tri_mat, tri_emitter, tri_flags;
  std::vector<float> emit_rad;   // [n][3]
  std::vector<Mat> mats;
  std::vector<float> uv;
  std::vector<uint32_t> uv_idx;
  std::vector<std::vector<float>> tex;
  std::vector<int> tex_w, tex_h;
  // derived (float)
  std::vector<F3> fn;            // face normals
  std::vector<float> area;
  std::vector<F3> vn;            // vertex normals
  std::vector<std::vector<int>> vfaces;  // vertex -> incident smooth faces (triangle order)
  std::vector<int> light_tri, emit_tri;  // emitter triangle table and inverse
  std::vector<float> light_f, light_cdf;
  float light_int = 0.0f;
  // simple binary BVH for the oracle's own ray queries
  std::vector<BNode> bvh;
  std::vector<int> bvh_tris;
};

static inline F3 P(const Scene& s, uint32_t v) { return F3{s.pos[3 * v], s.pos[3 * v + 1], s.pos[3 * v + 2]}; }

static void scene_derive(Scene& s) {
  // Primitive.cpp:299-321
  s.fn.resize(s.T);
  s.area.resize(s.T);
  for (int t = 0; t < s.T; ++t) {
    F3 p0 = P(s, s.idx[3 * t]), p1 = P(s, s.idx[3 * t + 1]), p2 = P(s, s.idx[3 * t + 2]);
    F3 c = cross(p1 - p0, p2 - p0);
    float l = length(c);
    s.fn[t] = c / l;
    s.area[t] = l * 0.5f;
  }
  // Primitive.cpp:323-340: normalize(sum of face normals / count)
  s.vn.assign(s.V, F3{0, 0, 0});
  s.vfaces.assign(s.V, {});
  for (int t = 0; t < s.T; ++t) {
    if (!(s.tri_flags[t] & 1)) continue;
    for (int j = 0; j < 3; ++j) s.vfaces[s.idx[3 * t + j]].push_back(t);
  }
  for (int v = 0; v < s.V; ++v) {
    if (s.vfaces[v].empty()) continue;
    F3 a{0, 0, 0};
    for (int t : s.vfaces[v]) a = a + s.fn[t];
    float c = (float)s.vfaces[v].size();
    s.vn[v] = normalize(F3{a.x / c, a.y / c, a.z / c});
  }
  // Scene.cpp:471-520 + Distribution.cpp:45-58
  s.light_tri.clear();
  s.emit_tri.assign(s.T, -1);
  for (int t = 0; t < s.T; ++t)
    if (s.tri_emitter[t] >= 0) {
      s.emit_tri[t] = (int)s.light_tri.size();
      s.light_tri.push_back(t);
    }
  int n = (int)s.light_tri.size();
  s.light_f.assign(n, 0.0f);
  s.light_cdf.assign(n, 0.0f);
  s.light_int = 0.0f;
  if (n > 0) {
    for (int i = 0; i < n; ++i) {
      int t = s.light_tri[i];
      const float* I = &s.emit_rad[3 * s.tri_emitter[t]];
      float lum = 0.212671f * I[0] + 0.715160f * I[1] + 0.072169f * I[2];
      s.light_f[i] = lum * s.area[t];
    }
    float inv = 1.0f / (float)n, run = 0.0f;
    for (int i = 0; i < n; ++i) {
      run += s.light_f[i];
      s.light_cdf[i] = run * inv;
    }
    s.light_int = s.light_cdf[n - 1];
    if (s.light_int > 0.0f)
      for (int i = 0; i < n; ++i) s.light_cdf[i] = s.light_cdf[i] / s.light_int;
  }
}

// Median-split binary BVH (independent of the product's SAH 4-wide BVH).
static int bvh_rec(Scene& s, std::vector<int>& tris, int first, int count, std::vector<float>& cent) {
  BNode n;
  for (int k = 0; k < 3; ++k) { n.lo[k] = FLT_MAX; n.hi[k] = -FLT_MAX; }
  for (int i = first; i < first + count; ++i)
    for (int j = 0; j < 3; ++j) {
      F3 p = P(s, s.idx[3 * tris[i] + j]);
      float c[3] = {p.x, p.y, p.z};
      for (int k = 0; k < 3; ++k) { n.lo[k] = std::min(n.lo[k], c[k]); n.hi[k] = std::max(n.hi[k], c[k]); }
    }
  // pad outward so the slab test is conservative
  for (int k = 0; k < 3; ++k) {
    n.lo[k] -= std::fabs(n.lo[k]) * 1e-6f + 1e-6f;
    n.hi[k] += std::fabs(n.hi[k]) * 1e-6f + 1e-6f;
  }
  int self = (int)s.bvh.size();
  s.bvh.push_back(n);
  if (count <= 4) {
    s.bvh[self].left = s.bvh[self].right = -1;
    s.bvh[self].first = first;
    s.bvh[self].count = count;
    return self;
  }
  int axis = 0;
  float ext = 0.0f;
  for (int k = 0; k < 3; ++k)
    if (n.hi[k] - n.lo[k] > ext) { ext = n.hi[k] - n.lo[k]; axis = k; }
  int mid = first + count / 2;
  std::nth_element(tris.begin() + first, tris.begin() + mid, tris.begin() + first + count,
                   [&](int a, int b) { return cent[3 * a + axis] < cent[3 * b + axis]; });
  int l = bvh_rec(s, tris, first, mid - first, cent);
  int r = bvh_rec(s, tris, mid, first + count - mid, cent);
  s.bvh[self].left = l;
  s.bvh[self].right = r;
  s.bvh[self].first = 0;
  s.bvh[self].count = 0;
  return self;
}

static void scene_build_bvh(Scene& s) {
  s.bvh.clear();
  s.bvh_tris.resize(s.T);
  std::vector<float> cent(3 * (size_t)s.T);
  for (int t = 0; t < s.T; ++t) {
    s.bvh_tris[t] = t;
    F3 c = (P(s, s.idx[3 * t]) + P(s, s.idx[3 * t + 1])) + P(s, s.idx[3 * t + 2]);
    cent[3 * t] = c.x; cent[3 * t + 1] = c.y; cent[3 * t + 2] = c.z;
  }
  bvh_rec(s, s.bvh_tris, 0, s.T, cent);
}

// ---------------------------------------------------------------------------
// Watertight ray / triangle test (Woop, Benthin, Wald, JCGT 2013)
// ---------------------------------------------------------------------------
struct Ray {
  F3 o, d;
  int kx, ky, kz;
  float Sx, Sy, Sz;
};
static inline float cmp(F3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
static inline Ray make_ray(F3 o, F3 d) {
  Ray r;
  r.o = o; r.d = d;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  r.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  r.kx = r.kz == 2 ? 0 : r.kz + 1;
  r.ky = r.kx == 2 ? 0 : r.kx + 1;
  float dz = cmp(d, r.kz);
  if (dz < 0.0f) std::swap(r.kx, r.ky);
  r.Sx = cmp(d, r.kx) / dz;
  r.Sy = cmp(d, r.ky) / dz;
  r.Sz = 1.0f / dz;
  return r;
}
static inline bool tri_test(const Ray& r, F3 p0, F3 p1, F3 p2, float* t, float* u, float* v) {
  F3 A = p0 - r.o, B = p1 - r.o, C = p2 - r.o;
  float Akz = cmp(A, r.kz), Bkz = cmp(B, r.kz), Ckz = cmp(C, r.kz);
  float Ax = cmp(A, r.kx) - r.Sx * Akz, Ay = cmp(A, r.ky) - r.Sy * Akz;
  float Bx = cmp(B, r.kx) - r.Sx * Bkz, By = cmp(B, r.ky) - r.Sy * Bkz;
  float Cx = cmp(C, r.kx) - r.Sx * Ckz, Cy = cmp(C, r.ky) - r.Sy * Ckz;
  float U = Cx * By - Cy * Bx, Vv = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
  if (U == 0.0f || Vv == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    Vv = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || Vv < 0.0f || W < 0.0f) && (U > 0.0f || Vv > 0.0f || W > 0.0f)) return false;
  float det = U + Vv + W;
  if (det == 0.0f) return false;
  float T = U * (r.Sz * Akz) + Vv * (r.Sz * Bkz) + W * (r.Sz * Ckz);
  float inv = 1.0f / det;
  *t = T * inv;
  *u = Vv * inv;
  *v = W * inv;
  return true;
}

struct Hit {
  int tri;
  float u, v, t;
};

// Closest hit: smallest t in (tmin, tmax), ties to the smallest triangle id.
static Hit intersect(const Scene& s, F3 o, F3 d, float tmin, float tmax, bool any, bool brute) {
  Ray r = make_ray(o, d);
  Hit h{-1, 0.0f, 0.0f, -1.0f};
  float best = tmax;
  auto test = [&](int t) {
    float tt, uu, vv;
    if (!tri_test(r, P(s, s.idx[3 * t]), P(s, s.idx[3 * t + 1]), P(s, s.idx[3 * t + 2]), &tt, &uu, &vv)) return false;
    if (!(tt > tmin && tt < tmax)) return false;
    if (h.tri < 0 || tt < best || (tt == best && t < h.tri)) {
      best = tt; h.tri = t; h.u = uu; h.v = vv; h.t = tt;
    }
    return true;
  };
  if (brute) {
    for (int t = 0; t < s.T; ++t)
      if (test(t) && any) return h;
    return h;
  }
  // robust slab test in double with generous padding (boxes are padded)
  double inv[3] = {1.0 / (double)d.x, 1.0 / (double)d.y, 1.0 / (double)d.z};
  double oo[3] = {o.x, o.y, o.z};
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BNode& n = s.bvh[stack[--sp]];
    double t0 = tmin, t1 = (double)best * (1.0 + 1e-6);
    bool miss = false;
    for (int k = 0; k < 3 && !miss; ++k) {
      if (!std::isfinite(inv[k])) {
        if (oo[k] < n.lo[k] || oo[k] > n.hi[k]) miss = true;
        continue;
      }
      double a = (n.lo[k] - oo[k]) * inv[k], b = (n.hi[k] - oo[k]) * inv[k];
      if (a > b) std::swap(a, b);
      t0 = std::max(t0, a);
      t1 = std::min(t1, b);
      if (t0 > t1) miss = true;
    }
    if (miss) continue;
    if (n.left < 0) {
      for (int i = n.first; i < n.first + n.count; ++i)
        if (test(s.bvh_tris[i]) && any) return h;
    } else {
      stack[sp++] = n.left;
      stack[sp++] = n.right;
    }
  }
  return h;
}

// ---------------------------------------------------------------------------
// Camera (reduced form of the reference matrices; see make_camera)
// ---------------------------------------------------------------------------
struct Cam {
  F3 eye, right, up, dir;
  float ax, bx, ay, by, tan_half, ratio, inv_area;
  int W, H;
};

static Cam make_camera(const float pos[3], const float target[3], const float upv[3], int w, int h, float fov) {
  // LookAt (Matrix.cpp:181-209) columns and Perspective (Matrix.cpp:211-220)
  // with the screen-to-raster map of BaseCamera::Resize (BaseCamera.cpp:73-88):
  // raster (rx, ry) lies on the z = 1 plane at
  // (ratio tan(fov/2) (1 - 2 rx / W), tan(fov/2) (1 - 2 ry / H), 1).
  double e[3] = {pos[0], pos[1], pos[2]};
  double d[3] = {(double)target[0] - e[0], (double)target[1] - e[1], (double)target[2] - e[2]};
  double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (double& x : d) x /= dl;
  double u[3] = {upv[0], upv[1], upv[2]};
  double ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (double& x : u) x /= ul;
  double r[3] = {u[1] * d[2] - u[2] * d[1], u[2] * d[0] - u[0] * d[2], u[0] * d[1] - u[1] * d[0]};
  double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  for (double& x : r) x /= rl;
  double nu[3] = {d[1] * r[2] - d[2] * r[1], d[2] * r[0] - d[0] * r[2], d[0] * r[1] - d[1] * r[0]};
  double th = std::tan(fov * 3.14159265358979323846 / 180.0 / 2.0);
  double ratio = (double)w / (double)h;
  Cam c;
  c.eye = F3{(float)e[0], (float)e[1], (float)e[2]};
  c.right = F3{(float)r[0], (float)r[1], (float)r[2]};
  c.up = F3{(float)nu[0], (float)nu[1], (float)nu[2]};
  c.dir = F3{(float)d[0], (float)d[1], (float)d[2]};
  c.ax = (float)(-2.0 * ratio * th / w);
  c.bx = (float)(ratio * th);
  c.ay = (float)(-2.0 * th / h);
  c.by = (float)th;
  c.tan_half = (float)th;
  c.ratio = (float)ratio;
  c.inv_area = (float)(1.0 / ((2.0 * ratio * th) * (2.0 * th)));
  c.W = w;
  c.H = h;
  return c;
}

static inline F3 camera_ray_dir(const Cam& c, float rx, float ry) {
  F3 v = normalize(F3{c.ax * rx + c.bx, c.ay * ry + c.by, 1.0f});
  return (v.x * c.right + v.y * c.up) + v.z * c.dir;
}

// ---------------------------------------------------------------------------
// Differentiable scene access.  In AD mode every vertex position read inside
// one path becomes a tape leaf (cached per path); material parameters too.
// ---------------------------------------------------------------------------
template <class S>
struct Access;

template <>
struct Access<float> {
  const Scene& s;
  explicit Access(const Scene& sc) : s(sc) {}
  F3 pos(uint32_t v) { return P(s, v); }
  float rough(int m) { return s.mats[m].roughness; }
  F3 albedo(int m) { return s.mats[m].albedo; }
  F3 spec(int m) { return s.mats[m].spec; }
  F3 texel(int ti, size_t k) { const float* p = &s.tex[ti][3 * k]; return F3{p[0], p[1], p[2]}; }
};

template <>
struct Access<Var> {
  const Scene& s;
  std::unordered_map<uint32_t, V3<Var>> vcache;
  std::unordered_map<int, Var> rcache;
  std::unordered_map<int, V3<Var>> acache, scache;
  std::unordered_map<uint64_t, V3<Var>> tcache;  // (texture << 32 | texel) -> leaves
  explicit Access(const Scene& sc) : s(sc) {}
  V3<Var> texel(int ti, size_t k) {
    const uint64_t key = ((uint64_t)ti << 32) | (uint64_t)k;
    auto it = tcache.find(key);
    if (it != tcache.end()) return it->second;
    const float* p = &s.tex[ti][3 * k];
    V3<Var> r{leaf(p[0]), leaf(p[1]), leaf(p[2])};
    tcache.emplace(key, r);
    return r;
  }
  V3<Var> pos(uint32_t v) {
    auto it = vcache.find(v);
    if (it != vcache.end()) return it->second;
    F3 p = P(s, v);
    V3<Var> r{leaf(p.x), leaf(p.y), leaf(p.z)};
    vcache.emplace(v, r);
    return r;
  }
  Var rough(int m) {
    auto it = rcache.find(m);
    if (it != rcache.end()) return it->second;
    Var r = leaf(s.mats[m].roughness);
    rcache.emplace(m, r);
    return r;
  }
  V3<Var> albedo(int m) {
    auto it = acache.find(m);
    if (it != acache.end()) return it->second;
    F3 a = s.mats[m].albedo;
    V3<Var> r{leaf(a.x), leaf(a.y), leaf(a.z)};
    acache.emplace(m, r);
    return r;
  }
  V3<Var> spec(int m) {
    auto it = scache.find(m);
    if (it != scache.end()) return it->second;
    F3 a = s.mats[m].spec;
    V3<Var> r{leaf(a.x), leaf(a.y), leaf(a.z)};
    scache.emplace(m, r);
    return r;
  }
};

// Twin access: one parameter shifted by `delta` in the perturbed lane.
// kind 0: vertex coordinate (index, comp); 1: roughness (index = material);
// 2: albedo (index, comp); 3: specular (index, comp); 4: texel channel
// (index = texture, comp = 3 * texel + channel).
template <>
struct Access<Pair> {
  const Scene& s;
  int kind, index, comp;
  double delta;
  Access(const Scene& sc, int k, int i, int c, double d) : s(sc), kind(k), index(i), comp(c), delta(d) {}
  V3<Pair> pos(uint32_t v) {
    F3 p = P(s, v);
    V3<Pair> r{Pair(p.x), Pair(p.y), Pair(p.z)};
    if (kind == 0 && (int)v == index) {
      Pair* c = comp == 0 ? &r.x : (comp == 1 ? &r.y : &r.z);
      c->s += delta;
    }
    return r;
  }
  Pair rough(int m) {
    Pair r(s.mats[m].roughness);
    if (kind == 1 && m == index) r.s += delta;
    return r;
  }
  V3<Pair> vec(F3 a, int m, int k) {
    V3<Pair> r{Pair(a.x), Pair(a.y), Pair(a.z)};
    if (kind == k && m == index) {
      Pair* c = comp == 0 ? &r.x : (comp == 1 ? &r.y : &r.z);
      c->s += delta;
    }
    return r;
  }
  V3<Pair> albedo(int m) { return vec(s.mats[m].albedo, m, 2); }
  V3<Pair> spec(int m) { return vec(s.mats[m].spec, m, 3); }
  V3<Pair> texel(int ti, size_t k) {
    const float* p = &s.tex[ti][3 * k];
    V3<Pair> r{Pair(p[0]), Pair(p[1]), Pair(p[2])};
    if (kind == 4 && ti == index && (size_t)comp / 3 == k) {
      Pair* c = comp % 3 == 0 ? &r.x : (comp % 3 == 1 ? &r.y : &r.z);
      c->s += delta;
    }
    return r;
  }
};

template <class S>
struct Its {
  V3<S> p, ng, ns, t, b;
  float tu, tv;
  int tri, mat, emitter;
};

// face normal and area from differentiable positions (Primitive.cpp:299-321)
template <class S>
static void face_nA(Access<S>& A, int t, V3<S>* n, S* area) {
  const Scene& s = A.s;
  V3<S> p0 = A.pos(s.idx[3 * t]), p1 = A.pos(s.idx[3 * t + 1]), p2 = A.pos(s.idx[3 * t + 2]);
  V3<S> c = cross(p1 - p0, p2 - p0);
  S l = length(c);
  *n = c / l;
  *area = l * S(0.5f);
}

// vertex normal (Primitive.cpp:323-340), differentiable
template <class S>
static V3<S> vertex_normal(Access<S>& A, uint32_t v) {
  const Scene& s = A.s;
  V3<S> a{S(0.0f), S(0.0f), S(0.0f)};
  for (int t : s.vfaces[v]) {
    V3<S> n;
    S ar;
    face_nA(A, t, &n, &ar);
    a = a + n;
  }
  S c((float)s.vfaces[v].size());
  return normalize(V3<S>{a.x / c, a.y / c, a.z / c});
}

// CoordinateSystem (Utils.cpp:118-140)
template <class S>
static void coordinate_system(V3<S> n, V3<S>* t, V3<S>* b) {
  bool cond = fabsf(val(n.x)) > fabsf(val(n.y));
  S q = cond ? n.x * n.x + n.z * n.z : n.y * n.y + n.z * n.z;
  S qm = val(q) > 1e-3f ? q : S(1e-3f);
  S il = S(1.0f) / sqrt_(qm);
  *t = cond ? V3<S>{-n.z * il, S(0.0f), n.x * il} : V3<S>{S(0.0f), n.z * il, -n.y * il};
  *b = cross(n, *t);
}

// Scene::PostIntersect (Scene.cpp:668-730)
template <class S>
static Its<S> post_intersect(Access<S>& A, int tri, float u, float v) {
  const Scene& s = A.s;
  Its<S> it;
  uint32_t i0 = s.idx[3 * tri], i1 = s.idx[3 * tri + 1], i2 = s.idx[3 * tri + 2];
  float w = 1.0f - u - v;
  it.p = (S(w) * A.pos(i0) + S(u) * A.pos(i1)) + S(v) * A.pos(i2);
  S ar;
  face_nA(A, tri, &it.ng, &ar);
  if (s.tri_flags[tri] & 1) {
    it.ns = (S(w) * vertex_normal(A, i0) + S(u) * vertex_normal(A, i1)) + S(v) * vertex_normal(A, i2);
  } else {
    it.ns = it.ng;
  }
  if (!s.uv.empty()) {
    uint32_t t0 = s.uv_idx[3 * tri], t1 = s.uv_idx[3 * tri + 1], t2 = s.uv_idx[3 * tri + 2];
    it.tu = w * s.uv[2 * t0] + u * s.uv[2 * t1] + v * s.uv[2 * t2];
    it.tv = w * s.uv[2 * t0 + 1] + u * s.uv[2 * t1 + 1] + v * s.uv[2 * t2 + 1];
  } else {
    it.tu = it.tv = 0.0f;
  }
  coordinate_system(it.ns, &it.t, &it.b);
  it.tri = tri;
  it.mat = s.tri_mat[tri];
  it.emitter = s.tri_emitter[tri];
  return it;
}

template <class S>
static S area_of(Access<S>& A, int tri) {
  V3<S> n;
  S ar;
  face_nA(A, tri, &n, &ar);
  return ar;
}

template <class S> static V3<S> to_local(const Its<S>& it, V3<S> v) { return {dot(v, it.t), dot(v, it.b), dot(v, it.ns)}; }
template <class S> static V3<S> to_world(const Its<S>& it, V3<S> l) { return (l.x * it.t + l.y * it.b) + l.z * it.ns; }

// ImageTexture::Eval (BSDF.cpp:72-98): nearest texel index at
// floor(frac(u) * (w-1)), floor(frac(v) * (h-1)).
static size_t tex_index(const Scene& s, int ti, float u, float v) {
  int w = s.tex_w[ti], h = s.tex_h[ti];
  float fu = u - floorf(u), fv = v - floorf(v);
  int x = (int)(fu * (float)(w - 1)), y = (int)(fv * (float)(h - 1));
  return (size_t)y * w + x;
}

// The texel is a differentiable leaf (BSDF::DiffTexture, BSDF.cpp:350-356):
// its gradient is the scatter of the lookup's adjoint.
template <class S>
static V3<S> albedo_at(Access<S>& A, const Its<S>& it) {
  const Mat& m = A.s.mats[it.mat];
  if (m.tex >= 0) return A.texel(m.tex, tex_index(A.s, m.tex, it.tu, it.tv));
  return A.albedo(it.mat);
}

// ---- sampling (Utils.cpp:70-115) ----
static F3 cosine_hemisphere(float s0, float s1) {
  float r1 = 2.0f * s0 - 1.0f, r2 = 2.0f * s1 - 1.0f;
  bool cond = (r1 * r1) > (r2 * r2);
  float r = cond ? r1 : r2;
  bool zero = (r1 == 0.0f) && (r2 == 0.0f);
  float phi = 0.0f;
  if (!zero) phi = cond ? 0.78539816339744830962f * (r2 / r1) : 1.57079632679489661923f - (r1 / r2) * 0.78539816339744830962f;
  float rr = zero ? 0.0f : r;
  float sp, cp;
  f_sincos(phi, &sp, &cp);
  float x = rr * cp, y = rr * sp;
  float z = sqrtf(std::max(1e-4f, 1.0f - x * x - y * y));
  return F3{x, y, z};
}
static void disk_concentric(float s0, float s1, float* ox, float* oy) {
  float x = 2.0f * s0 - 1.0f, y = 2.0f * s1 - 1.0f;
  bool zero = (x == 0.0f && y == 0.0f);
  bool q13 = fabsf(x) < fabsf(y);
  float r = q13 ? y : x, rp = q13 ? x : y;
  float phi = 0.0f;
  if (!zero) {
    float p_ = (0.25f * PI) * rp / r;
    phi = q13 ? 0.5f * PI - p_ : p_;
  }
  float sp, cp;
  f_sincos(phi, &sp, &cp);
  *ox = r * cp;
  *oy = r * sp;
}

// ---- GGX (BSDF.cpp:232-302), all float: only used for sampling and pdfs ----
static float tan2f_(F3 v) { float c2 = v.z * v.z; return std::max(1.0f - c2, 0.0f) / c2; }
static float sinth(F3 v) { return sqrtf(std::max(1.0f - v.z * v.z, 0.0f)); }
static float cosphi(F3 v) { float st = sinth(v); return st == 0.0f ? 1.0f : std::min(std::max(v.x / st, -1.0f), 1.0f); }
static float sinphi(F3 v) { float st = sinth(v); return st == 0.0f ? 0.0f : std::min(std::max(v.y / st, -1.0f), 1.0f); }
static float GGX_D(F3 wh, float a) {
  float t2 = tan2f_(wh);
  bool ok = t2 < 3.402823466e38f;
  float c2 = wh.z * wh.z;
  float e = 1.0f / (a * a) * t2;
  float root = a * c2 * (1.0f + e);
  float D = INV_PI / (root * root);
  return ok ? D : 0.0f;
}
static float Smith_G(F3 v, F3 wh, float a) {
  float t2 = tan2f_(wh);
  bool ok = t2 < 3.402823466e38f;
  float root = (a * a) * t2;
  float G = 2.0f / (1.0f + sqrtf(1.0f + root * root));
  G = ok ? G : 0.0f;
  float voh = dot(v, wh);
  G = t2 == 0.0f ? 1.0f : G;
  G = (voh * v.z) <= 0.0f ? 0.0f : G;
  return G;
}
static float VNDF_pdf(F3 v, F3 wh, float a) {
  return GGX_D(wh, a) * Smith_G(v, wh, a) * fabsf(dot(v, wh)) / fabsf(v.z);
}
static F3 VNDF_sample(F3 v, float a, float s0, float s1) {
  F3 Vv = normalize(F3{v.x * a, v.y * a, v.z * 1.0f});
  float sp = sinphi(Vv), cp = cosphi(Vv), ct = Vv.z;
  float px, py;
  disk_concentric(s0, s1, &px, &py);
  float sh = 0.5f * (1.0f + ct);
  float x = px;
  float aa = sqrtf(std::max(1.0f - x * x, 0.0f));
  float y = (1.0f - sh) * aa + sh * py;
  float z = sqrtf(std::max(1.0f - (px * px + py * py), 0.0f));
  float st = sqrtf(std::max(1.0f - ct * ct, 0.0f));
  float norm = 1.0f / (st * y + ct * z);
  float slx = (ct * y - st * z) * norm, sly = x * norm;
  float sx = (cp * slx - sp * sly) * a, sy = (sp * slx + cp * sly) * a;
  return normalize(F3{-sx, -sy, 1.0f});
}
static F3 reflect_(F3 i, F3 n) { return i + (2.0f * dot(-i, n)) * n; }

// ---- GGX and the conductor Fresnel term, templated for AD (BSDF.cpp:232-323) ----
template <class S> static S tan2_t(V3<S> v) {
  S c2 = v.z * v.z;
  S om = S(1.0f) - c2;
  if (!(val(om) > 0.0f)) om = S(0.0f);
  return om / c2;
}
template <class S> static S safe_sqrt_t(S x) { return val(x) > 0.0f ? sqrt_(x) : S(0.0f); }
template <class S> static S GGX_D_t(V3<S> wh, S a) {
  S t2 = tan2_t(wh);
  bool ok = val(t2) < 3.402823466e38f;
  S c2 = wh.z * wh.z;
  S e = S(1.0f) / (a * a) * t2;
  S root = a * c2 * (S(1.0f) + e);
  S D = S(INV_PI) / (root * root);
  return ok ? D : S(0.0f);
}
template <class S> static S Smith_G_t(V3<S> v, V3<S> wh, S a) {
  S t2 = tan2_t(wh);
  if (!(val(t2) < 3.402823466e38f)) return S(0.0f);
  if (val(t2) == 0.0f) return (val(dot(v, wh)) * val(v.z)) <= 0.0f ? S(0.0f) : S(1.0f);
  if ((val(dot(v, wh)) * val(v.z)) <= 0.0f) return S(0.0f);
  S root = (a * a) * t2;
  return S(2.0f) / (S(1.0f) + sqrt_(S(1.0f) + root * root));
}
// BSDF::FresnelConductor (BSDF.cpp:304-323), one channel.
template <class S> static S fresnel_conductor(S ci, S eta, S k) {
  S ci2 = ci * ci;
  S si2 = S(1.0f) - ci2;
  S si4 = si2 * si2;
  S eta2 = eta * eta, k2 = k * k;
  S temp1 = eta2 - k2 - si2;
  S a2pb2 = safe_sqrt_t(temp1 * temp1 + k2 * eta2 * S(4.0f));
  S a = safe_sqrt_t((a2pb2 + temp1) * S(0.5f));
  S term1 = a2pb2 + ci2;
  S term2 = a * (S(2.0f) * ci);
  S Rs2 = (term1 - term2) / (term1 + term2);
  S term3 = a2pb2 * ci2 + si4, term4 = term2 * si2;
  S Rp2 = Rs2 * (term3 - term4) / (term3 + term4);
  return S(0.5f) * (Rp2 + Rs2);
}

// Microfacet::EvalInner's three factors (Microfacet.cpp:49-63), as the
// reference writes them: GGX NDF with alpha = r^2 in its
// alpha / (cos^2 (alpha^2 - 1) + 1) form, Schlick's Fresnel with the
// spherical-Gaussian exponent 2^{vh (-5.55473 vh - 6.8316)}, and the
// Smith-Schlick G1 with k = (r + 1)^2 / 8.
template <class S> static S mf_ggx(S nh, S alpha) {
  S tmp = alpha / (nh * nh * (alpha * alpha - S(1.0f)) + S(1.0f));
  return tmp * tmp * S(INV_PI);
}
template <class S> static S mf_fresnel_weight(S vh) {
  S coeff = vh * (vh * S(-5.55473f) - S(6.8316f));
  return exp2_(coeff);
}
template <class S> static S mf_smith_g1(S ndot, S k) { return ndot / (ndot * (S(1.0f) - k) + k); }

// BSDF value (BSDF::Eval without correction), templated for AD.
template <class S>
static V3<S> bsdf_eval(Access<S>& A, const Its<S>& it, V3<S> wo_w, V3<S> wi_w) {
  const Mat& m = A.s.mats[it.mat];
  V3<S> wo = to_local(it, wo_w), wi = to_local(it, wi_w);
  if (m.type == 0) {
    bool refl = val(wo.z) > 0.0f && (val(wo.z) * val(wi.z)) > 0.0f;
    S inner = S(INV_PI) * S(refl ? 1.0f : 0.0f);
    return albedo_at(A, it) * inner;
  }
  if (m.type == 2) {
    // RoughConductor::EvalInner (Roughconductor.cpp:37-57); the albedo and
    // spec slots hold eta and k.
    V3<S> zero{S(0.0f), S(0.0f), S(0.0f)};
    if (!(val(wo.z) > 0.0f && val(wi.z) > 0.0f)) return zero;
    V3<S> wh = normalize(wo + wi);
    S r = A.rough(it.mat);
    S D = GGX_D_t(wh, r);
    if (val(D) == 0.0f) return zero;
    S G = Smith_G_t(wo, wh, r) * Smith_G_t(wi, wh, r);
    S coso = dot(wo, wh);
    V3<S> eta = A.albedo(it.mat), k = A.spec(it.mat);
    V3<S> Fr{fresnel_conductor(coso, eta.x, k.x), fresnel_conductor(coso, eta.y, k.y),
             fresnel_conductor(coso, eta.z, k.z)};
    S den = S(4.0f) * abs_(wi.z) * abs_(wo.z);
    return Fr * D * G / den;
  }
  // Microfacet::EvalInner (Microfacet.cpp:37-73)
  S nv = wo.z, nl = wi.z;
  bool active = val(nv) > 0.0f && val(nl) > 0.0f;
  V3<S> kd = albedo_at(A, it);
  V3<S> diff = kd * S(INV_PI);
  V3<S> wh = normalize(wo + wi);
  S nh = wh.z, vh = dot(wo, wh);
  S r = A.rough(it.mat);
  S alpha = r * r;
  S k = (r + S(1.0f)) * (r + S(1.0f)) * S(0.125f);
  S ggx = mf_ggx(nh, alpha);
  S e = mf_fresnel_weight(vh);
  V3<S> F0 = A.spec(it.mat);
  V3<S> one{S(1.0f), S(1.0f), S(1.0f)};
  V3<S> fres = F0 + (one - F0) * e;
  S g1 = mf_smith_g1(nv, k);
  S g2 = mf_smith_g1(nl, k);
  S G = g1 * g2;
  V3<S> num = (ggx * G) * fres;
  S den = S(4.0f) * abs_(nv) * abs_(nl);
  V3<S> spec = num / (den + S(1e-6f));
  V3<S> v = diff + spec;
  if (!active) return V3<S>{S(0.0f), S(0.0f), S(0.0f)};
  return v;
}

static float bsdf_pdf(const Scene& s, const Its<float>& it, F3 wo_w, F3 wi_w) {
  const Mat& m = s.mats[it.mat];
  F3 wo = to_local(it, wo_w), wi = to_local(it, wi_w);
  if (m.type == 0) {
    bool refl = wo.z > 0.0f && (wo.z * wi.z) > 0.0f;
    return fabsf(wi.z) / PI * (refl ? 1.0f : 0.0f);
  }
  F3 wh = normalize(wo + wi);
  float mp = VNDF_pdf(wo, wh, m.roughness);
  float pdf = fabsf(mp * (1.0f / (4.0f * dot(wi, wh))));
  bool refl = wo.z > 0.0f && mp != 0.0f && wi.z > 0.0f;
  return refl ? pdf : 0.0f;
}

static bool dir_valid(const Scene& s, const Its<float>& it, F3 w) {
  if (s.mats[it.mat].type == 0) return dot(it.ns, w) > EDGE_EPSILON_ && dot(it.ng, w) > EDGE_EPSILON_;
  return dot(it.ns, w) * dot(it.ng, w) > 0.0f;
}

static F3 bsdf_sample(const Scene& s, const Its<float>& it, F3 wo_w, float s0, float s1, float* pdf) {
  const Mat& m = s.mats[it.mat];
  F3 wo = to_local(it, wo_w);
  if (m.type == 0) {
    F3 wi = cosine_hemisphere(s0, s1);
    bool refl = wo.z > 0.0f && (wo.z * wi.z) > 0.0f;
    *pdf = fabsf(wi.z) / PI * (refl ? 1.0f : 0.0f);
    return to_world(it, wi);
  }
  bool refl = wo.z > 0.0f;
  F3 wh = VNDF_sample(wo, m.roughness, s0, s1);
  float mp = VNDF_pdf(wo, wh, m.roughness);
  F3 wi = reflect_(-wo, wh);
  refl = refl && mp != 0.0f && wi.z > 0.0f;
  float p = fabsf(mp * (1.0f / (4.0f * dot(wi, wh))));
  *pdf = refl ? p : 0.0f;
  return to_world(it, wi);
}

// Distribution1D sampling: first i with cdf[i] >= u, clamped.
static int cdf_find(const std::vector<float>& cdf, float u) {
  int n = (int)cdf.size();
  int i = (int)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
  return std::min(i, n - 1);
}

template <class S>
struct LightS {
  V3<S> p, n;
  S J;
  float pdf;
  int tri, emitter;
};

// AreaLight::Sample (Light.cpp:146-175)
template <class S>
static LightS<S> light_sample(Access<S>& A, float u0, float u1) {
  const Scene& s = A.s;
  LightS<S> L;
  int n = (int)s.light_tri.size();
  int i = cdf_find(s.light_cdf, u0);
  float pmf = s.light_f[i] / (s.light_int * (float)n);
  float lo = i > 0 ? s.light_cdf[i - 1] : 0.0f;
  float reuse = (u0 - lo) / (s.light_cdf[i] - lo);
  float t = sqrtf(reuse);
  float bu = 1.0f - t, bv = t * u1, bw = 1.0f - bu - bv;
  int tri = s.light_tri[i];
  L.p = (S(bw) * A.pos(s.idx[3 * tri]) + S(bu) * A.pos(s.idx[3 * tri + 1])) + S(bv) * A.pos(s.idx[3 * tri + 2]);
  S ar;
  face_nA(A, tri, &L.n, &ar);
  L.J = ar / detach(ar);
  L.pdf = pmf / s.area[tri];
  L.tri = tri;
  L.emitter = s.tri_emitter[tri];
  return L;
}

static float light_pdf(const Scene& s, int tri) {
  int n = (int)s.light_tri.size();
  float pmf = s.light_f[s.emit_tri[tri]] / (s.light_int * (float)n);
  return pmf / s.area[tri];
}

static F3 emitted(const Scene& s, int emitter, F3 ng, F3 w) {
  float on = dot(w, ng) > 0.0f ? 1.0f : 0.0f;
  const float* I = &s.emit_rad[3 * emitter];
  return F3{I[0], I[1], I[2]} * on;
}

// ---------------------------------------------------------------------------
// Bounce loop (PathTracer.cpp:55-73 + EvalRadianceDirect,
// Integrator.cpp:178-303) from a surface point `it` reached along `raydir`
// with throughput `beta`.  Random numbers of bounce b come from blocks
// block0 + 2b (light) and block0 + 2b + 1 (BSDF) of (seed, domain, index).
// per_bounce (optional): float value added by each executed bounce.
// Returns the number of executed bounces.
// ---------------------------------------------------------------------------
template <class S>
static int bounces(Access<S>& A, Its<S> it, V3<S> beta, V3<S> raydir, uint32_t seed, uint32_t domain,
                   uint64_t index, uint32_t block0, int B, V3<S>& contrib, F3* per_bounce) {
  const Scene& s = A.s;
  const V3<S> zero{S(0.0f), S(0.0f), S(0.0f)};
  int executed = 0;
  for (int b = 0; b < B; ++b) {
    R4 rl = rnd4(seed, domain, index, block0 + 2u * (uint32_t)b);
    R4 rb = rnd4(seed, domain, index, block0 + 2u * (uint32_t)b + 1u);
    ++executed;
    V3<S> ret = zero;
    V3<S> wo = -raydir;
    Its<float> itf;
    itf.p = val3(it.p); itf.ng = val3(it.ng); itf.ns = val3(it.ns); itf.t = val3(it.t); itf.b = val3(it.b);
    itf.tu = it.tu; itf.tv = it.tv; itf.tri = it.tri; itf.mat = it.mat; itf.emitter = it.emitter;
    // NEE (Integrator.cpp:183-229)
    LightS<S> ls = light_sample(A, rl.a, rl.b);
    {
      V3<S> wi = ls.p - it.p;
      S d2 = len2(wi);
      S dist = sqrt_(d2);
      wi = wi / dist;
      S dsh = dot(it.ns, wi), dgeo = dot(it.ng, wi);
      bool stable = val(dist) > SHADOW_EPSILON_ && fabsf(val(dsh)) > EDGE_EPSILON_;
      bool noleak = (val(dsh) * val(dgeo)) > 0.0f;
      bool visible = false;
      if (stable && noleak)
        visible = intersect(s, itf.p, val3(wi), SHADOW_EPSILON_, val(dist) - SHADOW_EPSILON_, true, false).tri < 0;
      S valid((stable && noleak && visible) ? 1.0f : 0.0f);
      S Gl = abs_(dot(ls.n, -wi)) / d2;
      V3<S> f = bsdf_eval(A, it, wo, wi) * abs_(dsh);
      V3<S> Le = cast3<S>(emitted(s, ls.emitter, val3(ls.n), val3(-wi)));
      V3<S> v = valid * beta * f * ls.J * Gl * Le / S(ls.pdf);
      float pb = bsdf_pdf(s, itf, val3(wo), val3(wi)) * val(Gl);
      float mis = (ls.pdf * ls.pdf) / (ls.pdf * ls.pdf + pb * pb);
      v = v * S(mis);
      ret = ret + v;
    }
    // BSDF sampling (Integrator.cpp:231-294)
    float pdfb;
    F3 wib = bsdf_sample(s, itf, val3(wo), rb.a, rb.b, &pdfb);
    Hit h2 = intersect(s, itf.p, wib, SHADOW_EPSILON_, 1e32f, false, false);
    bool ok = h2.tri >= 0 && pdfb > 0.0f && dir_valid(s, itf, wib);
    if (!ok) {
      if (finite3(ret)) {
        contrib = contrib + ret;
        if (per_bounce) per_bounce[b] = val3(ret);
      }
      break;
    }
    Its<S> it2 = post_intersect(A, h2.tri, h2.u, h2.v);
    V3<S> dd = it2.p - it.p;
    S dd2 = len2(dd);
    V3<S> nd = dd / sqrt_(dd2);
    S G2 = abs_(dot(it2.ng, -nd)) / dd2;
    V3<S> f2 = bsdf_eval(A, it, wo, nd) * abs_(dot(it.ns, nd));
    float ppdf = pdfb * val(G2);
    S ar2 = area_of(A, h2.tri);
    S J2 = ar2 / detach(ar2);
    V3<S> beta2 = beta * f2 * G2 * J2 / S(ppdf);
    if (!(val(G2) > 0.0f)) beta2 = zero;
    if (it2.emitter >= 0) {
      V3<S> v2 = beta2 * cast3<S>(emitted(s, it2.emitter, val3(it2.ng), val3(-nd)));
      float pn = light_pdf(s, h2.tri);
      float mis2 = (ppdf * ppdf) / (ppdf * ppdf + pn * pn);
      v2 = v2 * S(mis2);
      ret = ret + v2;
    }
    if (finite3(ret)) {
      contrib = contrib + ret;
      if (per_bounce) per_bounce[b] = val3(ret);
    }
    it = it2;
    beta = beta2;
    raydir = nd;
  }
  return executed;
}

// ---------------------------------------------------------------------------
// One path (PathTracer::Radiance), scalar or AD.
// ---------------------------------------------------------------------------
template <class S>
static V3<S> radiance(Access<S>& A, const Cam& C, uint32_t seed, int B, uint64_t path) {
  const Scene& s = A.s;
  const V3<S> zero{S(0.0f), S(0.0f), S(0.0f)};
  const uint64_t WH = (uint64_t)C.W * C.H;
  const uint32_t pix = (uint32_t)(path % WH);
  const uint32_t px = pix % C.W, py = pix / C.W;
  R4 rc = rnd4(seed, kDomInterior, path, 0u);
  F3 d0 = camera_ray_dir(C, (float)px + rc.a, (float)py + rc.b);
  Hit h = intersect(s, C.eye, d0, SHADOW_EPSILON_, 1e32f, false, false);
  if (h.tri < 0) return zero;
  Its<S> it = post_intersect(A, h.tri, h.u, h.v);
  V3<S> eye = cast3<S>(C.eye);
  // Camera::EvalFilter (box): G / detach(G), G = |cos_x| / (cos_y^3 d^2); times J
  S beta0;
  {
    V3<S> cd = it.p - eye;
    S dist = length(cd);
    cd = cd / dist;
    S cosy = dot(cd, cast3<S>(C.dir));
    S cosx = dot(-cd, it.ng);
    S G = abs_(cosx) / (cosy * cosy * cosy * (dist * dist));
    S ar = area_of(A, h.tri);
    beta0 = (G / detach(G)) * (ar / detach(ar));
  }
  V3<S> beta{beta0, beta0, beta0};
  V3<S> raydir = normalize(it.p - eye);
  V3<S> contrib = zero;
  if (it.emitter >= 0) {
    V3<S> e0 = beta * cast3<S>(emitted(s, it.emitter, val3(it.ng), val3(-raydir)));
    if (finite3(e0)) contrib = contrib + e0;
  }
  bounces(A, it, beta, raydir, seed, kDomInterior, path, 1u, B, contrib, nullptr);
  return contrib;
}

// ===========================================================================
// Boundary (edge) terms.  Each sample contributes only a gradient:
//   term = detach(K) * xDotN,  xDotN = n . u2,
//   u2 = barycentric point (detached vertices) of RayIntersectAD(dir, org, tri)
// (BoundaryPrimary.cpp:131-163, BoundaryPixel.cpp:108-148, Edge.cpp:265-287,
// BoundaryIndirect.cpp:117-131).  K (radiance, importance, baseVal / pdf) is
// evaluated in float; xDotN on the tape.
// ===========================================================================
constexpr uint32_t kDomPrimaryEdge = 0x22u;
constexpr uint32_t kDomPixelEdge = 0x33u;
constexpr uint32_t kDomDirectEdge = 0x44u;
constexpr uint32_t kDomIndirectEdge = 0x55u;
constexpr uint32_t kSensorBlock0 = 64u;

struct EdgeRec {
  uint32_t v0, v1, v2;
  int f0, f1;
  F3 p0, e1, n0, n1, p2;
  bool boundary;
};

// ObjMesh::processEdges over the flattened index buffer (meshes do not share
// vertices, and later meshes have larger vertex ids, so the global key order
// equals the per-mesh order concatenated).
static std::vector<EdgeRec> all_edges(const Scene& s) {
  std::map<std::pair<uint32_t, uint32_t>, std::vector<int>> m;
  for (int t = 0; t < s.T; ++t)
    for (int j = 0; j < 3; ++j) {
      uint32_t a = s.idx[3 * t + j], b = s.idx[3 * t + (j + 1) % 3], c = s.idx[3 * t + (j + 2) % 3];
      auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
      auto it = m.find(key);
      if (it == m.end()) it = m.emplace(key, std::vector<int>{(int)c}).first;
      it->second.push_back(t);
    }
  std::vector<EdgeRec> out;
  for (auto& kv : m) {
    EdgeRec e;
    e.v0 = kv.first.first;
    e.v1 = kv.first.second;
    e.v2 = (uint32_t)kv.second[0];
    e.f0 = kv.second[1];
    e.f1 = kv.second.size() == 3 ? kv.second[2] : -1;
    e.boundary = e.f1 < 0;
    e.p0 = P(s, e.v0);
    e.e1 = P(s, e.v1) - e.p0;
    e.p2 = P(s, e.v2);
    e.n0 = s.fn[e.f0];
    e.n1 = e.boundary ? F3{0, 0, 0} : s.fn[e.f1];
    out.push_back(e);
  }
  return out;
}

// ConstructSecEdgeList (Edge.cpp:42-88)
static std::vector<EdgeRec> sec_edges(const Scene& s) {
  std::vector<EdgeRec> out;
  for (const EdgeRec& e : all_edges(s)) {
    bool keep = e.boundary;
    if (!e.boundary) {
      bool v0 = dot(e.n0, e.n1) < 1.0f - 1e-6f;
      bool v1 = dot(e.n1, normalize(e.p2 - e.p0)) < -EDGE_EPSILON_;
      keep = v0 && v1;
    }
    if (keep) out.push_back(e);
  }
  return out;
}

// ConstructPrimEdgeList (Edge.cpp:313-348)
static std::vector<EdgeRec> prim_edges(const Scene& s, const Cam& c) {
  std::vector<EdgeRec> out;
  for (const EdgeRec& e : all_edges(s)) {
    F3 d = c.eye - e.p0;
    bool ff0 = dot(d, e.n0) > 0.0f, ff1 = dot(d, e.n1) > 0.0f;
    if (e.boundary || ff0 != ff1) out.push_back(e);
  }
  return out;
}

// Distribution1D over edge lengths (Distribution.cpp:45-58, Distribution.cu:37-48)
// and the guide's reductions over leaves use a blocked summation order (the
// order the device uses; the reference leaves it to its scan and reduction
// library, CUB): chunks
// of kSumChunk consecutive values are summed sequentially, then the chunk
// sums sequentially; an inclusive prefix is the sequential prefix of the
// preceding chunk sums plus the sequential prefix within the chunk.  Up to
// kSumChunk values this is the plain sequential order.
constexpr size_t kSumChunk = 1024;
static std::vector<float> chunk_offsets(const std::vector<float>& v, float* total) {
  const size_t nc = (v.size() + kSumChunk - 1) / kSumChunk;
  std::vector<float> off(nc);
  float run = 0.0f;
  for (size_t c = 0; c < nc; ++c) {
    float acc = 0.0f;
    for (size_t i = c * kSumChunk; i < std::min(v.size(), (c + 1) * kSumChunk); ++i) acc = acc + v[i];
    off[c] = run;
    run = run + acc;
  }
  *total = run;
  return off;
}
static float blocked_sum(const std::vector<float>& v) {
  float total;
  chunk_offsets(v, &total);
  return total;
}
struct Blocked {};

struct Dist1 {
  std::vector<float> f, cdf;
  float integral = 0.0f;
  explicit Dist1(const std::vector<float>& fn) : f(fn), cdf(fn.size()) {
    float inv = 1.0f / (float)fn.size(), run = 0.0f;
    for (size_t i = 0; i < fn.size(); ++i) { run += fn[i]; cdf[i] = run * inv; }
    integral = cdf.back();
    if (integral > 0.0f)
      for (float& c : cdf) c = c / integral;
  }
  // The same distribution with the CDF's prefix in the blocked order.
  Dist1(const std::vector<float>& fn, Blocked) : f(fn), cdf(fn.size()) {
    const float inv = 1.0f / (float)fn.size();
    float total;
    std::vector<float> off = chunk_offsets(fn, &total);
    for (size_t c = 0; c < off.size(); ++c) {
      float local = 0.0f;
      for (size_t i = c * kSumChunk; i < std::min(fn.size(), (c + 1) * kSumChunk); ++i) {
        local = local + fn[i];
        cdf[i] = (off[c] + local) * inv;
      }
    }
    integral = cdf.back();
    if (integral > 0.0f)
      for (float& c : cdf) c = c / integral;
  }
  int sample(float u, float* pmf) const {
    int i = cdf_find(cdf, u);
    *pmf = f[i] / (integral * (float)f.size());
    return i;
  }
  float reuse(float u, int i) const {
    float lo = i > 0 ? cdf[i - 1] : 0.0f;
    return (u - lo) / (cdf[i] - lo);
  }
};

static std::vector<float> edge_lengths(const std::vector<EdgeRec>& E) {
  std::vector<float> l(E.size());
  for (size_t i = 0; i < E.size(); ++i) l[i] = length(E[i].e1);
  return l;
}

// Camera::sampleDirect (Camera.cpp:367-384) with the in-front test.
struct Sensor {
  float qx, qy, value;
  bool valid;
  int pixel;
};
static Sensor sensor(const Cam& c, F3 p) {
  Sensor r;
  F3 d = p - c.eye;
  F3 l{dot(d, c.right), dot(d, c.up), dot(d, c.dir)};
  float sx = l.x / (l.z * c.ratio * c.tan_half);
  float sy = l.y / (l.z * c.tan_half);
  r.qx = (1.0f - sx) * 0.5f;
  r.qy = (1.0f - sy) * 0.5f;
  r.valid = l.z > 0.0f && r.qx >= 0.0f && r.qx < 1.0f && r.qy >= 0.0f && r.qy < 1.0f;
  r.pixel = (int)floorf(r.qy * (float)c.H) * c.W + (int)floorf(r.qx * (float)c.W);
  float d2 = len2(d);
  F3 dn = d / sqrtf(d2);
  float ct = dot(c.dir, dn);
  float ic = 1.0f / ct;
  r.value = 1.0f / d2 * ic * ic * ic * c.inv_area;
  return r;
}

// BSDF::Eval with the shading-normal correction (BSDF.cpp:172-195), float.
static F3 bsdf_eval_corr(const Scene& s, const Its<float>& it, F3 wo, F3 wi) {
  Access<float> A(s);
  F3 f = bsdf_eval(A, it, wo, wi);
  F3 lo = to_local(it, wo), li = to_local(it, wi);
  float corr = fabsf(lo.z * dot(it.ng, wi) / (li.z * dot(it.ng, wo)));
  return f * corr;
}

// EvalImportance (Integrator.cpp:75-176) for one sensor-side vertex: the
// importance contribution (pixel, value) and the next vertex.
struct ImpStep {
  bool has_contrib;
  int pixel;
  F3 value;
  bool next;
  Its<float> it;
  F3 dir, thr;
};
static ImpStep importance_step(const Scene& s, const Cam& c, const Its<float>& it, F3 rdir, F3 thr, R4 rnd) {
  ImpStep r;
  r.has_contrib = false;
  r.next = false;
  r.value = F3{0, 0, 0};
  F3 wo = -rdir;
  Sensor sd = sensor(c, it.p);
  if (sd.valid) {
    F3 cd = camera_ray_dir(c, sd.qx * (float)c.W, sd.qy * (float)c.H);
    float tmax = length(it.p - c.eye) - SHADOW_EPSILON_;
    bool vis = intersect(s, c.eye, cd, SHADOW_EPSILON_, tmax, true, false).tri < 0;
    if (vis) {
      F3 t = sd.value * thr;
      F3 bv{0, 0, 0};
      if (dir_valid(s, it, wo)) {
        F3 wi = normalize(c.eye - it.p);
        float wiok = dir_valid(s, it, wi) ? 1.0f : 0.0f;
        bv = bsdf_eval_corr(s, it, wo, wi) * fabsf(dot(it.ns, wi)) * wiok;
      }
      F3 v = t * bv;
      if (finite3(v)) {
        r.has_contrib = true;
        r.pixel = sd.pixel;
        r.value = v;
      }
    }
  }
  if (dir_valid(s, it, wo)) {
    float pdf;
    F3 wi = bsdf_sample(s, it, wo, rnd.a, rnd.b, &pdf);
    F3 bv = bsdf_eval_corr(s, it, wo, wi) * fabsf(dot(it.ns, wi));
    Hit h = intersect(s, it.p, wi, SHADOW_EPSILON_, 1e32f, false, false);
    if (h.tri >= 0 && pdf > 0.0f && dir_valid(s, it, wi)) {
      Access<float> A(s);
      r.next = true;
      r.it = post_intersect(A, h.tri, h.u, h.v);
      r.dir = wi;
      r.thr = thr * bv / pdf;
    }
  }
  return r;
}

// RayIntersectAD (Utils.cpp:288-297): Moller-Trumbore u, v.
template <class S>
static void ray_intersect_ad(V3<S> dir, V3<S> org, V3<S> p0, V3<S> e1, V3<S> e2, S* u, S* v) {
  V3<S> h = cross(dir, e2);
  S f = S(1.0f) / dot(e1, h);
  V3<S> s0 = org - p0;
  *u = f * dot(s0, h);
  V3<S> q = cross(s0, e1);
  *v = f * dot(dir, q);
}

// xDotN = n . ((1-u-v) P0d + u P1d + v P2d) for the triangle `tri`,
// (u, v) from the ray (org -> through) against the differentiable triangle.
template <class S>
static S x_dot_n(Access<S>& A, int tri, V3<S> org, V3<S> through, F3 n) {
  const Scene& s = A.s;
  V3<S> P0 = A.pos(s.idx[3 * tri]), P1 = A.pos(s.idx[3 * tri + 1]), P2 = A.pos(s.idx[3 * tri + 2]);
  S u, v;
  ray_intersect_ad(normalize(through - org), org, P0, P1 - P0, P2 - P0, &u, &v);
  S w = S(1.0f) - u - v;
  V3<S> u2 = (w * detach3(P0) + u * detach3(P1)) + v * detach3(P2);
  return dot(cast3<S>(n), u2);
}

// baseVal of the primary / pixel boundary (BoundaryPrimary.cpp:117-129)
static float base_val_camera(F3 itp, F3 ng, F3 raydir, F3 p_edge, F3 eye, F3 edge, F3 edge2, F3* n_out,
                             bool revert) {
  float dist = length(itp - eye);
  float dist1 = length(p_edge - eye);
  float cos2 = fabsf(dot(ng, -raydir));
  F3 e = cross(edge, raydir);
  float sinphi = length(e);
  F3 proj = normalize(cross(e, ng));
  float sinphi2 = length(cross(raydir, proj));
  F3 n = normalize(cross(ng, proj));
  bool sign0 = dot(e, edge2) > 0.0f, sign1 = dot(e, n) > 0.0f;
  float bv = (dist / dist1) * (sinphi / sinphi2) * cos2;
  bv = bv * (sinphi > EPSILON_ ? 1.0f : 0.0f) * (sinphi2 > EPSILON_ ? 1.0f : 0.0f);
  float sg = (sign0 == sign1) ? 1.0f : -1.0f;
  bv = bv * (revert ? -sg : sg);
  *n_out = n;
  return bv;
}

// baseVal of the secondary terms (Edge.cpp:245-257, BoundaryIndirect.cpp:100-112)
static float base_val_secondary(F3 its_p, F3 far_p, F3 far_n, F3 dir_to_near, F3 p_edge, F3 edge, F3 edge2,
                                F3* n_out) {
  // dir_to_near: ray direction from the edge point toward the near-side
  // surface (its_p); the far side is the emitter side.
  float dist = length(far_p - its_p);
  float cos2 = fabsf(dot(far_n, dir_to_near));
  F3 e = cross(edge, -dir_to_near);
  float sinphi = length(e);
  F3 proj = normalize(cross(e, far_n));
  float sinphi2 = length(cross(-dir_to_near, proj));
  float itsT = length(its_p - p_edge);
  F3 n = normalize(cross(far_n, proj));
  bool sign0 = dot(e, edge2) > 0.0f, sign1 = dot(e, n) > 0.0f;
  float bv = (itsT / dist) * (sinphi / sinphi2) * cos2;
  bv = bv * (sinphi > EPSILON_ ? 1.0f : 0.0f) * (sinphi2 > EPSILON_ ? 1.0f : 0.0f);
  bv = bv * ((sign0 == sign1) ? 1.0f : -1.0f);
  *n_out = n;
  return bv;
}

// Radiance along a camera boundary ray: EvalRadianceEmitted + bounces, all
// detached, throughput = sensor importance (GenerateBoundaryRays).
static F3 boundary_radiance(const Scene& s, const Its<float>& it, F3 raydir, float sensor_val, uint32_t seed,
                            uint32_t domain, uint64_t index, int B) {
  Access<float> A(s);
  F3 beta{sensor_val, sensor_val, sensor_val};
  F3 rad{0, 0, 0};
  if (it.emitter >= 0) {
    F3 e = beta * emitted(s, it.emitter, it.ng, -raydir);
    if (finite3(e)) rad = rad + e;
  }
  bounces(A, it, beta, raydir, seed, domain, index, 1u, B, rad, nullptr);
  return rad;
}

// One boundary sample: weight seeds on the tape.  Returns the scalar weight
// w = sum over splats of <dLdI[pixel], K> / spp and the xDotN Var (or sets
// *valid = false).
struct BSample {
  bool valid = false;
  int tri = -1;        // triangle of the RayIntersectAD
  F3 org, through, n;  // values (org / through may be differentiable, see kind)
  double w = 0.0;
  // differentiable org / through: edge (v0, v1, t) and / or material point (tri, u, v)
  bool through_edge = false;
  uint32_t ev0 = 0, ev1 = 0;
  float et = 0.0f;
  bool org_material = false;
  int otri = -1;
  float ou = 0.0f, ov = 0.0f;
};

static BSample primary_sample(const Scene& s, const Cam& c, const std::vector<EdgeRec>& E, const Dist1& D,
                              uint32_t seed, int B, int spp, uint64_t i, const float* dLdI) {
  BSample r;
  R4 rnd = rnd4(seed, kDomPrimaryEdge, i, 0u);
  float pmf;
  int ei = D.sample(rnd.a, &pmf);
  const EdgeRec& e = E[ei];
  float pdf = pmf / D.f[ei];
  F3 p = e.p0 + rnd.b * e.e1;
  F3 edge = normalize(e.e1);
  F3 edge2 = normalize(e.p2 - e.p0);
  F3 dir = normalize(p - c.eye);
  Hit h = intersect(s, p, dir, SHADOW_EPSILON_, 1e32f, false, false);
  if (h.tri < 0) return r;
  F3 hitP = p + h.t * dir;
  Sensor sd = sensor(c, hitP);
  if (!sd.valid) return r;
  F3 rd = camera_ray_dir(c, sd.qx * (float)c.W, sd.qy * (float)c.H);
  float tmax = length(p - c.eye) - SHADOW_EPSILON_;
  if (intersect(s, c.eye, rd, SHADOW_EPSILON_, tmax, true, false).tri >= 0) return r;
  Access<float> A(s);
  Its<float> it = post_intersect(A, h.tri, h.u, h.v);
  F3 n;
  float bv = base_val_camera(it.p, it.ng, rd, p, c.eye, edge, edge2, &n, false);
  F3 rad = boundary_radiance(s, it, rd, sd.value, seed, kDomPrimaryEdge, i, B);
  F3 K = rad * bv / pdf;
  if (!(length(K) > 0.0f) || !finite3(K)) return r;
  const uint64_t WH = (uint64_t)c.W * c.H;
  r.w = ((double)dLdI[sd.pixel] * K.x + (double)dLdI[WH + sd.pixel] * K.y + (double)dLdI[2 * WH + sd.pixel] * K.z) /
        (double)spp;
  r.valid = true;
  r.tri = h.tri;
  r.org = c.eye;
  r.through = p;
  r.n = n;
  r.through_edge = true;
  r.ev0 = e.v0;
  r.ev1 = e.v1;
  r.et = rnd.b;
  return r;
}

static BSample pixel_sample(const Scene& s, const Cam& c, uint32_t seed, int B, int spp, uint64_t i, uint64_t N,
                            const float* dLdI) {
  BSample r;
  const uint64_t per_anti = N / 4;
  const uint64_t j = i % per_anti;
  const uint64_t anti = i / per_anti;
  float rr = rnd4(seed, kDomPixelEdge, j, 0u).a;
  float rv;
  if (anti == 0) rv = rr;
  else if (anti == 1) rv = rr < 0.5f ? rr + 0.5f : rr - 0.5f;
  else if (anti == 2) rv = rr < 0.75f ? 0.75f - rr : 1.75f - rr;
  else rv = rr < 0.25f ? 0.25f - rr : 1.25f - rr;
  float r4 = rv * 4.0f;
  int side = (int)floorf(r4);
  float ox, oy;
  if (side == 0) { ox = 0.0f; oy = r4; }
  else if (side == 1) { ox = r4 - 1.0f; oy = 1.0f; }
  else if (side == 2) { ox = 1.0f; oy = 3.0f - r4; }
  else { ox = 4.0f - r4; oy = 0.0f; }
  const uint64_t WH = (uint64_t)c.W * c.H;
  const int pix = (int)(i % WH);
  float rx = (float)(pix % c.W) + ox, ry = (float)(pix / c.W) + oy;
  // RasterToWorld: camera point (ax rx + bx, ay ry + by, 1)
  F3 p0 = ((c.eye + (c.ax * rx + c.bx) * c.right) + (c.ay * ry + c.by) * c.up) + c.dir;
  // edge / visible directions of the pixel side (raster vectors through the
  // linear part of RasterToWorld)
  float plx = 2.0f * c.ratio * c.tan_half, ply = 2.0f * c.tan_half;
  F3 vdir = normalize(c.ay * c.up), hdir = normalize(c.ax * c.right);
  F3 edge = (side == 0 || side == 2) ? vdir : hdir;
  F3 edge2 = side == 0 ? hdir : (side == 1 ? -vdir : (side == 2 ? -hdir : vdir));
  float pdf = (side == 0 || side == 2) ? 0.25f / (ply * (float)c.W) : 0.25f / (plx * (float)c.H);
  F3 dir = normalize(p0 - c.eye);
  Hit h = intersect(s, p0, dir, SHADOW_EPSILON_, 1e32f, false, false);
  if (h.tri < 0) return r;
  F3 hitP = p0 + h.t * dir;
  Sensor sd = sensor(c, hitP);
  F3 rd = camera_ray_dir(c, sd.qx * (float)c.W, sd.qy * (float)c.H);
  Access<float> A(s);
  Its<float> it = post_intersect(A, h.tri, h.u, h.v);
  F3 n;
  float bv = base_val_camera(it.p, it.ng, rd, p0, c.eye, edge, edge2, &n, true);
  F3 rad = boundary_radiance(s, it, rd, sd.value, seed, kDomPixelEdge, j, B);
  F3 K = rad * bv / pdf;
  if (!finite3(K)) return r;
  r.w = ((double)dLdI[pix] * K.x + (double)dLdI[WH + pix] * K.y + (double)dLdI[2 * WH + pix] * K.z) / (double)spp;
  r.valid = r.w != 0.0;
  r.tri = h.tri;
  r.org = c.eye;
  r.through = p0;
  r.n = n;
  return r;
}

// ---------------------------------------------------------------------------
// Guiding of the direct boundary term (AQ_distrb.h:83-197, NEE.cu:70-89):
// one leaf per secondary edge, x in [cdf_{i-1}, cdf_i] of the edge-length
// CDF, y, z in [0, 1]; weight = mean |baseVal| over gspp samples (guiding
// mode of Sample/EvalBoundarySegmentDirect: only the hit / same-point tests
// apply), normalized, mixed eps + (1 - eps) w; leaves sampled with
// probability proportional to weight * volume.
// ---------------------------------------------------------------------------
constexpr uint32_t kDomGuide = 0x66u;

static void parallel_for(uint64_t n, int threads, const std::function<void(uint64_t, uint64_t, int)>& f);

struct Guide {
  std::vector<float> x0, x1, poly;
  std::unique_ptr<Dist1> dist;
  float fsum = 0.0f;
};

static float guide_value(const Scene& s, const std::vector<EdgeRec>& E, const Dist1& D, float x, float y, float z) {
  float pmf;
  int ei = D.sample(x, &pmf);
  const EdgeRec& e = E[ei];
  float t = D.reuse(x, ei);
  F3 p = e.p0 + t * e.e1;
  F3 edge = normalize(e.e1);
  F3 edge2 = normalize(e.p2 - e.p0);
  Access<float> A(s);
  LightS<float> ls = light_sample(A, y, z);
  F3 rdir = normalize(p - ls.p);
  Hit h0 = intersect(s, p, -rdir, SHADOW_EPSILON_, 1e32f, false, false);
  Hit h = intersect(s, p, rdir, SHADOW_EPSILON_, 1e32f, false, false);
  if (h0.tri < 0 || h.tri < 0) return 0.0f;
  F3 hitP = p + h0.t * (-rdir);
  if (!(length(hitP - ls.p) < SHADOW_EPSILON_)) return 0.0f;
  Its<float> it = post_intersect(A, h.tri, h.u, h.v);
  F3 n;
  float bv = base_val_secondary(it.p, ls.p, ls.n, rdir, p, edge, edge2, &n);
  float v = fabsf(bv);
  return std::isfinite(v) ? v : 0.0f;
}

static std::unique_ptr<Guide> build_guide(const Scene& s, const std::vector<EdgeRec>& E, const Dist1& D,
                                          uint32_t seed, int gspp, float eps, int threads) {
  std::unique_ptr<Guide> g(new Guide());
  const size_t n = E.size();
  g->x0.resize(n);
  g->x1.resize(n);
  g->poly.assign(n, 0.0f);
  for (size_t i = 0; i < n; ++i) {
    g->x0[i] = i > 0 ? D.cdf[i - 1] : 0.0f;
    g->x1[i] = D.cdf[i];
  }
  parallel_for(n, threads, [&](uint64_t a, uint64_t b, int) {
    for (uint64_t i = a; i < b; ++i) {
      float acc = 0.0f;
      for (int k = 0; k < gspp; ++k) {
        R4 r = rnd4(seed, kDomGuide, i * (uint64_t)gspp + (uint64_t)k, 0u);
        float x = g->x0[i] + r.a * (g->x1[i] - g->x0[i]);
        acc = acc + guide_value(s, E, D, x, r.b, r.c) / (float)gspp;
      }
      g->poly[i] = acc;
    }
  });
  const float sum = blocked_sum(g->poly);
  std::vector<float> f(n);
  for (size_t i = 0; i < n; ++i) {
    float w = sum > 0.0f ? g->poly[i] / sum : 0.0f;
    g->poly[i] = eps + (1.0f - eps) * w;
    f[i] = g->poly[i] * ((g->x1[i] - g->x0[i]) * 1.0f * 1.0f);
  }
  g->fsum = blocked_sum(f);
  g->dist.reset(new Dist1(f, Blocked{}));
  return g;
}

static BSample direct_sample(const Scene& s, const Cam& c, const std::vector<EdgeRec>& E, const Dist1& D,
                             const Guide* G, uint32_t seed, int B, int spp, uint64_t i, const float* dLdI) {
  BSample r;
  R4 rnd = rnd4(seed, kDomDirectEdge, i, 0u);
  float pdf_b = 1.0f;
  if (G) {
    // AdaptiveQuadratureDistribution::sample (AQ_distrb.h:182-197)
    float lp;
    int leaf = G->dist->sample(rnd.d, &lp);
    rnd.a = G->x0[leaf] + rnd.a * (G->x1[leaf] - G->x0[leaf]);
    pdf_b = G->poly[leaf] / G->fsum;
  }
  float pmf;
  int ei = D.sample(rnd.a, &pmf);
  const EdgeRec& e = E[ei];
  float pdf_e = pmf / D.f[ei];
  float t = D.reuse(rnd.a, ei);
  F3 p = e.p0 + t * e.e1;
  F3 edge = normalize(e.e1);
  F3 edge2 = normalize(e.p2 - e.p0);
  Access<float> A(s);
  LightS<float> ls = light_sample(A, rnd.b, rnd.c);
  F3 ev = ls.p - p;
  float d2 = len2(ev), dist = sqrtf(d2);
  F3 en = ev / dist;
  float cosT = dot(ls.n, -en);
  float c0 = dot(en, e.n0), c1 = dot(en, e.n1);
  bool valid0 = fabsf(c0) > EDGE_EPSILON_;
  bool valid1 = (c0 > EDGE_EPSILON_ && c1 < -EDGE_EPSILON_) || (c0 < -EDGE_EPSILON_ && c1 > EDGE_EPSILON_);
  bool right = (e.boundary && valid0) || (!e.boundary && valid1);
  if (!(right && cosT > EPSILON_ && dist > SHADOW_EPSILON_)) return r;
  float pdf = pdf_e * ls.pdf * d2 / cosT * pdf_b;
  // EvalBoundarySegmentDirect (Edge.cpp:201-309)
  F3 rdir = normalize(p - ls.p);
  Hit h0 = intersect(s, p, -rdir, SHADOW_EPSILON_, 1e32f, false, false);
  Hit h = intersect(s, p, rdir, SHADOW_EPSILON_, 1e32f, false, false);
  if (h0.tri < 0 || h.tri < 0) return r;
  F3 hitP = p + h0.t * (-rdir);
  if (!(length(hitP - ls.p) < SHADOW_EPSILON_)) return r;
  F3 Le{0, 0, 0};
  if (s.tri_emitter[h0.tri] >= 0) Le = emitted(s, s.tri_emitter[h0.tri], s.fn[h0.tri], rdir);
  Its<float> it = post_intersect(A, h.tri, h.u, h.v);
  F3 n;
  float bv = base_val_secondary(it.p, ls.p, ls.n, rdir, p, edge, edge2, &n);
  F3 K0 = Le * bv / pdf;
  // sensor side (EvalImportance for B bounces, throughput 1)
  const uint64_t WH = (uint64_t)c.W * c.H;
  double w = 0.0;
  Its<float> cur = it;
  F3 cdir = rdir, thr{1, 1, 1};
  for (int b = 0; b < B; ++b) {
    ImpStep st = importance_step(s, c, cur, cdir, thr, rnd4(seed, kDomDirectEdge, i, kSensorBlock0 + (uint32_t)b));
    if (st.has_contrib) {
      F3 v = K0 * st.value;
      if (length(v) > 0.0f)
        w += (double)dLdI[st.pixel] * v.x + (double)dLdI[WH + st.pixel] * v.y + (double)dLdI[2 * WH + st.pixel] * v.z;
    }
    if (!st.next) break;
    cur = st.it;
    cdir = st.dir;
    thr = st.thr;
  }
  if (w == 0.0 || !std::isfinite(w)) return r;
  r.valid = true;
  r.w = w / (double)spp;
  r.tri = h0.tri;
  r.org = it.p;
  r.through = p;
  r.n = n;
  r.through_edge = true;
  r.ev0 = e.v0;
  r.ev1 = e.v1;
  r.et = t;
  r.org_material = true;
  r.otri = h.tri;
  r.ou = h.u;
  r.ov = h.v;
  return r;
}

static BSample indirect_sample(const Scene& s, const Cam& c, const std::vector<EdgeRec>& E, const Dist1& D,
                               uint32_t seed, int B, int spp, uint64_t i, const float* dLdI) {
  BSample r;
  R4 rnd = rnd4(seed, kDomIndirectEdge, i, 0u);
  float pmf;
  int ei = D.sample(rnd.a, &pmf);
  const EdgeRec& e = E[ei];
  float pdf_e = pmf / D.f[ei];
  float t = D.reuse(rnd.a, ei);
  F3 p = e.p0 + t * e.e1;
  F3 edge = normalize(e.e1);
  F3 edge2 = normalize(e.p2 - e.p0);
  // SampleBoundarySegmentIndirect (Edge.cpp:146-199)
  F3 dir;
  float pdf_d;
  if (e.boundary) {
    F3 nrm = cross(e.n0, edge);
    Its<float> fr;
    coordinate_system(nrm, &fr.t, &fr.b);
    float z = std::min(1.0f - 2.0f * rnd.b, 1.0f - EDGE_EPSILON_);
    float rr = sqrtf(std::max(0.0f, 1.0f - z * z));
    float sp, cp;
    f_sincos(2.0f * PI * rnd.c, &sp, &cp);
    dir = ((rr * cp) * fr.t + (rr * sp) * fr.b) + z * nrm;
    pdf_d = 0.25f / PI;
  } else {
    float cosd = dot(e.n0, e.n1);
    float phi0 = acosf(std::min(1.0f, std::max(-1.0f, cosd)));
    pdf_d = 0.25f / phi0;
    F3 z = normalize(e.n0 + e.n1);
    F3 y = normalize(cross(e.n0, z));
    F3 x = cross(y, z);
    float phi = (rnd.b - 0.5f) * phi0;
    phi = std::min(std::max(phi, -0.5f * phi0 + EDGE_EPSILON_), 0.5f * phi0 - EDGE_EPSILON_);
    float sp, cp;
    f_sincos(phi, &sp, &cp);
    F3 x1 = x * cp + z * sp;
    bool m0 = rnd.c > 0.5f;
    float bb = 4.0f * rnd.c - (m0 ? 3.0f : 1.0f);
    float aa = sqrtf(1.0f - bb * bb) * (m0 ? -1.0f : 1.0f);
    dir = x1 * aa + y * bb;
  }
  float pdf = pdf_e * pdf_d;
  Hit he = intersect(s, p, -dir, SHADOW_EPSILON_, 1e32f, false, false);
  Hit hs = intersect(s, p, dir, SHADOW_EPSILON_, 1e32f, false, false);
  if (he.tri < 0 || hs.tri < 0) return r;
  Access<float> A(s);
  Its<float> ite = post_intersect(A, he.tri, he.u, he.v);
  Its<float> its = post_intersect(A, hs.tri, hs.u, hs.v);
  F3 n;
  float bv = base_val_secondary(its.p, ite.p, ite.ng, dir, p, edge, edge2, &n);
  float v0 = bv / pdf;
  // emitter side: cumulative radiance after 1..B-1 bounces from ite
  std::vector<F3> per(std::max(B - 1, 1), F3{0, 0, 0});
  F3 acc{0, 0, 0};
  if (B > 1) bounces(A, ite, F3{1, 1, 1}, -dir, seed, kDomIndirectEdge, i, 1u, B - 1, acc, per.data());
  std::vector<F3> cum(std::max(B - 1, 1));
  F3 run{0, 0, 0};
  for (int k = 0; k < B - 1; ++k) { run = run + per[k]; cum[k] = run; }
  // sensor side
  const uint64_t WH = (uint64_t)c.W * c.H;
  double w = 0.0;
  Its<float> cur = its;
  F3 cdir = dir, thr{1, 1, 1};
  for (int b = 0; b < B - 1; ++b) {
    ImpStep st = importance_step(s, c, cur, cdir, thr, rnd4(seed, kDomIndirectEdge, i, kSensorBlock0 + (uint32_t)b));
    int ri = B - b - 2;
    if (st.has_contrib && ri >= 0 && ri < B - 1) {
      F3 v = (v0 * cum[ri]) * st.value;
      if (length(v) > 0.0f)
        w += (double)dLdI[st.pixel] * v.x + (double)dLdI[WH + st.pixel] * v.y + (double)dLdI[2 * WH + st.pixel] * v.z;
    }
    if (!st.next) break;
    cur = st.it;
    cdir = st.dir;
    thr = st.thr;
  }
  if (w == 0.0 || !std::isfinite(w)) return r;
  r.valid = true;
  r.w = w / (double)spp;
  r.tri = he.tri;
  r.org = its.p;
  r.through = p;
  r.n = n;
  r.through_edge = true;
  r.ev0 = e.v0;
  r.ev1 = e.v1;
  r.et = t;
  r.org_material = true;
  r.otri = hs.tri;
  r.ou = hs.u;
  r.ov = hs.v;
  return r;
}

// Backpropagate one boundary sample's weight through xDotN into dpos.
static void boundary_backprop(const Scene& s, const BSample& b, Tape& tape, double* dpos) {
  tape.clear();
  Access<Var> A(s);
  V3<Var> through = cast3<Var>(b.through);
  if (b.through_edge) {
    V3<Var> q0 = A.pos(b.ev0), q1 = A.pos(b.ev1);
    through = q0 + Var(b.et) * (q1 - q0);
  }
  V3<Var> org = cast3<Var>(b.org);
  if (b.org_material) {
    float w = 1.0f - b.ou - b.ov;
    org = (Var(w) * A.pos(s.idx[3 * b.otri]) + Var(b.ou) * A.pos(s.idx[3 * b.otri + 1])) +
          Var(b.ov) * A.pos(s.idx[3 * b.otri + 2]);
  }
  Var x = x_dot_n(A, b.tri, org, through, b.n);
  if (x.i < 0) return;
  size_t nn = tape.nodes.size();
  tape.adj.assign(nn, 0.0);
  tape.adj[x.i] = b.w;
  for (size_t k = nn; k-- > 0;) {
    double a = tape.adj[k];
    if (a == 0.0) continue;
    const Node& nd = tape.nodes[k];
    if (nd.a >= 0) tape.adj[nd.a] += a * nd.da;
    if (nd.b >= 0) tape.adj[nd.b] += a * nd.db;
  }
  for (auto& kv : A.vcache) {
    double* d = &dpos[3 * (size_t)kv.first];
    d[0] += tape.adj[kv.second.x.i];
    d[1] += tape.adj[kv.second.y.i];
    d[2] += tape.adj[kv.second.z.i];
  }
}

static void parallel_for(uint64_t n, int threads, const std::function<void(uint64_t, uint64_t, int)>& f) {
  if (threads <= 1 || n < 64) {
    f(0, n, 0);
    return;
  }
  std::atomic<uint64_t> next(0);
  const uint64_t chunk = 256;
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; ++t)
    ts.emplace_back([&, t] {
      for (;;) {
        uint64_t a = next.fetch_add(chunk);
        if (a >= n) break;
        f(a, std::min(n, a + chunk), t);
      }
    });
  for (auto& t : ts) t.join();
}

}  // namespace oracle

// ---------------------------------------------------------------------------
// C entry points (ctypes)
// ---------------------------------------------------------------------------
using namespace oracle;

extern "C" {

void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  U4 r = philox(U4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

void oracle_sincos(const float* x, int n, float* s, float* c) {
  for (int i = 0; i < n; ++i) f_sincos(x[i], &s[i], &c[i]);
}
void oracle_exp2(const float* x, int n, float* y) {
  for (int i = 0; i < n; ++i) y[i] = f_exp2(x[i]);
}
// GGX_D at the half vector (sqrt(1 - z^2), 0, z) and roughness a.
void oracle_ggx_d(const float* z, int n, float a, float* out) {
  for (int i = 0; i < n; ++i) {
    F3 wh{sqrtf(std::max(0.0f, 1.0f - z[i] * z[i])), 0.0f, z[i]};
    out[i] = GGX_D_t<float>(wh, a);
  }
}
void oracle_fresnel_conductor(const float* ci, int n, float eta, float k, float* out) {
  for (int i = 0; i < n; ++i) out[i] = fresnel_conductor<float>(ci[i], eta, k);
}

// materials: [nmat][12] = type, tex, roughness, unused, albedo rgb, spec rgb, unused x2
void* oracle_scene_create(const float* pos, int nv, const uint32_t* idx, int nt, const int32_t* tri_mat,
                          const int32_t* tri_emitter, const int32_t* tri_flags, const float* emit_rad, int n_emit,
                          const float* mats, int nmat, const float* uv, int nuv, const uint32_t* uv_idx) {
  Scene* s = new Scene();
  s->V = nv;
  s->T = nt;
  s->pos.assign(pos, pos + 3 * (size_t)nv);
  s->idx.assign(idx, idx + 3 * (size_t)nt);
  s->tri_mat.assign(tri_mat, tri_mat + nt);
  s->tri_emitter.assign(tri_emitter, tri_emitter + nt);
  s->tri_flags.assign(tri_flags, tri_flags + nt);
  s->emit_rad.assign(emit_rad, emit_rad + 3 * (size_t)n_emit);
  for (int m = 0; m < nmat; ++m) {
    const float* q = mats + 12 * m;
    Mat mm;
    mm.type = (int)q[0];
    mm.tex = (int)q[1];
    mm.roughness = q[2];
    mm.albedo = F3{q[4], q[5], q[6]};
    mm.spec = F3{q[7], q[8], q[9]};
    s->mats.push_back(mm);
  }
  if (uv && nuv > 0) {
    s->uv.assign(uv, uv + 2 * (size_t)nuv);
    s->uv_idx.assign(uv_idx, uv_idx + 3 * (size_t)nt);
  }
  scene_derive(*s);
  scene_build_bvh(*s);
  return s;
}

void oracle_scene_add_texture(void* sp, const float* texels, int w, int h) {
  Scene* s = (Scene*)sp;
  s->tex.emplace_back(texels, texels + 3 * (size_t)w * h);
  s->tex_w.push_back(w);
  s->tex_h.push_back(h);
}

void oracle_scene_destroy(void* s) { delete (Scene*)s; }

void oracle_scene_derived(void* sp, float* fn, float* area, float* vn, float* light_cdf, float* light_f,
                          float* light_int) {
  Scene* s = (Scene*)sp;
  for (int t = 0; t < s->T; ++t) {
    if (fn) { fn[3 * t] = s->fn[t].x; fn[3 * t + 1] = s->fn[t].y; fn[3 * t + 2] = s->fn[t].z; }
    if (area) area[t] = s->area[t];
  }
  if (vn)
    for (int v = 0; v < s->V; ++v) { vn[3 * v] = s->vn[v].x; vn[3 * v + 1] = s->vn[v].y; vn[3 * v + 2] = s->vn[v].z; }
  for (size_t i = 0; i < s->light_tri.size(); ++i) {
    if (light_cdf) light_cdf[i] = s->light_cdf[i];
    if (light_f) light_f[i] = s->light_f[i];
  }
  if (light_int) *light_int = s->light_int;
}

// rays [n][8] (ox,oy,oz,tmin,dx,dy,dz,tmax)
void oracle_intersect(void* sp, const float* rays, int n, int brute, int any_hit, int32_t* tri, float* uvt,
                      int threads) {
  const Scene& s = *(Scene*)sp;
  parallel_for((uint64_t)n, threads, [&](uint64_t a, uint64_t b, int) {
    for (uint64_t i = a; i < b; ++i) {
      const float* r = rays + 8 * i;
      Hit h = intersect(s, F3{r[0], r[1], r[2]}, F3{r[4], r[5], r[6]}, r[3], r[7], any_hit != 0, brute != 0);
      tri[i] = h.tri;
      uvt[3 * i] = h.u;
      uvt[3 * i + 1] = h.v;
      uvt[3 * i + 2] = h.t;
    }
  });
}

// camera: [pos3, target3, up3, fov] + w, h.  out: [12] = eye, right, up, dir
// as used by the kernels plus ax, bx, ay, by, tan_half, ratio, inv_area.
void oracle_camera(const float* cam, int w, int h, float* out) {
  Cam c = make_camera(cam, cam + 3, cam + 6, w, h, cam[9]);
  float o[19] = {c.eye.x, c.eye.y, c.eye.z, c.right.x, c.right.y, c.right.z, c.up.x, c.up.y, c.up.z,
                 c.dir.x, c.dir.y, c.dir.z, c.ax, c.bx, c.ay, c.by, c.tan_half, c.ratio, c.inv_area};
  std::memcpy(out, o, sizeof(o));
}

void oracle_camera_dirs(const float* cam, int w, int h, const float* raster, int n, float* dirs) {
  Cam c = make_camera(cam, cam + 3, cam + 6, w, h, cam[9]);
  for (int i = 0; i < n; ++i) {
    F3 d = camera_ray_dir(c, raster[2 * i], raster[2 * i + 1]);
    dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
  }
}

// Per-path radiance for paths [first, first + count): out [count][3].
void oracle_render_paths(void* sp, const float* cam, int w, int h, uint32_t seed, int max_bounces,
                         uint64_t first, uint64_t count, float* out, int threads) {
  const Scene& s = *(Scene*)sp;
  Cam c = make_camera(cam, cam + 3, cam + 6, w, h, cam[9]);
  parallel_for(count, threads, [&](uint64_t a, uint64_t b, int) {
    Access<float> A(s);
    for (uint64_t i = a; i < b; ++i) {
      F3 r = radiance(A, c, seed, max_bounces, first + i);
      out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
    }
  });
}

// Film [3][H*W] = (sum_s path(s, pix)) / spp, samples summed in order.
void oracle_render_film(void* sp, const float* cam, int w, int h, uint32_t seed, int max_bounces, int spp,
                        float* film, int threads) {
  const Scene& s = *(Scene*)sp;
  Cam c = make_camera(cam, cam + 3, cam + 6, w, h, cam[9]);
  const uint64_t WH = (uint64_t)w * h;
  parallel_for(WH, threads, [&](uint64_t a, uint64_t b, int) {
    Access<float> A(s);
    for (uint64_t pix = a; pix < b; ++pix) {
      float sx = 0, sy = 0, sz = 0;
      for (int k = 0; k < spp; ++k) {
        F3 r = radiance(A, c, seed, max_bounces, (uint64_t)k * WH + pix);
        sx += r.x; sy += r.y; sz += r.z;
      }
      float inv = 1.0f / (float)spp;
      film[pix] = sx * inv;
      film[WH + pix] = sy * inv;
      film[2 * WH + pix] = sz * inv;
    }
  });
}

// Central difference of one path's material-form estimator: out[0..2] =
// F(+delta), out[3..5] = F(-delta) (see Pair).
void oracle_path_twin(void* sp, const float* cam, int w, int h, uint32_t seed, int max_bounces, uint64_t path,
                      int kind, int index, int comp, double delta, double* out) {
  const Scene& s = *(Scene*)sp;
  Cam c = make_camera(cam, cam + 3, cam + 6, w, h, cam[9]);
  for (int sgn = 0; sgn < 2; ++sgn) {
    Access<Pair> A(s, kind, index, comp, sgn == 0 ? delta : -delta);
    V3<Pair> r = radiance(A, c, seed, max_bounces, path);
    out[3 * sgn] = r.x.s;
    out[3 * sgn + 1] = r.y.s;
    out[3 * sgn + 2] = r.z.s;
  }
}

// Boundary gradients: kind 0 primary edges, 1 pixel boundary, 2 secondary
// direct, 3 secondary indirect.  dpos [V][3] += d/dV of the boundary terms
// for samples [first, first + count) of spp * W * H.
void oracle_boundary_adj_gs(void* sp, const float* cam, int w, int h, uint32_t seed, int max_bounces, int kind,
                            int spp, const float* dLdI, uint64_t first, uint64_t count, double* dpos, int threads,
                            int g_direct, int g_spp, float g_eps, uint32_t g_seed) {
  const Scene& s = *(Scene*)sp;
  Cam c = make_camera(cam, cam + 3, cam + 6, w, h, cam[9]);
  std::vector<EdgeRec> E;
  if (kind == 0) E = prim_edges(s, c);
  else if (kind >= 2) E = sec_edges(s);
  if (kind != 1 && E.empty()) return;
  std::unique_ptr<Dist1> D;
  if (kind != 1) D.reset(new Dist1(edge_lengths(E), Blocked{}));
  std::unique_ptr<Guide> G;
  if (kind == 2 && g_direct) G = build_guide(s, E, *D, g_seed, g_spp, g_eps, threads);
  const uint64_t N = (uint64_t)spp * w * h;
  int nt = std::max(1, threads);
  std::vector<std::vector<double>> tp(nt, std::vector<double>(3 * (size_t)s.V, 0.0));
  parallel_for(count, threads, [&](uint64_t a, uint64_t b, int tid) {
    Tape tape;
    g_tape = &tape;
    for (uint64_t k = a; k < b; ++k) {
      uint64_t i = first + k;
      BSample bs;
      if (kind == 0) bs = primary_sample(s, c, E, *D, seed, max_bounces, spp, i, dLdI);
      else if (kind == 1) bs = pixel_sample(s, c, seed, max_bounces, spp, i, N, dLdI);
      else if (kind == 2) bs = direct_sample(s, c, E, *D, G.get(), seed, max_bounces, spp, i, dLdI);
      else bs = indirect_sample(s, c, E, *D, seed, max_bounces, spp, i, dLdI);
      if (bs.valid) boundary_backprop(s, bs, tape, tp[tid].data());
    }
    g_tape = nullptr;
  });
  for (int t = 0; t < nt; ++t)
    for (size_t k = 0; k < tp[t].size(); ++k) dpos[k] += tp[t][k];
}

// The guide built from the samples' own seed (the reference's behaviour).
void oracle_boundary_adj(void* sp, const float* cam, int w, int h, uint32_t seed, int max_bounces, int kind, int spp,
                         const float* dLdI, uint64_t first, uint64_t count, double* dpos, int threads, int g_direct,
                         int g_spp, float g_eps) {
  oracle_boundary_adj_gs(sp, cam, w, h, seed, max_bounces, kind, spp, dLdI, first, count, dpos, threads, g_direct,
                         g_spp, g_eps, seed);
}

// Guiding weights of the direct boundary term: poly [E_sec] (after the eps
// mix) and the leaf x ranges.
void oracle_guide(void* sp, uint32_t seed, int g_spp, float g_eps, float* x0, float* x1, float* poly, int* n_out) {
  const Scene& s = *(Scene*)sp;
  std::vector<EdgeRec> E = sec_edges(s);
  *n_out = (int)E.size();
  if (E.empty()) return;
  Dist1 D(edge_lengths(E), Blocked{});
  std::unique_ptr<Guide> G = build_guide(s, E, D, seed, g_spp, g_eps, 8);
  for (size_t i = 0; i < E.size(); ++i) {
    if (x0) x0[i] = G->x0[i];
    if (x1) x1[i] = G->x1[i];
    if (poly) poly[i] = G->poly[i];
  }
}

// Reverse mode: dpos [V][3] += d/dV sum_paths <dLdI[pix], path(pix)/spp>,
// dmat [M][8] (roughness, albedo rgb, spec rgb).  Accumulated in double.
void oracle_render_adj(void* sp, const float* cam, int w, int h, uint32_t seed, int max_bounces, int spp,
                       const float* dLdI, uint64_t first, uint64_t count, double* dpos, double* dmat,
                       int threads, double* dtex) {
  const Scene& s = *(Scene*)sp;
  Cam c = make_camera(cam, cam + 3, cam + 6, w, h, cam[9]);
  const uint64_t WH = (uint64_t)w * h;
  int nt = std::max(1, threads);
  std::vector<std::vector<double>> tp(nt, std::vector<double>(3 * (size_t)s.V, 0.0));
  std::vector<std::vector<double>> tm(nt, std::vector<double>(8 * s.mats.size(), 0.0));
  // texture gradients: textures concatenated in index order
  std::vector<size_t> toff(s.tex.size() + 1, 0);
  for (size_t t = 0; t < s.tex.size(); ++t) toff[t + 1] = toff[t] + s.tex[t].size();
  std::vector<std::vector<double>> tt(nt, std::vector<double>(dtex ? toff.back() : 0, 0.0));
  const float inv_spp = 1.0f / (float)spp;
  parallel_for(count, threads, [&](uint64_t a, uint64_t b, int tid) {
    Tape tape;
    g_tape = &tape;
    for (uint64_t i = a; i < b; ++i) {
      uint64_t path = first + i;
      uint32_t pix = (uint32_t)(path % WH);
      float g[3] = {dLdI[pix] * inv_spp, dLdI[WH + pix] * inv_spp, dLdI[2 * WH + pix] * inv_spp};
      if (g[0] == 0.0f && g[1] == 0.0f && g[2] == 0.0f) continue;
      tape.clear();
      Access<Var> A(s);
      V3<Var> r = radiance(A, c, seed, max_bounces, path);
      size_t nn = tape.nodes.size();
      tape.adj.assign(nn, 0.0);
      if (r.x.i >= 0) tape.adj[r.x.i] += g[0];
      if (r.y.i >= 0) tape.adj[r.y.i] += g[1];
      if (r.z.i >= 0) tape.adj[r.z.i] += g[2];
      for (size_t k = nn; k-- > 0;) {
        double a_ = tape.adj[k];
        if (a_ == 0.0) continue;
        const Node& nd = tape.nodes[k];
        if (nd.a >= 0) tape.adj[nd.a] += a_ * nd.da;
        if (nd.b >= 0) tape.adj[nd.b] += a_ * nd.db;
      }
      for (auto& kv : A.vcache) {
        const V3<Var>& lv = kv.second;
        double* d = &tp[tid][3 * (size_t)kv.first];
        d[0] += tape.adj[lv.x.i]; d[1] += tape.adj[lv.y.i]; d[2] += tape.adj[lv.z.i];
      }
      for (auto& kv : A.rcache) tm[tid][8 * kv.first] += tape.adj[kv.second.i];
      for (auto& kv : A.acache) {
        tm[tid][8 * kv.first + 1] += tape.adj[kv.second.x.i];
        tm[tid][8 * kv.first + 2] += tape.adj[kv.second.y.i];
        tm[tid][8 * kv.first + 3] += tape.adj[kv.second.z.i];
      }
      for (auto& kv : A.scache) {
        tm[tid][8 * kv.first + 4] += tape.adj[kv.second.x.i];
        tm[tid][8 * kv.first + 5] += tape.adj[kv.second.y.i];
        tm[tid][8 * kv.first + 6] += tape.adj[kv.second.z.i];
      }
      if (dtex)
        for (auto& kv : A.tcache) {
          double* d = &tt[tid][toff[kv.first >> 32] + 3 * (kv.first & 0xFFFFFFFFull)];
          d[0] += tape.adj[kv.second.x.i]; d[1] += tape.adj[kv.second.y.i]; d[2] += tape.adj[kv.second.z.i];
        }
    }
    g_tape = nullptr;
  });
  for (int t = 0; t < nt; ++t) {
    for (size_t k = 0; k < tp[t].size(); ++k) dpos[k] += tp[t][k];
    if (dmat)
      for (size_t k = 0; k < tm[t].size(); ++k) dmat[k] += tm[t][k];
    if (dtex)
      for (size_t k = 0; k < tt[t].size(); ++k) dtex[k] += tt[t][k];
  }
}

}  // extern "C"

// ---- BSDFs in a bare local frame (shading frame = world axes, normal +z):
// checks of the BSDF restatement against closed forms ----
static Scene one_material_scene(int type, float rough, const float* albedo, const float* spec) {
  Scene s;
  Mat m;
  m.type = type;
  m.tex = -1;
  m.roughness = rough;
  m.albedo = F3{albedo[0], albedo[1], albedo[2]};
  m.spec = F3{spec[0], spec[1], spec[2]};
  s.mats.push_back(m);
  return s;
}
static Its<float> local_its() {
  Its<float> it;
  it.p = F3{0.0f, 0.0f, 0.0f};
  it.ng = it.ns = F3{0.0f, 0.0f, 1.0f};
  it.t = F3{1.0f, 0.0f, 0.0f};
  it.b = F3{0.0f, 1.0f, 0.0f};
  it.tu = it.tv = 0.0f;
  it.tri = 0;
  it.mat = 0;
  it.emitter = -1;
  return it;
}

extern "C" {
// f [n][3] = BSDF::Eval (no correction), pdf [n] = BSDF::Pdf for local wo, wi.
void oracle_bsdf_local(int type, float rough, const float* albedo, const float* spec, const float* wo,
                       const float* wi, int n, float* f, float* pdf) {
  Scene s = one_material_scene(type, rough, albedo, spec);
  Access<float> A(s);
  Its<float> it = local_its();
  for (int i = 0; i < n; ++i) {
    F3 o{wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]}, w{wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]};
    F3 v = bsdf_eval(A, it, o, w);
    f[3 * i] = v.x;
    f[3 * i + 1] = v.y;
    f[3 * i + 2] = v.z;
    pdf[i] = bsdf_pdf(s, it, o, w);
  }
}
// BSDF::SampleOnly for local wo and sample pairs u [n][2]: wi [n][3], pdf [n].
void oracle_bsdf_sample_local(int type, float rough, const float* wo, const float* u, int n, float* wi,
                              float* pdf) {
  const float zero[3] = {0.0f, 0.0f, 0.0f};
  Scene s = one_material_scene(type, rough, zero, zero);
  Its<float> it = local_its();
  for (int i = 0; i < n; ++i) {
    F3 o{wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]};
    F3 w = bsdf_sample(s, it, o, u[2 * i], u[2 * i + 1], &pdf[i]);
    wi[3 * i] = w.x;
    wi[3 * i + 1] = w.y;
    wi[3 * i + 2] = w.z;
  }
}
// GGX_VNDF_Sample (half vector) and GGX_VNDF_Pdf (BSDF.cpp:262-302) with
// alpha = a, for local v and sample pairs u [n][2]: wh [n][3], pdf [n].
void oracle_vndf(const float* v, float a, const float* u, int n, float* wh, float* pdf) {
  for (int i = 0; i < n; ++i) {
    F3 vv{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    F3 h = VNDF_sample(vv, a, u[2 * i], u[2 * i + 1]);
    wh[3 * i] = h.x;
    wh[3 * i + 1] = h.y;
    wh[3 * i + 2] = h.z;
    pdf[i] = VNDF_pdf(vv, h, a);
  }
}
// BSDF::Smith_G(v, wh, a) (BSDF.cpp:243-256) for local v [n][3], wh [n][3].
void oracle_smith_g(const float* v, const float* wh, float a, int n, float* g) {
  for (int i = 0; i < n; ++i)
    g[i] = Smith_G(F3{v[3 * i], v[3 * i + 1], v[3 * i + 2]}, F3{wh[3 * i], wh[3 * i + 1], wh[3 * i + 2]}, a);
}
// The Microfacet factors: in [n][3] = (cos_h, vh, cos) and roughness r;
// out [n][3] = (GGX D with alpha = r^2, Fresnel weight 2^{...}, Smith-Schlick G1).
void oracle_mf_parts(const float* in, int n, float r, float* out) {
  const float alpha = r * r, k = (r + 1.0f) * (r + 1.0f) * 0.125f;
  for (int i = 0; i < n; ++i) {
    out[3 * i] = mf_ggx(in[3 * i], alpha);
    out[3 * i + 1] = mf_fresnel_weight(in[3 * i + 1]);
    out[3 * i + 2] = mf_smith_g1(in[3 * i + 2], k);
  }
}
}  // extern "C"
