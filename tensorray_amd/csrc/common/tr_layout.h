// Device data layout shared by the host builder and the HIP kernels.
//
// Everything a kernel reads for one scene lives in a handful of flat HBM
// arrays owned by the library (scene residency, needed by the serving loop):
//
//   pos      float[V][3]      vertex positions (world space), AoS
//   idx      uint32[T][3]     triangle vertex indices into pos
//   vnrm     float[V][3]      vertex normals (smooth-shaded meshes only)
//   shade    TriShade[T]      per-triangle face normal, area, material, emitter
//   tris     TriRecord[T]     triangle vertices in BVH leaf order (48 B)
//   nodes    Bvh4Node[N]      4-wide BVH, 128 B per node (one cache line)
//   mats     Material[M]
//   light_*  emitter triangles, their CDF and sampling weights
//
// The reference keeps the same information in SoA tensors
// (LibTensorRay/Renderer/Scene.cpp:46-124) and an opaque OptiX GAS
// (Scene.cpp:171-366); here the acceleration structure is explicit.
#pragma once
#include <stdint.h>
#include "tr_math.h"

namespace trc {

// Longest path the integrators record for replay (bounces after the primary hit).
constexpr int kMaxBounces = 8;

// 4-wide BVH node.  Child boxes are stored SoA so one node is read with
// eight 16-byte loads; `child` encodes inner nodes (index), leaves
// (kLeafBit | (count-1) << 27 | first) and empty slots (kEmpty).
struct alignas(128) Bvh4Node {
  float lox[4], loy[4], loz[4];
  float hix[4], hiy[4], hiz[4];
  uint32_t child[4];
  uint32_t pad[4];
};
static_assert(sizeof(Bvh4Node) == 128, "node must be one 128-byte line");

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafFirstMask = (1u << 27) - 1u;
constexpr int kMaxLeafTris = 8;

TR_HD bool child_is_leaf(uint32_t c) { return (c & kLeafBit) != 0u && c != kEmpty; }
TR_HD uint32_t leaf_first(uint32_t c) { return c & kLeafFirstMask; }
TR_HD uint32_t leaf_count(uint32_t c) { return ((c >> 27) & 0xFu) + 1u; }
TR_HD uint32_t make_leaf(uint32_t first, uint32_t count) {
  return kLeafBit | ((count - 1u) << 27) | first;
}

// SAH costs of a 4-wide node visit and of a triangle test, used by the
// builder and by the cost that decides when a refit tree is rebuilt.
// Forward time on the 20k-triangle metric scene against the traversal cost
// (leaves of at most 4 triangles): 0.25-0.35 fastest (every leaf then holds
// one triangle); 0.1 and 0.5 +0.5-1 %; 1 +4 %; 1.5 +6 %; 2 +11 % (also
// with leaves of up to 8); 3 with leaves of up to 8 +23 %.
constexpr double kSahTraverse = 0.25;
constexpr double kSahIntersect = 1.0;

// Triangle in leaf order for the watertight test: three world-space vertices,
// the original triangle id carried in v0.w.
struct alignas(16) TriRecord {
  float v0[4];
  float v1[4];
  float v2[4];
};

// Per-triangle shading data (Scene::PostIntersect inputs, Scene.cpp:668-730).
struct alignas(16) TriShade {
  float n[3];        // geometric (face) normal, normalize(cross(p1-p0, p2-p0))
  float area;        // 0.5 * |cross|
  int32_t material;  // BSDF index (mSbtIndexBuffer)
  int32_t emitter;   // -1 or emitter shape index k (mEmitterIDBuffer)
  int32_t emit_tri;  // index into the light arrays or -1 (mTriIdToEmitTriIdBuffer)
  uint32_t flags;    // bit0 smooth shading, bit1 vertices differentiable
};
constexpr uint32_t kTriSmooth = 1u;
constexpr uint32_t kTriDiff = 2u;

enum MaterialType : int32_t {
  kMatDiffuse = 0,
  kMatMicrofacet = 1,
  kMatRoughConductor = 2,
};

// Material record.  Diffuse: albedo = reflectance (Diffuse.cpp:37-85).
// Microfacet: albedo = diffuse reflectance kd, spec = F0, roughness
// (Microfacet.cpp:37-165).  Rough conductor: spec = specular reflectance,
// eta / k complex IOR (Roughconductor.cpp:38-127).
struct alignas(16) Material {
  int32_t type;
  int32_t tex;        // texture index for the albedo / kd slot, -1 = constant
  float roughness;
  float pad0;
  float albedo[3];
  float pad1;
  float spec[3];
  float pad2;
  float eta[3];
  float pad3;
  float k[3];
  float pad4;
};

// Mesh edge for the boundary terms (ObjMesh::processEdges semantics): end
// vertices v0 < v1, the opposite vertex v2 of face f0, the second face f1
// (-1 on a boundary edge).  Geometry is gathered from `pos` / face normals.
struct alignas(16) EdgeDev {
  uint32_t v0, v1, v2;
  int32_t f0, f1;
  uint32_t pad[3];
};

// Size and length integral of an edge list built on the device.
struct EdgeCount {
  int32_t n;
  float integral;
};

// A sampled edge list with its length distribution (Distribution1D over
// |p1 - p0|, Edge.cpp:90-106 / 350-367).  When `dyn` is set, the list was
// selected on the device and its size and integral are read from there
// (n and integral are then unused).
struct EdgeSet {
  const EdgeDev* e;
  const float* len;
  const float* cdf;
  int32_t n;
  float integral;
  const EdgeCount* dyn;
};

// Texture descriptor: FP32 RGB texels, row-major [h][w][3], nearest lookup
// (ImageTexture::Eval, BSDF.cpp:72-98).
struct TexDesc {
  const float* texels;
  int32_t w, h;
  int64_t grad_off;  // float offset of this texture in the texture-gradient buffer
};

// Parameters of a pinhole camera (Camera::Init/Resize, Camera.cpp:89-137 and
// EDXUtil/Graphics/BaseCamera.cpp:45-88) reduced to what the kernels need:
//   raster (rx, ry) -> camera-space point (ax*rx + bx, ay*ry + by, 1)
//   camera -> world rotation columns right/up/dir, origin = eye.
struct CameraDev {
  float eye[3];
  float right[3];
  float up[3];
  float dir[3];
  float ax, bx, ay, by;  // raster -> camera plane z = 1
  float tan_half;        // tan(fov/2)
  float ratio;           // width / height
  float inv_area;        // 1 / area of the image plane at z = 1
  int32_t width, height;
};

struct SceneDev {
  const float* pos;
  const uint32_t* idx;
  const float* vnrm;
  const float* uv;          // float[nuv][2] or null
  const uint32_t* uv_idx;   // uint32[T][3] or null
  const TriShade* shade;
  const TriRecord* tris;
  const Bvh4Node* nodes;
  const Material* mats;
  const TexDesc* texs;
  const int32_t* light_tri;   // emitter triangle -> global triangle id
  const float* light_cdf;     // normalized CDF over emitter triangles
  const float* light_f;       // luminance * area per emitter triangle
  const float* emit_rad;      // [n_emitters][3]
  EdgeSet sec;                // secondary edges (ConstructSecEdgeList, Edge.cpp:42-88)
  int32_t n_light;
  float light_integral;       // sum(f) / n_light (Distribution1D::mIntegralVal)
  int32_t n_tris;
  int32_t n_mats;
  int32_t stack_need;         // worst-case traversal stack depth of this BVH
  int32_t n_verts;
  // smooth-shaded faces incident to each vertex, in triangle order (CSR:
  // faces of vertex v are vf_list[vf_off[v] .. vf_off[v+1])), for the
  // adjoint of the vertex normals; null when no differentiable mesh is smooth
  const uint32_t* vf_off;
  const uint32_t* vf_list;
};

}  // namespace trc
