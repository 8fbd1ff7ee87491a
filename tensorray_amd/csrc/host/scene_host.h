// Host-side scene model: meshes, materials, emitters, their flattening into
// global buffers, derived per-triangle data, emitter sampling tables, edge
// lists and the SAH 4-wide BVH.
//
// Reference counterparts:
//   Scene::ExtractGeometry           LibTensorRay/Renderer/Scene.cpp:46-124
//   Scene::ConfigureLight            Scene.cpp:471-520
//   TriangleMesh::ComputeFaceNormalAndDistrib / ComputeVertexNormal
//                                    Primitive.cpp:299-340
//   ObjMesh::processEdges            EDXUtil/Graphics/ObjMesh.cpp:584-632
//   OptiX GAS build                  Scene.cpp:171-366 (replaced by bvh_build)
#pragma once
#include <stdint.h>
#include <string>
#include <vector>
#include "../common/tr_layout.h"

namespace trc {

struct MeshHost {
  std::vector<float> pos;       // [V][3]
  std::vector<uint32_t> idx;    // [T][3]
  std::vector<float> uv;        // [nuv][2] (optional)
  std::vector<uint32_t> uv_idx; // [T][3]   (optional)
  int32_t material = 0;
  int32_t emitter = -1;         // emitter shape index, assigned at flatten time
  float radiance[3] = {0, 0, 0};
  bool is_emitter = false;
  bool smooth = false;
  bool diff = false;            // vertex positions are differentiable
};

// Edge of a mesh (ObjMesh::processEdges): vertices v0 < v1, the opposite
// vertex v2 of face0, the two incident faces (face1 = -1 on a boundary).
struct EdgeHost {
  uint32_t v0, v1, v2;
  int32_t f0, f1;
};

struct BvhBuildStats {
  int32_t nodes = 0;
  int32_t leaves = 0;
  int32_t max_depth = 0;
  int32_t stack_need = 0;
  double sah_cost = 0.0;
};

// Refit order of a 4-wide BVH: node ids grouped by depth, deepest level
// first; level k holds nodes[level_off[k] .. level_off[k+1]).  Every child
// of a node sits one level deeper, so the levels can be refit in turn.
struct BvhRefitPlan {
  std::vector<uint32_t> nodes;
  std::vector<uint32_t> level_off;
};

struct SceneHost {
  std::vector<MeshHost> meshes;
  std::vector<Material> materials;
  std::vector<std::vector<float>> textures;  // texels [h][w][3]
  std::vector<int32_t> tex_w, tex_h;

  // ---- flattened (configure) ----
  std::vector<float> pos;        // [V][3]
  std::vector<uint32_t> idx;     // [T][3]
  std::vector<float> vnrm;       // [V][3]
  std::vector<float> uv;
  std::vector<uint32_t> uv_idx;
  std::vector<TriShade> shade;   // [T]
  std::vector<uint32_t> mesh_vert_offset, mesh_tri_offset;
  // emitters
  std::vector<int32_t> light_tri;
  std::vector<float> light_f, light_cdf;
  std::vector<float> emit_rad;   // [n_emitters][3]
  float light_integral = 0.0f;
  // edges (global vertex / triangle ids)
  std::vector<EdgeHost> edges;
  std::vector<uint32_t> mesh_edge_offset;
  // secondary (scene-wide) edge list and its length distribution
  std::vector<EdgeDev> sec_edges;
  std::vector<float> sec_len, sec_cdf;
  float sec_integral = 0.0f;
  // BVH
  std::vector<Bvh4Node> nodes;
  std::vector<TriRecord> tris;
  std::vector<uint32_t> tri_order;
  BvhBuildStats bvh_stats;
  BvhRefitPlan refit_plan;

  int32_t num_vertices() const { return (int32_t)(pos.size() / 3); }
  int32_t num_triangles() const { return (int32_t)(idx.size() / 3); }

  // Flatten meshes, compute face normals / areas / vertex normals, emitter
  // tables and edge lists.  Throws std::runtime_error on invalid input.
  void flatten();
  // Recompute derived data after vertex positions changed (same topology).
  void update_positions(const float* new_pos);
  // Build the BVH over the current positions.
  void build_bvh();
  // Refit: keep the topology of the BVH, recompute boxes / triangle records.
  void refit_bvh();
};

// Camera reduction (pos, target, up, w, h, fov in degrees).
CameraDev make_camera(const float pos[3], const float target[3], const float up[3], int w, int h,
                      float fov_deg);

// Edge filters of the boundary terms.  Secondary: boundary edges, or
// non-coplanar convex ones (Edge.cpp:42-88).  The per-camera primary list
// (Edge.cpp:313-348) is selected on the device (kernels/update.hip).
void select_sec_edges(const SceneHost& s, std::vector<EdgeDev>& out);
// Distribution1D over edge lengths: len, normalized cdf, integral.  The
// prefix sum runs in a blocked order (sequential within chunks of kCdfChunk
// edges, then over the chunk sums), which the device reproduces in parallel.
constexpr int kCdfChunk = 1024;
float edge_distribution(const SceneHost& s, const std::vector<EdgeDev>& e, std::vector<float>& len,
                        std::vector<float>& cdf);

// Mesh edge list with the reference's ordering (std::map keyed by (min v,
// max v), lexicographic).  Returns edges with local vertex / face ids.
std::vector<EdgeHost> build_edge_list(const std::vector<uint32_t>& idx);

// BVH build (binned SAH, collapsed to 4-wide).
// Subtrees are built on up to 16 threads; the result does not depend on the
// thread count.  `plan` (optional) receives the refit order.
void bvh_build(const std::vector<float>& pos, const std::vector<uint32_t>& idx,
               std::vector<Bvh4Node>& nodes, std::vector<TriRecord>& tris,
               std::vector<uint32_t>& order, BvhBuildStats& stats, BvhRefitPlan* plan = nullptr);
void bvh_refit(const std::vector<float>& pos, const std::vector<uint32_t>& idx,
               const std::vector<uint32_t>& order, std::vector<Bvh4Node>& nodes,
               std::vector<TriRecord>& tris);

}  // namespace trc
