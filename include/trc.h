/*
 * trc.h -- C ABI of the MI355X (gfx950) differentiable path-space integrator.
 *
 * This is the drop-in boundary that the TensorRay host layer (the
 * `TensorRay` Python module, TensorRay.cpp:48-222 in the reference, and its
 * C++ Scene / Integrator classes) binds.  Every entry point takes plain
 * pointers, sizes and an explicit HIP stream (`trc_stream`, a hipStream_t;
 * NULL = the default stream), returns a status code and never throws.  The
 * message of the last failure is available from trc_last_error().
 *
 * Memory: scenes (vertex / index buffers, BVH, emitter tables, textures) are
 * owned by the library and stay resident in HBM until trc_scene_destroy.
 * Films, adjoint images and gradient buffers are caller-owned device memory.
 *
 * Image layout (as the reference Tensorf image, Shape({W*H}, Vec3)): three
 * planes [3][H*W] of float, pixel index = y * W + x.
 * Vertex gradients: [V][3] float, V = total vertex count of the scene, in
 * shape order.  Material gradients: [M][8] float per BSDF, slots
 * {roughness, albedo/kd r, g, b, specular/F0 r, g, b, unused}; for a rough
 * conductor {roughness, eta r, g, b, k r, g, b, unused}.
 */
#ifndef TRC_H
#define TRC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t trc_status;
#define TRC_OK 0
#define TRC_ERR_INVALID 1   /* bad argument or scene */
#define TRC_ERR_HIP 2       /* HIP runtime failure */
#define TRC_ERR_NOMEM 3     /* device allocation failed or budget exceeded */
#define TRC_ERR_STATE 4     /* call order (e.g. render before configure) */
#define TRC_ERR_OVERFLOW 5  /* the scene's BVH is deeper than the traversal stack (16 LDS + 64 spill entries) */

typedef struct trc_ctx trc_ctx;
typedef struct trc_scene trc_scene;
typedef void* trc_stream;

/* Material types (BSDF classes of the reference). */
#define TRC_MAT_DIFFUSE 0      /* Diffuse.cpp:37-85 */
#define TRC_MAT_MICROFACET 1   /* Microfacet.cpp:37-165 */
#define TRC_MAT_ROUGHCONDUCTOR 2 /* Roughconductor.cpp:37-121; albedo = eta, spec = k, no texture */

/* Mesh flags. */
#define TRC_MESH_SMOOTH 1      /* smooth shading from averaged face normals (Primitive.cpp:323-340) */
#define TRC_MESH_DIFF 2        /* vertex positions receive gradients (Primitive::DiffAllVertexPos) */

typedef struct trc_camera {
  float pos[3], target[3], up[3];
  int32_t width, height;
  float fov; /* vertical field of view, degrees */
} trc_camera;

/* RenderOptions (Integrator.h:47-78; bound at TensorRay.cpp:165-188). */
typedef struct trc_options {
  uint32_t seed;         /* counter-RNG seed: same seed => same sample paths */
  int32_t max_bounces;   /* B: NEE + BSDF steps after the primary hit */
  int32_t spp;           /* interior samples per pixel */
  int32_t sppe;          /* primary-edge samples per pixel */
  int32_t sppse0;        /* secondary (direct) edge samples per pixel */
  int32_t sppse1;        /* secondary (indirect) edge samples per pixel */
  int32_t sppe0;         /* pixel-boundary samples per pixel */
  int32_t quiet;
  /* Guiding of the secondary direct term (Integrator.h:72-76, AQ_distrb.h):
   * one leaf per secondary edge over [cdf_{i-1}, cdf_i] x [0,1]^2, weight =
   * mean |baseVal| over g_direct_spp samples, mixed eps + (1 - eps) w. */
  int32_t g_direct;          /* 0 = off */
  int32_t g_direct_depth;    /* accepted and ignored: the reference's refinement loop is disabled */
  int32_t g_direct_max_size; /* leaf capacity; more secondary edges than this is an error (<= 0: no limit) */
  int32_t g_direct_spp;      /* guide samples per leaf, 1..4096 */
  float g_direct_thold;      /* accepted and ignored, as in the reference */
  float g_eps;               /* uniform mix in [0, 1]; the reference uses 0.01 */
  /* Guide reuse.  No counterpart in the reference, which rebuilds the guide
   * in every RenderD.  The guide's leaf weights depend on the geometry and
   * the lights only, not on the camera, so the views of one scene can share
   * one guide: with g_keep != 0 the guide is built from g_seed (not seed),
   * kept on the scene, and reused by later calls with the same g_seed,
   * g_direct_spp and g_eps until the scene changes.  0 = rebuild per call. */
  int32_t g_keep;
  uint32_t g_seed;
} trc_options;

typedef struct trc_scene_info {
  int32_t num_vertices, num_triangles, num_materials, num_emitter_triangles;
  int32_t num_edges;
  int32_t bvh_nodes, bvh_leaves, bvh_max_depth, bvh_stack_need;
  double bvh_sah_cost;
  int64_t device_bytes;
  int64_t texture_floats; /* size of the texture-gradient buffer of trc_render_adj */
  int32_t bvh_refits;     /* trc_scene_set_vertices refits (on the device) since the scene was created */
  int32_t bvh_rebuilds;   /* full rebuilds: refit = 0, or a refit tree whose SAH cost grew past 1.25x its build's */
} trc_scene_info;

/* ---- context: replaces EnvCreate / EnvRelease (pybind_utils.h:38-61) ---- */
trc_status trc_context_create(int32_t device, trc_ctx** out);
/* A context without a device: scene assembly and trc_scene_prepare only
 * (host-side validation of the flattened scene, edge lists and BVH). */
trc_status trc_context_create_host(trc_ctx** out);
trc_status trc_context_destroy(trc_ctx* ctx);
const char* trc_last_error(const trc_ctx* ctx);
/* Number of compute units of the context's device. */
int32_t trc_device_cus(const trc_ctx* ctx);

/* ---- scene assembly: replaces the SceneLoader shape/BSDF/emitter setup and
 *      Scene::ExtractGeometry / ConfigureLight (Scene.cpp:46-124, 471-520) ---- */
trc_status trc_scene_create(trc_ctx* ctx, trc_scene** out);
trc_status trc_scene_destroy(trc_scene* scene);
/* Returns the material index (>= 0) or -1 on error. */
int32_t trc_scene_add_material(trc_scene* scene, int32_t type, const float albedo[3], const float spec[3],
                               float roughness, int32_t texture);
/* texels: [h][w][3] float (host).  Returns the texture index or -1. */
int32_t trc_scene_add_texture(trc_scene* scene, const float* texels, int32_t w, int32_t h);
/* pos [nv][3], idx [nt][3] (host).  uv / uv_idx optional (NULL).  emit: RGB
 * radiance of an area emitter or NULL.  Returns the shape index or -1. */
int32_t trc_scene_add_mesh(trc_scene* scene, const float* pos, int32_t nv, const uint32_t* idx, int32_t nt,
                           const float* uv, int32_t nuv, const uint32_t* uv_idx, int32_t material,
                           const float* emit, int32_t flags);
/* Scene::Configure (Scene.cpp:522-533): normals, areas, emitter CDF, edge
 * lists, SAH 4-wide BVH, upload.  Synchronous with respect to `stream`. */
trc_status trc_scene_configure(trc_scene* scene, trc_stream stream);
/* Host half of trc_scene_configure (flatten, derived data, edges, BVH)
 * without any device work. */
trc_status trc_scene_prepare(trc_scene* scene);
/* Scene::Update (Scene.cpp:535-546): replace all vertex positions ([V][3],
 * host or device pointer) and refit (refit = 1) or rebuild (0) the BVH.
 * A refit runs on the device (normals, areas, emitter table, secondary
 * edges, BVH boxes; the positions never travel to the host) and rebuilds the
 * tree on the host when its SAH cost has grown past 1.25x its cost at the
 * last build.  Synchronous with respect to `stream`. */
trc_status trc_scene_set_vertices(trc_scene* scene, const float* pos, int32_t nv, int32_t refit,
                                  trc_stream stream);
/* Set a material's roughness / albedo / specular (BSDF parameters). */
trc_status trc_scene_set_material(trc_scene* scene, int32_t material, const float albedo[3],
                                  const float spec[3], float roughness);
trc_status trc_scene_get_info(const trc_scene* scene, trc_scene_info* out);
/* Copies of the flattened host arrays (sizes from trc_scene_get_info). */
trc_status trc_scene_get_vertices(const trc_scene* scene, float* out /* [V][3] */);
trc_status trc_scene_get_indices(const trc_scene* scene, uint32_t* out /* [T][3] */);
/* BVH arrays for validation: nodes [n][32] words, triangle order [T]. */
trc_status trc_scene_get_bvh(const trc_scene* scene, uint32_t* nodes, uint32_t* tri_order);
/* Edge list: [E][5] int32 (v0, v1, v2, f0, f1), global ids (ObjMesh::processEdges). */
trc_status trc_scene_get_edges(const trc_scene* scene, int32_t* out);

/* ---- ray queries: replaces RayIntersectOptix (Optix.cpp:252-305) ----
 * rays [n][8] = (ox, oy, oz, tmin, dx, dy, dz, tmax), device memory.
 * tri [n] = hit triangle (-1 miss), uvt [n][3] = (u, v, t) (0, 0, -1 on miss).
 * any_hit = 1 returns some hit in range (occlusion queries). */
trc_status trc_intersect(trc_scene* scene, const float* rays, int32_t n, int32_t* tri, float* uvt,
                         int32_t any_hit, trc_stream stream);

/* ---- interior path tracer: PathTracer::RenderC / RenderD ----
 * (PathTracer.h:60-68, Integrator.h:106-121, PathTracer.cpp:37-123)
 * film: [3][H*W] device, overwritten.  */
trc_status trc_render_fwd(trc_scene* scene, const trc_camera* cam, const trc_options* opt, float* film,
                          trc_stream stream);
/* trc_render_fwd without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The film
 * is bit-identical to trc_render_fwd's.  Its per-path radiance and traversal
 * spill live in one fixed 64 MiB device array per GPU, held until the
 * process ends; the samples are rendered in chunks that fit it.  Fails with
 * TRC_ERR_INVALID when one sample of every pixel (12 B x W x H) plus the
 * scene's spill slice does not fit (about 5.5M pixels).  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams,
 * other contexts or threads) race and give wrong films. */
trc_status trc_render_fwd_noalloc(trc_scene* scene, const trc_camera* cam, const trc_options* opt, float* film,
                                  trc_stream stream);
/* Per-path radiance (before the 1/spp factor), [spp*H*W][3], path index =
 * sample * H*W + pixel.  Used by tests and the serving loop. */
trc_status trc_render_paths(trc_scene* scene, const trc_camera* cam, const trc_options* opt, float* paths,
                            trc_stream stream);
/* film[c][pix] = (sum_s paths[s*wh + pix][c]) / spp, summed in sample order
 * (deterministic); the reduction that follows trc_render_paths. */
trc_status trc_film_from_paths(trc_ctx* ctx, const float* paths, int32_t spp, int32_t wh, float* film,
                               trc_stream stream);
/* Adam step on n device floats (Adam::Step, Optimizer.h:95-122 of the
 * reference's C++ examples): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
 * x -= lr (m / (1-b1^step)) / (sqrt(v / (1-b2^step)) + eps).  step counts
 * from 1; m and v start zeroed; non-finite entries of grad count as zero. */
trc_status trc_adam_step(trc_ctx* ctx, float* x, const float* grad, float* m, float* v, uint64_t n, int32_t step,
                         float lr, float beta1, float beta2, float eps, trc_stream stream);
/* Uniform Adam step, the update of the reference's LargeSteps optimizer
 * (pyTensorRay/optimizer.py:40-63; AdamUniform of Nicolet et al. 2021,
 * "Large Steps in Inverse Rendering of Geometry"): moments as in
 * trc_adam_step, then x -= lr / ((1-b1^step) sqrt(max(v) / (1-b2^step))) m,
 * one scale for every coordinate (no update while max(v) = 0). */
trc_status trc_adam_uniform_step(trc_ctx* ctx, float* x, const float* grad, float* m, float* v, uint64_t n,
                                 int32_t step, float lr, float beta1, float beta2, trc_stream stream);
/* Solver for M = I + lambda L, L the uniform graph Laplacian of a triangle
 * mesh (L_ii = number of neighbours, L_ij = -1 per mesh edge) -- the
 * matrix of largesteps' compute_matrix used by LargeSteps
 * (pyTensorRay/optimizer.py:147-158).  Vectors are [nv][3] device floats;
 * solve runs conjugate gradients on the three columns (x holds the initial
 * guess) until |r| <= tol |b| per column or max_iter iterations, without
 * host synchronization; status reads the iteration count and relative
 * residuals back (synchronous).  Solves on one solver must not overlap. */
typedef struct trc_lsolver trc_lsolver;
trc_status trc_lsolver_create(trc_ctx* ctx, const uint32_t* faces /* host [nf][3] */, int32_t nf, int32_t nv,
                              float lambda, trc_lsolver** out);
trc_status trc_lsolver_apply(trc_lsolver* s, const float* x, float* y, trc_stream stream);  /* y = M x */
trc_status trc_lsolver_solve(trc_lsolver* s, const float* b, float* x, int32_t max_iter, float tol,
                             trc_stream stream);
trc_status trc_lsolver_status(trc_lsolver* s, int32_t* iters, float rel_residual[3]);
trc_status trc_lsolver_destroy(trc_lsolver* s);
/* Diagnostic: work counters of the forward paths, counters[4] = (paths,
 * ray queries, BVH node visits, triangle tests).  Synchronous. */
trc_status trc_render_stats(trc_scene* scene, const trc_camera* cam, const trc_options* opt, uint64_t* counters,
                            trc_stream stream);
/* Boundary (edge) terms of the gradient, accumulated into dpos ([V][3]):
 *   kind 0  primary edges      (PrimaryBoundaryIntegrator::RenderD, BoundaryPrimary.cpp:36-181), opt->sppe
 *   kind 1  pixel boundary     (PixelBoundaryIntegrator::RenderD, BoundaryPixel.cpp:36-159), opt->sppe0
 *   kind 2  secondary, direct  (DirectBoundaryIntegrator::RenderD, BoundaryDirect.cpp:38-143), opt->sppse0
 *   kind 3  secondary, indirect (IndirectBoundaryIntegrator::RenderD, BoundaryIndirect.cpp:38-192), opt->sppse1
 * Samples per pixel 0 skips the term.  Their forward value is zero. */
trc_status trc_render_boundary(trc_scene* scene, const trc_camera* cam, const trc_options* opt, int32_t kind,
                               const float* dLdI, float* dpos, trc_stream stream);
/* Forward-mode derivative image of one integrator along a scalar parameter
 * theta (RenderD with RenderOptions.export_deriv, Integrator.h:106-121, and
 * GradiantImageHandler::AccumulateDeriv, Integrator.cpp:338-373):
 * dimg [3][H*W] += d(image)/d(theta).  The parameter enters through its
 * tangents: tpos [V][3] = d(vertex position)/d(theta) and tmat [M][8] =
 * d(material slot)/d(theta) (NULL: none; used by the interior term only),
 * both device memory.  kind -1 = the interior path tracer (opt->spp;
 * smooth-shaded differentiable meshes included through the vertex-normal
 * tangents), 0..3 = the boundary terms of trc_render_boundary.  Texels
 * carry no tangent.  Non-finite derivatives are dropped, as the reference
 * filters its derivative images.  The sum over all five integrators is the
 * derivative image the reference's validation compares with finite
 * differences (validate_gpu.py:69-101). */
trc_status trc_render_deriv(trc_scene* scene, const trc_camera* cam, const trc_options* opt, int32_t kind,
                            const float* tpos, const float* tmat, float* dimg, trc_stream stream);
/* Path checkpoint (adjoint without re-tracing).  The checkpoint holds, per
 * path, the outcome of every ray query -- hit triangle and barycentrics of
 * each vertex and the NEE visibility bits -- in
 * trc_checkpoint_bytes = spp*W*H * (4 + 12 * (max_bounces + 1)) bytes of
 * caller-owned device memory (-1 for invalid arguments).  The adjoint then
 * re-evaluates each path from its recorded hits instead of replaying the
 * traversal (PathTracer::RenderD replays the whole forward pass,
 * PathTracer.cpp:107-111); the result equals trc_render_adj's. */
int64_t trc_checkpoint_bytes(const trc_camera* cam, const trc_options* opt);
/* trc_render_fwd that also writes the checkpoint. */
trc_status trc_render_fwd_ckpt(trc_scene* scene, const trc_camera* cam, const trc_options* opt, float* film,
                               void* ckpt, int64_t ckpt_bytes, trc_stream stream);
/* trc_render_adj from a checkpoint written by trc_render_fwd_ckpt with the
 * same scene state, camera and options. */
trc_status trc_render_adj_ckpt(trc_scene* scene, const trc_camera* cam, const trc_options* opt, const void* ckpt,
                               int64_t ckpt_bytes, const float* dLdI, float* dpos, float* dmat, float* dtex,
                               trc_stream stream);
/* ---- batched rendering: many requests (scenes, cameras, sizes, depths) in
 * one launch.  No counterpart in the reference, which renders one request
 * per call (TensorRay.cpp:192-221); this is the device side of the serving
 * loop.  A job covers paths [first_path, first_path + num_paths) of its
 * request's spp*W*H paths (path index = sample * W*H + pixel), so a request
 * can be split over several batches.  Each path's result is the same as in
 * a trc_render_fwd_ckpt / trc_render_adj_ckpt call of the whole request. */
typedef struct trc_job {
  trc_scene* scene;
  trc_camera cam;
  trc_options opt;       /* seed, max_bounces, spp */
  uint64_t first_path, num_paths;
  float* paths;          /* forward: per-path radiance [spp*W*H][3] of the request */
  void* ckpt;            /* the request's checkpoint (trc_checkpoint_bytes): forward writes, adjoint reads */
  const float* dLdI;     /* adjoint: [3][H*W] */
  float* dpos;           /* adjoint: [V][3] */
  float* dmat;           /* adjoint: [M][8] or NULL */
  float* dtex;           /* adjoint: texel gradients or NULL */
  float* vn_scratch;     /* adjoint: zeroed [V][3] scratch when the scene has smooth differentiable meshes
                            (left zeroed on return), else NULL */
} trc_job;
/* adjoint = 0: forward of every job (paths and checkpoint); 1: adjoint of
 * every job from its checkpoint.  Stream-ordered; jobs of one request may
 * appear in different batches but not twice in one. */
trc_status trc_render_batch(trc_ctx* ctx, const trc_job* jobs, int32_t n, int32_t adjoint, trc_stream stream);

/* Prepared batch (multi-view rendering): trc_plan_create validates the jobs,
 * uploads the job table once and allocates what the launch needs
 * (privatized gradient copies), so that trc_plan_launch only launches
 * kernels -- no allocation, no host synchronization -- and may be captured
 * into a hipGraph and replayed.  The jobs' buffers and scenes must stay
 * valid for the plan's lifetime.  The plan snapshots its scenes' device
 * state: after trc_scene_set_vertices / trc_scene_set_material /
 * trc_scene_configure on one of them, launching or tuning fails with
 * TRC_ERR_STATE until the plan is created again.  No counterpart in the reference, which renders views one
 * at a time (renderer_multi_view.py:55-78). */
typedef struct trc_plan trc_plan;
trc_status trc_plan_create(trc_ctx* ctx, const trc_job* jobs, int32_t n, int32_t adjoint, trc_plan** out);
/* Paths per tile of the persistent launch (multiple of 64; 0 = built-in heuristic). */
trc_status trc_plan_set_tile(trc_plan* plan, uint32_t tile_paths);
uint32_t trc_plan_tile(const trc_plan* plan);
/* A plan owns one spill buffer and one set of gradient copies: launches of
 * the same plan must not overlap (launch it on one stream, or order the
 * streams); different plans may run concurrently. */
trc_status trc_plan_launch(trc_plan* plan, trc_stream stream);
/* Autotuner: launches the plan `reps` times per candidate tile size, keeps
 * the fastest (median of HIP-event times) and returns it in *best.  The
 * launches write the outputs; an adjoint plan accumulates into its gradient
 * buffers, which the caller clears afterwards. */
trc_status trc_plan_tune(trc_plan* plan, const uint32_t* candidates, int32_t n, int32_t reps, uint32_t* best,
                         trc_stream stream);
trc_status trc_plan_destroy(trc_plan* plan);

/* Adjoint: accumulates dL/dV += J^T dLdI into dpos ([V][3] device), the
 * material gradients into dmat ([M][8] device, may be NULL) and the texel
 * gradients of albedo / kd textures into dtex (may be NULL; all textures'
 * [h][w][3] texels concatenated in texture order, texture_floats floats;
 * BSDF::DiffTexture, BSDF.cpp:350-356).  Calls on one scene are ordered on
 * one stream (smooth-shaded meshes use scene-owned scratch). */
trc_status trc_render_adj(trc_scene* scene, const trc_camera* cam, const trc_options* opt, const float* dLdI,
                          float* dpos, float* dmat, float* dtex, trc_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* TRC_H */
