/*
 * srt_hip.h - C ABI of the MI355X (gfx950) path-tracing hot path.
 *
 * The reference renderer (woAIxuexiSR/SRT) has no C ABI: its hot path is the
 * C++ `RenderPass` interface (src/passes/renderpass.h:10-30) whose
 * `PathTracer` pass (src/passes/integrator/path_tracer.cu:5-33) launches an
 * OptiX megakernel (src/shader/integrator/path_tracer.cu:50-169) over
 * acceleration structures built by `OptixRayTracer`
 * (src/optix/optix_ray_tracer.cu:213-463).  Each entry point below names the
 * reference interface it replaces.  The C++ passes in srt_amd/host call these
 * functions; so do the Python bindings.
 *
 * Conventions
 *   - Every function returns SRT_OK (0) or a negative SRT_E_* code and never
 *     exits the process (the reference prints and calls exit(-1)).
 *     srt_last_error() returns the message of the last failure on a context.
 *   - Host arrays passed in are copied; the caller keeps ownership.
 *     Scene handles are owned by the context.
 *   - A context is used from one host thread at a time, except srt_submit /
 *     srt_poll / srt_wait / srt_serve_stats, which are thread-safe.
 *   - The film is RGBA float32, row-major, pixel index = iy * width + ix, row 0
 *     at the bottom of the image (src/passes/film.h:7-20).
 */
#ifndef SRT_HIP_H
#define SRT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SRT_OK 0
#define SRT_E_INVALID (-1)   /* bad argument or unknown name */
#define SRT_E_NO_DEVICE (-2) /* no HIP device with that ordinal */
#define SRT_E_HIP (-3)       /* HIP runtime error */
#define SRT_E_OOM (-4)       /* exceeds the context's HBM budget or allocation failed */
#define SRT_E_NOT_FOUND (-5) /* unknown scene / ticket handle */
#define SRT_E_LIMIT (-6)     /* exceeds a documented limit (BVH depth, leaf size, sizes) */
#define SRT_E_STATE (-7)     /* call not valid in the current state */

typedef struct srt_ctx srt_ctx;
typedef uint64_t srt_scene;

/* BxDF kinds, src/device_include/scene/bxdf.h:62 */
#define SRT_BXDF_DIFFUSE 0
#define SRT_BXDF_DIFFUSE_TRANSMISSION 1
#define SRT_BXDF_DIELECTRIC 2
#define SRT_BXDF_DISNEY 3

/* Material: BxDF parameters (bxdf.h:64-76) + GMaterial colours (gmaterial.h:9-12).
 * 80 bytes.  A material is emissive iff intensity > 0. */
typedef struct srt_material {
    int32_t type;
    float ior, metallic, subsurface, roughness, specular, specular_tint, anisotropic, sheen, sheen_tint,
        clearcoat, clearcoat_gloss, spec_trans;
    float base_color[3];
    float emission_color[3];
    float intensity;
} srt_material;

/* World-space triangle meshes (one identity instance per mesh).  Replaces the
 * host Scene -> GScene upload (src/scene/scene.cu:84-288). */
typedef struct srt_scene_desc {
    int32_t num_vertices;
    const float* positions;           /* 3 * num_vertices */
    const float* normals;             /* 3 * num_vertices, or NULL */
    const float* tangents;            /* 3 * num_vertices, or NULL */
    int32_t num_triangles;
    const uint32_t* indices;          /* 3 * num_triangles */
    int32_t num_meshes;
    const int32_t* mesh_first_triangle; /* num_meshes + 1 offsets into the triangles */
    const int32_t* mesh_material;     /* num_meshes */
    int32_t num_materials;
    const srt_material* materials;
    float env_color[3];               /* constant environment (light.h:27-44) */
} srt_scene_desc;

/* Pinhole camera, src/renderer/renderer.cu:125-140 (LookAt + vertical fov). */
typedef struct srt_camera {
    float position[3], target[3], up[3];
    float fov_deg;
} srt_camera;

/* PathTracer pass parameters, src/passes/integrator/path_tracer.h:34-38
 * (defaults 1, 16, 4, 1, 1).  `seed` replaces the host random frame seed
 * (path_tracer.cu:17).  mlp_mode: 0 plain tracer, 1 query the radiance MLP
 * after vertex 2 (needs srt_mlp_*). */
typedef struct srt_pt_params {
    int32_t samples_per_pixel, max_depth, rr_depth, use_nee, use_mis;
    uint32_t seed;
    int32_t mlp_mode;
    int32_t reserved;
} srt_pt_params;

typedef struct srt_bvh_opts {
    int32_t max_leaf;      /* 1..8; 0: the default, 2 */
    int32_t cluster_size;  /* 0: automatic (max(4, n/262144) triangles per cluster under the binned-SAH top);
                            * at least max_leaf either way */
} srt_bvh_opts;

/* ---- context (replaces OptixRayTracer's context/stream, optix_ray_tracer.cu:37-56) ---- */
int srt_create(int device, size_t hbm_budget_bytes, srt_ctx** out);
void srt_destroy(srt_ctx* ctx);
const char* srt_last_error(const srt_ctx* ctx);
/* tunables: "path_pool" (path slots per chunk, default 2^27), "profile" (0/1 per-kernel HIP events),
 * "count_visits" (0/1 count BVH node visits and triangle tests), "hbm_budget"
 * (bytes; at least srt_hbm_used), "trace_stack_lds" (traversal stack entries per
 * lane in LDS, 4..64, default 16), "trace_refill" / "shadow_refill" (finished
 * lanes that make a traversal wave fetch new rays, 1..64, defaults 40 / 32),
 * "trace_blocks_per_cu"
 * (0: as many as fit), "trace_bvh_width" (4: traverse the 4-wide tree collapsed
 * from the binary one, default; 2: the binary tree; results are identical),
 * "trace_inner_exit" (a traversal wave leaves its inner-node loop once at most
 * this many lanes still search for a leaf, 0..63, default 18), "max_wide_nodes"
 * (a BVH built while this is set keeps its 4-wide form only up to this many
 * 4-wide nodes, else it is traversed in binary form with the same results;
 * 0..16777215, the default: 4-wide nodes are addressed with 32-bit byte
 * offsets), "serve_age" (serving order, below: the weight of the time a request
 * has waited, per mille, 0..1000, default 250), "trace_packet" (camera rays of a
 * flat scene with a 4-wide tree are traversed as 64-ray packets, one node
 * sequence per wave; the value is the number of lanes holding postponed leaves
 * that makes the wave test them, 1..64; 0, the default: per-ray traversal; the
 * results are identical; packets gain 1 % of the closest-hit time on a scene
 * that fills the frame with deep traversals and lose 10 % where camera rays
 * visit few nodes), "keep_build_scratch" (1, the
 * default: the device scratch of a BVH build stays allocated for the next build
 * and counts in srt_hbm_used; it is freed when an allocation would exceed the
 * budget or fail; 0: freed now and at the end of every build).  Limits: 2^28 - 1 triangles per scene; "path_pool" and the pool
 * of srt_serve_start in [64, 2^31 - 1]. */
int srt_set_option(srt_ctx* ctx, const char* key, int64_t value);
size_t srt_hbm_used(const srt_ctx* ctx);

/* ---- scenes (replaces Scene::build_gscene, scene.cu:84-288) ---- */
int srt_scene_upload(srt_ctx* ctx, const srt_scene_desc* desc, srt_scene* out);
int srt_scene_release(srt_ctx* ctx, srt_scene scene);
/* light CDF and per-mesh area as built on upload (scene.cu:240-264, mesh.cpp:3-18);
 * any pointer may be NULL.  Returns the number of lights or a negative code. */
int srt_scene_lights(srt_ctx* ctx, srt_scene scene, float* light_cdf, int32_t* light_mesh, float* area_cdf,
                     float* mesh_area, float* weight_sum);

/* ---- instances (replace the IAS over per-mesh GAS with 3x4 instance
 *      transforms, optix_ray_tracer.cu:337-359, and the instance data of
 *      Scene::build_gscene_instances, scene.cu:225-288) ----
 * After srt_scene_upload: the scene's meshes become object space and
 * instance i places mesh inst_mesh[i] with the row-major 3x4 transform
 * xforms_3x4[12 i .. 12 i + 11] (a mesh may be placed several times).  The
 * lights become the emissive instances, weighted by intensity x the mesh's
 * object-space area (scene.cu:240-264).  srt_scene_build_bvh then builds one
 * tree per mesh plus a top-level tree over the instances; hits apply the
 * transform as hit_info.h:58-62 does (position by apply_point, normal and
 * shading frame by apply_vector, not renormalised), and so do light samples
 * (gmesh.h:77-78).  Closest hit: (t, primitive id, instance) lexicographic.
 * Instanced scenes take the path tracer, serving and the radiance MLP; not
 * skinning, the Simple or LightTracer passes, nor srt_scene_bvh_download.
 * Limits: fewer than 2^27 triangles, fewer than 2^30 - 1 instances. */
int srt_scene_set_instances(srt_ctx* ctx, srt_scene scene, int32_t n, const int32_t* inst_mesh, const float* xforms_3x4);
/* node animation (Scene::update_gscene's instance transforms, scene.cu:309-334,
 * then update_as, optix_ray_tracer.cu:443-463): new transforms for all
 * instances, then the top-level tree is refit (the mesh trees are unchanged) */
int srt_scene_update_instances(srt_ctx* ctx, srt_scene scene, const float* xforms_3x4);

/* ---- acceleration structure (replaces OptixRayTracer::build_gas/build_ias,
 *      optix_ray_tracer.cu:213-359, and update_as, :443-463) ---- */
int srt_scene_build_bvh(srt_ctx* ctx, srt_scene scene, const srt_bvh_opts* opts);
/* info[0] nodes, [1] root ref, [2] max depth, [3] clusters, [4] top-level nodes,
 * [5] cluster size, [6] triangles, [7] max leaf */
int srt_scene_bvh_info(srt_ctx* ctx, srt_scene scene, int32_t info[8]);
/* host copies: nodes (64 B each), sorted->primitive order, sorted triangles (12 floats each) */
int srt_scene_bvh_download(srt_ctx* ctx, srt_scene scene, void* nodes, uint32_t* order, float* tris);
/* linear blend skinning (scene.cu:336-395): 4 (bone id, weight) per vertex; id -1 ends the list */
int srt_scene_set_skin(srt_ctx* ctx, srt_scene scene, const int32_t* bone_ids, const float* bone_weights);
/* per frame: skin with bone matrices (3x4 row-major, n_bones of them), then refit the BVH */
int srt_scene_skin_refit(srt_ctx* ctx, srt_scene scene, const float* bones_3x4, int32_t n_bones);
int srt_scene_download_vertices(srt_ctx* ctx, srt_scene scene, float* positions, float* normals, float* tangents);

/* ---- rays (replaces optixTrace closest / shadow, path_tracer.cu:75,124) ----
 * Host arrays.  any = 0: closest hit, prim = -1 on miss; any = 1: prim = 1 if
 * anything is hit in [tmin, tmax[i]], else 0. */
int srt_trace_rays(srt_ctx* ctx, srt_scene scene, int32_t n, const float* origins, const float* dirs, float tmin,
                   const float* tmax, int32_t any, float* t, float* u, float* v, int32_t* prim);

/* As srt_trace_rays, with the instance hit (closest hits of instanced
 * scenes; -1 on a miss and for scenes without instances). */
int srt_trace_rays_inst(srt_ctx* ctx, srt_scene scene, int32_t n, const float* origins, const float* dirs, float tmin,
                        const float* tmax, int32_t any, float* t, float* u, float* v, int32_t* prim, int32_t* inst);

/* ---- ray queries that allocate nothing ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The same queries and results as srt_trace_rays, but
 * origins, dirs, tmax and t, u, v, prim are DEVICE arrays, and the call makes
 * no allocation: the traversal's fetch counters and stack spill area live in
 * a file-scope device array of 32 MiB per GPU (held until the process ends),
 * looked up once per device by srt_create.  The persistent grid is sized to
 * fit its stack spill in that array; a tree whose one-workgroup spill exceeds
 * it returns SRT_E_LIMIT.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (from several contexts or threads) race. */
int srt_trace_rays_noalloc(srt_ctx* ctx, srt_scene scene, int32_t n, const float* origins, const float* dirs,
                           float tmin, const float* tmax, int32_t any, float* t, float* u, float* v, int32_t* prim);

/* ---- render (replaces PathTracer::render + the megakernel, path_tracer.cu:13-33) ----
 * Writes float4 RGBA (alpha 1) for width*height pixels.  out_on_device = 1:
 * rgba_out is device memory; 0: host memory; rgba_out == NULL keeps the film in
 * the context (see srt_film_download). */
int srt_render(srt_ctx* ctx, srt_scene scene, const srt_camera* camera, const srt_pt_params* params, int32_t width,
               int32_t height, float* rgba_out, int32_t out_on_device);

/* ---- Simple NPR pass (replaces Simple::render + its OptiX programs,
 * passes/npr/simple.cu:14-33, shader/npr/simple.cu:42-97) ----
 * samples_per_pixel jittered camera rays per pixel (the path tracer's camera
 * ray streams for the same seed), closest hit, and one first-hit colouring:
 * type 0 Depth (distance from the camera scaled between the scene box's
 * nearest and farthest distance, 0.2 .. 1.0), 1 Normal ((n + 1) / 2), 2
 * BaseColor, 3 Ambient (base colour x (|n . d| / 2 + 1/2)), 4 FaceOrientation;
 * misses are white.  Pixel = sum over samples of L / samples_per_pixel, alpha
 * 1.  Output as for srt_render (rgba_out == NULL is an error here). */
int srt_render_simple(srt_ctx* ctx, srt_scene scene, const srt_camera* camera, int32_t type, int32_t samples_per_pixel,
                      uint32_t seed, int32_t width, int32_t height, float* rgba_out, int32_t out_on_device);

/* ---- LightTracer pass (replaces LightTracer::render + its OptiX programs,
 * passes/integrator/light_tracer.cu:13-37, shader/integrator/light_tracer.cu:57-176) ----
 * width*height*samples_per_pixel particles from the area lights, splatted
 * through the pinhole camera by camera connections at non-specular vertices
 * (the reference's weights, aspect ratio and previous-direction pdf
 * included); uses samples_per_pixel, max_depth, rr_depth and seed of params.
 * The film is cleared to (0, 0, 0, 1) first.  Splats are float atomic adds:
 * a pixel's value can differ in the last bits between runs.  A scene without
 * area lights is an error.  Output as for srt_render (rgba_out != NULL). */
int srt_render_light(srt_ctx* ctx, srt_scene scene, const srt_camera* camera, const srt_pt_params* params, int32_t width,
                     int32_t height, float* rgba_out, int32_t out_on_device);
int srt_film_download(srt_ctx* ctx, float* rgba_host, int64_t num_pixels);

/* ---- post-processing (Accumulate::render accumulate.cu:10-29,
 *      ToneMapping::render tonemapping.cu:5-70) on device or host buffers ----
 * tonemap type: 0 None, 1 Clamp, 2 Reinhard, 3 Uncharted2, 4 ACES */
int srt_tonemap(srt_ctx* ctx, float* rgba, int64_t num_pixels, int32_t type, float exposure, int32_t use_gamma,
                int32_t on_device);
int srt_accumulate(srt_ctx* ctx, float* acc, float* rgba, int64_t num_pixels, int32_t frame_count, int32_t on_device);
/* Denoise pass (replaces Denoise::render, passes/postprocess/denoise.cu:54-114,
 * which runs the closed OptiX HDR denoiser on the colour alone).  In place on
 * width*height float4 pixels: five passes of the edge-avoiding a-trous
 * wavelet filter (Dammertz et al. 2010, B3-spline taps at spacing 1..16, the
 * colour-difference weight on log(1 + c)); alpha is kept.
 * on_device as srt_accumulate. */
int srt_denoise(srt_ctx* ctx, float* rgba, int32_t width, int32_t height, int32_t on_device);

/* ---- radiance MLP (replaces the NeuralRadiosity network, src/passes/npr/
 *      neural_radiosity.cu:6-52, whose training is a stub at :165-169) ----
 * 9 raw inputs in [0,1] (position normalised to the scene box, (n+1)/2,
 * (w_o+1)/2) -> frequency encoding (3 octaves, width 63) -> 4 x 64 ReLU ->
 * 3 outputs; BF16 MFMA with FP32 accumulation; relative-L2 loss; Adam.
 * Parameters in logical order: W0[64][63] b0[64] W1..W3[64][64]+b W4[3][64] b4[3]. */
int srt_mlp_num_params(void);
int srt_mlp_create(srt_ctx* ctx, uint32_t seed, float lr);
/* Selects the MLP that the other srt_mlp_* calls act on: the one bound to
 * `scene`, or with scene 0 the context's own (the default).  srt_mlp_create
 * creates (or replaces) the selected one.  Renders with mlp_mode 1, training
 * frames and serving requests of a scene use the scene's own MLP when it has
 * one, else the context's own.  Releasing a scene frees its MLP.  (The
 * reference's NeuralRadiosity pass holds one network per pass,
 * neural_radiosity.cu:76-82.) */
int srt_mlp_select(srt_ctx* ctx, srt_scene scene);
int srt_mlp_set_params(srt_ctx* ctx, const float* params, int32_t n);
int srt_mlp_get_params(srt_ctx* ctx, float* params, int32_t n);
int srt_mlp_set_hparams(srt_ctx* ctx, float lr, float beta1, float beta2, float eps);
int srt_mlp_set_bounds(srt_ctx* ctx, const float lo[3], const float hi[3]);
/* host arrays: in9 [n][9], out3 [n][3] */
int srt_mlp_infer(srt_ctx* ctx, int32_t n, const float* in9, float* out3);
/* one Adam step on (in9, target3) (apply = 1) or gradients only (apply = 0);
 * loss: mean relative L2 before the step; grads: logical order (may be NULL) */
int srt_mlp_train_step(srt_ctx* ctx, int32_t n, const float* in9, const float* target3, int32_t apply, float* loss,
                       float* grads);
/* Online training from the path tracer itself (config: radiance samples per
 * frame): renders width x height x spp training paths with the plain tracer,
 * records at vertex 2 (the second surface hit) the features and throughput,
 * takes (radiance gathered after vertex 2's NEE) / throughput as the target,
 * keeps the first max_samples samples in path order, and runs `steps` Adam
 * steps on equal consecutive mini-batches.  Render with mlp_mode = 1 to stop
 * paths at vertex 2 and add throughput * max(MLP, 0) after its NEE. */
int srt_mlp_train_frame(srt_ctx* ctx, srt_scene scene, const srt_camera* camera, const srt_pt_params* params,
                        int32_t width, int32_t height, int64_t max_samples, int32_t steps, float* loss,
                        int64_t* samples);
/* checkpoint: weights + Adam state + step + bounds (the reference saves weights
 * only, neural_radiosity.cu:54-72) */
int srt_mlp_save(srt_ctx* ctx, const char* path);
int srt_mlp_load(srt_ctx* ctx, const char* path);

/* ---- device buffers owned by the context (the reference's GPUMemory<T> for the
 *      Film, src/passes/film.h:7-20); budget-checked ----
 * kind: 0 host->device, 1 device->host, 2 device->device */
int srt_device_alloc(srt_ctx* ctx, size_t bytes, void** out);
int srt_device_free(srt_ctx* ctx, void* ptr);
int srt_memcpy(srt_ctx* ctx, void* dst, const void* src, size_t bytes, int32_t kind);

/* ---- serving (no counterpart in the reference, whose renderers render one
 *      frame per loop iteration synchronously, src/renderer/renderer.cu:176-275) ----
 * srt_serve_start starts a dispatcher thread that owns the context's stream and
 * path pool until srt_serve_stop.  Requests are admitted first-come first-served
 * while their device film (28 bytes per pixel) fits in the HBM budget (a request
 * that can never fit ends with status SRT_E_OOM).  Each chunk of at most
 * `path_pool` paths (0: the smaller of the "path_pool" option and 2^24) runs
 * the scene of the admitted request with the smallest key, remaining paths at
 * 2 Gpaths/s + serve_age / 1000 x submit time (0: shortest remaining first;
 * 1000: earliest due time), filled with that scene's admitted requests in key
 * order; a large request is cut into chunks, and newer small ones overtake it
 * between its chunks until its waiting time outweighs their sizes.
 * Requests with mlp_mode 1 query their scene's MLP (srt_mlp_select), else the
 * context's own.
 * A request's image is bit-identical to srt_render with the same arguments.
 * While serving, only srt_submit, srt_poll, srt_wait, srt_serve_stats,
 * srt_serve_stop, srt_last_error and srt_hbm_used may be called; every other
 * entry point returns SRT_E_STATE.  Scenes must have a BVH when submitted. */
typedef struct srt_request {
    srt_scene scene;
    srt_camera camera;
    srt_pt_params params;      /* mlp_mode 0 or 1 */
    int32_t width, height;
    float* rgba_out;           /* host, width * height * 4 floats; written before the status turns 2 */
} srt_request;

typedef struct srt_result {
    int32_t status;            /* 0 queued, 1 admitted, 2 done, < 0 error code */
    int32_t reserved;
    double submit_ms, admit_ms, done_ms;   /* since srt_serve_start */
    int64_t paths;             /* width * height * spp */
    int64_t chunks;            /* chunks the request took part in */
} srt_result;

int srt_serve_start(srt_ctx* ctx, int64_t path_pool);
/* drain = 1: waits until every queued and admitted request is done; drain = 0:
 * cancels (status SRT_E_STATE) the requests that have no path scheduled yet and
 * waits for the others.  Then stops the thread. */
int srt_serve_stop(srt_ctx* ctx, int32_t drain);
/* thread-safe; copies the request; the ticket stays valid until srt_serve_stop */
int srt_submit(srt_ctx* ctx, const srt_request* request, uint64_t* ticket);
int srt_poll(srt_ctx* ctx, uint64_t ticket, srt_result* out);
/* blocks until the request is done or failed, or timeout_ms passes (< 0: no limit) */
int srt_wait(srt_ctx* ctx, uint64_t ticket, int32_t timeout_ms, srt_result* out);
/* paths served and chunks run since start; requests queued and admitted now */
int srt_serve_stats(srt_ctx* ctx, int64_t* served_paths, int64_t* chunks, int32_t* queued, int32_t* active);

/* ---- statistics (the reference's PROFILE scopes, src/utils/profiler.h:156-219) ----
 * Per-kernel totals since the last reset, filled in order of kernel name.
 * Returns the number of kernels; names are written NUL-separated into names. */
int srt_stats(srt_ctx* ctx, char* names, int32_t names_len, double* ms, int64_t* launches, int64_t* items, int32_t max);
void srt_stats_reset(srt_ctx* ctx);
/* segments traced and shadow rays of the last srt_render */
int srt_last_counts(srt_ctx* ctx, int64_t* segments, int64_t* shadow_rays);
/* node visits and triangle tests of the last srt_render made with the option
 * "count_visits" = 1: out[0] closest-hit node visits, [1] closest-hit triangle
 * tests, [2] shadow node visits, [3] shadow triangle tests (0 otherwise) */
int srt_last_visits(srt_ctx* ctx, int64_t out[4]);

#ifdef __cplusplus
}
#endif
#endif /* SRT_HIP_H */
