/*
 * eleven.h — C ABI of the MI355X-native path-tracing core.
 *
 * Plain C: opaque handles, POD structs, pointers and sizes.  No C++ or HIP
 * types cross this boundary.  Every entry point returns an int status
 * (EL_OK == 0, negative on error) unless stated otherwise; the message of the
 * last error on a context is returned by el_last_error().
 *
 * Threading: one el_ctx per host thread; a context owns three HIP streams
 * on one device (its launches and the two shadow launches that may run
 * beside them).  No host thread it starts outlives a call: el_upload_scene
 * builds on up to the hardware's thread count, capped by OMP_NUM_THREADS
 * when that is set.  Calls are synchronous on return unless stated
 * otherwise.
 *
 * Reference interfaces replaced (paths relative to src/tfg-pathtracer/):
 *   renderSetup(Scene*)                       kernel.h:80,  kernel.cu:566-661
 *   renderCuda(Scene*, int sampleTarget)      kernel.h:78,  kernel.cu:665-686
 *   getBuffers(RenderData&, int*, int)        kernel.h:82,  kernel.cu:688-710
 *   getSamples()                              kernel.h:84,  kernel.cu:712-724
 *   loadScene(std::string)                    SceneLoader.hpp:7-143
 *   BVH::transverse (closest hit)             BVH.hpp:120-157, kernel.cu:152-173
 */
#ifndef ELEVEN_H
#define ELEVEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define EL_OK          0
#define EL_ERR_ARG    -1   /* bad argument (null pointer, zero size, ...)   */
#define EL_ERR_IO     -2   /* missing / unreadable / corrupt input file     */
#define EL_ERR_HIP    -3   /* HIP runtime error (no device, launch failure) */
#define EL_ERR_STATE  -4   /* call out of order (e.g. render before upload) */
#define EL_ERR_RANGE  -5   /* size beyond a documented limit                */
#define EL_ERR_NOMEM  -6   /* host or device allocation failed              */

/* ---- render modes / flags --------------------------------------------- */
#define EL_MODE_PLAIN      0  /* one thread per pixel, full path per thread  */
#define EL_MODE_WAVEFRONT  1  /* per-bounce ray queues, staged kernels       */

#define EL_FLAG_AOV        1u /* also accumulate NORMAL/TANGENT/BITANGENT    */
#define EL_FLAG_NO_CULL    4u /* trace zero-throughput paths to the end, as
                                 the reference does; by default a path whose
                                 throughput is exactly zero stops when no
                                 remaining bounce can make it NaN (the image
                                 and sample counts are the same either way) */
#define EL_FLAG_COUNT_TRAVERSAL 2u /* wavefront: tally node fetches and
                                      triangle tests of the closest-hit stage */
#define EL_FLAG_RECORD_PATHS 8u /* opt-in, O(pixels x spp x depth) memory:
                                   keep each path's radiance (16 B per path,
                                   when that fits half of the free device
                                   memory) so that el_adjoint walks every
                                   path once instead of twice; needs a cleared
                                   accumulation (the record covers this one
                                   render; el_clear and el_set_camera release
                                   it).  Wavefront renders with max_depth
                                   <= 16 also keep every vertex's hit and
                                   shadow answers (16 B per vertex + 4 B per
                                   path, when that fits a quarter of the free
                                   device memory): el_adjoint then traces no
                                   ray.  The image is the same either way.   */

#define EL_FLAG_SERIAL_SHADOWS 16u /* wavefront: run every shadow launch after
                                     the bounce's shade on the one stream.  By
                                     default the two shadow launches of a
                                     bounce run on two more streams, beside
                                     each other and the next bounce's
                                     closest-hit launch, when the launches are
                                     small (generations under 16 M paths, or
                                     from the third bounce on); the image is
                                     the same either way                       */

#define EL_FLAG_KERNEL_TIMES 32u /* wavefront: time every closest-hit, shade
                                     and shadow launch with events on its
                                     stream (el_stats extend_ms, shade_ms,
                                     shadow_ms; 0 without the flag).  The
                                     events add a few microseconds between
                                     launches: a 128x128 x 32 spp render takes
                                     ~10 % longer with them                   */

/* Passes, same numbering as the reference's `enum Passes` (kernel.h:7). */
#define EL_PASS_BEAUTY     0
#define EL_PASS_DENOISE    1  /* not produced (denoiser is out of scope)     */
#define EL_PASS_NORMAL     2
#define EL_PASS_TANGENT    3
#define EL_PASS_BITANGENT  4

/* ---- scene description (POD mirror of the reference structs) ---------- */
typedef struct { float x, y, z; } el_float3;           /* Vector.hpp:7-12  */

/* Material.hpp:8-37 field for field (name dropped). */
typedef struct {
    int32_t albedo_tex, emission_tex, roughness_tex, metallic_tex,
            normal_tex, opacity_tex;                   /* -1 = none */
    el_float3 albedo, emission, opacity;
    float roughness, metallic, clearcoat_gloss, clearcoat, anisotropic, eta,
          transmission, specular, specular_tint, sheen_tint, subsurface, sheen;
} el_material;

/* Texture.hpp:19-33: FP32 RGB interleaved, data[3*(y*width+x)+c]. */
typedef struct {
    int32_t width, height;
    const float* data;
    float x_tile, y_tile, x_offset, y_offset;
    int32_t filter;                                    /* 0 nearest, 1 bilinear */
} el_texture;

typedef struct { el_float3 position, radiance; } el_point_light; /* PointLight.hpp */

/* Camera.hpp:10-23 */
typedef struct {
    uint32_t x_res, y_res;
    float focal_length, sensor_width, sensor_height, aperture, focus_distance;
    el_float3 rotation;                                /* Euler XYZ, degrees */
    el_float3 position;
    int32_t bokeh;
} el_camera;

/* A whole scene as flat arrays.  The library copies everything it needs
 * during el_upload_scene; the caller keeps ownership of every pointer. */
typedef struct {
    uint32_t tri_count;
    const float*   vertices;      /* [tri_count][3][3]  Tri::vertices        */
    const float*   uvs;           /* [tri_count][3][2]  Tri::uv (x, y)       */
    const float*   normals;       /* [tri_count][3][3]  Tri::normals         */
    const float*   tangents;      /* [tri_count][3][3]  Tri::tangents        */
    const float*   tangent_sign;  /* [tri_count]        Tri::tangentsSign    */
    const int32_t* object_id;     /* [tri_count]        Tri::objectID        */
    uint32_t mesh_count;
    const int32_t* mesh_material; /* [mesh_count]       MeshObject::materialID */
    uint32_t material_count;
    const el_material* materials;
    uint32_t texture_count;
    const el_texture* textures;
    el_texture hdri;              /* environment map (HDRI::texture)          */
    uint32_t light_count;
    const el_point_light* lights;
    el_camera camera;
} el_scene_view;

typedef struct { el_float3 origin, direction; } el_ray; /* Ray.hpp:8-17 */

typedef struct {
    uint32_t spp;          /* samples per pixel added by this call          */
    uint32_t max_depth;    /* MAXBOUNCES (Definitions.h:9); 0 -> 5          */
    uint64_t seed;         /* RNG key                                       */
    int32_t  mode;         /* EL_MODE_*                                     */
    uint32_t block;        /* traversal threads per block: 64..256 step 64,
                              0 -> tuned default                            */
    uint32_t queue_cap;    /* wavefront paths in flight, 0 -> default       */
    uint32_t bvh_width;    /* 4 or 8, 0 -> current                          */
    uint32_t flags;        /* EL_FLAG_*                                     */
    uint32_t refill;       /* wavefront traversal: idle lanes (1..64) that
                              trigger an in-wave ray refill, 0 -> default   */
    uint32_t rng_scene;    /* scene word of the RNG counter (a batch uses
                              the scene's index there), 0 by default        */
} el_render_params;

typedef struct {
    uint64_t rays;             /* BVH queries: camera+extension+shadow       */
    uint64_t vertices;         /* shaded path vertices (reference kPaths)    */
    uint64_t paths;            /* paths traced                               */
    double   render_ms;        /* wall time of the last el_render (device)   */
    double   extend_ms;        /* summed closest-hit kernel time (wavefront,
                                  EL_FLAG_KERNEL_TIMES; shadow_ms, shade_ms
                                  likewise)                                  */
    double   shadow_ms;        /* summed shadow kernel time (wavefront); el_render
                                  runs a bounce's two shadow launches on two
                                  more streams while the next bounce's
                                  closest-hit launch runs, so the times
                                  overlap and their sum exceeds render_ms    */
    double   shade_ms;         /* summed shading kernel time (wavefront)     */
    uint64_t extend_launches;  /* number of closest-hit launches timed       */
    uint64_t extend_rays;      /* rays processed by those launches           */
    uint64_t node_fetches;     /* closest-hit node fetches (COUNT_TRAVERSAL) */
    uint64_t tri_tests;        /* closest-hit triangle tests (COUNT_TRAVERSAL) */
    uint64_t node_bytes;       /* size of the device BVH node array          */
    uint64_t tri_bytes;        /* size of the device triangle stream         */
    uint32_t bvh_width;
    uint32_t stack_bound;      /* worst-case traversal stack depth           */
    uint32_t node_stride;      /* bytes per BVH node as stored (64 or 96)    */
    uint32_t tri_stride;       /* bytes per triangle record as stored (48)   */
    double   adjoint_ms;       /* device time of the last el_adjoint         */
    uint64_t adjoint_rays;     /* BVH queries of the last el_adjoint (primal
                                  walk + replay)                             */
    uint64_t texels_8bit;      /* texels read from the 8-bit copies (textures
                                  from 8-bit files, not since set_texture)   */
    uint64_t adjoint_bytes;    /* device memory the last el_adjoint used beyond
                                  the scene and the accumulation: per-thread
                                  vertex cache, adjoint image, gradients     */
    uint64_t record_bytes;     /* device memory held by the last recording
                                  render (EL_FLAG_RECORD_PATHS), 0 if none   */
} el_stats;

/* Adjoint of the accumulated image (no counterpart in the reference, which
 * renders forward only). */
typedef struct {
    int32_t  grad_texture; /* texture whose texels receive gradients, -1: none */
    uint32_t block;        /* threads per block, 0 -> default               */
    uint32_t flags;        /* EL_ADJ_*                                      */
} el_adjoint_params;
#define EL_ADJ_ACCUMULATE 1u /* add to the gradients of the previous el_adjoint
                                calls (same grad_texture) instead of starting
                                from zero, e.g. to sum the views of one
                                scene; the outputs receive the running sums */
#define EL_ADJ_DEVICE_RESIDUAL 2u /* d_image is not read (may be NULL): the
                                adjoint weights are the residual that the last
                                el_residual_l2 left on the device */

/* ---- host-side scene loading (SceneLoader.hpp:7 loadScene) ------------ */
typedef struct el_scene el_scene;

/* Load <dir>/scene.json, scene.mtl, scene.obj, HDRI/<name>.hdr and the .bmp
 * maps named in scene.mtl.  Accepts '/' and '\\' separators. */
int  el_scene_load(const char* dir, el_scene** out);
/* Flat view of a loaded scene; pointers stay valid until el_scene_free. */
int  el_scene_view_get(const el_scene* scene, el_scene_view* out);
/* HDRI importance-sampling data the reference builds in HDRI::generateCDF
 * (HDRI.hpp:107-128): cdf[w*h+1] and the FP32 radiance sum. */
int  el_scene_hdri_cdf(const el_scene* scene, const float** cdf, float* radiance_sum);
void el_scene_free(el_scene* scene);
/* The renderer's environment-pixel sampler evaluated on the host: for each
 * r[i] in [0, 1), pixel[i] = HDRI::sample's index binarySearch(cdf, r, W*H)
 * (HDRI.hpp:130-162), computed with the same guide-table code the kernels
 * run.  For inspection and tests; el_upload_scene builds the same tables. */
int  el_hdri_lookup(const el_texture* hdri, const float* r, uint32_t n, int32_t* pixel);
const char* el_scene_last_error(void);

/* ---- device context ---------------------------------------------------- */
typedef struct el_ctx el_ctx;

int  el_create(el_ctx** out, int device);
void el_destroy(el_ctx* ctx);
const char* el_last_error(const el_ctx* ctx);
int  el_device_count(int* out);

/* renderSetup: builds the BVH, uploads the scene, clears the accumulation. */
int  el_upload_scene(el_ctx* ctx, const el_scene_view* view);
/* renderCuda: adds params->spp samples per pixel to the accumulation. */
int  el_render(el_ctx* ctx, const el_render_params* params);
/* Clears the accumulation buffers and the sample counter. */
int  el_clear(el_ctx* ctx);
/* getBuffers/getSamples: mean beauty RGBA [H][W][4] (row 0 = top, as the
 * reference's idx = W*(H-y-1)+x), per-pixel accepted sample counts [H][W],
 * and the ray / vertex tallies.  Any output pointer may be NULL. */
int  el_read_passes(el_ctx* ctx, float* rgba, uint32_t* samples,
                    uint64_t* rays, uint64_t* vertices);
/* AOV pass (EL_PASS_NORMAL/TANGENT/BITANGENT) as mean RGBA [H][W][4]. */
int  el_read_aov(el_ctx* ctx, int pass, float* rgba);
int  el_get_stats(el_ctx* ctx, el_stats* out);

/* Replace material parameters on the device without re-uploading geometry
 * (used by the inverse-rendering loop). */
int  el_set_materials(el_ctx* ctx, const el_material* materials, uint32_t count);
/* Replace the texels of texture `tex_id` (same size as uploaded). */
int  el_set_texture(el_ctx* ctx, int32_t tex_id, const float* rgb);
/* Replace the camera (the el_camera fields of Camera.hpp:10-23; the
 * resolution may change) and clear the accumulation, so one uploaded scene
 * is rendered from several views (multi-view inverse rendering).  The
 * reference fixes its camera when the scene is loaded (SceneLoader.hpp:24-45);
 * this entry point has no counterpart there. */
int  el_set_camera(el_ctx* ctx, const el_camera* camera);

/* Path-replay adjoint (Vicini et al. 2021, detached sampling) of the image
 * currently accumulated in the context (every el_render since el_clear,
 * which must share one seed and depth <= 16): gradients of
 *     sum over pixels and channels of d_image * I
 * with respect to each material's (albedo r, g, b, roughness, metallic,
 * specular) - the values stored in el_material, before the pow 2.2 of
 * generateHitData (kernel.cu:103-104) - written to d_materials[M][6], and to
 * the texels of texture params->grad_texture, written to d_texture[h][w][3]
 * (may be NULL if grad_texture < 0).  d_image is [H][W][3] in the layout of
 * el_read_passes.  Each path is replayed from its RNG counters, nothing of a
 * path outlives the thread that replays it; a material whose albedo comes
 * from a texture gets no albedo gradient (its texels do, if selected).
 * Without a recording render each path is walked twice: once traced, for its
 * radiance, keeping the closest hits of its vertices in the walking thread's
 * own slots (a buffer of resident threads x depth x 16 B, independent of the
 * pixel and sample counts), then replayed from those without tracing.  If
 * the render used EL_FLAG_RECORD_PATHS, each path is walked once (the
 * recorded radiance gives the radiance after each vertex), from its recorded
 * vertices when the render kept them.  With `device_ptrs` != 0 the three
 * arrays are device pointers.  d_materials and d_texture may be NULL:
 * nothing is copied out, and the sums stay in the context for following
 * EL_ADJ_ACCUMULATE calls. */
int  el_adjoint(el_ctx* ctx, const el_adjoint_params* params, const float* d_image,
                float* d_materials, float* d_texture, int device_ptrs);

/* L2 residual of the accumulated image on the device (no counterpart in the
 * reference; for inverse rendering): with I the mean image (the rgb of
 * el_read_passes) and target [H][W][3] in host memory, keeps
 * r = 2 (I - target) (computed in FP64, stored as FP32) for a following
 * el_adjoint with EL_ADJ_DEVICE_RESIDUAL and writes sum (I - target)^2 to
 * *loss (may be NULL).  The image never leaves the device.  The residual
 * stays until the next el_residual_l2, el_set_camera or scene upload, so it
 * may be taken from one render and the adjoint run on another (the
 * decorrelated multi-view gradient). */
int  el_residual_l2(el_ctx* ctx, const float* target, double* loss);

/* Closest hit for n rays: tri = original triangle index or -1, t = hit
 * distance along the (caller-normalised) direction, uv = barycentrics
 * (may be NULL).  Closest means minimum t, ties to the lower triangle
 * index.  With `device_ptrs` != 0 all pointers are device pointers and the
 * call is ordered on the context's stream (still synchronous on return). */
int  el_trace_closest(el_ctx* ctx, const el_ray* rays, uint32_t n,
                      int32_t* tri, float* t, float* uv, int device_ptrs);
/* el_trace_closest for rays and outputs in device memory, allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  The traversal's scratch -- its tallies and the per-thread
 * spill area of its stack -- is a static device array of the library,
 * 16 MiB + 512 B on each GPU the library runs on, held until the process
 * ends; the call allocates nothing, waits for its stream before returning
 * (so it cannot be captured into a graph), and its results equal
 * el_trace_closest's with device_ptrs = 1 (uv may be NULL).  The array
 * bounds the resident threads to 16 MiB / (4 B x the scene's spill slots per
 * thread), so large scenes trace with fewer threads than el_trace_closest.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (two contexts on two host threads) race.  Fails with EL_ERR_RANGE when the
 * scene's stack needs more spill slots per thread than the array holds for
 * one block (65,536 at the default 64-thread block). */
int  el_trace_closest_noalloc(el_ctx* ctx, const el_ray* d_rays, uint32_t n,
                              int32_t* d_tri, float* d_t, float* d_uv);
/* Any hit with 0 <= t < tmax[i]: occluded[i] = 1 / 0. */
int  el_trace_any(el_ctx* ctx, const el_ray* rays, const float* tmax,
                  uint32_t n, uint8_t* occluded, int device_ptrs);
/* Node fetches and triangle tests of the closest-hit traversal of n rays
 * (same kernel, counting variant): the algorithmic traffic of a batch. */
int  el_trace_count(el_ctx* ctx, const el_ray* rays, uint32_t n,
                    uint64_t* node_fetches, uint64_t* tri_tests, int device_ptrs);
/* Mean duration (ms) of the closest-hit kernel over the last `iters`
 * launches of el_trace_bench (HIP events on the context's stream). */
int  el_trace_bench(el_ctx* ctx, const el_ray* d_rays, uint32_t n,
                    int32_t* d_tri, float* d_t, int iters, double* mean_ms);

/* Device memory helpers for callers without their own allocator. */
int  el_device_alloc(el_ctx* ctx, size_t bytes, void** out);
int  el_device_free(el_ctx* ctx, void* ptr);
int  el_memcpy_htod(el_ctx* ctx, void* dst, const void* src, size_t bytes);
int  el_memcpy_dtoh(el_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- batches: many scenes through one wavefront loop ------------------
 * No counterpart in the reference (it renders one scene per process).  Each
 * scene keeps its own spp budget, seed and camera; the batch enumerates the
 * work of all scenes, fills each generation of queue_cap paths across scene
 * boundaries and runs the camera / extend / shade / shadow / finalize
 * sequence over the mixed queues.  Scene i of a batch draws exactly the
 * random numbers of el_render with seed = its seed and rng_scene = i. */
typedef struct el_batch el_batch;

#define EL_BATCH_GRAPH 1u   /* capture the launch sequence in a hipGraph once,
                               replay it on later renders with equal plans  */
#define EL_BATCH_NO_CULL 4u /* as EL_FLAG_NO_CULL                            */
#define EL_BATCH_SERIAL_SHADOWS 8u /* as EL_FLAG_SERIAL_SHADOWS (batches overlap
                                      a bounce's shadow launches with the next
                                      extension from the third bounce on, or
                                      always for generations under 16 M paths,
                                      also inside a captured graph)           */

typedef struct {
    uint32_t max_depth;    /* 0 -> 5                                        */
    uint32_t block;        /* traversal threads per block, 0 -> default     */
    uint32_t queue_cap;    /* paths in flight, 0 -> a quarter of free HBM   */
    uint32_t bvh_width;    /* 4 or 8, 0 -> current                          */
    uint32_t refill;       /* in-wave refill threshold, 0 -> default        */
    uint32_t flags;        /* EL_BATCH_*                                    */
} el_batch_params;

int  el_batch_create(el_batch** out, int device);
void el_batch_destroy(el_batch* batch);
const char* el_batch_last_error(const el_batch* batch);
/* Trials and choice of the last el_batch_tune, one line per trial. */
const char* el_batch_log(const el_batch* batch);
/* Uploads a scene (the library copies it) with its spp budget and seed;
 * *index receives its position in the batch. */
int  el_batch_add(el_batch* batch, const el_scene_view* view, uint32_t spp,
                  uint64_t seed, int32_t* index);
/* Renders every scene's spp from cleared accumulations (synchronous). */
int  el_batch_render(el_batch* batch, const el_batch_params* params);
/* Autotuner: coordinate descent over bvh_width, queue_cap, block and refill,
 * each trial rendering every scene at min(spp, trial_spp) samples; a value
 * replaces the current one only if its trial is at least 2 % faster.  Writes
 * the chosen setting into *params (flags are kept). */
int  el_batch_tune(el_batch* batch, el_batch_params* params, uint32_t trial_spp);
/* Mean RGBA [H][W][4] and accepted-sample counts of scene `scene`. */
int  el_batch_read(el_batch* batch, int32_t scene, float* rgba, uint32_t* samples);
int  el_batch_stats(el_batch* batch, el_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* ELEVEN_H */
