/*
 * oracle.h — CPU restatement of the reference path tracer, used ONLY as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * It is test infrastructure: the product (libeleven.so, the CLI) never links,
 * loads or calls it.
 *
 * Every function restates the reference (src/tfg-pathtracer/, read as text)
 * and cites the lines it follows.  Nothing here is copied from the reference.
 * The reference's estimator cannot be built in this image (kernel.cu needs
 * nvcc and cuRAND, and its host headers need the CUDA runtime header, which
 * is absent; stand-ins are not allowed), so the oracle is pinned by closed
 * forms, brute force, invariants, exhaustive checks and exact expectations
 * in tests/test_oracle_*.py.  The reference's vendored MikkTSpace (plain C)
 * does build where it lies: oracle/Makefile compiles it into oracle/_ref/
 * behind mikk_ref_glue.cpp, and tests/test_tangents.py pins the loader's
 * tangents against it.
 *
 * Scene input uses the product's plain-C el_scene_view struct (a POD layout
 * shared with include/eleven.h; no product code is called).
 */
#ifndef ELEVEN_ORACLE_H
#define ELEVEN_ORACLE_H

#include <stdint.h>

#include "../include/eleven.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Tri::hit (Tri.hpp:38-68): returns 1 on hit; t, u, v out. */
int or_tri_hit(const float* v9, const float* o3, const float* d3, float* t, float* u, float* v);

/* Brute-force closest hit over all triangles: min t, ties to the lower index
 * (the build's definition of "closest").  tri = -1 on a miss. */
void or_closest_brute(const float* verts, int n, const float* rays, int nrays, int32_t* tri, float* t, float* u,
                      float* v, int threads);

/* Reference-faithful BVH: complete binary tree of depth `depth` with
 * `bins`-bin SAH splits (BVH.hpp:187-211, 291-330, 373-460, 492-578), 44-byte
 * nodes in DFS order and a triIndices array.  Returns an opaque handle. */
void* or_refbvh_build(const float* verts, int n, int depth, int bins);
void  or_refbvh_free(void* h);
/* Traversal as BVH.hpp:120-157 (left-first, right pushed, right leaf tested
 * first at depth-1, no t culling) with the closest rule above. */
void  or_refbvh_closest(void* h, const float* rays, int nrays, int32_t* tri, float* t, int threads);
/* Any hit with 0 <= t < tmax[i]. */
void  or_refbvh_any(void* h, const float* rays, const float* tmax, int nrays, uint8_t* occ, int threads);
int   or_refbvh_node_count(void* h);

/* Philox4x32-10 (Salmon et al. SC'11): out[4] for counter c[4], key k[2]. */
void or_philox(const uint32_t* c, const uint32_t* k, uint32_t* out);

/* Disney BRDF pieces (Disney.hpp:41-253).  hd: 12 scalars in HitData order
 * (metallic, roughness, clearcoatGloss, clearcoat, anisotropic, eta,
 * transmission, specular, specularTint, sheenTint, subsurface, sheen) then
 * emission[3], albedo[3], normal[3]; `dir` is the incoming ray direction.
 * precision: 0 = FP32 as the reference, 1 = FP64. */
void   or_disney_eval(const double* hd, const double* dir, const double* L, int precision, double* out3);
double or_disney_pdf(const double* hd, const double* dir, const double* L, int precision);
void   or_disney_sample(const double* hd, const double* dir, double r1, double r2, double r3, int precision,
                        double* out3);
void   or_cosine_sample_hemisphere(double u1, double u2, double* out3);
void   or_importance_sample_ggx(double rgh, double r1, double r2, double* out3);
void   or_uniform_circle(double u1, double u2, double u3, double* x, double* y);

/* HDRI (HDRI.hpp:107-162): CDF and radiance sum at construction, the
 * reference binarySearch, and the pdf of pixel (x, y). */
void  or_hdri_cdf(const float* rgb, int w, int h, float* cdf, float* radiance_sum);
int   or_hdri_search(const float* cdf, float r, int length);
float or_hdri_pdf(const float* rgb, int w, int h, float x_offset, float radiance_sum, int x, int y);

/* calculateCameraRay (kernel.cu:260-337). */
void or_camera_ray(const el_camera* cam, int x, int y, const float* r5, float* o3, float* d3);

/* Full estimator (kernel.cu:369-481), HDRI pixel by the reference's
 * binarySearch over the CDF (HDRI.hpp:130-162).
 *   precision   0: FP32 as the reference, 1: FP64 shading and accumulation
 *   shadow_mode 0: reference point-light visibility (closest hit, smooth
 *               shading position closer than the light); 1: any hit with
 *               t < dist - 0.001 (the build's definition)
 * Pixels [x0,x1) x [y0,y1) (image coordinates as the kernel's x, y), samples
 * [s0, s0+spp).  Outputs per output pixel idx = W*(H-y-1)+x of the full image:
 * sum[3] and sumsq[3] (double) of accepted clamped samples, count, and the
 * tallies rays / vertices over all paths. */
typedef struct {
    uint32_t spp, s0, depth;
    uint64_t seed;
    int precision, shadow_mode, threads;
    int x0, x1, y0, y1;
    /* NULL, or materials whose BRDF values DisneyEval uses while sampling,
     * pdfs and MIS weights keep the scene's: the detached-sampling estimator
     * whose derivative or_adjoint computes (test use only). */
    const el_material* eval_materials;
    /* NULL, or [W*H][9] (output pixel idx): sums over the accepted samples of
     * the first hit's normal, tangent and bitangent (the NORMAL / TANGENT /
     * BITANGENT passes, kernel.cu:435-440, 451-476). */
    double* aov;
} or_render_params;

int or_render(const el_scene_view* scene, void* refbvh, const or_render_params* p, double* sum, double* sumsq,
              uint32_t* count, uint64_t* rays, uint64_t* vertices);

/* d DisneyEval / d (albedo.x, albedo.y, albedo.z, roughness, metallic,
 * specular) of HitData (effective values), FP64, analytic: out18[3k + c].
 * Not part of the reference (it has no adjoint). */
void or_disney_eval_grad(const double* hd, const double* dir, const double* L, double* out18);

/* Path-replay adjoint of the estimator (detached sampling): gradients of
 * sum_p sum_c dimage[idx][c] * I[idx][c] with respect to every material's
 * (albedo r, g, b, roughness, metallic, specular) - raw values, before the
 * pow 2.2 of generateHitData - added into dmat[M][6], and to the texels of
 * texture `grad_texture` added into dtex[h*w*3] (may be NULL).  I is the mean
 * of the accepted samples s0 .. s0+spp-1; image (optional) receives it.
 * dimage / image are indexed by the reference's idx = W*(H-y-1)+x. */
typedef struct {
    uint32_t spp, s0, depth;
    uint64_t seed;
    int precision, shadow_mode, threads;
    int grad_texture;
} or_adjoint_params;
int or_adjoint(const el_scene_view* scene, void* refbvh, const or_adjoint_params* p, const float* dimage,
               double* dmat, double* dtex, double* image);

#ifdef __cplusplus
}
#endif
#endif
