/*
 * CPU ORACLE for the ray-traversal + path-integration hot path.
 *
 * This is TEST INFRASTRUCTURE. It restates, on the CPU and in plain C++, the
 * algorithms of the reference renderer (IlluminaPRT) that the GPU library
 * replaces, so that tests can check the GPU results against it. Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and
 * only as the checker or as the timed CPU baseline. The product (the HIP
 * library under illuminaprt_amd/) never calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference tree root). Functions whose results no golden vector,
 * closed form, invariant or library routine pins are marked
 * "parity unpinned" here and in the test that uses them.
 */
#ifndef ILM_ORACLE_H
#define ILM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- leaves (known-answer tests) ------------------------------------- */

/* Shirley ray/triangle test, Core/Shape/IndexedTriangle.cpp:92-133 (closest)
 * and :256-289 (any). Returns 1 on hit and fills t, beta (weight of p1),
 * gamma (weight of p2). Ray.Min is ignored, as in the reference. */
int orc_tri_intersect(const float p0[3], const float p1[3], const float p2[3],
                      const float o[3], const float d[3], float ray_max,
                      float* t, float* beta, float* gamma);

/* Slab test, Core/Geometry/BoundingBox.cpp:157-240. Uses the exact 1/d
 * inverse cache (Core/Geometry/Ray.cpp:27-34, Vector3.inl:321-339). */
int orc_aabb_intersect(const float bmin[3], const float bmax[3],
                       const float o[3], const float d[3],
                       float ray_min, float ray_max, float* tin, float* tout);

/* MT19937 with 69069 seeding, Core/Maths/Random.cpp:36-110; NextFloat
 * :117-120. Writes n raw words (or floats) after seeding with `seed`. */
void orc_mt_words(uint32_t seed, uint32_t* out, int n);
void orc_mt_floats(uint32_t seed, float* out, int n);

/* Philox4x64-10 (Salmon et al., SC'11) — the counter-based generator both
 * the oracle ("repro" mode) and the GPU path use. out[4]. */
void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);

/* Warps, Core/Maths/Montecarlo.h:30-37 (cosine hemisphere), :60-64 (triangle). */
void orc_cosine_hemisphere(float u, float v, float out[3]);
void orc_uniform_triangle(float u, float v, float* b1, float* b2);
/* Quad::SamplePoint (Shape/Quad.cpp:146-153) for the quad v1..v4 (quad =
 * 12 floats): point and the normalised winding normal cross(v3-v1, v2-v1). */
void orc_quad_sample(const float quad[12], float u, float v, float point[3], float normal[3]);

/* Camera ray, Core/Camera/PerspectiveCamera.cpp:44-51 with the frame of
 * Camera.cpp:30-44/:72-78, Basis.cpp:28-44 and the half-pixel shift of
 * Filter/BoxFilter.cpp:20-28. cam = {pos[3], target[3], up[3], fov_deg, aspect}. */
void orc_camera_ray(const float cam[11], int width, int height, int px, int py,
                    float u, float v, float o[3], float d[3]);

/* ---- meshes and BVHMesh ---------------------------------------------- */

typedef struct orc_mesh orc_mesh;

/* Triangle mesh (Core/Shape/TriangleMesh.h:27-95): vertex positions and
 * triangle indices; the hit index of a triangle is its position in idx. */
orc_mesh* orc_mesh_create(const float* pos, int64_t nv, const int32_t* idx, int64_t nt);
void orc_mesh_destroy(orc_mesh* m);

/* BVHMesh::Compile, Core/Shape/BVHMesh.cpp:78-107,203-269 with
 * TreeMesh.cpp:44-85,159-212. Returns the node count. */
int64_t orc_bvh_build(orc_mesh* m, int max_leaf, int max_depth);

/* stats: internal, leaves, max depth, refs, max leaf size, empty leaves */
void orc_bvh_stats(const orc_mesh* m, int64_t out[6]);

/* Export nodes (pre-order) for invariant tests: box[6*n] = min xyz, max xyz;
 * child[2*n] (-1 for leaves); leaf_first/leaf_count index into refs. */
void orc_bvh_export(const orc_mesh* m, float* box, int32_t* child,
                    int32_t* leaf_first, int32_t* leaf_count, int32_t* refs);

/* KDTreeMesh (kind 1; Core/Shape/KDTreeMesh.cpp:114-135,941-1005) or
 * KDTreeMeshEx (kind 2; Core/Shape/KDTreeMeshEx.cpp:36-112,118-146) over the
 * same triangles. KDTreeMeshEx takes its depth limit from max_leaf, as the
 * reference's constructors do. Returns the node count (-1 for a bad kind). */
int64_t orc_kd_build(orc_mesh* m, int kind, int max_leaf, int max_depth);
/* nodes, leaves, max depth, triangle references */
void orc_kd_stats(const orc_mesh* m, int64_t out[4]);
/* box[6n], child[2n] (-1 for leaves), split[2n] = (axis, plane), leaf ranges
 * into refs (duplicated triangle references included). */
void orc_kd_export(const orc_mesh* m, float* box, int32_t* child, float* split, int32_t* leaf_first,
                   int32_t* leaf_count, int32_t* refs);

/* Trace n rays; rays are 8 floats each: ox oy oz tmin dx dy dz tmax.
 * mode 0: closest hit (BVHMesh.cpp:114-119,169-198 / BasicMesh.cpp:16-45),
 * mode 1: any hit (BVHMesh.cpp:125-130,143-165 / BasicMesh.cpp:47-64).
 * accel 0: brute force (BasicMesh restatement), 1: BVHMesh restatement,
 * 2: the kd-tree built last by orc_kd_build (KDTreeMesh::Intersect_New,
 * KDTreeMesh.cpp:313-577, or KDTreeMeshEx::Intersects_Stack,
 * KDTreeMeshEx.cpp:170-390).
 * prim[i] = hit triangle (-1 on miss); t and bary (2 floats) may be NULL.
 * nthreads <= 0 uses all hardware threads. Returns the threads used. */
int orc_trace(const orc_mesh* m, const float* rays, int64_t n, int mode, int accel,
              int32_t* prim, float* t, float* bary, int nthreads);

/* ---- spaces ------------------------------------------------------------ */

/* BasicSpace over GeometricPrimitives (Core/Space/BasicSpace.cpp:58-114,
 * Core/Scene/GeometricPrimitive.cpp:68-200). A primitive is a mesh (its
 * BVHMesh restatement is used when built) placed by the reference's
 * Transformation: rot = Matrix3x3 elements _00.._22 in storage order (applied
 * as Matrix3x3::Product, Geometry/Matrix3x3.h:586-597), scl, trn; NULL means
 * identity. Rays are transformed to object space (Ray::ApplyInverse). */
typedef struct orc_space orc_space;
orc_space* orc_space_create(void);
void orc_space_destroy(orc_space* s);
int orc_space_add(orc_space* s, const orc_mesh* m, const float* rot, const float* scl, const float* trn);
/* prim[i] = primitive hit (-1 on miss), tri[i] = triangle within it, t may be NULL. */
int orc_space_trace(const orc_space* s, const float* rays, int64_t n, int mode, int32_t* prim, int32_t* tri,
                    float* t, int nthreads);

/* ---- scenes and the path integrator ---------------------------------- */

typedef struct orc_scene orc_scene;

/* Material record, 8 floats: kd rgb, ks rgb, phong exponent, type
 * (0 = Lambertian (Material/LambertianBxDF.cpp), 1 = Lambertian + Phong lobe;
 * the Phong lobe has no counterpart in the reference). */
/* Light record, 4 floats: Le rgb, sampling mode. The triangles whose
 * tri_light equals the light's index form its emitter: mode 0 samples it as
 * ITriangleMesh (a triangle by index, the normal turned toward the shaded
 * point; Shape/TriangleMesh.cpp:143-161), mode 1 as the reference's Quad
 * from its two triangles (v1, v3, v2), (v1, v4, v3) (orc_quad_sample). */
orc_scene* orc_scene_create(const float* pos, const float* nrm, int64_t nv,
                            const int32_t* idx, const int32_t* tri_mat,
                            const int32_t* tri_light, int64_t nt,
                            const float* mats, int nmat,
                            const float* lights, int nlight);
void orc_scene_destroy(orc_scene* s);

typedef struct {
    float cam[11];      /* pos, target, up, fov_deg, aspect */
    int32_t width, height;
    int32_t spp;        /* samples per pixel rendered by this call */
    int32_t spp_offset; /* first sample index (progressive rendering) */
    int32_t max_depth;  /* PathIntegrator RayDepth */
    int32_t rr_depth;   /* RR applies when depth >= rr_depth (reference: 4) */
    float eps;          /* reflect epsilon (reference factory: 1e-4) */
    uint32_t seed;
    int32_t rng_mode;   /* 0: Philox counter-based, 1: MT19937 per tile */
    int32_t x0, y0, x1, y1; /* pixel window (x1/y1 exclusive; 0 = full) */
    int64_t pix0;       /* added to the pixel index of the counter-based streams
                           (a view's first pixel in a batch of images) */
} orc_render_params;

/* RenderThread-style pool (Utility/Sandbox/Multithreaded.h:28-264) over
 * 32x32 tiles, each pixel rendered as in BaseRenderer::RenderRegion
 * (Core/Renderer/BaseRenderer.cpp:62-137) with PathIntegrator::Radiance
 * (Core/Integrator/PathIntegrator.cpp:44-201,357-367). Adds the radiance
 * SUM over the spp samples of each pixel into accum (width*height*3 doubles,
 * row-major, y = camera row). ray_counts (may be NULL) receives
 * {extension+camera rays, shadow rays}. Returns the threads used. */
int orc_render(const orc_scene* s, const orc_render_params* p, double* accum,
               int64_t ray_counts[2], int nthreads);

/* ICIntegrator on the host (Integrator/ICIntegrator.cpp:306-642): camera
 * samples as orc_render draws them; a hit is shaded with SampleAllLights plus
 * the cache's irradiance x albedo x indirect_factor (the reference: 1/pi,
 * :394); records are computed and inserted on a cache miss, shared by all
 * threads (a reader-writer lock; the reference inserts without one); misses
 * return miss_radiance. counts = {camera rays, gather path rays, shadow
 * rays, records created}. Returns the threads used. */
typedef struct {
    int32_t cache_depth;       /* CacheDepth */
    float error_threshold;     /* ErrorThreshold */
    float ambient_resolution;  /* RMin = longest scene edge x this */
    float ambient_multiplier;  /* RMax = longest scene edge x this */
    int32_t azimuth_strata, altitude_strata;
    int32_t ray_depth;         /* RayDepth of the gather paths */
    float eps;                 /* Epsilon */
    float indirect_factor;
    float miss_radiance[3];
} orc_ic_params;
int orc_ic_render(const orc_scene* s, const orc_render_params* p, const orc_ic_params* icp, double* accum,
                  int64_t counts[4], int nthreads);

/* ---- materials -------------------------------------------------------- */

/* The BSDF of a material record (8 floats: kd rgb, ks rgb, Phong exponent,
 * type 0 Lambertian / 1 Lambertian + Phong) in its local frame (z = normal):
 * f(wo, wi) as BSDF::F (Core/Material/BSDF.cpp:215-237, reflection only) and
 * the pdf BSDF::SampleF reports for a direction it sampled (BSDF.cpp:141-211,
 * LambertianBxDF.cpp:27-55). The Phong lobe (normalised modified Phong,
 * f = ks (n+2)/(2 pi) cos^n a about the mirror direction, sampled by cos^n)
 * has no reference counterpart. n directions each. */
void orc_bsdf_eval(const float mat[8], int64_t n, const float* wo, const float* wi, float* f, float* pdf);
/* BSDF::SampleF for n outgoing directions and (u, v, w) triples: wi, f, pdf. */
void orc_bsdf_sample(const float mat[8], int64_t n, const float* wo, const float* uvw, float* wi, float* f,
                     float* pdf);

#ifdef __cplusplus
}
#endif
#endif
