/*
 * ilm_gpu.h — C ABI of the MI355X ray-traversal and path-integration library.
 *
 * The library replaces the CPU hot path of the IlluminaPRT renderer: the mesh
 * accelerator behind IShape::Intersects, the linear space behind
 * ISpace::Intersects, and the path integrator behind
 * IIntegrator::Radiance / IRenderer::RenderRegion. Each entry point below
 * names the reference interface it replaces (paths relative to the reference
 * tree root). All signatures use plain pointers and sizes.
 *
 * Conventions
 *  - Every call returns ILM_OK (0) or a negative ILM_ERR_* code; nothing
 *    throws across the ABI (the reference throws heap-allocated exceptions:
 *    System/FactoryManager.inl:69,128-129). ilm_last_error() returns the
 *    message of the calling thread's last failure.
 *  - "dev" pointers are device (HBM) addresses, e.g. from ilm_alloc.
 *    "host" pointers are ordinary host memory owned by the caller.
 *  - A context owns its device allocations, its stream and its scratch
 *    buffers. Use one context per host thread, or synchronise externally
 *    (the reference calls Intersects concurrently from tile threads:
 *    Utility/Sandbox/Multithreaded.h:207-263; batch those calls instead).
 *  - stream == NULL means the context's own stream. Calls are asynchronous
 *    on that stream unless documented otherwise.
 *  - Traversal launches of one context share its ray-fetch counters and
 *    stack-spill area: launches given other streams must still be ordered
 *    after one another (same stream, or events). Launches that overlap
 *    would trace some of their rays twice and skip others.
 */
#ifndef ILM_GPU_H
#define ILM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ILM_OK 0
#define ILM_ERR_INVALID (-1)  /* bad argument */
#define ILM_ERR_HIP (-2)      /* HIP runtime failure (message has the HIP error) */
#define ILM_ERR_NOMEM (-3)    /* allocation failed */
#define ILM_ERR_STATE (-4)    /* object not ready (e.g. BVH not built) */
#define ILM_ERR_NODEV (-5)    /* no HIP device */

typedef struct ilm_ctx ilm_ctx;
typedef struct ilm_mesh ilm_mesh;
typedef struct ilm_scene ilm_scene;

/* A ray: 32 bytes, two 16-byte halves. Replaces Geometry/Ray.h:17-54.
 * tmin is kept for the caller but, as in the reference triangle test
 * (Shape/IndexedTriangle.cpp:114-115), hits are accepted for
 * 1e-5 <= t <= tmax - 1e-5 (absolute, t in units of |d|). */
typedef struct {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} ilm_ray;

/* A closest-hit record: 16 bytes. prim = triangle index in the mesh as
 * uploaded (the reference's hit index is the triangle's position in
 * ITriangleMesh::TriangleList, IndexedTriangle.cpp:225 SetShape), -1 on a
 * miss. u, v = barycentric weights of vertices 1 and 2 (the reference's
 * beta, gamma). */
typedef struct {
    float t;
    int32_t prim;
    float u, v;
} ilm_hit;

/* ---------------------------------------------------------------- context */

/* Creates a context on HIP device `device`. No reference counterpart (the
 * reference runs in the host process; EngineKernel, System/EngineKernel.h:34-57,
 * is the closest owner of global state). */
int ilm_ctx_create(ilm_ctx** out, int device);
int ilm_ctx_destroy(ilm_ctx* ctx);
const char* ilm_last_error(void);
/* Number of visible HIP devices (0 when none). */
int ilm_device_count(int* count);
/* Device name and compute-unit count. */
int ilm_device_info(ilm_ctx* ctx, char* name, int name_len, int* cu_count);

/* Free and total device memory of the context's device (hipMemGetInfo);
 * used to report a workload's HBM footprint. */
int ilm_device_memory(ilm_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);

/* Device buffers owned by the context (freed by ilm_ctx_destroy). */
int ilm_alloc(ilm_ctx* ctx, size_t bytes, void** dev);
int ilm_free(ilm_ctx* ctx, void* dev);
int ilm_upload(ilm_ctx* ctx, void* dev, const void* host, size_t bytes);     /* synchronous */
int ilm_download(ilm_ctx* ctx, void* host, const void* dev, size_t bytes);   /* synchronous */
int ilm_memset(ilm_ctx* ctx, void* dev, int value, size_t bytes);            /* synchronous */
/* Waits for the context stream; reports ILM_ERR_STATE if a traversal hit
 * its stack bound since context creation (never expected: the bound is the
 * tree's exact worst case). */
int ilm_sync(ilm_ctx* ctx);
/* Device-to-device streaming copy kernel (bytes a multiple of 16). No
 * reference counterpart: it re-measures HBM bandwidth and calibrates the
 * profiler's byte counters against a transfer of known size. */
int ilm_device_copy(ilm_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);
/* Cache-bandwidth probe: blocks_per_cu workgroups per CU each read the
 * whole buffer (bytes a multiple of 16) `passes` times with 16-byte loads,
 * each workgroup starting 64 KiB after the previous one. A buffer larger
 * than a CU's L1 and smaller than an XCD's L2 (e.g. 2 MiB) measures the
 * aggregate L2 read bandwidth (bytes read = CUs x blocks_per_cu x passes x
 * bytes). sink (16 bytes, device) is written only for an impossible fold of
 * the data. No reference counterpart. */
int ilm_read_probe(ilm_ctx* ctx, const void* src, size_t bytes, int passes, int blocks_per_cu, void* sink,
                   void* stream);
/* Event timing on the context stream: ilm_timer_mark records an event and
 * returns its slot; ilm_timer_elapsed gives milliseconds between two slots
 * (synchronises). Used by benchmarks to time kernels on the stream they
 * are launched on. */
int ilm_timer_mark(ilm_ctx* ctx, int* slot);
int ilm_timer_elapsed(ilm_ctx* ctx, int slot0, int slot1, float* ms);

/* ---------------------------------------------------------------- meshes */

/* Creates a triangle mesh from host arrays: pos = nv*3 floats, idx = nt*3
 * vertex indices. Replaces ITriangleMesh::AddVertex/AddIndexedTriangle
 * (Shape/TriangleMesh.h:41-94). The triangle order defines hit indices.
 * ctx may be NULL: the mesh can then be built and exported on the host but
 * not traced (tracing calls return ILM_ERR_STATE). */
int ilm_mesh_create(ilm_ctx* ctx, const float* pos, int64_t nv,
                    const int32_t* idx, int64_t nt, ilm_mesh** out);
int ilm_mesh_destroy(ilm_mesh* mesh);

#define ILM_BUILDER_SAH 0        /* binned SAH over centroids (default) */
#define ILM_BUILDER_REFERENCE 1  /* the reference BVHMesh topology */
/* The reference kd-tree topologies, traversed as trees of cells whose leaves
 * may share triangles: KDTreeMesh (Shape/KDTreeMesh.cpp:114-135,941-1005)
 * and KDTreeMeshEx (Shape/KDTreeMeshEx.cpp:59-146; its depth limit is
 * max_leaf, as the reference's constructors set it). Hits keep the ray window
 * of every other topology (KDTreeMeshEx's own traversal widens it to the
 * root cell's exit: KDTreeMeshEx.cpp:258-259,371-372). */
#define ILM_BUILDER_KDTREE 2
#define ILM_BUILDER_KDTREE_EX 3
/* Built on the device: a linear BVH over 63-bit Morton codes (Karras 2012),
 * refitted and collapsed to BVH4 on the GPU; max_leaf = largest subtree kept
 * as one leaf (default 4). Needs a mesh created with a context; build_ms is
 * the device time. Faster to build, slower to trace than ILM_BUILDER_SAH. */
#define ILM_BUILDER_LBVH 4

/* Node formats of the BVH4 the traversal kernels read. */
#define ILM_NODES_FLOAT 0      /* 112 B: float child boxes, SoA, then four child words */
#define ILM_NODES_QUANTIZED 1  /* 64 B: child boxes as bytes on a per-node power-of-two grid,
                                  rounded outward (conservative) */

typedef struct {
    int32_t builder;     /* ILM_BUILDER_* */
    int32_t width;       /* children per node: 2, 4 (default when 0) or 8 with float nodes;
                            quantised nodes and the device builder are 4 */
    int32_t max_leaf;    /* SAH, LBVH: max triangles per leaf (default 4);
                            REFERENCE, KDTREE*: MaximumLeafObjects (default 10) */
    int32_t max_depth;   /* REFERENCE, KDTREE: MaximumTreeDepth (default 35, KDTREE at most 40) */
    int32_t threads;     /* host build threads (0 = all) */
    int32_t node_format; /* ILM_NODES_* */
} ilm_bvh_params;

typedef struct {
    int64_t nodes;        /* wide nodes */
    int64_t node_bytes;   /* bytes of the node array in HBM */
    int64_t tri_bytes;    /* bytes of the triangle array in HBM */
    int32_t width;
    int32_t depth;        /* wide-tree depth */
    int32_t max_leaf;     /* largest leaf (triangles) */
    int32_t stack_need;   /* worst-case traversal stack entries */
    double build_ms;      /* host build + flatten time */
    double sah_cost;      /* SAH cost of the wide tree (Ci = 1, Ct = 1) */
    int32_t node_format;  /* ILM_NODES_* */
    int32_t node_stride;  /* bytes per node */
} ilm_bvh_info;

/* Builds the hierarchy on the host, flattens it to a wide BVH and uploads
 * node and triangle arrays. Replaces BVHMesh::Compile
 * (Shape/BVHMesh.cpp:78-107,203-269) and ITreeMesh's split helpers
 * (Shape/TreeMesh.cpp:44-212). Synchronous. */
int ilm_bvh_build(ilm_mesh* mesh, const ilm_bvh_params* params);
int ilm_bvh_get_info(ilm_mesh* mesh, ilm_bvh_info* info);

/* Host copy of the flattened tree for validation: nodes as in HBM
 * (node_bytes), and the original triangle index of each slot of the
 * reordered triangle array (tri_bytes / 48 int32: ntri, or more when kd-tree
 * leaves share triangles). Either pointer may be NULL. */
int ilm_bvh_export(ilm_mesh* mesh, void* nodes, int32_t* tri_order);

/* ---------------------------------------------------------------- tracing */

/* Sets how many idle lanes (1..64, default 24) a traversal wave waits for
 * before it fetches new rays for them: fewer means more refills (one atomic
 * and a divergent ray load each), more means idle lanes for longer. Applies
 * to later traversal launches of the context. No reference counterpart. */
int ilm_ctx_set_trace_refill(ilm_ctx* ctx, int idle_lanes);

/* Closest hit for n rays. Replaces IShape::Intersects(const Ray&,
 * DifferentialSurface&) of BVHMesh (Shape/BVHMesh.cpp:114-119,169-198) and
 * ISpace::Intersects(const Ray&, Intersection&) (Space/BasicSpace.cpp:58-73),
 * batched. rays and hits are dev pointers. */
int ilm_trace_closest(ilm_mesh* mesh, const ilm_ray* rays, int64_t n, ilm_hit* hits, void* stream);

/* Any hit (occlusion) for n rays: occluded[i] = 1 if some triangle is hit
 * in [1e-5, tmax - 1e-5]. Replaces IShape::Intersects(const Ray&)
 * (Shape/BVHMesh.cpp:125-130,143-165) as used by VisibilityQuery::IsOccluded
 * (Scene/Visibility.cpp:33-37). */
int ilm_trace_any(ilm_mesh* mesh, const ilm_ray* rays, int64_t n, uint8_t* occluded, void* stream);

/* ilm_trace_closest that allocates nothing: for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops). No
 * reference counterpart. Same arguments and results as ilm_trace_closest;
 * the ray-fetch counters and the traversal-stack spill area live in a fixed
 * 64 MiB device array per GPU (held until the process ends) instead of the
 * context's scratch, which ilm_trace_closest grows with hipMalloc. The
 * persistent grid shrinks to what that spill area holds; a tree whose stack
 * need (ilm_bvh_info.stack_need) leaves fewer than one workgroup per CU is
 * rejected with ILM_ERR_INVALID. Every call on a GPU shares its array: calls
 * that overlap on one GPU (different streams, threads or contexts) race. */
int ilm_trace_closest_noalloc(ilm_mesh* mesh, const ilm_ray* rays, int64_t n, ilm_hit* hits, void* stream);

/* Host-side traversal statistics (no GPU needed, no hits produced): per ray
 * (visits, host, 2 int32 each) the node visits and triangle tests of a
 * deterministic ordered traversal of the mesh's flattened tree (float or
 * quantised nodes, each box decided by the kernels' own formula), the
 * traversal kernels' without speculation (nearest child first by the
 * kernels' child keys and exchange network, a closest hit shrinks
 * the window, an any hit ends the ray). rays are host memory. threads 0 =
 * all host cores. Prices a launch by algorithmic bytes per ray
 * (32 + 16 + visits * node bytes + tests * 48); the measured counterpart of
 * the speculative kernels is ilm_trace_stats. No reference counterpart
 * (the reference's BVHMesh traversal is unordered, BVHMesh.cpp:169-198). */
int ilm_bvh_visit_counts(ilm_mesh* mesh, const ilm_ray* rays, int64_t n, int any_hit, int32_t* visits,
                         int threads);
/* Same traversal with per-ray counters: visits[2*i] = wide nodes visited,
 * visits[2*i+1] = triangles tested. Used to state algorithmic bytes. */
int ilm_trace_stats(ilm_mesh* mesh, const ilm_ray* rays, int64_t n, int any_hit,
                    int32_t* visits, void* stream);

/* ---------------------------------------------------------------- spaces */

/* A space holds the primitives of a scene: a triangle shape placed by the
 * reference's Transformation, with a material index and a light index (-1:
 * not emissive). Building flattens every primitive to world space into one
 * mesh with one BVH, which ilm_trace_closest/any traverse in a single launch.
 * Replaces BasicSpace::Intersects (Space/BasicSpace.cpp:58-114), a linear loop
 * that transforms each ray into each primitive's object space
 * (GeometricPrimitive::Intersects, Scene/GeometricPrimitive.cpp:68-200; the
 * ray parameter t is the same in both spaces, so hits keep the reference's
 * t window). The exclude-primitive overloads (BasicSpace.cpp:75-114) are not
 * provided: the integrator on this path does not use them
 * (Integrator.cpp:124). */
typedef struct ilm_space ilm_space;

/* ctx may be NULL: the space can then be flattened and built on the host
 * (for export) but not traced or resolved. */
int ilm_space_create(ilm_ctx* ctx, ilm_space** out);
int ilm_space_destroy(ilm_space* space);
/* Adds a primitive (GeometricPrimitive: shape + WorldTransform + material,
 * EmissivePrimitive when light >= 0). Arrays are copied. nrm = object-space
 * vertex normals or NULL (area-weighted, as ITriangleMesh::UpdateNormals).
 * rotation = the 9 elements of the reference's Matrix3x3 in storage order,
 * applied as Matrix3x3::Product (Geometry/Matrix3x3.h:586-597), i.e.
 * world_i = sum_j rotation[3j+i] * (scaling_j * p_j) + translation_i
 * (Transformation::Apply, Geometry/Transform.cpp:133-145); NULL = identity.
 * Normals are scaled and rotated (Transformation::Rotate, :178-194) and
 * normalised. *prim_id (may be NULL) receives the primitive's index. */
int ilm_space_add(ilm_space* space, const float* pos, int64_t nv, const int32_t* idx, int64_t nt,
                  const float* nrm, const float* rotation, const float* scaling, const float* translation,
                  int32_t material, int32_t light, int32_t* prim_id);
/* Flattens and builds (ISpace::Build). Synchronous. */
int ilm_space_build(ilm_space* space, const ilm_bvh_params* params);
/* The flattened world-space mesh (owned by the space, valid until the next
 * build or destroy); hits report global triangle indices. */
int ilm_space_mesh(ilm_space* space, ilm_mesh** mesh);
/* Primitive count and flattened vertex / triangle counts. */
int ilm_space_size(ilm_space* space, int64_t* nprims, int64_t* nv, int64_t* nt);
/* Host copies of the flattened shading arrays, ready for ilm_scene_create:
 * world vertex normals (nv*3), per-triangle material and light (nt each).
 * Any pointer may be NULL. */
int ilm_space_shading(ilm_space* space, float* nrm, int32_t* tri_mat, int32_t* tri_light);
/* Host copies of the flattened world positions (nv*3) and triangle indices
 * (nt*3, global vertex indices). Either pointer may be NULL. */
int ilm_space_geometry(ilm_space* space, float* pos, int32_t* idx);
/* prim[i], local[i] = primitive and triangle within it of hits[i] (-1, -1 on
 * a miss), as Intersection::GetPrimitive and the surface's triangle report
 * them. hits, prim and local are dev pointers. */
int ilm_space_resolve(ilm_space* space, const ilm_hit* hits, int64_t n, int32_t* prim, int32_t* local,
                      void* stream);

/* ---------------------------------------------------------------- instances */

/* The flat instance table: a two-level BVH over shared meshes, the
 * instancing counterpart of a space. Each mesh's tree and triangles are
 * stored once however many instances place it; an instance is a record
 * (world-to-object transform, its mesh's root node, its index) under a
 * top-level BVH4 over the instances' world boxes. Replaces BasicSpace's
 * linear loop over primitives (Space/BasicSpace.cpp:58-114), each testing
 * the ray transformed into its object space (GeometricPrimitive::Intersects,
 * Scene/GeometricPrimitive.cpp:68-200; Ray::ApplyInverse, Geometry/Ray.cpp:
 * 50-60): the object-space direction is not normalised, so t is the same in
 * both spaces and hits keep the reference's t window. */
typedef struct ilm_instances ilm_instances;
typedef struct {
    int32_t mesh;     /* index into the meshes of ilm_instances_create */
    float rot[9];     /* Matrix3x3 in storage order, applied as in ilm_space_add */
    float scl[3];     /* scaling (non-zero) */
    float trn[3];     /* translation */
} ilm_instance;
/* meshes: built meshes with float BVH4 nodes (ILM_NODES_FLOAT, width 4),
 * whose trees and triangles are copied to the table's device memory once;
 * the meshes may be destroyed afterwards. */
int ilm_instances_create(ilm_ctx* ctx, ilm_mesh* const* meshes, int nmeshes, ilm_instances** out);
int ilm_instances_destroy(ilm_instances* table);
/* Sets the instance table (or replaces it: ISpace::Update when transforms
 * change, as the reference's frame loop calls it, Utility/Sandbox/
 * Multithreaded.h:399) and rebuilds the top-level tree on the host; the
 * meshes' trees are not touched. Synchronous. */
int ilm_instances_set(ilm_instances* table, const ilm_instance* inst, int64_t n);
/* Closest hit of n rays (dev pointers): hits[i] = (t, triangle index within
 * the instance's mesh, u, v) and inst[i] = the instance's index in the last
 * ilm_instances_set (-1 on a miss), as BasicSpace::Intersects reports the
 * primitive and its surface. The launch uses the context's ray-fetch
 * counters and spill area, like ilm_trace_closest: a context's traversal
 * launches must be ordered on one stream. */
int ilm_instances_trace_closest(ilm_instances* table, const ilm_ray* rays, int64_t n, ilm_hit* hits, int32_t* inst,
                                void* stream);
/* Any hit (BasicSpace::Intersects(const Ray&), BasicSpace.cpp:92-101). */
int ilm_instances_trace_any(ilm_instances* table, const ilm_ray* rays, int64_t n, uint8_t* occluded, void* stream);
/* Instance count and device bytes: meshes' trees and triangles (stored
 * once), instance records (64 B each), top-level tree; the stack entries a
 * ray may need. Any pointer may be NULL. */
int ilm_instances_info(ilm_instances* table, int64_t* ninst, int64_t* mesh_bytes, int64_t* table_bytes,
                       int64_t* tlas_bytes, int32_t* stack_need);
/* Waits for the context stream and reports ILM_ERR_STATE (once) if a
 * traversal on the table's context exceeded its stack bound (never expected:
 * the spill area is sized from the table's stack need), as ilm_sync does. */
int ilm_instances_check(ilm_instances* table);

/* ---------------------------------------------------------------- ray sets */

/* Camera description: PerspectiveCamera with the factory's frame
 * (Camera/CameraFactories.h:40-78, Camera/Camera.cpp:30-44). */
typedef struct {
    float pos[3], target[3], up[3];
    float fov_deg, aspect;
} ilm_camera;

/* Camera rays for a width x height image, one per pixel, ordered row-major
 * (y = camera row as in BaseRenderer::RenderRegion, BaseRenderer.cpp:92-118).
 * jitter = 0: the pixel's own sample position (0.5, 0.5); jitter = 1:
 * counter-based sample (seed, pixel, sample_index). Replaces
 * PerspectiveCamera::GetRay (Camera/PerspectiveCamera.cpp:53-60). */
int ilm_raygen_camera(ilm_ctx* ctx, const ilm_camera* cam, int width, int height,
                      int jitter, uint32_t seed, uint32_t sample_index, ilm_ray* rays, void* stream);

/* out[i] = camera ray through pixel pixels[i % npix] (pixel = y*width + x):
 * the first npix rays use the pixel's own sample position (0.5, 0.5),
 * later ones a counter-based jittered sample (seed, pixel, i / npix). Used
 * to draw a primary ray set made only of rays that hit geometry. */
int ilm_raygen_camera_list(ilm_ctx* ctx, const ilm_camera* cam, int width, int height,
                           const int32_t* pixels, int64_t npix, int64_t n, uint32_t seed,
                           ilm_ray* rays, void* stream);

#define ILM_SPAWN_SHADOW 0   /* segment to a point on a parallelogram light */
#define ILM_SPAWN_COSINE 1   /* cosine-distributed bounce about the geometric normal */

/* Secondary rays from closest hits: rays_out[i] starts at the hit of
 * rays_in[src[i]] (src may be NULL: identity). Shadow rays follow
 * DiffuseAreaLight::SampleRadiance + VisibilityQuery::SetSegment
 * (Light/DiffuseAreaLight.cpp:52-69, Scene/Visibility.cpp:50-53): origin =
 * hit point, direction = light point - origin, tmin 1e-4, tmax 1 - 1e-4;
 * light = {corner[3], edge_u[3], edge_v[3]}. Cosine rays follow
 * PathIntegrator::Radiance's bounce (Integrator/PathIntegrator.cpp:138-187):
 * origin offset by eps along the geometric normal (faced toward the
 * incoming ray), tmin eps. */
int ilm_raygen_secondary(ilm_mesh* mesh, const ilm_ray* rays_in, const ilm_hit* hits,
                         const int32_t* src, int64_t n, int kind, const float* light,
                         float eps, uint32_t seed, ilm_ray* rays_out, void* stream);

/* Stream compaction of rays whose hit is valid: writes indices of hits with
 * prim >= 0 to idx_out and their count to *count_dev (dev int64). Built on
 * wavefront ballot + prefix counts. */
int ilm_compact_hits(ilm_ctx* ctx, const ilm_hit* hits, int64_t n, int32_t* idx_out,
                     int64_t* count_dev, void* stream);

/* ---------------------------------------------------------------- scene generators */

/* Deterministic procedural scenes (host arrays; sizes from the *_size
 * queries). No reference counterpart: the reference reads .ilm scripts
 * (Scene/EnvironmentLoader.cpp:136-931); these generators build the
 * benchmark scenes procedurally. */
int ilm_scenegen_icosphere_size(int level, int64_t* nv, int64_t* nt);
/* Icosphere of the given subdivision level, displaced radially by
 * octaves of seeded value noise: r = 1 + amplitude * noise(p). */
int ilm_scenegen_icosphere(int level, float amplitude, int octaves, uint32_t seed,
                           float* pos, int32_t* idx);
/* n boxes (24 vertices, 12 triangles each, outward winding), edge lengths
 * uniform in [emin, emax], uniform random rotation, centres uniform in a
 * ball of radius `radius`. Appends: vertex/triangle offsets are given. */
int ilm_scenegen_boxes(int n, float emin, float emax, float radius, uint32_t seed,
                       float* pos, int32_t* idx, int64_t vertex_offset);
/* Area-weighted vertex normals (ITriangleMesh::UpdateNormals,
 * Shape/TriangleMesh.cpp:96-133). */
int ilm_scenegen_vertex_normals(const float* pos, int64_t nv, const int32_t* idx, int64_t nt,
                                float* nrm);

/* ---------------------------------------------------------------- PersistentMesh files */

/* The reference's on-disk mesh and kd-tree (Shape/PersistentMesh.h:22-36,
 * PersistentMesh.cpp:77-89, 171-245; written by KDTreeMesh's export,
 * KDTreeMesh.cpp:143-257): <trunk>.ver (48-byte Vertex records: position,
 * normal, uv), <trunk>.tri (48-byte records: three vertex ids, group id, two
 * edges), <trunk>.tre (kd-tree nodes: float partition, uint32 axis | count
 * << 2, then the children's byte offsets or the leaf's triangle indices).
 * Load validates all three files on the host (offsets inside the file and
 * 4-byte aligned, every node reached once, indices in range, at most 64
 * levels). */
typedef struct ilm_persistent ilm_persistent;
int ilm_persistent_load(const char* trunk, ilm_persistent** out);
/* Vertex and triangle counts, the tree's node count (as a binary tree of
 * cells) and its triangle references. Any pointer may be NULL. */
int ilm_persistent_counts(ilm_persistent* p, int64_t* nv, int64_t* nt, int64_t* tree_nodes, int64_t* tree_refs);
/* Host copies: positions and normals (nv*3), triangle vertex ids (nt*3) and
 * group ids (nt, the material group of each triangle). Any pointer may be NULL. */
int ilm_persistent_arrays(ilm_persistent* p, float* pos, float* nrm, int32_t* idx, int32_t* group);
/* Gives mesh (created from this file's positions and indices, with or
 * without a context) the file's kd-tree as its BVH: the cells collapsed to
 * BVH4 nodes of the given ILM_NODES_* format, as ILM_BUILDER_KDTREE trees
 * are, instead of building one. */
int ilm_persistent_build(ilm_persistent* p, ilm_mesh* mesh, int node_format);
int ilm_persistent_destroy(ilm_persistent* p);

/* ---------------------------------------------------------------- path integrator */

/* Material: 8 floats — kd rgb, ks rgb, phong exponent, type (0 = Lambertian
 * MatteMaterial, Material/Matte.h:16-65; 1 = Lambertian + normalised Phong
 * lobe, which has no reference counterpart). Light: 4 floats — Le rgb,
 * sampling mode; its emitter is every triangle whose tri_light equals its
 * index (DiffuseAreaLight, Light/DiffuseAreaLight.cpp:17-149). Mode 0: the
 * emitter is sampled as ITriangleMesh::SamplePoint does (a triangle chosen
 * uniformly by index, a uniform point on it, the normal turned toward the
 * shaded point; Shape/TriangleMesh.cpp:143-161). Mode 1: the emitter is the
 * reference's Quad, exactly two triangles (v1, v3, v2) and (v1, v4, v3) in
 * that order (Shape/Quad.cpp:11-31), sampled by Quad::SamplePoint's map
 * v1 + (v3-v1) u + (v4-v3) v + (v1-v3+v2-v4) u v with the one-sided winding
 * normal cross(v3-v1, v2-v1) (Quad.cpp:146-153); the map does not cover the
 * quad unless it is degenerate, which the reference's renders share. */
int ilm_scene_create(ilm_ctx* ctx, ilm_mesh* mesh, const float* nrm,
                     const int32_t* tri_mat, const int32_t* tri_light,
                     const float* mats, int nmat, const float* lights, int nlight,
                     ilm_scene** out);
int ilm_scene_destroy(ilm_scene* scene);

typedef struct {
    ilm_camera cam;
    int32_t width, height;
    int32_t spp;          /* samples per pixel in this call */
    int32_t spp_offset;   /* first sample index (progressive rendering) */
    int32_t max_depth;    /* PathTracing RayDepth (IntegratorFactories.h:43-52) */
    int32_t rr_depth;     /* roulette when depth >= rr_depth (reference: 4) */
    float eps;            /* PathTracing Epsilon (reference: 1e-4) */
    uint32_t seed;
    int32_t batch_paths;  /* paths per wavefront batch (0 = auto) */
} ilm_render_params;

typedef struct {
    int64_t camera_rays, extension_rays, shadow_rays;
    double ms;            /* GPU time of the call (events on the stream) */
    int64_t train_records; /* cache frames: training records made (training paths that reached
                              the query vertex); 0 otherwise */
} ilm_render_stats;

/* Renders spp samples per pixel with the wavefront path integrator and
 * ADDS the per-pixel radiance sums into accum (dev, width*height*3 doubles).
 * Replaces PathIntegrator::Radiance (Integrator/PathIntegrator.cpp:44-367)
 * driven by BaseRenderer::RenderRegion (Renderer/BaseRenderer.cpp:62-137)
 * over the RenderThread tile pool (Sandbox/Multithreaded.h:28-264).
 * Synchronous; stats may be NULL. */
int ilm_render(ilm_scene* scene, const ilm_render_params* params, double* accum,
               ilm_render_stats* stats);

/* enable = 1: ilm_render / ilm_render_nrc capture a frame's launch sequence
 * (generate, the bounce loop's trace / shade / shadow launches, cache
 * inference, accumulation) in a HIP graph on the first frame with given
 * settings, and replay it for later frames with the same settings (image
 * size, spp, depths, accumulator, cache and tree); only the seed and sample
 * offset may change between replays. The cache's training steps run
 * outside the graph. No reference counterpart: it replaces the per-tile
 * thread loop of RenderThread (Sandbox/Multithreaded.h:207-263). */
int ilm_scene_set_graph(ilm_scene* scene, int enable);

/* ---------------------------------------------------------------- scene files */

/* Wavefront OBJ (+ MTL) loading with the reference loader's rules
 * (Scene/WavefrontSceneLoader.cpp:203-561): v / vn / vt / f / usemtl /
 * mtllib, newmtl / Kd / Ks / Ke / Ns; faces of 3 or 4 corners (quads split
 * (0,1,2),(0,2,3), larger polygons skipped); corners "p", "p/t/n", "p//n"
 * (a two-field corner is position/normal); equal (p, t, n) corners share a
 * vertex. Vertices without a normal get area-weighted normals. Materials
 * become 8-float records (kd, ks, Ns, type 1 if ks is not black); a material
 * with non-black Ke is also a light with Le = Ke. The arrays feed
 * ilm_mesh_create / ilm_space_add / ilm_scene_create directly. */
typedef struct ilm_obj ilm_obj;
int ilm_obj_load(const char* path, ilm_obj** out);
int ilm_obj_counts(ilm_obj* obj, int64_t* nv, int64_t* nt, int32_t* nmat, int32_t* nlight);
/* Host copies: pos, nrm (nv*3), idx (nt*3), tri_mat, tri_light (nt),
 * mats (nmat*8), lights (nlight*4). Any pointer may be NULL. */
int ilm_obj_arrays(ilm_obj* obj, float* pos, float* nrm, int32_t* idx, int32_t* tri_mat, int32_t* tri_light,
                   float* mats, float* lights);
/* Per material (nmat entries): the MTL illumination model ("illum", 2 when
 * absent; the reference reads 4 as glass and 5 as mirror,
 * Scene/WavefrontSceneLoader.cpp:214-228), and the file name of the first
 * "mtllib" (empty when none; the reference names the material group after
 * it, :306-307). Either pointer may be NULL. */
int ilm_obj_material_info(ilm_obj* obj, int32_t* illum, char* mtllib, int mtllib_len);
int ilm_obj_destroy(ilm_obj* obj);

/* ---------------------------------------------------------------- frames and images */

/* Mean radiance per pixel from an FP64 accumulator (width*height*3 doubles,
 * dev): out (dev, width*height*3 floats) = accum * inv_samples, as
 * AccumulationBuffer::Apply averages frames (Postproc/AccumulationBuffer.h:59-88).
 * flip = 1 writes frame pixel (x, y) to (width-1-x, height-1-y), the mirrored
 * order ImageDevice writes to its image (Device/ImageDevice.cpp:137-148). */
int ilm_image_resolve(ilm_ctx* ctx, const double* accum, int width, int height, double inv_samples, int flip,
                      float* out, void* stream);

#define ILM_TONE_GLOBAL 0  /* GlobalTone: c / (c + 1) per channel (Postproc/GlobalTone.h:37-53) */
#define ILM_TONE_DRAGO 1   /* DragoTone with its fixed bias 0.8 and scale 95%, per-channel maximum
                              over the image, clamped to [0, 1] (Postproc/DragoTone.h:37-80) */
/* Tone maps width*height RGB floats (dev) into out (dev; may equal in). */
int ilm_image_tonemap(ilm_ctx* ctx, const float* in, int width, int height, int op, float* out, void* stream);

#define ILM_IMAGE_PFM 0  /* "PF w h -1 " + little-endian RGB floats (Image/ImagePFM.h:89-111) */
#define ILM_IMAGE_PPM 1  /* "P6 w h 255 " + bytes min(255, (int)(256 v)) (Image/ImagePPM.h:85-110);
                            negative values write 0 */
/* Writes host RGB floats (rows in the given order) to path. */
int ilm_image_write(const char* path, const float* rgb, int width, int height, int format);

/* ---------------------------------------------------------------- batches of scenes */

/* Several scenes rendered by one frame's launch sequence: rays of every
 * scene share the traversal launches (each starts at its own scene's BVH root
 * in one packed node array) and paths of every view share the shading
 * launches. Replaces running one RenderThread pool per scene
 * (Sandbox/Multithreaded.h:336-534); no reference counterpart otherwise.
 * The scenes' meshes must be built with ILM_NODES_FLOAT (the batch also keeps
 * a quantised copy of the packed nodes). The scenes stay owned by the caller
 * and may be destroyed after the batch is created. */
typedef struct ilm_batch ilm_batch;
int ilm_batch_create(ilm_ctx* ctx, ilm_scene* const* scenes, int nscenes, ilm_batch** out);
int ilm_batch_destroy(ilm_batch* batch);

typedef struct {
    int32_t scene;        /* index into the scenes given to ilm_batch_create */
    ilm_camera cam;
    int32_t width, height, spp;
} ilm_batch_view;

/* Renders every view and ADDS their radiance sums into accum: the views'
 * images one after another (sum of width*height*3 doubles). params gives
 * max_depth, rr_depth, eps, seed, spp_offset and batch_paths (0 = the tuned or
 * default value); its camera, width, height and spp are not used. */
int ilm_render_batch(ilm_batch* batch, const ilm_batch_view* views, int nviews, const ilm_render_params* params,
                     double* accum, ilm_render_stats* stats);
/* The batch's packed tree, for validation: the packed mesh (owned by the
 * batch; ilm_bvh_get_info / ilm_bvh_export read its float BVH4 nodes and its
 * triangle slots, whose indices count the scenes' triangles one scene after
 * another) and the root node of each scene (roots: nscenes uint32, may be
 * NULL). ILM_ERR_STATE while the batch uses another node format or width.
 * No reference counterpart. */
int ilm_batch_tree(ilm_batch* batch, ilm_mesh** mesh, uint32_t* roots);
/* As ilm_scene_set_graph, for the batch's frames. */
int ilm_batch_set_graph(ilm_batch* batch, int enable);
/* Fixes the tree of later frames -- node format (ILM_NODES_*) and BVH width
 * (4 when 0; 2 and 8 with float nodes: the batch's BVH4 trees regrouped over
 * the same leaves, built on first use) -- and the paths per wavefront batch
 * (0 = default). */
int ilm_batch_configure(ilm_batch* batch, int node_format, int width, int64_t batch_paths);

typedef struct {
    int32_t node_format;     /* chosen ILM_NODES_* */
    int32_t width;           /* chosen BVH width */
    int64_t batch_paths;     /* chosen paths per wavefront batch */
    double best_ms;          /* its timing frame */
    int32_t candidates;
    int32_t candidate_format[16];
    int32_t candidate_width[16];
    int64_t candidate_batch[16];
    double candidate_ms[16];
} ilm_batch_tuning;

/* Times frames of the given views (at most 2 spp) for each tree (float
 * BVH4, quantised BVH4, float BVH2, float BVH8) and batch size (2^20, 2^22,
 * 2^24, 2^26 paths: the wavefront state of 2^26 paths takes ~27 GiB), keeps
 * the fastest for later frames and reports the table (out may be NULL). */
int ilm_batch_autotune(ilm_batch* batch, const ilm_batch_view* views, int nviews, const ilm_render_params* params,
                       ilm_batch_tuning* out);

/* ---------------------------------------------------------------- irradiance cache */

/* The irradiance cache of ICIntegrator (Integrator/ICIntegrator.cpp:29-668)
 * on the GPU, the comparator of the neural radiance cache: an octree of
 * irradiance records (position, shading normal, irradiance, harmonic mean
 * distance), Ward's weights for lookups, records gathered with stratified
 * cosine-distributed PathLi paths. Zero fields take the ICIntegrator
 * factory's defaults (Integrator/IntegratorFactories.h:159-183). */
typedef struct {
    int32_t cache_depth;       /* CacheDepth (16) */
    float error_threshold;     /* ErrorThreshold (1) */
    float ambient_resolution;  /* AmbientResolution (1e-3): RMin = longest scene edge * this */
    float ambient_multiplier;  /* AmbientMultiplier (1e3): RMax = longest scene edge * this */
    int32_t azimuth_strata;    /* AzimuthStrata (12) */
    int32_t altitude_strata;   /* AltitudeStrata (4) */
    int32_t ray_depth;         /* RayDepth of the gather paths (6) */
    float eps;                 /* Epsilon (1e-4) */
    float indirect_factor;     /* indirect = irradiance * albedo * this; 0 = 1/pi as the reference
                                  (ICIntegrator.cpp:394); 1 makes it the radiance a Lambertian
                                  surface reflects, consistent with the path integrator */
    float miss_radiance[3];    /* camera rays that hit nothing (the reference returns (20, 50, 100),
                                  ICIntegrator.cpp:405) */
    int32_t placement_stride;  /* coarsest pixel stride of the record placement passes (16) */
    int32_t variant;           /* ILM_IC_WARD: ICIntegrator; ILM_IC_TABELION: MLICIntegrator as the
                                  reference compiles it (MLICIntegrator.cpp:28-34): Tabelion weights,
                                  Ri = nearest gather hit, insertion by the half-size along x,
                                  gather rays from P; its ComputeRecord/PathLi are otherwise ICIntegrator's */
} ilm_ic_params;
#define ILM_IC_WARD 0
#define ILM_IC_TABELION 1

typedef struct {
    int64_t records_created, records_total;
    int64_t gather_rays;   /* PathLi paths traced for new records */
    int64_t lookups;       /* samples shaded from the cache (not from their own record) */
    double ms;             /* wall time of the call */
} ilm_ic_stats;

typedef struct ilm_ic ilm_ic;
/* ICIntegrator::Initialise: the octree's root cell from the scene's bounds. */
int ilm_ic_create(ilm_scene* scene, const ilm_ic_params* params, ilm_ic** out);
int ilm_ic_destroy(ilm_ic* ic);
/* Drops every record. */
int ilm_ic_reset(ilm_ic* ic);
/* One frame of ICIntegrator::Radiance for params (camera, size, spp,
 * spp_offset, seed; max_depth is ignored): direct light at the first hit,
 * new records where the cache does not cover a sample (placement passes at
 * pixel strides placement_stride, ..., 2, 1, then every sample), indirect
 * light from the cache. ADDS the samples' radiance sums per pixel to accum.
 * The records stay for later frames. Synchronous. */
int ilm_ic_render(ilm_ic* ic, const ilm_render_params* params, double* accum, ilm_ic_stats* stats);
/* Inserts records given on the host: n x 10 floats (P, N, E, Ri); RiClamp
 * is Ri clamped to [RMin, RMax] (IrradianceCache::Insert). */
int ilm_ic_insert(ilm_ic* ic, const float* records, int64_t n);
/* Record, octree node and node-reference counts, RMin, RMax and the root
 * cell (6 floats: lo, hi). Any pointer may be NULL. */
int ilm_ic_info(ilm_ic* ic, int64_t* records, int64_t* nodes, int64_t* refs, float* rmin, float* rmax,
                float* root_lo_hi);
/* Host copy of the records: n x 11 floats (P, N, E, Ri, RiClamp). */
int ilm_ic_records(ilm_ic* ic, float* out);
/* One octree node: its record count, first child (-1: none; children are 8
 * consecutive nodes in SetBounds order) and up to cap record indices. */
int ilm_ic_node_records(ilm_ic* ic, int64_t node, int32_t* count, int32_t* child, int32_t* records, int32_t cap);
/* Cache lookups at n points (dev, 6 floats each: P, N): out (dev, 4 floats
 * each) = Ward-weighted mean irradiance rgb and the weight sum (0: no
 * record covers the point) (IrradianceCache::FindRecords, ICIntegrator.cpp:142-200). */
int ilm_ic_lookup(ilm_ic* ic, const float* points, int64_t n, float* out, void* stream);

/* ---------------------------------------------------------------- neural radiance cache */

/* A small MLP (96 -> 5 x 64 ReLU -> 16, bf16 weights and activations, FP32
 * accumulation and FP32 master weights, Adam) trained online on radiance
 * produced by the integrator's own paths. No reference counterpart: it
 * stands in for the irradiance cache (Integrator/ICIntegrator.cpp:29-668)
 * at the far end of each path.
 * A query is 16 floats: position xyz, direction (theta/pi, phi/2pi), normal
 * (theta/pi, phi/2pi), roughness, diffuse albedo rgb, specular albedo rgb,
 * two unused. The position is normalised by the bounds given at creation. */
typedef struct ilm_nrc ilm_nrc;

typedef struct {
    float bounds_lo[3], bounds_hi[3];
    float lr;      /* Adam learning rate (default 1e-3 when 0) */
    float beta1;   /* default 0.9 */
    float beta2;   /* default 0.99 */
    float eps;     /* default 1e-8 */
    uint32_t seed; /* weight initialisation */
} ilm_nrc_params;

/* Checks the MFMA operand layout once per process (fails with ILM_ERR_STATE
 * if it differs from the one the kernels assume), then allocates and
 * initialises the network. */
int ilm_nrc_create(ilm_ctx* ctx, const ilm_nrc_params* params, ilm_nrc** out);
int ilm_nrc_destroy(ilm_nrc* nrc);
/* Number of FP32 master weights (row-major per layer, padding included). */
int ilm_nrc_param_count(int* count);
/* Host copies of the FP32 master weights (param_count floats). Setting
 * them resets the Adam moments. */
int ilm_nrc_set_weights(ilm_nrc* nrc, const float* weights);
int ilm_nrc_get_weights(ilm_nrc* nrc, float* weights);
/* Summed loss gradient of the last training step (param_count floats),
 * kept so the backward kernels can be checked against a CPU oracle. */
int ilm_nrc_get_gradients(ilm_nrc* nrc, float* grads);
/* Radiance predictions for n queries (dev): out = n x 4 floats (rgb, 0). */
int ilm_nrc_infer(ilm_nrc* nrc, const float* queries, int64_t n, float* out, void* stream);
/* One Adam step on n (query, target) pairs (dev). targets = n x 4 floats
 * (rgb, weight). Loss: mean over samples and channels of
 * weight * (y - t)^2 / (y^2 + 0.01), the denominator held constant in the
 * gradient. loss (dev, n floats) receives per-sample losses, or NULL.
 * n is at most 2^26. */
int ilm_nrc_train_step(ilm_nrc* nrc, const float* queries, const float* targets, int64_t n,
                       float* loss, void* stream);

typedef struct {
    int32_t prefix_depth;   /* depth of the vertex answered by the cache (default 2, >= 1) */
    int32_t suffix_depth;   /* further bounces of training paths (default 8) */
    float train_fraction;   /* fraction of the frame's paths extended as training paths (default
                               1/16): exactly floor(paths * fraction) of them, spread evenly over
                               the path order with a per-frame offset from the seed; every
                               training path's record is kept, in a slot of its own */
    int32_t train_batch;    /* samples per Adam step (default 65536); the frame's records form
                               ceil(records / train_batch) batches */
    int32_t train_steps;    /* Adam steps per frame (0 = no training, negative = 1); successive
                               steps, also across frames, take the batches in turn */
    int32_t accumulate;     /* add this frame's radiance to accum */
} ilm_nrc_render_params;

/* Renders like ilm_render, but paths that are not training paths stop at
 * prefix_depth and take their remaining radiance from the cache (batched
 * inference at the end of each wavefront batch); training paths (a
 * fixed fraction) are traced prefix_depth + suffix_depth deep, and the
 * radiance they measure leaving their query vertex trains the cache
 * (train_steps Adam steps on train_batch samples each after the frame). The
 * image adds the training paths' own estimates and the cache's answers.
 * Synchronous; accum may be NULL when accumulate is 0. */
int ilm_render_nrc(ilm_scene* scene, ilm_nrc* nrc, const ilm_render_params* params,
                   const ilm_nrc_render_params* nrc_params, double* accum, ilm_render_stats* stats);
/* The cache renderer over a batch's views (as ilm_render_nrc). All views
 * share the one network, so this is meant for several views of one scene
 * (views of scenes that overlap in space would share cache entries). */
int ilm_render_batch_nrc(ilm_batch* batch, ilm_nrc* nrc, const ilm_batch_view* views, int nviews,
                         const ilm_render_params* params, const ilm_nrc_render_params* nrc_params, double* accum,
                         ilm_render_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
