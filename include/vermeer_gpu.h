/*
 * vermeer_gpu.h — C ABI of the MI355X (gfx950) ray-traversal library.
 *
 * This is the drop-in boundary for the reference's traversal hot path.
 * Each entry point names the reference interface it replaces (paths are
 * relative to the reference checkout).  The ABI uses plain pointers and
 * sizes only, so it can be bound from cgo, ctypes or C/C++.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative VGPU_ERR_* code on
 *     failure; vgpu_last_error() then describes the failure (per host thread).
 *     Nothing aborts across the boundary (the reference panics instead).
 *   - The caller owns every host buffer.  The library copies what it needs
 *     and keeps no caller pointer after a call returns (cgo pointer rules).
 *   - Device memory is owned by the scene object.
 *   - Calls on one scene (and on the renderers and path tracers made over it)
 *     are serialised by a per-scene lock: host-memory calls may come from many
 *     host threads at once.  Work enqueued with VGPU_TRACE_DEVICE_PTRS shares
 *     the scene's traversal spill area and error flag while it runs, so all
 *     device-pointer work on one scene must be enqueued on one stream.
 */
#ifndef VERMEER_GPU_H
#define VERMEER_GPU_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VGPU_OK 0
#define VGPU_ERR_ARG (-1)      /* invalid argument or size */
#define VGPU_ERR_HIP (-2)      /* HIP runtime error */
#define VGPU_ERR_BUILD (-3)    /* BVH builder failure (degenerate input the reference recurses on forever) */
#define VGPU_ERR_STACK (-4)    /* a ray needed more traversal stack than VGPU_MAX_STACK */
#define VGPU_ERR_STATE (-5)    /* call out of order (e.g. trace before commit) */
#define VGPU_ERR_NOMEM (-6)    /* host or device allocation failed */

/* Traversal stack bound per ray.  The reference holds 90 entries
 * (core/ray.go:158) shared by the scene and mesh levels. */
#define VGPU_MAX_STACK 128

/* Ray record, 32 bytes.  P/D as core.Ray.P/D (core/ray.go:28), tmax is the
 * initial core.Ray.Tclosest (Ray.Init maxdist, core/ray.go:59), time is
 * core.Ray.Time in [0,1] for motion blur (a time outside it is held to the
 * first or last key, where the reference would index out of range). */
typedef struct vgpu_ray {
    float P[3];
    float tmax;
    float D[3];
    float time;
} vgpu_ray;

/* Hit record, 16 bytes.  prim = ShaderContext.ElemID, the triangle index
 * after the BVH leaf reorder (builtin/geom/polymesh/trace.go:513), or -1 on
 * a miss; t = Ray.Tclosest after the hit; u,v = ShaderContext.Bu/Bv, the
 * barycentric weights of the triangle's first and second vertex
 * (trace.go:343-344).  On a miss t holds the ray's tmax. */
typedef struct vgpu_hit {
    int32_t prim;
    float t;
    float u;
    float v;
} vgpu_hit;

/* qbvh.Node, 128 bytes, the reference layout (qbvh/qbvh.go:31-36):
 * boxes[child + i*12 + axis*4] (i=0 min, 1 max); children >= 0 are nodes,
 * -1 is an empty leaf, otherwise 1<<31 | first<<4 | (count-1). */
typedef struct vgpu_qbvh_node {
    float boxes[24];
    uint32_t axis0, axis1, axis2;
    int32_t children[4];
    int32_t parent;
} vgpu_qbvh_node;

/* qbvh.MotionNode topology (qbvh/mqbvh.go:24-29) reduced to what traversal
 * reads: axis0..2, children[4], one pad word. */
typedef struct vgpu_mqbvh_node {
    int32_t axis0, axis1, axis2;
    int32_t children[4];
    int32_t pad;
} vgpu_mqbvh_node;

/* vgpu_trace flags */
#define VGPU_TRACE_SHADOW 1u        /* RayTypeShadow: any-hit early out (qbvh/intersect.go:232-236) */
#define VGPU_TRACE_DEVICE_PTRS 2u   /* rays/hits/geoms/counters are device pointers; the call only
                                       enqueues work on `stream` and returns without synchronising */

#define VGPU_GRAPH 32u                /* vgpu_render_pass / vgpu_pt_pass with VGPU_TRACE_DEVICE_PTRS: the
                                       pass's launches are captured into a HIP graph and replayed as one
                                       (the graph is updated in place from pass to pass; a pass whose
                                       launch sequence differs re-instantiates it).  Not with timing on. */

/* vgpu_scene_add_mesh flags */
#define VGPU_MESH_MOTION_REFERENCE_INDEXING 1u
    /* motion meshes only: test the triangle idxp[leaf slot] as the reference does
       (builtin/geom/polymesh/trace.go:532-537, idxp is never reordered for motion
       meshes, buildqbvh.go:19-57).  Default: the face the BVH leaf bounds, accel.idx[slot]. */

/* ShaderStd parameters (builtin/shader/std.go:25-47).  has_* = 1 marks a
 * parameter that is set (the reference's nil-vs-set distinction, which
 * selects its defaults: diffuse roughness 0.5, spec roughness 0.5, IOR 1.7). */
typedef struct vgpu_material {
    float emission[3]; float emission_strength; int32_t has_emission_colour, has_emission_strength;
    float diffuse[3]; int32_t has_diffuse_colour;
    float diffuse_strength; int32_t has_diffuse_strength;
    float diffuse_roughness; int32_t has_diffuse_roughness;
    float spec1[3]; int32_t has_spec1_colour;
    float spec1_strength; int32_t has_spec1_strength;
    float spec1_roughness; int32_t has_spec1_roughness;
    float ior; int32_t has_ior;
    int32_t fresnel_model;   /* Spec1FresnelModel: 0 (unset, the reference's default) or
                                VGPU_FRESNEL_DIELECTRIC = fresnel/dielectric.go with ior;
                                VGPU_FRESNEL_METAL = fresnel/conductor.go (Gulbrandsen) */
    float spec1_fresnel_refl[3]; int32_t has_spec1_fresnel_refl;   /* Spec1FresnelRefl (default 0.5 grey) */
    float spec1_fresnel_edge[3]; int32_t has_spec1_fresnel_edge;   /* Spec1FresnelEdge: as in std.go:185-187
                                it replaces the reflectivity, and the edge tint stays 0.5 grey */
} vgpu_material;
#define VGPU_FRESNEL_DIELECTRIC 1
#define VGPU_FRESNEL_METAL 2

/* A light node and the geometry it inserts into the scene (geom = insertion
 * index of that geometry; the renderer drops a light while shading its own
 * geometry, scene.go:100-116):
 *   VGPU_LIGHT_TRI    TriLight (builtin/light/triangle.go:18-28): P0, P1, P2 the
 *                     vertices; geometry: a one-triangle mesh.
 *   VGPU_LIGHT_DISK   DiskLight (disk.go:17-31): P0 = P, P1 = Up, P2 = LookAt,
 *                     radius; geometry: the Segments-triangle fan of
 *                     Disk.createMesh (disk.go:224-262), added by the caller.
 *   VGPU_LIGHT_SPHERE SphereLight (sphere.go:16-28): P0 = centre, radius;
 *                     geometry: vgpu_scene_add_sphere. */
typedef struct vgpu_light {
    float P0[3], P1[3], P2[3];
    int32_t material;
    int32_t samples;         /* NumSamples = 1 << samples (default 1) */
    int32_t geom;
    int32_t type;            /* VGPU_LIGHT_* */
    float radius;            /* disk and sphere */
} vgpu_light;
#define VGPU_LIGHT_TRI 0
#define VGPU_LIGHT_DISK 1
#define VGPU_LIGHT_SPHERE 2

/* Camera for one frame: the 4x4 column-major camera-to-world matrix M
 * (camera.go:225-236), TanThetaFocal = tan(Fov/2)*Focal, Aspect, Focal. */
typedef struct vgpu_camera {
    float M[16];
    float tan_theta_focal;
    float aspect;
    float focal;
    float radius;            /* lens radius: > 0 samples a thin lens with UniformDisk2D
                                (camera.go:265-283, math/sample/sample.go:105-129); path mode needs 0 */
} vgpu_camera;

typedef struct vgpu_scene vgpu_scene;
typedef struct vgpu_render vgpu_render;

/* Description of the last failure on the calling host thread (never NULL). */
const char* vgpu_last_error(void);

/* Number of visible GPUs. */
int vgpu_device_count(int* count);

/* Host builder.  Replaces qbvh.BuildAccel (qbvh/build.go:293-307) with an
 * output identical node for node: boxes (n*6: min xyz, max xyz), centroids
 * (n*3) and idx (n) are permuted in place exactly as the reference permutes
 * them; nodes_out receives up to max_nodes nodes; *nnodes_out the count;
 * bounds_out (6 floats, nullable) the root box.  nthreads <= 0 uses all
 * hardware threads; the output does not depend on it. */
int vgpu_qbvh_build(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leaf_max,
                    vgpu_qbvh_node* nodes_out, int32_t max_nodes, int32_t* nnodes_out,
                    float* bounds_out, int32_t nthreads);

/* GPU builder.  Replaces qbvh.BuildAccel (qbvh/build.go:293-307) as
 * vgpu_qbvh_build does, with the same output node for node and the same
 * permutation of boxes / centroids / idx, computed on the current device with
 * work enqueued on `stream` (a hipStream_t, NULL = the null stream; the call
 * returns after the results are in the host arrays).  Non-finite centroids
 * fail with VGPU_ERR_BUILD as on the host; so does a NaN box coordinate,
 * which vgpu_qbvh_build accepts.  A tree more than 192 levels deep below
 * its 2048-element subtrees fails with VGPU_ERR_STACK (vgpu_qbvh_build has no
 * such limit).  Transfers of 4 MiB or more go through 128 MiB of pinned host
 * memory per device, allocated by the first such call and kept until the
 * process ends (calls on one device take turns on it). */
int vgpu_qbvh_build_gpu(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leaf_max,
                        vgpu_qbvh_node* nodes_out, int32_t max_nodes, int32_t* nnodes_out,
                        float* bounds_out, void* stream);

/* Replaces qbvh.BuildAccelMotion (qbvh/motionbuild.go:108-122): topology only. */
int vgpu_qbvh_build_motion(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leaf_max,
                           vgpu_mqbvh_node* nodes_out, int32_t max_nodes, int32_t* nnodes_out,
                           int32_t nthreads);

/* Scene lifetime.  Replaces core.Init(scene.New()) (core/core.go:26,
 * builtin/scene/scene.go:24) with a scene whose geometry lives on `device`. */
int vgpu_scene_create(vgpu_scene** out, int device);
int vgpu_scene_destroy(vgpu_scene* scene);

/* Adds a triangle mesh (a PolyMesh with FaceIdx and no PolyCount,
 * builtin/geom/polymesh/init.go:71-94).  verts: nkeys*nverts*3 floats,
 * key-major (param.PointArray, core/param/array.go:19-28); idx: ntris*3
 * vertex indices.  The per-mesh accel is built at commit exactly as
 * PolyMesh.initAccel (buildqbvh.go:14-144).  *geom_out receives the
 * insertion index (the post-commit slot is reported by vgpu_scene_info). */
int vgpu_scene_add_mesh(vgpu_scene* scene, const float* verts, uint32_t nverts, uint32_t nkeys,
                        const uint32_t* idx, uint32_t ntris, float ray_bias, uint32_t flags,
                        int32_t* geom_out);

/* Adds a static mesh whose accel the caller already built with the
 * reference's own builder (the cgo path: Go's PolyMesh.initAccel output).
 * idxp_sorted: ntris*3 post-sort indices; nodes: mesh.accel.qbvh; bounds:
 * mesh.bounds (6 floats). */
int vgpu_scene_add_mesh_prebuilt(vgpu_scene* scene, const float* verts, uint32_t nverts,
                                 const uint32_t* idxp_sorted, uint32_t ntris,
                                 const vgpu_qbvh_node* nodes, uint32_t nnodes,
                                 const float* bounds, float ray_bias, int32_t* geom_out);

/* Adds a moving mesh whose accel the caller already built with the
 * reference's own code (the cgo path: Go's PolyMesh.initMotionAccel output,
 * buildqbvh.go:19-57, and initMotionBoxes, :148-212).  verts: nkeys*nverts*3
 * floats, key-major; idxp: ntris*3 vertex indices in face order (mesh.idxp,
 * which the reference leaves unsorted for moving meshes); accel_idx: ntris
 * faces in leaf order (mesh.accel.idx); nodes: mesh.accel.mqbvh.Nodes
 * (axes and children; Parent is not used); boxes: mesh.accel.mqbvh.Boxes,
 * nkeys*nnodes*24 floats, key-major.  mesh.motionBounds and mesh.bounds
 * are derived from the root's boxes as initMotionBoxes derives them.  flags:
 * as vgpu_scene_add_mesh (VGPU_MESH_MOTION_REFERENCE_INDEXING reads idxp by
 * leaf slot, as TraceMotionElems does). */
int vgpu_scene_add_motion_mesh_prebuilt(vgpu_scene* scene, const float* verts, uint32_t nverts,
                                        uint32_t nkeys, const uint32_t* idxp, uint32_t ntris,
                                        const int32_t* accel_idx, const vgpu_mqbvh_node* nodes,
                                        uint32_t nnodes, const float* boxes, float ray_bias,
                                        uint32_t flags, int32_t* geom_out);

/* A PolyMesh's Transform (builtin/geom/polymesh/trace.go:20-74): nkeys
 * column-major 4x4 matrices.  Rays are moved into the mesh's space by the
 * inverse of the transform at their time (one key: the matrix as given; more:
 * the TransformDecomp keys interpolated, math/animdecomp.go:66-80).  As in the
 * reference, the mesh's bounds ignore it (bounds.go:26), so the scene culls the
 * mesh by its untransformed box.  nkeys 0 removes it.  Before commit. */
int vgpu_scene_set_transform(vgpu_scene* scene, int32_t geom, const float* matrices, uint32_t nkeys);

/* Leaves geom (insertion index) out of the scene's top level: it is traced only
 * through the instances and procs that hold it (the geoms a Proc handler
 * creates, builtin/geom/proc/vnf/vnf.go:33-47).  Before commit. */
int vgpu_scene_set_hidden(vgpu_scene* scene, int32_t geom, int32_t hidden);

/* GeomInstance (builtin/geom/instance/instance.go:36-160): geom (a mesh or
 * sphere, insertion index) traced under nkeys transform keys (TransformSRT,
 * interpolated at the ray's time; one key recomposed from its decomposition),
 * with bounds the box of bmin, bmax (BMin[0], BMax[0]; the reference ignores
 * later keys).  Hits report the instance's slot.  Scenes holding instances,
 * procs or mesh transforms are traced by a per-lane kernel without per-ray
 * counters.  *geom_out: insertion index. */
int vgpu_scene_add_instance(vgpu_scene* scene, int32_t geom, const float* matrices, uint32_t nkeys, const float* bmin,
                            const float* bmax, int32_t* geom_out);

/* Proc (builtin/geom/proc/proc.go:40-196) over geoms[0..ngeoms) (meshes or
 * spheres; usually hidden), in that order, under nkeys transform keys (0: the
 * identity).  As the reference does, a closest-hit ray takes the first of its
 * geoms that reports a hit, closer or not (proc.go:112-120); shadow rays try
 * them all.  Bounds as for instances. */
int vgpu_scene_add_proc(vgpu_scene* scene, const int32_t* geoms, uint32_t ngeoms, const float* matrices, uint32_t nkeys,
                        const float* bmin, const float* bmax, int32_t* geom_out);

/* Builds the per-mesh accels not yet built, the scene top level
 * (Scene.initAccel, scene.go:135-203, leafMax 1) and uploads everything to
 * the device.  Replaces Scene.PreRender (scene.go:131-133). */
int vgpu_scene_commit(vgpu_scene* scene);

/* After commit: geom_order[slot] = insertion index of the geom in post-sort
 * slot `slot` (the permutation of scene.go:164-171); geom_nodes / geom_tris
 * per slot (nullable arrays of ngeoms); *top_nodes the top-level node count. */
int vgpu_scene_info(const vgpu_scene* scene, int32_t* ngeoms, int32_t* top_nodes,
                    int32_t* geom_order, int32_t* geom_nodes, int32_t* geom_tris);

/* After commit: copies a static geom's host-side accel (for checking the
 * builder): nodes (geom_nodes[slot] nodes), post-sort idxp (ntris*3) and
 * accel.idx (ntris).  Any output may be NULL. */
int vgpu_scene_mesh_accel(const vgpu_scene* scene, int32_t slot, vgpu_qbvh_node* nodes,
                          uint32_t* idxp_sorted, int32_t* perm);

/* Traces n rays against the committed scene.  Replaces core.TraceProbe ->
 * Scene.Trace -> qbvh.Trace/TraceMotion -> PolyMesh.Trace ->
 * qbvh.Trace/TraceMotion -> PolyMesh.TraceElems/TraceMotionElems
 * (core/trace.go:26-35, builtin/scene/scene.go:30-97,
 * qbvh/intersect.go:91-246, qbvh/motionintersect.go:22-127,
 * builtin/geom/polymesh/trace.go:15-194, :520-669) for a batch of rays.
 * hits: n records.  geoms (nullable): n post-sort geom slots
 * (ShaderContext.Geom, scene.go:67) or -1.  counters (nullable): 4 int32
 * per ray — node visits, leaf visits, triangle tests, maximum stack depth.
 * Without VGPU_TRACE_DEVICE_PTRS all arrays are host memory and the call is
 * synchronous; with it they are device memory and the work is enqueued on
 * `stream` (a hipStream_t, NULL = the null stream). */
int vgpu_trace(vgpu_scene* scene, uint32_t flags, const vgpu_ray* rays, uint64_t n,
               vgpu_hit* hits, int32_t* geoms, int32_t* counters, void* stream);

/* Allocation-free variant of vgpu_trace.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (work under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * array is device memory; the work is enqueued on `stream` and the call
 * returns without synchronising (flags: VGPU_TRACE_SHADOW, and
 * VGPU_TRACE_DEVICE_PTRS, which is implied).  Results equal vgpu_trace's.
 * The traversal stack spill lives in a 128 MiB array that the library holds
 * on every GPU it runs on until the process ends; the launch takes the
 * scene's shape with fewer resident waves where the spill would not fit, and
 * rays are never sorted.  Any n below 2^32 fits (the persistent grid reuses
 * its spill columns); larger n is rejected with VGPU_ERR_ARG.  d_status
 * (nullable, one device int32) is cleared by a kernel at the start of the
 * call and set non-zero if a ray needed more than VGPU_MAX_STACK entries (its
 * hit is written as a miss).  The array is shared by every call on one GPU
 * and keeps the last call's contents: calls that overlap on one GPU, from any
 * scene or stream, race.  It calls no hipMalloc, hipFree or other allocating
 * routine. */
int vgpu_trace_noalloc(vgpu_scene* scene, uint32_t flags, const vgpu_ray* d_rays, uint64_t n, vgpu_hit* d_hits,
                       int32_t* d_geoms, int32_t* d_counters, int32_t* d_status, void* stream);

/* Synchronises `stream` and reports (then clears) an error latched by
 * traversal work enqueued with VGPU_TRACE_DEVICE_PTRS: VGPU_ERR_STACK if a
 * ray overflowed VGPU_MAX_STACK (its hit was written as a miss).  No
 * reference counterpart: the reference panics on stack overflow. */
int vgpu_scene_status(vgpu_scene* scene, void* stream);

/* Adds a static triangle mesh with shading data: normals (nverts*3 floats,
 * per vertex, nullable: N = Ng), shaderidx (one byte per triangle, nullable:
 * all 0) and the mesh's shader table (material ids; PolyMesh.Shader,
 * polymesh.go:293-306).  As vgpu_scene_add_mesh otherwise. */
int vgpu_scene_add_shaded_mesh(vgpu_scene* scene, const float* verts, uint32_t nverts, const uint32_t* idx,
                               uint32_t ntris, const float* normals, const uint8_t* shaderidx,
                               const int32_t* shaders, uint32_t nshaders, int32_t* geom_out);

/* Adds a moving triangle mesh with shading data (a PolyMesh whose P has
 * nkeys >= 2 motion keys, builtin/geom/polymesh/trace.go:520-719): verts
 * nkeys*nverts*3 floats, key-major, as vgpu_scene_add_mesh.  normals
 * (nullable) hold normal_keys*ntris*3*3 floats, one normal per face corner
 * and key, as TraceMotionElems indexes Normals (trace.go:694-703); they are
 * used only when normal_keys == nkeys (otherwise N = Ng, as trace.go:694
 * decides).  shaderidx (nullable) is per triangle in insertion order.  flags:
 * 0 or VGPU_MESH_MOTION_REFERENCE_INDEXING (the leaf slot indexes the
 * vertices, shader index and normals, trace.go:527-545).  A hit is shaded at
 * the ray's time: vertices and normals interpolated between the two keys that
 * bracket it, P, Ng, dPdu = v1 - v0, dPdv = v2 - v0 and the offset as
 * trace.go:625-669 computes them. */
int vgpu_scene_add_shaded_motion_mesh(vgpu_scene* scene, const float* verts, uint32_t nverts, uint32_t nkeys,
                                      const uint32_t* idx, uint32_t ntris, const float* normals,
                                      uint32_t normal_keys, const uint8_t* shaderidx, const int32_t* shaders,
                                      uint32_t nshaders, uint32_t flags, int32_t* geom_out);

/* Adds the analytic sphere a SphereLight inserts (builtin/geom/sphere:
 * sphere.go:15-95, trace.go:13-52) with its shader (a material id).  It is a
 * scene geom like a mesh: bounds P -/+ radius in the top-level BVH, hit by
 * the closed-form quadratic with no ray bias; a hit reports prim 0, u = v = 0. */
int vgpu_scene_add_sphere(vgpu_scene* scene, const float* centre, float radius, int32_t material, int32_t* geom_out);

/* Renderer over a committed scene: materials, lights, camera and image
 * size.  Replaces the body of core.Render's pass loop (core/render.go:66-137)
 * in the reference semantics: per pixel one camera ray, ShaderStd.Eval
 * direct lighting with MIS (std.go:77-296, core/shader.go:203-402), perfect
 * mirrors (Spec1Roughness 0) by recursion to level 3 with one light sample
 * per light below level 0, spectral hero-wavelength colour, film
 * fb = (fb*iter + c)/(iter+1).  Each sample has the time
 * VanDerCorput(iter, time scramble) (render.go:93,118) that its camera,
 * mirror and shadow rays carry; moving meshes are traced and shaded at it. */
int vgpu_render_create(vgpu_render** out, vgpu_scene* scene, const vgpu_material* materials, uint32_t nmaterials,
                       const vgpu_light* lights, uint32_t nlights, const vgpu_camera* camera,
                       uint32_t width, uint32_t height);

/* Pixel filter (core.PixelFilter, builtin/filter): VGPU_FILTER_BOX (no
 * filter node: samples stay where RasterXY puts them), VGPU_FILTER_GAUSSIAN
 * (gauss.go; defaults Width 2, Res 17) or VGPU_FILTER_AIRY (airy.go; defaults
 * Width 6, Res 49, Peak 4).  width/res/peak <= 0 take the defaults.  The
 * filter's Res x Res table is built on the host in float64 as
 * filter.CreateSampler does (filter.go:87-173) and samples are warped as
 * Sampler.WarpSample (filter.go:40-85), quirks included. */
int vgpu_render_set_filter(vgpu_render* render, int32_t type, float width, int32_t res, float peak);

/* Camera motion keys: nkeys column-major LocalToWorld matrices (the
 * camera's LocalToWorld.Elems as calcLookatMatrices or matrixCalc leave
 * them, builtin/camera/camera.go:109-193).  Each is decomposed on the host as
 * TransformDecompMatrix4 (math/animdecomp.go:20-63).  With two or more keys
 * every sample's camera matrix is the lerp of the two keys bracketing its
 * time, recomposed as ComputeRay does (camera.go:224-236); one key sets the
 * single matrix of every sample to that recomposition (replacing the
 * vgpu_camera matrix); 0 keys go back to the vgpu_camera matrix given at
 * creation.  At most 64 keys. */
int vgpu_render_set_camera_keys(vgpu_render* render, const float* local_to_world, uint32_t nkeys);
/* Path mode (no reference counterpart; core.Render and ShaderStd.Eval,
 * core/render.go:66-137 and builtin/shader/std.go:77-296, shade one camera
 * hit with direct light and recurse only through perfect mirrors).  depth 0
 * keeps the reference's shading.  depth D in [1,16] traces paths of up to D
 * segments: every vertex is shaded by ShaderStd as at level 0 of the
 * reference (EvaluateLightSamples per light and lobe, one sample per light
 * below the camera vertex) and, before the last segment, continues with one
 * ray: the diffuse (Oren-Nayar), mirror or GGX lobe is picked with its
 * normalised weight and sampled with the lobe's own Sample; the continuation
 * carries (Eval / pdf).ToRGB() x the lobe colour (for GGX, pdf is the density
 * of the sampled direction), clamped at 0 as the reference clamps its mirror
 * ray.  Emission of a light's own geometry reached by an Oren-Nayar or GGX
 * continuation is not added (the light's EvaluateLightSamples at the vertex
 * before covers it).  QMC: the camera vertex's light samples use the pixel's
 * scramble0/1 as the reference does; every other level and dimension pair uses
 * a SplitMix64 hash of them.  Depth 1 gives the reference's pixels for scenes
 * without mirrors. */
int vgpu_render_set_path_depth(vgpu_render* render, uint32_t depth);

/* Rays traced by the renderer's passes since creation, as core.RenderStats
 * counts core.TraceProbe calls (core/stats.go:14-45, core/trace.go:26-35):
 * closest hit (camera rays, mirror rays, path mode's continuations) and
 * shadow rays.  Synchronises the device. */
int vgpu_render_stats(vgpu_render* render, uint64_t* closest_rays, uint64_t* shadow_rays);
#define VGPU_FILTER_BOX 0
#define VGPU_FILTER_GAUSSIAN 1
#define VGPU_FILTER_AIRY 2
int vgpu_render_destroy(vgpu_render* render);

/* One pass over every pixel.  scrambles: 6 uint64 per pixel (lensU, lensV,
 * time, lambda, scramble0, scramble1; core/render.go:166-176); film: width*
 * height*3 floats, updated in place; sample_out (nullable): this pass's
 * sample colour per pixel.  iter: the 1-based pass number.  With
 * VGPU_TRACE_DEVICE_PTRS the arrays are device memory and the pass is
 * enqueued on `stream`; otherwise host memory and synchronous.
 * *shadow_rays_out (nullable, host mode only): shadow rays traced. */
int vgpu_render_pass(vgpu_render* render, uint32_t flags, const uint64_t* scrambles, uint32_t iter, float* film,
                     float* sample_out, void* stream, uint64_t* shadow_rays_out);

/* Before commit: the leafMax argument BuildAccel gets for static meshes (1..16).
 * The reference uses 16 (builtin/geom/polymesh/buildqbvh.go:32), the default;
 * hit IDs match the reference only with it.  Path mode, which has no
 * reference counterpart, may use smaller leaves for fewer triangle tests. */
int vgpu_scene_set_leaf_size(vgpu_scene* scene, int32_t leaf_max);

/* Kernel timing, a diagnostic with no reference counterpart (the reference
 * reports only ray counts, core/stats.go:14-45).  While enabled, every
 * traversal records HIP events on its stream around the ray sort and the
 * traversal kernel (do not enable it while capturing a graph).
 * vgpu_scene_timing synchronises on them, returns the summed traversal and
 * sort milliseconds and the number of traversals since the last call, and
 * starts a new interval. */
int vgpu_scene_set_timing(vgpu_scene* scene, int32_t enable);

/* Streaming diagnostic, no reference counterpart: measures the HBM bandwidth a
 * plain access pattern reaches and calibrates the FETCH_SIZE / WRITE_SIZE
 * counters.  Enqueues on `stream` (current device) one kernel of
 * CUs x blocks_per_cu workgroups of 256 threads over `bytes` (a multiple of
 * 16; 16-byte aligned device buffers).  VGPU_DIAG_COPY copies src to dst;
 * VGPU_DIAG_READ reads src and writes one float per thread to dst
 * (CUs x blocks_per_cu x 256 floats). */
#define VGPU_DIAG_COPY 0
#define VGPU_DIAG_READ 1
int vgpu_diag_stream(int32_t mode, const void* src, void* dst, uint64_t bytes, int32_t blocks_per_cu, void* stream);
int vgpu_scene_timing(vgpu_scene* scene, float* trace_ms, float* sort_ms, int32_t* launches);

/* Traversal launch configuration (the autotuner's knobs).  Until the caller
 * sets one (or runs the autotuner), vgpu_scene_commit picks block 128, LDS
 * stack 12, 24 waves per CU, no sort for a static scene, and 64 / 8 / 20 /
 * no sort for a scene with motion keys. */
typedef struct vgpu_trace_config {
    int32_t block;        /* threads per workgroup: 64, 128 or 256 */
    int32_t lds_stack;    /* stack entries per ray held in LDS: 8, 12, 16, 24 or 32; the rest spill to HBM */
    int32_t waves_per_cu; /* resident-grid sizing: workgroups = CUs * waves_per_cu * 64 / block */
    int32_t sort;         /* 0: trace in the given order; k in 1..6: first reorder the rays by bucket =
                           * (Morton cell of the origin in a 2^k-per-axis grid over the scene box) x
                           * (direction octant); results still land in the callers' order */
    int32_t refill;       /* 1..64: a wavefront's finished lanes wait until this many are idle, then take
                           * new rays together (ray set-up runs once for them); 0: the default, 48 */
} vgpu_trace_config;

int vgpu_scene_set_trace_config(vgpu_scene* scene, const vgpu_trace_config* cfg);
int vgpu_scene_get_trace_config(const vgpu_scene* scene, vgpu_trace_config* cfg);

/* Runs the autotuner: times each candidate configuration (block x LDS depth x
 * waves per CU in {8, 16, 20, 24, 32} x sort in {0, 4, 5, 6}, the sort
 * included in the time) on the given rays
 * (device pointers, n rays), then the refill threshold (16, 32, 40, 48, 56, 64)
 * in the fastest of those, and keeps the fastest.  The results of a trace never
 * depend on its configuration.  best_out (nullable) receives the chosen
 * configuration; ms_out (nullable) its time. */
int vgpu_scene_autotune(vgpu_scene* scene, uint32_t flags, const vgpu_ray* d_rays, uint64_t n,
                        vgpu_hit* d_hits, void* stream, vgpu_trace_config* best_out, float* ms_out);

/* ------------------------------------------------------------------------
 * Path-guiding network.  No reference counterpart: the reference samples
 * only its BSDFs (the Go files of builtin/shader/bsdf).  A small MLP maps a hit
 * position (normalised by the scene bounds) and shading normal, with a
 * 12-frequency sin/cos encoding (144 features), through 4 hidden layers of
 * 64 (ReLU) to 8 von Mises-Fisher lobes (mu, kappa, weight).  Weights are
 * multiplied in bf16 on the matrix cores with fp32 accumulation; it is
 * trained online with Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the loss
 * (1/B) sum_i -(c_i/cbar) log q(d_i), cbar = mean c.  Parameters: 24,360
 * floats in the order W1 (64x144, row = output), b1, W2, b2, W3, b3, W4, b4,
 * W5 (40x64), b5; output k*5+0..2 = lobe direction (unnormalised),
 * k*5+3 = log kappa (clamped to [-5,7]), k*5+4 = mixture logit.
 * ------------------------------------------------------------------------ */
typedef struct vgpu_guide vgpu_guide;

#define VGPU_GUIDE_NPARAMS 24360
#define VGPU_GUIDE_RAW 4        /* vgpu_guide_infer: write the 40 raw outputs */
#define VGPU_GUIDE_NO_UPDATE 8  /* vgpu_guide_train_step: compute the gradient only */

/* bounds: scene box (lo xyz, hi xyz) used to normalise positions; seed:
 * weight initialisation; max_batch: largest training batch (a multiple of
 * 1024, at most 2^22). */
int vgpu_guide_create(vgpu_guide** out, int device, const float* bounds, uint64_t seed, uint32_t max_batch);
int vgpu_guide_destroy(vgpu_guide* guide);
/* Replaces the parameters (24,360 floats) and resets the Adam state. */
int vgpu_guide_set_params(vgpu_guide* guide, const float* params);
int vgpu_guide_get_params(vgpu_guide* guide, float* params);
/* The gradient of the last training step's loss (24,360 floats). */
int vgpu_guide_get_grad(vgpu_guide* guide, float* grad);
/* Evaluates n queries (8 floats each: position xyz, pad, normal xyz, pad).
 * out: n x 40 floats, either 8 decoded lobes of (mu xyz, kappa, weight) or,
 * with VGPU_GUIDE_RAW, the raw outputs in parameter order.  Host pointers
 * and synchronous unless VGPU_TRACE_DEVICE_PTRS. */
int vgpu_guide_infer(vgpu_guide* guide, uint32_t flags, const float* queries, uint64_t n, float* out, void* stream);
/* One training step on n samples (12 floats each: position xyz, c, normal
 * xyz, pad, unit direction xyz, pad); n a multiple of 1024 up to max_batch.
 * The loss is sum_i c_i (-log q(d_i)) / sum_i c_i.  Samples with c = 0
 * contribute no gradient; a batch whose c sum is not positive makes no update
 * and leaves a zero gradient.  *loss_out (nullable, host mode only): the
 * batch loss. */
int vgpu_guide_train_step(vgpu_guide* guide, uint32_t flags, const float* samples, uint32_t n, float lr,
                          void* stream, float* loss_out);

/* ------------------------------------------------------------------------
 * Path mode.  No reference counterpart: core.Render shades one camera hit
 * with direct light (core/render.go:66-137, builtin/shader/std.go:77-296).
 * Path mode traces one path per pixel per pass, up to max_depth segments,
 * over Lambertian surfaces with emission: next-event estimation to every
 * emissive triangle (chosen by area x luminance) and one-sample MIS between
 * cosine-weighted BSDF sampling and, when a guide is attached, the guide's
 * vMF mixture, p = (1-alpha) cos/pi + alpha q; NEE and the mixture are
 * combined by the balance heuristic.  Each pass adds its radiance to an
 * FP64 accumulation buffer (the image is accum / passes).
 * ------------------------------------------------------------------------ */
typedef struct vgpu_pt vgpu_pt;

typedef struct vgpu_pt_material {
    float albedo[3];     /* Lambertian reflectance, each in [0,1] */
    float emission[3];   /* emitted radiance */
    int32_t one_sided;   /* 1: emits from the front face only (v0, v1, v2 counter-clockwise) */
    int32_t pad;
} vgpu_pt_material;

#define VGPU_PT_NO_TRAIN 16   /* vgpu_pt_pass: use the guide without training it */

/* Scene meshes refer to materials through their shader tables
 * (vgpu_scene_add_shaded_mesh).  camera: as for vgpu_render_create. */
int vgpu_pt_create(vgpu_pt** out, vgpu_scene* scene, const vgpu_pt_material* materials, uint32_t nmaterials,
                   const vgpu_camera* camera, uint32_t width, uint32_t height, uint32_t max_depth, uint64_t seed);
int vgpu_pt_destroy(vgpu_pt* pt);
/* Attaches a guide (NULL detaches: plain BSDF sampling).  alpha in (0,1):
 * probability of sampling the guide.  With train_steps > 0, each pass
 * records the paths of one pixel in record_stride, turns their vertices
 * into training samples (c = luminance of the incident radiance estimate x
 * cos / sampling pdf) and runs train_steps Adam steps on batches of `batch`
 * samples drawn from them.  Attaching the same guide again changes only these
 * settings: its weights and Adam state stay, so train_steps = 0 freezes it. */
int vgpu_pt_set_guide(vgpu_pt* pt, vgpu_guide* guide, float alpha, uint32_t record_stride, uint32_t train_steps,
                      uint32_t batch, float lr);
/* One pass (pass index selects the random numbers).  accum: width*height*3
 * doubles, updated in place; device memory with VGPU_TRACE_DEVICE_PTRS
 * (enqueued on stream), else host memory (synchronous). */
int vgpu_pt_pass(vgpu_pt* pt, uint32_t flags, uint32_t pass, double* accum, void* stream);
/* Rays traced since creation (closest + shadow) and the training samples of
 * the last pass.  Synchronises the device. */
int vgpu_pt_stats(vgpu_pt* pt, uint64_t* rays, uint32_t* train_samples);

/* ------------------------------------------------------------------------
 * Scene files.  A .vnf file (docs/quickstart.rst) is parsed as nodes.Parse
 * does (nodes/lex.go, nodes/parser.go), quirks included, and built as
 * core.PreRender builds it (core/core.go:38-61; PolyMesh.init's fan
 * triangulation of PolyCount polygons, builtin/geom/polymesh/init.go:12-134;
 * the light nodes' geometry; ShaderStd; the LookAt or matrix camera; the
 * pixel filter) into a scene and a renderer.  Node types whose rendering this
 * library does not provide (DebugShader, QuadLight, Proc, GeomInstance,
 * rgbtex textures, non-identity mesh transforms) fail the build.
 * ------------------------------------------------------------------------ */
typedef struct vgpu_vnf vgpu_vnf;

#define VGPU_VNF_MESSAGES 0   /* vgpu_vnf_text: the parse errors the reference prints */
#define VGPU_VNF_NODES 1      /* vgpu_vnf_text: every parsed node's present parameters (floats as bits) */
#define VGPU_OUTPUT_HDR 0     /* an OutputHDR node */
#define VGPU_OUTPUT_FLOAT 1   /* an OutputFloat node */

/* Replaces nodes.Parse (nodes/parser.go:108-130) plus core.AddNode: the file's
 * nodes, and the frame size of its last Globals node (1024 x 1024 without
 * one).  Parse errors are reported as the reference prints them (see
 * vgpu_vnf_text) and parsing goes on, as it does; more than 10 errors fail
 * the load (the reference exits). */
int vgpu_vnf_load(const char* path, vgpu_vnf** out);
int vgpu_vnf_destroy(vgpu_vnf* vnf);
/* Frame size, Globals MaxIter and the number of parse errors. */
int vgpu_vnf_info(const vgpu_vnf* vnf, uint32_t* width, uint32_t* height, int32_t* max_iter, int32_t* nerrors);
/* Text of VGPU_VNF_MESSAGES or VGPU_VNF_NODES into buf (len bytes, NUL
 * terminated, truncated); *needed (nullable) the full size. */
int vgpu_vnf_text(const vgpu_vnf* vnf, int32_t what, char* buf, uint64_t len, uint64_t* needed);
/* Replaces core.PreRender and the setup of core.Render (core/render.go:140-176):
 * runs the nodes' PreRender in the reference's rounds (Include nodes parse
 * their file, lights add their geometry), creates the scene on `device` with
 * every geom in core.AddNode order, commits it and creates the renderer (its
 * camera from Globals Camera or "camera").  Once per loaded file. */
int vgpu_vnf_build(vgpu_vnf* vnf, int device, vgpu_scene** scene_out, vgpu_render** render_out);
/* After build: *count (nullable) the output nodes; for i >= 0 the i-th one's
 * type (VGPU_OUTPUT_*) and file name (len bytes, NUL terminated). */
int vgpu_vnf_output(const vgpu_vnf* vnf, int32_t i, int32_t* type, char* path, uint64_t len, int32_t* count);

/* ------------------------------------------------------------------------
 * Output drivers.  The film is width*height*3 floats in pixel order (row 0
 * first), as core.FrameBuf() holds it; with VGPU_TRACE_DEVICE_PTRS it is
 * device memory on `device` (work on `stream`), otherwise host memory.
 * ------------------------------------------------------------------------ */

/* The Radiance RGBE bytes of the film as image/hdr/writer.go:62-107 and
 * hdr.go:24-48 produce them on amd64: 4 bytes per pixel, the bottom film row
 * first; rgbe (width*height*4 bytes) is device memory with
 * VGPU_TRACE_DEVICE_PTRS, else host memory.  Encoded on the GPU. */
int vgpu_film_to_rgbe(int device, uint32_t flags, const float* film, uint32_t width, uint32_t height, uint8_t* rgbe,
                      void* stream);

/* Replaces OutputHDR.PostRender (builtin/driver/outputhdr.go:30-57): writes
 * the file byte for byte as the reference does (its header, no run-length
 * encoding).  Synchronous. */
int vgpu_write_hdr(int device, uint32_t flags, const float* film, uint32_t width, uint32_t height, const char* path,
                   void* stream);

/* Replaces OutputFloat.PostRender (builtin/driver/outputfloat.go:30-42): the
 * film's float32 values, little-endian, in pixel order.  Synchronous. */
int vgpu_write_float(int device, uint32_t flags, const float* film, uint32_t width, uint32_t height, const char* path,
                     void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VERMEER_GPU_H */
