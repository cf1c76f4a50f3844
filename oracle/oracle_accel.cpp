// ORACLE — test infrastructure only.
//
// CPU restatement of the reference's ray-traversal hot path, used as the
// checker for the HIP implementation.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load this library; the product path
// (vermeer_amd/) never links, imports or calls it.
//
// Compiled with -O2 -ffp-contract=off -fno-fast-math -msse4.1 so that every
// float32 operation rounds exactly as the reference's Go code on amd64
// (GOAMD64=v1: no fused multiply-add).
//
// Functions restated here (paths relative to the reference checkout):
//   math/dim_amd64.s:8-21            Max/Min (MAXSS/MINSS operand semantics)
//   math/abs.go:13-15                Abs (sign-bit clear)
//   math/unsafe.go:15-37             SignMask / Xorf
//   math/ferror.go:16-24             EpsilonFloat32, MachineEpsilon32, Gamma
//   math/boundingbox.go:8-127        BoundingBox Reset/Grow/GrowBox/Dim/SurfaceArea/LongestAxis/Centroid
//   math/vec3.go:52-58               Vec3Lerp
//   core/ray.go:103-148              Ray.Setup
//   qbvh/qbvh.go:29-113              Node layout, SetLeaf/LeafBase/LeafCount
//   qbvh/intersect_amd64.s:13-100    intersectBoxes (1 ray x 4 slabs)
//   qbvh/build.go:28-307             calcMinCost / calcBox / binarySplit / buildAccelRec / BuildAccel
//   qbvh/intersect.go:91-246         Trace (closest-hit / shadow any-hit)
//   qbvh/mqbvh.go:20-98              MotionQBVH layout
//   qbvh/motionbuild.go:12-122       BuildAccelMotion
//   qbvh/motionintersect.go:22-127   TraceMotion
//   builtin/geom/polymesh/buildqbvh.go:14-212   initAccel / initMotionBoxes
//   builtin/geom/polymesh/trace.go:15-102       PolyMesh.Trace (static / motion dispatch)
//   builtin/geom/polymesh/trace.go:108-194      TraceElems watertight test (idxp branch)
//   builtin/geom/polymesh/trace.go:276-513      TraceElems hit attributes (Ng, Poffset, P)
//   builtin/geom/polymesh/trace.go:520-719      TraceMotionElems
//   builtin/scene/scene.go:30-268               Scene.Trace / TraceElems / initAccel / initMotionBoxes
//
// Mesh transforms (trace.go:22-55), instancing and procedurals are not part
// of this restatement: the workloads it checks have none.

#include "oracle_core.h"


// =================================================================== C ABI

using namespace orc;

extern "C" {

struct orc_ray { float P[3]; float tmax; float D[3]; float time; };
struct orc_hit { int32_t geom; int32_t prim; float t; float u; float v; };

// Restated qbvh.BuildAccel on caller arrays.  boxes: n*6 (min xyz, max xyz),
// centroids: n*3, idx: n — all permuted in place exactly as the reference
// permutes them.  nodes_out holds max_nodes 128-byte nodes.  Returns the node
// count, or -1 on builder failure, -2 if max_nodes is too small.
int orc_build_qbvh(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leafMax,
                   void* nodes_out, int32_t max_nodes, float* bounds_out) {
    std::vector<Box> bx(n); std::vector<Vec3> ce(n);
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) { bx[i].b[0][k] = boxes[i * 6 + k]; bx[i].b[1][k] = boxes[i * 6 + 3 + k]; ce[i][k] = centroids[i * 3 + k]; }
    }
    std::vector<Node> nodes;
    Box b;
    try { b = BuildAccel(nodes, bx.data(), ce.data(), idx, n, leafMax); } catch (BuildError&) { return -1; }
    if ((int)nodes.size() > max_nodes) return -2;
    std::memcpy(nodes_out, nodes.data(), nodes.size() * sizeof(Node));
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) { boxes[i * 6 + k] = bx[i].b[0][k]; boxes[i * 6 + 3 + k] = bx[i].b[1][k]; centroids[i * 3 + k] = ce[i][k]; }
    }
    if (bounds_out) for (int k = 0; k < 3; k++) { bounds_out[k] = b.b[0][k]; bounds_out[3 + k] = b.b[1][k]; }
    return (int)nodes.size();
}

// Restated intersectBoxes for known-answer tests: ray = P[3], D[3] (Dinv from Ray.Setup).
void orc_intersect_boxes(const float* P, const float* D, const float* boxes24, int32_t* hits, float* t) {
    Ray r{}; for (int k = 0; k < 3; k++) { r.P[k] = P[k]; r.D[k] = D[k]; }
    RaySetup(r);
    intersectBoxes(r, boxes24, hits, t);
}

// Restated Ray.Setup: out = S[3], Dinv[3] as float, K = Kx,Ky,Kz.
void orc_ray_setup(const float* D, float* S_out, float* Dinv_out, int32_t* K_out) {
    Ray r{}; for (int k = 0; k < 3; k++) r.D[k] = D[k];
    RaySetup(r);
    for (int k = 0; k < 3; k++) { S_out[k] = r.S[k]; Dinv_out[k] = r.Dinv[k]; }
    K_out[0] = r.Kx; K_out[1] = r.Ky; K_out[2] = r.Kz;
}

// 0: RSQRTSS estimate in Vec3Normalize (the reference on this CPU); 1: IEEE 1/sqrt estimate.
void orc_set_normalize_ieee(int32_t on) { normalize_ieee() = on != 0; }

void* orc_scene_new() { return new Scene(); }
void orc_scene_free(void* s) { delete (Scene*)s; }

// Adds a triangle mesh (PolyMesh with FaceIdx, no PolyCount). verts: nkeys*nverts*3 floats,
// key-major.  idx: ntris*3.  Returns insertion index.
int orc_scene_add_mesh(void* sp, const float* verts, int32_t nverts, int32_t nkeys, const uint32_t* idx, int32_t ntris,
                       float rayBias, int32_t motion_bug) {
    Scene* s = (Scene*)sp;
    Mesh* m = new Mesh();
    m->nkeys = nkeys; m->elemsPerKey = nverts;
    m->verts.resize((size_t)nverts * nkeys);
    std::memcpy(m->verts.data(), verts, sizeof(float) * 3 * (size_t)nverts * nkeys);
    m->idxp.assign(idx, idx + (size_t)ntris * 3);
    m->facecount = ntris;
    m->RayBias = rayBias;
    m->motionBug = motion_bug != 0;
    s->geoms.push_back(m);
    return (int)s->geoms.size() - 1;
}

// Adds the analytic sphere of a SphereLight (builtin/geom/sphere) with shader table {material}.
int orc_scene_add_sphere(void* sp, const float* centre, float radius, int32_t material) {
    Scene* s = (Scene*)sp;
    Mesh* m = new Mesh();
    m->sphere = true;
    m->nkeys = 1;
    for (int k = 0; k < 3; k++) m->sphereC[k] = centre[k];
    m->sphereR = radius;
    // Sphere.Bounds (sphere.go:68-77)
    for (int k = 0; k < 3; k++) { m->bounds.b[0][k] = centre[k] - radius; m->bounds.b[1][k] = centre[k] + radius; }
    m->shaders.assign(1, material);
    s->geoms.push_back(m);
    return (int)s->geoms.size() - 1;
}

// A geom's transform keys (nkeys column-major 4x4 matrices) and their TransformDecompMatrix4.
static void set_keys(Mesh* m, const float* mats, int32_t nkeys) {
    m->xform.clear(); m->xformSRT.clear();
    for (int k = 0; k < nkeys; k++) {
        m->xform.emplace_back(mats + 16 * k, mats + 16 * k + 16);
        m->xformSRT.push_back(DecompM4(mats + 16 * k));
    }
}
static Box user_box(const float* bmin, const float* bmax) {
    Box b; b.Reset();
    Vec3 a, c;
    for (int k = 0; k < 3; k++) { a[k] = bmin[k]; c[k] = bmax[k]; }
    b.GrowVec3(a); b.GrowVec3(c);
    return b;
}

// A PolyMesh's Transform (builtin/geom/polymesh/trace.go:20-74): nkeys matrices (one: used as given;
// more: TransformDecomp keys interpolated at the ray's time).
void orc_scene_set_transform(void* sp, int32_t geom, const float* mats, int32_t nkeys) {
    set_keys(((Scene*)sp)->geoms[geom], mats, nkeys);
}
// A geom left out of the scene's top level (reached only through instances or procs).
void orc_scene_set_hidden(void* sp, int32_t geom, int32_t hidden) { ((Scene*)sp)->geoms[geom]->hidden = hidden != 0; }

// GeomInstance (builtin/geom/instance/instance.go): geom `proto` (insertion index) under nkeys
// transform keys; bounds the box of bmin, bmax.  Returns the insertion index.
int orc_scene_add_instance(void* sp, int32_t proto, const float* mats, int32_t nkeys, const float* bmin, const float* bmax) {
    Scene* s = (Scene*)sp;
    Mesh* m = new Mesh();
    m->kind = GEOM_INSTANCE;
    m->children.push_back(s->geoms[proto]);
    set_keys(m, mats, nkeys);
    m->userBounds = user_box(bmin, bmax);
    s->geoms.push_back(m);
    return (int)s->geoms.size() - 1;
}
// Proc (builtin/geom/proc/proc.go) holding geoms children[0..nchild) in that order, under nkeys
// transform keys (0: the identity, proc.go:154-157); bounds the box of bmin, bmax.
int orc_scene_add_proc(void* sp, const int32_t* children, int32_t nchild, const float* mats, int32_t nkeys,
                       const float* bmin, const float* bmax) {
    Scene* s = (Scene*)sp;
    Mesh* m = new Mesh();
    m->kind = GEOM_PROC;
    for (int i = 0; i < nchild; i++) m->children.push_back(s->geoms[children[i]]);
    if (nkeys > 0) set_keys(m, mats, nkeys);
    else { M4 I; M4Identity(I); set_keys(m, I, 1); }
    m->userBounds = user_box(bmin, bmax);
    s->geoms.push_back(m);
    return (int)s->geoms.size() - 1;
}

// PreRender: per-mesh accel, then the scene's top level.
// leafMax of the static meshes' build (16: the reference's); call before orc_scene_commit.
void orc_scene_set_leaf_size(void* sp, int32_t leaf_max) {
    for (auto* g : ((Scene*)sp)->geoms) g->leafMax = leaf_max;
}

int orc_scene_commit(void* sp) {
    Scene* s = (Scene*)sp;
    try {
        for (auto* g : s->geoms) if (g->kind == GEOM_MESH) g->initAccel();
        s->initAccel();
    } catch (BuildError&) { return -1; }
    return 0;
}

// Query sizes: out[0]=top nodes, then per post-sort geom: nodes, ntris.
int orc_scene_info(void* sp, int32_t* n_top, int32_t* geom_order, int32_t* geom_nodes, int32_t* geom_tris, int32_t* geom_keys) {
    Scene* s = (Scene*)sp;
    *n_top = s->motion ? (int32_t)s->mnodes.size() : (int32_t)s->qbvh.size();
    for (size_t i = 0; i < s->geoms.size(); i++) {
        geom_order[i] = s->geomOrder[i];
        geom_nodes[i] = s->geoms[i]->nkeys == 1 ? (int32_t)s->geoms[i]->qbvh.size() : (int32_t)s->geoms[i]->mnodes.size();
        geom_tris[i] = s->geoms[i]->facecount;
        geom_keys[i] = s->geoms[i]->MotionKeys();
    }
    return (int)s->geoms.size();
}

// Copies a static geom's accel: nodes (128 B each), post-sort idxp (ntris*3), accel.idx (ntris).
int orc_scene_mesh_accel(void* sp, int32_t g, void* nodes_out, uint32_t* idxp_out, int32_t* perm_out) {
    Scene* s = (Scene*)sp;
    Mesh* m = s->geoms[g];
    if (m->nkeys != 1) return -1;
    std::memcpy(nodes_out, m->qbvh.data(), m->qbvh.size() * sizeof(Node));
    std::memcpy(idxp_out, m->idxp.data(), m->idxp.size() * sizeof(uint32_t));
    std::memcpy(perm_out, m->accelIdx.data(), m->accelIdx.size() * sizeof(int32_t));
    return 0;
}

// Copies a motion geom's accel: topology nodes (children[4], axis[3] as 8 int32 per node:
// axis0,axis1,axis2,c0,c1,c2,c3,0), boxes (nkeys*nnodes*24), accel.idx (ntris).
int orc_scene_mesh_motion_accel(void* sp, int32_t g, int32_t* topo_out, float* boxes_out, int32_t* perm_out) {
    Scene* s = (Scene*)sp;
    Mesh* m = s->geoms[g];
    if (m->nkeys == 1) return -1;
    for (size_t i = 0; i < m->mnodes.size(); i++) {
        const MNode& n = m->mnodes[i];
        int32_t* o = topo_out + i * 8;
        o[0] = n.axis0; o[1] = n.axis1; o[2] = n.axis2;
        for (int c = 0; c < 4; c++) o[3 + c] = n.children[c];
        o[7] = 0;
    }
    for (int k = 0; k < m->nkeys; k++) std::memcpy(boxes_out + (size_t)k * m->mnodes.size() * 24, m->mboxes[k].data(), m->mnodes.size() * 24 * sizeof(float));
    std::memcpy(perm_out, m->accelIdx.data(), m->accelIdx.size() * sizeof(int32_t));
    return 0;
}

// Copies the top-level (static) nodes.
int orc_scene_top_accel(void* sp, void* nodes_out) {
    Scene* s = (Scene*)sp;
    if (s->motion) return -1;
    std::memcpy(nodes_out, s->qbvh.data(), s->qbvh.size() * sizeof(Node));
    return 0;
}

int orc_scene_top_motion_accel(void* sp, int32_t* topo_out, float* boxes_out) {
    Scene* s = (Scene*)sp;
    if (!s->motion) return -1;
    for (size_t i = 0; i < s->mnodes.size(); i++) {
        const MNode& n = s->mnodes[i];
        int32_t* o = topo_out + i * 8;
        o[0] = n.axis0; o[1] = n.axis1; o[2] = n.axis2;
        for (int c = 0; c < 4; c++) o[3 + c] = n.children[c];
        o[7] = 0;
    }
    for (size_t k = 0; k < s->mboxes.size(); k++) std::memcpy(boxes_out + k * s->mnodes.size() * 24, s->mboxes[k].data(), s->mnodes.size() * 24 * sizeof(float));
    return 0;
}

// Traces n rays (core.TraceProbe semantics).  shadow != 0 sets RayTypeShadow.
// hits: n records.  counters (nullable): 4 int32 per ray (nodes, leaves, tris, max stack depth).
// nthreads <= 0: all OpenMP threads.  Returns 0.
int orc_trace(void* sp, const orc_ray* rays, int64_t n, int32_t shadow, orc_hit* hits, int32_t* counters, int32_t nthreads) {
    const Scene* s = (const Scene*)sp;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
    {
        Task* tk = new Task();
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1024)
#endif
        for (int64_t i = 0; i < n; i++) {
            Ray r{};
            for (int k = 0; k < 3; k++) { r.P[k] = rays[i].P[k]; r.D[k] = rays[i].D[k]; }
            r.Tclosest = rays[i].tmax;
            r.Time = rays[i].time;
            r.Type = shadow ? RayTypeShadow : RayTypeCamera;
            RaySetup(r);
            tk->StackTop = 0; tk->maxDepth = 0;
            r.task = tk;
            SceneHit sh{}; sh.geom = -1; sh.hr.elem = -1;
            bool hit = s->Trace(r, sh);
            orc_hit& h = hits[i];
            if (hit) { h.geom = sh.geom; h.prim = sh.hr.elem; h.t = r.Tclosest; h.u = sh.hr.Bu; h.v = sh.hr.Bv; }
            else { h.geom = -1; h.prim = -1; h.t = r.Tclosest; h.u = 0; h.v = 0; }
            if (counters) {
                counters[i * 4 + 0] = (int32_t)r.NodesT; counters[i * 4 + 1] = (int32_t)r.LeafsT;
                counters[i * 4 + 2] = (int32_t)r.TrisT; counters[i * 4 + 3] = tk->maxDepth;
            }
        }
        delete tk;
    }
    (void)nthreads;
    return 0;
}

// BVH-independent brute force: every triangle of every geom, in original face order, with the
// same watertight test (static: idxp branch bounds; motion: lerped vertices).  Hits report the
// post-sort geom slot and the post-sort triangle index of the winning face (ties: last tested wins).
int orc_brute(void* sp, const orc_ray* rays, int64_t n, orc_hit* hits, int32_t nthreads) {
    const Scene* s = (const Scene*)sp;
    // inverse permutations: original face -> post-sort slot
    std::vector<std::vector<int32_t>> inv(s->geoms.size());
    for (size_t g = 0; g < s->geoms.size(); g++) {
        const Mesh* m = s->geoms[g];
        inv[g].resize(m->facecount);
        for (int i = 0; i < m->facecount; i++) inv[g][m->accelIdx[i]] = i;
    }
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
#endif
    for (int64_t i = 0; i < n; i++) {
        Ray r{};
        for (int k = 0; k < 3; k++) { r.P[k] = rays[i].P[k]; r.D[k] = rays[i].D[k]; }
        r.Tclosest = rays[i].tmax; r.Time = rays[i].time;
        RaySetup(r);
        orc_hit h{-1, -1, r.Tclosest, 0, 0};
        for (size_t g = 0; g < s->geoms.size(); g++) {
            const Mesh* m = s->geoms[g];
            // build a one-face-at-a-time view through the post-sort slot of each original face
            for (int f = 0; f < m->facecount; f++) {
                HitRecord hr{};
                int slot = inv[g][f];
                bool ok;
                if (m->nkeys == 1) {
                    ok = m->TraceElems(r, hr, slot, 1);
                } else {
                    float k = r.Time * (float)(m->nkeys - 1);
                    float time = k - Floor(k);
                    int key = (int)Floor(k), key2 = (int)Ceil(k);
                    ok = m->TraceMotionElems(time, key, key2, r, hr, slot, 1);
                }
                if (ok) { h.geom = (int32_t)g; h.prim = hr.elem; h.t = r.Tclosest; h.u = hr.Bu; h.v = hr.Bv; }
            }
        }
        hits[i] = h;
    }
    (void)nthreads;
    return 0;
}

}  // extern "C"
