/* mlod.h — C ABI of the MI355X-native MLoD / APSS projection library (libmlod.so).
 *
 * The library replaces the reference's projection plugin interface, `class APSS`
 * (CorentinMercier/MlodSurfaces, sources/apss/GPUUtils.h:22-67), implemented there by
 * `Fastapss` (sources/apss/Fastapss.h:32-51) and `GlobalAPSS` (sources/apss/GlobalAPSS.h:19-43),
 * together with the octree it projects on, `APSSOctree` (sources/apss/CPUUtils.h:166-229).
 * Each entry point below names the reference member it stands in for.
 *
 * Conventions: plain pointers and sizes, no C++ or torch types.  Every call returns
 * MLOD_OK (0) or a negative status; mlod_last_error() gives the message of the calling
 * thread's last failure (the reference instead calls exit() from gpuErrchk,
 * GPUUtils.cu:32-41).  `stream` is a hipStream_t passed as void* (NULL = the default
 * stream).  One handle must not be used from two threads at once.
 */
#ifndef MLOD_H
#define MLOD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlod_tree mlod_tree;

enum {
  MLOD_OK = 0,
  MLOD_ERR_INVALID = -1,     /* bad argument */
  MLOD_ERR_HIP = -2,         /* HIP runtime error */
  MLOD_ERR_OOM = -3,         /* device allocation failed */
  MLOD_ERR_UNSUPPORTED = -4, /* option not implemented on the GPU path */
  MLOD_ERR_STATE = -5        /* call out of order (e.g. project before queries) */
};

/* Kernel::KernelType numbering (sources/apss/Kernel.h:37), as persisted in .params files. */
enum { MLOD_KERNEL_GAUSSIAN = 0, MLOD_KERNEL_SINGULAR = 1, MLOD_KERNEL_GAUSSIAN_MULTIPLE = 2 };

/* Weight kernel (Kernel fields, Kernel.h:38-46).  sigma / sigmas are in the point-cloud
 * frame, i.e. already multiplied by the octree bbox radius as scene::setNewKernel does
 * (scene.cpp:1025-1033); constant_shift is not scaled. */
#define MLOD_MAX_SIGMAS 32
typedef struct {
  int32_t type;
  float sigma;          /* GAUSSIAN */
  float exponent;       /* SINGULAR ("rational") */
  float constant_shift; /* GAUSSIAN / SINGULAR */
  int32_t n_sigmas;     /* GAUSSIAN_MULTIPLE */
  float sigmas[MLOD_MAX_SIGMAS];
} mlod_kernel;

/* APSSOctree::build options (CPUUtils.h:176-180).  Fields not set in an initializer are 0. */
typedef struct {
  uint32_t max_depth;        /* <= 12 (MAX_DEPTH, CPUUtils.h:20) */
  uint32_t max_pts_per_leaf; /* maxNumberOfPointsPerLeaf: a node holding at most this many points
                                is a leaf (CPUUtils.cpp:279-283); the reference default is 1 */
  int32_t tie_policy;        /* exact-midpoint ties: 0 = low child, 1 = high child
                                (the reference draws rand(), CPUUtils.cpp:299-308) */
  uint32_t use_root_box;     /* updateRootBoundingBox = false (CPUUtils.cpp:547): the root's box is
                                root_box instead of the points' bounding box */
  float root_box[6];         /* {bb.x, bb.y, bb.z, BB.x, BB.y, BB.z}, finite, bb <= BB */
} mlod_build_opts;

/* Which of the reference's two projection functions to compute (they differ,
 * Fastapss.cu vs CPUUtils.cpp: leaf points, parent share, plane threshold). */
enum { MLOD_SEM_REF_GPU = 0, MLOD_SEM_REF_CPU = 1 };

/* Arithmetic form of the projection. */
enum {
  MLOD_PREC_FP32 = 0,      /* plain: global-frame FP32 moments and FP32 fit (the reference's form) */
  MLOD_PREC_FP32C = 1,     /* precision: node-centred moments, query-recentred compensated FP32
                              accumulation, FP64 sphere fit */
  MLOD_PREC_FP64 = 2       /* plain FP64 everywhere */
};

/* How the MLoD walk is laid on the GPU (no counterpart in the reference, whose kernel walks one
 * thread per query, Fastapss.cu:160-333).  LANE: each lane walks its own query's tree path
 * (the default, and the faster on every workload measured).  WAVE: one walk per wavefront over
 * the union of its 64 lanes' paths, node records in scalar registers, each lane taking part in
 * the nodes its own walk visits (near/far decided per lane and combined by ballot).  Both add
 * every lane's terms in the same order: the FP64 form gives the same bits, the FP32 forms the
 * same results to their rounding (the two loops may fuse a multiply-add differently).  Serving
 * launches use the lane walk. */
enum { MLOD_TRAVERSAL_LANE = 0, MLOD_TRAVERSAL_WAVE = 1 };

/* Neighbourhood: MLoD octree (Fastapss), all points (GlobalAPSS), k nearest or ball
 * (APSSOctree::knnMode / ballMode, CPUUtils.h:191-192). */
enum { MLOD_MODE_MLOD = 0, MLOD_MODE_GLOBAL = 1, MLOD_MODE_KNN = 2, MLOD_MODE_BALL = 3 };

typedef struct {
  uint32_t n_points;
  uint32_t n_nodes;
  uint32_t depth;           /* deepest node depth (APSSOctree computeDepth) */
  uint32_t level_offset[14];/* BFS index of the first node of each depth; [depth+1] = n_nodes */
  float bbox[6];            /* root bb (xyz), BB (xyz) */
  float radius;             /* root bbox radius: default step bound and sigma scale */
  uint64_t ties;            /* exact-midpoint ties met by the build */
  uint64_t device_bytes;    /* HBM held by the tree (nodes, points, moments) */
} mlod_tree_info;

const char* mlod_last_error(void);

/* APSSOctree::build + Fastapss::copyToGPU (CPUUtils.cpp:540-569, Fastapss.cu:457-484):
 * builds the octree, point areas and node moments on the GPU from host arrays
 * (n x 3 floats each).  opts may be NULL (max_depth 12, 1 point per leaf, ties low). */
int mlod_tree_create(const float* xyz, const float* nrm, uint32_t n, const mlod_build_opts* opts, void* stream,
                     mlod_tree** out);
/* Same from device-resident arrays. */
int mlod_tree_create_device(const float* d_xyz, const float* d_nrm, uint32_t n, const mlod_build_opts* opts,
                            void* stream, mlod_tree** out);
/* Fastapss::eraseFromGPU + APSS destructor (Fastapss.cu:486-511, GPUUtils.h:24). */
int mlod_tree_destroy(mlod_tree* t);
int mlod_tree_get_info(const mlod_tree* t, mlod_tree_info* out);

/* APSS::updateKernel (GPUUtils.cu:466-472). */
int mlod_set_kernel(mlod_tree* t, const mlod_kernel* k);
/* APSS::setScalingProtectionSphere / setMinDepth (GPUUtils.h:40-43) plus the semantics
 * and arithmetic form. */
int mlod_set_params(mlod_tree* t, float cutoff, uint32_t min_depth, int32_t semantics, int32_t precision);
/* APSSOctree::knnMode / ballMode (CPUUtils.cpp:674-688); ball_radius < 0 => 3 * sigmas[0]. */
int mlod_set_mode(mlod_tree* t, int32_t mode, uint32_t knn, float ball_radius);
/* No counterpart in the reference: MLOD_TRAVERSAL_LANE (default) or MLOD_TRAVERSAL_WAVE for the
 * MLoD projection steps of this tree (mlod_project, mlod_project_points*, mlod_count_visits). */
int mlod_set_traversal(mlod_tree* t, int32_t traversal);

/* APSS::copyPointsToGPU (GPUUtils.cu:406-411): uploads m query points (host, m x 3). */
int mlod_queries_upload(mlod_tree* t, const float* xyz, uint32_t m, void* stream);
/* Same from a device array (m x 3 floats). */
int mlod_queries_set_device(mlod_tree* t, const float* d_xyz, uint32_t m, void* stream);
/* APSS::erasePointsFromGPU (GPUUtils.cu:452-464). */
int mlod_queries_erase(mlod_tree* t);

/* APSS::project (GPUUtils.h:26-27, Fastapss.cu:439-455): n_iter projection steps of the
 * first n uploaded queries, in place, so the next call continues from the projected
 * positions.  step_max < 0 => root bbox radius.  out_pts / out_nmls (host, n x 3) are
 * written when non-NULL (then the call waits for the stream). */
int mlod_project(mlod_tree* t, uint32_t n, uint32_t n_iter, float step_max, float* out_pts, float* out_nmls,
                 void* stream);
/* Copies the current positions / normals of the first n queries to host arrays. */
int mlod_queries_download(mlod_tree* t, uint32_t n, float* out_pts, float* out_nmls, void* stream);

/* Which branch of the sphere fit (Fastapss.cu:354-395 / CPUUtils.cpp:739-782) the last MLoD or
 * global projection step of each of the first n queries took: 0 sphere, 1 plane (|u4| below the
 * semantics' threshold), 2 invalid fit (query kept, zero normal).  Diagnostic, for counting
 * plane/sphere switches between arithmetic forms. */
int mlod_queries_fit_kinds(mlod_tree* t, uint32_t n, uint8_t* out, void* stream);

/* APSS::projectCPU (GPUUtils.h:33-34; APSSOctree::projectCPU, CPUUtils.cpp:621-636), batched:
 * n_iter steps of m host points (m x 3) with the tree's current settings, results to out_pts /
 * out_nmls (host, m x 3).  The uploaded queries are left untouched.  The reference's CPU function
 * is selected with MLOD_SEM_REF_CPU in mlod_set_params. */
int mlod_project_points(mlod_tree* t, const float* xyz, uint32_t m, uint32_t n_iter, float step_max, float* out_pts,
                        float* out_nmls, void* stream);

/* No counterpart in the reference.  mlod_project_points for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the points and results are device
 * arrays (m x 3 floats each; an output may be NULL) and m <= MLOD_NOALLOC_MAX_QUERIES (larger m
 * fails with MLOD_ERR_INVALID).  The call allocates nothing: its working buffers are a fixed
 * 16 MiB device array per GPU, held until the process ends, so calls that overlap on one GPU
 * race (calls on one stream are ordered).  It only launches work on `stream`: the results are
 * in place once the stream has completed.  Same results as mlod_project_points. */
#define MLOD_NOALLOC_MAX_QUERIES 262144u
int mlod_project_points_noalloc(mlod_tree* t, const float* d_xyz, uint32_t m, uint32_t n_iter, float step_max,
                                float* d_out_pts, float* d_out_nmls, void* stream);

/* APSS::reorganizePoints (GPUUtils.cu:436-450): for i = 0 .. n-1 in order, the uploaded queries
 * invalid[i] and end_valid - 1 - i (caller's indices) exchange positions; normals stay with their
 * index, as in the reference (which exchanges positions only). */
int mlod_queries_reorganize(mlod_tree* t, const uint32_t* invalid, uint32_t n, uint32_t end_valid, void* stream);

/* Octree dump in BFS order (the flattening of Fastapss copy(), Fastapss.cu:97-150, via
 * getRelations, CPUUtils.cpp:213-234), for parity checks.  Any pointer may be NULL.
 * leaf_idx receives, for each leaf in BFS order, its point indices in input order. */
int mlod_tree_dump(const mlod_tree* t, int32_t* depth, float* bb6, float* center3, float* radius, int32_t* father,
                   int32_t* first_child, int32_t* next_brother, int32_t* num_children, uint32_t* leaf_count,
                   uint32_t* leaf_idx, float* areas, float* moments9);

/* Traversal counters of one step over the uploaded queries (untimed).  out6 receives: far
 * nodes accepted, leaf points summed, child nodes tested, near nodes adding their own share,
 * traversal loop iterations summed over queries, and the lane slots those iterations were
 * issued in (64 x the wavefronts' traversal loop iterations), so out6[4] / out6[5] is the
 * SIMD lane efficiency. */
int mlod_count_visits(mlod_tree* t, uint32_t n, uint64_t* out6, void* stream);

/* ---------------------------------------------------------------------------------------
 * Serving.  No counterpart in the reference (its harness projects one batch at a time with
 * blocking calls, scene.cpp:916-966).  A server keeps registered octrees resident, takes
 * projection requests from any thread and schedules them a projection step at a time
 * (continuous batching): requests are cut into chunks of at most chunk_queries queries and
 * queued by class (interactive: at most 2M projected points, ahead of bulk, which still gets a
 * launch at least every 20 ms) and deadline (arrival time plus 1 ms per 20k projected points);
 * each of `inflight` streams keeps one step launch of at most max_batch_queries queries on the
 * device and, when it completes, takes the most urgent resident or queued work for the next one.
 * Queries stay in a device pool between steps; bulk launches take two steps, interactive ones
 * up to 16 at raised wavefront priority; a request's results reach its buffers as soon as its
 * own workgroups finish.  Chunks are admitted to the pool by memory footprint (36 B per query
 * slot, 68 B in the FP64 form) and wait in their queue while it is full.  Each query gives the
 * same result alone, in any mix and through mlod_project.  Registered trees must outlive the
 * server; request buffers must stay valid until the request's completion is polled. */
typedef struct mlod_server mlod_server;
typedef struct {
  uint64_t ticket;
  double latency_ms;   /* submit -> all results copied to the caller's buffers */
  double queue_ms;     /* submit -> first step launch of the request */
  double completed_s;  /* completion time since server creation */
  int32_t status;      /* 0 completed, 1 cancelled (mlod_server_cancel) */
  int32_t reserved;
} mlod_server_done;

/* chunk_queries = 0: max_batch_queries / 4 (rounded to 128); precision: any MLOD_PREC_* form */
int mlod_server_create(uint32_t max_batch_queries, uint32_t chunk_queries, uint32_t inflight, int32_t semantics,
                       int32_t precision, mlod_server** out);
int mlod_server_destroy(mlod_server* s);
/* Registers a resident tree; MLOD_ERR_OOM when it would exceed the server's HBM budget. */
int mlod_server_add_tree(mlod_server* s, mlod_tree* t, uint32_t* tree_id);
/* HBM budget of the server (0 = none): the registered trees, the staging buffers and the query
 * pool together hold at most hbm_bytes of device memory.  The pool is reallocated to what the
 * budget leaves (at most its default size); MLOD_ERR_OOM when that is under two chunks, or when
 * the trees already exceed it.  Only while no request is outstanding (else MLOD_ERR_STATE). */
int mlod_server_set_budget(mlod_server* s, uint64_t hbm_bytes);
/* [budget, tree bytes, staging bytes, pool bytes, pool query slots] */
int mlod_server_memory(mlod_server* s, uint64_t* out5);
/* n_iter projection steps of m queries (host, m x 3, any m >= 1) on tree tree_id with kernel k
 * (checked as by mlod_set_kernel, up to 32 sigmas); results go to out_pts / out_nmls (host, m x 3,
 * either may be NULL).  cutoff >= 0; step_max < 0 => the tree's radius.  n_iter = 0 completes at
 * once with the queries unchanged and zero normals.  MLOD_ERR_OOM when a chunk of the request can
 * never be resident in the (budgeted) pool. */
int mlod_server_submit(mlod_server* s, uint32_t tree_id, const mlod_kernel* k, float cutoff, uint32_t min_depth,
                       uint32_t n_iter, float step_max, const float* queries, uint32_t m, float* out_pts,
                       float* out_nmls, uint64_t* ticket);
/* Cancels an outstanding request: its queued chunks and its runs waiting in the pool are dropped
 * at the next launch boundary, runs in a launch are discarded when it completes; the request then
 * completes with status 1 (its buffers hold the runs handed back before, if any). */
int mlod_server_cancel(mlod_server* s, uint64_t ticket);
/* Completed requests since the last poll (waits up to timeout_s seconds for one). */
int mlod_server_poll(mlod_server* s, mlod_server_done* out, uint32_t cap, uint32_t* n_out, double timeout_s);
/* step launches, requests completed, queries completed, chunks queued */
int mlod_server_stats(mlod_server* s, uint64_t* out4);

/* ---------------------------------------------------------------------------------------
 * Adaptive octree: OctreeGenerator (sources/scene/upsample.h:26-52).  Cells are octree ids
 * (depth, x, y, z) (octree_id, sources/dual-contouring/octree.h:18-44) returned as int32 rows in
 * the reference's node_comparator order (depth, then z, y, x; octree.h:147-154).  generate works
 * in the [-1, 1]^3 frame of the reference (upsample.cpp:319-322) and projects with the tree's
 * current kernel, parameters and neighbourhood mode; it overwrites the tree's uploaded queries
 * (the reference re-uploads them too, upsample.cpp:331-332). */
typedef struct mlod_generator mlod_generator;
/* OctreeGenerator(APSS*, const Kernel&) (upsample.h:28); t may be NULL for generate_from_points. */
int mlod_generator_create(mlod_tree* t, mlod_generator** out);
int mlod_generator_destroy(mlod_generator* g);
/* OctreeGenerator::generate (upsample.cpp:228-251), max_depth <= 12.  keep_total != 0 also
 * collects total_tree, the cells crossed (scene.cpp:162-164). */
int mlod_generate(mlod_generator* g, uint32_t max_depth, int32_t keep_total, void* stream);
/* OctreeGenerator::generateFromPoints (upsample.cpp:128-144): the max_depth cells enclosing n
 * points (host array, or device when device_src != 0) of box6 = (min xyz, max xyz); NULL or an
 * empty box means [-1, 1]^3. */
int mlod_generate_from_points(mlod_generator* g, const float* xyz, uint32_t n, int32_t device_src, uint32_t max_depth,
                              const float* box6, void* stream);
/* [leaves, total cells, max relative depth, centres projected, projection calls,
 *  projections with a unit normal but a non-finite position (mapped to the root)] */
int mlod_generator_info(const mlod_generator* g, uint64_t* out6);
/* which = 0: the generated cells (generate's result); 1: total_tree.  out_ids4: n x 4 int32. */
int mlod_generator_cells(mlod_generator* g, int32_t which, int32_t* out_ids4, void* stream);
/* Centres of the generated cells (aabb_octree::center, octree.h:952-955) in the generator's
 * frame: copied to out_xyz (host, n x 3) when non-NULL; *d_xyz receives the device array, valid
 * until the next call on g.  These are the query points of the reference's upsampling. */
int mlod_generator_centres(mlod_generator* g, float* out_xyz, const float** d_xyz, void* stream);

/* ---------------------------------------------------------------------------------------
 * Dual contouring of the adaptive octree: scene::computeDualContouring (scene.cpp:1255-1400),
 * the reference's mesh output from the projection.  Stages:
 *   mlod_dc_build    uniformize(max_depth, cells, ..., cut) (upsample.cpp:262-272): re-root,
 *                    grade, regularize up and siblings, optional cut; cells are n x (depth, x, y, z)
 *                    int32 (host), usually mlod_generator_cells(g, 1, ...);
 *   corner values    mlod_dc_project_corners: every leaf corner projected one step on tree t with
 *                    its current kernel and settings, value (c - p')·n', NaN where |n'|^2 < 0.5
 *                    (scene.cpp:1346-1387); or mlod_dc_set_corner_values (host, one per corner in
 *                    mlod_dc_corners order);
 *   mlod_dc_extract  dual_contour::extract (dual_contouring.h:59-87) over the minimal edges:
 *                    a vertex per cell of a sign-changing minimal edge (mean of its edges' zero
 *                    crossings, the cell centre when there is none), two triangles per edge.
 * Positions are in the frame of the projection ([-1, 1]^3 of the adaptive octree: new2bbox);
 * unit positions in the uniformized octree's [0, 1]^3.  Vertices are numbered in node order and
 * triangles in (leaf, edge) order; the reference numbers them in its traversal's order. */
typedef struct mlod_dc mlod_dc;
int mlod_dc_create(mlod_dc** out);
int mlod_dc_destroy(mlod_dc* dc);
int mlod_dc_build(mlod_dc* dc, const int32_t* ids4, uint32_t n, uint32_t max_depth, int32_t cut, void* stream);
/* uniformizeFromPoints(cells, ...) (upsample.cpp:253-259), the variant scene::computeDualContouring
 * takes for an octree built on points (scene.cpp:1279-1299: OctreeGenerator::generateFromPoints
 * cells, mlod_generate_from_points): grade, regularize up and siblings, in the [-1, 1]^3 frame
 * (no re-rooting, no cut). */
int mlod_dc_build_from_points(mlod_dc* dc, const int32_t* ids4, uint32_t n, void* stream);
/* [nodes, leaves, corners, max depth, diag, origin x, origin y, origin z, extracted] */
int mlod_dc_info(const mlod_dc* dc, double* out9);
/* the uniformized octree in node_comparator order: ids4 (n x 4) and leaf flags (either may be NULL) */
int mlod_dc_nodes(const mlod_dc* dc, int32_t* ids4, uint8_t* is_leaf, void* stream);
/* the distinct leaf corners: integer positions at the finest depth (n x 3) and positions in the
 * projection frame (n x 3 floats); either may be NULL */
int mlod_dc_corners(const mlod_dc* dc, int32_t* xyz3, float* pos, void* stream);
int mlod_dc_project_corners(mlod_dc* dc, mlod_tree* t, void* stream);
int mlod_dc_set_corner_values(mlod_dc* dc, const float* f, void* stream);
int mlod_dc_corner_values(const mlod_dc* dc, float* f, void* stream);
int mlod_dc_extract(mlod_dc* dc, uint32_t* n_verts, uint32_t* n_tris, void* stream);
/* vertices (projection frame and unit frame, n x 3 floats), the octree node of each vertex, and
 * triangles (n x 3 vertex indices); any may be NULL */
int mlod_dc_mesh(const mlod_dc* dc, float* verts, float* unit_verts, uint32_t* vert_node, uint32_t* tris, void* stream);

/* The surface mesh of scene::computeDualContouring after extraction (scene.cpp:1417-1452), kept on
 * the device: the extracted vertices projected `steps` steps on tree t (its kernel, settings and
 * mode; the default step bound), then `rounds` times smoothing, subdivision and projection again
 * (the reference: steps 3, rounds 2).  The same arithmetic as mlod_mesh_smooth,
 * mlod_mesh_subdivide and mlod_project_points on the host arrays, so the same bits.
 * mlod_dc_refined_mesh copies vertices, normals (n x 3 floats, projection frame) and triangles
 * (n x 3); any may be NULL. */
int mlod_dc_refine(mlod_dc* dc, mlod_tree* t, uint32_t steps, uint32_t rounds, uint32_t* n_verts, uint32_t* n_tris,
                   void* stream);
int mlod_dc_refined_mesh(const mlod_dc* dc, float* verts, float* normals, uint32_t* tris, void* stream);

/* Mesh passes of the surface output (host arrays in and out), the reference's results bit for bit:
 * mesh::smoothMesh (upsample.h:55-76): every vertex moved to the mean of the other vertices of
 * its triangles (NaN for a vertex no triangle uses); mesh::subdivideMesh (upsample.h:78-133): a
 * vertex at the middle of every edge, numbered in order of first use, each triangle split in
 * four (out_tris: 4 nt triangles; out_verts must hold nv + 3 nt vertices; *out_nv receives the
 * count). */
int mlod_mesh_smooth(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, float* out_verts, void* stream);
int mlod_mesh_subdivide(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, float* out_verts,
                        uint32_t out_cap, uint32_t* out_tris, uint32_t* out_nv, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MLOD_H */
