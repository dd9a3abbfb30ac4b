// mlod_internal.h — shared declarations of the HIP implementation of libmlod.so (gfx950).
//
// Data layout in HBM (one octree):
//   nodes, breadth-first order (depth, then Morton prefix), so the children of a node are
//   contiguous and the nodes of one depth form one range [level_off[d], level_off[d+1]):
//     geo  float4 {cx, cy, cz, r}            centre and radius of the cell box
//     hdr  uint4  {first, count|flags, A, father}
//            first  = first child (internal) or first sorted point (leaf)
//            count  = #children or #points, bits 27..30 depth, bit 31 leaf
//            A      = sum of point areas (float bits, reference summation order)
//     momP float4 x2 {ap, apn}, {an, app}     plain moments (global frame, FP32)
//     momC float4 x2 {pbar, C}, {N, V}        centred moments: mean point, sum a n,
//                                              sum a (p - pbar).n, sum a |p - pbar|^2
//     momD double4 x2 + double                plain moments in FP64
//   points, in leaf (sorted) order: pts float4 {p, a}, nrm float4 {n, 0}; order maps a
//   sorted position to the input index.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/mlod.h"

namespace mlod {

constexpr int kMaxDepth = 12;          // MAX_DEPTH, sources/apss/CPUUtils.h:20
constexpr uint32_t kLeafBit = 1u << 31;
constexpr uint32_t kCountMask = (1u << 27) - 1;

__host__ __device__ inline uint32_t hdr_count(uint32_t c) { return c & kCountMask; }
__host__ __device__ inline uint32_t hdr_depth(uint32_t c) { return (c >> 27) & 15u; }
__host__ __device__ inline bool hdr_leaf(uint32_t c) { return (c & kLeafBit) != 0; }

// Device-side view of one tree, passed by value to kernels.
struct TreeView {
  const float4* geo;
  const uint4* hdr;
  const float4* momP;
  const float4* momC;
  const double4* momD;
  const double* AD;
  const float4* pts;
  const float4* nrm;
  const uint2* span;  // per node: sorted points [x, y) of its subtree (leaf order is point order)
  uint32_t n, nn;
};

// Weight kernel in the form the kernels evaluate (derived on the host from mlod_kernel).
//   multi-Gaussian  w = sum_s exp2(d2 * a[s] + b[s])   (a = -log2(e)/(2 s^2), b = -3 log2 s)
//   Gaussian        w = exp2(d2 * a[0]) + c
//   rational        w = (c' / (d2 + c))^e,  c' = max(c, 1e-4)
struct alignas(16) KParams {
  float a[MLOD_MAX_SIGMAS];     // FP32: -log2(e) / (2 s^2)  (first, 16-byte aligned: read four at a time)
  float b[MLOD_MAX_SIGMAS];     // FP32: 1 / s^3
  int32_t type;
  int32_t ns;
  int32_t eint;      // integer exponent (1..16) or -1
  float c;           // constant_shift
  float cp;          // max(c, 1e-4)
  float e;           // exponent
  double ad[MLOD_MAX_SIGMAS];  // FP64: -1 / (2 s^2)
  double bd[MLOD_MAX_SIGMAS];  // FP64: 1 / s^3
};

// One run of a serving step launch (a resident range of one request's queries): its tree,
// kernel, parameters and place in the server's query pool.
struct ReqDesc {
  TreeView tv;
  KParams kp;       // read by reference from the launch table (wave-uniform scalar loads)
  float cutoff;
  float stepMax;
  uint32_t minDepth;
  uint32_t nIter;   // steps in this launch
  uint32_t slot0;   // first slot of the run in the admission staging (runs packed back to back)
  uint32_t count;   // queries
  uint32_t pad0;    // first pool slot of the run (a multiple of kBatchWG)
  uint32_t outOff;  // last steps: first result row of the run in the launch's host staging; else kNoOut
  float org[3];     // tree box origin and Morton scale
  float scale[3];
};
constexpr uint32_t kNoOut = 0xffffffffu;

// Where the runs taking their last step put their results: rows outOff + (index within the run,
// pidx of the pool slot) of `out` (positions, then normals from row cap on), in pinned,
// host-coherent memory; each workgroup of such a run then stores `seq` to flag[its launch block],
// so the host sees a run finish while the launch still runs others.
struct BatchOut {
  const uint32_t* pidx;
  float* out;
  uint32_t cap;
  uint32_t* flag;
  uint32_t seq;
};

struct QueryState {
  float4* q;         // sorted query positions (in place across project calls)
  float4* qn;        // normals of the last step
  uint32_t* perm;    // sorted position -> caller's index
  float* stage;      // m x 3 staging for un-permuted download
  float* hstage;     // pinned host staging (m x 3 x 2)
  uint32_t m = 0, cap = 0;
};

}  // namespace mlod

struct mlod_tree {
  int device = 0;
  uint32_t n = 0, nn = 0, depth = 0;
  uint32_t level_off[14] = {0};
  float bb[3] = {0, 0, 0}, BB[3] = {0, 0, 0};
  float radius = 0;
  uint64_t ties = 0;
  uint64_t bytes = 0;
  mlod_build_opts opts{12, 1, 0};
  // device buffers
  float4* geo = nullptr;
  uint4* hdr = nullptr;
  float4* momP = nullptr;
  float4* momC = nullptr;
  double4* momD = nullptr;
  double* AD = nullptr;
  float4* pts = nullptr;
  float4* nrm = nullptr;
  uint32_t* order = nullptr;
  float* box = nullptr;   // per node {bb, BB} (parity dumps, neighbour pruning)
  uint2* span = nullptr;  // per node [first, end) of its subtree's sorted points
  float* area = nullptr;  // per point area, input order (parity dumps)
  // parameters
  mlod_kernel kernel{};
  bool kernel_set = false;
  float cutoff = 2.0f;      // m_scalingProtectionSphere default, GPUUtils.h:63
  uint32_t min_depth = 0;   // m_minDepth default, GPUUtils.h:64
  int32_t semantics = MLOD_SEM_REF_GPU;
  int32_t traversal = MLOD_TRAVERSAL_LANE;
  int32_t precision = MLOD_PREC_FP32;
  int32_t mode = MLOD_MODE_MLOD;
  uint32_t knn = 20;
  float ball_radius = 0.1f;
  mlod::QueryState qs;
  mlod::TreeView view() const {
    return mlod::TreeView{geo, hdr, momP, momC, momD, AD, pts, nrm, span, n, nn};
  }
};

namespace mlod {

// error plumbing (api.cpp)
int fail(int code, const std::string& msg);
// argument checks of a weight kernel (Kernel.h:37-46): type, sigma count, positive sigmas (api.cpp)
int validate_kernel(const mlod_kernel* k);
std::string g_last_error();  // the calling thread's last message
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define MLOD_HIP(expr)                                                         \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) return ::mlod::hip_fail(e_, #expr, __FILE__, __LINE__); \
  } while (0)

// Workgroups are dispatched round-robin over the 8 XCDs of the chip, each with its own L2.
// Queries are Morton-sorted, so consecutive workgroups touch the same octree nodes; this maps
// the hardware workgroup index to a logical one such that each XCD walks one contiguous share
// of the query range (a bijection on [0, gridDim.x)), keeping each L2's working set disjoint.
#if defined(__HIPCC__)
__device__ __forceinline__ uint32_t xcd_block() {
  constexpr uint32_t kXcds = 8;
  const uint32_t b = blockIdx.x, nb = gridDim.x;
  const uint32_t x = b % kXcds, idx = b / kXcds, q = nb / kXcds, r = nb % kXcds;
  return x * q + (x < r ? x : r) + idx;
}
#endif

// build.hip
int build_tree(mlod_tree* t, const float* d_xyz, const float* d_nrm, hipStream_t s);
int dump_tree(const mlod_tree* t, int32_t* depth, float* bb6, float* center3, float* radius, int32_t* father,
              int32_t* first_child, int32_t* next_brother, int32_t* num_children, uint32_t* leaf_count,
              uint32_t* leaf_idx, float* areas, float* moments9);

// project.hip
KParams make_kparams(const mlod_kernel& k);
int launch_project(mlod_tree* t, uint32_t n, uint32_t n_iter, float step_max, hipStream_t s);
int launch_count(mlod_tree* t, uint32_t n, unsigned long long* d_out4, hipStream_t s);

// serving step launches (project.hip): the server's query pool is cut into blocks of kBatchWG
// slots, one workgroup each; runs of every kernel type share one launch
constexpr uint32_t kBatchWG = 128;
// One launch over nblk pool blocks: block b holds queries [d_blocks[b].y * kBatchWG, +kBatchWG) of run
// d_reqs[d_blocks[b].x] (those below its pad0 + count); prio != 0 raises the wavefronts' priority.
// d_qd: FP64 positions of the pool slots (the FP64 form carries them between launches), else NULL.
int launch_project_batch(const ReqDesc* d_reqs, const uint2* d_blocks, uint32_t nblk, float4* d_q, float4* d_qn,
                         double4* d_qd, const BatchOut& out, uint32_t prio, int semantics, int precision,
                         hipStream_t s);

// neighbors.hip
int launch_neighbors(mlod_tree* t, uint32_t n, uint32_t n_iter, hipStream_t s);

// api.cpp: n_iter steps of the first n queries of t->qs in the tree's neighbourhood mode
// (launches only)
int project_uploaded(mlod_tree* t, uint32_t n, uint32_t n_iter, float step_max, hipStream_t s);

// queries.hip
// Buffers of the query sort: Morton codes (two arrays of m), caller indices (m) and rocPRIM's
// temporary storage (tmpBytes, at least queries_sort_bytes(m)).
struct SortBufs {
  uint32_t* code;
  uint32_t* code2;
  uint32_t* idx;
  void* tmp;
  size_t tmpBytes;
};
size_t queries_sort_bytes(uint32_t m);  // 0 if the size query fails
// Morton order of m device points in the tree's box: q / qn (m float4 each) receive the sorted
// points and zero normals, perm (m) the caller index of each sorted slot.  Launches only (no
// allocation, no synchronisation).
int queries_order(const mlod_tree* t, const float* d_xyz, uint32_t m, float4* q, float4* qn, uint32_t* perm,
                  const SortBufs& sb, hipStream_t s);
int queries_set(mlod_tree* t, const float* xyz, uint32_t m, bool device_src, hipStream_t s);
int queries_get(mlod_tree* t, uint32_t n, float* out_pts, float* out_nmls, hipStream_t s);
int queries_scatter(mlod_tree* t, uint32_t n, float* d_pts, float* d_nmls, hipStream_t s);
void queries_free(mlod_tree* t);
int queries_reorganize(mlod_tree* t, const uint32_t* invalid, uint32_t n, uint32_t end_valid, hipStream_t s);
int queries_fit_kinds(mlod_tree* t, uint32_t n, uint8_t* out, hipStream_t s);

}  // namespace mlod
