// api.cpp — the extern "C" entry points of libmlod.so (declared in include/mlod.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include <vector>

#include "mlod_internal.h"

namespace mlod {

static thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

std::string g_last_error() { return g_err; }

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  g_err = std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) + ")";
  return e == hipErrorOutOfMemory ? MLOD_ERR_OOM : MLOD_ERR_HIP;
}

static void free_tree(mlod_tree* t) {
  if (!t) return;
  queries_free(t);
  void* bufs[] = {t->geo, t->hdr, t->momP, t->momC, t->momD, t->AD, t->pts, t->nrm, t->order, t->box, t->area, t->span};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  delete t;
}

// BBOX::radius (sources/apss/CPUUtils.h:83-85): sqrt(distance2(BB, bb)) / 2 in float, glm
// dot order, no contraction.
static float bbox_radius(const float bb[3], const float BB[3]) {
  volatile float dx = bb[0] - BB[0], dy = bb[1] - BB[1], dz = bb[2] - BB[2];
  volatile float x2 = dx * dx, y2 = dy * dy, z2 = dz * dz;
  volatile float s = x2 + y2;
  volatile float d2 = s + z2;
  return std::sqrt((float)d2) / 2.f;
}

int validate_kernel(const mlod_kernel* k) {
  if (!k) return fail(MLOD_ERR_INVALID, "kernel is NULL");
  if (k->type < 0 || k->type > 2) return fail(MLOD_ERR_INVALID, "kernel type must be 0, 1 or 2 (Kernel.h:37)");
  if (k->type == MLOD_KERNEL_GAUSSIAN_MULTIPLE && (k->n_sigmas < 1 || k->n_sigmas > MLOD_MAX_SIGMAS))
    return fail(MLOD_ERR_INVALID, "multi-Gaussian kernel needs 1..32 sigmas");
  if (k->type == MLOD_KERNEL_GAUSSIAN_MULTIPLE)
    for (int s = 0; s < k->n_sigmas; ++s)
      if (!(k->sigmas[s] > 0.f) || !std::isfinite(k->sigmas[s]))
        return fail(MLOD_ERR_INVALID, "sigmas must be positive and finite");
  if (k->type == MLOD_KERNEL_GAUSSIAN && (!(k->sigma > 0.f) || !std::isfinite(k->sigma)))
    return fail(MLOD_ERR_INVALID, "sigma must be positive and finite");
  if (k->type == MLOD_KERNEL_SINGULAR && !std::isfinite(k->exponent))
    return fail(MLOD_ERR_INVALID, "exponent must be finite");
  return MLOD_OK;
}

}  // namespace mlod

using namespace mlod;

extern "C" {

const char* mlod_last_error(void) { return g_err.c_str(); }

int mlod_tree_create_device(const float* d_xyz, const float* d_nrm, uint32_t n, const mlod_build_opts* opts, void* stream,
                            mlod_tree** out) {
  if (!out) return fail(MLOD_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (n == 0) return fail(MLOD_ERR_INVALID, "empty point set (APSSOctree::build asserts n > 0, CPUUtils.cpp:546)");
  if (!d_xyz || !d_nrm) return fail(MLOD_ERR_INVALID, "points or normals are NULL");
  if (n > (1u << 27) - 1) return fail(MLOD_ERR_UNSUPPORTED, "more than 2^27 - 1 points");
  mlod_build_opts o{12, 1, 0, 0, {0, 0, 0, 0, 0, 0}};
  if (opts) o = *opts;
  if (o.max_depth < 1 || o.max_depth > (uint32_t)kMaxDepth)
    return fail(MLOD_ERR_INVALID, "max_depth must be in 1..12 (MAX_DEPTH, CPUUtils.h:20)");
  if (o.tie_policy != 0 && o.tie_policy != 1) return fail(MLOD_ERR_INVALID, "tie_policy must be 0 or 1");
  if (o.use_root_box) {
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(o.root_box[a]) || !std::isfinite(o.root_box[3 + a]) || o.root_box[a] > o.root_box[3 + a])
        return fail(MLOD_ERR_INVALID, "root_box must be finite with bb <= BB on every axis");
  }
  mlod_tree* t = new mlod_tree();
  t->n = n;
  t->opts = o;
  (void)hipGetDevice(&t->device);
  const int r = build_tree(t, d_xyz, d_nrm, (hipStream_t)stream);
  if (r != MLOD_OK) {
    free_tree(t);
    return r;
  }
  t->radius = bbox_radius(t->bb, t->BB);
  *out = t;
  return MLOD_OK;
}

int mlod_tree_create(const float* xyz, const float* nrm, uint32_t n, const mlod_build_opts* opts, void* stream,
                     mlod_tree** out) {
  if (!out) return fail(MLOD_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (n == 0) return fail(MLOD_ERR_INVALID, "empty point set (APSSOctree::build asserts n > 0, CPUUtils.cpp:546)");
  if (!xyz || !nrm) return fail(MLOD_ERR_INVALID, "points or normals are NULL");
  float *dx = nullptr, *dn = nullptr;
  const size_t b = (size_t)n * 3 * sizeof(float);
  if (hipMalloc((void**)&dx, b) != hipSuccess || hipMalloc((void**)&dn, b) != hipSuccess) {
    (void)hipFree(dx), (void)hipFree(dn);
    return fail(MLOD_ERR_OOM, "input staging: hipMalloc failed");
  }
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(dx, xyz, b, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dn, nrm, b, hipMemcpyHostToDevice, s);
  int r = e == hipSuccess ? mlod_tree_create_device(dx, dn, n, opts, stream, out) : hip_fail(e, "input upload", __FILE__, __LINE__);
  (void)hipStreamSynchronize(s);
  (void)hipFree(dx), (void)hipFree(dn);
  return r;
}

int mlod_tree_destroy(mlod_tree* t) {
  if (!t) return MLOD_OK;
  (void)hipDeviceSynchronize();
  free_tree(t);
  return MLOD_OK;
}

int mlod_tree_get_info(const mlod_tree* t, mlod_tree_info* out) {
  if (!t || !out) return fail(MLOD_ERR_INVALID, "NULL argument");
  out->n_points = t->n;
  out->n_nodes = t->nn;
  out->depth = t->depth;
  for (int d = 0; d < 14; ++d) out->level_offset[d] = t->level_off[d];
  for (int a = 0; a < 3; ++a) {
    out->bbox[a] = t->bb[a];
    out->bbox[3 + a] = t->BB[a];
  }
  out->radius = t->radius;
  out->ties = t->ties;
  out->device_bytes = t->bytes;
  return MLOD_OK;
}

int mlod_set_kernel(mlod_tree* t, const mlod_kernel* k) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  const int r = validate_kernel(k);
  if (r) return r;
  t->kernel = *k;
  t->kernel_set = true;
  return MLOD_OK;
}

int mlod_set_params(mlod_tree* t, float cutoff, uint32_t min_depth, int32_t semantics, int32_t precision) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  if (!(cutoff >= 0.f)) return fail(MLOD_ERR_INVALID, "cutoff must be >= 0");
  if (semantics != MLOD_SEM_REF_GPU && semantics != MLOD_SEM_REF_CPU) return fail(MLOD_ERR_INVALID, "bad semantics");
  if (precision < MLOD_PREC_FP32 || precision > MLOD_PREC_FP64) return fail(MLOD_ERR_INVALID, "bad precision");
  t->cutoff = cutoff;
  t->min_depth = min_depth;
  t->semantics = semantics;
  t->precision = precision;
  return MLOD_OK;
}

int mlod_set_traversal(mlod_tree* t, int32_t traversal) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  if (traversal != MLOD_TRAVERSAL_LANE && traversal != MLOD_TRAVERSAL_WAVE) return fail(MLOD_ERR_INVALID, "bad traversal");
  t->traversal = traversal;
  return MLOD_OK;
}

int mlod_set_mode(mlod_tree* t, int32_t mode, uint32_t knn, float ball_radius) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  if (mode < MLOD_MODE_MLOD || mode > MLOD_MODE_BALL) return fail(MLOD_ERR_INVALID, "bad mode");
  if (mode == MLOD_MODE_KNN && knn == 0) return fail(MLOD_ERR_INVALID, "k must be >= 1");
  t->mode = mode;
  t->knn = knn;
  t->ball_radius = ball_radius;
  return MLOD_OK;
}

int mlod_queries_upload(mlod_tree* t, const float* xyz, uint32_t m, void* stream) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  if (m && !xyz) return fail(MLOD_ERR_INVALID, "queries are NULL");
  return queries_set(t, xyz, m, false, (hipStream_t)stream);
}

int mlod_queries_set_device(mlod_tree* t, const float* d_xyz, uint32_t m, void* stream) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  if (m && !d_xyz) return fail(MLOD_ERR_INVALID, "queries are NULL");
  return queries_set(t, d_xyz, m, true, (hipStream_t)stream);
}

int mlod_queries_erase(mlod_tree* t) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  (void)hipDeviceSynchronize();
  queries_free(t);
  return MLOD_OK;
}

}  // extern "C"

namespace mlod {
// n_iter steps of the first n uploaded queries in the tree's neighbourhood mode.  kNN / ball
// APSS take no step bound: their fit is called with q aliased to the output
// (CPUUtils.cpp:735), which makes the clamp a no-op.
int project_uploaded(mlod_tree* t, uint32_t n, uint32_t n_iter, float step_max, hipStream_t s) {
  if (!n_iter) return MLOD_OK;
  return (t->mode == MLOD_MODE_KNN || t->mode == MLOD_MODE_BALL) ? launch_neighbors(t, n, n_iter, s)
                                                                 : launch_project(t, n, n_iter, step_max, s);
}
}  // namespace mlod

extern "C" {

int mlod_project(mlod_tree* t, uint32_t n, uint32_t n_iter, float step_max, float* out_pts, float* out_nmls, void* stream) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  if (n == 0) return MLOD_OK;  // Fastapss::project returns on an empty batch (Fastapss.cu:441-445)
  if (!t->kernel_set) return fail(MLOD_ERR_STATE, "no kernel set (mlod_set_kernel)");
  if (n > t->qs.m) return fail(MLOD_ERR_INVALID, "n exceeds the number of uploaded queries");
  if (step_max < 0.f) step_max = t->radius;  // Fastapss.cu:448-449
  hipStream_t s = (hipStream_t)stream;
  if (int r = project_uploaded(t, n, n_iter, step_max, s)) return r;
  if (out_pts || out_nmls) return queries_get(t, n, out_pts, out_nmls, s);
  return MLOD_OK;
}

int mlod_project_points(mlod_tree* t, const float* xyz, uint32_t m, uint32_t n_iter, float step_max, float* out_pts,
                        float* out_nmls, void* stream) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  if (m == 0) return MLOD_OK;
  if (!xyz) return fail(MLOD_ERR_INVALID, "points are NULL");
  if (!t->kernel_set) return fail(MLOD_ERR_STATE, "no kernel set (mlod_set_kernel)");
  if (step_max < 0.f) step_max = t->radius;
  hipStream_t s = (hipStream_t)stream;
  // the uploaded queries are set aside and restored: these points use their own buffers
  const QueryState saved = t->qs;
  t->qs = QueryState{};
  int r = queries_set(t, xyz, m, false, s);
  if (!r) r = project_uploaded(t, m, n_iter, step_max, s);
  if (!r && (out_pts || out_nmls)) r = queries_get(t, m, out_pts, out_nmls, s);
  (void)hipStreamSynchronize(s);
  queries_free(t);
  t->qs = saved;
  return r;
}

int mlod_queries_fit_kinds(mlod_tree* t, uint32_t n, uint8_t* out, void* stream) {
  if (!t || !out) return fail(MLOD_ERR_INVALID, "NULL argument");
  return queries_fit_kinds(t, n, out, (hipStream_t)stream);
}

int mlod_queries_reorganize(mlod_tree* t, const uint32_t* invalid, uint32_t n, uint32_t end_valid, void* stream) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  if (n && !invalid) return fail(MLOD_ERR_INVALID, "invalid list is NULL");
  return queries_reorganize(t, invalid, n, end_valid, (hipStream_t)stream);
}

int mlod_queries_download(mlod_tree* t, uint32_t n, float* out_pts, float* out_nmls, void* stream) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  if (n > t->qs.m) return fail(MLOD_ERR_INVALID, "n exceeds the number of uploaded queries");
  if (n == 0) return MLOD_OK;
  return queries_get(t, n, out_pts, out_nmls, (hipStream_t)stream);
}

int mlod_tree_dump(const mlod_tree* t, int32_t* depth, float* bb6, float* center3, float* radius, int32_t* father,
                   int32_t* first_child, int32_t* next_brother, int32_t* num_children, uint32_t* leaf_count,
                   uint32_t* leaf_idx, float* areas, float* moments9) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  return dump_tree(t, depth, bb6, center3, radius, father, first_child, next_brother, num_children, leaf_count, leaf_idx,
                   areas, moments9);
}

int mlod_count_visits(mlod_tree* t, uint32_t n, uint64_t* out6, void* stream) {
  if (!t || !out6) return fail(MLOD_ERR_INVALID, "NULL argument");
  if (!t->kernel_set) return fail(MLOD_ERR_STATE, "no kernel set (mlod_set_kernel)");
  if (n > t->qs.m) return fail(MLOD_ERR_INVALID, "n exceeds the number of uploaded queries");
  hipStream_t s = (hipStream_t)stream;
  const uint32_t m = t->qs.m;
  // far, leaf points, tested, near, lane iterations, wavefront iterations (the longest lane's)
  const size_t cbytes = 6 * sizeof(unsigned long long);
  float4 *bq = nullptr, *bn = nullptr;
  unsigned long long* dc = nullptr;
  if (hipMalloc((void**)&bq, (size_t)m * sizeof(float4) + 16) != hipSuccess ||
      hipMalloc((void**)&bn, (size_t)m * sizeof(float4) + 16) != hipSuccess ||
      hipMalloc((void**)&dc, cbytes) != hipSuccess) {
    (void)hipFree(bq), (void)hipFree(bn), (void)hipFree(dc);
    return fail(MLOD_ERR_OOM, "counter pass buffers");
  }
  hipError_t e = hipMemcpyAsync(bq, t->qs.q, (size_t)m * sizeof(float4), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(bn, t->qs.qn, (size_t)m * sizeof(float4), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(dc, 0, cbytes, s);
  int r = e == hipSuccess ? launch_count(t, n, dc, s) : hip_fail(e, "counter pass setup", __FILE__, __LINE__);
  unsigned long long h[6] = {0, 0, 0, 0, 0, 0};
  if (r == MLOD_OK) {
    e = hipMemcpyAsync(h, dc, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(t->qs.q, bq, (size_t)m * sizeof(float4), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(t->qs.qn, bn, (size_t)m * sizeof(float4), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) r = hip_fail(e, "counter pass", __FILE__, __LINE__);
  }
  (void)hipFree(bq), (void)hipFree(bn), (void)hipFree(dc);
  const uint64_t laneSlots = 64ull * h[5];
  for (int k = 0; k < 5; ++k) out6[k] = h[k];
  out6[5] = laneSlots;
  return r;
}

}  // extern "C"
