// tfs_flow_api.hip -- the C-ABI of include/tfs_flow.h.
//
// Owns device copies of one save's sources (the reference's x_segs_device,
// f_segs_device, ... of flow_field_evaluator.cu:265-320, without the per-GPU
// split of the tracer points) and, for host-memory callers, staging buffers
// for the tracer points and velocities that grow to the largest request.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/tfs_flow.h"
#include "tfs_api_internal.h"

using namespace tfs;

struct tfs_flow {
    tfs_flow_desc desc{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int block = 256;
    double* src = nullptr;     // x_segs | f_segs | x_blobs | f_blobs | x_bodies | net
    double *x_segs = nullptr, *f_segs = nullptr, *x_blobs = nullptr, *f_blobs = nullptr, *x_bodies = nullptr,
           *net = nullptr;
    double* pts = nullptr;     // staging for host callers: points | velocities
    int64_t pts_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    tfs::F32Stage f32;         // FP32 entry points (tfs_flow_*_f32)
};

namespace {
int check(hipError_t e, const char* where) { return e == hipSuccess ? TFS_OK : hip_fail(e, where); }
}  // namespace

extern "C" {

int tfs_flow_create(const tfs_flow_desc* desc, tfs_flow** out) {
    g_last_error.clear();
    if (!desc || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_create: null argument");
    *out = nullptr;
    const tfs_flow_desc& D = *desc;
    if (D.num_bodies < 1 || D.segs_per_body < 0 || D.blobs_per_body < 0 ||
        (int64_t)D.num_bodies * (D.segs_per_body + D.blobs_per_body) > INT32_MAX)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_create: bad sizes");
    if (D.mode != TFS_FLOW_STOKESLET_ROTLET && D.mode != TFS_FLOW_RPY)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_create: unknown mode");
    if (D.precision != TFS_PRECISION_FP64 && D.precision != TFS_PRECISION_FP32_ACC64)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_create: unknown precision");
    if (D.memory != TFS_MEMORY_HOST && D.memory != TFS_MEMORY_DEVICE)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_create: unknown memory space");
    if (!(D.mu > 0) || (D.mode == TFS_FLOW_RPY && (!(D.seg_radius > 0) || !(D.blob_radius > 0))))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_create: radii and mu must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(TFS_ERR_NO_DEVICE, "tfs_flow_create: no HIP device visible");
    tfs_flow* h = new tfs_flow();
    h->desc = D;
    hipError_t e = hipSuccess;
    if (D.stream) {
        h->stream = (hipStream_t)D.stream;
    } else {
        e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
        h->own_stream = e == hipSuccess;
    }
    const int64_t ns = (int64_t)D.num_bodies * D.segs_per_body, nb = (int64_t)D.num_bodies * D.blobs_per_body;
    const int64_t lens[6] = {3 * ns, 6 * ns, 3 * nb, 3 * nb, 3LL * D.num_bodies, 3LL * D.num_bodies};
    int64_t total = 0;
    for (int64_t l : lens) total += ((std::max<int64_t>(l, 1) + 31) / 32) * 32;
    if (e == hipSuccess) e = hipMalloc(&h->src, total * 8);
    if (e == hipSuccess) e = hipMemsetAsync(h->src, 0, total * 8, h->stream);
    if (e == hipSuccess) e = hipEventCreate(&h->ev0);
    if (e == hipSuccess) e = hipEventCreate(&h->ev1);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        tfs_flow_destroy(h);
        return hip_fail(e, "tfs_flow_create");
    }
    double* p = h->src;
    double** dst[6] = {&h->x_segs, &h->f_segs, &h->x_blobs, &h->f_blobs, &h->x_bodies, &h->net};
    for (int i = 0; i < 6; ++i) {
        *dst[i] = p;
        p += ((std::max<int64_t>(lens[i], 1) + 31) / 32) * 32;
    }
    *out = h;
    return TFS_OK;
}

int tfs_flow_destroy(tfs_flow* h) {
    if (!h) return TFS_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->src) hipFree(h->src);
    if (h->pts) hipFree(h->pts);
    free_stage(h->f32);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    if (h->own_stream) hipStreamDestroy(h->stream);
    delete h;
    return TFS_OK;
}

int tfs_flow_set_sources(tfs_flow* h, const double* x_segs, const double* f_segs, const double* x_blobs,
                         const double* f_blobs, const double* x_bodies) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_set_sources: null handle");
    const tfs_flow_desc& D = h->desc;
    const int64_t ns = (int64_t)D.num_bodies * D.segs_per_body, nb = (int64_t)D.num_bodies * D.blobs_per_body;
    if ((ns && (!x_segs || !f_segs)) || (nb && (!x_blobs || !f_blobs)) || (D.remove_net_force && !x_bodies))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_set_sources: missing array");
    const hipMemcpyKind k = D.memory == TFS_MEMORY_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    hipError_t e = hipSuccess;
    if (ns) e = hipMemcpyAsync(h->x_segs, x_segs, 24 * ns, k, h->stream);
    if (e == hipSuccess && ns) e = hipMemcpyAsync(h->f_segs, f_segs, 48 * ns, k, h->stream);
    if (e == hipSuccess && nb) e = hipMemcpyAsync(h->x_blobs, x_blobs, 24 * nb, k, h->stream);
    if (e == hipSuccess && nb) e = hipMemcpyAsync(h->f_blobs, f_blobs, 24 * nb, k, h->stream);
    if (e == hipSuccess && x_bodies) e = hipMemcpyAsync(h->x_bodies, x_bodies, 24LL * D.num_bodies, k, h->stream);
    if (e == hipSuccess && D.memory == TFS_MEMORY_HOST) e = hipStreamSynchronize(h->stream);
    return check(e, "tfs_flow_set_sources");
}

}  // extern "C"

namespace {

// Staging for host callers' points and velocities (grown to the largest request).
int ensure_points(tfs_flow* h, int64_t num_points) {
    if (num_points <= h->pts_cap) return TFS_OK;
    if (h->pts) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipFree(h->pts);
    }
    h->pts = nullptr;
    h->pts_cap = 0;
    if (hipMalloc(&h->pts, 48 * num_points) != hipSuccess)
        return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_flow_evaluate: staging");
    h->pts_cap = num_points;
    return TFS_OK;
}

int check_points(tfs_flow* h, int64_t num_points, bool have_arrays, const char* where) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, std::string(where) + ": null handle");
    if (num_points < 0 || (num_points > 0 && !have_arrays))
        return fail(TFS_ERR_INVALID_ARGUMENT, std::string(where) + ": bad points");
    if (num_points > ((int64_t)1 << 31) / 4)
        return fail(TFS_ERR_INVALID_ARGUMENT, std::string(where) + ": too many points for one call");
    return TFS_OK;
}

// The velocities() kernels on device points xp, velocities vp.
hipError_t run_flow(tfs_flow* h, int64_t num_points, const double* xp, double* vp) {
    const tfs_flow_desc& D = h->desc;
    FlowArgs a;
    a.xt = {xp, 3, 1};
    a.vt = {vp, 3, 1};
    a.nt = num_points;
    a.xs = {h->x_segs, 3, 1};
    a.fs = {h->f_segs, 6, 1};
    a.xb = {h->x_blobs, 3, 1};
    a.fb = {h->f_blobs, 3, 1};
    a.ns_per_body = D.segs_per_body;
    a.nb_per_body = D.blobs_per_body;
    a.num_bodies = D.num_bodies;
    a.ns_total = D.num_bodies * D.segs_per_body;
    a.nb_total = D.num_bodies * D.blobs_per_body;
    a.x_bodies = h->x_bodies;
    a.remove_net_force = D.mode == TFS_FLOW_STOKESLET_ROTLET && D.remove_net_force ? 1 : 0;
    a.mode = D.mode;
    a.a_seg = D.seg_radius;
    a.a_blob = D.blob_radius;
    a.mu = D.mu;
    hipError_t e = hipSuccess;
    if (a.remove_net_force) e = launch_flow_net(a, h->net, D.num_bodies, h->stream);
    if (e == hipSuccess) e = hipEventRecord(h->ev0, h->stream);
    if (e == hipSuccess) e = launch_flow_eval(D.precision, a, h->net, h->block, h->stream);
    if (e == hipSuccess) e = hipEventRecord(h->ev1, h->stream);
    return e;
}

}  // namespace

extern "C" {

int tfs_flow_evaluate(tfs_flow* h, int64_t num_points, const double* x_points, double* v_points) {
    int rc = check_points(h, num_points, x_points && v_points, "tfs_flow_evaluate");
    if (rc || num_points == 0) return rc;
    const double* xp = x_points;
    double* vp = v_points;
    hipError_t e = hipSuccess;
    const bool host = h->desc.memory == TFS_MEMORY_HOST;
    if (host) {
        rc = ensure_points(h, num_points);
        if (rc) return rc;
        e = hipMemcpyAsync(h->pts, x_points, 24 * num_points, hipMemcpyHostToDevice, h->stream);
        xp = h->pts;
        vp = h->pts + 3 * num_points;
    }
    if (e == hipSuccess) e = run_flow(h, num_points, xp, vp);
    if (e == hipSuccess && host) {
        e = hipMemcpyAsync(v_points, vp, 24 * num_points, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    return check(e, "tfs_flow_evaluate");
}

int tfs_flow_set_sources_f32(tfs_flow* h, const float* x_segs, const float* f_segs, const float* x_blobs,
                             const float* f_blobs, const float* x_bodies) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_set_sources_f32: null handle");
    const tfs_flow_desc& D = h->desc;
    const int64_t ns = (int64_t)D.num_bodies * D.segs_per_body, nb = (int64_t)D.num_bodies * D.blobs_per_body;
    if ((ns && (!x_segs || !f_segs)) || (nb && (!x_blobs || !f_blobs)) || (D.remove_net_force && !x_bodies))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_set_sources_f32: missing array");
    const bool host = D.memory == TFS_MEMORY_HOST;
    const char* w = "tfs_flow_set_sources_f32";
    int rc = TFS_OK;
    if (ns) rc = widen_in(h->f32, x_segs, h->x_segs, 3 * ns, host, h->stream, w);
    if (!rc && ns) rc = widen_in(h->f32, f_segs, h->f_segs, 6 * ns, host, h->stream, w);
    if (!rc && nb) rc = widen_in(h->f32, x_blobs, h->x_blobs, 3 * nb, host, h->stream, w);
    if (!rc && nb) rc = widen_in(h->f32, f_blobs, h->f_blobs, 3 * nb, host, h->stream, w);
    if (!rc && x_bodies) rc = widen_in(h->f32, x_bodies, h->x_bodies, 3LL * D.num_bodies, host, h->stream, w);
    if (!rc && host) rc = check(hipStreamSynchronize(h->stream), w);
    return rc;
}

int tfs_flow_evaluate_f32(tfs_flow* h, int64_t num_points, const float* x_points, float* v_points) {
    int rc = check_points(h, num_points, x_points && v_points, "tfs_flow_evaluate_f32");
    if (rc || num_points == 0) return rc;
    const bool host = h->desc.memory == TFS_MEMORY_HOST;
    rc = ensure_points(h, num_points);
    if (!rc) rc = widen_in(h->f32, x_points, h->pts, 3 * num_points, host, h->stream, "tfs_flow_evaluate_f32");
    if (rc) return rc;
    hipError_t e = run_flow(h, num_points, h->pts, h->pts + 3 * num_points);
    if (e != hipSuccess) return check(e, "tfs_flow_evaluate_f32");
    rc = narrow_out(h->f32, h->pts + 3 * num_points, v_points, 3 * num_points, host, h->stream,
                    "tfs_flow_evaluate_f32");
    if (!rc && host) rc = check(hipStreamSynchronize(h->stream), "tfs_flow_evaluate_f32");
    return rc;
}

int tfs_flow_sync(tfs_flow* h) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_sync: null handle");
    return check(hipStreamSynchronize(h->stream), "tfs_flow_sync");
}

int tfs_flow_set_block(tfs_flow* h, int32_t block) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_set_block: null handle");
    if (block != 64 && block != 128 && block != 256)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_set_block: block must be 64, 128 or 256");
    h->block = block;
    return TFS_OK;
}

int tfs_flow_last_kernel_ms(tfs_flow* h, double* ms) {
    if (!h || !ms) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_flow_last_kernel_ms: null argument");
    hipError_t e = hipEventSynchronize(h->ev1);
    float f = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, h->ev0, h->ev1);
    *ms = f;
    return check(e, "tfs_flow_last_kernel_ms");
}

}  // extern "C"
