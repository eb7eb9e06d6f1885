/*
 * ndt_api.cpp - host side of the C ABI (include/ndt_hip.h): contexts, device
 * memory, grid-build and alignment orchestration, and the compat-mode host
 * controller.  Compiled by g++ with -ffp-contract=off: the compat controller
 * restates the reference's host arithmetic (src/CPU/ndt_mapping/kernel.cpp) with
 * the same libm calls, so its decisions match the reference given the same sums.
 */
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "ndt_hip.h"
#include "ndt_internal.h"
#include "ndt_host.h"

using namespace ndt;

/* ------------------------------------------------------------------ errors */

static thread_local std::string g_last_error;

int ndt::fail(int code, const std::string& msg) {
    g_last_error = msg;
    /* the launch wrappers report hipGetLastError(): clear this failure's error so
     * that it is not reported again by the thread's next launch */
    (void)hipGetLastError();
    return code;
}

int ndt::check_poses(const double* poses, int64_t n, const char* what) {
    for (int64_t i = 0; i < 16 * n; i++)
        if (!std::isfinite(poses[i]))
            return fail(NDT_ERR_INVALID, std::string(what) + " " + std::to_string(i / 16) + " has a non-finite entry");
    return 0;
}

std::atomic<int64_t>& ndt::device_bytes_held() {
    static std::atomic<int64_t> held{0};
    return held;
}

int64_t ndt::device_memory_limit() {
    const char* v = std::getenv("NDT_DEVICE_MEMORY_LIMIT");
    if (!v || !*v) return -1;
    char* end = nullptr;
    const long long x = std::strtoll(v, &end, 10);
    return (end && *end == 0 && x >= 0) ? (int64_t)x : -1;
}

extern "C" const char* ndt_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* ndt_version(void) { return "ndt-mi355x 0.1 (gfx950)"; }

/* ------------------------------------------------------------------ params */

extern "C" void ndt_params_default(ndt_params* p, int32_t mode) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->mode = mode;
    p->resolution = 1.0f;
    p->step_size = 0.1;
    p->outlier_ratio = 0.55;
    if (mode == NDT_MODE_COMPAT) {
        /* src/CPU/ndt_mapping/kernel.cpp:37-41,56-57 */
        p->trans_eps = 0.1;
        p->max_iterations = 0;
    } else {
        p->trans_eps = 0.01;
        p->max_iterations = 30;
    }
    p->min_points = 6;
    p->min_eig_ratio = 0.01;
    /* float64 per-point arithmetic by default: NDT trajectories on sparse scenes are
     * sensitive enough that float32 per-pair rounding moves some results by more
     * than a centimetre; float32 modes remain selectable */
    p->point_precision = NDT_POINT_FP64;
    p->sum_precision = NDT_SUM_FP64;
    p->fine_pass = DEFAULT_FINE_PASS;
}

/* first pass of a run in fine chunks */
static int32_t fine_pass_of(const ndt_params* p) { return p->fine_pass > 0 ? p->fine_pass : DEFAULT_FINE_PASS; }

static int check_params(const ndt_params* p) {
    if (!p) return fail(NDT_ERR_INVALID, "params is null");
    if (p->mode != NDT_MODE_COMPAT && p->mode != NDT_MODE_NDT) return fail(NDT_ERR_INVALID, "unknown mode");
    if (!(p->resolution > 0.0f) || !std::isfinite(p->resolution)) return fail(NDT_ERR_INVALID, "resolution must be > 0");
    if (!(p->step_size > 0) || !(p->trans_eps > 0)) return fail(NDT_ERR_INVALID, "step_size and trans_eps must be > 0");
    if (p->max_iterations < 0) return fail(NDT_ERR_INVALID, "max_iterations must be >= 0");
    if (!(p->outlier_ratio > 0) || !(p->outlier_ratio < 1)) return fail(NDT_ERR_INVALID, "outlier_ratio must be in (0, 1)");
    if (p->sum_precision < 0 || p->sum_precision > 2) return fail(NDT_ERR_INVALID, "bad sum_precision");
    if (p->point_precision < 0 || p->point_precision > 1) return fail(NDT_ERR_INVALID, "bad point_precision");
    if (p->block_size != 0 && p->block_size != 128 && p->block_size != 256 && p->block_size != 512)
        return fail(NDT_ERR_INVALID, "block_size must be 0, 128, 256 or 512");
    if (p->points_per_wg < 0 || p->points_per_wg > (1 << 24))
        return fail(NDT_ERR_INVALID, "points_per_wg must be in 0 .. 2^24");
    if (p->scheduler < NDT_SCHED_DEFAULT || p->scheduler > NDT_SCHED_QUEUE) return fail(NDT_ERR_INVALID, "bad scheduler");
    if (p->hessian_at_accept != 0 && p->hessian_at_accept != 1)
        return fail(NDT_ERR_INVALID, "hessian_at_accept must be 0 or 1");
    if (p->fine_pass < 0 || p->fine_pass > (1 << 20)) return fail(NDT_ERR_INVALID, "fine_pass must be in 0 .. 2^20");
    return 0;
}

/* ------------------------------------------------------------------ context */

extern "C" int ndt_ctx_create(int32_t device, ndt_ctx** out) {
    if (!out) return fail(NDT_ERR_INVALID, "out is null");
    *out = nullptr;
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(NDT_ERR_INVALID, "no such device");
    HIP_TRY(hipSetDevice(device));
    ndt_ctx* c = new (std::nothrow) ndt_ctx();
    if (!c) return fail(NDT_ERR_NOMEM, "host allocation failed");
    c->device = device;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete c;
        return fail(NDT_ERR_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    }
    c->cus = prop.multiProcessorCount;
    std::snprintf(c->name, sizeof(c->name), "%.120s (%.120s)", prop.name, prop.gcnArchName);
    if (!prop.cooperativeLaunch) {
        delete c;
        return fail(NDT_ERR_UNSUPPORTED, "device lacks cooperative launch");
    }
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess ||
        (e = hipEventCreate(&c->ev2)) != hipSuccess ||
        (e = hipHostMalloc(&c->hsmall, HSMALL_BYTES, hipHostMallocDefault)) != hipSuccess) {
        delete c;
        return fail(NDT_ERR_HIP, std::string("stream/event creation: ") + hipGetErrorString(e));
    }
    /* once per device, with the device current: the address of the workspace array
     * of ndt_align_scans_noalloc (the other entry points do not need it) */
    c->scratch_err = noalloc_scratch(device, &c->scratch);
    *out = c;
    return 0;
}

void ndt::ctx_retain(ndt_ctx* c) { c->refs.fetch_add(1); }

void ndt::ctx_release(ndt_ctx* c) {
    if (c->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->pin) (void)hipHostFree(c->pin);
    if (c->hsmall) (void)hipHostFree(c->hsmall);
    delete c;
    (void)hipGetLastError();
}

extern "C" int ndt_ctx_destroy(ndt_ctx* c) {
    if (!c) return 0;
    if (c->compat_grid) {
        /* the cached grid holds a reference to its context */
        ndt_grid* g = c->compat_grid;
        c->compat_grid = nullptr;
        ndt_grid_destroy(g);
    }
    ctx_release(c);
    return 0;
}

extern "C" int ndt_ctx_device_info(ndt_ctx* c, int32_t* cus, char* name, int32_t name_len) {
    if (!c) return fail(NDT_ERR_INVALID, "ctx is null");
    if (cus) *cus = c->cus;
    if (name && name_len > 0) std::snprintf(name, (size_t)name_len, "%s", c->name);
    return 0;
}

extern "C" int ndt_ctx_last_profile(ndt_ctx* c, double phases_us[4]) {
    if (!c || !phases_us) return fail(NDT_ERR_INVALID, "null argument");
    for (int k = 0; k < 4; k++) phases_us[k] = c->last_profile[k];
    return 0;
}

extern "C" int ndt_ctx_last_step_time(ndt_ctx* c, double* device_ms) {
    if (!c || !device_ms) return fail(NDT_ERR_INVALID, "null argument");
    *device_ms = c->last_step_ms;
    return 0;
}

extern "C" int ndt_ctx_last_timing(ndt_ctx* c, double* ms, int32_t* launches, int64_t* pairs) {
    if (!c) return fail(NDT_ERR_INVALID, "ctx is null");
    if (ms) *ms = c->last_ms;
    if (launches) *launches = c->last_launches;
    if (pairs) *pairs = c->last_pairs;
    return 0;
}

static int ceil_log2(int64_t v) {
    int b = 0;
    while ((1ll << b) < v) b++;
    return b;
}

static inline float ord2f(uint32_t u) {
    uint32_t bits = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

/* runs of equal keys in a sorted array: flags -> ranks -> starts; returns count */
static int sorted_runs(ndt_ctx* c, const uint64_t* keys, int64_t n, int shift, uint64_t limit, DevBuf& flags,
                       DevBuf& ranks, DevBuf& starts, DevBuf& scratch, int64_t* nruns) {
    BUF_TRY(flags, sizeof(uint32_t) * (n + 1));
    BUF_TRY(ranks, sizeof(uint32_t) * (n + 1));
    BUF_TRY(scratch, scan_scratch_bytes(n) + 512);
    LAUNCH_TRY(launch_run_heads(keys, n, shift, limit, flags.as<uint32_t>(), c->stream));
    uint32_t* total = (uint32_t*)((char*)scratch.p + ((scan_scratch_bytes(n) + 255) & ~(size_t)255));
    LAUNCH_TRY(exclusive_scan_u32(flags.as<uint32_t>(), ranks.as<uint32_t>(), n, total, scratch.p, c->stream));
    HIP_TRY(hipMemcpyAsync(c->hsmall, total, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const uint32_t h_total = *(const uint32_t*)c->hsmall;
    BUF_TRY(starts, sizeof(uint32_t) * (h_total + 1));
    LAUNCH_TRY(launch_run_starts(flags.as<uint32_t>(), ranks.as<uint32_t>(), n, starts.as<uint32_t>(), c->stream));
    *nruns = h_total;
    return 0;
}

/* count of sorted keys below limit (non-sentinel prefix) */
static int count_below(ndt_ctx* c, const uint64_t* keys, int64_t n, uint64_t limit, int64_t* out) {
    /* binary search on the host over a few device reads */
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        uint64_t k = 0;
        HIP_TRY(hipMemcpyAsync(&k, keys + mid, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (k < limit) lo = mid + 1;
        else hi = mid;
    }
    *out = lo;
    return 0;
}

/* ------------------------------------------------------------------ grid build */

/* The context's pinned staging buffer, grown (in powers of two, at least 4 MiB)
 * to hold `bytes`.  Before it is replaced, the copies in flight from it complete. */
static int pinned_reserve(ndt_ctx* c, size_t bytes) {
    if (c->pin && c->pin_bytes >= bytes) return 0;
    if (c->pin_pending) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->pin_pending = false;
    }
    if (c->pin) {
        (void)hipHostFree(c->pin);
        c->pin = nullptr;
        c->pin_bytes = 0;
    }
    size_t want = (size_t)4 << 20;
    while (want < bytes) want <<= 1;
    hipError_t e = hipHostMalloc(&c->pin, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        c->pin = nullptr;
        return fail(NDT_ERR_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    c->pin_bytes = want;
    return 0;
}

/* Host-to-device copies (one or two) through the pinned staging buffer: a pageable
 * hipMemcpyAsync is staged by the runtime through buffers shared by every stream,
 * which serialises contexts that upload concurrently.  The staging buffer is
 * refilled only after the previous copies from it have completed. */
static int upload_staged2(ndt_ctx* c, void* d1, const void* s1, size_t b1, void* d2, const void* s2, size_t b2) {
    if (c->pin_pending) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->pin_pending = false;
    }
    const size_t off2 = (b1 + 255) & ~(size_t)255;
    int e = pinned_reserve(c, off2 + b2);
    if (e) return e;
    std::memcpy(c->pin, s1, b1);
    HIP_TRY(hipMemcpyAsync(d1, c->pin, b1, hipMemcpyHostToDevice, c->stream));
    if (b2) {
        std::memcpy((char*)c->pin + off2, s2, b2);
        HIP_TRY(hipMemcpyAsync(d2, (char*)c->pin + off2, b2, hipMemcpyHostToDevice, c->stream));
    }
    c->pin_pending = true;
    return 0;
}

static int upload_staged(ndt_ctx* c, void* dst, const void* src, size_t bytes) {
    return upload_staged2(c, dst, src, bytes, nullptr, nullptr, 0);
}

/* m points (x, y, z, w floats) to dst through the pinned staging buffer in chunks
 * of 16,384 points: each chunk is copied and checked on the host, then its copy to
 * the device is queued, so the transfer of one chunk overlaps the host work on the
 * next.  Returns NDT_ERR_INVALID at the first chunk holding a point whose x, y or
 * z is not finite (chunks before it may already be on the device). */
static int upload_points_finite(ndt_ctx* c, void* dst, const float* src, int64_t m) {
    if (c->pin_pending) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->pin_pending = false;
    }
    int e = pinned_reserve(c, sizeof(float4) * (size_t)m);
    if (e) return e;
    constexpr int64_t CHUNK = 16384;
    for (int64_t p0 = 0; p0 < m; p0 += CHUNK) {
        const int64_t n = std::min(CHUNK, m - p0);
        const float* sp = src + 4 * p0;
        float* hp = (float*)c->pin + 4 * p0;
        bool finite = true;
        for (int64_t i = 0; i < 4 * n; i += 4) {
            const float x = sp[i], y = sp[i + 1], z = sp[i + 2];
            hp[i] = x;
            hp[i + 1] = y;
            hp[i + 2] = z;
            hp[i + 3] = sp[i + 3];
            /* v - v is 0 for a finite v and NaN for an infinite or NaN one */
            finite &= (x - x == 0.0f) & (y - y == 0.0f) & (z - z == 0.0f);
        }
        if (!finite) {
            c->pin_pending = p0 > 0;
            return fail(NDT_ERR_INVALID, "compat mode needs finite map points");
        }
        HIP_TRY(hipMemcpyAsync((char*)dst + sizeof(float4) * p0, hp, sizeof(float4) * n, hipMemcpyHostToDevice,
                               c->stream));
        c->pin_pending = true;
    }
    return 0;
}

static int grid_build_ndt(ndt_ctx* c, const float* map, int64_t m, ndt_grid* g) {
    const double res = (double)g->params.resolution;
    BUF_TRY(c->pts, sizeof(float4) * m);
    {
        const int eu = upload_staged(c, c->pts.p, map, sizeof(float4) * m);
        if (eu) return eu;
    }
    DevBuf& mm = c->bk_mm;
    BUF_TRY(mm, 64);
    uint32_t init[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    HIP_TRY(hipMemcpyAsync(mm.p, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    LAUNCH_TRY(launch_bbox(c->pts.as<float4>(), m, mm.as<uint32_t>(), c->stream));
    uint32_t h_mm[6];
    HIP_TRY(hipMemcpyAsync(h_mm, mm.p, sizeof(h_mm), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (h_mm[0] == 0xFFFFFFFFu) return fail(NDT_ERR_INVALID, "map has no finite points");
    int64_t lo_cell[3], hi_cell[3];
    for (int k = 0; k < 3; k++) {
        lo_cell[k] = (int64_t)std::floor((double)ord2f(h_mm[k]) / res);
        hi_cell[k] = (int64_t)std::floor((double)ord2f(h_mm[3 + k]) / res);
        g->origin_cell[k] = lo_cell[k] - 1;
        g->bits[k] = std::max(1, ceil_log2(hi_cell[k] - lo_cell[k] + 3));
        g->dims[k] = (int32_t)std::min<int64_t>(hi_cell[k] - lo_cell[k] + 3, 0x7fffffff);
    }
    const int B = g->bits[0] + g->bits[1] + g->bits[2];
    if (B + 5 > 62) return fail(NDT_ERR_UNSUPPORTED, "map extent too large for the cell key");
    DevBuf &keys = c->bk_keys, &vals = c->bk_vals, &keys_t = c->bk_keys_t, &vals_t = c->bk_vals_t;
    DevBuf &scratch = c->bk_scratch, &flags = c->bk_flags, &ranks = c->bk_ranks, &starts = c->bk_starts;
    DevBuf& sc2 = c->bk_sc2;
    BUF_TRY(keys, sizeof(uint64_t) * m);
    BUF_TRY(vals, sizeof(uint32_t) * m);
    BUF_TRY(keys_t, sizeof(uint64_t) * m);
    BUF_TRY(vals_t, sizeof(uint32_t) * m);
    BUF_TRY(scratch, radix_sort_scratch_bytes(m));
    LAUNCH_TRY(launch_keys_ndt(c->pts.as<float4>(), m, res, g->origin_cell, g->bits, keys.as<uint64_t>(),
                               vals.as<uint32_t>(), c->stream));
    LAUNCH_TRY(radix_sort_pairs(keys.as<uint64_t>(), vals.as<uint32_t>(), keys_t.as<uint64_t>(), vals_t.as<uint32_t>(), m,
                                B + 1, scratch.p, c->stream));
    const uint64_t limit = 1ull << B;
    int64_t nvox = 0, nvalid_pts = 0;
    int e = sorted_runs(c, keys.as<uint64_t>(), m, 0, limit, flags, ranks, starts, sc2, &nvox);
    if (e) return e;
    e = count_below(c, keys.as<uint64_t>(), m, limit, &nvalid_pts);
    if (e) return e;
    if (nvox == 0) return fail(NDT_ERR_INVALID, "map has no finite points");
    BUF_TRY(g->all_mean, sizeof(double) * 3 * nvox);
    BUF_TRY(g->all_icov, sizeof(double) * 9 * nvox);
    BUF_TRY(g->all_cov, sizeof(double) * 9 * nvox);
    BUF_TRY(g->all_count, sizeof(int32_t) * nvox);
    BUF_TRY(g->all_valid, sizeof(uint32_t) * (nvox + 1));
    BUF_TRY(g->all_key, sizeof(uint64_t) * nvox);
    LAUNCH_TRY(launch_voxel_stats_ndt(c->pts.as<float4>(), vals.as<uint32_t>(), keys.as<uint64_t>(), starts.as<uint32_t>(),
                                      nvox, nvalid_pts, g->params.min_points, g->params.min_eig_ratio, g->bits,
                                      g->all_mean.as<double>(), g->all_icov.as<double>(), g->all_cov.as<double>(),
                                      g->all_count.as<int32_t>(), g->all_valid.as<uint32_t>(), g->all_key.as<uint64_t>(),
                                      c->stream));
    /* map points sorted by voxel and a hash of the occupied voxels, for nearest-point
     * queries (fitness score) */
    BUF_TRY(g->fit_pts, sizeof(float4) * (size_t)std::max<int64_t>(nvalid_pts, 1));
    LAUNCH_TRY(launch_gather_points(c->pts.as<float4>(), vals.as<uint32_t>(), nvalid_pts, g->fit_pts.as<float4>(),
                                    c->stream));
    const int flog2 = std::max(4, ceil_log2(2 * nvox));
    BUF_TRY(g->fit_slots, sizeof(HashSlot) << flog2);
    LAUNCH_TRY(launch_hash_clear(g->fit_slots.as<HashSlot>(), 1ll << flog2, c->stream));
    LAUNCH_TRY(launch_hash_insert(keys.as<uint64_t>(), starts.as<uint32_t>(), nvox, nvalid_pts, 0,
                                  g->fit_slots.as<HashSlot>(), flog2, c->stream));
    g->fdev.slots = g->fit_slots.as<HashSlot>();
    g->fdev.pts = g->fit_pts.as<float4>();
    g->fdev.log2cap = flog2;
    for (int k = 0; k < 3; k++) {
        g->fdev.bits[k] = g->bits[k];
        g->fdev.origin_cell[k] = g->origin_cell[k];
    }
    g->fdev.resolution = res;
    /* compact the valid voxels */
    DevBuf vrank, sc3;
    BUF_TRY(vrank, sizeof(uint32_t) * (nvox + 1));
    BUF_TRY(sc3, scan_scratch_bytes(nvox) + 512);
    uint32_t* vtotal = (uint32_t*)((char*)sc3.p + ((scan_scratch_bytes(nvox) + 255) & ~(size_t)255));
    LAUNCH_TRY(exclusive_scan_u32(g->all_valid.as<uint32_t>(), vrank.as<uint32_t>(), nvox, vtotal, sc3.p, c->stream));
    uint32_t h_nvalid = 0;
    HIP_TRY(hipMemcpyAsync(&h_nvalid, vtotal, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int64_t nv = h_nvalid;
    g->info.map_points = m;
    g->info.voxels = nvox;
    g->info.valid_voxels = nv;
    BUF_TRY(g->mean_c, sizeof(double) * 3 * (nv + 1));
    BUF_TRY(g->icov_c, sizeof(double) * 6 * (nv + 1));
    BUF_TRY(g->icov_f, sizeof(float4) * 2 * (nv + 1));
    DevBuf vkey_c;
    BUF_TRY(vkey_c, sizeof(uint64_t) * (nv + 1));
    LAUNCH_TRY(launch_compact_voxels(g->all_valid.as<uint32_t>(), vrank.as<uint32_t>(), nvox, g->all_mean.as<double>(),
                                     g->all_icov.as<double>(), g->all_key.as<uint64_t>(), g->mean_c.as<double>(),
                                     g->icov_c.as<double>(), g->icov_f.as<float4>(), vkey_c.as<uint64_t>(), c->stream));
    /* dilation: candidate lists of every query cell */
    const int64_t ne = 27 * nv;
    int64_t nq = 0;
    if (ne > 0) {
        BUF_TRY(keys, sizeof(uint64_t) * ne);
        BUF_TRY(vals, sizeof(uint32_t) * ne);
        BUF_TRY(keys_t, sizeof(uint64_t) * ne);
        BUF_TRY(vals_t, sizeof(uint32_t) * ne);
        BUF_TRY(scratch, radix_sort_scratch_bytes(ne));
        LAUNCH_TRY(launch_dilate_emit(vkey_c.as<uint64_t>(), nv, g->bits, keys.as<uint64_t>(), vals.as<uint32_t>(), c->stream));
        LAUNCH_TRY(radix_sort_pairs(keys.as<uint64_t>(), vals.as<uint32_t>(), keys_t.as<uint64_t>(), vals_t.as<uint32_t>(),
                                    ne, B + 5, scratch.p, c->stream));
        e = sorted_runs(c, keys.as<uint64_t>(), ne, 5, ~0ull, flags, ranks, starts, sc2, &nq);
        if (e) return e;
    }
    int log2cap = std::max(4, ceil_log2(2 * std::max<int64_t>(nq, 1)));
    int64_t cap = 1ll << log2cap;
    BUF_TRY(g->slots, sizeof(HashSlot) * cap);
    LAUNCH_TRY(launch_hash_clear(g->slots.as<HashSlot>(), cap, c->stream));
    BUF_TRY(g->cand_head, sizeof(float4) * (ne + 1));
    BUF_TRY(g->cand_s32, 2 * sizeof(float4) * (ne + 1));
    BUF_TRY(g->vs64, VS64_STRIDE * sizeof(double2) * (nv + 1));
    LAUNCH_TRY(launch_voxel_s64(g->mean_c.as<double>(), g->icov_c.as<double>(), nv, g->vs64.as<double2>(), c->stream));
    /* dense query-cell index when the map's cell box is at most 2^28 cells (2 GiB);
     * the environment variable NDT_FORCE_HASHED_INDEX=1 makes any map use the hashed
     * index, so that tests run that path on maps of every size */
    const double box = (double)g->dims[0] * (double)g->dims[1] * (double)g->dims[2];
    const char* force_hash = std::getenv("NDT_FORCE_HASHED_INDEX");
    const bool dense = box <= (double)(1 << 28) && !(force_hash && std::strcmp(force_hash, "1") == 0);
    if (dense) {
        const size_t cells = (size_t)g->dims[0] * (size_t)g->dims[1] * (size_t)g->dims[2];
        BUF_TRY(g->dense, sizeof(uint2) * cells);
        HIP_TRY(hipMemsetAsync(g->dense.p, 0, sizeof(uint2) * cells, c->stream));
        if (nq > 0)
            LAUNCH_TRY(launch_dense_fill(keys.as<uint64_t>(), starts.as<uint32_t>(), nq, ne, g->bits, g->dims,
                                         g->dense.as<uint2>(), c->stream));
    }
    if (nq > 0) {
        LAUNCH_TRY(launch_hash_insert(keys.as<uint64_t>(), starts.as<uint32_t>(), nq, ne, 5, g->slots.as<HashSlot>(),
                                      log2cap, c->stream));
        LAUNCH_TRY(launch_cand_fill(keys.as<uint64_t>(), vals.as<uint32_t>(), ne, g->mean_c.as<double>(),
                                    g->icov_c.as<double>(), g->bits, g->origin_cell, (double)g->params.resolution,
                                    g->cand_head.as<float4>(), g->cand_s32.as<float4>(), c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    g->info.query_cells = nq;
    g->info.candidates = ne;
    g->info.hash_capacity = cap;
    g->info.dense_cells = dense ? (int64_t)box : 0;
    NdtGridDev& d = g->dev;
    d.dense = dense ? g->dense.as<uint2>() : nullptr;
    for (int k = 0; k < 3; k++) d.dims[k] = g->dims[k];
    d.slots = g->slots.as<HashSlot>();
    d.head = g->cand_head.as<float4>();
    d.s32 = g->cand_s32.as<float4>();
    d.vs64 = g->vs64.as<double2>();
    d.log2cap = log2cap;
    for (int k = 0; k < 3; k++) {
        d.bits[k] = g->bits[k];
        d.origin_cell[k] = g->origin_cell[k];
        d.origin_cell_d[k] = (double)g->origin_cell[k];
    }
    d.resolution = g->params.resolution;
    d.inv_resolution = 1.0f / g->params.resolution;
    d.resolution_d = res;
    int ex = 0;
    d.res_pow2 = std::frexp(res, &ex) == 0.5 ? 1 : 0;
    d.inv_resolution_d = 1.0 / res;
    return 0;
}

static uint64_t next_grid_serial() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

static int grid_build_compat(ndt_ctx* c, const float* map, int64_t m, ndt_grid* g) {
    /* initCompute, src/CPU/ndt_mapping/kernel.cpp:1302-1373 */
    const float res = g->params.resolution;
    BUF_TRY(c->pts, sizeof(float4) * m);
    int e = upload_points_finite(c, c->pts.p, map, m);
    if (e) return e;
    DevBuf& mm = c->bk_mm;
    BUF_TRY(mm, 64);
    uint32_t* hs = (uint32_t*)c->hsmall;
    const uint32_t init[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    std::memcpy(hs, init, sizeof(init));
    HIP_TRY(hipMemcpyAsync(mm.p, hs, sizeof(init), hipMemcpyHostToDevice, c->stream));
    LAUNCH_TRY(launch_bbox(c->pts.as<float4>(), m, mm.as<uint32_t>(), c->stream));
    HIP_TRY(hipMemcpyAsync(hs + 16, mm.p, 6 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    uint32_t h_mm[6];
    std::memcpy(h_mm, hs + 16, sizeof(h_mm));
    CompatGridDev& d = g->cdev;
    int64_t ncells = 1;
    for (int k = 0; k < 3; k++) {
        float mn = ord2f(h_mm[k]), mx = ord2f(h_mm[3 + k]);
        d.minv[k] = mn - 0.01f;
        d.maxv[k] = mx + 0.01f;
        float span = (d.maxv[k] - d.minv[k]) / res + 1;
        if (!(span < 2.0e9f)) return fail(NDT_ERR_UNSUPPORTED, "dense grid too large");
        d.dims[k] = (int)span;
        ncells *= d.dims[k];
    }
    if (ncells > 0x7fffffffll) return fail(NDT_ERR_UNSUPPORTED, "dense grid exceeds the reference's int cell index");
    d.resolution = res;
    g->info.dense_cells = ncells;
    for (int k = 0; k < 3; k++) {
        g->info.dims[k] = d.dims[k];
        g->info.min_corner[k] = d.minv[k];
        g->info.max_corner[k] = d.maxv[k];
    }
    DevBuf &keys = c->bk_keys, &vals = c->bk_vals, &keys_t = c->bk_keys_t, &vals_t = c->bk_vals_t;
    DevBuf &scratch = c->bk_scratch, &flags = c->bk_flags, &ranks = c->bk_ranks, &starts = c->bk_starts;
    DevBuf& sc2 = c->bk_sc2;
    BUF_TRY(keys, sizeof(uint64_t) * m);
    BUF_TRY(vals, sizeof(uint32_t) * m);
    BUF_TRY(keys_t, sizeof(uint64_t) * m);
    BUF_TRY(vals_t, sizeof(uint32_t) * m);
    BUF_TRY(scratch, radix_sort_scratch_bytes(m));
    LAUNCH_TRY(launch_keys_compat(c->pts.as<float4>(), m, d.minv, d.dims, res, keys.as<uint64_t>(), vals.as<uint32_t>(),
                                  c->stream));
    LAUNCH_TRY(radix_sort_pairs(keys.as<uint64_t>(), vals.as<uint32_t>(), keys_t.as<uint64_t>(), vals_t.as<uint32_t>(), m,
                                std::max(1, ceil_log2(ncells + 1)), scratch.p, c->stream));
    int64_t nvox = 0;
    e = sorted_runs(c, keys.as<uint64_t>(), m, 0, (uint64_t)ncells, flags, ranks, starts, sc2, &nvox);
    if (e) return e;
    BUF_TRY(g->all_mean, sizeof(double) * 3 * nvox);
    BUF_TRY(g->all_icov, sizeof(double) * 9 * nvox);
    BUF_TRY(g->all_count, sizeof(int32_t) * nvox);
    BUF_TRY(g->all_key, sizeof(uint64_t) * nvox);
    LAUNCH_TRY(launch_voxel_stats_compat(c->pts.as<float4>(), vals.as<uint32_t>(), keys.as<uint64_t>(),
                                         starts.as<uint32_t>(), nvox, m, (double)(size_t)ncells, g->all_mean.as<double>(),
                                         g->all_icov.as<double>(), g->all_count.as<int32_t>(), g->all_key.as<uint64_t>(),
                                         c->stream));
    int log2cap = std::max(4, ceil_log2(2 * std::max<int64_t>(nvox, 1)));
    int64_t cap = 1ll << log2cap;
    BUF_TRY(g->slots, sizeof(HashSlot) * cap);
    LAUNCH_TRY(launch_hash_clear(g->slots.as<HashSlot>(), cap, c->stream));
    LAUNCH_TRY(launch_hash_insert_ids(g->all_key.as<uint64_t>(), nvox, g->slots.as<HashSlot>(), log2cap, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    d.slots = g->slots.as<HashSlot>();
    d.mean = g->all_mean.as<double>();
    d.icov = g->all_icov.as<double>();
    d.log2cap = log2cap;
    g->info.map_points = m;
    g->info.voxels = nvox;
    g->info.valid_voxels = nvox;
    g->info.hash_capacity = cap;
    return 0;
}

extern "C" int ndt_grid_build(ndt_ctx* c, const float* map, int64_t m, const ndt_params* params, ndt_grid** out) {
    if (!c || !map || !out) return fail(NDT_ERR_INVALID, "null argument");
    *out = nullptr;
    if (m <= 0) return fail(NDT_ERR_INVALID, "empty map");
    if (m > 0x7fffffffll) return fail(NDT_ERR_UNSUPPORTED, "map larger than 2^31 points");
    int e = check_params(params);
    if (e) return e;
    HIP_TRY(hipSetDevice(c->device));
    ndt_grid* g = new (std::nothrow) ndt_grid();
    if (g) g->serial = next_grid_serial();
    if (!g) return fail(NDT_ERR_NOMEM, "host allocation failed");
    g->ctx = c;
    g->params = *params;
    std::memset(&g->info, 0, sizeof(g->info));
    std::memset(&g->dev, 0, sizeof(g->dev));
    std::memset(&g->cdev, 0, sizeof(g->cdev));
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    e = (params->mode == NDT_MODE_COMPAT) ? grid_build_compat(c, map, m, g) : grid_build_ndt(c, map, m, g);
    if (e) {
        delete g;
        return e;
    }
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    g->info.build_ms = ms;
    *out = g;
    return 0;
}

extern "C" int ndt_grid_info_get(const ndt_grid* g, ndt_grid_info* info) {
    if (!g || !info) return fail(NDT_ERR_INVALID, "null argument");
    *info = g->info;
    return 0;
}

extern "C" int ndt_grid_destroy(ndt_grid* g) {
    if (!g) return 0;
    (void)hipSetDevice(g->ctx->device);
    delete g;
    (void)hipGetLastError();
    return 0;
}

extern "C" int ndt_grid_voxels(const ndt_grid* g, int64_t cap, int64_t* keys, int32_t* n, double* mean, double* icov,
                               int32_t* valid) {
    if (!g) return fail(NDT_ERR_INVALID, "grid is null");
    ndt_ctx* c = g->ctx;
    HIP_TRY(hipSetDevice(c->device));
    const int64_t k = std::min<int64_t>(cap, g->info.voxels);
    if (k <= 0) return 0;
    std::vector<uint64_t> hk(k);
    HIP_TRY(hipMemcpy(hk.data(), g->all_key.p, sizeof(uint64_t) * k, hipMemcpyDeviceToHost));
    if (keys) {
        for (int64_t i = 0; i < k; i++) {
            if (g->params.mode == NDT_MODE_COMPAT) {
                keys[3 * i] = (int64_t)hk[i];
                keys[3 * i + 1] = keys[3 * i + 2] = 0;
            } else {
                uint64_t key = hk[i];
                keys[3 * i] = (int64_t)(key & ((1ull << g->bits[0]) - 1)) + g->origin_cell[0];
                keys[3 * i + 1] = (int64_t)((key >> g->bits[0]) & ((1ull << g->bits[1]) - 1)) + g->origin_cell[1];
                keys[3 * i + 2] = (int64_t)(key >> (g->bits[0] + g->bits[1])) + g->origin_cell[2];
            }
        }
    }
    if (n) HIP_TRY(hipMemcpy(n, g->all_count.p, sizeof(int32_t) * k, hipMemcpyDeviceToHost));
    if (mean) HIP_TRY(hipMemcpy(mean, g->all_mean.p, sizeof(double) * 3 * k, hipMemcpyDeviceToHost));
    if (icov) HIP_TRY(hipMemcpy(icov, g->all_icov.p, sizeof(double) * 9 * k, hipMemcpyDeviceToHost));
    if (valid) {
        if (g->params.mode == NDT_MODE_COMPAT) {
            for (int64_t i = 0; i < k; i++) valid[i] = 1;
        } else {
            std::vector<uint32_t> hv(k);
            HIP_TRY(hipMemcpy(hv.data(), g->all_valid.p, sizeof(uint32_t) * k, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < k; i++) valid[i] = (int32_t)hv[i];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ ndt alignment */

/* Bound of one device-side wait (grid barrier, work-queue claim), in ticks of the
 * 100 MHz s_memrealtime counter: 2 s, or the environment variable
 * NDT_WAIT_LIMIT_US (microseconds, for failure-path tests) when set.  An expired
 * wait aborts the launch, which drains and reports NDT_ERR_TIMEOUT. */
static uint64_t wait_limit_ticks() {
    const char* v = std::getenv("NDT_WAIT_LIMIT_US");
    if (v && *v) {
        char* end = nullptr;
        const unsigned long long us = std::strtoull(v, &end, 10);
        if (end && *end == 0 && us <= 3600000000ull) return (uint64_t)us * 100ull;
    }
    return 200000000ull;
}

static AlignParams make_align_params(const ndt_params* p, int32_t S) {
    AlignParams a;
    std::memset(&a, 0, sizeof(a));
    a.step_size = p->step_size;
    a.trans_eps = p->trans_eps;
    a.resolution_d = (double)p->resolution;
    gauss_consts((double)p->resolution, p->outlier_ratio, &a.d1, &a.d2);
    for (int k = 0; k < 3; k++) a.origin[k] = p->origin[k];
    a.max_iterations = p->max_iterations;
    a.point_fp64 = p->point_precision == NDT_POINT_FP64;
    a.sum_mode = a.point_fp64 ? NDT_SUM_FP64 : p->sum_precision;
    a.nscans = S;
    a.hess_at_accept = p->hessian_at_accept;
    a.wait_ticks = wait_limit_ticks();
    return a;
}

/* Fills results from the outputs of S scans (host copy); aborted marks scans that
 * never finished as timed out. */
static int results_from_out(const ScanOut* out, const int64_t* counts, int32_t S, int aborted, ndt_ctx* c,
                            ndt_result* results, int32_t* used_alt = nullptr) {
    int64_t pairs = 0;
    for (int32_t s = 0; s < S; s++) {
        ndt_result& r = results[s];
        std::memset(&r, 0, sizeof(r));
        for (int k = 0; k < 16; k++) {
            r.final_transformation_f64[k] = out[s].final_m[k];
            r.final_transformation[k] = (float)out[s].final_m[k];
        }
        r.score = out[s].score;
        r.trans_probability = counts[s] > 0 ? out[s].score / (double)counts[s] : 0.0;
        r.iterations = out[s].iterations;
        r.converged = out[s].converged;
        r.passes = out[s].passes;
        r.status = out[s].status;
        /* a scan that never finished has no final pose */
        if (aborted && out[s].final_m[15] == 0.0) r.status = NDT_ERR_TIMEOUT;
        if (used_alt) used_alt[s] = out[s].used_alt;
        pairs += (int64_t)out[s].pairs;
    }
    c->last_pairs = pairs;
    if (aborted) return fail(NDT_ERR_TIMEOUT, "a device-side wait exceeded its bound");
    return 0;
}

/* Fills results from the device outputs of S scans in the context's buffers. */
static int collect_results(ndt_ctx* c, const int64_t* counts, int32_t S, int aborted, ndt_result* results,
                           int32_t* used_alt = nullptr) {
    /* read back through the pinned staging buffer (stream order puts this copy
     * after the uploads from it) */
    const size_t bytes = sizeof(ScanOut) * (size_t)S;
    int e = pinned_reserve(c, bytes);
    if (e) return e;
    HIP_TRY(hipMemcpyAsync(c->pin, c->out.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->pin_pending = false;
    return results_from_out((const ScanOut*)c->pin, counts, S, aborted, c, results, used_alt);
}

/* Chunk descriptors of a batch: chunk size, count and first partial slot per scan. */
static void make_descs(const int64_t* offsets, const int64_t* counts, int32_t S, const ndt_params* p,
                       std::vector<ScanDesc>& desc, int64_t* slots_out, int64_t* max_chunks_out,
                       int64_t* first_pass_out) {
    desc.resize(S);
    int64_t slots = 0, max_chunks = 0, first = 0;
    for (int32_t s = 0; s < S; s++) {
        const int64_t chunk = p->points_per_wg > 0 ? p->points_per_wg : align_chunk_points(counts[s]);
        desc[s].offset = offsets[s];
        desc[s].n = (int32_t)counts[s];
        desc[s].chunk = (int32_t)chunk;
        desc[s].nchunks = (int32_t)((counts[s] + chunk - 1) / chunk);
        const int64_t fine = fine_chunk((int32_t)chunk);
        desc[s].fnchunks = (int32_t)((counts[s] + fine - 1) / fine);
        desc[s].chunk_base = slots;
        /* first-pass items are packed in scan order; slots cover the finer chunking */
        desc[s].qbase = first;
        desc[s].fine_pass = fine_pass_of(p);
        desc[s].pad = 0;
        slots += desc[s].fnchunks;
        first += desc[s].nchunks;
        max_chunks = std::max<int64_t>(max_chunks, desc[s].fnchunks);
    }
    *slots_out = slots;
    *max_chunks_out = max_chunks;
    *first_pass_out = first;
}

/* Work-queue items needed by a launch: every scan's chunks for each of its
 * possible passes. */
static uint64_t queue_capacity(int64_t total_chunks, int32_t max_iterations, bool fallback = false) {
    return (uint64_t)total_chunks *
           (uint64_t)(fallback ? max_passes_fallback(max_iterations) : max_passes(max_iterations));
}
constexpr uint64_t MAX_QUEUE_ITEMS = 1ull << 28;

/* Aligns scans already packed in `pts` (offsets/counts) with guesses on the host.
 * The chunk size (points per partial sum) depends only on the scan's own size
 * unless the caller sets points_per_wg, and partial sums are reduced in a fixed
 * order, so a scan's result does not depend on the other scans of the batch, on
 * the number of workgroups the device runs, nor on the scheduler (it does depend
 * on the block size, which groups the points into waves).  The work-queue scheduler runs the whole
 * batch in one launch; the round scheduler runs groups of up to MAX_BATCH_SCANS
 * scans and is also used when the queue would exceed MAX_QUEUE_ITEMS. */
static int align_packed(ndt_ctx* c, const ndt_grid* g, const float4* pts, const int64_t* offsets,
                        const int64_t* counts, int32_t S, const double* guesses, const ndt_params* p,
                        ndt_result* results, const double* alt = nullptr, int32_t* used_alt = nullptr) {
    if (g->params.mode != NDT_MODE_NDT) return fail(NDT_ERR_INVALID, "grid was not built in ndt mode");
    if (S <= 0) return fail(NDT_ERR_INVALID, "no scans");
    const int block = p->block_size ? p->block_size : 256;
    AlignParams prm = make_align_params(p, S);
    std::vector<ScanDesc> desc;
    int64_t slots = 0, max_chunks = 0;
    int64_t first_pass = 0;
    make_descs(offsets, counts, S, p, desc, &slots, &max_chunks, &first_pass);
    const uint64_t cap = queue_capacity(slots, p->max_iterations, alt != nullptr);
    const bool queue = p->scheduler != NDT_SCHED_ROUNDS && cap <= MAX_QUEUE_ITEMS && S < (1 << ITEM_SCAN_BITS) &&
                       max_chunks < (1 << ITEM_CHUNK_BITS);
    if (p->scheduler == NDT_SCHED_QUEUE && !queue)
        return fail(NDT_ERR_UNSUPPORTED, "batch too large for the work-queue scheduler");
    int bpc = 0;
    if (queue) LAUNCH_TRY(align_wq_max_blocks(block, prm.point_fp64, prm.sum_mode, &bpc));
    else LAUNCH_TRY(align_kernel_max_blocks(block, prm.point_fp64, prm.sum_mode, &bpc));
    if (bpc <= 0) return fail(NDT_ERR_UNSUPPORTED, "align kernel does not fit on a compute unit");
    const int32_t nwg = bpc * c->cus;
    BUF_TRY(c->guesses, sizeof(double) * 16 * S * (alt ? 2 : 1));
    BUF_TRY(c->states, sizeof(ScanState) * S);
    BUF_TRY(c->sync, sizeof(RoundSync) * ((S + MAX_BATCH_SCANS - 1) / MAX_BATCH_SCANS));
    BUF_TRY(c->out, sizeof(ScanOut) * S);
    BUF_TRY(c->desc, sizeof(ScanDesc) * S);
    BUF_TRY(c->partials, sizeof(double) * 32 * slots);
    BUF_TRY(c->small, 64 * sizeof(double));
    HIP_TRY(hipMemsetAsync(c->out.p, 0, sizeof(ScanOut) * S, c->stream));
    {
        const int eu = upload_staged2(c, c->guesses.p, guesses, sizeof(double) * 16 * S, c->desc.p, desc.data(),
                                      sizeof(ScanDesc) * S);
        if (eu) return eu;
        if (alt) {
            const int ea = upload_staged(c, c->guesses.as<double>() + 16 * (size_t)S, alt, sizeof(double) * 16 * S);
            if (ea) return ea;
        }
    }
    const double* d_alt = alt ? c->guesses.as<double>() + 16 * (size_t)S : nullptr;
    int32_t launches = 0;
    double prof_sum[4] = {0, 0, 0, 0};
    int aborted = 0;
    float ms_total = 0;
    if (queue) {
        BUF_TRY(c->queue, sizeof(WorkQueue));
        BUF_TRY(c->items, sizeof(uint64_t) * cap);
        BUF_TRY(c->done, sizeof(uint32_t) * S);
        HIP_TRY(hipMemsetAsync(c->items.p, 0, sizeof(uint64_t) * cap, c->stream));
        LAUNCH_TRY(launch_init_states(c->guesses.as<double>(), d_alt, S, prm, c->states.as<ScanState>(),
                                      c->sync.as<RoundSync>(), c->stream));
        LAUNCH_TRY(launch_init_queue(c->desc.as<ScanDesc>(), S, (uint64_t)first_pass, c->queue.as<WorkQueue>(),
                                     c->items.as<uint64_t>(), c->done.as<uint32_t>(), c->stream));
        HIP_TRY(hipEventRecord(c->ev0, c->stream));
        LAUNCH_TRY(launch_align_wq(g->dev, pts, c->desc.as<ScanDesc>(), nwg, block, prm,
                                   c->states.as<ScanState>(), c->queue.as<WorkQueue>(), c->items.as<uint64_t>(), cap,
                                   c->done.as<uint32_t>(), c->partials.as<double>(), c->out.as<ScanOut>(),
                                   c->small.as<double>(), 1, c->stream));
        HIP_TRY(hipEventRecord(c->ev1, c->stream));
        double prof[5];
        HIP_TRY(hipMemcpyAsync(prof, c->small.p, sizeof(prof), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipEventElapsedTime(&ms_total, c->ev0, c->ev1));
        for (int k = 0; k < 4; k++) prof_sum[k] = prof[k];
        aborted = prof[4] != 0.0;
        launches = 1;
    } else {
        for (int32_t s0 = 0; s0 < S; s0 += MAX_BATCH_SCANS) {
            const int32_t n = std::min<int32_t>(MAX_BATCH_SCANS, S - s0);
            RoundSync* rs = c->sync.as<RoundSync>() + s0 / MAX_BATCH_SCANS;
            LAUNCH_TRY(launch_init_states(c->guesses.as<double>() + 16 * (size_t)s0,
                                          d_alt ? d_alt + 16 * (size_t)s0 : nullptr, n, prm,
                                          c->states.as<ScanState>() + s0, rs, c->stream));
            HIP_TRY(hipEventRecord(c->ev0, c->stream));
            LAUNCH_TRY(launch_align(g->dev, pts, c->desc.as<ScanDesc>() + s0, n, nwg, block, prm,
                                    c->states.as<ScanState>() + s0, rs, c->partials.as<double>(),
                                    c->out.as<ScanOut>() + s0, c->small.as<double>(), c->stream));
            HIP_TRY(hipEventRecord(c->ev1, c->stream));
            double prof[5];
            HIP_TRY(hipMemcpyAsync(prof, c->small.p, sizeof(prof), hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            ms_total += ms;
            for (int k = 0; k < 4; k++) prof_sum[k] += prof[k];
            if (prof[4] != 0.0) aborted = 1;
            launches++;
        }
    }
    c->last_ms = ms_total;
    c->last_step_ms = c->last_ms;
    c->last_launches = launches;
    for (int k = 0; k < 4; k++) c->last_profile[k] = prof_sum[k];
    return collect_results(c, counts, S, aborted, results, used_alt);
}

/* Spreads the bits of a 10-bit value to every third bit (Morton interleave). */
static inline uint32_t spread3(uint32_t v) {
    v &= 0x3ff;
    v = (v | (v << 16)) & 0x030000FF;
    v = (v | (v << 8)) & 0x0300F00F;
    v = (v | (v << 4)) & 0x030C30C3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

/* Copies scans into one device buffer, each scan's points reordered along a Morton
 * curve of 1 m cells of its own (sensor-frame) coordinates: consecutive points,
 * and so the points of one workgroup, then look up neighbouring voxels and hit the
 * same cache lines.  A rigid transform keeps that locality.  The alignment does not
 * depend on the point order beyond floating-point rounding of the sums. */
static int pack_scans(ndt_ctx* c, DevBuf& dst, const float* const* scans, const int64_t* counts, int32_t S,
                      std::vector<int64_t>& offsets, int64_t* total) {
    offsets.assign(S, 0);
    int64_t t = 0;
    for (int32_t s = 0; s < S; s++) {
        if (!scans[s] || counts[s] <= 0) return fail(NDT_ERR_INVALID, "empty or null scan");
        if (counts[s] > 0x7fffffffll) return fail(NDT_ERR_UNSUPPORTED, "scan larger than 2^31 points");
        offsets[s] = t;
        t += counts[s];
    }
    BUF_TRY(dst, sizeof(float4) * t);
    std::vector<float> staging((size_t)(4 * t));
    auto work = [&](int32_t s0, int32_t s1) {
        /* stable LSD radix sort of the 30-bit Morton keys, three passes of 10 bits */
        std::vector<uint32_t> key, idx, tmp, cnt(1025);
        for (int32_t s = s0; s < s1; s++) {
            const float* src = scans[s];
            const int64_t n = counts[s];
            key.resize((size_t)n);
            idx.resize((size_t)n);
            tmp.resize((size_t)n);
            for (int64_t i = 0; i < n; i++) {
                uint32_t q[3];
                for (int k = 0; k < 3; k++) {
                    float v = src[4 * i + k];
                    double cell = std::isfinite(v) ? std::floor((double)v) + 512.0 : 0.0;
                    q[k] = (uint32_t)std::min(1023.0, std::max(0.0, cell));
                }
                key[(size_t)i] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
                idx[(size_t)i] = (uint32_t)i;
            }
            for (int sh = 0; sh < 30; sh += 10) {
                std::fill(cnt.begin(), cnt.end(), 0u);
                for (int64_t i = 0; i < n; i++) cnt[((key[idx[(size_t)i]] >> sh) & 1023u) + 1]++;
                for (int b = 0; b < 1024; b++) cnt[(size_t)b + 1] += cnt[(size_t)b];
                for (int64_t i = 0; i < n; i++) tmp[cnt[(key[idx[(size_t)i]] >> sh) & 1023u]++] = idx[(size_t)i];
                idx.swap(tmp);
            }
            float* d = staging.data() + 4 * offsets[s];
            for (int64_t i = 0; i < n; i++) std::memcpy(d + 4 * i, src + 4 * (size_t)idx[(size_t)i], 16);
        }
    };
    /* scans are independent: reorder them on up to 16 host threads */
    const int32_t nthreads = (int32_t)std::max(1u, std::min(16u, std::min(std::thread::hardware_concurrency(),
                                                                           (unsigned)((S + 3) / 4))));
    if (nthreads <= 1 || t < (1 << 16)) {
        work(0, S);
    } else {
        std::vector<std::thread> pool;
        for (int32_t k = 0; k < nthreads; k++)
            pool.emplace_back(work, (int32_t)((int64_t)S * k / nthreads), (int32_t)((int64_t)S * (k + 1) / nthreads));
        for (auto& th : pool) th.join();
    }
    HIP_TRY(hipMemcpyAsync(dst.p, staging.data(), sizeof(float4) * t, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *total = t;
    return 0;
}

extern "C" int ndt_align_batch(ndt_ctx* c, const ndt_grid* g, const float* const* scans, const int64_t* counts,
                               int32_t S, const double* guesses, const ndt_params* p, ndt_result* results) {
    if (!c || !g || !scans || !counts || !guesses || !results) return fail(NDT_ERR_INVALID, "null argument");
    if (S <= 0) return fail(NDT_ERR_INVALID, "no scans");
    int e = check_params(p);
    if (e) return e;
    if (p->mode != NDT_MODE_NDT) return fail(NDT_ERR_INVALID, "batched alignment runs in ndt mode");
    e = check_poses(guesses, S, "guess");
    if (e) return e;
    HIP_TRY(hipSetDevice(c->device));
    std::vector<int64_t> offsets;
    int64_t total = 0;
    e = pack_scans(c, c->pts, scans, counts, S, offsets, &total);
    if (e) return e;
    return align_packed(c, g, c->pts.as<float4>(), offsets.data(), counts, S, guesses, p, results);
}

extern "C" int ndt_align(ndt_ctx* c, const ndt_grid* g, const float* scan, int64_t n, const double guess[16],
                         const ndt_params* p, ndt_result* result) {
    if (!scan) return fail(NDT_ERR_INVALID, "scan is null");
    const float* scans[1] = {scan};
    int64_t counts[1] = {n};
    return ndt_align_batch(c, g, scans, counts, 1, guess, p, result);
}

extern "C" int ndt_scans_upload(ndt_ctx* c, const float* const* scans, const int64_t* counts, int32_t S,
                                ndt_scans** out) {
    if (!c || !scans || !counts || !out) return fail(NDT_ERR_INVALID, "null argument");
    *out = nullptr;
    if (S <= 0) return fail(NDT_ERR_INVALID, "no scans");
    HIP_TRY(hipSetDevice(c->device));
    ndt_scans* sc = new (std::nothrow) ndt_scans();
    if (sc) sc->serial = next_grid_serial();
    if (!sc) return fail(NDT_ERR_NOMEM, "host allocation failed");
    sc->ctx = c;
    sc->S = S;
    sc->counts.assign(counts, counts + S);
    int e = pack_scans(c, sc->pts, scans, counts, S, sc->offsets, &sc->total);
    if (!e) {
        hipError_t he = hipStreamSynchronize(c->stream);
        if (he != hipSuccess) e = fail(NDT_ERR_HIP, hipGetErrorString(he));
    }
    if (e) {
        delete sc;
        return e;
    }
    *out = sc;
    return 0;
}

extern "C" int ndt_scans_destroy(ndt_scans* sc) {
    if (!sc) return 0;
    (void)hipSetDevice(sc->ctx->device);
    delete sc;
    (void)hipGetLastError();
    return 0;
}

/* ------------------------------------------------------------------ allocation-free alignment */

/* ndt_align_scans without any allocation: the launch sequence of the captured
 * submission, with every buffer carved from the device's workspace array (its
 * address looked up when the context was created), the chunk descriptors made on
 * the device from the batch's point counts, and the results converted on the device
 * and copied straight into the caller's array.  Calls on one device share the
 * array, so calls that overlap on a device race; every region a launch reads
 * before writing is cleared by fill kernels first. */
extern "C" int ndt_align_scans_noalloc(ndt_ctx* c, const ndt_grid* g, const ndt_scans* sc, const double* guesses,
                                       const ndt_params* p, ndt_result* results) {
    if (!c || !g || !sc || !guesses || !results) return fail(NDT_ERR_INVALID, "null argument");
    int e = check_params(p);
    if (e) return e;
    if (p->mode != NDT_MODE_NDT) return fail(NDT_ERR_INVALID, "batched alignment runs in ndt mode");
    if (g->params.mode != NDT_MODE_NDT) return fail(NDT_ERR_INVALID, "grid was not built in ndt mode");
    e = check_poses(guesses, sc->S, "guess");
    if (e) return e;
    if (!c->scratch)
        return fail(NDT_ERR_HIP, std::string("workspace address of the device unknown: ") +
                                     hipGetErrorString((hipError_t)c->scratch_err));
    if (p->scheduler == NDT_SCHED_ROUNDS)
        return fail(NDT_ERR_UNSUPPORTED, "the allocation-free entry point runs the work-queue scheduler only");
    HIP_TRY(hipSetDevice(c->device));
    const int32_t S = sc->S;
    const int32_t chunk = p->points_per_wg > 0 ? p->points_per_wg : ALIGN_CHUNK_POINTS;
    const int32_t fine = fine_chunk(chunk);
    int64_t slots = 0, first_pass = 0, max_chunks = 0;
    for (int32_t s = 0; s < S; s++) {
        const int64_t n = sc->counts[(size_t)s];
        const int64_t fn = (n + fine - 1) / fine;
        slots += fn;
        first_pass += (n + chunk - 1) / chunk;
        max_chunks = std::max(max_chunks, fn);
    }
    const uint64_t cap = queue_capacity(slots, p->max_iterations);
    if (cap > MAX_QUEUE_ITEMS || S >= (1 << ITEM_SCAN_BITS) || max_chunks >= (1 << ITEM_CHUNK_BITS))
        return fail(NDT_ERR_UNSUPPORTED, "batch too large for the work-queue scheduler");
    /* workspace layout, 256-byte aligned regions */
    uint64_t used = 0;
    auto take = [&used](uint64_t bytes) {
        const uint64_t at = used;
        used = (used + bytes + 255) & ~(uint64_t)255;
        return at;
    };
    const uint64_t o_counts = take(sizeof(int64_t) * (uint64_t)S);
    const uint64_t o_desc = take(sizeof(ScanDesc) * (uint64_t)S);
    const uint64_t o_guess = take(sizeof(double) * 16 * (uint64_t)S);
    const uint64_t o_states = take(sizeof(ScanState) * (uint64_t)S);
    const uint64_t o_sync = take(sizeof(RoundSync));
    const uint64_t o_out = take(sizeof(ScanOut) * (uint64_t)S);
    const uint64_t o_part = take(sizeof(double) * 32 * (uint64_t)slots);
    const uint64_t o_queue = take(sizeof(WorkQueue));
    const uint64_t o_items = take(sizeof(uint64_t) * cap);
    const uint64_t o_done = take(sizeof(uint32_t) * (uint64_t)S);
    const uint64_t o_prof = take(sizeof(double) * 8);
    const uint64_t o_res = take(sizeof(ndt_result) * (uint64_t)S);
    if (used > NOALLOC_SCRATCH_BYTES)
        return fail(NDT_ERR_UNSUPPORTED, "batch needs " + std::to_string(used) +
                                             " bytes of workspace; the allocation-free entry point has " +
                                             std::to_string(NOALLOC_SCRATCH_BYTES) + " per device");
    const int block = p->block_size ? p->block_size : 256;
    AlignParams prm = make_align_params(p, S);
    int bpc = 0;
    LAUNCH_TRY(align_wq_max_blocks(block, prm.point_fp64, prm.sum_mode, &bpc));
    if (bpc <= 0) return fail(NDT_ERR_UNSUPPORTED, "align kernel does not fit on a compute unit");
    const int32_t nwg = bpc * c->cus;
    char* ws = (char*)c->scratch;
    int64_t* d_counts = (int64_t*)(ws + o_counts);
    ScanDesc* d_desc = (ScanDesc*)(ws + o_desc);
    double* d_guess = (double*)(ws + o_guess);
    ScanState* d_states = (ScanState*)(ws + o_states);
    RoundSync* d_sync = (RoundSync*)(ws + o_sync);
    ScanOut* d_out = (ScanOut*)(ws + o_out);
    double* d_part = (double*)(ws + o_part);
    WorkQueue* d_queue = (WorkQueue*)(ws + o_queue);
    uint64_t* d_items = (uint64_t*)(ws + o_items);
    uint32_t* d_done = (uint32_t*)(ws + o_done);
    double* d_prof = (double*)(ws + o_prof);
    ndt_result* d_res = (ndt_result*)(ws + o_res);
    HIP_TRY(hipMemcpyAsync(d_counts, sc->counts.data(), sizeof(int64_t) * (size_t)S, hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(d_guess, guesses, sizeof(double) * 16 * (size_t)S, hipMemcpyHostToDevice, c->stream));
    LAUNCH_TRY(launch_make_descs(d_counts, S, chunk, fine_pass_of(p), d_desc, c->stream));
    LAUNCH_TRY(launch_zero_u32(d_out, (int64_t)(sizeof(ScanOut) * (size_t)S / 4), c->stream));
    LAUNCH_TRY(launch_zero_u32(d_items, (int64_t)(2 * cap), c->stream));
    LAUNCH_TRY(launch_zero_u32(d_prof, 16, c->stream));
    LAUNCH_TRY(launch_init_states(d_guess, nullptr, S, prm, d_states, d_sync, c->stream));
    LAUNCH_TRY(launch_init_queue(d_desc, S, (uint64_t)first_pass, d_queue, d_items, d_done, c->stream));
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    LAUNCH_TRY(launch_align_wq(g->dev, sc->pts.as<float4>(), d_desc, nwg, block, prm, d_states, d_queue, d_items, cap,
                               d_done, d_part, d_out, d_prof, 0, c->stream));
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    LAUNCH_TRY(launch_noalloc_results(d_out, d_desc, S, d_prof, d_res, c->stream));
    double prof[8];
    HIP_TRY(hipMemcpyAsync(results, d_res, sizeof(ndt_result) * (size_t)S, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(prof, d_prof, sizeof(prof), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_step_ms = c->last_ms;
    c->last_launches = 1;
    c->last_pairs = (int64_t)prof[5];
    for (int k = 0; k < 4; k++) c->last_profile[k] = prof[k];
    if (prof[4] != 0.0) return fail(NDT_ERR_TIMEOUT, "a device-side wait exceeded its bound");
    return 0;
}

/* ------------------------------------------------------------------ submissions for callers' graphs */

int ndt::align_sub_prepare(ndt_ctx* c, const ndt_grid* g, const ndt_scans* sc, const ndt_params* p, bool fallback,
                           AlignSub& a) {
    if (g->params.mode != NDT_MODE_NDT) return fail(NDT_ERR_INVALID, "grid was not built in ndt mode");
    const int32_t S = sc->S;
    std::vector<ScanDesc> desc;
    int64_t slots = 0, max_chunks = 0, first_pass = 0;
    make_descs(sc->offsets.data(), sc->counts.data(), S, p, desc, &slots, &max_chunks, &first_pass);
    const uint64_t cap = queue_capacity(slots, p->max_iterations, fallback);
    if (p->scheduler == NDT_SCHED_ROUNDS || cap > MAX_QUEUE_ITEMS || S >= (1 << ITEM_SCAN_BITS) ||
        max_chunks >= (1 << ITEM_CHUNK_BITS))
        return NDT_ERR_UNSUPPORTED;
    a.block = p->block_size ? p->block_size : 256;
    a.prm = make_align_params(p, S);
    int bpc = 0;
    LAUNCH_TRY(align_wq_max_blocks(a.block, a.prm.point_fp64, a.prm.sum_mode, &bpc));
    if (bpc <= 0) return fail(NDT_ERR_UNSUPPORTED, "align kernel does not fit on a compute unit");
    a.nwg = bpc * c->cus;
    a.S = S;
    a.cap = cap;
    a.first_pass = first_pass;
    a.fallback = fallback;
    BUF_TRY(a.desc, sizeof(ScanDesc) * S);
    BUF_TRY(a.states, sizeof(ScanState) * S);
    BUF_TRY(a.sync, sizeof(RoundSync));
    BUF_TRY(a.out, sizeof(ScanOut) * S);
    BUF_TRY(a.partials, sizeof(double) * 32 * slots);
    BUF_TRY(a.queue, sizeof(WorkQueue));
    BUF_TRY(a.items, sizeof(uint64_t) * cap);
    BUF_TRY(a.done, sizeof(uint32_t) * S);
    BUF_TRY(a.prof, sizeof(double) * 8);
    a.release_host();
    HIP_TRY(hipHostMalloc(&a.h_out, sizeof(ScanOut) * S, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&a.h_prof, sizeof(double) * 8, hipHostMallocDefault));
    HIP_TRY(hipMemcpyAsync(a.desc.p, desc.data(), sizeof(ScanDesc) * S, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int ndt::align_sub_enqueue(ndt_ctx* c, const ndt_grid* g, const ndt_scans* sc, AlignSub& a, const double* d_guesses,
                           const double* d_alt) {
    const int32_t S = a.S;
    LAUNCH_TRY(launch_zero_u32(a.out.p, (int64_t)(sizeof(ScanOut) * S / 4), c->stream));
    LAUNCH_TRY(launch_zero_u32(a.items.p, (int64_t)(2 * a.cap), c->stream));
    LAUNCH_TRY(launch_zero_u32(a.prof.p, 16, c->stream));
    LAUNCH_TRY(launch_init_states(d_guesses, a.fallback ? d_alt : nullptr, S, a.prm, a.states.as<ScanState>(),
                                  a.sync.as<RoundSync>(), c->stream));
    LAUNCH_TRY(launch_init_queue(a.desc.as<ScanDesc>(), S, (uint64_t)a.first_pass, a.queue.as<WorkQueue>(),
                                 a.items.as<uint64_t>(), a.done.as<uint32_t>(), c->stream));
    LAUNCH_TRY(launch_align_wq(g->dev, sc->pts.as<float4>(), a.desc.as<ScanDesc>(), a.nwg, a.block, a.prm,
                               a.states.as<ScanState>(), a.queue.as<WorkQueue>(), a.items.as<uint64_t>(), a.cap,
                               a.done.as<uint32_t>(), a.partials.as<double>(), a.out.as<ScanOut>(),
                               a.prof.as<double>(), 0, c->stream));
    HIP_TRY(hipMemcpyAsync(a.h_out, a.out.p, sizeof(ScanOut) * S, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(a.h_prof, a.prof.p, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
    return 0;
}

int ndt::align_sub_results(ndt_ctx* c, const ndt_scans* sc, const AlignSub& a, ndt_result* results,
                           int32_t* used_alt) {
    for (int k = 0; k < 4; k++) c->last_profile[k] = a.h_prof[k];
    const int aborted = a.h_prof[4] != 0.0;
    return results_from_out((const ScanOut*)a.h_out, sc->counts.data(), a.S, aborted, c, results, used_alt);
}

/* ------------------------------------------------------------------ captured submission */

static bool same_align_params(const ndt_params& a, const ndt_params& b) {
    return a.mode == b.mode && a.resolution == b.resolution && a.step_size == b.step_size &&
           a.trans_eps == b.trans_eps && a.max_iterations == b.max_iterations && a.outlier_ratio == b.outlier_ratio &&
           a.point_precision == b.point_precision && a.sum_precision == b.sum_precision &&
           a.origin[0] == b.origin[0] && a.origin[1] == b.origin[1] && a.origin[2] == b.origin[2] &&
           a.block_size == b.block_size && a.points_per_wg == b.points_per_wg && a.scheduler == b.scheduler &&
           a.hessian_at_accept == b.hessian_at_accept && fine_pass_of(&a) == fine_pass_of(&b);
}

bool ndt::same_params(const ndt_params& a, const ndt_params& b) { return same_align_params(a, b); }
uint64_t ndt::wait_limit() { return wait_limit_ticks(); }

/* the captured launch sequence; the caller ends the capture whatever this returns */
static int capture_body(ndt_ctx* c, const ndt_grid* g, ndt_scans* sc, GraphPlan& gp, const AlignParams& prm,
                        int64_t first_pass, uint64_t cap, int32_t nwg, int block) {
    const int32_t S = sc->S;
    HIP_TRY(hipMemcpyAsync(gp.guesses.p, gp.h_guess, sizeof(double) * 16 * S, hipMemcpyHostToDevice, c->stream));
    LAUNCH_TRY(launch_zero_u32(gp.out.p, (int64_t)(sizeof(ScanOut) * S / 4), c->stream));
    LAUNCH_TRY(launch_zero_u32(gp.items.p, (int64_t)(2 * cap), c->stream));
    LAUNCH_TRY(launch_init_states(gp.guesses.as<double>(), nullptr, S, prm, gp.states.as<ScanState>(),
                                  gp.sync.as<RoundSync>(), c->stream));
    LAUNCH_TRY(launch_init_queue(gp.desc.as<ScanDesc>(), S, (uint64_t)first_pass, gp.queue.as<WorkQueue>(),
                                 gp.items.as<uint64_t>(), gp.done.as<uint32_t>(), c->stream));
    LAUNCH_TRY(launch_align_wq(g->dev, sc->pts.as<float4>(), gp.desc.as<ScanDesc>(), nwg, block, prm,
                               gp.states.as<ScanState>(), gp.queue.as<WorkQueue>(), gp.items.as<uint64_t>(), cap,
                               gp.done.as<uint32_t>(), gp.partials.as<double>(), gp.out.as<ScanOut>(),
                               gp.prof.as<double>(), 0, c->stream));
    HIP_TRY(hipMemcpyAsync(gp.h_out, gp.out.p, sizeof(ScanOut) * S, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(gp.h_prof, gp.prof.p, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
    return 0;
}

/* Captures the submission of a device-resident batch into a graph (once per scans
 * handle, grid and parameters).  Returns NDT_ERR_UNSUPPORTED when the batch needs
 * the round scheduler. */
static int graph_prepare(ndt_ctx* c, const ndt_grid* g, ndt_scans* sc, const ndt_params* p) {
    GraphPlan& gp = sc->plan;
    if (gp.valid && gp.grid_serial == g->serial && gp.device == c->device && same_align_params(gp.params, *p) &&
        gp.wait_ticks == wait_limit_ticks())
        return 0;
    gp.release();
    const int32_t S = sc->S;
    std::vector<ScanDesc> desc;
    int64_t slots = 0, max_chunks = 0;
    int64_t first_pass = 0;
    make_descs(sc->offsets.data(), sc->counts.data(), S, p, desc, &slots, &max_chunks, &first_pass);
    const uint64_t cap = queue_capacity(slots, p->max_iterations);
    if (p->scheduler == NDT_SCHED_ROUNDS || cap > MAX_QUEUE_ITEMS || S >= (1 << ITEM_SCAN_BITS) ||
        max_chunks >= (1 << ITEM_CHUNK_BITS))
        return NDT_ERR_UNSUPPORTED;
    const int block = p->block_size ? p->block_size : 256;
    AlignParams prm = make_align_params(p, S);
    int bpc = 0;
    LAUNCH_TRY(align_wq_max_blocks(block, prm.point_fp64, prm.sum_mode, &bpc));
    if (bpc <= 0) return fail(NDT_ERR_UNSUPPORTED, "align kernel does not fit on a compute unit");
    const int32_t nwg = bpc * c->cus;
    BUF_TRY(gp.desc, sizeof(ScanDesc) * S);
    BUF_TRY(gp.states, sizeof(ScanState) * S);
    BUF_TRY(gp.sync, sizeof(RoundSync));
    BUF_TRY(gp.out, sizeof(ScanOut) * S);
    BUF_TRY(gp.partials, sizeof(double) * 32 * slots);
    BUF_TRY(gp.queue, sizeof(WorkQueue));
    BUF_TRY(gp.items, sizeof(uint64_t) * cap);
    BUF_TRY(gp.done, sizeof(uint32_t) * S);
    BUF_TRY(gp.guesses, sizeof(double) * 16 * S);
    BUF_TRY(gp.prof, sizeof(double) * 8);
    HIP_TRY(hipHostMalloc((void**)&gp.h_guess, sizeof(double) * 16 * S, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(&gp.h_out, sizeof(ScanOut) * S, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&gp.h_prof, sizeof(double) * 8, hipHostMallocDefault));
    HIP_TRY(hipMemcpy(gp.desc.p, desc.data(), sizeof(ScanDesc) * S, hipMemcpyHostToDevice));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int e = capture_body(c, g, sc, gp, prm, first_pass, cap, nwg, block);
    hipGraph_t graph = nullptr;
    hipError_t he = hipStreamEndCapture(c->stream, &graph);
    if (e) {
        if (graph) (void)hipGraphDestroy(graph);
        return e;
    }
    if (he != hipSuccess) return fail(NDT_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(he));
    gp.graph = graph;
    HIP_TRY(hipGraphInstantiate(&gp.exec, gp.graph, nullptr, nullptr, 0));
    gp.valid = true;
    gp.wait_ticks = prm.wait_ticks;
    gp.grid_serial = g->serial;
    gp.device = c->device;
    gp.params = *p;
    return 0;
}

/* Replays the captured submission with new guesses. */
static int graph_align(ndt_ctx* c, ndt_scans* sc, const double* guesses, ndt_result* results) {
    GraphPlan& gp = sc->plan;
    const int32_t S = sc->S;
    std::memcpy(gp.h_guess, guesses, sizeof(double) * 16 * S);
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HIP_TRY(hipGraphLaunch(gp.exec, c->stream));
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_step_ms = c->last_ms;
    c->last_launches = 1;
    for (int k = 0; k < 4; k++) c->last_profile[k] = gp.h_prof[k];
    const int aborted = gp.h_prof[4] != 0.0;
    return results_from_out((const ScanOut*)gp.h_out, sc->counts.data(), S, aborted, c, results);
}

extern "C" int ndt_align_scans(ndt_ctx* c, const ndt_grid* g, const ndt_scans* sc, const double* guesses,
                               const ndt_params* p, ndt_result* results) {
    if (!c || !g || !sc || !guesses || !results) return fail(NDT_ERR_INVALID, "null argument");
    int e = check_params(p);
    if (e) return e;
    if (p->mode != NDT_MODE_NDT) return fail(NDT_ERR_INVALID, "batched alignment runs in ndt mode");
    e = check_poses(guesses, sc->S, "guess");
    if (e) return e;
    HIP_TRY(hipSetDevice(c->device));
    if (g->params.mode != NDT_MODE_NDT) return fail(NDT_ERR_INVALID, "grid was not built in ndt mode");
    ndt_scans* msc = const_cast<ndt_scans*>(sc);
    e = graph_prepare(c, g, msc, p);
    if (e == 0) return graph_align(c, msc, guesses, results);
    if (e != NDT_ERR_UNSUPPORTED) return e;
    return align_packed(c, g, sc->pts.as<float4>(), sc->offsets.data(), sc->counts.data(), sc->S, guesses, p, results);
}

/* Aligns the scans idx[0..n) of a device-resident batch (ndt mode). */
int ndt::align_scan_subset(ndt_ctx* c, const ndt_grid* g, const ndt_scans* sc, const int32_t* idx, int32_t n,
                           const double* guesses, const ndt_params* p, ndt_result* results, const double* alt,
                           int32_t* used_alt) {
    if (!c || !g || !sc || !idx || !guesses || !results) return fail(NDT_ERR_INVALID, "null argument");
    int e = check_params(p);
    if (e) return e;
    if (p->mode != NDT_MODE_NDT) return fail(NDT_ERR_INVALID, "batched alignment runs in ndt mode");
    std::vector<int64_t> off((size_t)n), cnt((size_t)n);
    for (int32_t k = 0; k < n; k++) {
        if (idx[k] < 0 || idx[k] >= sc->S) return fail(NDT_ERR_INVALID, "scan index out of range");
        off[(size_t)k] = sc->offsets[(size_t)idx[k]];
        cnt[(size_t)k] = sc->counts[(size_t)idx[k]];
    }
    e = check_poses(guesses, n, "guess");
    if (e) return e;
    if (alt && (e = check_poses(alt, n, "refined guess"))) return e;
    HIP_TRY(hipSetDevice(c->device));
    return align_packed(c, g, sc->pts.as<float4>(), off.data(), cnt.data(), n, guesses, p, results, alt, used_alt);
}

/* ------------------------------------------------------------------ autotuner */

static std::string tune_key(ndt_ctx* c, const ndt_scans* sc, const ndt_params* p) {
    char buf[512];
    double total = (double)sc->total, mag = 1.0;
    while (total >= 1000.0 * mag) mag *= 10.0;
    const double rounded = std::round(total / mag) * mag;
    std::snprintf(buf, sizeof(buf), "dev=%.120s;cus=%d;pp=%d;sp=%d;ha=%d;fp=%d;S=%d;N=%.0f", c->name, c->cus,
                  p->point_precision, p->point_precision == NDT_POINT_FP64 ? NDT_SUM_FP64 : p->sum_precision,
                  p->hessian_at_accept, fine_pass_of(p), sc->S, rounded);
    std::string k(buf);
    for (auto& ch : k)
        if (ch == ' ' || ch == '\n' || ch == '\t') ch = '_';
    return k;
}

extern "C" int ndt_autotune(ndt_ctx* c, const ndt_grid* g, const ndt_scans* sc, const double* guesses,
                            ndt_params* params, const char* cache_path, double* best_ms) {
    if (!c || !g || !sc || !guesses || !params) return fail(NDT_ERR_INVALID, "null argument");
    int e = check_params(params);
    if (e) return e;
    if (params->mode != NDT_MODE_NDT) return fail(NDT_ERR_INVALID, "the autotuner tunes ndt-mode alignment");
    e = check_poses(guesses, sc->S, "guess");
    if (e) return e;
    HIP_TRY(hipSetDevice(c->device));
    const std::string key = tune_key(c, sc, params);
    if (cache_path) {
        if (FILE* f = std::fopen(cache_path, "r")) {
            char line[1024];
            while (std::fgets(line, sizeof(line), f)) {
                char k[700];
                int blk = 0, ch = 0;
                double ms = 0.0;
                if (std::sscanf(line, "%699s %d %d %lf", k, &blk, &ch, &ms) == 4 && key == k) {
                    std::fclose(f);
                    params->block_size = blk;
                    params->points_per_wg = ch;
                    if (best_ms) *best_ms = ms;
                    return 0;
                }
            }
            std::fclose(f);
        }
    }
    static const int blocks[] = {128, 256, 512};
    static const int chunks[] = {1024, 2048, 4096};
    std::vector<ndt_result> res((size_t)sc->S);
    double best = 1e300;
    int best_blk = 256, best_ch = 2048;
    for (int blk : blocks)
        for (int ch : chunks) {
            ndt_params q = *params;
            q.block_size = blk;
            q.points_per_wg = ch;
            q.scheduler = NDT_SCHED_QUEUE;
            double t = 1e300;
            for (int rep = 0; rep < 3; rep++) { /* the first run captures the submission */
                e = ndt_align_scans(c, g, sc, guesses, &q, res.data());
                if (e) return e;
                if (rep > 0) t = std::min(t, c->last_ms);
            }
            if (t < best) {
                best = t;
                best_blk = blk;
                best_ch = ch;
            }
        }
    params->block_size = best_blk;
    params->points_per_wg = best_ch;
    if (best_ms) *best_ms = best;
    if (cache_path) {
        if (FILE* f = std::fopen(cache_path, "a")) {
            std::fprintf(f, "%s %d %d %.4f\n", key.c_str(), best_blk, best_ch, best);
            std::fclose(f);
        } else {
            return fail(NDT_ERR_INVALID, std::string("cannot write the autotuner cache ") + cache_path);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ fitness score */

extern "C" int ndt_fitness_score(ndt_ctx* c, const ndt_grid* g, const ndt_scans* sc, const double* poses,
                                 double max_range, double* fitness, int64_t* counts) {
    if (!c || !g || !sc || !poses || !fitness) return fail(NDT_ERR_INVALID, "null argument");
    if (g->params.mode != NDT_MODE_NDT) return fail(NDT_ERR_INVALID, "fitness needs an ndt-mode grid");
    if (max_range != max_range) return fail(NDT_ERR_INVALID, "max_range is NaN");
    int e = check_poses(poses, sc->S, "pose");
    if (e) return e;
    HIP_TRY(hipSetDevice(c->device));
    const int32_t S = sc->S;
    const double res = (double)g->params.resolution;
    int shells = FIT_MAX_SHELLS;
    double max_d2 = INFINITY;
    if (max_range > 0 && std::isfinite(max_range)) {
        shells = (int)std::min<double>(FIT_MAX_SHELLS, std::ceil(max_range / res) + 1.0);
        max_d2 = max_range * max_range;
    }
    std::vector<double> local(poses, poses + 16 * (size_t)S);
    for (int32_t s = 0; s < S; s++) {
        local[16 * (size_t)s + 3] -= g->params.origin[0];
        local[16 * (size_t)s + 7] -= g->params.origin[1];
        local[16 * (size_t)s + 11] -= g->params.origin[2];
    }
    std::vector<int32_t> cnt((size_t)S);
    int64_t maxn = 1;
    for (int32_t s = 0; s < S; s++) {
        cnt[(size_t)s] = (int32_t)sc->counts[(size_t)s];
        maxn = std::max<int64_t>(maxn, sc->counts[(size_t)s]);
    }
    const int nblk = (int)std::min<int64_t>(64, (maxn + 255) / 256);
    DevBuf d_off, d_cnt, d_pose, d_part;
    BUF_TRY(d_off, 8 * (size_t)S);
    BUF_TRY(d_cnt, 4 * (size_t)S);
    BUF_TRY(d_pose, 8 * 16 * (size_t)S);
    BUF_TRY(d_part, 16 * (size_t)S * nblk);
    HIP_TRY(hipMemcpyAsync(d_off.p, sc->offsets.data(), 8 * (size_t)S, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d_cnt.p, cnt.data(), 4 * (size_t)S, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d_pose.p, local.data(), 8 * local.size(), hipMemcpyHostToDevice, c->stream));
    LAUNCH_TRY(launch_fitness(g->fdev, sc->pts.as<float4>(), d_off.as<int64_t>(), d_cnt.as<int32_t>(), S,
                              d_pose.as<double>(), max_d2, shells, nblk, d_part.as<double>(), c->stream));
    std::vector<double> part(2 * (size_t)S * nblk);
    HIP_TRY(hipMemcpyAsync(part.data(), d_part.p, 8 * part.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int32_t s = 0; s < S; s++) {
        double sum = 0.0, n = 0.0;
        for (int b = 0; b < nblk; b++) {
            sum += part[((size_t)s * nblk + b) * 2];
            n += part[((size_t)s * nblk + b) * 2 + 1];
        }
        /* PCL returns the largest double when no point has a correspondence */
        fitness[s] = n > 0 ? sum / n : 1.7976931348623157e308;
        if (counts) counts[s] = (int64_t)n;
    }
    return 0;
}

/* ------------------------------------------------------------------ single derivative pass */

static int compat_pass(ndt_ctx* c, const ndt_grid* g, const float4* scan, int64_t n, const float M[16], const double p[6],
                       double d1, double d2, double* red /* NRED_COMPAT */);

extern "C" int ndt_derivatives(ndt_ctx* c, const ndt_grid* g, const float* scan, int64_t n, const double p[6],
                               const float compat_transform[16], const ndt_params* params, double* score,
                               double gradient[6], double hessian[36], int64_t* pairs) {
    if (!c || !g || !scan || !p || !params) return fail(NDT_ERR_INVALID, "null argument");
    if (n <= 0) return fail(NDT_ERR_INVALID, "empty scan");
    int e = check_params(params);
    if (e) return e;
    if (params->mode != g->params.mode) return fail(NDT_ERR_INVALID, "params mode differs from the grid's");
    HIP_TRY(hipSetDevice(c->device));
    BUF_TRY(c->pts, sizeof(float4) * n);
    HIP_TRY(hipMemcpyAsync(c->pts.p, scan, sizeof(float4) * n, hipMemcpyHostToDevice, c->stream));
    if (params->mode == NDT_MODE_COMPAT) {
        if (!compat_transform) return fail(NDT_ERR_INVALID, "compat mode needs the transform matrix");
        double d1, d2;
        gauss_consts((double)params->resolution, params->outlier_ratio, &d1, &d2);
        double red[NRED_COMPAT];
        e = compat_pass(c, g, c->pts.as<float4>(), n, compat_transform, p, d1, d2, red);
        if (e) return e;
        if (score) *score = red[0];
        if (gradient)
            for (int i = 0; i < 6; i++) gradient[i] = red[1 + i];
        if (hessian)
            for (int i = 0; i < 36; i++) hessian[i] = red[7 + i];
        if (pairs) *pairs = (int64_t)red[NACC_COMPAT];
        return 0;
    }
    AlignParams prm = make_align_params(params, 1);
    const int nblocks = (int)std::min<int64_t>(1024, (n + 255) / 256);
    BUF_TRY(c->partials2, sizeof(double) * NRED * nblocks + 64 * sizeof(double));
    BUF_TRY(c->small, 64 * sizeof(double));
    HIP_TRY(hipMemcpyAsync(c->small.p, p, sizeof(double) * 6, hipMemcpyHostToDevice, c->stream));
    double* red_dev = c->partials2.as<double>() + (size_t)NRED * nblocks;
    LAUNCH_TRY(launch_deriv_ndt(g->dev, c->pts.as<float4>(), n, c->small.as<double>(), prm, c->partials2.as<double>(),
                                nblocks, c->stream));
    LAUNCH_TRY(launch_reduce_partials(c->partials2.as<double>(), nblocks, NRED, red_dev, c->stream));
    double red[NRED];
    HIP_TRY(hipMemcpyAsync(red, red_dev, sizeof(red), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (score) *score = red[0];
    if (gradient)
        for (int i = 0; i < 6; i++) gradient[i] = red[1 + i];
    if (hessian)
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) hessian[i * 6 + j] = red[7 + (i <= j ? tri_index(i, j) : tri_index(j, i))];
    if (pairs) *pairs = (int64_t)red[NACC];
    return 0;
}

/* ------------------------------------------------------------------ compat alignment (host controller) */

static int compat_pass(ndt_ctx* c, const ndt_grid* g, const float4* scan, int64_t n, const float M[16], const double p[6],
                       double d1, double d2, double* red) {
    double jd[8][3], hd[15][3];
    angle_tables(p, jd, hd); /* host libm, as the reference */
    double jh[23 * 3];
    for (int k = 0; k < 8; k++)
        for (int q = 0; q < 3; q++) jh[k * 3 + q] = jd[k][q];
    for (int k = 0; k < 15; k++)
        for (int q = 0; q < 3; q++) jh[(8 + k) * 3 + q] = hd[k][q];
    /* sixteen threads per point (k_deriv_compat) */
    const int nblocks = (int)std::min<int64_t>(512, (16 * n + 255) / 256);
    BUF_TRY(c->partials2, sizeof(double) * NRED_COMPAT * (nblocks + 1));
    double* red_dev = c->partials2.as<double>() + (size_t)NRED_COMPAT * nblocks;
    LAUNCH_TRY(launch_deriv_compat(g->cdev, scan, n, M, jh, d1, d2, c->partials2.as<double>(), nblocks, c->stream));
    LAUNCH_TRY(launch_reduce_partials(c->partials2.as<double>(), nblocks, NRED_COMPAT, red_dev, c->stream));
    static_assert(sizeof(double) * NRED_COMPAT <= HSMALL_BYTES, "pinned readback buffer too small");
    HIP_TRY(hipMemcpyAsync(c->hsmall, red_dev, sizeof(double) * NRED_COMPAT, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::memcpy(red, c->hsmall, sizeof(double) * NRED_COMPAT);
    return 0;
}

/* buildTransformationMatrix, kernel.cpp:949-1013 (rotation only; translation kept) */
static void compat_build_matrix(float M[16], const double* t) {
    const float q_ha = (float)(0.5f * t[3]);
    const float q_w = (float)std::cos((double)q_ha);
    const float q_x = (float)std::sin((double)q_ha);
    const float q_y = 0.0f, q_z = 0.0f;
    const float q_ha2 = (float)(0.5f * t[4]);
    const float q_w2 = (float)std::cos((double)q_ha2);
    const float q_x2 = 0.0f;
    const float q_y2 = (float)std::sin((double)q_ha2);
    const float q_z2 = 0.0f;
    const float q_ha3 = (float)(0.5f * t[5]);
    const float q_w3 = (float)std::cos((double)q_ha3);
    const float q_x3 = 0.0f, q_y3 = 0.0f;
    const float q_z3 = (float)std::sin((double)q_ha3);
    const float r_x = q_w * q_w2 - q_x * q_x2 - q_y * q_y2 - q_z * q_z2;
    const float r_y = q_w * q_x2 + q_x * q_w2 + q_y * q_z2 - q_z * q_y2;
    const float r_z = q_w * q_y2 + q_y * q_w2 + q_z * q_x2 - q_x * q_z2;
    const float r_w = q_w * q_z2 + q_z * q_w2 + q_x * q_y2 - q_y * q_x2;
    const float r2_x = r_w * q_w3 - r_x * q_x3 - r_y * q_y3 - r_z * q_z3;
    const float r2_y = r_w * q_x3 + r_x * q_w3 + r_y * q_z3 - r_z * q_y3;
    const float r2_z = r_w * q_y3 + r_y * q_w3 + r_z * q_x3 - r_x * q_z3;
    const float r2_w = r_w * q_z3 + r_z * q_w3 + r_x * q_y3 - r_y * q_x3;
    const float tx = 2.0f * r2_x, ty = 2.0f * r2_y, tz = 2.0f * r2_z;
    const float twx = tx * r2_w, twy = ty * r2_w, twz = tz * r2_w;
    const float txx = tx * r2_x, txy = ty * r2_x, txz = tz * r2_x;
    const float tyy = ty * r2_y, tyz = tz * r2_y, tzz = tz * r2_z;
    M[12] = 0.0f; M[13] = 0.0f; M[14] = 0.0f; M[15] = 1.0f;
    M[0] = 1.0f - (tyy + tzz);
    M[1] = txy - twz;
    M[2] = txz + twy;
    M[4] = txy + twz;
    M[5] = 1.0f - (txx + tzz);
    M[6] = tyz - twx;
    M[8] = txz - twy;
    M[9] = tyz + twx;
    M[10] = 1.0f - (txx + tyy);
}

/* eulerAngles, kernel.cpp:1146-1173 */
static void compat_euler(const float M[16], double* result) {
    const double PI_REF = 3.1415926535897932384626433832795;
    double res[3];
    res[0] = std::atan2((double)M[6], (double)M[10]);
    double n1 = M[0], n2 = M[1];
    double c2 = std::sqrt(n1 * n1 + n2 * n2);
    if (res[0] > 0.0) {
        res[0] -= PI_REF;
        res[1] = std::atan2((double)(-M[2]), -c2);
    } else {
        res[1] = std::atan2((double)(-M[2]), c2);
    }
    double s1 = std::sin(res[0]), c1 = std::cos(res[0]);
    res[2] = std::atan2(s1 * M[8] - c1 * M[4], c1 * M[5] - s1 * M[9]);
    result[0] = -res[0];
    result[1] = -res[1];
    result[2] = -res[2];
}

extern "C" int ndt_compat_callback(ndt_ctx* c, const float* scan, int64_t n, const float guess[16], const float* map,
                                   int64_t m, ndt_result* result) {
    /* partial_points_callback -> ndt_align -> computeTransformation with the
     * reference's parameters (kernel.cpp:1405-1414, :1375-1392, :1175-1272) */
    if (!c || !scan || !guess || !map || !result) return fail(NDT_ERR_INVALID, "null argument");
    if (n <= 0 || m <= 0) return fail(NDT_ERR_INVALID, "empty scan or map");
    HIP_TRY(hipSetDevice(c->device));
    ndt_params prm;
    ndt_params_default(&prm, NDT_MODE_COMPAT);
    if (m > 0x7fffffffll) return fail(NDT_ERR_UNSUPPORTED, "map larger than 2^31 points");
    /* the context's compat grid is rebuilt in place: its buffers only grow */
    ndt_grid* g = c->compat_grid;
    if (!g) {
        g = new (std::nothrow) ndt_grid();
        if (!g) return fail(NDT_ERR_NOMEM, "host allocation failed");
        g->ctx = c;
        c->compat_grid = g;
    }
    g->serial = next_grid_serial();
    g->params = prm;
    std::memset(&g->info, 0, sizeof(g->info));
    std::memset(&g->dev, 0, sizeof(g->dev));
    std::memset(&g->cdev, 0, sizeof(g->cdev));
    int e = grid_build_compat(c, map, m, g);
    if (e) return e;
    DevBuf& scan_dev = c->compat_scan;
    BUF_TRY(scan_dev, sizeof(float4) * n);
    e = upload_staged(c, scan_dev.p, scan, sizeof(float4) * n);
    if (e) return e;
    double d1, d2;
    gauss_consts((double)prm.resolution, prm.outlier_ratio, &d1, &d2);
    float final_m[16];
    std::memcpy(final_m, guess, sizeof(final_m));
    double p[6], grad[6], H[36], red[NRED_COMPAT];
    p[0] = final_m[3];
    p[1] = final_m[7];
    p[2] = final_m[11];
    double ea[3];
    compat_euler(final_m, ea);
    p[3] = ea[0];
    p[4] = ea[1];
    p[5] = ea[2];
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    e = compat_pass(c, g, scan_dev.as<float4>(), n, final_m, p, d1, d2, red);
    if (e) return e;
    int passes = 1;
    double pairs = red[NACC_COMPAT];
    double score = red[0];
    for (int i = 0; i < 6; i++) grad[i] = red[1 + i];
    for (int i = 0; i < 36; i++) H[i] = red[7 + i];
    const double step_size = prm.step_size, trans_eps = prm.trans_eps;
    int converged = 0, nr_iterations = 0;
    while (!converged) {
        double neg_grad[6], delta_p[6];
        for (int i = 0; i < 6; i++) neg_grad[i] = -grad[i];
        solve6(delta_p, H, neg_grad);
        double delta_p_norm = 1; /* kernel.cpp:1231 */
        for (int i = 0; i < 6; i++) delta_p[i] /= delta_p_norm;
        /* computeStepLengthMT (kernel.cpp:1015-1144): the interval starts converged,
         * so only the first trial step a_t = clamp(1, 0.05, 0.1) is evaluated */
        double step_init = delta_p_norm, step_max = step_size, step_min = trans_eps / 2;
        double phi_0 = -score;
        double d_phi_0 = -dot6(grad, delta_p);
        double a_t = 0;
        if (!(d_phi_0 >= 0 && d_phi_0 == 0)) {
            if (d_phi_0 >= 0) {
                d_phi_0 *= -1;
                for (int i = 0; i < 6; i++) delta_p[i] *= -1;
            }
            (void)phi_0;
            a_t = step_init;
            a_t = dmin_std(a_t, step_max);
            a_t = dmax_std(a_t, step_min);
            double x_t[6];
            for (int i = 0; i < 6; i++) x_t[i] = p[i] + delta_p[i] * a_t;
            compat_build_matrix(final_m, x_t);
            e = compat_pass(c, g, scan_dev.as<float4>(), n, final_m, x_t, d1, d2, red);
            if (e) return e;
            passes++;
            pairs += red[NACC_COMPAT];
            score = red[0];
            for (int i = 0; i < 6; i++) grad[i] = red[1 + i];
            for (int i = 0; i < 36; i++) H[i] = red[7 + i];
        }
        delta_p_norm = a_t;
        for (int i = 0; i < 6; i++) delta_p[i] *= delta_p_norm;
        for (int i = 0; i < 6; i++) p[i] = p[i] + delta_p[i];
        if (nr_iterations > prm.max_iterations || (nr_iterations && (std::fabs(delta_p_norm) < trans_eps))) converged = 1;
        nr_iterations++;
    }
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_step_ms = c->last_ms;
    c->last_launches = 2 * passes;
    c->last_pairs = (int64_t)pairs;
    std::memset(result, 0, sizeof(*result));
    for (int k = 0; k < 16; k++) {
        result->final_transformation[k] = final_m[k];
        result->final_transformation_f64[k] = final_m[k];
    }
    result->score = score;
    result->trans_probability = score / (double)n;
    result->iterations = nr_iterations;
    result->converged = converged;
    result->passes = passes;
    return 0;
}
