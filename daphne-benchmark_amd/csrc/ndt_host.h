/*
 * ndt_host.h - host-side helpers shared by the C ABI translation units
 * (ndt_api.cpp, ndt_learn.cpp): error recording, device buffers and the opaque
 * handle structures.
 */
#ifndef NDT_HOST_H
#define NDT_HOST_H

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "ndt_hip.h"
#include "ndt_internal.h"

namespace ndt {

/* ------------------------------------------------------------------ errors */

/* records the message of a failure for ndt_last_error() and returns code */
int fail(int code, const std::string& msg);

/* NDT_ERR_INVALID unless every entry of the n row-major 4x4 poses is finite. */
int check_poses(const double* poses, int64_t n, const char* what);

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (hipError_t)(expr);                                                        \
        if (e_ != hipSuccess)                                                                      \
            return fail(e_ == hipErrorOutOfMemory ? NDT_ERR_NOMEM : NDT_ERR_HIP,                   \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                        \
    } while (0)

#define LAUNCH_TRY(expr)                                                                           \
    do {                                                                                           \
        int e_ = (expr);                                                                           \
        if (e_ != 0)                                                                               \
            return fail(NDT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString((hipError_t)e_)); \
    } while (0)

/* ------------------------------------------------------------------ device buffers */

/* Bytes of device memory held by all DevBufs of the process. */
std::atomic<int64_t>& device_bytes_held();

/* Injected device-memory limit for failure-path tests: with the environment
 * variable NDT_DEVICE_MEMORY_LIMIT set to a byte count, a DevBuf allocation that
 * would take device_bytes_held() above it fails with NDT_ERR_NOMEM exactly as an
 * exhausted device would.  -1 when unset.  Read at every allocation. */
int64_t device_memory_limit();

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            (void)hipFree(p);
            device_bytes_held() -= (int64_t)bytes;
        }
        p = nullptr;
        bytes = 0;
    }
    int ensure(size_t want) {
        if (want <= bytes && p) return 0;
        release();
        size_t b = want ? want : 1;
        const int64_t limit = device_memory_limit();
        if (limit >= 0 && device_bytes_held().load() + (int64_t)b > limit)
            return fail(NDT_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(b) +
                                           "): above NDT_DEVICE_MEMORY_LIMIT=" + std::to_string(limit));
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(e == hipErrorOutOfMemory ? NDT_ERR_NOMEM : NDT_ERR_HIP,
                        std::string("hipMalloc(") + std::to_string(b) + "): " + hipGetErrorString(e));
        }
        bytes = b;
        device_bytes_held() += (int64_t)b;
        return 0;
    }
    template <class T> T* as() const { return (T*)p; }
};

#define BUF_TRY(buf, bytes)                \
    do {                                   \
        int e_ = (buf).ensure(bytes);      \
        if (e_) return e_;                 \
    } while (0)

} // namespace ndt

using ndt::DevBuf;

namespace ndt {
/* aligns the scans idx[0..n) of a device-resident batch from guesses[0..n) (ndt mode);
 * with alt (refined starts, n x 16) in one launch: a probe pass at each given guess,
 * NDT from the refined guess where it scores at least as high as the given guess,
 * and plain NDT from the given guess otherwise or where the refined run ended with
 * a lower score than the given guess has (used_alt[k] = 1 where the refined run is
 * the result; passes count every pass spent) */
int align_scan_subset(ndt_ctx* c, const ndt_grid* g, const ndt_scans* sc, const int32_t* idx, int32_t n,
                      const double* guesses, const ndt_params* p, ndt_result* results, const double* alt = nullptr,
                      int32_t* used_alt = nullptr);

/* A context lives until ndt_ctx_destroy has been called and every handle made on it
 * is destroyed: ndt_ctx_destroy drops the creator's reference, each handle holds
 * one (CtxRef).  The last release synchronises and destroys the stream. */
void ctx_retain(ndt_ctx* c);
void ctx_release(ndt_ctx* c);

/* A handle's reference to its context.  Declared as the handle's first member, it
 * is released after the handle's device buffers are freed. */
struct CtxRef {
    ndt_ctx* p = nullptr;
    CtxRef() = default;
    CtxRef(const CtxRef&) = delete;
    CtxRef& operator=(const CtxRef&) = delete;
    CtxRef& operator=(ndt_ctx* c) {
        if (c) ctx_retain(c);
        if (p) ctx_release(p);
        p = c;
        return *this;
    }
    ~CtxRef() {
        if (p) ctx_release(p);
    }
    operator ndt_ctx*() const { return p; }
    ndt_ctx* operator->() const { return p; }
};
} // namespace ndt

/* bytes of a context's small pinned readback buffer */
constexpr size_t HSMALL_BYTES = 4096;

/* opaque handles of the C ABI */
struct ndt_ctx {
    std::atomic<int> refs{1};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev2 = nullptr; /* between the two graphs of a captured learned step */
    int cus = 0;
    /* device address of the allocation-free entry point's workspace on this
     * context's device (looked up at creation), or null with the lookup's error */
    void* scratch = nullptr;
    int scratch_err = 0;
    char name[256] = {0};
    /* workspaces */
    DevBuf pts, desc, states, sync, partials, out, guesses, small, partials2, queue, items, done;
    /* grid-build temporaries, and the compat callback's grid and scan, kept for
     * reuse: freeing device memory synchronises the device, which would serialise
     * contexts that work concurrently (one per host thread) */
    DevBuf bk_keys, bk_vals, bk_keys_t, bk_vals_t, bk_scratch, bk_flags, bk_ranks, bk_starts, bk_sc2, bk_mm;
    DevBuf compat_scan;
    ndt_grid* compat_grid = nullptr;
    /* pinned staging buffer for host-to-device uploads; pin_pending while a copy
     * from it may still be in flight on the stream */
    void* pin = nullptr;
    size_t pin_bytes = 0;
    bool pin_pending = false;
    /* small pinned buffer for scalars read back by host-side control (bounding
     * box, run counts, a compat pass's sums); each use ends with a stream
     * synchronisation before the next */
    void* hsmall = nullptr;
    double last_ms = 0.0;
    double last_step_ms = 0.0; /* a captured learned step: the whole graph */
    int32_t last_launches = 0;
    int64_t last_pairs = 0;
    double last_profile[4] = {0, 0, 0, 0};
};

struct ndt_grid {
    ndt::CtxRef ctx;
    uint64_t serial = 0; /* unique per built grid (captured submissions refer to it) */
    ndt_params params;
    ndt_grid_info info;
    /* ndt */
    ndt::NdtGridDev dev;
    DevBuf slots, dense, cand_head, cand_s32, vs64, icov_f, mean_c, icov_c;
    int32_t dims[3] = {0, 0, 0};
    /* map points sorted by voxel + occupied-voxel hash (fitness score) */
    ndt::FitGridDev fdev;
    DevBuf fit_pts, fit_slots;
    /* all occupied voxels (diagnostics) */
    DevBuf all_mean, all_icov, all_cov, all_count, all_valid, all_key;
    /* compat */
    ndt::CompatGridDev cdev;
    int bits[3] = {0, 0, 0};
    int64_t origin_cell[3] = {0, 0, 0};
};

/* A captured submission of a device-resident batch (ndt_align_scans): guesses in
 * from pinned memory, state and queue set-up, the work-queue launch, results out to
 * pinned memory.  Its buffers are its own, so other calls cannot disturb it. */
struct GraphPlan {
    bool valid = false;
    uint64_t grid_serial = 0;
    uint64_t wait_ticks = 0;
    int device = -1;
    ndt_params params{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf desc, states, sync, out, partials, queue, items, done, guesses, prof;
    double* h_guess = nullptr;
    void* h_out = nullptr;
    double* h_prof = nullptr;
    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (h_guess) (void)hipHostFree(h_guess);
        if (h_out) (void)hipHostFree(h_out);
        if (h_prof) (void)hipHostFree(h_prof);
        exec = nullptr;
        graph = nullptr;
        h_guess = nullptr;
        h_out = nullptr;
        h_prof = nullptr;
        valid = false;
    }
    ~GraphPlan() { release(); }
};

namespace ndt {
/* The buffers of one alignment submission of a device-resident batch in the work
 * queue, for a caller that records the submission in its own graph: device
 * workspace, the chunk descriptors (uploaded by align_sub_prepare) and pinned
 * host memory for the results. */
struct AlignSub {
    DevBuf desc, states, sync, out, partials, queue, items, done, prof;
    void* h_out = nullptr;
    double* h_prof = nullptr;
    AlignParams prm{};
    uint64_t cap = 0;
    int64_t first_pass = 0;
    int32_t nwg = 0, block = 256, S = 0;
    bool fallback = false;
    void release_host() {
        if (h_out) (void)hipHostFree(h_out);
        if (h_prof) (void)hipHostFree(h_prof);
        h_out = nullptr;
        h_prof = nullptr;
    }
    ~AlignSub() { release_host(); }
};
/* Sizes and allocates a submission for the whole batch (fallback: with refined
 * starts); NDT_ERR_UNSUPPORTED when the batch needs the round scheduler. */
int align_sub_prepare(ndt_ctx* c, const ndt_grid* g, const ndt_scans* sc, const ndt_params* p, bool fallback,
                      AlignSub& a);
/* Enqueues the submission on the context's stream without synchronising (it may
 * be recorded into a graph): set-up kernels, the work-queue launch from the
 * device-resident guesses (and refined starts), results to the pinned buffers. */
int align_sub_enqueue(ndt_ctx* c, const ndt_grid* g, const ndt_scans* sc, AlignSub& a, const double* d_guesses,
                      const double* d_alt);
/* Results of a completed submission (after the stream was synchronised). */
int align_sub_results(ndt_ctx* c, const ndt_scans* sc, const AlignSub& a, ndt_result* results, int32_t* used_alt);
/* parameters that give the same alignment launch */
bool same_params(const ndt_params& a, const ndt_params& b);
/* bound of a device-side wait in s_memrealtime ticks (NDT_WAIT_LIMIT_US or 2 s) */
uint64_t wait_limit();
} // namespace ndt

struct ndt_scans {
    ndt::CtxRef ctx;
    uint64_t serial = 0; /* unique per uploaded batch (derived data cached elsewhere refers to it) */
    int32_t S = 0;
    std::vector<int64_t> counts, offsets;
    int64_t total = 0;
    DevBuf pts;
    GraphPlan plan;
};

#endif
