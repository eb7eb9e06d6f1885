/*
 * ndt_learn.cpp - host side of the learned warm start (include/ndt_learn.h):
 * map tiling, model storage and initialisation, the training loop (features,
 * forward, backward and Adam on the GPU), prediction and batched NDT from the
 * refined guesses with its fallback to plain NDT.
 */
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ndt_hip.h"
#include "ndt_learn.h"
#include "ndt_host.h"
#include "ndt_learn_internal.h"

using namespace ndt;

namespace {

/* correction scales: x, y, z in metres, roll, pitch, yaw in radians */
const double CORR_SCALE[6] = {3.0, 3.0, 1.0, 0.05, 0.05, 0.2};

uint16_t f2bf_host(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

/* splitmix64 stream with Box-Muller normals: reproducible on every platform */
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

/* correction parameters of target relative to guess: C = guess^-1 target */
void correction_params(const double* G, const double* T, double* out6) {
    double C[16];
    /* guess^-1 = [R^T, -R^T t] */
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 4; c++) {
            double v = 0.0;
            for (int k = 0; k < 3; k++) v += G[k * 4 + r] * (T[k * 4 + c] - (c == 3 ? G[k * 4 + 3] : 0.0));
            C[r * 4 + c] = v;
        }
    }
    C[12] = C[13] = C[14] = 0.0;
    C[15] = 1.0;
    double e[3];
    pose_euler(C, e);
    out6[0] = C[3];
    out6[1] = C[7];
    out6[2] = C[11];
    out6[3] = e[0];
    out6[4] = e[1];
    out6[5] = e[2];
}

int pick_splits(int M, int N, int K, int cus) {
    const int tiles = (M / GEMM_BM) * (N / GEMM_BN);
    int s = 1;
    while (tiles * s < 2 * cus && K % (GEMM_BK * s * 2) == 0 && s < 64) s *= 2;
    return s;
}

} // namespace

/* ------------------------------------------------------------------ registration settings */

static bool reg_on(const ndt_register_params& rp) { return rp.radius > 0; }

static int reg_check(const ndt_register_params* rp) {
    if (!rp) return fail(NDT_ERR_INVALID, "null registration parameters");
    if (rp->radius < 0 || rp->radius > NDT_REGISTER_MAX_RADIUS)
        return fail(NDT_ERR_INVALID, "registration radius must be 0..8 cells");
    if (rp->radius == 0) return 0;
    if (rp->yaws < 1 || rp->yaws > NDT_REGISTER_MAX_YAWS || rp->yaws % 2 == 0)
        return fail(NDT_ERR_INVALID, "registration yaws must be odd, 1..11");
    if (rp->yaws > 1 && !(rp->yaw_step > 0.0 && rp->yaw_step <= 0.05))
        return fail(NDT_ERR_INVALID, "registration yaw_step must be in (0, 0.05] rad");
    if (rp->refine_yaw != 0 && rp->refine_yaw != 1) return fail(NDT_ERR_INVALID, "refine_yaw must be 0 or 1");
    if (rp->z_search != 0 && rp->z_search != 1) return fail(NDT_ERR_INVALID, "z_search must be 0 or 1");
    if (rp->level != 0 && rp->level != 1) return fail(NDT_ERR_INVALID, "level must be 0 or 1");
    return 0;
}

/* the kernel's view: cosine, sine and tangent of each searched yaw (k - (yaws - 1) / 2) * step */
static RegisterDev reg_dev(const ndt_register_params& rp) {
    RegisterDev d{};
    d.radius = rp.radius;
    d.yaws = rp.yaws;
    d.refine_yaw = rp.refine_yaw;
    d.z_search = rp.z_search;
    d.level = rp.level;
    d.yaw_step = rp.yaw_step;
    for (int k = 0; k < rp.yaws; k++) {
        const double th = (double)(k - (rp.yaws - 1) / 2) * rp.yaw_step;
        d.c[k] = std::cos(th);
        d.s[k] = std::sin(th);
        d.t[k] = std::tan(th);
    }
    return d;
}

static bool same_reg(const ndt_register_params& a, const ndt_register_params& b) {
    return a.radius == b.radius && a.yaws == b.yaws && a.yaw_step == b.yaw_step && a.refine_yaw == b.refine_yaw &&
           a.z_search == b.z_search && a.level == b.level;
}

/* ------------------------------------------------------------------ handles */

struct ndt_bevmap {
    ndt::CtxRef ctx;
    std::atomic<int> refs{1}; /* the creator's, plus one per dataset made on it */
    uint64_t serial = 0;      /* unique per BEV map */
    DevBuf pts, start, lut;
    TileGridDev dev{};
    double origin[3] = {0, 0, 0};
    int64_t m = 0;
};

struct ndt_model {
    ndt::CtxRef ctx;
    int f0 = 0, f1 = 0, f2 = 0;
    DevBuf w[6], mom[6], vel[6], g[6];
    DevBuf wb1, wb2, wb2t, wb3;
    int64_t step = 0;
    /* batch workspace */
    int wsB = 0, splits1 = 1, splits2 = 1;
    DevBuf X, XT, P1, H1, H1T, P2, H2, Y, dY, lossb, dZ2, dZ2T, dH1, dZ1T, target, weights, lossacc, bpart;
    /* prediction inputs: the scan grids of the batch being predicted for, its scan
     * offsets and counts, the feature rows' scan indices, the world and map-local
     * guesses, and the refined guesses */
    DevBuf sbev, idx, gl, off, cnt, gw, refined;
    /* the registration before the model (radius 0: none) and the registered world
     * guesses the refinement starts from */
    ndt_register_params reg{};
    DevBuf greg;
    /* the captured learned step of the last (scans, grid, BEV map, parameters) */
    struct LearnedPlan {
        bool valid = false;
        uint64_t scans = 0, grid = 0, bm = 0, wait = 0;
        int device = -1, keep = 0;
        ndt_params params{};
        ndt_register_params reg{};
        std::vector<const void*> bufs; /* workspace the graphs refer to */
        /* two graphs launched back to back: the prediction, then the alignment
         * (events between them time the align kernel) */
        hipGraph_t graph[2] = {nullptr, nullptr};
        hipGraphExec_t exec[2] = {nullptr, nullptr};
        double* h_guess = nullptr;
        ndt::AlignSub sub;
        void release() {
            for (int k = 0; k < 2; k++) {
                if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
                if (graph[k]) (void)hipGraphDestroy(graph[k]);
                exec[k] = nullptr;
                graph[k] = nullptr;
            }
            if (h_guess) (void)hipHostFree(h_guess);
            h_guess = nullptr;
            valid = false;
        }
        ~LearnedPlan() { release(); }
    } plan;
};

/* Drops one reference to a BEV map; the last frees it. */
static void bevmap_release(ndt_bevmap* bm) {
    if (bm->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(bm->ctx->device);
    delete bm;
    (void)hipGetLastError();
}

struct ndt_dataset {
    ndt::CtxRef ctx;
    const ndt_bevmap* bm = nullptr; /* retained: training crops its map grids */
    ~ndt_dataset() {
        if (bm) bevmap_release(const_cast<ndt_bevmap*>(bm));
    }
    int64_t n = 0;
    int64_t nscans = 0;
    DevBuf scan_bev; /* nscans grids; capacity grows by doubling */
    std::vector<int32_t> scan_index;
    std::vector<double> guesses;  /* map-local, 16 per sample */
    std::vector<float> targets;   /* scaled corrections, 6 per sample */
    DevBuf d_idx, d_guess, d_target;
};

static int64_t tensor_size(const ndt_model* md, int which) {
    switch (which) {
    case NDT_W1: return (int64_t)md->f1 * md->f0;
    case NDT_B1: return md->f1;
    case NDT_W2: return (int64_t)md->f2 * md->f1;
    case NDT_B2: return md->f2;
    case NDT_W3: return (int64_t)NOUT * md->f2;
    case NDT_B3: return NOUT;
    default: return -1;
    }
}

/* ------------------------------------------------------------------ map tiles */

extern "C" int ndt_bevmap_build(ndt_ctx* c, const float* map, int64_t m, const double origin[3], ndt_bevmap** out) {
    if (!c || !map || !out) return fail(NDT_ERR_INVALID, "null argument");
    if (m <= 0 || m > 0xffffffffll) return fail(NDT_ERR_INVALID, "map must have 1 .. 2^32-1 points");
    *out = nullptr;
    HIP_TRY(hipSetDevice(c->device));
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
    int64_t finite = 0;
    for (int64_t i = 0; i < m; i++) {
        const float* p = map + 4 * i;
        if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) continue;
        finite++;
        for (int k = 0; k < 2; k++) {
            lo[k] = std::min(lo[k], (double)p[k]);
            hi[k] = std::max(hi[k], (double)p[k]);
        }
    }
    if (!finite) return fail(NDT_ERR_INVALID, "map has no finite points");
    const int64_t ntx = (int64_t)std::floor((hi[0] - lo[0]) / MAP_TILE) + 1;
    const int64_t nty = (int64_t)std::floor((hi[1] - lo[1]) / MAP_TILE) + 1;
    if (ntx * nty > (1ll << 26)) return fail(NDT_ERR_UNSUPPORTED, "map footprint too large for the tile grid");
    /* counting sort by tile, row-major over (ty, tx) */
    std::vector<uint32_t> start((size_t)(ntx * nty + 1), 0u);
    std::vector<uint32_t> tile((size_t)m, 0xffffffffu);
    for (int64_t i = 0; i < m; i++) {
        const float* p = map + 4 * i;
        if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) continue;
        int64_t tx = (int64_t)std::floor(((double)p[0] - lo[0]) / MAP_TILE);
        int64_t ty = (int64_t)std::floor(((double)p[1] - lo[1]) / MAP_TILE);
        tx = std::min(std::max<int64_t>(tx, 0), ntx - 1);
        ty = std::min(std::max<int64_t>(ty, 0), nty - 1);
        tile[(size_t)i] = (uint32_t)(ty * ntx + tx);
        start[(size_t)tile[(size_t)i] + 1]++;
    }
    for (size_t t = 1; t < start.size(); t++) start[t] += start[t - 1];
    std::vector<float> sorted((size_t)(4 * finite));
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < m; i++) {
        if (tile[(size_t)i] == 0xffffffffu) continue;
        const uint32_t d = fill[tile[(size_t)i]]++;
        std::memcpy(&sorted[(size_t)d * 4], map + 4 * i, 16);
    }
    std::vector<uint16_t> lut(65536);
    for (int k = 0; k < 65536; k++) lut[(size_t)k] = f2bf_host((float)std::log1p((double)k));
    ndt_bevmap* bm = new ndt_bevmap();
    {
        static std::atomic<uint64_t> next_bm{1};
        bm->serial = next_bm.fetch_add(1);
    }
    bm->ctx = c;
    bm->m = finite;
    for (int k = 0; k < 3; k++) bm->origin[k] = origin ? origin[k] : 0.0;
    int e = 0;
    if ((e = bm->pts.ensure(sorted.size() * 4)) || (e = bm->start.ensure(start.size() * 4)) ||
        (e = bm->lut.ensure(lut.size() * 2))) {
        delete bm;
        return e;
    }
    hipError_t he = hipMemcpy(bm->pts.p, sorted.data(), sorted.size() * 4, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(bm->start.p, start.data(), start.size() * 4, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(bm->lut.p, lut.data(), lut.size() * 2, hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        delete bm;
        return fail(NDT_ERR_HIP, std::string("bevmap upload: ") + hipGetErrorString(he));
    }
    bm->dev.pts = bm->pts.as<float4>();
    bm->dev.start = bm->start.as<uint32_t>();
    bm->dev.ntx = (int)ntx;
    bm->dev.nty = (int)nty;
    bm->dev.x0 = lo[0];
    bm->dev.y0 = lo[1];
    *out = bm;
    return 0;
}

extern "C" int ndt_bevmap_destroy(ndt_bevmap* bm) {
    if (!bm) return 0;
    bevmap_release(bm);
    return 0;
}

/* ------------------------------------------------------------------ model */

static int ensure_workspace(ndt_model* md, int B) {
    const int64_t f0 = md->f0, f1 = md->f1, f2 = md->f2;
    /* the split-K counts depend on this call's rows only, so that a result does not
     * depend on the sizes of earlier calls */
    md->splits1 = pick_splits(B, md->f1, md->f0, md->ctx->cus);
    md->splits2 = pick_splits(B, md->f2, md->f1, md->ctx->cus);
    BUF_TRY(md->P1, 4 * (size_t)md->splits1 * B * f1);
    BUF_TRY(md->P2, 4 * (size_t)md->splits2 * B * f2);
    if (md->wsB >= B) return 0;
    BUF_TRY(md->X, 2 * (size_t)B * f0);
    BUF_TRY(md->XT, 2 * (size_t)B * f0);
    BUF_TRY(md->H1, 2 * (size_t)B * f1);
    BUF_TRY(md->H1T, 2 * (size_t)B * f1);
    BUF_TRY(md->H2, 2 * (size_t)B * f2);
    BUF_TRY(md->Y, 4 * (size_t)B * NOUT);
    BUF_TRY(md->dY, 4 * (size_t)B * NOUT);
    BUF_TRY(md->target, 4 * (size_t)B * NOUT);
    BUF_TRY(md->weights, 4 * NOUT);
    BUF_TRY(md->lossb, 4 * (size_t)B);
    BUF_TRY(md->dZ2, 2 * (size_t)B * f2);
    BUF_TRY(md->dZ2T, 2 * (size_t)B * f2);
    BUF_TRY(md->dH1, 4 * (size_t)B * f1);
    BUF_TRY(md->dZ1T, 2 * (size_t)B * f1);
    BUF_TRY(md->lossacc, 4 * 64);
    BUF_TRY(md->bpart, 4 * (size_t)bwd_tiles(B) * (size_t)std::max<int64_t>(f1, (NOUT + 1) * f2 + NOUT));
    md->wsB = B;
    return 0;
}

/* bf16 copies of the weight matrices (and the transpose of W2) from the masters */
static int refresh_bf16(ndt_model* md) {
    hipStream_t st = md->ctx->stream;
    LAUNCH_TRY(launch_to_bf16(md->w[NDT_W1].as<float>(), tensor_size(md, NDT_W1), md->wb1.as<uint16_t>(), st));
    LAUNCH_TRY(launch_to_bf16(md->w[NDT_W2].as<float>(), tensor_size(md, NDT_W2), md->wb2.as<uint16_t>(), st));
    LAUNCH_TRY(launch_to_bf16(md->w[NDT_W3].as<float>(), tensor_size(md, NDT_W3), md->wb3.as<uint16_t>(), st));
    LAUNCH_TRY(launch_transpose_u16(md->wb2.as<uint16_t>(), md->f2, md->f1, md->wb2t.as<uint16_t>(), st));
    return 0;
}

static int model_alloc(ndt_ctx* c, int f0, int f1, int f2, ndt_model** out) {
    if (f0 <= 0 || f1 <= 0 || f2 <= 0 || f0 % 128 || f1 % 128 || f2 % 128)
        return fail(NDT_ERR_INVALID, "model dimensions must be positive multiples of 128");
    ndt_model* md = new ndt_model();
    md->ctx = c;
    md->f0 = f0;
    md->f1 = f1;
    md->f2 = f2;
    int e = 0;
    for (int k = 0; k < 6 && !e; k++) {
        const size_t bytes = 4 * (size_t)tensor_size(md, k);
        if ((e = md->w[k].ensure(bytes)) || (e = md->mom[k].ensure(bytes)) || (e = md->vel[k].ensure(bytes)) ||
            (e = md->g[k].ensure(bytes)))
            break;
        hipError_t he = hipMemset(md->mom[k].p, 0, bytes);
        if (he == hipSuccess) he = hipMemset(md->vel[k].p, 0, bytes);
        if (he == hipSuccess) he = hipMemset(md->g[k].p, 0, bytes);
        if (he != hipSuccess) e = fail(NDT_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(he));
    }
    if (!e) e = md->wb1.ensure(2 * (size_t)tensor_size(md, NDT_W1));
    if (!e) e = md->wb2.ensure(2 * (size_t)tensor_size(md, NDT_W2));
    if (!e) e = md->wb2t.ensure(2 * (size_t)tensor_size(md, NDT_W2));
    if (!e) e = md->wb3.ensure(2 * (size_t)tensor_size(md, NDT_W3));
    if (e) {
        delete md;
        return e;
    }
    *out = md;
    return 0;
}

static int upload_tensor(ndt_model* md, int which, const float* src) {
    const size_t bytes = 4 * (size_t)tensor_size(md, which);
    HIP_TRY(hipMemcpyAsync(md->w[which].p, src, bytes, hipMemcpyHostToDevice, md->ctx->stream));
    HIP_TRY(hipStreamSynchronize(md->ctx->stream));
    return 0;
}

extern "C" int ndt_model_create(ndt_ctx* c, int32_t f0, int32_t f1, int32_t f2, uint64_t seed, ndt_model** out) {
    if (!c || !out) return fail(NDT_ERR_INVALID, "null argument");
    *out = nullptr;
    HIP_TRY(hipSetDevice(c->device));
    ndt_model* md = nullptr;
    int e = model_alloc(c, f0, f1, f2, &md);
    if (e) return e;
    Rng rng(seed);
    const int fan_in[6] = {f0, 0, f1, 0, f2, 0};
    for (int k = 0; k < 6 && !e; k++) {
        std::vector<float> h((size_t)tensor_size(md, k), 0.0f);
        if (k == NDT_W1 || k == NDT_W2) {
            const double sd = std::sqrt(2.0 / fan_in[k]); /* He initialisation for ReLU layers */
            for (auto& v : h) v = (float)(sd * rng.normal());
        } else if (k == NDT_W3) {
            const double sd = 0.1 * std::sqrt(1.0 / fan_in[k]); /* start near "no correction" */
            for (auto& v : h) v = (float)(sd * rng.normal());
        }
        e = upload_tensor(md, k, h.data());
    }
    if (!e) e = refresh_bf16(md);
    if (!e) {
        hipError_t he = hipStreamSynchronize(c->stream);
        if (he != hipSuccess) e = fail(NDT_ERR_HIP, hipGetErrorString(he));
    }
    if (e) {
        delete md;
        return e;
    }
    *out = md;
    return 0;
}

extern "C" int ndt_model_destroy(ndt_model* md) {
    if (!md) return 0;
    if (md->ctx) (void)hipSetDevice(md->ctx->device);
    delete md;
    (void)hipGetLastError();
    return 0;
}

extern "C" int ndt_model_dims(const ndt_model* md, int32_t dims[3]) {
    if (!md || !dims) return fail(NDT_ERR_INVALID, "null argument");
    dims[0] = md->f0;
    dims[1] = md->f1;
    dims[2] = md->f2;
    return 0;
}

extern "C" void ndt_register_params_default(ndt_register_params* rp) {
    if (!rp) return;
    *rp = ndt_register_params{};
    rp->radius = 6;
    rp->yaws = 7;
    rp->yaw_step = 3.14159265358979323846 / 180.0;
    rp->refine_yaw = 1;
    rp->z_search = 1;
    rp->level = 1;
}

extern "C" int ndt_model_set_register(ndt_model* md, const ndt_register_params* rp) {
    if (!md) return fail(NDT_ERR_INVALID, "null model");
    if (int e = reg_check(rp)) return e;
    md->reg = *rp;
    return 0;
}

extern "C" int ndt_model_get_register(const ndt_model* md, ndt_register_params* rp) {
    if (!md || !rp) return fail(NDT_ERR_INVALID, "null argument");
    *rp = md->reg;
    return 0;
}

extern "C" int ndt_model_get(const ndt_model* md, int32_t which, float* out, int64_t n) {
    if (!md || !out) return fail(NDT_ERR_INVALID, "null argument");
    if (which < 0 || which > 5 || n != tensor_size(md, which)) return fail(NDT_ERR_INVALID, "bad tensor or size");
    HIP_TRY(hipSetDevice(md->ctx->device));
    HIP_TRY(hipStreamSynchronize(md->ctx->stream));
    HIP_TRY(hipMemcpy(out, md->w[which].p, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int ndt_model_grad(const ndt_model* md, int32_t which, float* out, int64_t n) {
    if (!md || !out) return fail(NDT_ERR_INVALID, "null argument");
    if (which < 0 || which > 5 || n != tensor_size(md, which)) return fail(NDT_ERR_INVALID, "bad tensor or size");
    HIP_TRY(hipSetDevice(md->ctx->device));
    HIP_TRY(hipStreamSynchronize(md->ctx->stream));
    HIP_TRY(hipMemcpy(out, md->g[which].p, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int ndt_model_set(ndt_model* md, int32_t which, const float* in, int64_t n) {
    if (!md || !in) return fail(NDT_ERR_INVALID, "null argument");
    if (which < 0 || which > 5 || n != tensor_size(md, which)) return fail(NDT_ERR_INVALID, "bad tensor or size");
    HIP_TRY(hipSetDevice(md->ctx->device));
    int e = upload_tensor(md, which, in);
    if (e) return e;
    e = refresh_bf16(md);
    if (e) return e;
    HIP_TRY(hipStreamSynchronize(md->ctx->stream));
    return 0;
}

static const char MODEL_MAGIC[8] = {'N', 'D', 'T', 'M', 'L', 'P', '0', '1'};
/* optional trailer after the tensors: the registration setting (radius, yaws,
 * refine_yaw, z_search and level as int32, then yaw_step; files with the earlier
 * trailer, without level, load with level 0) */
static const char REG_MAGIC[8] = {'N', 'D', 'T', 'R', 'E', 'G', '0', '3'};
static const char REG_MAGIC_02[8] = {'N', 'D', 'T', 'R', 'E', 'G', '0', '2'};

extern "C" int ndt_model_save(const ndt_model* md, const char* path) {
    if (!md || !path) return fail(NDT_ERR_INVALID, "null argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(NDT_ERR_INVALID, std::string("cannot write ") + path);
    int32_t dims[3] = {md->f0, md->f1, md->f2};
    bool ok = std::fwrite(MODEL_MAGIC, 1, 8, f) == 8 && std::fwrite(dims, 4, 3, f) == 3;
    for (int k = 0; k < 6 && ok; k++) {
        std::vector<float> h((size_t)tensor_size(md, k));
        if (ndt_model_get(md, k, h.data(), (int64_t)h.size())) {
            std::fclose(f);
            return NDT_ERR_HIP;
        }
        ok = std::fwrite(h.data(), 4, h.size(), f) == h.size();
    }
    if (ok && reg_on(md->reg)) {
        const int32_t ri[5] = {md->reg.radius, md->reg.yaws, md->reg.refine_yaw, md->reg.z_search, md->reg.level};
        ok = std::fwrite(REG_MAGIC, 1, 8, f) == 8 && std::fwrite(ri, 4, 5, f) == 5 &&
             std::fwrite(&md->reg.yaw_step, 8, 1, f) == 1;
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : fail(NDT_ERR_INVALID, std::string("short write to ") + path);
}

extern "C" int ndt_model_load(ndt_ctx* c, const char* path, ndt_model** out) {
    if (!c || !path || !out) return fail(NDT_ERR_INVALID, "null argument");
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(NDT_ERR_INVALID, std::string("cannot read ") + path);
    char magic[8];
    int32_t dims[3];
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, MODEL_MAGIC, 8) || std::fread(dims, 4, 3, f) != 3) {
        std::fclose(f);
        return fail(NDT_ERR_INVALID, std::string("not a model file: ") + path);
    }
    HIP_TRY(hipSetDevice(c->device));
    ndt_model* md = nullptr;
    int e = model_alloc(c, dims[0], dims[1], dims[2], &md);
    if (e) {
        std::fclose(f);
        return e;
    }
    for (int k = 0; k < 6 && !e; k++) {
        std::vector<float> h((size_t)tensor_size(md, k));
        if (std::fread(h.data(), 4, h.size(), f) != h.size()) e = fail(NDT_ERR_INVALID, "truncated model file");
        else e = upload_tensor(md, k, h.data());
    }
    char tag[8];
    if (!e && std::fread(tag, 1, 8, f) == 8) {
        int32_t ri[5] = {0, 0, 0, 0, 0};
        ndt_register_params rp{};
        const int nint = !std::memcmp(tag, REG_MAGIC, 8) ? 5 : !std::memcmp(tag, REG_MAGIC_02, 8) ? 4 : 0;
        if (!nint || std::fread(ri, 4, (size_t)nint, f) != (size_t)nint || std::fread(&rp.yaw_step, 8, 1, f) != 1)
            e = fail(NDT_ERR_INVALID, "bad registration trailer in model file");
        if (!e) {
            rp.radius = ri[0];
            rp.yaws = ri[1];
            rp.refine_yaw = ri[2];
            rp.z_search = ri[3];
            rp.level = ri[4];
            if (!(e = reg_check(&rp))) md->reg = rp;
        }
    }
    std::fclose(f);
    if (!e) e = refresh_bf16(md);
    if (!e) {
        hipError_t he = hipStreamSynchronize(c->stream);
        if (he != hipSuccess) e = fail(NDT_ERR_HIP, hipGetErrorString(he));
    }
    if (e) {
        delete md;
        return e;
    }
    *out = md;
    return 0;
}

/* Forward over the B rows in md->X; with `train`, also the loss terms against
 * md->target and the transposed activations the backward pass reads. */
static int forward_dev(ndt_model* md, int B, bool train) {
    hipStream_t st = md->ctx->stream;
    const int f0 = md->f0, f1 = md->f1, f2 = md->f2;
    LAUNCH_TRY(launch_gemm_nt(md->X.as<uint16_t>(), f0, md->wb1.as<uint16_t>(), f0, md->P1.as<float>(), f1, B, f1, f0,
                              md->splits1, st));
    LAUNCH_TRY(launch_bias_act(md->P1.as<float>(), md->splits1, B, f1, md->w[NDT_B1].as<float>(), 1,
                               md->H1.as<uint16_t>(), train ? md->H1T.as<uint16_t>() : nullptr, st));
    LAUNCH_TRY(launch_gemm_nt(md->H1.as<uint16_t>(), f1, md->wb2.as<uint16_t>(), f1, md->P2.as<float>(), f2, B, f2, f1,
                              md->splits2, st));
    LAUNCH_TRY(launch_bias_act(md->P2.as<float>(), md->splits2, B, f2, md->w[NDT_B2].as<float>(), 1,
                               md->H2.as<uint16_t>(), nullptr, st));
    LAUNCH_TRY(launch_head_fwd(md->H2.as<uint16_t>(), md->wb3.as<uint16_t>(), md->w[NDT_B3].as<float>(), B, f2,
                               md->Y.as<float>(), train ? md->target.as<float>() : nullptr, md->weights.as<float>(),
                               md->dY.as<float>(), md->lossb.as<float>(), st));
    return 0;
}

/* Backward of the loss of the last training forward: gradients into md->g. */
static int backward_dev(ndt_model* md, int B) {
    hipStream_t st = md->ctx->stream;
    const int f0 = md->f0, f1 = md->f1, f2 = md->f2;
    LAUNCH_TRY(launch_head_bwd(md->dY.as<float>(), md->H2.as<uint16_t>(), md->wb3.as<uint16_t>(), B, f2,
                               md->g[NDT_W3].as<float>(), md->g[NDT_B3].as<float>(), md->dZ2.as<uint16_t>(),
                               md->dZ2T.as<uint16_t>(), md->g[NDT_B2].as<float>(), md->bpart.as<float>(), st));
    /* dH1 [B x f1] = dZ2 [B x f2] W2 = dZ2 (W2^T)^T */
    LAUNCH_TRY(launch_gemm_nt(md->dZ2.as<uint16_t>(), f2, md->wb2t.as<uint16_t>(), f2, md->dH1.as<float>(), f1, B, f1,
                              f2, 1, st));
    LAUNCH_TRY(launch_relu_bwd(md->dH1.as<float>(), md->H1.as<uint16_t>(), B, f1, md->dZ1T.as<uint16_t>(),
                               md->g[NDT_B1].as<float>(), md->bpart.as<float>(), st));
    /* dW2 [f2 x f1] = dZ2^T H1 */
    LAUNCH_TRY(launch_gemm_nt(md->dZ2T.as<uint16_t>(), B, md->H1T.as<uint16_t>(), B, md->g[NDT_W2].as<float>(), f1, f2,
                              f1, B, 1, st));
    /* dW1 [f1 x f0] = dZ1^T X */
    LAUNCH_TRY(launch_transpose_u16(md->X.as<uint16_t>(), B, f0, md->XT.as<uint16_t>(), st));
    LAUNCH_TRY(launch_gemm_nt(md->dZ1T.as<uint16_t>(), B, md->XT.as<uint16_t>(), B, md->g[NDT_W1].as<float>(), f0, f1,
                              f0, B, 1, st));
    return 0;
}

static int adam_dev(ndt_model* md, float lr, float b1, float b2, float eps, float wd) {
    hipStream_t st = md->ctx->stream;
    md->step++;
    const float bc1 = (float)(1.0 - std::pow((double)b1, (double)md->step));
    const float bc2 = (float)(1.0 - std::pow((double)b2, (double)md->step));
    for (int k = 0; k < 6; k++) {
        uint16_t* wb = k == NDT_W1 ? md->wb1.as<uint16_t>()
                       : k == NDT_W2 ? md->wb2.as<uint16_t>()
                       : k == NDT_W3 ? md->wb3.as<uint16_t>()
                                     : nullptr;
        const float wdk = (k == NDT_W1 || k == NDT_W2 || k == NDT_W3) ? wd : 0.0f;
        LAUNCH_TRY(launch_adam(md->w[k].as<float>(), md->mom[k].as<float>(), md->vel[k].as<float>(),
                               md->g[k].as<float>(), tensor_size(md, k), lr, b1, b2, eps, bc1, bc2, wdk, wb, st));
    }
    LAUNCH_TRY(launch_transpose_u16(md->wb2.as<uint16_t>(), md->f2, md->f1, md->wb2t.as<uint16_t>(), st));
    return 0;
}

static int check_batch(const ndt_model* md, int32_t B) {
    if (B <= 0 || B % 128) return fail(NDT_ERR_INVALID, "batch must be a positive multiple of 128");
    (void)md;
    return 0;
}

extern "C" int ndt_model_forward(ndt_ctx* c, ndt_model* md, const uint16_t* x, int32_t B, float* y) {
    if (!c || !md || !x || !y) return fail(NDT_ERR_INVALID, "null argument");
    int e = check_batch(md, B);
    if (e) return e;
    HIP_TRY(hipSetDevice(c->device));
    if ((e = ensure_workspace(md, B))) return e;
    HIP_TRY(hipMemcpyAsync(md->X.p, x, 2 * (size_t)B * md->f0, hipMemcpyHostToDevice, c->stream));
    if ((e = forward_dev(md, B, false))) return e;
    HIP_TRY(hipMemcpyAsync(y, md->Y.p, 4 * (size_t)B * NOUT, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int ndt_model_backward(ndt_ctx* c, ndt_model* md, const uint16_t* x, const float* target,
                                  const float* weights, int32_t B, float* loss) {
    if (!c || !md || !x || !target) return fail(NDT_ERR_INVALID, "null argument");
    int e = check_batch(md, B);
    if (e) return e;
    HIP_TRY(hipSetDevice(c->device));
    if ((e = ensure_workspace(md, B))) return e;
    float w[NOUT] = {1, 1, 1, 1, 1, 1};
    if (weights) std::memcpy(w, weights, sizeof(w));
    HIP_TRY(hipMemcpyAsync(md->X.p, x, 2 * (size_t)B * md->f0, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(md->target.p, target, 4 * (size_t)B * NOUT, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(md->weights.p, w, sizeof(w), hipMemcpyHostToDevice, c->stream));
    if ((e = forward_dev(md, B, true))) return e;
    if ((e = backward_dev(md, B))) return e;
    std::vector<float> lb((size_t)B);
    HIP_TRY(hipMemcpyAsync(lb.data(), md->lossb.p, 4 * (size_t)B, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (loss) {
        double s = 0.0;
        for (float v : lb) s += v;
        *loss = (float)(s / B);
    }
    return 0;
}

extern "C" int ndt_model_adam_step(ndt_ctx* c, ndt_model* md, float lr, float beta1, float beta2, float eps,
                                   float weight_decay) {
    if (!c || !md) return fail(NDT_ERR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    int e = adam_dev(md, lr, beta1, beta2, eps, weight_decay);
    if (e) return e;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

/* ------------------------------------------------------------------ features */

/* scan grids of every scan of a batch into md->sbev */
static int scan_grids(ndt_model* md, const ndt_bevmap* bm, const ndt_scans* sc, DevBuf& dst) {
    hipStream_t st = md->ctx->stream;
    BUF_TRY(dst, 2 * (size_t)sc->S * BEV_CELLS);
    BUF_TRY(md->off, 8 * (size_t)sc->S);
    BUF_TRY(md->cnt, 4 * (size_t)sc->S);
    std::vector<int32_t> cnt((size_t)sc->S);
    for (int32_t s = 0; s < sc->S; s++) cnt[(size_t)s] = (int32_t)sc->counts[(size_t)s];
    HIP_TRY(hipMemcpyAsync(md->off.p, sc->offsets.data(), 8 * (size_t)sc->S, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(md->cnt.p, cnt.data(), 4 * (size_t)sc->S, hipMemcpyHostToDevice, st));
    LAUNCH_TRY(launch_bev_scan(sc->pts.as<float4>(), md->off.as<int64_t>(), md->cnt.as<int32_t>(), sc->S,
                               bm->lut.as<uint16_t>(), dst.as<uint16_t>(), st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

/* md->X rows 0..B: scan grid rows idx[b] of `sbev` and map grids at local guesses */
static int batch_features(ndt_model* md, const ndt_bevmap* bm, const uint16_t* sbev, const int32_t* d_idx,
                          const double* d_guess_local, int B) {
    hipStream_t st = md->ctx->stream;
    LAUNCH_TRY(launch_gather_rows(sbev, d_idx, B, BEV_CELLS, md->X.as<uint16_t>(), md->f0, st));
    LAUNCH_TRY(launch_bev_crop(bm->dev, d_guess_local, B, bm->lut.as<uint16_t>(), md->X.as<uint16_t>(), md->f0,
                               BEV_CELLS, st));
    return 0;
}

static void guesses_local(const ndt_bevmap* bm, const double* world, int64_t n, std::vector<double>& out) {
    out.assign(world, world + 16 * n);
    for (int64_t i = 0; i < n; i++) {
        out[(size_t)(16 * i + 3)] -= bm->origin[0];
        out[(size_t)(16 * i + 7)] -= bm->origin[1];
        out[(size_t)(16 * i + 11)] -= bm->origin[2];
    }
}

extern "C" int ndt_features(ndt_ctx* c, const ndt_bevmap* bm, const ndt_scans* sc, const int32_t* scan_index,
                            const double* guesses, int32_t B, uint16_t* out) {
    if (!c || !bm || !sc || !scan_index || !guesses || !out) return fail(NDT_ERR_INVALID, "null argument");
    if (B <= 0) return fail(NDT_ERR_INVALID, "B must be positive");
    for (int32_t b = 0; b < B; b++)
        if (scan_index[b] < 0 || scan_index[b] >= sc->S) return fail(NDT_ERR_INVALID, "scan index out of range");
    if (int e = check_poses(guesses, B, "guess")) return e;
    HIP_TRY(hipSetDevice(c->device));
    /* a throw-away model-less workspace: features only need X */
    DevBuf sbev, X, idx, gl, off, cnt;
    hipStream_t st = c->stream;
    BUF_TRY(sbev, 2 * (size_t)sc->S * BEV_CELLS);
    BUF_TRY(off, 8 * (size_t)sc->S);
    BUF_TRY(cnt, 4 * (size_t)sc->S);
    std::vector<int32_t> counts((size_t)sc->S);
    for (int32_t s = 0; s < sc->S; s++) counts[(size_t)s] = (int32_t)sc->counts[(size_t)s];
    HIP_TRY(hipMemcpyAsync(off.p, sc->offsets.data(), 8 * (size_t)sc->S, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(cnt.p, counts.data(), 4 * (size_t)sc->S, hipMemcpyHostToDevice, st));
    LAUNCH_TRY(launch_bev_scan(sc->pts.as<float4>(), off.as<int64_t>(), cnt.as<int32_t>(), sc->S,
                               bm->lut.as<uint16_t>(), sbev.as<uint16_t>(), st));
    std::vector<double> gl_h;
    guesses_local(bm, guesses, B, gl_h);
    BUF_TRY(X, 2 * (size_t)B * NDT_FEATURES);
    BUF_TRY(idx, 4 * (size_t)B);
    BUF_TRY(gl, 8 * gl_h.size());
    HIP_TRY(hipMemcpyAsync(idx.p, scan_index, 4 * (size_t)B, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(gl.p, gl_h.data(), 8 * gl_h.size(), hipMemcpyHostToDevice, st));
    LAUNCH_TRY(launch_gather_rows(sbev.as<uint16_t>(), idx.as<int32_t>(), B, BEV_CELLS, X.as<uint16_t>(), NDT_FEATURES,
                                  st));
    LAUNCH_TRY(launch_bev_crop(bm->dev, gl.as<double>(), B, bm->lut.as<uint16_t>(), X.as<uint16_t>(), NDT_FEATURES,
                               BEV_CELLS, st));
    HIP_TRY(hipMemcpyAsync(out, X.p, 2 * (size_t)B * NDT_FEATURES, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

extern "C" int ndt_bev_register(ndt_ctx* c, const ndt_bevmap* bm, const ndt_scans* sc, const int32_t* scan_index,
                                const double* guesses, int32_t B, const ndt_register_params* rp, double* out_guesses,
                                int32_t* out_score) {
    if (!c || !bm || !sc || !scan_index || !guesses || !out_guesses) return fail(NDT_ERR_INVALID, "null argument");
    if (B <= 0) return fail(NDT_ERR_INVALID, "B must be positive");
    if (int e = reg_check(rp)) return e;
    for (int32_t b = 0; b < B; b++)
        if (scan_index[b] < 0 || scan_index[b] >= sc->S) return fail(NDT_ERR_INVALID, "scan index out of range");
    if (int e = check_poses(guesses, B, "guess")) return e;
    if (!reg_on(*rp)) {
        std::memcpy(out_guesses, guesses, 16 * sizeof(double) * (size_t)B);
        if (out_score) std::memset(out_score, 0, 4 * (size_t)B);
        return 0;
    }
    HIP_TRY(hipSetDevice(c->device));
    DevBuf sbev, off, cnt, idx, gl, gw, score;
    hipStream_t st = c->stream;
    BUF_TRY(sbev, 2 * (size_t)sc->S * BEV_CELLS);
    BUF_TRY(off, 8 * (size_t)sc->S);
    BUF_TRY(cnt, 4 * (size_t)sc->S);
    std::vector<int32_t> counts((size_t)sc->S);
    for (int32_t s = 0; s < sc->S; s++) counts[(size_t)s] = (int32_t)sc->counts[(size_t)s];
    HIP_TRY(hipMemcpyAsync(off.p, sc->offsets.data(), 8 * (size_t)sc->S, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(cnt.p, counts.data(), 4 * (size_t)sc->S, hipMemcpyHostToDevice, st));
    LAUNCH_TRY(launch_bev_scan(sc->pts.as<float4>(), off.as<int64_t>(), cnt.as<int32_t>(), sc->S,
                               bm->lut.as<uint16_t>(), sbev.as<uint16_t>(), st));
    std::vector<double> gl_h;
    guesses_local(bm, guesses, B, gl_h);
    BUF_TRY(idx, 4 * (size_t)B);
    BUF_TRY(gl, 8 * gl_h.size());
    BUF_TRY(gw, 8 * gl_h.size());
    BUF_TRY(score, 4 * (size_t)B);
    HIP_TRY(hipMemcpyAsync(idx.p, scan_index, 4 * (size_t)B, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(gl.p, gl_h.data(), 8 * gl_h.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(gw.p, guesses, 8 * gl_h.size(), hipMemcpyHostToDevice, st));
    LAUNCH_TRY(launch_bev_register(bm->dev, sbev.as<uint16_t>(), sc->pts.as<float4>(), off.as<int64_t>(),
                                   cnt.as<int32_t>(), idx.as<int32_t>(), gl.as<double>(), gw.as<double>(), B,
                                   reg_dev(*rp), gl.as<double>(), gw.as<double>(), score.as<int32_t>(), st));
    HIP_TRY(hipMemcpyAsync(out_guesses, gw.p, 8 * gl_h.size(), hipMemcpyDeviceToHost, st));
    if (out_score) HIP_TRY(hipMemcpyAsync(out_score, score.p, 4 * (size_t)B, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

/* ------------------------------------------------------------------ training */

extern "C" void ndt_train_params_default(ndt_train_params* tp) {
    if (!tp) return;
    std::memset(tp, 0, sizeof(*tp));
    tp->epochs = 8;
    tp->batch = 512;
    tp->lr = 2e-4f;
    tp->beta1 = 0.9f;
    tp->beta2 = 0.999f;
    tp->eps = 1e-8f;
    tp->seed = 1;
    for (int k = 0; k < 6; k++) tp->weights[k] = 1.0f;
    tp->weight_decay = 0.0f;
    tp->cosine = 1;
}

/* appends the grids of every scan of `sc` to the dataset's grid store */
static int dataset_add_grids(ndt_dataset* ds, const ndt_scans* sc) {
    const size_t row = 2 * (size_t)BEV_CELLS;
    const size_t need = row * (size_t)(ds->nscans + sc->S);
    if (need > ds->scan_bev.bytes) {
        DevBuf grown;
        BUF_TRY(grown, std::max(need, 2 * ds->scan_bev.bytes));
        if (ds->nscans)
            HIP_TRY(hipMemcpyAsync(grown.p, ds->scan_bev.p, row * (size_t)ds->nscans, hipMemcpyDeviceToDevice,
                                   ds->ctx->stream));
        HIP_TRY(hipStreamSynchronize(ds->ctx->stream));
        std::swap(grown.p, ds->scan_bev.p);
        std::swap(grown.bytes, ds->scan_bev.bytes);
    }
    ndt_model tmp;
    tmp.ctx = (ndt_ctx*)ds->ctx;
    DevBuf dst; /* staging for the new scans' grids */
    int e = scan_grids(&tmp, ds->bm, sc, dst);
    if (e) return e;
    HIP_TRY(hipMemcpyAsync((char*)ds->scan_bev.p + row * (size_t)ds->nscans, dst.p, row * (size_t)sc->S,
                           hipMemcpyDeviceToDevice, ds->ctx->stream));
    HIP_TRY(hipStreamSynchronize(ds->ctx->stream));
    return 0;
}

extern "C" int ndt_dataset_add(ndt_dataset* ds, const ndt_scans* sc, const int32_t* scan_index, const double* guesses,
                               const double* targets, int64_t n) {
    if (!ds || !sc || !scan_index || !guesses || !targets) return fail(NDT_ERR_INVALID, "null argument");
    if (n <= 0) return fail(NDT_ERR_INVALID, "no samples");
    if (ds->nscans + sc->S > 0x7fffffff) return fail(NDT_ERR_UNSUPPORTED, "too many scans in one dataset");
    for (int64_t i = 0; i < n; i++)
        if (scan_index[i] < 0 || scan_index[i] >= sc->S) return fail(NDT_ERR_INVALID, "scan index out of range");
    int e = check_poses(guesses, n, "guess");
    if (!e) e = check_poses(targets, n, "target");
    if (e) return e;
    HIP_TRY(hipSetDevice(ds->ctx->device));
    e = dataset_add_grids(ds, sc);
    if (e) return e;
    std::vector<double> gl;
    guesses_local(ds->bm, guesses, n, gl);
    for (int64_t i = 0; i < n; i++) {
        ds->scan_index.push_back((int32_t)(ds->nscans + scan_index[i]));
        ds->guesses.insert(ds->guesses.end(), gl.begin() + 16 * i, gl.begin() + 16 * (i + 1));
        double p6[6];
        correction_params(guesses + 16 * i, targets + 16 * i, p6);
        for (int k = 0; k < 6; k++) ds->targets.push_back((float)(p6[k] / CORR_SCALE[k]));
    }
    ds->nscans += sc->S;
    ds->n += n;
    return 0;
}

extern "C" int ndt_dataset_create(ndt_ctx* c, const ndt_bevmap* bm, const ndt_scans* sc, const int32_t* scan_index,
                                  const double* guesses, const double* targets, int64_t n, ndt_dataset** out) {
    if (!c || !bm || !sc || !scan_index || !guesses || !targets || !out) return fail(NDT_ERR_INVALID, "null argument");
    *out = nullptr;
    ndt_dataset* ds = new ndt_dataset();
    ds->ctx = c;
    const_cast<ndt_bevmap*>(bm)->refs.fetch_add(1);
    ds->bm = bm;
    int e = ndt_dataset_add(ds, sc, scan_index, guesses, targets, n);
    if (e) {
        delete ds;
        return e;
    }
    *out = ds;
    return 0;
}

extern "C" int ndt_dataset_size(const ndt_dataset* ds, int64_t* samples, int64_t* scans) {
    if (!ds) return fail(NDT_ERR_INVALID, "null argument");
    if (samples) *samples = ds->n;
    if (scans) *scans = ds->nscans;
    return 0;
}

extern "C" int ndt_dataset_destroy(ndt_dataset* ds) {
    if (!ds) return 0;
    if (ds->ctx) (void)hipSetDevice(ds->ctx->device);
    delete ds;
    (void)hipGetLastError();
    return 0;
}

extern "C" int ndt_model_train(ndt_ctx* c, ndt_model* md, const ndt_dataset* dsc, const ndt_train_params* tp,
                               float* epoch_loss) {
    if (!c || !md || !dsc || !tp) return fail(NDT_ERR_INVALID, "null argument");
    ndt_dataset* ds = const_cast<ndt_dataset*>(dsc);
    const int B = tp->batch;
    int e = check_batch(md, B);
    if (e) return e;
    if (md->f0 != NDT_FEATURES) return fail(NDT_ERR_INVALID, "training on features needs f0 = 65536");
    if (tp->epochs <= 0 || tp->epochs > 64) return fail(NDT_ERR_INVALID, "epochs must be in 1..64");
    if (ds->n < B) return fail(NDT_ERR_INVALID, "dataset smaller than one batch");
    HIP_TRY(hipSetDevice(c->device));
    if ((e = ensure_workspace(md, B))) return e;
    hipStream_t st = c->stream;
    HIP_TRY(hipMemcpyAsync(md->weights.p, tp->weights, 4 * NOUT, hipMemcpyHostToDevice, st));
    const int64_t n = ds->n, nb = n / B;
    BUF_TRY(ds->d_idx, 4 * (size_t)n);
    BUF_TRY(ds->d_guess, 8 * 16 * (size_t)n);
    BUF_TRY(ds->d_target, 4 * 6 * (size_t)n);
    std::vector<int32_t> idx((size_t)n);
    std::vector<double> gs((size_t)(16 * n));
    std::vector<float> tg((size_t)(6 * n));
    std::vector<int64_t> order((size_t)n);
    Rng rng(tp->seed);
    const int64_t total_steps = (int64_t)tp->epochs * nb;
    int64_t step = 0;
    for (int ep = 0; ep < tp->epochs; ep++) {
        for (int64_t i = 0; i < n; i++) order[(size_t)i] = i;
        for (int64_t i = n - 1; i > 0; i--) { /* Fisher-Yates */
            const int64_t j = (int64_t)(rng.next() % (uint64_t)(i + 1));
            std::swap(order[(size_t)i], order[(size_t)j]);
        }
        for (int64_t i = 0; i < n; i++) {
            const int64_t o = order[(size_t)i];
            idx[(size_t)i] = ds->scan_index[(size_t)o];
            std::memcpy(&gs[(size_t)(16 * i)], &ds->guesses[(size_t)(16 * o)], 16 * sizeof(double));
            std::memcpy(&tg[(size_t)(6 * i)], &ds->targets[(size_t)(6 * o)], 6 * sizeof(float));
        }
        HIP_TRY(hipMemcpyAsync(ds->d_idx.p, idx.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(ds->d_guess.p, gs.data(), 8 * gs.size(), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(ds->d_target.p, tg.data(), 4 * tg.size(), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemsetAsync(md->lossacc.p, 0, 4, st));
        for (int64_t k = 0; k < nb; k++) {
            if ((e = batch_features(md, ds->bm, ds->scan_bev.as<uint16_t>(), ds->d_idx.as<int32_t>() + k * B,
                                    ds->d_guess.as<double>() + 16 * k * B, B)))
                return e;
            HIP_TRY(hipMemcpyAsync(md->target.p, ds->d_target.as<float>() + 6 * k * B, 4 * 6 * (size_t)B,
                                   hipMemcpyDeviceToDevice, st));
            if ((e = forward_dev(md, B, true))) return e;
            if ((e = backward_dev(md, B))) return e;
            float lr = tp->lr;
            if (tp->cosine) {
                const double lo = tp->lr / 20.0;
                lr = (float)(lo + (tp->lr - lo) * 0.5 * (1.0 + std::cos(3.141592653589793 * (double)step / (double)total_steps)));
            }
            step++;
            if ((e = adam_dev(md, lr, tp->beta1, tp->beta2, tp->eps, tp->weight_decay))) return e;
            LAUNCH_TRY(launch_sum_ordered(md->lossb.as<float>(), B, md->lossacc.as<float>(), st));
        }
        float acc = 0.f;
        HIP_TRY(hipMemcpyAsync(&acc, md->lossacc.p, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (epoch_loss) epoch_loss[ep] = acc / (float)(nb * B);
    }
    return 0;
}

/* ------------------------------------------------------------------ prediction */

static int predict_rows(int S) { return ((S + 127) / 128) * 128; }

/* Sets up the prediction inputs of a batch: workspace for its feature rows, the
 * scans' offsets and counts, the rows' scan indices (rows past S repeat scan 0). */
static int predict_prepare(ndt_model* md, const ndt_scans* sc) {
    if (md->f0 != NDT_FEATURES) return fail(NDT_ERR_INVALID, "prediction from features needs f0 = 65536");
    const int S = sc->S, Bp = predict_rows(S);
    int e = ensure_workspace(md, Bp);
    if (e) return e;
    hipStream_t st = md->ctx->stream;
    BUF_TRY(md->sbev, 2 * (size_t)S * BEV_CELLS);
    BUF_TRY(md->off, 8 * (size_t)S);
    BUF_TRY(md->cnt, 4 * (size_t)S);
    BUF_TRY(md->idx, 4 * (size_t)Bp);
    BUF_TRY(md->gl, 8 * 16 * (size_t)Bp);
    BUF_TRY(md->gw, 8 * 16 * (size_t)S);
    BUF_TRY(md->refined, 8 * 16 * (size_t)S);
    BUF_TRY(md->greg, 8 * 16 * (size_t)S);
    std::vector<int32_t> cnt((size_t)S), idx((size_t)Bp);
    for (int32_t k = 0; k < S; k++) cnt[(size_t)k] = (int32_t)sc->counts[(size_t)k];
    for (int b = 0; b < Bp; b++) idx[(size_t)b] = b < S ? b : 0;
    HIP_TRY(hipMemcpyAsync(md->off.p, sc->offsets.data(), 8 * (size_t)S, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(md->cnt.p, cnt.data(), 4 * (size_t)S, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(md->idx.p, idx.data(), 4 * (size_t)Bp, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

/* The prediction of a prepared batch from the world guesses in md->gw: scan grids,
 * map grids at the guesses, the forward pass, refined guesses into md->refined.
 * Kernels only, no synchronisation: a graph may record it. */
static int predict_enqueue(ndt_model* md, const ndt_bevmap* bm, const ndt_scans* sc, int keep_rotation) {
    hipStream_t st = md->ctx->stream;
    const int S = sc->S, Bp = predict_rows(S);
    LAUNCH_TRY(launch_bev_scan(sc->pts.as<float4>(), md->off.as<int64_t>(), md->cnt.as<int32_t>(), S,
                               bm->lut.as<uint16_t>(), md->sbev.as<uint16_t>(), st));
    LAUNCH_TRY(launch_guesses_local(md->gw.as<double>(), S, Bp, bm->origin[0], bm->origin[1], bm->origin[2],
                                    md->gl.as<double>(), st));
    /* registration: the map-local guesses of the real rows in place, the world
     * guesses into greg (gw keeps the given guesses, plain NDT's starts) */
    const bool reg = reg_on(md->reg);
    if (reg)
        LAUNCH_TRY(launch_bev_register(bm->dev, md->sbev.as<uint16_t>(), sc->pts.as<float4>(), md->off.as<int64_t>(),
                                       md->cnt.as<int32_t>(), md->idx.as<int32_t>(), md->gl.as<double>(),
                                       md->gw.as<double>(), S, reg_dev(md->reg), md->gl.as<double>(),
                                       md->greg.as<double>(), nullptr, st));
    int e = batch_features(md, bm, md->sbev.as<uint16_t>(), md->idx.as<int32_t>(), md->gl.as<double>(), Bp);
    if (e) return e;
    if ((e = forward_dev(md, Bp, false))) return e;
    CorrScale scale;
    for (int k = 0; k < NOUT; k++) scale.v[k] = CORR_SCALE[k];
    LAUNCH_TRY(launch_refine(md->Y.as<float>(), reg ? md->greg.as<double>() : md->gw.as<double>(), S, scale,
                             keep_rotation, md->refined.as<double>(), st));
    return 0;
}

/* Refined guesses of a batch into host memory (the scan grids are made at every
 * prediction: a batch's scans may be predicted for once only). */
static int predict_dev(ndt_ctx* c, ndt_model* md, const ndt_bevmap* bm, const ndt_scans* sc, const double* guesses,
                       double* out_guesses, int keep_rotation = 0) {
    int e = predict_prepare(md, sc);
    if (e) return e;
    const int S = sc->S;
    HIP_TRY(hipMemcpyAsync(md->gw.p, guesses, 8 * 16 * (size_t)S, hipMemcpyHostToDevice, c->stream));
    if ((e = predict_enqueue(md, bm, sc, keep_rotation))) return e;
    HIP_TRY(hipMemcpyAsync(out_guesses, md->refined.p, 8 * 16 * (size_t)S, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

/* The workspace a captured learned step refers to. */
static std::vector<const void*> plan_bufs(const ndt_model* md) {
    return {md->X.p, md->P1.p, md->H1.p, md->P2.p, md->H2.p, md->Y.p, md->wb1.p, md->wb2.p, md->wb3.p,
            md->w[NDT_B1].p, md->w[NDT_B2].p, md->w[NDT_B3].p, md->sbev.p, md->off.p, md->cnt.p, md->idx.p,
            md->gl.p, md->gw.p, md->refined.p, md->greg.p, (const void*)(intptr_t)md->splits1,
            (const void*)(intptr_t)md->splits2};
}

/* Captures the learned step with the fallback policy for a batch (once per scans,
 * grid, BEV map, parameters and rotation choice, and again when the model's
 * workspace moved): the guesses in from pinned memory, the prediction, the refined
 * guesses, the set-up kernels and the work-queue launch with refined starts, the
 * results out to pinned memory.  NDT_ERR_UNSUPPORTED when the batch needs the
 * round scheduler. */
static int learned_plan_prepare(ndt_ctx* c, const ndt_grid* g, const ndt_bevmap* bm, ndt_model* md,
                                const ndt_scans* sc, const ndt_params* p, int keep) {
    auto& lp = md->plan;
    if (lp.valid && lp.scans == sc->serial && lp.grid == g->serial && lp.bm == bm->serial && lp.device == c->device &&
        lp.keep == keep && ndt::same_params(lp.params, *p) && lp.wait == ndt::wait_limit() && lp.bufs == plan_bufs(md) &&
        same_reg(lp.reg, md->reg))
        return 0;
    lp.release();
    int e = predict_prepare(md, sc);
    if (e) return e;
    if ((e = ndt::align_sub_prepare(c, g, sc, p, true, lp.sub))) return e;
    HIP_TRY(hipHostMalloc((void**)&lp.h_guess, 8 * 16 * (size_t)sc->S, hipHostMallocDefault));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 2; k++) {
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        if (k == 0) {
            hipError_t he = hipMemcpyAsync(md->gw.p, lp.h_guess, 8 * 16 * (size_t)sc->S, hipMemcpyHostToDevice,
                                           c->stream);
            e = he == hipSuccess ? 0 : fail(NDT_ERR_HIP, std::string("guess upload: ") + hipGetErrorString(he));
            if (!e) e = predict_enqueue(md, bm, sc, keep);
        } else {
            e = ndt::align_sub_enqueue(c, g, sc, lp.sub, md->gw.as<double>(), md->refined.as<double>());
        }
        hipGraph_t graph = nullptr;
        const hipError_t he = hipStreamEndCapture(c->stream, &graph);
        if (e) {
            if (graph) (void)hipGraphDestroy(graph);
            return e;
        }
        if (he != hipSuccess) return fail(NDT_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(he));
        lp.graph[k] = graph;
        HIP_TRY(hipGraphInstantiate(&lp.exec[k], lp.graph[k], nullptr, nullptr, 0));
    }
    lp.valid = true;
    lp.scans = sc->serial;
    lp.grid = g->serial;
    lp.bm = bm->serial;
    lp.device = c->device;
    lp.keep = keep;
    lp.params = *p;
    lp.wait = ndt::wait_limit();
    lp.bufs = plan_bufs(md);
    lp.reg = md->reg;
    return 0;
}

/* Replays the captured learned step with new guesses. */
static int learned_plan_run(ndt_ctx* c, ndt_model* md, const ndt_scans* sc, const double* guesses,
                            ndt_result* results, int32_t* used_learned) {
    auto& lp = md->plan;
    std::memcpy(lp.h_guess, guesses, 8 * 16 * (size_t)sc->S);
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HIP_TRY(hipGraphLaunch(lp.exec[0], c->stream));
    HIP_TRY(hipEventRecord(c->ev2, c->stream));
    HIP_TRY(hipGraphLaunch(lp.exec[1], c->stream));
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.f, align_ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    HIP_TRY(hipEventElapsedTime(&align_ms, c->ev2, c->ev1));
    std::vector<int32_t> used((size_t)sc->S);
    int e = ndt::align_sub_results(c, sc, lp.sub, results, used.data());
    c->last_ms = align_ms;
    c->last_step_ms = ms;
    c->last_launches = 1;
    if (used_learned)
        for (int32_t s = 0; s < sc->S; s++) used_learned[s] = used[(size_t)s];
    return e;
}

extern "C" int ndt_model_predict(ndt_ctx* c, ndt_model* md, const ndt_bevmap* bm, const ndt_scans* sc,
                                 const double* guesses, double* out_guesses) {
    if (!c || !md || !bm || !sc || !guesses || !out_guesses) return fail(NDT_ERR_INVALID, "null argument");
    if (int e = check_poses(guesses, sc->S, "guess")) return e;
    HIP_TRY(hipSetDevice(c->device));
    return predict_dev(c, md, bm, sc, guesses, out_guesses);
}

extern "C" int ndt_align_scans_learned(ndt_ctx* c, const ndt_grid* g, const ndt_bevmap* bm, ndt_model* md,
                                       const ndt_scans* sc, const double* guesses, const ndt_params* p, int32_t policy,
                                       ndt_result* results, int32_t* used_learned) {
    if (!c || !g || !bm || !md || !sc || !guesses || !p || !results) return fail(NDT_ERR_INVALID, "null argument");
    const bool keep_rotation = (policy & NDT_LEARNED_KEEP_ROTATION) != 0;
    policy &= ~NDT_LEARNED_KEEP_ROTATION;
    if (policy != NDT_LEARNED_RACE && policy != NDT_LEARNED_FALLBACK) return fail(NDT_ERR_INVALID, "unknown policy");
    if (int e = check_poses(guesses, sc->S, "guess")) return e;
    HIP_TRY(hipSetDevice(c->device));
    const int S = sc->S;
    double ms_total = 0.0, ms_align = 0.0;
    int32_t launches = 0;
    int64_t pairs = 0;
    auto account = [&]() {
        ms_total += c->last_ms;
        ms_align += c->last_ms;
        launches += c->last_launches;
        pairs += c->last_pairs;
    };
    if (policy == NDT_LEARNED_FALLBACK) {
        /* the whole step as one captured graph */
        int e = learned_plan_prepare(c, g, bm, md, sc, p, keep_rotation ? 1 : 0);
        if (e == 0) return learned_plan_run(c, md, sc, guesses, results, used_learned);
        if (e != NDT_ERR_UNSUPPORTED) return e;
    }
    /* refined guesses */
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    std::vector<double> refined((size_t)(16 * S));
    int e = predict_dev(c, md, bm, sc, guesses, refined.data(), keep_rotation ? 1 : 0);
    if (e) return e;
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    float fms = 0.f;
    HIP_TRY(hipEventElapsedTime(&fms, c->ev0, c->ev1));
    ms_total += fms;
    launches += 1;
    if (policy == NDT_LEARNED_RACE) {
        /* both starts of every scan in one batch: plain first, learned second */
        std::vector<int32_t> idx((size_t)(2 * S));
        std::vector<double> gs((size_t)(32 * S));
        for (int s = 0; s < S; s++) {
            idx[(size_t)s] = s;
            idx[(size_t)(S + s)] = s;
        }
        std::memcpy(gs.data(), guesses, 16 * sizeof(double) * (size_t)S);
        std::memcpy(gs.data() + 16 * (size_t)S, refined.data(), 16 * sizeof(double) * (size_t)S);
        std::vector<ndt_result> r2((size_t)(2 * S));
        if ((e = align_scan_subset(c, g, sc, idx.data(), 2 * S, gs.data(), p, r2.data()))) return e;
        account();
        for (int s = 0; s < S; s++) {
            const ndt_result& a = r2[(size_t)s];
            const ndt_result& b = r2[(size_t)(S + s)];
            const bool learned = (b.converged && !a.converged) || (b.converged == a.converged && b.score > a.score);
            results[s] = learned ? b : a;
            results[s].passes = a.passes + b.passes;
            if (used_learned) used_learned[s] = learned ? 1 : 0;
        }
    } else {
        /* one launch: per scan a probe pass at the given guess, NDT from the refined
         * guess where it scores at least as high, and plain NDT from the given guess
         * otherwise or where the refined run ended below the given guess's score */
        std::vector<int32_t> idx((size_t)S), used((size_t)S);
        for (int s = 0; s < S; s++) idx[(size_t)s] = s;
        if ((e = align_scan_subset(c, g, sc, idx.data(), S, guesses, p, results, refined.data(), used.data())))
            return e;
        account();
        if (used_learned)
            for (int s = 0; s < S; s++) used_learned[s] = used[(size_t)s];
    }
    c->last_ms = ms_align;
    c->last_step_ms = ms_total;
    c->last_launches = launches;
    c->last_pairs = pairs;
    return 0;
}
