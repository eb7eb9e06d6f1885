/*
 * ndt_kernels.hip - gfx950 kernels of the NDT scan-matching engine.
 *
 *  grid build   bbox reduce, voxel keys, stable LSD radix sort, run heads/starts,
 *               per-voxel float64 statistics (ndt: two-pass covariance, Jacobi
 *               eigenvalue clamp, inverse; compat: the reference's single-pass
 *               normalisation), neighbourhood dilation, hash build
 *  align        fused per-point kernel: transform, hashed cell lookup, candidate
 *               list scan, score / gradient / Hessian per pair, wave64 shuffle +
 *               LDS reductions; persistent cooperative kernel that keeps every
 *               scan's Newton / More-Thuente state on the device
 *  compat       the reference's per-pair arithmetic in float64 on the GPU
 *
 * The file is compiled with -ffp-contract=off so that float64 control code and
 * the compat path keep the reference's rounding; the float32 hot loop opts into
 * contraction explicitly.
 */
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

#include <mutex>

#include "ndt_internal.h"

namespace ndt {

constexpr int WAVE = 64;

/* ------------------------------------------------------------------ helpers */

__device__ inline uint64_t ld_agent_u64(const void* p) {
    return __hip_atomic_load((uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_agent_u64(void* p, uint64_t v) {
    __hip_atomic_store((uint64_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double ld_agent_f64(const double* p) { return __longlong_as_double((long long)ld_agent_u64(p)); }
__device__ inline void st_agent_f64(double* p, double v) { st_agent_u64(p, (uint64_t)__double_as_longlong(v)); }
__device__ inline int32_t ld_agent_i32(const int32_t* p) {
    return __hip_atomic_load((int32_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

static inline int grid_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

/* ------------------------------------------------------------------ bbox */

/* Bounding box of the finite points as order-preserving keys (mm[0..2] minima,
 * mm[3..5] maxima, merged into what mm holds).  Each thread keeps four loads in
 * flight; a workgroup reduces over its waves in LDS and merges once, so the
 * atomics on the six words number six per workgroup (at most 256 workgroups). */
constexpr int BBOX_BLOCK = 512;

__global__ __launch_bounds__(BBOX_BLOCK) void k_bbox(const float4* __restrict__ pts, int64_t n,
                                                      uint32_t* __restrict__ mm) {
    __shared__ uint32_t red[BBOX_BLOCK / WAVE][6];
    uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
    const int64_t stride = (int64_t)gridDim.x * BBOX_BLOCK;
    for (int64_t i0 = blockIdx.x * (int64_t)BBOX_BLOCK + threadIdx.x; i0 < n; i0 += 4 * stride) {
        float4 p[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t i = i0 + u * stride;
            p[u] = i < n ? pts[i] : make_float4(NAN, NAN, NAN, NAN);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (!isfinite(p[u].x) || !isfinite(p[u].y) || !isfinite(p[u].z)) continue;
            const uint32_t o[3] = {f2ord(p[u].x), f2ord(p[u].y), f2ord(p[u].z)};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                lo[k] = o[k] < lo[k] ? o[k] : lo[k];
                hi[k] = o[k] > hi[k] ? o[k] : hi[k];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            uint32_t a = __shfl_xor(lo[k], off, WAVE), b = __shfl_xor(hi[k], off, WAVE);
            lo[k] = a < lo[k] ? a : lo[k];
            hi[k] = b > hi[k] ? b : hi[k];
        }
    }
    const int wave = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        for (int k = 0; k < 3; k++) {
            red[wave][k] = lo[k];
            red[wave][3 + k] = hi[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        uint32_t v = red[0][k];
        for (int w = 1; w < BBOX_BLOCK / WAVE; w++) v = k < 3 ? min(v, red[w][k]) : max(v, red[w][k]);
        if (k < 3) atomicMin(&mm[k], v);
        else atomicMax(&mm[k], v);
    }
}

int launch_bbox(const float4* pts, int64_t n, uint32_t* mm, hipStream_t stream) {
    int64_t blocks = (n + 4 * (int64_t)BBOX_BLOCK - 1) / (4 * (int64_t)BBOX_BLOCK);
    if (blocks > 256) blocks = 256;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_bbox, dim3((unsigned)blocks), dim3(BBOX_BLOCK), 0, stream, pts, n, mm);
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------ keys */

struct KeyLayout {
    int64_t o[3];
    int b[3];
};

__global__ void k_keys_ndt(const float4* __restrict__ pts, int64_t n, double res, KeyLayout L,
                           uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pts[i];
    const uint64_t sentinel = 1ull << (L.b[0] + L.b[1] + L.b[2]);
    uint64_t key = sentinel;
    if (isfinite(p.x) && isfinite(p.y) && isfinite(p.z)) {
        int64_t c[3] = {(int64_t)floor((double)p.x / res) - L.o[0], (int64_t)floor((double)p.y / res) - L.o[1],
                        (int64_t)floor((double)p.z / res) - L.o[2]};
        bool ok = true;
        for (int k = 0; k < 3; k++) ok = ok && c[k] >= 0 && c[k] < (1ll << L.b[k]);
        if (ok) key = ((uint64_t)c[2] << (L.b[0] + L.b[1])) | ((uint64_t)c[1] << L.b[0]) | (uint64_t)c[0];
    }
    keys[i] = key;
    vals[i] = (uint32_t)i;
}

int launch_keys_ndt(const float4* pts, int64_t n, double resolution, const int64_t* origin_cell, const int* bits,
                    uint64_t* keys, uint32_t* vals, hipStream_t stream) {
    KeyLayout L;
    for (int k = 0; k < 3; k++) {
        L.o[k] = origin_cell[k];
        L.b[k] = bits[k];
    }
    hipLaunchKernelGGL(k_keys_ndt, dim3(grid_for(n, 256)), dim3(256), 0, stream, pts, n, resolution, L, keys, vals);
    return (int)hipGetLastError();
}

struct CompatLayout {
    float minv[3];
    int dims[3];
    float res;
};

/* linearizeCoord / linearizeAddr, src/CPU/ndt_mapping/kernel.cpp:231-244 */
__device__ inline int compat_linearize(const float* minv, const int* dims, float res, float x, float y, float z) {
    int ix = (int)((x - minv[0]) / res);
    int iy = (int)((y - minv[1]) / res);
    int iz = (int)((z - minv[2]) / res);
    return ix + dims[0] * (iy + dims[1] * iz);
}

__global__ void k_keys_compat(const float4* __restrict__ pts, int64_t n, CompatLayout L, uint64_t* __restrict__ keys,
                              uint32_t* __restrict__ vals) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pts[i];
    keys[i] = (uint64_t)(uint32_t)compat_linearize(L.minv, L.dims, L.res, p.x, p.y, p.z);
    vals[i] = (uint32_t)i;
}

int launch_keys_compat(const float4* pts, int64_t n, const float* minv, const int* dims, float resolution,
                       uint64_t* keys, uint32_t* vals, hipStream_t stream) {
    CompatLayout L;
    for (int k = 0; k < 3; k++) {
        L.minv[k] = minv[k];
        L.dims[k] = dims[k];
    }
    L.res = resolution;
    hipLaunchKernelGGL(k_keys_compat, dim3(grid_for(n, 256)), dim3(256), 0, stream, pts, n, L, keys, vals);
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------ radix sort
 * Stable LSD radix sort, 8-bit digits.  Per pass: per-tile digit histograms
 * (digit-major), one exclusive scan, and a stable scatter in which each tile
 * ranks its keys round by round with wave ballots. */

constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;

__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                      uint32_t* __restrict__ hist, int nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; r++) {
        int64_t idx = base + r * RS_BLOCK + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                         uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                         int64_t n, int shift, const uint32_t* __restrict__ offs,
                                                         int nblocks) {
    __shared__ uint32_t base[256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[RS_BLOCK / WAVE][256];
    const int tid = threadIdx.x;
    const int wave = tid / WAVE, lane = tid % WAVE;
    base[tid] = offs[(int64_t)tid * nblocks + blockIdx.x];
    run[tid] = 0;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int64_t tbase = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; r++) {
        int64_t idx = tbase + r * RS_BLOCK + tid;
        bool valid = idx < n;
        uint64_t key = valid ? kin[idx] : 0;
        uint32_t val = valid ? vin[idx] : 0;
        uint32_t d = valid ? (uint32_t)((key >> shift) & 255u) : 256u;
        uint64_t peers = __ballot(1);
#pragma unroll
        for (int b = 0; b < 9; b++) {
            uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        for (int w = 0; w < RS_BLOCK / WAVE; w++) wcnt[w][tid] = 0;
        __syncthreads();
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t acc = run[tid];
            for (int w = 0; w < RS_BLOCK / WAVE; w++) {
                uint32_t c = wcnt[w][tid];
                wcnt[w][tid] = acc;
                acc += c;
            }
            run[tid] = acc;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + wcnt[wave][d] + rank;
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
    }
}

/* ------------------------------------------------------------------ scan */

constexpr int SC_BLOCK = 256;
constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = SC_BLOCK * SC_ITEMS;

/* exclusive block scan of one value per thread; returns the block total */
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* excl, uint32_t* lds_waves) {
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        uint32_t t = __shfl_up(incl, off, WAVE);
        if (lane >= off) incl += t;
    }
    if (lane == WAVE - 1) lds_waves[wave] = incl;
    __syncthreads();
    uint32_t wave_off = 0, total = 0;
    for (int w = 0; w < (int)(blockDim.x / WAVE); w++) {
        uint32_t c = lds_waves[w];
        if (w < wave) wave_off += c;
        total += c;
    }
    __syncthreads();
    *excl = wave_off + incl - v;
    return total;
}

__global__ __launch_bounds__(SC_BLOCK) void k_scan_reduce(const uint32_t* __restrict__ in, int64_t n,
                                                          uint32_t* __restrict__ bsum) {
    __shared__ uint32_t lw[SC_BLOCK / WAVE];
    int64_t base = (int64_t)blockIdx.x * SC_TILE;
    uint32_t s = 0;
    for (int r = 0; r < SC_ITEMS; r++) {
        int64_t idx = base + r * SC_BLOCK + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    uint32_t ex;
    uint32_t total = block_exclusive_scan(s, &ex, lw);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* __restrict__ bsum, int64_t nb, uint32_t* __restrict__ total) {
    __shared__ uint32_t lw[1024 / WAVE];
    const int64_t per = (nb + blockDim.x - 1) / blockDim.x;
    const int64_t b = threadIdx.x * per;
    uint32_t s = 0;
    for (int64_t i = b; i < b + per && i < nb; i++) s += bsum[i];
    uint32_t ex;
    uint32_t tot = block_exclusive_scan(s, &ex, lw);
    uint32_t run = ex;
    for (int64_t i = b; i < b + per && i < nb; i++) {
        uint32_t v = bsum[i];
        bsum[i] = run;
        run += v;
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

__global__ __launch_bounds__(SC_BLOCK) void k_scan_down(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                        int64_t n, const uint32_t* __restrict__ bsum) {
    __shared__ uint32_t lw[SC_BLOCK / WAVE];
    /* blocked arrangement: thread t owns items [t*16, t*16+16) of the tile */
    int64_t base = (int64_t)blockIdx.x * SC_TILE + (int64_t)threadIdx.x * SC_ITEMS;
    uint32_t v[SC_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t ex;
    block_exclusive_scan(s, &ex, lw);
    uint32_t run = ex + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

size_t scan_scratch_bytes(int64_t n) {
    int64_t nb = (n + SC_TILE - 1) / SC_TILE;
    return (size_t)(nb + 1) * sizeof(uint32_t);
}

/* Exclusive scan of a short array in one workgroup (in may equal out): thread t
 * scans its contiguous run of ceil(n / 1024) values after a block scan of the
 * run sums.  One launch instead of three for the radix sort's histograms. */
constexpr int64_t SCAN_SMALL_MAX = 16384;
__global__ __launch_bounds__(1024) void k_scan_small(const uint32_t* in, uint32_t* out, int64_t n,
                                                     uint32_t* __restrict__ total) {
    __shared__ uint32_t lw[1024 / WAVE];
    const int64_t per = (n + blockDim.x - 1) / blockDim.x;
    const int64_t b = threadIdx.x * per;
    uint32_t s = 0;
    for (int64_t i = b; i < b + per && i < n; i++) s += in[i];
    uint32_t ex;
    const uint32_t tot = block_exclusive_scan(s, &ex, lw);
    uint32_t run = ex;
    for (int64_t i = b; i < b + per && i < n; i++) {
        const uint32_t v = in[i];
        out[i] = run;
        run += v;
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev, void* scratch,
                       hipStream_t stream) {
    if (n <= 0) return 0;
    if (n <= SCAN_SMALL_MAX) {
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, stream, in, out, n, total_dev);
        return (int)hipGetLastError();
    }
    int64_t nb = (n + SC_TILE - 1) / SC_TILE;
    uint32_t* bsum = (uint32_t*)scratch;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SC_BLOCK), 0, stream, in, n, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, stream, bsum, nb, total_dev);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb), dim3(SC_BLOCK), 0, stream, in, out, n, bsum);
    return (int)hipGetLastError();
}

size_t radix_sort_scratch_bytes(int64_t n) {
    int64_t nb = (n + RS_TILE - 1) / RS_TILE;
    int64_t h = 256 * nb;
    return (size_t)h * sizeof(uint32_t) + scan_scratch_bytes(h) + 256;
}

int radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_tmp, uint32_t* vals_tmp, int64_t n, int nbits,
                     void* scratch, hipStream_t stream) {
    if (n <= 1 || nbits <= 0) return 0;
    int nb = (int)((n + RS_TILE - 1) / RS_TILE);
    int64_t h = 256ll * nb;
    uint32_t* hist = (uint32_t*)scratch;
    void* sc = (void*)(((uintptr_t)(hist + h) + 255) & ~(uintptr_t)255);
    uint64_t *ki = keys, *ko = keys_tmp;
    uint32_t *vi = vals, *vo = vals_tmp;
    int passes = (nbits + 7) / 8;
    for (int p = 0; p < passes; p++) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_BLOCK), 0, stream, ki, n, shift, hist, nb);
        int e = exclusive_scan_u32(hist, hist, h, nullptr, sc, stream);
        if (e) return e;
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_BLOCK), 0, stream, ki, vi, ko, vo, n, shift, hist, nb);
        uint64_t* tk = ki; ki = ko; ko = tk;
        uint32_t* tv = vi; vi = vo; vo = tv;
    }
    if (ki != keys) {
        hipError_t e = hipMemcpyAsync(keys, ki, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(vals, vi, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess) return (int)e;
    }
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------ runs */

__global__ void k_run_heads(const uint64_t* __restrict__ keys, int64_t n, int shift, uint64_t limit,
                            uint32_t* __restrict__ flags) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = keys[i];
    uint32_t f = 0;
    if (k < limit) f = (i == 0) || ((keys[i - 1] >> shift) != (k >> shift));
    flags[i] = f;
}

int launch_run_heads(const uint64_t* keys, int64_t n, int shift, uint64_t limit, uint32_t* flags,
                     hipStream_t stream) {
    hipLaunchKernelGGL(k_run_heads, dim3(grid_for(n, 256)), dim3(256), 0, stream, keys, n, shift, limit, flags);
    return (int)hipGetLastError();
}

__global__ void k_run_starts(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ ranks, int64_t n,
                             uint32_t* __restrict__ starts) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flags[i]) starts[ranks[i]] = (uint32_t)i;
}

int launch_run_starts(const uint32_t* flags, const uint32_t* ranks, int64_t n, uint32_t* starts, hipStream_t stream) {
    hipLaunchKernelGGL(k_run_starts, dim3(grid_for(n, 256)), dim3(256), 0, stream, flags, ranks, n, starts);
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------ voxel statistics */

/* cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix: ascending
 * eigenvalues, eigenvectors in the columns of V */
__device__ void sym3_eig(const double* Ain, double* ev, double* V) {
    double a[3][3], v[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            a[r][c] = Ain[r * 3 + c];
            v[r][c] = (r == c) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 50; sweep++) {
        double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (off <= 1e-36 * diag || off == 0.0) break;
        for (int pq = 0; pq < 3; pq++) {
            int p = (pq == 2) ? 1 : 0;
            int q = (pq == 0) ? 1 : 2;
            if (a[p][q] == 0.0) continue;
            double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
            double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            double c = 1.0 / sqrt(t * t + 1.0);
            double s = t * c;
            for (int k = 0; k < 3; k++) {
                double akp = a[k][p], akq = a[k][q];
                a[k][p] = c * akp - s * akq;
                a[k][q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; k++) {
                double apk = a[p][k], aqk = a[q][k];
                a[p][k] = c * apk - s * aqk;
                a[q][k] = s * apk + c * aqk;
            }
            for (int k = 0; k < 3; k++) {
                double vkp = v[k][p], vkq = v[k][q];
                v[k][p] = c * vkp - s * vkq;
                v[k][q] = s * vkp + c * vkq;
            }
        }
    }
    int idx[3] = {0, 1, 2};
    double d[3] = {a[0][0], a[1][1], a[2][2]};
    for (int i = 0; i < 2; i++)
        for (int j = i + 1; j < 3; j++)
            if (d[idx[j]] < d[idx[i]]) {
                int t = idx[i];
                idx[i] = idx[j];
                idx[j] = t;
            }
    for (int i = 0; i < 3; i++) {
        ev[i] = d[idx[i]];
        for (int r = 0; r < 3; r++) V[r * 3 + i] = v[r][idx[i]];
    }
}

__global__ void k_voxel_stats_ndt(const float4* __restrict__ pts, const uint32_t* __restrict__ sidx,
                                  const uint64_t* __restrict__ skeys, const uint32_t* __restrict__ seg_start,
                                  int64_t nseg, int64_t nvalid_pts, int min_points, double eig_ratio,
                                  double* __restrict__ mean_out, double* __restrict__ icov_out,
                                  double* __restrict__ cov_out, int32_t* __restrict__ count,
                                  uint32_t* __restrict__ valid, uint64_t* __restrict__ vkey) {
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    int64_t b = seg_start[s];
    int64_t e = (s + 1 < nseg) ? (int64_t)seg_start[s + 1] : nvalid_pts;
    int64_t n = e - b;
    double sum[3] = {0, 0, 0};
    for (int64_t k = b; k < e; k++) {
        float4 p = pts[sidx[k]];
        sum[0] += p.x;
        sum[1] += p.y;
        sum[2] += p.z;
    }
    double mean[3] = {sum[0] / (double)n, sum[1] / (double)n, sum[2] / (double)n};
    double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t k = b; k < e; k++) {
        float4 p = pts[sidx[k]];
        double d[3] = {p.x - mean[0], p.y - mean[1], p.z - mean[2]};
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) C[r * 3 + c] += d[r] * d[c];
    }
    uint32_t ok = 0;
    double icov[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, cov[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (n >= min_points && n >= 2) {
        for (int k = 0; k < 9; k++) C[k] /= (double)(n - 1);
        double ev[3], V[9];
        sym3_eig(C, ev, V);
        if (ev[2] > 0 && isfinite(ev[2])) {
            double lo = eig_ratio * ev[2];
            for (int k = 0; k < 3; k++)
                if (ev[k] < lo) ev[k] = lo;
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    double a = 0, bb = 0;
                    for (int k = 0; k < 3; k++) {
                        a += V[r * 3 + k] * ev[k] * V[c * 3 + k];
                        bb += V[r * 3 + k] * (1.0 / ev[k]) * V[c * 3 + k];
                    }
                    cov[r * 3 + c] = a;
                    icov[r * 3 + c] = bb;
                }
            ok = 1;
        }
    }
    for (int c = 0; c < 3; c++) mean_out[3 * s + c] = mean[c];
    for (int k = 0; k < 9; k++) {
        icov_out[9 * s + k] = icov[k];
        cov_out[9 * s + k] = cov[k];
    }
    count[s] = (int32_t)n;
    valid[s] = ok;
    vkey[s] = skeys[b];
}

int launch_voxel_stats_ndt(const float4* pts, const uint32_t* sorted_idx, const uint64_t* sorted_keys,
                           const uint32_t* seg_start, int64_t nseg, int64_t nvalid_pts, int min_points,
                           double min_eig_ratio, const int* bits, double* mean, double* icov, double* cov,
                           int32_t* count, uint32_t* valid, uint64_t* vkey, hipStream_t stream) {
    (void)bits;
    hipLaunchKernelGGL(k_voxel_stats_ndt, dim3(grid_for(nseg, 128)), dim3(128), 0, stream, pts, sorted_idx,
                       sorted_keys, seg_start, nseg, nvalid_pts, min_points, min_eig_ratio, mean, icov, cov, count,
                       valid, vkey);
    return (int)hipGetLastError();
}

__global__ void k_compact_voxels(const uint32_t* __restrict__ valid, const uint32_t* __restrict__ rank, int64_t nvox,
                                 const double* __restrict__ mean, const double* __restrict__ icov,
                                 const uint64_t* __restrict__ vkey, double* __restrict__ mean_out,
                                 double* __restrict__ icov_out, float4* __restrict__ icov_f,
                                 uint64_t* __restrict__ vkey_out) {
    int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= nvox || !valid[v]) return;
    int64_t r = rank[v];
    for (int c = 0; c < 3; c++) mean_out[3 * r + c] = mean[3 * v + c];
    const double* C = icov + 9 * v;
    double u[6] = {C[0], C[1], C[2], C[4], C[5], C[8]};
    for (int k = 0; k < 6; k++) icov_out[6 * r + k] = u[k];
    icov_f[2 * r] = make_float4((float)u[0], (float)u[1], (float)u[2], (float)u[3]);
    icov_f[2 * r + 1] = make_float4((float)u[4], (float)u[5], 0.0f, 0.0f);
    vkey_out[r] = vkey[v];
}

int launch_compact_voxels(const uint32_t* valid, const uint32_t* rank, int64_t nvox, const double* mean,
                          const double* icov, const uint64_t* vkey, double* mean_out, double* icov_out, float4* icov_f,
                          uint64_t* vkey_out, hipStream_t stream) {
    hipLaunchKernelGGL(k_compact_voxels, dim3(grid_for(nvox, 256)), dim3(256), 0, stream, valid, rank, nvox, mean,
                       icov, vkey, mean_out, icov_out, icov_f, vkey_out);
    return (int)hipGetLastError();
}

struct Bits3 {
    int b[3];
};

/* each valid voxel announces itself to the 27 query cells that see it; the
 * composite key orders a query cell's list by probe index (z, then y, then x) */
__global__ void k_dilate_emit(const uint64_t* __restrict__ vkey, int64_t nvalid, Bits3 B, uint64_t* __restrict__ keys,
                              uint32_t* __restrict__ vals) {
    int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= nvalid) return;
    uint64_t k = vkey[r];
    int64_t c[3];
    c[0] = (int64_t)(k & ((1ull << B.b[0]) - 1));
    c[1] = (int64_t)((k >> B.b[0]) & ((1ull << B.b[1]) - 1));
    c[2] = (int64_t)(k >> (B.b[0] + B.b[1]));
    int o = 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++, o++) {
                /* voxel = query + (dx, dy, dz) */
                uint64_t qx = (uint64_t)(c[0] - dx), qy = (uint64_t)(c[1] - dy), qz = (uint64_t)(c[2] - dz);
                uint64_t q = (qz << (B.b[0] + B.b[1])) | (qy << B.b[0]) | qx;
                keys[27 * r + o] = (q << 5) | (uint64_t)o;
                vals[27 * r + o] = (uint32_t)r;
            }
}

int launch_dilate_emit(const uint64_t* vkey, int64_t nvalid, const int* bits, uint64_t* keys, uint32_t* vals,
                       hipStream_t stream) {
    Bits3 B;
    for (int k = 0; k < 3; k++) B.b[k] = bits[k];
    hipLaunchKernelGGL(k_dilate_emit, dim3(grid_for(nvalid, 256)), dim3(256), 0, stream, vkey, nvalid, B, keys, vals);
    return (int)hipGetLastError();
}

__global__ void k_hash_clear(HashSlot* slots, int64_t cap) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= cap) return;
    slots[i].key = EMPTY_KEY;
    slots[i].begin = 0;
    slots[i].count = 0;
}

int launch_hash_clear(HashSlot* slots, int64_t cap, hipStream_t stream) {
    hipLaunchKernelGGL(k_hash_clear, dim3(grid_for(cap, 256)), dim3(256), 0, stream, slots, cap);
    return (int)hipGetLastError();
}

__global__ void k_hash_insert(const uint64_t* __restrict__ skeys, const uint32_t* __restrict__ starts, int64_t nq,
                              int64_t nentries, int shift, HashSlot* __restrict__ slots, int log2cap) {
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint32_t b = starts[q];
    uint32_t e = (q + 1 < nq) ? starts[q + 1] : (uint32_t)nentries;
    uint64_t key = skeys[b] >> shift;
    uint64_t mask = (1ull << log2cap) - 1;
    uint64_t h = hash_slot(key, log2cap);
    for (uint64_t probe = 0; probe <= mask; probe++) {
        unsigned long long prev = atomicCAS((unsigned long long*)&slots[h].key, (unsigned long long)EMPTY_KEY,
                                            (unsigned long long)key);
        if (prev == EMPTY_KEY) {
            slots[h].begin = b;
            slots[h].count = e - b;
            return;
        }
        h = (h + 1) & mask;
    }
}

/* Dense query-cell index: entry (c2 * D1 + c1) * D0 + c0 = (begin, count) of the
 * candidate run of query cell (c0, c1, c2); cells without candidates stay (0, 0). */
__global__ void k_dense_fill(const uint64_t* __restrict__ skeys, const uint32_t* __restrict__ starts, int64_t nq,
                             int64_t nentries, Bits3 B, int32_t d0, int32_t d1, uint2* __restrict__ dense) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const uint32_t b = starts[q];
    const uint32_t e = (q + 1 < nq) ? starts[q + 1] : (uint32_t)nentries;
    const uint64_t key = skeys[b] >> 5;
    const int64_t c0 = (int64_t)(key & ((1ull << B.b[0]) - 1));
    const int64_t c1 = (int64_t)((key >> B.b[0]) & ((1ull << B.b[1]) - 1));
    const int64_t c2 = (int64_t)(key >> (B.b[0] + B.b[1]));
    dense[(c2 * d1 + c1) * d0 + c0] = make_uint2(b, e - b);
}

int launch_dense_fill(const uint64_t* sorted_keys, const uint32_t* starts, int64_t nq, int64_t nentries,
                      const int* bits, const int32_t* dims, uint2* dense, hipStream_t stream) {
    Bits3 B;
    for (int k = 0; k < 3; k++) B.b[k] = bits[k];
    hipLaunchKernelGGL(k_dense_fill, dim3(grid_for(nq, 256)), dim3(256), 0, stream, sorted_keys, starts, nq, nentries, B,
                       dims[0], dims[1], dense);
    return (int)hipGetLastError();
}

int launch_hash_insert(const uint64_t* sorted_keys, const uint32_t* starts, int64_t nq, int64_t nentries, int shift,
                       HashSlot* slots, int log2cap, hipStream_t stream) {
    hipLaunchKernelGGL(k_hash_insert, dim3(grid_for(nq, 256)), dim3(256), 0, stream, sorted_keys, starts, nq, nentries,
                       shift, slots, log2cap);
    return (int)hipGetLastError();
}

__global__ void k_hash_insert_ids(const uint64_t* __restrict__ vkey, int64_t nvox, HashSlot* __restrict__ slots,
                                  int log2cap) {
    int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    uint64_t key = vkey[v];
    uint64_t mask = (1ull << log2cap) - 1;
    uint64_t h = hash_slot(key, log2cap);
    for (uint64_t probe = 0; probe <= mask; probe++) {
        unsigned long long prev = atomicCAS((unsigned long long*)&slots[h].key, (unsigned long long)EMPTY_KEY,
                                            (unsigned long long)key);
        if (prev == EMPTY_KEY) {
            slots[h].begin = (uint32_t)v;
            slots[h].count = 1;
            return;
        }
        h = (h + 1) & mask;
    }
}

int launch_hash_insert_ids(const uint64_t* vkey, int64_t nvox, HashSlot* slots, int log2cap, hipStream_t stream) {
    hipLaunchKernelGGL(k_hash_insert_ids, dim3(grid_for(nvox, 256)), dim3(256), 0, stream, vkey, nvox, slots, log2cap);
    return (int)hipGetLastError();
}

/* Candidate i of the sorted dilation keys (query cell << 5 | probe): the voxel's
 * mean relative to the lower corner of the query cell in float32 (values in
 * [-res, 2 res), kept to about 1e-7 m at any map extent) with the voxel id, and
 * the voxel's statistics in float32. */
__global__ void k_cand_fill(const uint64_t* __restrict__ skeys, const uint32_t* __restrict__ svals, int64_t n,
                            const double* __restrict__ mean, const double* __restrict__ icov, Bits3 B, int64_t o0,
                            int64_t o1, int64_t o2, double res, float4* __restrict__ head, float4* __restrict__ s32) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = svals[i];
    const uint64_t q = skeys[i] >> 5;
    const int64_t c0 = (int64_t)(q & ((1ull << B.b[0]) - 1)) + o0;
    const int64_t c1 = (int64_t)((q >> B.b[0]) & ((1ull << B.b[1]) - 1)) + o1;
    const int64_t c2 = (int64_t)(q >> (B.b[0] + B.b[1])) + o2;
    const double* m = mean + 3 * (int64_t)v;
    const double* C = icov + 6 * (int64_t)v;
    head[i] = make_float4((float)(m[0] - (double)c0 * res), (float)(m[1] - (double)c1 * res),
                          (float)(m[2] - (double)c2 * res), __uint_as_float(v));
    s32[2 * i] = make_float4((float)C[0], (float)C[1], (float)C[2], (float)C[3]);
    s32[2 * i + 1] = make_float4((float)C[4], (float)C[5], 0.0f, 0.0f);
}

int launch_cand_fill(const uint64_t* sorted_keys, const uint32_t* sorted_vals, int64_t nentries, const double* mean,
                     const double* icov, const int* bits, const int64_t* origin_cell, double res, float4* head,
                     float4* s32, hipStream_t stream) {
    Bits3 B;
    for (int k = 0; k < 3; k++) B.b[k] = bits[k];
    hipLaunchKernelGGL(k_cand_fill, dim3(grid_for(nentries, 256)), dim3(256), 0, stream, sorted_keys, sorted_vals,
                       nentries, mean, icov, B, origin_cell[0], origin_cell[1], origin_cell[2], res, head, s32);
    return (int)hipGetLastError();
}

/* Float64 statistics of valid voxel v: mean and inverse covariance (upper
 * triangle), 80 bytes. */
__global__ void k_voxel_s64(const double* __restrict__ mean, const double* __restrict__ icov, int64_t n,
                            double2* __restrict__ vs64) {
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    const double* m = mean + 3 * v;
    const double* C = icov + 6 * v;
    double2* o = vs64 + VS64_STRIDE * v;
    o[0] = make_double2(m[0], m[1]);
    o[1] = make_double2(m[2], C[0]);
    o[2] = make_double2(C[1], C[2]);
    o[3] = make_double2(C[3], C[4]);
    o[4] = make_double2(C[5], 0.0);
}

int launch_voxel_s64(const double* mean, const double* icov, int64_t nvalid, double2* vs64, hipStream_t stream) {
    if (nvalid <= 0) return 0;
    hipLaunchKernelGGL(k_voxel_s64, dim3(grid_for(nvalid, 256)), dim3(256), 0, stream, mean, icov, nvalid, vs64);
    return (int)hipGetLastError();
}

/* initCompute's scatter and normalisation (src/CPU/ndt_mapping/kernel.cpp:1335-1372)
 * per occupied cell, summing its points in index order as the reference does. */
__global__ void k_voxel_stats_compat(const float4* __restrict__ pts, const uint32_t* __restrict__ sidx,
                                     const uint64_t* __restrict__ skeys, const uint32_t* __restrict__ seg_start,
                                     int64_t nseg, int64_t npts, double ncells, double* __restrict__ mean_out,
                                     double* __restrict__ icov_out, int32_t* __restrict__ count,
                                     uint64_t* __restrict__ vkey) {
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    int64_t b = seg_start[s];
    int64_t e = (s + 1 < nseg) ? (int64_t)seg_start[s + 1] : npts;
    double m[3] = {0, 0, 0};
    double S[3][3] = {{0, 0, 1.0}, {0, 1.0, 0}, {1.0, 0, 0}};
    int n = 0;
    for (int64_t k = b; k < e; k++) {
        float4 p4 = pts[sidx[k]];
        float p[3] = {p4.x, p4.y, p4.z};
        m[0] += p[0];
        m[1] += p[1];
        m[2] += p[2];
        n++;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) S[r][c] += (double)(p[r] * p[c]);
    }
    double ps[3] = {m[0], m[1], m[2]};
    m[0] /= n;
    m[1] /= n;
    m[2] /= n;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            S[r][c] = (S[r][c] - 2 * (ps[r] * m[c])) / ncells + m[r] * m[c];
            S[r][c] *= (ncells - 1.0) / n;
        }
    /* invertMatrix, kernel.cpp:1277-1300 */
    double t[3][3];
    double det = S[0][0] * (S[2][2] * S[1][1] - S[2][1] * S[1][2]) - S[1][0] * (S[2][2] * S[0][1] - S[2][1] * S[0][2]) +
                 S[2][0] * (S[1][2] * S[0][1] - S[1][1] * S[0][2]);
    double invDet = 1.0 / det;
    t[0][0] = S[2][2] * S[1][1] - S[2][1] * S[1][2];
    t[0][1] = -(S[2][2] * S[0][1] - S[2][1] * S[0][2]);
    t[0][2] = S[1][2] * S[0][1] - S[1][1] * S[0][2];
    t[1][0] = -(S[2][2] * S[0][1] - S[2][0] * S[1][2]);
    t[1][1] = S[2][2] * S[0][0] - S[2][1] * S[0][2];
    t[1][2] = -(S[1][2] * S[0][0] - S[1][0] * S[0][2]);
    t[2][0] = S[2][1] * S[1][0] - S[2][0] * S[1][1];
    t[2][1] = -(S[2][1] * S[0][0] - S[2][0] * S[0][1]);
    t[2][2] = S[1][1] * S[0][0] - S[1][0] * S[0][1];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) icov_out[9 * s + 3 * r + c] = t[r][c] * invDet;
    for (int c = 0; c < 3; c++) mean_out[3 * s + c] = m[c];
    count[s] = n;
    vkey[s] = skeys[b];
}

int launch_voxel_stats_compat(const float4* pts, const uint32_t* sorted_idx, const uint64_t* sorted_keys,
                              const uint32_t* seg_start, int64_t nseg, int64_t npts, double ncells, double* mean,
                              double* icov, int32_t* count, uint64_t* vkey, hipStream_t stream) {
    hipLaunchKernelGGL(k_voxel_stats_compat, dim3(grid_for(nseg, 128)), dim3(128), 0, stream, pts, sorted_idx,
                       sorted_keys, seg_start, nseg, npts, ncells, mean, icov, count, vkey);
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------ ndt per-point math */

/* Per-pass values shared by every point of a scan.  The angular derivatives of
 * T(p) x = Rx(roll) Ry(pitch) Rz(yaw) x + t are written with the rotation axes
 * a3 = e_x, a4 = Rx e_y, a5 = Rx Ry e_z and the rotated point y = R x:
 *   dT/dtheta_k = a_k x y,   d2T/dtheta_i dtheta_j = a_i x (a_j x y)  (i <= j),
 * which is eq. 6.19 / 6.21 (the j_ang / h_ang tables, kernel.cpp:707-825) in
 * closed form.  The tables' small-angle rule (cos = 1, sin = 0 below 1e-4) is kept
 * by evaluating the axes and Ra (used for y in the derivatives) at the angles
 * with small components set to zero; R and t (the transform) stay exact. */
template <typename T>
struct PassParams {
    double R[9], t[3];         /* the transform, float64 in every precision */
    T Ra[9], a4[3], a5[3];
    T res, res2, d1, d2, nhd2; /* nhd2 = -d2 / 2 */
    /* res2 (1 -+ 2^-48): outside [r2lo, r2hi) the float64 range test
     * sqrt(|d|^2) < res is decided by |d|^2 alone (set by load_uniform) */
    T r2lo, r2hi;
    /* parameter k of the 32-value layout written by pass_params_f64 */
    __device__ void set(int k, double v) {
        if (k < 9) R[k] = v;
        else if (k < 12) t[k - 9] = v;
        else if (k < 21) Ra[k - 12] = (T)v;
        else if (k < 24) a4[k - 21] = (T)v;
        else if (k < 27) a5[k - 24] = (T)v;
        else if (k == 27) res = (T)v;
        else if (k == 28) res2 = (T)v;
        else if (k == 29) d1 = (T)v;
        else if (k == 30) d2 = (T)v;
        else nhd2 = (T)v;
    }
};

/* The 32 pass parameters of pose x from the cosines c[0..2] and sines sn[0..2] of
 * its angles x[3..5]; the small-angle rule's zeroed angles have cosine 1 and sine 0
 * exactly. */
__device__ inline void pass_params_from_trig(const double* x, const double* c, const double* sn,
                                             const AlignParams& prm, double res, double* out /* 32 */) {
    double ca[3], sa[3];
    double M[16];
    pose_matrix_cs(x, c, sn, M);
    double xa[6];
    for (int k = 0; k < 6; k++) xa[k] = (k >= 3 && fabs(x[k]) < 10e-5) ? 0.0 : x[k];
    for (int k = 0; k < 3; k++) {
        const bool zero = fabs(x[3 + k]) < 10e-5;
        ca[k] = zero ? 1.0 : c[k];
        sa[k] = zero ? 0.0 : sn[k];
    }
    double Ma[16];
    pose_matrix_cs(xa, ca, sa, Ma);
    for (int r = 0; r < 3; r++) {
        for (int cc = 0; cc < 3; cc++) {
            out[r * 3 + cc] = M[r * 4 + cc];
            out[12 + r * 3 + cc] = Ma[r * 4 + cc];
        }
        out[9 + r] = M[r * 4 + 3];
    }
    const double c3 = ca[0], s3 = sa[0], c4 = ca[1], s4 = sa[1];
    out[21] = 0.0;
    out[22] = c3;
    out[23] = s3;
    out[24] = s4;
    out[25] = -s3 * c4;
    out[26] = c3 * c4;
    out[27] = res;
    out[28] = res * res;
    out[29] = prm.d1;
    out[30] = prm.d2;
    out[31] = -prm.d2 / 2.0;
}

__device__ __attribute__((noinline)) void pass_params_f64(const double* x, const AlignParams& prm, double res,
                                                          double* out /* 32 */) {
    double c[3], sn[3];
    for (int k = 0; k < 3; k++) {
        c[k] = cos(x[3 + k]);
        sn[k] = sin(x[3 + k]);
    }
    pass_params_from_trig(x, c, sn, prm, res, out);
}

/* The pass parameters of a scan's trial pose st.x_t into st.pp, by the first wave
 * of the workgroup after the controller (every lane of wave 0 must call this):
 * lanes 0-2 take the angles' cosines and lanes 3-5 their sines at once, lane 0
 * the rest.  Between two passes of a scan this is on its critical path, where
 * six transcendentals one after the other on one lane took about 2 us. */
__device__ __attribute__((noinline)) void wave_pass_params(ScanState& st, const AlignParams& prm,
                                                           double* trig /* LDS, 6 */) {
    const int lane = threadIdx.x % WAVE;
    if (lane < 3) trig[lane] = cos(st.x_t[3 + lane]);
    else if (lane < 6) trig[lane] = sin(st.x_t[lane]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane == 0) pass_params_from_trig(st.x_t, trig, trig + 3, prm, prm.resolution_d, st.pp);
}

template <typename T>
__device__ inline void make_pass_params(const double* x, const AlignParams& prm, double res, PassParams<T>& pp) {
    double v[32];
    pass_params_f64(x, prm, res, v);
    for (int k = 0; k < 32; k++) pp.set(k, v[k]);
}

__device__ inline float uniform_value(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ inline double uniform_value(double v) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(uint32_t)lo);
}

/* copies the LDS pass parameters into wave-uniform (scalar) registers */
template <typename T>
__device__ inline PassParams<T> load_uniform(const PassParams<T>& lds) {
    PassParams<T> u;
#pragma unroll
    for (int k = 0; k < 9; k++) u.R[k] = uniform_value(lds.R[k]);
#pragma unroll
    for (int k = 0; k < 3; k++) u.t[k] = uniform_value(lds.t[k]);
#pragma unroll
    for (int k = 0; k < 9; k++) u.Ra[k] = uniform_value(lds.Ra[k]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        u.a4[k] = uniform_value(lds.a4[k]);
        u.a5[k] = uniform_value(lds.a5[k]);
    }
    u.res = uniform_value(lds.res);
    u.res2 = uniform_value(lds.res2);
    u.d1 = uniform_value(lds.d1);
    u.d2 = uniform_value(lds.d2);
    u.nhd2 = uniform_value(lds.nhd2);
    u.r2lo = u.res2 * (T)(1.0 - 0x1p-48);
    u.r2hi = u.res2 * (T)(1.0 + 0x1p-48);
    return u;
}

/* Accumulator of the 28 sums (score, gradient, Hessian upper triangle). */
template <typename T, int SUM>
struct Acc;

template <typename T>
struct Acc<T, NDT_SUM_FP64> {
    double s[NACC];
    __device__ void zero() {
        for (int k = 0; k < NACC; k++) s[k] = 0.0;
    }
    __device__ void add(int k, T v) { s[k] += (double)v; }
    /* s[k] += a * b with one rounding */
    __device__ void fma(int k, double a, double b) { s[k] = ::fma(a, b, s[k]); }
    __device__ double get(int k) const { return s[k]; }
};

template <typename T>
struct Acc<T, NDT_SUM_FP32> {
    float s[NACC];
    __device__ void zero() {
        for (int k = 0; k < NACC; k++) s[k] = 0.0f;
    }
    __device__ void add(int k, T v) { s[k] += (float)v; }
    __device__ double get(int k) const { return (double)s[k]; }
};

template <typename T>
struct Acc<T, NDT_SUM_COMP32> {
    float s[NACC], c[NACC];
    __device__ void zero() {
        for (int k = 0; k < NACC; k++) s[k] = c[k] = 0.0f;
    }
    __device__ void add(int k, T v) {
        /* Neumaier's compensated summation */
        float x = (float)v;
        float t = s[k] + x;
        if (fabsf(s[k]) >= fabsf(x)) c[k] += (s[k] - t) + x;
        else c[k] += (x - t) + s[k];
        s[k] = t;
    }
    __device__ double get(int k) const { return (double)s[k] + (double)c[k]; }
};

__device__ inline bool hash_find(const HashSlot* __restrict__ slots, int log2cap, uint64_t key, uint32_t& begin,
                                 uint32_t& count) {
    uint64_t mask = (1ull << log2cap) - 1;
    uint64_t h = hash_slot(key, log2cap);
    for (uint64_t probe = 0; probe <= mask; probe++) {
        HashSlot s = slots[h];
        if (s.key == key) {
            begin = s.begin;
            count = s.count;
            return true;
        }
        if (s.key == EMPTY_KEY) return false;
        h = (h + 1) & mask;
    }
    return false;
}

/* Index lookup of one scan point: the transformed point (float64), its position
 * relative to its cell's lower corner (float32, where the candidate means are
 * stored) and its cell's candidate run. */
struct PointLook {
    double x0, x1, x2;
    float r0, r1, r2;
    uint32_t begin, count;
};

/* x / resolution in float64, as the oracle computes a cell coordinate.  With a
 * power-of-two resolution the reciprocal is exact and one multiplication gives the
 * same value as the division. */
__device__ inline double cell_quot(const NdtGridDev& g, const double x) {
    return g.res_pow2 ? x * g.inv_resolution_d : x / g.resolution_d;
}

/* Transform, cell and index lookup; false when the point has no candidates. */
__device__ inline bool point_lookup(const NdtGridDev& g, const float4 p, const double* R, const double* t,
                                    PointLook& L) {
#pragma clang fp contract(fast)
    const double X0 = p.x, X1 = p.y, X2 = p.z;
    const double xd0 = R[0] * X0 + R[1] * X1 + R[2] * X2 + t[0];
    const double xd1 = R[3] * X0 + R[4] * X1 + R[5] * X2 + t[1];
    const double xd2 = R[6] * X0 + R[7] * X1 + R[8] * X2 + t[2];
    if (!(fabs(xd0) < 1e15) || !(fabs(xd1) < 1e15) || !(fabs(xd2) < 1e15)) return false;
    const int64_t a0 = (int64_t)floor(cell_quot(g, xd0));
    const int64_t a1 = (int64_t)floor(cell_quot(g, xd1));
    const int64_t a2 = (int64_t)floor(cell_quot(g, xd2));
    const int64_t c0 = a0 - g.origin_cell[0], c1 = a1 - g.origin_cell[1], c2 = a2 - g.origin_cell[2];
    uint32_t begin, count;
    if (g.dense) {
        /* direct index: neighbouring points read neighbouring entries */
        if (c0 < 0 || c1 < 0 || c2 < 0 || c0 >= g.dims[0] || c1 >= g.dims[1] || c2 >= g.dims[2]) return false;
        const uint2 e = g.dense[(c2 * g.dims[1] + c1) * g.dims[0] + c0];
        begin = e.x;
        count = e.y;
        if (!count) return false;
    } else {
        if (c0 < 0 || c1 < 0 || c2 < 0 || c0 >= (1ll << g.bits[0]) || c1 >= (1ll << g.bits[1]) ||
            c2 >= (1ll << g.bits[2]))
            return false;
        const uint64_t key = ((uint64_t)c2 << (g.bits[0] + g.bits[1])) | ((uint64_t)c1 << g.bits[0]) | (uint64_t)c0;
        if (!hash_find(g.slots, g.log2cap, key, begin, count)) return false;
    }
    L.x0 = xd0;
    L.x1 = xd1;
    L.x2 = xd2;
    L.r0 = (float)(xd0 - (double)a0 * g.resolution_d);
    L.r1 = (float)(xd1 - (double)a1 * g.resolution_d);
    L.r2 = (float)(xd2 - (double)a2 * g.resolution_d);
    L.begin = begin;
    L.count = count;
    return true;
}

/* The same lookup split in two for software pipelining over a thread's points:
 * lookup_issue computes the cell and starts the dense-index load; the entry is
 * consumed one iteration later, when its latency has passed under the previous
 * point's work.  Only for grids with a dense index. */
__device__ inline void lookup_issue(const NdtGridDev& g, const float4 p, const bool valid, const double* R,
                                    const double* t, PointLook& L, uint2& e) {
#pragma clang fp contract(fast)
    const double X0 = p.x, X1 = p.y, X2 = p.z;
    const double xd0 = R[0] * X0 + R[1] * X1 + R[2] * X2 + t[0];
    const double xd1 = R[3] * X0 + R[4] * X1 + R[5] * X2 + t[1];
    const double xd2 = R[6] * X0 + R[7] * X1 + R[8] * X2 + t[2];
    const bool finite = (fabs(xd0) < 1e15) && (fabs(xd1) < 1e15) && (fabs(xd2) < 1e15);
    /* Cell coordinates in float64: floor() of the quotient and its difference from
     * the index origin are exact integers there, so no 64-bit integer arithmetic
     * is needed; inside the index box they convert exactly to int32, and the box
     * (at most 2^28 cells) keeps the linear index in int32. */
    const double a0 = floor(cell_quot(g, xd0)), a1 = floor(cell_quot(g, xd1)), a2 = floor(cell_quot(g, xd2));
    const double c0 = a0 - g.origin_cell_d[0], c1 = a1 - g.origin_cell_d[1], c2 = a2 - g.origin_cell_d[2];
    L.x0 = xd0;
    L.x1 = xd1;
    L.x2 = xd2;
    L.r0 = (float)(xd0 - a0 * g.resolution_d);
    L.r1 = (float)(xd1 - a1 * g.resolution_d);
    L.r2 = (float)(xd2 - a2 * g.resolution_d);
    const bool ok = valid && finite && c0 >= 0.0 && c1 >= 0.0 && c2 >= 0.0 && c0 < (double)g.dims[0] &&
                    c1 < (double)g.dims[1] && c2 < (double)g.dims[2];
    const int i0 = ok ? (int)c0 : 0, i1 = ok ? (int)c1 : 0, i2 = ok ? (int)c2 : 0;
    /* the load is unconditional (entry 0 stands in for points without a cell), so
     * that its wait is counted exactly where the entry is used */
    const uint2 v = g.dense[(i2 * g.dims[1] + i1) * g.dims[0] + i0];
    e = ok ? v : make_uint2(0u, 0u);
}

/* range prefilter of candidate k of a point: bit k of the result */
__device__ inline uint32_t prefilter_bit(const NdtGridDev& g, const PointLook& L, uint32_t k, float lim) {
    const float4 cd = g.head[L.begin + k];
    const float e0 = L.r0 - cd.x, e1 = L.r1 - cd.y, e2 = L.r2 - cd.z;
    return (e0 * e0 + e1 * e1 + e2 * e2 < lim ? 1u : 0u) << k;
}

/* The float prefilter keeps every candidate the exact test can keep: in fp64 mode
 * it carries a margin and the exact test decides. */
template <typename T>
__device__ inline float prefilter_limit(const PassParams<T>& u) {
    return (sizeof(T) == 4) ? (float)u.res2 : (float)u.res2 * 1.0001f + 1e-5f;
}

/* Score / gradient / Hessian terms (Magnusson 2009 eqs. 6.9-6.13) of one
 * (point, candidate) pair, added when the candidate's mean lies within one
 * resolution of the point: xd is the transformed point (float64), r its position
 * relative to its cell's lower corner (float32), p the point in the sensor frame.
 * idx is the candidate entry (float32 points) or the voxel id (float64 points).
 * Branch-free: a pair that is not valid, not in range or fails the value guard
 * still runs every instruction, on zeroed inputs and with zero weights, and adds
 * +-0 to every sum (the sums are unchanged), so the accumulators need no copies at
 * control-flow joins.  Returns 1 when the pair contributes. */
template <typename T, int SUM, bool HESS = true>
__device__ inline int pair_add(const NdtGridDev& g, const bool valid, const double xd0, const double xd1,
                               const double xd2, const float r0, const float r1, const float r2, const float4 p,
                               uint32_t idx, const PassParams<T>& u, Acc<T, SUM>& acc) {
#pragma clang fp contract(fast)
    idx = valid ? idx : 0u; /* entry / voxel 0 always exists */
    T d0, d1, d2;
    T c00, c01, c02, c11, c12, c22;
    bool in;
    if constexpr (sizeof(T) == 4) {
        const float4 h = g.head[idx];
        const float4* r = g.s32 + 2 * (int64_t)idx;
        const float4 a = r[0], b = r[1];
        d0 = r0 - h.x;
        d1 = r1 - h.y;
        d2 = r2 - h.z;
        in = valid && (d0 * d0 + d1 * d1 + d2 * d2 < u.res2);
        c00 = a.x; c01 = a.y; c02 = a.z; c11 = a.w; c12 = b.x; c22 = b.y;
    } else {
        const double2* r = g.vs64 + VS64_STRIDE * (int64_t)idx;
        const double2 m01 = r[0], m2c0 = r[1], c12v = r[2], c34 = r[3], c5p = r[4];
        d0 = xd0 - m01.x;
        d1 = xd1 - m01.y;
        d2 = xd2 - m2c0.x;
        /* sqrt(dd) < res as the oracle tests it: a correctly rounded square root
         * is below res for every dd < r2lo and not below it for dd >= r2hi, so it
         * is computed only for the rare dd in between */
        const double dd = d0 * d0 + d1 * d1 + d2 * d2;
        in = dd < u.r2lo;
        if (!(dd < u.r2lo) && dd < u.r2hi) in = sqrt(dd) < u.res;
        in = valid && in;
        c00 = m2c0.y; c01 = c12v.x; c02 = c12v.y; c11 = c34.x; c12 = c34.y; c22 = c5p.x;
    }
    d0 = in ? d0 : (T)0;
    d1 = in ? d1 : (T)0;
    d2 = in ? d2 : (T)0;
    /* point Jacobian columns J3 = e_x x ya, J4 = a4 x ya, J5 = a5 x ya */
    const T x0 = in ? (T)p.x : (T)0, x1 = in ? (T)p.y : (T)0, x2 = in ? (T)p.z : (T)0;
    const T y0 = u.Ra[0] * x0 + u.Ra[1] * x1 + u.Ra[2] * x2;
    const T y1 = u.Ra[3] * x0 + u.Ra[4] * x1 + u.Ra[5] * x2;
    const T y2 = u.Ra[6] * x0 + u.Ra[7] * x1 + u.Ra[8] * x2;
    const T J3y = -y2, J3z = y1;
    const T J4x = u.a4[1] * y2 - u.a4[2] * y1, J4y = u.a4[2] * y0 - u.a4[0] * y2, J4z = u.a4[0] * y1 - u.a4[1] * y0;
    const T J5x = u.a5[1] * y2 - u.a5[2] * y1, J5y = u.a5[2] * y0 - u.a5[0] * y2, J5z = u.a5[0] * y1 - u.a5[1] * y0;
    /* q = C x' */
    const T q0 = c00 * d0 + c01 * d1 + c02 * d2;
    const T q1 = c01 * d0 + c11 * d1 + c12 * d2;
    const T q2 = c02 * d0 + c12 * d1 + c22 * d2;
    const T xCx = d0 * q0 + d1 * q1 + d2 * q2;
    T e;
    if constexpr (sizeof(T) == 4) e = expf(u.nhd2 * xCx);
    else e = exp(u.nhd2 * xCx);
    const T e2 = u.d2 * e;
    /* invalid-value guard, kernel.cpp:456-457 */
    const bool ok = in && !(e2 > (T)1 || e2 < (T)0 || e2 != e2);
    const T f = ok ? e2 * u.d1 : (T)0;
    const T m = -u.d2;
    /* x'^T C J_k */
    const T u3 = q1 * J3y + q2 * J3z;
    const T u4 = q0 * J4x + q1 * J4y + q2 * J4z;
    const T u5 = q0 * J5x + q1 * J5y + q2 * J5z;
    /* C J_k */
    const T k30 = c01 * J3y + c02 * J3z, k31 = c11 * J3y + c12 * J3z, k32 = c12 * J3y + c22 * J3z;
    const T k40 = c00 * J4x + c01 * J4y + c02 * J4z;
    const T k41 = c01 * J4x + c11 * J4y + c12 * J4z;
    const T k42 = c02 * J4x + c12 * J4y + c22 * J4z;
    const T k50 = c00 * J5x + c01 * J5y + c02 * J5z;
    const T k51 = c01 * J5x + c11 * J5y + c12 * J5z;
    const T k52 = c02 * J5x + c12 * J5y + c22 * J5z;
    /* J_b^T C J_a */
    const T j33 = J3y * k31 + J3z * k32;
    const T j34 = J4x * k30 + J4y * k31 + J4z * k32;
    const T j35 = J5x * k30 + J5y * k31 + J5z * k32;
    const T j44 = J4x * k40 + J4y * k41 + J4z * k42;
    const T j45 = J5x * k40 + J5y * k41 + J5z * k42;
    const T j55 = J5x * k50 + J5y * k51 + J5z * k52;
    /* x'^T C H_ab = J_b . (q x a_a) */
    const T h33 = q2 * J3y - q1 * J3z;
    const T h34 = q2 * J4y - q1 * J4z;
    const T h35 = q2 * J5y - q1 * J5z;
    const T qa40 = q1 * u.a4[2] - q2 * u.a4[1], qa41 = q2 * u.a4[0] - q0 * u.a4[2], qa42 = q0 * u.a4[1] - q1 * u.a4[0];
    const T qa50 = q1 * u.a5[2] - q2 * u.a5[1], qa51 = q2 * u.a5[0] - q0 * u.a5[2], qa52 = q0 * u.a5[1] - q1 * u.a5[0];
    const T h44 = qa40 * J4x + qa41 * J4y + qa42 * J4z;
    const T h45 = qa40 * J5x + qa41 * J5y + qa42 * J5z;
    const T h55 = qa50 * J5x + qa51 * J5y + qa52 * J5z;
    acc.add(0, ok ? -u.d1 * e : (T)0);
    if constexpr (sizeof(T) == 8 && SUM == NDT_SUM_FP64) {
        /* H_ij += f m u_i u_j + f K_ij and g_i += f u_i as fused multiply-adds
         * into the sums: (f m u_i) u_j and f K_ij, one rounding each */
        acc.fma(1, f, q0);
        acc.fma(2, f, q1);
        acc.fma(3, f, q2);
        acc.fma(4, f, u3);
        acc.fma(5, f, u4);
        acc.fma(6, f, u5);
        if constexpr (!HESS) return ok ? 1 : 0;
        const T fm = f * m;
        const T w0 = fm * q0, w1 = fm * q1, w2 = fm * q2, w3 = fm * u3, w4 = fm * u4, w5 = fm * u5;
        acc.fma(7, w0, q0);   acc.fma(7, f, c00);
        acc.fma(8, w0, q1);   acc.fma(8, f, c01);
        acc.fma(9, w0, q2);   acc.fma(9, f, c02);
        acc.fma(10, w0, u3);  acc.fma(10, f, k30);
        acc.fma(11, w0, u4);  acc.fma(11, f, k40);
        acc.fma(12, w0, u5);  acc.fma(12, f, k50);
        acc.fma(13, w1, q1);  acc.fma(13, f, c11);
        acc.fma(14, w1, q2);  acc.fma(14, f, c12);
        acc.fma(15, w1, u3);  acc.fma(15, f, k31);
        acc.fma(16, w1, u4);  acc.fma(16, f, k41);
        acc.fma(17, w1, u5);  acc.fma(17, f, k51);
        acc.fma(18, w2, q2);  acc.fma(18, f, c22);
        acc.fma(19, w2, u3);  acc.fma(19, f, k32);
        acc.fma(20, w2, u4);  acc.fma(20, f, k42);
        acc.fma(21, w2, u5);  acc.fma(21, f, k52);
        acc.fma(22, w3, u3);  acc.fma(22, f, h33 + j33);
        acc.fma(23, w3, u4);  acc.fma(23, f, h34 + j34);
        acc.fma(24, w3, u5);  acc.fma(24, f, h35 + j35);
        acc.fma(25, w4, u4);  acc.fma(25, f, h44 + j44);
        acc.fma(26, w4, u5);  acc.fma(26, f, h45 + j45);
        acc.fma(27, w5, u5);  acc.fma(27, f, h55 + j55);
    } else {
        acc.add(1, f * q0);
        acc.add(2, f * q1);
        acc.add(3, f * q2);
        acc.add(4, f * u3);
        acc.add(5, f * u4);
        acc.add(6, f * u5);
        if constexpr (!HESS) return ok ? 1 : 0;
        const T mu0 = m * q0, mu1 = m * q1, mu2 = m * q2, mu3 = m * u3, mu4 = m * u4;
        acc.add(7, f * (mu0 * q0 + c00));
        acc.add(8, f * (mu0 * q1 + c01));
        acc.add(9, f * (mu0 * q2 + c02));
        acc.add(10, f * (mu0 * u3 + k30));
        acc.add(11, f * (mu0 * u4 + k40));
        acc.add(12, f * (mu0 * u5 + k50));
        acc.add(13, f * (mu1 * q1 + c11));
        acc.add(14, f * (mu1 * q2 + c12));
        acc.add(15, f * (mu1 * u3 + k31));
        acc.add(16, f * (mu1 * u4 + k41));
        acc.add(17, f * (mu1 * u5 + k51));
        acc.add(18, f * (mu2 * q2 + c22));
        acc.add(19, f * (mu2 * u3 + k32));
        acc.add(20, f * (mu2 * u4 + k42));
        acc.add(21, f * (mu2 * u5 + k52));
        acc.add(22, f * (mu3 * u3 + h33 + j33));
        acc.add(23, f * (mu3 * u4 + h34 + j34));
        acc.add(24, f * (mu3 * u5 + h35 + j35));
        acc.add(25, f * (mu4 * u4 + h44 + j44));
        acc.add(26, f * (mu4 * u5 + h45 + j45));
        acc.add(27, f * (m * u5 * u5 + h55 + j55));
    }
    return ok ? 1 : 0;
}

/* The candidates of one looked-up point selected by mask.  Returns the number of
 * pairs. */
template <typename T, int SUM, bool HESS = true>
__device__ inline int point_pairs(const NdtGridDev& g, const float4 p, const PointLook& L, uint32_t mask,
                                  const PassParams<T>& u, Acc<T, SUM>& acc) {
    int pairs = 0;
    while (mask) {
        const uint32_t k = (uint32_t)__builtin_ctz(mask);
        mask &= mask - 1;
        const uint32_t idx = (sizeof(T) == 4) ? L.begin + k : __float_as_uint(g.head[L.begin + k].w);
        pairs += pair_add<T, SUM, HESS>(g, true, L.x0, L.x1, L.x2, L.r0, L.r1, L.r2, p, idx, u, acc);
    }
    return pairs;
}

/* A query cell has at most 27 candidates (its 3 x 3 x 3 neighbourhood). */
constexpr int MAX_CANDIDATES = 27;

/* One wave's pair table in LDS: candidate j of lane l that passes the prefilter
 * at pick[64 j + l] (its pair_add index; lanes interleaved so that the lanes of
 * one store hit distinct banks; 32 columns hold the 27 candidates plus the
 * branch-free prefilter's writes past the last one), the wave's pairs in
 * evaluation order as positions in pick, and the pairs carried over to the next
 * round (fewer than 64: the point's transformed position, or its cell-relative
 * position with float32 points, its sensor-frame position and the pair_add
 * index). */
constexpr int PICK_ROW = 32;
struct WavePairs {
    uint32_t pick[WAVE * PICK_ROW];
    uint16_t order[WAVE * MAX_CANDIDATES];
    double cx[3][WAVE];
    float cp[3][WAVE];
    uint32_t cidx[WAVE];
};

__device__ inline uint32_t bperm_u32(int src_byte, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src_byte, (int)v);
}
__device__ inline float bperm_f32(int src_byte, float v) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src_byte, __float_as_int(v)));
}
__device__ inline double bperm_f64(int src_byte, double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = bperm_u32(src_byte, (uint32_t)b), hi = bperm_u32(src_byte, (uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

/* candidate heads a lane loads at once in the prefilter: all eight loads are in
 * flight together, so most points wait for one round trip (a query cell has
 * 3.7 candidates on average on the 1 M-point map) */
#ifndef HEAD_BATCH
#define HEAD_BATCH 8
#endif

/* Inputs of pair n of the wave's current round.  Pairs 0 .. carry-1 are those
 * carried over from the previous round (slot n of the carry table, read only when
 * `read_carry`), pairs carry .. all-1 are this round's, in order, with their
 * point data taken from the lane that looked the point up. */
template <typename T>
__device__ inline void pair_source(const WavePairs& wp, const uint32_t n, const uint32_t carry, const uint32_t all,
                                   const bool read_carry, const float4 p, const PointLook& L, bool& valid,
                                   double& x0, double& x1, double& x2, float& r0, float& r1, float& r2, float4& ps,
                                   uint32_t& idx) {
    const uint32_t lane = threadIdx.x % WAVE;
    const bool fresh = n >= carry && n < all;
    const uint32_t d = fresh ? (uint32_t)wp.order[n - carry] : lane;
    idx = wp.pick[d];
    const int src = (int)(d % WAVE) << 2;
    ps = make_float4(bperm_f32(src, p.x), bperm_f32(src, p.y), bperm_f32(src, p.z), 0.f);
    x0 = x1 = x2 = 0.0;
    r0 = r1 = r2 = 0.f;
    if constexpr (sizeof(T) == 4) {
        r0 = bperm_f32(src, L.r0);
        r1 = bperm_f32(src, L.r1);
        r2 = bperm_f32(src, L.r2);
    } else {
        x0 = bperm_f64(src, L.x0);
        x1 = bperm_f64(src, L.x1);
        x2 = bperm_f64(src, L.x2);
    }
    if (read_carry) {
        const uint32_t c = n % WAVE;
        const bool held = n < carry;
        ps.x = held ? wp.cp[0][c] : ps.x;
        ps.y = held ? wp.cp[1][c] : ps.y;
        ps.z = held ? wp.cp[2][c] : ps.z;
        idx = held ? wp.cidx[c] : idx;
        if constexpr (sizeof(T) == 4) {
            r0 = held ? (float)wp.cx[0][c] : r0;
            r1 = held ? (float)wp.cx[1][c] : r1;
            r2 = held ? (float)wp.cx[2][c] : r2;
        } else {
            x0 = held ? wp.cx[0][c] : x0;
            x1 = held ? wp.cx[1][c] : x1;
            x2 = held ? wp.cx[2][c] : x2;
        }
    }
    valid = n < all;
}

/* The pairs of one point per lane, spread evenly over the wave in full batches of
 * 64.  Lane l holds its point's lookup L and has written its cnt prefiltered
 * candidates to its column of wp.pick (cnt <= 27; 0 without a point).  The
 * round's pairs are numbered after the `carry` pairs left from the previous
 * round: lane by lane, in prefilter order within a lane.  Pair n is evaluated by
 * lane n mod 64 in the round's batch n / 64.  Only full batches are evaluated;
 * the rest (fewer than 64) is copied to the carry table and leads the next round,
 * except in the wave's last round of the chunk (`last`), which evaluates
 * everything.  A wave thus spends about pairs / 64 batches on the pair terms.
 * Each lane adds the terms of the pairs it evaluates to its own accumulators; the
 * assignment depends only on the wave's points, so results are reproducible.
 * Every lane of the wave must call this. */
template <typename T, int SUM, bool HESS = true>
__device__ inline int wave_pairs(const NdtGridDev& g, const float4 p, const PointLook& L, const uint32_t cnt,
                                 const PassParams<T>& u, Acc<T, SUM>& acc, WavePairs& wp, uint32_t& carry,
                                 const bool last) {
    const uint32_t lane = threadIdx.x % WAVE;
    /* exclusive prefix of the per-lane counts (< 32) from five bit-plane ballots */
    uint32_t excl = 0, total = 0;
#pragma unroll
    for (int b = 0; b < 5; b++) {
        const uint64_t bal = __ballot((cnt >> b) & 1u);
        excl += __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u)) << b;
        total += (uint32_t)__popcll(bal) << b;
    }
    for (uint32_t j = 0; j < cnt; j++) wp.order[excl + j] = (uint16_t)(WAVE * j + lane);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t all = carry + total;
    const uint32_t full = last ? all : all - all % WAVE;
    int pairs = 0;
    for (uint32_t base = 0; base < full; base += WAVE) {
        const uint32_t n = base + lane;
        bool valid;
        double x0, x1, x2;
        float r0, r1, r2;
        float4 ps;
        uint32_t idx;
        pair_source<T>(wp, n, carry, all, base < carry, p, L, valid, x0, x1, x2, r0, r1, r2, ps, idx);
        pairs += pair_add<T, SUM, HESS>(g, valid, x0, x1, x2, r0, r1, r2, ps, idx, u, acc);
    }
    if (full < all) {
        /* the remainder moves to carry slots 0 ..: pair full + j to slot j.  When no
         * batch ran (full = 0) the carried pairs stay in their slots; otherwise
         * full >= 64 > carry and every remaining pair is new.  Slot j is only ever
         * written and read by lane j. */
        const uint32_t n = full + lane;
        bool valid;
        double x0, x1, x2;
        float r0, r1, r2;
        float4 ps;
        uint32_t idx;
        pair_source<T>(wp, n, carry, all, false, p, L, valid, x0, x1, x2, r0, r1, r2, ps, idx);
        if (valid && n >= carry) {
            if constexpr (sizeof(T) == 4) {
                wp.cx[0][lane] = (double)r0;
                wp.cx[1][lane] = (double)r1;
                wp.cx[2][lane] = (double)r2;
            } else {
                wp.cx[0][lane] = x0;
                wp.cx[1][lane] = x1;
                wp.cx[2][lane] = x2;
            }
            wp.cp[0][lane] = ps.x;
            wp.cp[1][lane] = ps.y;
            wp.cp[2][lane] = ps.z;
            wp.cidx[lane] = idx;
        }
    }
    carry = all - full;
    /* the next round rewrites the tables: its writes stay after these reads */
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return pairs;
}

/* One scan point: lookup, prefilter over the (at most 27) candidate heads with
 * their loads in flight together, then the pairs.  Returns the number of pairs. */
template <typename T, int SUM, bool HESS = true>
__device__ inline int point_accumulate(const NdtGridDev& g, const float4 p, const PassParams<T>& u, Acc<T, SUM>& acc) {
    PointLook L;
    if (!point_lookup(g, p, u.R, u.t, L)) return 0;
    const float lim = prefilter_limit(u);
    uint32_t mask = 0;
#pragma unroll 4
    for (uint32_t k = 0; k < L.count; k++) mask |= prefilter_bit(g, L, k, lim);
    return point_pairs<T, SUM, HESS>(g, p, L, mask, u, acc);
}

/* Points [pb, pe) of one chunk at one trial pose, added to this thread's
 * accumulators; returns this thread's pair count.  With a dense index each wave
 * takes 64 consecutive points per round (wave-uniform trip count), the next
 * round's point load and index lookup are in flight during this round's pairs, and
 * the pairs are spread over the wave in full batches of 64, a round's remainder
 * leading the next round (wave_pairs).  With the hashed index each
 * thread handles its own points.  Both schedulers run their chunks through this,
 * so their partial sums are bit-identical. */
template <typename T, int SUM, int BLOCK, bool HESS = true>
__device__ inline int chunk_accumulate(const NdtGridDev& g, const float4* __restrict__ pts, const int64_t pb,
                                       const int64_t pe, const PassParams<T>& u, Acc<T, SUM>& acc,
                                       WavePairs& wp) {
    int pairs = 0;
    if (g.dense) {
        const int lane = threadIdx.x % WAVE;
        const int64_t wb = pb + (threadIdx.x - lane);
        const float lim = prefilter_limit(u);
        /* point loads are unconditional, clamped to the chunk's last point, so
         * that their waits are counted exactly */
        const bool vc = wb + lane < pe;
        float4 pc = pts[min(wb + lane, pe - 1)];
        PointLook lc;
        uint2 ec;
        lookup_issue(g, pc, vc, u.R, u.t, lc, ec);
        uint32_t carry = 0;
        for (int64_t ib = wb; ib < pe; ib += BLOCK) {
            const bool vn = ib + BLOCK + lane < pe;
            const float4 pn = pts[min(ib + BLOCK + lane, pe - 1)];
            lc.begin = ec.x;
            lc.count = ec.y;
            /* prefilter: the passing candidates go to this lane's column of the
             * table.  Heads are loaded HEAD_BATCH at a time, all loads issued before
             * the first is used (indices past the run repeat its last entry, whose
             * cache line is already on its way, and are not counted), so a point
             * waits for one round trip per HEAD_BATCH candidates.  Branch-free:
             * every candidate is written at the next free position, which a
             * failing one leaves free. */
            uint32_t* mine = wp.pick + lane;
            uint32_t cnt = 0;
            for (uint32_t k0 = 0; k0 < lc.count; k0 += HEAD_BATCH) {
                const uint32_t last = lc.begin + lc.count - 1;
                const uint32_t i0 = lc.begin + k0;
                float4 h[HEAD_BATCH];
#pragma unroll
                for (int q = 0; q < HEAD_BATCH; q++) h[q] = g.head[min(i0 + q, last)];
#pragma unroll
                for (int q = 0; q < HEAD_BATCH; q++) {
                    const float e0 = lc.r0 - h[q].x, e1 = lc.r1 - h[q].y, e2 = lc.r2 - h[q].z;
                    const bool in = (k0 + q < lc.count) && (fmaf(e2, e2, fmaf(e1, e1, e0 * e0)) < lim);
                    mine[WAVE * cnt] = (sizeof(T) == 4) ? i0 + q : __float_as_uint(h[q].w);
                    cnt += in ? 1u : 0u;
                }
            }
            PointLook ln;
            uint2 en;
            lookup_issue(g, pn, vn, u.R, u.t, ln, en);
            pairs += wave_pairs<T, SUM, HESS>(g, pc, lc, cnt, u, acc, wp, carry, ib + BLOCK >= pe);
            pc = pn;
            lc = ln;
            ec = en;
        }
    } else {
        for (int64_t i = pb + threadIdx.x; i < pe; i += BLOCK)
            pairs += point_accumulate<T, SUM, HESS>(g, pts[i], u, acc);
    }
    return pairs;
}


/* Reduces one value per thread per accumulator over the workgroup in a fixed
 * order; thread k < NACC ends with sum k in out[k] (LDS). */
template <int BLOCK>
__device__ inline void block_reduce_acc(double* vals /* NRED per thread, registers */, double (*lds)[NRED],
                                        double* out) {
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < NRED; k++) {
        double v = wave_sum(vals[k]);
        if (lane == 0) lds[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < NRED) {
        double s = 0.0;
        for (int w = 0; w < BLOCK / WAVE; w++) s += lds[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

/* ------------------------------------------------------------------ the Newton / line-search controller
 * Runs on one thread of the workgroup that completes a scan's pass.  Consumes the
 * 28 sums of the pass just evaluated at st.x_t and either chooses the next pose
 * to evaluate or finishes the scan.  Same semantics as computeTransformation /
 * computeStepLengthMT (src/CPU/ndt_mapping/kernel.cpp:1015-1144, :1175-1272) in
 * their textbook form (line search live, translation updated, true convergence). */

/* Gaussian elimination with partial pivoting (rows of A and b swapped together),
 * unrolled so that the system stays in registers.  Same arithmetic as the
 * reference's solve (kernel.cpp:293-343) apart from the right-hand-side swap,
 * which the reference omits. */
__device__ inline void solve6_pivot(double* result, const double* Ain, const double* bin) {
    double A[6][6], b[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        b[i] = bin[i];
#pragma unroll
        for (int k = 0; k < 6; k++) A[i][k] = Ain[i * 6 + k];
    }
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double mx = fabs(A[j][j]);
        int mi = j;
#pragma unroll
        for (int i = j + 1; i < 6; i++) {
            double v = fabs(A[i][j]);
            if (v > mx) {
                mi = i;
                mx = v;
            }
        }
#pragma unroll
        for (int i = j + 1; i < 6; i++) {
            const bool sw = (mi == i);
#pragma unroll
            for (int k = 0; k < 6; k++) {
                double t = A[i][k];
                A[i][k] = sw ? A[j][k] : t;
                A[j][k] = sw ? t : A[j][k];
            }
            double t = b[i];
            b[i] = sw ? b[j] : t;
            b[j] = sw ? t : b[j];
        }
        if (mx == 0.0) A[j][j] = 0.001;
#pragma unroll
        for (int i = j + 1; i < 6; i++) {
            double pivot = A[i][j] / A[j][j];
#pragma unroll
            for (int k = 0; k < 6; k++) A[i][k] = A[i][k] - pivot * A[j][k];
            b[i] = b[i] - pivot * b[j];
        }
    }
    double r[6];
    r[5] = b[5] / A[5][5];
#pragma unroll
    for (int i = 4; i >= 0; i--) {
        double sum = 0.0;
#pragma unroll
        for (int j = i + 1; j < 6; j++) sum = sum + A[i][j] * r[j];
        r[i] = (b[i] - sum) / A[i][i];
    }
#pragma unroll
    for (int i = 0; i < 6; i++) result[i] = r[i];
}

__device__ __attribute__((noinline)) void ctl_finish(ScanState& st, const AlignParams& prm, int converged, ScanOut* out) {
    st.converged = converged;
    st.done = 1;
    st.phase = PH_DONE;
    double M[16];
    pose_matrix(st.p, M);
    M[3] += prm.origin[0];
    M[7] += prm.origin[1];
    M[11] += prm.origin[2];
    for (int k = 0; k < 16; k++) {
        st.final_m[k] = M[k];
        out->final_m[k] = M[k];
    }
    out->score = st.score;
    out->iterations = st.iter;
    out->converged = converged;
    out->passes = st.passes;
    out->status = st.status;
    out->pairs = st.pairs;
    out->used_alt = st.used_alt;
}

/* Plain NDT from the given guess in a launch with refined starts.  The probe pass
 * was at the given guess, so it is the run's first pass: the run starts from its
 * score, gradient and Hessian, and its later passes are numbered (and chunked) as
 * those of a launch of its own. */
__device__ inline void ctl_fallback(ScanState& st) {
    st.fb_stage = FB_FALLBACK;
    st.used_alt = 0;
    st.iter = 0;
    st.converged = 0;
    for (int k = 0; k < 6; k++) {
        st.x_t[k] = st.alt[k];
        st.p[k] = st.alt[k];
    }
    st.score = st.probe_score;
    for (int i = 0; i < 6; i++) st.grad[i] = st.probe_d[i];
    for (int k = 0; k < 21; k++) st.hess[k] = st.probe_d[6 + k];
    st.run_base = st.passes - 1;
}

/* The end of a run.  In a launch with refined starts, a run from the refined guess
 * that ended with a lower score than the given guess has gives way to plain NDT
 * from the given guess: returns 1, and the caller starts a Newton iteration from
 * the state.  Otherwise the scan finishes (returns 0).  The test is the score
 * alone: a run that used all its iterations without converging but ended above
 * the given guess's score is kept (restarting it would put a second full run
 * behind the batch's longest ones). */
__device__ inline int ctl_end(ScanState& st, const AlignParams& prm, int converged, ScanOut* out) {
    if (st.fb_stage == FB_REFINED && st.score < st.probe_score) {
        ctl_fallback(st);
        return 1;
    }
    st.used_alt = st.fb_stage == FB_REFINED;
    ctl_finish(st, prm, converged, out);
    return 0;
}

__device__ __attribute__((noinline)) void ctl_newton_start(ScanState& __restrict__ st, const AlignParams& __restrict__ prm,
                                                           ScanOut* __restrict__ out) {
    const double mu = 1.e-4;
    for (;;) {
        if (st.iter >= prm.max_iterations) {
            if (ctl_end(st, prm, 0, out)) continue;
            return;
        }
        double H[36], ng[6], delta[6];
        for (int i = 0; i < 6; i++) {
            ng[i] = -st.grad[i];
            for (int j = 0; j < 6; j++) H[i * 6 + j] = st.hess[i <= j ? tri_index(i, j) : tri_index(j, i)];
        }
        solve6_pivot(delta, H, ng);
        double norm = sqrt(delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2] + delta[3] * delta[3] +
                           delta[4] * delta[4] + delta[5] * delta[5]);
        if (norm == 0 || norm != norm) {
            if (ctl_end(st, prm, norm == norm, out)) continue;
            return;
        }
        for (int i = 0; i < 6; i++) st.dir[i] = delta[i] / norm;
        st.phi_0 = -st.score;
        st.d_phi_0 = -dot6(st.grad, st.dir);
        if (st.d_phi_0 >= 0) {
            if (st.d_phi_0 == 0) {
                /* no descent direction: zero step */
                st.iter++;
                if (st.iter > 1 && 0.0 < prm.trans_eps) {
                    if (ctl_end(st, prm, 1, out)) continue;
                    return;
                }
                continue;
            }
            st.d_phi_0 *= -1;
            for (int i = 0; i < 6; i++) st.dir[i] *= -1;
        }
        st.a_l = 0;
        st.a_u = 0;
        st.f_l = psi_mt(st.a_l, st.phi_0, st.phi_0, st.d_phi_0, mu);
        st.g_l = dpsi_mt(st.d_phi_0, st.d_phi_0, mu);
        st.f_u = psi_mt(st.a_u, st.phi_0, st.phi_0, st.d_phi_0, mu);
        st.g_u = dpsi_mt(st.d_phi_0, st.d_phi_0, mu);
        st.interval_converged = 0;
        st.open_interval = 1;
        st.step_iterations = 0;
        double a_t = norm;
        a_t = dmin_std(a_t, prm.step_size);
        a_t = dmax_std(a_t, prm.trans_eps / 2);
        st.a_t = a_t;
        for (int i = 0; i < 6; i++) st.x_t[i] = st.p[i] + st.dir[i] * a_t;
        st.phase = PH_MT_FIRST;
        return;
    }
}

/* accepts the line search's step x_t = p + a_t dir, whose score, gradient and
 * Hessian are in the state, and starts the next Newton iteration */
__device__ inline void ctl_accept(ScanState& st, const AlignParams& prm, ScanOut* out) {
    for (int i = 0; i < 6; i++) st.p[i] = st.p[i] + st.dir[i] * st.a_t;
    st.iter++;
    if (st.iter > 1 && fabs(st.a_t) < prm.trans_eps) {
        if (!ctl_end(st, prm, 1, out)) return;
    }
    ctl_newton_start(st, prm, out);
}

__device__ __attribute__((noinline)) void ctl_step(ScanState& __restrict__ st, const double* __restrict__ red,
                                                   const AlignParams& __restrict__ prm, ScanOut* __restrict__ out) {
    const double mu = 1.e-4, nu = 0.9;
    st.passes++;
    st.pairs += red[NACC];
    if (st.phase == PH_PROBE) {
        /* the given guess's score and derivatives; the run starts at the refined
         * guess */
        st.probe_score = red[0];
        for (int k = 0; k < 27; k++) st.probe_d[k] = red[1 + k];
        for (int k = 0; k < 6; k++) {
            const double given = st.x_t[k];
            st.x_t[k] = st.alt[k];
            st.p[k] = st.alt[k];
            st.alt[k] = given;
        }
        st.fb_stage = FB_REFINED;
        st.phase = PH_INIT;
        st.run_base = st.passes;
        return;
    }
    if (st.phase == PH_MT_HESS) {
        /* the Hessian at the accepted step of a line search whose loop ran
         * (computeHessian, kernel.cpp:1141-1142); score and gradient are the last
         * trial's */
        for (int k = 0; k < 21; k++) st.hess[k] = red[7 + k];
        ctl_accept(st, prm, out);
        return;
    }
    if (st.phase == PH_INIT) {
        if (st.fb_stage == FB_REFINED && red[0] < st.probe_score) {
            /* NDT scores the refined guess below the given one: the scan runs from
             * the given guess instead */
            ctl_fallback(st);
            ctl_newton_start(st, prm, out);
            return;
        }
        st.score = red[0];
        for (int i = 0; i < 6; i++) st.grad[i] = red[1 + i];
        for (int k = 0; k < 21; k++) st.hess[k] = red[7 + k];
        for (int i = 0; i < 6; i++) st.p[i] = st.x_t[i];
        ctl_newton_start(st, prm, out);
        return;
    }
    /* a line-search evaluation at x_t = p + a_t dir */
    const bool red_has_hess = !(prm.hess_at_accept && st.phase == PH_MT_LOOP);
    double g_t[6];
    for (int i = 0; i < 6; i++) g_t[i] = red[1 + i];
    const double phi_t = -red[0];
    const double d_phi_t = -dot6(g_t, st.dir);
    const double psi_t = psi_mt(st.a_t, phi_t, st.phi_0, st.d_phi_0, mu);
    const double d_psi_t = dpsi_mt(d_phi_t, st.d_phi_0, mu);
    /* A trial at the step of the trial before it (the trial value clamped to the
     * same bound again) is at the same pose, whose pass would reproduce `red` bit
     * for bit: the search takes `red` again without a pass, for as long as it
     * repeats the step. */
    for (;;) {
        if (st.phase == PH_MT_LOOP) {
            if (st.open_interval && (psi_t <= 0 && d_psi_t >= 0)) {
                st.open_interval = 0;
                st.f_l = st.f_l + st.phi_0 - mu * st.d_phi_0 * st.a_l;
                st.g_l = st.g_l + mu * st.d_phi_0;
                st.f_u = st.f_u + st.phi_0 - mu * st.d_phi_0 * st.a_u;
                st.g_u = st.g_u + mu * st.d_phi_0;
            }
            if (st.open_interval)
                st.interval_converged =
                    mt_update_interval(st.a_l, st.f_l, st.g_l, st.a_u, st.f_u, st.g_u, st.a_t, psi_t, d_psi_t);
            else
                st.interval_converged =
                    mt_update_interval(st.a_l, st.f_l, st.g_l, st.a_u, st.f_u, st.g_u, st.a_t, phi_t, d_phi_t);
            st.step_iterations++;
        }
        if (st.interval_converged || st.step_iterations >= 10 || (psi_t <= 0 && d_phi_t <= -nu * st.d_phi_0))
            break;
        double a_t;
        if (st.open_interval)
            a_t = mt_trial_value(st.a_l, st.f_l, st.g_l, st.a_u, st.f_u, st.g_u, st.a_t, psi_t, d_psi_t);
        else
            a_t = mt_trial_value(st.a_l, st.f_l, st.g_l, st.a_u, st.f_u, st.g_u, st.a_t, phi_t, d_phi_t);
        a_t = dmin_std(a_t, prm.step_size);
        a_t = dmax_std(a_t, prm.trans_eps / 2);
        const bool repeat = a_t == st.a_t;
        st.phase = PH_MT_LOOP;
        if (repeat) continue;
        st.a_t = a_t;
        for (int i = 0; i < 6; i++) st.x_t[i] = st.p[i] + st.dir[i] * a_t;
        return;
    }
    /* accept x_t: its pass supplies score and gradient, and the Hessian unless the
     * loop's passes skipped it (then one more pass at x_t computes it, unless the
     * pass in `red`, at the same pose, did) */
    st.score = red[0];
    for (int i = 0; i < 6; i++) st.grad[i] = g_t[i];
    if (prm.hess_at_accept && st.phase == PH_MT_LOOP && !red_has_hess) {
        st.phase = PH_MT_HESS;
        return;
    }
    for (int k = 0; k < 21; k++) st.hess[k] = red[7 + k];
    ctl_accept(st, prm, out);
}

/* pose parameters (x, y, z, roll, pitch, yaw) of a row-major world pose in the
 * map-local frame */
__device__ inline void pose_params(const double* G, const AlignParams& prm, double* x) {
    double M[16];
    for (int k = 0; k < 16; k++) M[k] = G[k];
    M[3] -= prm.origin[0];
    M[7] -= prm.origin[1];
    M[11] -= prm.origin[2];
    double ea[3];
    pose_euler(M, ea);
    x[0] = M[3];
    x[1] = M[7];
    x[2] = M[11];
    x[3] = ea[0];
    x[4] = ea[1];
    x[5] = ea[2];
}

__global__ void k_init_states(const double* __restrict__ guesses, const double* __restrict__ alt, int32_t S,
                              AlignParams prm, ScanState* states, RoundSync* rs) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    ScanState st;
    uint64_t* w = (uint64_t*)&st;
    for (int k = 0; k < STATE_WORDS; k++) w[k] = 0;
    pose_params(guesses + 16 * s, prm, st.x_t);
    for (int k = 0; k < 6; k++) st.p[k] = st.x_t[k];
    pass_params_f64(st.x_t, prm, prm.resolution_d, st.pp);
    st.phase = PH_INIT;
    if (alt) {
        pose_params(alt + 16 * s, prm, st.alt);
        st.fb_stage = FB_PROBE;
        st.phase = PH_PROBE;
    }
    states[s] = st;
    if (s == 0) {
        rs->count = 0;
        rs->abort = 0;
    }
}

int launch_init_states(const double* guesses, const double* alt, int32_t S, AlignParams prm, ScanState* states,
                       RoundSync* rs, hipStream_t stream) {
    hipLaunchKernelGGL(k_init_states, dim3(grid_for(S, 64)), dim3(64), 0, stream, guesses, alt, S, prm, states, rs);
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------ persistent align kernel
 * One cooperative launch aligns every scan of a batch in rounds.  In a round each
 * still-active scan is evaluated once at its current trial pose: the points of all
 * active scans are cut into chunks of a fixed size, the co-resident workgroups
 * take contiguous ranges of chunks, and every chunk leaves its 28 partial sums (and
 * pair count) in its own slot.  After a grid barrier the scans are spread over the
 * workgroups: each reduces its chunks' partials in chunk order (deterministic, and
 * independent of the number of workgroups) and runs the controller, which chooses
 * the next trial pose or finishes the scan.  A second grid barrier closes the
 * round.  Scans that finish drop out, so the last slow scans get the whole GPU.
 * All waits are bounded. */

constexpr int MAX_BATCH = 2048; /* scans per launch (LDS tables) */
/* waves per SIMD the register allocation must allow: float64 points or float64
 * sums need up to 256 VGPRs to keep the point loop free of spills, float32 points
 * with float32 (plain or compensated) sums fit in 168 */
#ifndef ALIGN_MIN_WAVES_F64
#define ALIGN_MIN_WAVES_F64 2
#endif
#ifndef ALIGN_MIN_WAVES_F32
#define ALIGN_MIN_WAVES_F32 3
#endif

/* Grid-wide barrier of a cooperative launch.
 *
 * Every value that one workgroup hands to another (scan states, pass parameters,
 * partial sums, done flags) is written and read with device-scope relaxed atomics,
 * which gfx950 issues as sc1 stores and loads: they complete at, and read from,
 * the device coherence point and never leave a stale copy in an XCD's L2.  The
 * barrier therefore needs no cache maintenance: each wave waits until its own
 * stores have completed (s_waitcnt vmcnt(0)), then one lane per workgroup bumps a
 * monotonically growing arrival counter and polls it until all workgroups of this
 * barrier have arrived.  A release/acquire pair would add an L2 write-back and, in
 * every poll, an L2 invalidate that evicts the read-only voxel data the computing
 * workgroups are using.  `target` is nwg * (barrier number + 1).  The wait is
 * bounded in time; on expiry the workgroup raises the abort flag, which every
 * other workgroup sees in its own wait.  Returns false after an abort. */
__device__ inline bool grid_barrier(RoundSync* rs, uint32_t target, uint64_t wait_ticks, int* s_flag) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __hip_atomic_fetch_add(&rs->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        uint32_t spins = 0;
        while (__hip_atomic_load(&rs->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(8);
            if ((++spins & 63u) == 0) {
                if (__hip_atomic_load(&rs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    ok = 0;
                    break;
                }
                if (__builtin_amdgcn_s_memrealtime() - t0 > wait_ticks) {
                    __hip_atomic_store(&rs->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
        }
        if (__hip_atomic_load(&rs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = 0;
        *s_flag = ok;
    }
    __syncthreads();
    return *s_flag != 0;
}

/* exclusive scan over n <= MAX_BATCH values held in LDS (in place); returns the total */
template <int BLOCK>
__device__ inline uint32_t lds_exclusive_scan(uint32_t* v, int n, uint32_t* lw) {
    const int per = (n + BLOCK - 1) / BLOCK;
    const int b = threadIdx.x * per;
    uint32_t sum = 0;
    for (int i = b; i < b + per && i < n; i++) sum += v[i];
    uint32_t ex;
    uint32_t total = block_exclusive_scan(sum, &ex, lw);
    uint32_t run = ex;
    for (int i = b; i < b + per && i < n; i++) {
        uint32_t t = v[i];
        v[i] = run;
        run += t;
    }
    __syncthreads();
    return total;
}

/* Sum of a scan's chunk slots in a fixed order, the same in every kernel and for
 * every block size: four interleaved groups (group q sums chunks q, q + 4, ...)
 * added in group order.  Threads 0..127 take part; red[0..NRED) receives the sums. */
constexpr int SLOT_GROUPS = 4;
__device__ inline void reduce_slots(const double* partials, int64_t base, int nchunks, double (*gred)[32],
                                    double* red) {
    const int grp = threadIdx.x >> 5, j = threadIdx.x & 31;
    if (threadIdx.x < SLOT_GROUPS * 32) {
        double a = 0.0;
        if (j < NRED) {
            /* eight slots' loads in flight at once, added in chunk order */
            for (int c0 = grp; c0 < nchunks; c0 += 8 * SLOT_GROUPS) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int c = c0 + u * SLOT_GROUPS;
                    v[u] = c < nchunks ? ld_agent_f64(partials + (base + c) * 32 + j) : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (c0 + u * SLOT_GROUPS < nchunks) a += v[u];
            }
        }
        gred[grp][j] = a;
    }
    __syncthreads();
    if (threadIdx.x < NRED) {
        double r = gred[0][threadIdx.x];
        for (int q = 1; q < SLOT_GROUPS; q++) r += gred[q][threadIdx.x];
        red[threadIdx.x] = r;
    }
    __syncthreads();
}

template <typename T, int SUM, int BLOCK, bool HA>
__global__ __launch_bounds__(BLOCK, (sizeof(T) == 8 || SUM == NDT_SUM_FP64) ? ALIGN_MIN_WAVES_F64 : ALIGN_MIN_WAVES_F32) void k_align(
    NdtGridDev g, const float4* __restrict__ pts, const ScanDesc* __restrict__ scans, int32_t S,
    AlignParams prm, ScanState* states, RoundSync* rs, double* partials, ScanOut* out, double* prof_out) {
    __shared__ PassParams<T> pp;
    __shared__ double lds_red[BLOCK / WAVE][NRED];
    __shared__ double red[NRED];
    __shared__ uint64_t sstate[STATE_WORDS];
    __shared__ double s_trig[6];
    __shared__ double gred[SLOT_GROUPS][32];
    __shared__ WavePairs wslots[BLOCK / WAVE];
    __shared__ uint32_t s_active[MAX_BATCH];
    __shared__ uint32_t s_prefix[MAX_BATCH + 1];
    __shared__ uint32_t lw[BLOCK / WAVE];
    __shared__ int s_nactive, s_pp_scan, s_flag, s_grad_only;
    const int tid = threadIdx.x;
    const int wg = blockIdx.x;
    const int nwg = gridDim.x;
    uint32_t target = 0;
    double prof[4] = {0, 0, 0, 0};
    bool ok = true;
    if (tid == 0) s_pp_scan = -1;
    for (;;) {
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        /* active scans and their chunk counts, identical in every workgroup */
        for (int i = tid; i < S; i += BLOCK) s_prefix[i] = ld_agent_i32(&states[i].done) ? 0u : 1u;
        __syncthreads();
        const uint32_t nactive = lds_exclusive_scan<BLOCK>(s_prefix, S, lw);
        for (int i = tid; i < S; i += BLOCK) {
            const uint32_t a = s_prefix[i];
            const uint32_t next = (i + 1 < S) ? s_prefix[i + 1] : nactive;
            if (next != a) s_active[a] = (uint32_t)i;
        }
        __syncthreads();
        if (nactive == 0) break;
        for (int a = tid; a < (int)nactive; a += BLOCK) {
            const int s = (int)s_active[a];
            s_prefix[a] = (uint32_t)pass_nchunks(scans[s], (uint32_t)(ld_agent_i32(&states[s].passes) -
                                                                       ld_agent_i32(&states[s].run_base)));
        }
        __syncthreads();
        const uint32_t nchunks = lds_exclusive_scan<BLOCK>(s_prefix, (int)nactive, lw);
        if (tid == 0) s_prefix[nactive] = nchunks;
        __syncthreads();
        const uint64_t t1 = __builtin_amdgcn_s_memrealtime();
        /* this workgroup's contiguous range of chunks */
        const uint32_t c0 = (uint32_t)(((uint64_t)nchunks * wg) / nwg);
        const uint32_t c1 = (uint32_t)(((uint64_t)nchunks * (wg + 1)) / nwg);
        uint32_t a = 0;
        for (uint32_t c = c0; c < c1; c++) {
            /* active index of chunk c: last a with prefix[a] <= c */
            uint32_t lo = a, hi = nactive;
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if (s_prefix[mid] <= c) lo = mid;
                else hi = mid;
            }
            a = lo;
            const int s = (int)s_active[a];
            const ScanDesc sd = scans[s];
            const int64_t k = c - s_prefix[a];
            if (s != s_pp_scan) {
                __syncthreads();
                if (tid < 32) pp.set(tid, ld_agent_f64(&states[s].pp[tid]));
                if (tid == 0) {
                    s_pp_scan = s;
                    s_grad_only = HA && grad_only_pass(ld_agent_i32(&states[s].phase));
                }
                __syncthreads();
            }
            Acc<T, SUM> acc;
            acc.zero();
            int pairs = 0;
            {
                const PassParams<T> u = load_uniform(pp);
                const int64_t cs = pass_chunk(sd, (uint32_t)(ld_agent_i32(&states[s].passes) -
                                                            ld_agent_i32(&states[s].run_base)));
                const int64_t pb = sd.offset + k * cs;
                const int64_t pe = sd.offset + min((int64_t)sd.n, (k + 1) * cs);
                if (HA && s_grad_only)
                    pairs = chunk_accumulate<T, SUM, BLOCK, false>(g, pts, pb, pe, u, acc, wslots[tid / WAVE]);
                else
                    pairs = chunk_accumulate<T, SUM, BLOCK, true>(g, pts, pb, pe, u, acc, wslots[tid / WAVE]);
            }
            double v[NRED];
#pragma unroll
            for (int q = 0; q < NACC; q++) v[q] = acc.get(q);
            v[NACC] = (double)pairs;
            block_reduce_acc<BLOCK>(v, lds_red, red);
            if (tid < NRED) st_agent_f64(partials + (sd.chunk_base + k) * 32 + tid, red[tid]);
        }
        const uint64_t t2 = __builtin_amdgcn_s_memrealtime();
        target += (uint32_t)nwg;
        ok = grid_barrier(rs, target, prm.wait_ticks, &s_flag);
        if (!ok) break;
        const uint64_t t3 = __builtin_amdgcn_s_memrealtime();
        /* reductions and controllers, scans spread over the workgroups */
        for (uint32_t ai = (uint32_t)wg; ai < nactive; ai += (uint32_t)nwg) {
            const int s = (int)s_active[ai];
            const ScanDesc sd = scans[s];
            for (int q = tid; q < STATE_WORDS; q += BLOCK) sstate[q] = ld_agent_u64((const uint64_t*)(states + s) + q);
            __syncthreads();
            reduce_slots(partials, sd.chunk_base, pass_nchunks(sd, run_pass(*(ScanState*)sstate)), gred, red);
            if (tid == 0) ctl_step(*(ScanState*)sstate, red, prm, out + s);
            __syncthreads();
            if (tid < WAVE && !((ScanState*)sstate)->done) wave_pass_params(*(ScanState*)sstate, prm, s_trig);
            __syncthreads();
            for (int q = tid; q < STATE_WORDS; q += BLOCK) st_agent_u64((uint64_t*)(states + s) + q, sstate[q]);
            __syncthreads();
        }
        if (tid == 0) s_pp_scan = -1;
        const uint64_t t4 = __builtin_amdgcn_s_memrealtime();
        target += (uint32_t)nwg;
        ok = grid_barrier(rs, target, prm.wait_ticks, &s_flag);
        if (!ok) break;
        const uint64_t t5 = __builtin_amdgcn_s_memrealtime();
        if (wg == 0 && tid == 0) {
            prof[0] += (double)(t1 - t0);
            prof[1] += (double)(t2 - t1);
            prof[2] += (double)(t3 - t2) + (double)(t5 - t4);
            prof[3] += (double)(t4 - t3);
        }
    }
    if (wg == 0 && tid == 0) {
        /* s_memrealtime counts at 100 MHz */
        for (int q = 0; q < 4; q++) prof_out[q] = prof[q] * 0.01;
        prof_out[4] = ok ? 0.0 : 1.0;
    }
}

template <typename T, int SUM, int BLOCK>
static int launch_align_t(NdtGridDev g, const float4* pts, const ScanDesc* scans, int32_t S,
                          int32_t nwg, AlignParams prm, ScanState* states, RoundSync* rs, double* partials,
                          ScanOut* out, double* prof, hipStream_t stream) {
    void* args[] = {&g, &pts, &scans, &S, &prm, &states, &rs, &partials, &out, &prof};
    const void* kern = prm.hess_at_accept ? (const void*)k_align<T, SUM, BLOCK, true>
                                          : (const void*)k_align<T, SUM, BLOCK, false>;
    return (int)hipLaunchCooperativeKernel(kern, dim3(nwg), dim3(BLOCK), args, 0, stream);
}

/* occupancy of the kernel without the gradient-only passes (the other one holds the
 * same registers and LDS, checked by the launches that use it) */
template <typename T, int SUM, int BLOCK>
static int max_blocks_t(int* bpc) {
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, (const void*)k_align<T, SUM, BLOCK, false>, BLOCK,
                                                             0);
}

#define NDT_DISPATCH(BLOCKSZ, FP64, SUMMODE, CALL)                                         \
    do {                                                                                  \
        if (FP64) {                                                                       \
            if (BLOCKSZ == 128) return CALL(double, NDT_SUM_FP64, 128);                    \
            if (BLOCKSZ == 256) return CALL(double, NDT_SUM_FP64, 256);                    \
            if (BLOCKSZ == 512) return CALL(double, NDT_SUM_FP64, 512);                    \
        } else if (SUMMODE == NDT_SUM_FP32) {                                             \
            if (BLOCKSZ == 128) return CALL(float, NDT_SUM_FP32, 128);                     \
            if (BLOCKSZ == 256) return CALL(float, NDT_SUM_FP32, 256);                     \
            if (BLOCKSZ == 512) return CALL(float, NDT_SUM_FP32, 512);                     \
        } else if (SUMMODE == NDT_SUM_COMP32) {                                           \
            if (BLOCKSZ == 128) return CALL(float, NDT_SUM_COMP32, 128);                   \
            if (BLOCKSZ == 256) return CALL(float, NDT_SUM_COMP32, 256);                   \
            if (BLOCKSZ == 512) return CALL(float, NDT_SUM_COMP32, 512);                   \
        } else {                                                                          \
            if (BLOCKSZ == 128) return CALL(float, NDT_SUM_FP64, 128);                     \
            if (BLOCKSZ == 256) return CALL(float, NDT_SUM_FP64, 256);                     \
            if (BLOCKSZ == 512) return CALL(float, NDT_SUM_FP64, 512);                     \
        }                                                                                 \
        return (int)hipErrorInvalidValue;                                                 \
    } while (0)

int align_kernel_max_blocks(int block_size, int point_fp64, int sum_mode, int* blocks_per_cu) {
#define NDT_MB(T, S, B) max_blocks_t<T, S, B>(blocks_per_cu)
    NDT_DISPATCH(block_size, point_fp64, sum_mode, NDT_MB);
#undef NDT_MB
}

int launch_align(NdtGridDev grid, const float4* pts, const ScanDesc* scans, int32_t S, int32_t nwg,
                 int block_size, AlignParams prm, ScanState* states, RoundSync* rs, double* partials, ScanOut* out,
                 double* prof, hipStream_t stream) {
    if (S > MAX_BATCH) return (int)hipErrorInvalidValue;
#define NDT_LA(T, S_, B) \
    launch_align_t<T, S_, B>(grid, pts, scans, S, nwg, prm, states, rs, partials, out, prof, stream)
    NDT_DISPATCH(block_size, prm.point_fp64, prm.sum_mode, NDT_LA);
#undef NDT_LA
}

/* ------------------------------------------------------------------ asynchronous align kernel
 * One launch aligns every scan of a batch without grid-wide synchronisation.  A
 * scan's pass is cut into chunks; each chunk is an item of a work queue.  Every
 * workgroup repeatedly claims the next queue position, waits until its item is
 * written, evaluates the chunk at the scan's current trial pose and leaves the 28
 * sums (and pair count) in the chunk's slot.  The workgroup that completes a
 * scan's last chunk (a per-scan counter) reduces the slots in chunk order, runs
 * the controller and appends the chunks of the scan's next pass to the queue, or
 * retires the scan.  Scans therefore advance independently: nothing waits for
 * the slowest chunk of other scans.  Values handed between workgroups use
 * device-scope relaxed atomics (sc1 loads and stores); each wave waits for its
 * stores to complete before the counter or queue update that publishes them.
 * Every wait is bounded in time and ends when all scans are finished. */

__device__ inline void stores_complete() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__global__ void k_init_queue(const ScanDesc* __restrict__ scans, int32_t S, uint64_t total, WorkQueue* q,
                             uint64_t* __restrict__ items, uint32_t* __restrict__ done_cnt) {
    const int s = blockIdx.x;
    if (s >= S) return;
    const ScanDesc sd = scans[s];
    for (int k = threadIdx.x; k < sd.nchunks; k += blockDim.x) items[sd.qbase + k] = make_item(0, s, k);
    if (threadIdx.x == 0) done_cnt[s] = 0u;
    if (s == 0 && threadIdx.x == 0) {
        q->head = 0u;
        q->tail = (uint32_t)total;
        q->remaining = (uint32_t)S;
        q->abort = 0u;
    }
}

int launch_init_queue(const ScanDesc* scans, int32_t S, uint64_t total_chunks, WorkQueue* q, uint64_t* items,
                      uint32_t* done_cnt, hipStream_t stream) {
    hipLaunchKernelGGL(k_init_queue, dim3(S), dim3(256), 0, stream, scans, S, total_chunks, q, items, done_cnt);
    return (int)hipGetLastError();
}

template <typename T, int SUM, int BLOCK, bool HA>
__global__ __launch_bounds__(BLOCK, (sizeof(T) == 8 || SUM == NDT_SUM_FP64) ? ALIGN_MIN_WAVES_F64 : ALIGN_MIN_WAVES_F32) void k_align_wq(
    NdtGridDev g, const float4* __restrict__ pts, const ScanDesc* __restrict__ scans, AlignParams prm,
    ScanState* states, WorkQueue* q, uint64_t* items, uint64_t cap, uint32_t* done_cnt, double* partials,
    ScanOut* out, double* prof_out) {
    __shared__ PassParams<T> pp;
    __shared__ double lds_red[BLOCK / WAVE][NRED];
    __shared__ double red[NRED];
    __shared__ double gred[SLOT_GROUPS][32];
    __shared__ uint64_t sstate[STATE_WORDS];
    __shared__ double s_trig[6];
    __shared__ WavePairs wslots[BLOCK / WAVE];
    __shared__ uint64_t s_item, s_ppkey;
    __shared__ uint32_t s_base;
    __shared__ int s_last, s_done, s_grad_only;
    __shared__ uint32_t s_run_base, s_next_rel;
    const int tid = threadIdx.x;
    double prof[3] = {0, 0, 0};
    int ok = 1;
    if (tid == 0) s_ppkey = ~0ull;
    for (;;) {
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        if (tid == 0) {
            uint64_t v = 0;
            const uint32_t pos = __hip_atomic_fetch_add(&q->head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((uint64_t)pos < cap) {
                uint32_t spins = 0;
                for (;;) {
                    v = ld_agent_u64(items + pos);
                    if (v) break;
                    if (__hip_atomic_load(&q->remaining, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u ||
                        __hip_atomic_load(&q->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
                        break;
                    __builtin_amdgcn_s_sleep(4);
                    if ((++spins & 63u) == 0 && __builtin_amdgcn_s_memrealtime() - t0 > prm.wait_ticks) {
                        __hip_atomic_store(&q->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
            s_item = v;
        }
        __syncthreads();
        const uint64_t item = s_item;
        const uint64_t t1 = __builtin_amdgcn_s_memrealtime();
        if (!item) break;
        const int s = (int)((item >> ITEM_CHUNK_BITS) & ((1u << ITEM_SCAN_BITS) - 1));
        const int k = (int)(item & ((1u << ITEM_CHUNK_BITS) - 1));
        const uint32_t pass = (uint32_t)((item >> (ITEM_CHUNK_BITS + ITEM_SCAN_BITS)) & ((1u << ITEM_PASS_BITS) - 1));
        const uint64_t ppkey = item >> ITEM_CHUNK_BITS;
        const ScanDesc sd = scans[s];
        if (ppkey != s_ppkey) {
            /* the parameters, the run's first pass and the phase in one round trip */
            if (tid < 32) pp.set(tid, ld_agent_f64(&states[s].pp[tid]));
            else if (tid == 32) s_run_base = (uint32_t)ld_agent_i32(&states[s].run_base);
            else if (tid == 33) s_grad_only = HA && grad_only_pass(ld_agent_i32(&states[s].phase));
            __syncthreads();
            if (tid == 0) s_ppkey = ppkey;
            __syncthreads();
        }
        Acc<T, SUM> acc;
        acc.zero();
        int pairs = 0;
        const uint32_t rel = pass - s_run_base; /* pass number within the scan's current run */
        const int32_t nck = pass_nchunks(sd, rel);
        {
            const PassParams<T> u = load_uniform(pp);
            const int64_t cs = pass_chunk(sd, rel);
            const int64_t pb = sd.offset + (int64_t)k * cs;
            const int64_t pe = sd.offset + min((int64_t)sd.n, (int64_t)(k + 1) * cs);
            if (HA && s_grad_only)
                pairs = chunk_accumulate<T, SUM, BLOCK, false>(g, pts, pb, pe, u, acc, wslots[tid / WAVE]);
            else
                pairs = chunk_accumulate<T, SUM, BLOCK, true>(g, pts, pb, pe, u, acc, wslots[tid / WAVE]);
        }
        double v[NRED];
#pragma unroll
        for (int qq = 0; qq < NACC; qq++) v[qq] = acc.get(qq);
        v[NACC] = (double)pairs;
        block_reduce_acc<BLOCK>(v, lds_red, red);
        if (tid < NRED) st_agent_f64(partials + (sd.chunk_base + k) * 32 + tid, red[tid]);
        stores_complete();
        __syncthreads();
        if (tid == 0) {
            const uint32_t old = __hip_atomic_fetch_add(&done_cnt[s], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_last = (old + 1u == (uint32_t)nck) ? 1 : 0;
        }
        __syncthreads();
        const uint64_t t2 = __builtin_amdgcn_s_memrealtime();
        if (s_last) {
            /* every chunk of this pass is in: reduce in chunk order, G interleaved groups */
            if (tid == 0) __hip_atomic_store(&done_cnt[s], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (int w = tid; w < STATE_WORDS; w += BLOCK) sstate[w] = ld_agent_u64((const uint64_t*)(states + s) + w);
            reduce_slots(partials, sd.chunk_base, nck, gred, red);
            if (tid == 0) {
                ScanState& st = *(ScanState*)sstate;
                ctl_step(st, red, prm, out + s);
                if (!st.done && (uint64_t)pass + 1 >= (1ull << ITEM_PASS_BITS)) {
                    st.status = NDT_ERR_UNSUPPORTED;
                    ctl_finish(st, prm, 0, out + s);
                }
                s_done = st.done;
                s_next_rel = run_pass(st);
            }
            __syncthreads();
            if (tid < WAVE && !s_done) wave_pass_params(*(ScanState*)sstate, prm, s_trig);
            __syncthreads();
            for (int w = tid; w < STATE_WORDS; w += BLOCK) st_agent_u64((uint64_t*)(states + s) + w, sstate[w]);
            stores_complete();
            __syncthreads();
            if (!s_done) {
                const int32_t nnext = pass_nchunks(sd, s_next_rel);
                if (tid == 0)
                    s_base = __hip_atomic_fetch_add(&q->tail, (uint32_t)nnext, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
                __syncthreads();
                const uint64_t base = s_base;
                if (base + (uint64_t)nnext > cap) {
                    if (tid == 0) __hip_atomic_store(&q->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    for (int c = tid; c < nnext; c += BLOCK) st_agent_u64(items + base + c, make_item(pass + 1, s, c));
                }
            } else if (tid == 0) {
                __hip_atomic_fetch_sub(&q->remaining, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        const uint64_t t3 = __builtin_amdgcn_s_memrealtime();
        if (blockIdx.x == 0 && tid == 0) {
            prof[0] += (double)(t1 - t0);
            prof[1] += (double)(t2 - t1);
            prof[2] += (double)(t3 - t2);
        }
    }
    if (tid == 0 && __hip_atomic_load(&q->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = 0;
    if (blockIdx.x == 0 && tid == 0) {
        /* s_memrealtime counts at 100 MHz: active-list time is not used here */
        prof_out[0] = 0.0;
        prof_out[1] = prof[1] * 0.01;
        prof_out[2] = prof[0] * 0.01;
        prof_out[3] = prof[2] * 0.01;
        prof_out[4] = ok ? 0.0 : 1.0;
    }
}

/* The work queue needs no co-residency: a workgroup only waits for an item that a
 * running workgroup will produce (the claims are taken in order by running
 * workgroups and no running workgroup waits while holding an item), so the kernel
 * may also be launched normally, which graphs can capture. */
template <typename T, int SUM, int BLOCK>
static int launch_align_wq_t(NdtGridDev g, const float4* pts, const ScanDesc* scans, int32_t nwg,
                             AlignParams prm, ScanState* states, WorkQueue* q, uint64_t* items, uint64_t cap,
                             uint32_t* done_cnt, double* partials, ScanOut* out, double* prof, int cooperative,
                             hipStream_t stream) {
    if (!cooperative) {
        if (prm.hess_at_accept)
            hipLaunchKernelGGL((k_align_wq<T, SUM, BLOCK, true>), dim3(nwg), dim3(BLOCK), 0, stream, g, pts, scans, prm,
                               states, q, items, cap, done_cnt, partials, out, prof);
        else
            hipLaunchKernelGGL((k_align_wq<T, SUM, BLOCK, false>), dim3(nwg), dim3(BLOCK), 0, stream, g, pts, scans, prm,
                               states, q, items, cap, done_cnt, partials, out, prof);
        return (int)hipGetLastError();
    }
    void* args[] = {&g, &pts, &scans, &prm, &states, &q, &items, &cap, &done_cnt, &partials, &out, &prof};
    const void* kern = prm.hess_at_accept ? (const void*)k_align_wq<T, SUM, BLOCK, true>
                                          : (const void*)k_align_wq<T, SUM, BLOCK, false>;
    return (int)hipLaunchCooperativeKernel(kern, dim3(nwg), dim3(BLOCK), args, 0, stream);
}

__global__ void k_zero_u32(uint32_t* __restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = 0u;
}

int launch_zero_u32(void* dst, int64_t n, hipStream_t stream) {
    if (n <= 0) return 0;
    const int64_t want = (n + 255) / 256;
    const int blocks = (int)(want < 4096 ? want : 4096);
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks), dim3(256), 0, stream, (uint32_t*)dst, n);
    return (int)hipGetLastError();
}

template <typename T, int SUM, int BLOCK>
static int max_blocks_wq_t(int* bpc) {
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, (const void*)k_align_wq<T, SUM, BLOCK, false>,
                                                             BLOCK, 0);
}

int align_wq_max_blocks(int block_size, int point_fp64, int sum_mode, int* blocks_per_cu) {
#define NDT_MB(T, S, B) max_blocks_wq_t<T, S, B>(blocks_per_cu)
    NDT_DISPATCH(block_size, point_fp64, sum_mode, NDT_MB);
#undef NDT_MB
}

int launch_align_wq(NdtGridDev g, const float4* pts, const ScanDesc* scans, int32_t nwg, int block_size,
                    AlignParams prm, ScanState* states, WorkQueue* q, uint64_t* items, uint64_t cap, uint32_t* done_cnt,
                    double* partials, ScanOut* out, double* prof, int cooperative, hipStream_t stream) {
#define NDT_LA(T, S_, B)                                                                                        \
    launch_align_wq_t<T, S_, B>(g, pts, scans, nwg, prm, states, q, items, cap, done_cnt, partials, out, prof, \
                                cooperative, stream)
    NDT_DISPATCH(block_size, prm.point_fp64, prm.sum_mode, NDT_LA);
#undef NDT_LA
}

/* ------------------------------------------------------------------ fitness score
 * Mean squared distance from the transformed scan points to their nearest map
 * points, over the points whose nearest map point lies within max_range (PCL's
 * Registration::getFitnessScore, which Autoware's ndt_matching reports and which
 * the reference declares as ndt_getFitnessScore, src/Cuda/ndt_mapping/kernel.cu:124).
 * The search visits shells of cells at growing Chebyshev distance r around the
 * point's cell and stops once no cell of the next shell can hold a closer point:
 * every point of shell r >= 1 is at least (r - 1) res + dmin away, dmin being the
 * point's distance to the nearest face of its own cell. */

__global__ void k_gather_points(const float4* __restrict__ pts, const uint32_t* __restrict__ idx, int64_t n,
                                float4* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = pts[idx[i]];
}

int launch_gather_points(const float4* pts, const uint32_t* idx, int64_t n, float4* out, hipStream_t stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_gather_points, dim3(grid_for(n, 256)), dim3(256), 0, stream, pts, idx, n, out);
    return (int)hipGetLastError();
}

__device__ inline void fit_cell(const FitGridDev& fg, int64_t c0, int64_t c1, int64_t c2, double x0, double x1, double x2,
                                double& best) {
    c0 -= fg.origin_cell[0];
    c1 -= fg.origin_cell[1];
    c2 -= fg.origin_cell[2];
    if (c0 < 0 || c1 < 0 || c2 < 0 || c0 >= (1ll << fg.bits[0]) || c1 >= (1ll << fg.bits[1]) || c2 >= (1ll << fg.bits[2]))
        return;
    const uint64_t key = ((uint64_t)c2 << (fg.bits[0] + fg.bits[1])) | ((uint64_t)c1 << fg.bits[0]) | (uint64_t)c0;
    uint32_t b, n;
    if (!hash_find(fg.slots, fg.log2cap, key, b, n)) return;
    for (uint32_t k = 0; k < n; k++) {
        const float4 m = fg.pts[b + k];
        const double d0 = x0 - (double)m.x, d1 = x1 - (double)m.y, d2 = x2 - (double)m.z;
        const double d = d0 * d0 + d1 * d1 + d2 * d2;
        best = d < best ? d : best;
    }
}

__global__ __launch_bounds__(256) void k_fitness(FitGridDev fg, const float4* __restrict__ pts,
                                                 const int64_t* __restrict__ offsets,
                                                 const int32_t* __restrict__ counts, const double* __restrict__ poses,
                                                 double max_d2, int shells, double* __restrict__ partials) {
    __shared__ double red[2][256 / WAVE];
    const int s = blockIdx.y;
    const int64_t off = offsets[s];
    const int n = counts[s];
    const double* P = poses + 16 * (int64_t)s;
    const double res = fg.resolution;
    double sum = 0.0, cnt = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float4 p = pts[off + i];
        const double X0 = p.x, X1 = p.y, X2 = p.z;
        const double x0 = P[0] * X0 + P[1] * X1 + P[2] * X2 + P[3];
        const double x1 = P[4] * X0 + P[5] * X1 + P[6] * X2 + P[7];
        const double x2 = P[8] * X0 + P[9] * X1 + P[10] * X2 + P[11];
        if (!(fabs(x0) < 1e15) || !(fabs(x1) < 1e15) || !(fabs(x2) < 1e15)) continue;
        const int64_t a0 = (int64_t)floor(x0 / res), a1 = (int64_t)floor(x1 / res), a2 = (int64_t)floor(x2 / res);
        double dmin = fmin(fmin(x0 - a0 * res, (a0 + 1) * res - x0),
                           fmin(fmin(x1 - a1 * res, (a1 + 1) * res - x1), fmin(x2 - a2 * res, (a2 + 1) * res - x2)));
        dmin = fmax(dmin, 0.0);
        double best = INFINITY;
        fit_cell(fg, a0, a1, a2, x0, x1, x2, best);
        for (int r = 1; r <= shells; r++) {
            const double lb = (r - 1) * res + dmin;
            if (lb * lb >= best) break;
            for (int dz = -r; dz <= r; dz++)
                for (int dy = -r; dy <= r; dy++) {
                    const bool face = (dz == -r || dz == r || dy == -r || dy == r);
                    for (int dx = -r; dx <= r; dx += face ? 1 : 2 * r)
                        fit_cell(fg, a0 + dx, a1 + dy, a2 + dz, x0, x1, x2, best);
                }
        }
        if (best <= max_d2) {
            sum += best;
            cnt += 1.0;
        }
    }
    sum = wave_sum(sum);
    cnt = wave_sum(cnt);
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    if (lane == 0) {
        red[0][wave] = sum;
        red[1][wave] = cnt;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double t = 0.0;
        for (int w = 0; w < 256 / WAVE; w++) t += red[threadIdx.x][w];
        partials[((int64_t)s * gridDim.x + blockIdx.x) * 2 + threadIdx.x] = t;
    }
}

int launch_fitness(FitGridDev fg, const float4* pts, const int64_t* offsets, const int32_t* counts, int32_t S,
                   const double* poses, double max_d2, int shells, int nblk, double* partials, hipStream_t stream) {
    if (S <= 0) return 0;
    hipLaunchKernelGGL(k_fitness, dim3(nblk, S), dim3(256), 0, stream, fg, pts, offsets, counts, poses, max_d2, shells,
                       partials);
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------ single derivative pass (ndt) */

template <typename T, int SUM>
__global__ __launch_bounds__(256, (sizeof(T) == 8 || SUM == NDT_SUM_FP64) ? ALIGN_MIN_WAVES_F64 : ALIGN_MIN_WAVES_F32) void k_deriv_ndt(NdtGridDev g, const float4* __restrict__ pts, int64_t n,
                                                   const double* __restrict__ pose, AlignParams prm,
                                                   double* __restrict__ partials) {
    __shared__ PassParams<T> pp;
    __shared__ double lds_red[256 / WAVE][NRED];
    __shared__ double red[NRED];
    if (threadIdx.x == 0) {
        double x[6];
        for (int k = 0; k < 6; k++) x[k] = pose[k];
        make_pass_params<T>(x, prm, g.resolution_d, pp);
    }
    __syncthreads();
    Acc<T, SUM> acc;
    acc.zero();
    int pairs = 0;
    const PassParams<T> u = load_uniform(pp);
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        pairs += point_accumulate<T, SUM>(g, pts[i], u, acc);
    double v[NRED];
#pragma unroll
    for (int k = 0; k < NACC; k++) v[k] = acc.get(k);
    v[NACC] = (double)pairs;
    block_reduce_acc<256>(v, lds_red, red);
    if (threadIdx.x < NRED) partials[(int64_t)blockIdx.x * NRED + threadIdx.x] = red[threadIdx.x];
}

int launch_deriv_ndt(NdtGridDev grid, const float4* pts, int64_t n, const double* pose_params, AlignParams prm,
                     double* partials, int nblocks, hipStream_t stream) {
    if (prm.point_fp64)
        hipLaunchKernelGGL((k_deriv_ndt<double, NDT_SUM_FP64>), dim3(nblocks), dim3(256), 0, stream, grid, pts, n,
                           pose_params, prm, partials);
    else if (prm.sum_mode == NDT_SUM_FP32)
        hipLaunchKernelGGL((k_deriv_ndt<float, NDT_SUM_FP32>), dim3(nblocks), dim3(256), 0, stream, grid, pts, n,
                           pose_params, prm, partials);
    else if (prm.sum_mode == NDT_SUM_COMP32)
        hipLaunchKernelGGL((k_deriv_ndt<float, NDT_SUM_COMP32>), dim3(nblocks), dim3(256), 0, stream, grid, pts, n,
                           pose_params, prm, partials);
    else
        hipLaunchKernelGGL((k_deriv_ndt<float, NDT_SUM_FP64>), dim3(nblocks), dim3(256), 0, stream, grid, pts, n,
                           pose_params, prm, partials);
    return (int)hipGetLastError();
}

/* out[k] = sum over b of partials[b][k], one wave per k in a fixed order: lane l
 * adds b = l, l + 64, ... in turn, then the lanes are added by wave_sum. */
__global__ __launch_bounds__(WAVE) void k_reduce_partials(const double* __restrict__ partials, int nblocks, int width,
                                                          double* __restrict__ out) {
    const int k = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += WAVE) s += partials[(int64_t)b * width + k];
    s = wave_sum(s);
    if (threadIdx.x == 0) out[k] = s;
}

int launch_reduce_partials(const double* partials, int nblocks, int width, double* out, hipStream_t stream) {
    if (width <= 0) return 0;
    hipLaunchKernelGGL(k_reduce_partials, dim3(width), dim3(WAVE), 0, stream, partials, nblocks, width, out);
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------ compat derivative pass
 * voxelRadiusSearch + computePointDerivatives + updateDerivatives
 * (src/CPU/ndt_mapping/kernel.cpp:246-286, :438-560) per scan point, in float64
 * with the reference's operation order and its Hessian term. */

struct CompatPassArgs {
    float M[16];
    double jh[23][3];
    double d1, d2;
};

__device__ inline double cdot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ inline bool compat_lookup(const CompatGridDev& g, uint64_t key, uint32_t& vox) {
    uint32_t count;
    return hash_find(g.slots, g.log2cap, key, vox, count);
}

__global__ __launch_bounds__(256) void k_deriv_compat(CompatGridDev g, const float4* __restrict__ scan, int64_t n,
                                                      CompatPassArgs a, double* __restrict__ partials) {
    __shared__ double lds[256 / WAVE][NRED_COMPAT];
    double acc[NRED_COMPAT];
    for (int k = 0; k < NRED_COMPAT; k++) acc[k] = 0.0;
    const double radius = (double)g.resolution;
    /* which angular second-derivative vector sits in block (i, j), i, j >= 3
     * (point_hessian_ layout, kernel.cpp:528-558): a b c / b d e / c e f */
    const int hsel[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    /* sixteen threads per point, one per (z, y) probe row of its search (at most
     * four steps per axis), so that a small scan still spreads over the GPU */
    for (int64_t t = blockIdx.x * 256ll + threadIdx.x; t < 16 * n; t += (int64_t)gridDim.x * 256) {
        const int64_t idx = t >> 4;
        const int kzs = (int)(t >> 2) & 3, kys = (int)t & 3;
        const float4 p4 = scan[idx];
        const float in[3] = {p4.x, p4.y, p4.z};
        float pt[3];
        for (int row = 0; row < 3; row++)
            pt[row] = a.M[row * 4 + 0] * in[0] + a.M[row * 4 + 1] * in[1] + a.M[row * 4 + 2] * in[2] + a.M[row * 4 + 3];
        const double x[3] = {in[0], in[1], in[2]};
        double pg[3][6];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 6; c++) pg[r][c] = 0.0;
        pg[0][0] = 1.0;
        pg[1][1] = 1.0;
        pg[2][2] = 1.0;
        pg[1][3] = cdot(x, a.jh[0]);
        pg[2][3] = cdot(x, a.jh[1]);
        pg[0][4] = cdot(x, a.jh[2]);
        pg[1][4] = cdot(x, a.jh[3]);
        pg[2][4] = cdot(x, a.jh[4]);
        pg[0][5] = cdot(x, a.jh[5]);
        pg[1][5] = cdot(x, a.jh[6]);
        pg[2][5] = cdot(x, a.jh[7]);
        double hv[6][3];
        hv[0][0] = 0; hv[0][1] = cdot(x, a.jh[8]); hv[0][2] = cdot(x, a.jh[9]);
        hv[1][0] = 0; hv[1][1] = cdot(x, a.jh[10]); hv[1][2] = cdot(x, a.jh[11]);
        hv[2][0] = 0; hv[2][1] = cdot(x, a.jh[12]); hv[2][2] = cdot(x, a.jh[13]);
        hv[3][0] = cdot(x, a.jh[14]); hv[3][1] = cdot(x, a.jh[15]); hv[3][2] = cdot(x, a.jh[16]);
        hv[4][0] = cdot(x, a.jh[17]); hv[4][1] = cdot(x, a.jh[18]); hv[4][2] = cdot(x, a.jh[19]);
        hv[5][0] = cdot(x, a.jh[20]); hv[5][1] = cdot(x, a.jh[21]); hv[5][2] = cdot(x, a.jh[22]);
        const float radiusFinal = (float)(radius + 0.001f);
        /* the probe loops step in float like the reference; the step counters only
         * bound them for coordinates so large that adding the resolution no longer
         * changes the float (where the reference would not terminate).  This thread
         * takes step kzs of the z loop and step kys of the y loop, reached by the
         * same sequence of float additions, if the loops get that far. */
        float z = (float)(pt[2] - radius);
        bool zin = z <= pt[2] + radiusFinal;
        for (int k = 0; k < kzs && zin; k++) {
            z += g.resolution;
            zin = z <= pt[2] + radiusFinal;
        }
        float y = (float)(pt[1] - radius);
        bool yin = y <= pt[1] + radiusFinal;
        for (int k = 0; k < kys && yin; k++) {
            y += g.resolution;
            yin = y <= pt[1] + radiusFinal;
        }
        if (zin && yin) {
            {
                int kx = 0;
                for (float xx = (float)(pt[0] - radius); xx <= pt[0] + radiusFinal && kx < 4; xx += g.resolution, kx++) {
                    if ((xx < g.minv[0]) || (xx > g.maxv[0]) || (y < g.minv[1]) || (y > g.maxv[1]) || (z < g.minv[2]) ||
                        (z > g.maxv[2]))
                        continue;
                    int lin = compat_linearize(g.minv, g.dims, g.resolution, xx, y, z);
                    uint32_t v;
                    if (!compat_lookup(g, (uint64_t)(uint32_t)lin, v)) continue; /* empty cell: NaN mean */
                    const double* c = g.mean + 3 * v;
                    float dx = (float)(c[0] - pt[0]);
                    float dy = (float)(c[1] - pt[1]);
                    float dz = (float)(c[2] - pt[2]);
                    float dist = (float)sqrt((double)(dx * dx + dy * dy + dz * dz));
                    if (!(dist < radius)) continue;
                    acc[NACC_COMPAT] += 1.0;
                    double xt[3] = {pt[0], pt[1], pt[2]};
                    xt[0] -= c[0];
                    xt[1] -= c[1];
                    xt[2] -= c[2];
                    const double* ci = g.icov + 9 * v;
                    double xCx = ci[0] * xt[0] * xt[0] + ci[4] * xt[1] * xt[1] + ci[8] * xt[2] * xt[2] +
                                 (ci[1] + ci[3]) * xt[0] * xt[1] + (ci[2] + ci[6]) * xt[0] * xt[2] +
                                 (ci[5] + ci[7]) * xt[1] * xt[2];
                    double e = exp(-a.d2 * (xCx) / 2);
                    double score_inc = -a.d1 * e;
                    e = a.d2 * e;
                    if (e > 1 || e < 0 || e != e) continue;
                    acc[0] += score_inc;
                    e *= a.d1;
                    for (int i = 0; i < 6; i++) {
                        double cov[3];
                        for (int r = 0; r < 3; r++) {
                            cov[r] = 0;
                            for (int cc = 0; cc < 3; cc++) cov[r] += ci[r * 3 + cc] * pg[cc][i];
                        }
                        double xc = cdot(xt, cov);
                        acc[1 + i] += xc * e;
                        for (int j = 0; j < 6; j++) {
                            double colVec[3] = {pg[0][j], pg[1][j], pg[2][j]};
                            double hb[3] = {0.0, 0.0, 0.0};
                            if (i >= 3 && j >= 3) {
                                const int hs = hsel[i - 3][j - 3];
                                hb[0] = hv[hs][0];
                                hb[1] = hv[hs][1];
                                hb[2] = hv[hs][2];
                            }
                            double cvh[3] = {colVec[0] + hb[0], colVec[1] + hb[1], colVec[2] + hb[2]};
                            double mp[3];
                            for (int r = 0; r < 3; r++) {
                                mp[r] = 0;
                                for (int cc = 0; cc < 3; cc++) mp[r] += ci[r * 3 + cc] * cvh[cc];
                            }
                            acc[7 + i * 6 + j] += e * (-a.d2 * xc * cdot(xt, mp) + cdot(colVec, cov));
                        }
                    }
                }
            }
        }
    }
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    for (int k = 0; k < NRED_COMPAT; k++) {
        double v = wave_sum(acc[k]);
        if (lane == 0) lds[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < NRED_COMPAT) {
        double s = 0.0;
        for (int w = 0; w < 256 / WAVE; w++) s += lds[w][threadIdx.x];
        partials[(int64_t)blockIdx.x * NRED_COMPAT + threadIdx.x] = s;
    }
}

int launch_deriv_compat(CompatGridDev grid, const float4* scan, int64_t n, const float* M16, const double* jh,
                        double d1, double d2, double* partials, int nblocks, hipStream_t stream) {
    CompatPassArgs a;
    for (int k = 0; k < 16; k++) a.M[k] = M16[k];
    for (int k = 0; k < 23; k++)
        for (int c = 0; c < 3; c++) a.jh[k][c] = jh[k * 3 + c];
    a.d1 = d1;
    a.d2 = d2;
    hipLaunchKernelGGL(k_deriv_compat, dim3(nblocks), dim3(256), 0, stream, grid, scan, n, a, partials);
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------ allocation-free alignment
 * Workspace of ndt_align_scans_noalloc, the entry point for callers that cannot
 * allocate device memory.  It is a file-scope device array: every device the
 * library runs on holds its own copy from the loading of the code object until the
 * process ends.  On the host the array's name stands for a host-side shadow, not
 * for device memory, so host code never passes it to a kernel or a copy:
 * noalloc_scratch() looks up the device address once per device (with that device
 * current) and every launch and copy of the entry point uses that address. */
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES];

int noalloc_scratch(int device, void** ptr) {
    *ptr = nullptr;
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return (int)hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return 0;
}

/* Inclusive prefix sum over the 64 lanes of a wave. */
__device__ inline int64_t wave_inclusive_scan(int64_t v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const int64_t u = __shfl_up(v, d, WAVE);
        if (lane >= d) v += u;
    }
    return v;
}

constexpr int DESC_BLOCK = 1024;

/* Chunk descriptors of a device-resident batch from its point counts, as the host
 * makes them for the other entry points: every scan cut into `chunk`-point chunks
 * (and fine_chunk(chunk)-point chunks for late passes), points, partial-sum slots
 * and first-pass queue items laid out in scan order.  One workgroup; exclusive
 * prefix sums tile by tile with a running carry. */
__global__ __launch_bounds__(DESC_BLOCK) void k_make_descs(const int64_t* __restrict__ counts, int32_t S,
                                                           int32_t chunk, int32_t fine_pass,
                                                           ScanDesc* __restrict__ desc) {
    __shared__ int64_t wsum[3][DESC_BLOCK / WAVE];
    __shared__ int64_t carry[3];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    const int32_t fine = fine_chunk(chunk);
    if (tid < 3) carry[tid] = 0;
    __syncthreads();
    for (int32_t base = 0; base < S; base += DESC_BLOCK) {
        const int32_t s = base + tid;
        const int64_t n = s < S ? counts[s] : 0;
        int64_t v[3] = {n, (n + fine - 1) / fine, (n + chunk - 1) / chunk};
        int64_t inc[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            inc[k] = wave_inclusive_scan(v[k]);
            if (lane == WAVE - 1) wsum[k][wave] = inc[k];
        }
        __syncthreads();
        int64_t ex[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            int64_t before = carry[k];
            for (int w = 0; w < wave; w++) before += wsum[k][w];
            ex[k] = before + inc[k] - v[k];
        }
        if (s < S) {
            ScanDesc d;
            d.offset = ex[0];
            d.n = (int32_t)n;
            d.chunk = chunk;
            d.nchunks = (int32_t)v[2];
            d.fnchunks = (int32_t)v[1];
            d.chunk_base = ex[1];
            d.qbase = ex[2];
            d.fine_pass = fine_pass;
            d.pad = 0;
            desc[s] = d;
        }
        __syncthreads();
        if (tid < 3) {
            int64_t t = 0;
            for (int w = 0; w < DESC_BLOCK / WAVE; w++) t += wsum[tid][w];
            carry[tid] += t;
        }
        __syncthreads();
    }
}

int launch_make_descs(const int64_t* counts, int32_t S, int32_t chunk, int32_t fine_pass, ScanDesc* desc,
                      hipStream_t stream) {
    hipLaunchKernelGGL(k_make_descs, dim3(1), dim3(DESC_BLOCK), 0, stream, counts, S, chunk, fine_pass, desc);
    return (int)hipGetLastError();
}

/* Results of a finished launch in the C ABI's layout (the host conversion of the
 * other entry points, on the device): a scan that never finished in an aborted
 * launch is marked as timed out; prof[5] receives the pairs of all scans. */
__global__ __launch_bounds__(DESC_BLOCK) void k_noalloc_results(const ScanOut* __restrict__ out,
                                                                const ScanDesc* __restrict__ desc, int32_t S,
                                                                double* prof, ndt_result* __restrict__ res) {
    __shared__ double wp[DESC_BLOCK / WAVE];
    const int tid = threadIdx.x;
    const bool aborted = prof[4] != 0.0;
    double pairs = 0.0;
    for (int32_t s = tid; s < S; s += DESC_BLOCK) {
        const ScanOut o = out[s];
        ndt_result r;
        for (int k = 0; k < 16; k++) {
            r.final_transformation_f64[k] = o.final_m[k];
            r.final_transformation[k] = (float)o.final_m[k];
        }
        r.fitness_score = 0.0;
        r.score = o.score;
        r.trans_probability = desc[s].n > 0 ? o.score / (double)desc[s].n : 0.0;
        r.iterations = o.iterations;
        r.converged = o.converged;
        r.passes = o.passes;
        r.status = (aborted && o.final_m[15] == 0.0) ? NDT_ERR_TIMEOUT : o.status;
        res[s] = r;
        pairs += o.pairs;
    }
    pairs = wave_sum(pairs);
    if ((tid & (WAVE - 1)) == 0) wp[tid / WAVE] = pairs;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < DESC_BLOCK / WAVE; w++) t += wp[w];
        prof[5] = t;
    }
}

int launch_noalloc_results(const ScanOut* out, const ScanDesc* desc, int32_t S, double* prof, ndt_result* res,
                           hipStream_t stream) {
    hipLaunchKernelGGL(k_noalloc_results, dim3(1), dim3(DESC_BLOCK), 0, stream, out, desc, S, prof, res);
    return (int)hipGetLastError();
}

}  // namespace ndt
