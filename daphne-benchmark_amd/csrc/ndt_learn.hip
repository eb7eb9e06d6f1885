/*
 * ndt_learn.hip - gfx950 kernels of the learned warm start.
 *
 *  features   bird's-eye-view occupancy histograms in LDS: the scan in its sensor
 *             frame, the map around a guess in the guess frame (map points
 *             bucketed by 4 m tiles), log1p counts through a bf16 lookup table
 *  model      bf16 GEMM C = A B^T on the matrix cores (v_mfma_f32_16x16x32_bf16,
 *             128 x 128 tiles, 64-deep K steps double-buffered through LDS,
 *             split-K with float32 partials), bias + ReLU epilogues that also
 *             write the transposed activations the backward GEMMs read, the
 *             6-wide output layer, its loss and backward, Adam
 *
 * All reductions run in a fixed order, so training is deterministic.  There is no
 * counterpart in the reference: the guess it refines enters at
 * src/CPU/ndt_mapping/kernel.cpp:1187-1206.
 */
#include <hip/hip_runtime.h>

#include "ndt_internal.h"
#include "ndt_learn_internal.h"

namespace ndt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

static inline int blocks_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

/* float32 -> bfloat16, round to nearest even (NaN stays NaN) */
__device__ inline uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

/* ------------------------------------------------------------------ features */

/* BEV cell of a point given in the grid's own frame, or -1 outside the grid
 * (NaN coordinates fall outside). */
__device__ inline int bev_index(double qx, double qy, double qz) {
    const double fx = floor((qx - BEV_X0) * 2.0);
    const double fy = floor((qy - BEV_Y0) * 2.0);
    if (!(fx >= 0.0 && fx < (double)BEV_NX && fy >= 0.0 && fy < (double)BEV_NY)) return -1;
    /* the division only for points inside the footprint */
    const double fz = floor((qz - BEV_Z0) / BEV_DZ);
    if (!(fz >= 0.0 && fz < (double)BEV_NZ)) return -1;
    return ((int)fz * BEV_NY + (int)fy) * BEV_NX + (int)fx;
}

constexpr int BEV_BLOCK = 1024;
constexpr int BEV_HALF = BEV_CELLS / 2; /* z planes 0-3, then 4-7 */
/* points one workgroup may count with 16-bit counters: no count can then carry */
constexpr int64_t BEV_PACKED_MAX = 65535;

/* A grid's histogram lives in 64 KiB of LDS (BEV_HALF words).  With at most
 * BEV_PACKED_MAX points (every scan and map crop of the benchmark) the counters are
 * the 16-bit halves of the words, cell c in the low half of word c and cell
 * c + BEV_HALF in the high half (cells that share a word lie four z planes apart,
 * so that nearby points rarely contend for a word), and the points are read once.
 * Larger inputs take two passes over the points with exact 32-bit counters, one
 * pass per half of the cells.  Counts saturate at 65535 in the output table. */
__device__ inline void bev_clear(uint32_t* hist) {
    for (int i = threadIdx.x; i < BEV_HALF; i += BEV_BLOCK) hist[i] = 0u;
}
__device__ inline void bev_add(uint32_t* hist, const int c, const bool packed, const int half) {
    if (packed) {
        if (c >= 0) atomicAdd(&hist[c & (BEV_HALF - 1)], c >= BEV_HALF ? 65536u : 1u);
    } else {
        const int h = c - half * BEV_HALF;
        if (c >= 0 && h >= 0 && h < BEV_HALF) atomicAdd(&hist[h], 1u);
    }
}
__device__ inline void bev_emit(const uint32_t* hist, const bool packed, const int half,
                                const uint16_t* __restrict__ lut, uint16_t* __restrict__ out) {
    for (int i = threadIdx.x; i < BEV_HALF; i += BEV_BLOCK) {
        const uint32_t w = hist[i];
        if (packed) {
            out[i] = lut[w & 0xffffu];
            out[BEV_HALF + i] = lut[w >> 16];
        } else {
            out[half * BEV_HALF + i] = lut[w > 65535u ? 65535u : w];
        }
    }
}

/* One workgroup per scan: histogram of the scan's points in the sensor frame. */
__global__ __launch_bounds__(BEV_BLOCK) void k_bev_scan(const float4* __restrict__ pts,
                                                        const int64_t* __restrict__ offsets,
                                                        const int32_t* __restrict__ counts,
                                                        const uint16_t* __restrict__ lut, uint16_t* __restrict__ out) {
    __shared__ uint32_t hist[BEV_HALF];
    const int s = blockIdx.x;
    const float4* p = pts + offsets[s];
    const int n = counts[s];
    uint16_t* o = out + (int64_t)s * BEV_CELLS;
    const bool packed = n <= BEV_PACKED_MAX;
    for (int half = 0; half < (packed ? 1 : 2); half++) {
        bev_clear(hist);
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += BEV_BLOCK) {
            const float4 q = p[i];
            bev_add(hist, bev_index((double)q.x, (double)q.y, (double)q.z), packed, half);
        }
        __syncthreads();
        bev_emit(hist, packed, half, lut, o);
        __syncthreads();
    }
}

int launch_bev_scan(const float4* pts, const int64_t* offsets, const int32_t* counts, int32_t S, const uint16_t* lut,
                    uint16_t* out, hipStream_t stream) {
    if (S <= 0) return 0;
    hipLaunchKernelGGL(k_bev_scan, dim3(S), dim3(BEV_BLOCK), 0, stream, pts, offsets, counts, lut, out);
    return (int)hipGetLastError();
}

/* One workgroup per guess: histogram of the map points around the guess, in the
 * guess frame q = R^T (p - t).  Only the tiles under the grid's footprint are
 * visited: the footprint's corners at both z limits, mapped to the map frame,
 * bound the tile rows and columns. */
__global__ __launch_bounds__(BEV_BLOCK) void k_bev_crop(TileGridDev tg, const double* __restrict__ guesses,
                                                        const uint16_t* __restrict__ lut, uint16_t* __restrict__ X,
                                                        int64_t ldx, int64_t col_off) {
    __shared__ uint32_t hist[BEV_HALF];
    __shared__ int s_rng[4];
    __shared__ int64_t s_n;
    const int b = blockIdx.x;
    const double* G = guesses + 16 * (int64_t)b;
    const double r00 = G[0], r01 = G[1], r02 = G[2], t0 = G[3];
    const double r10 = G[4], r11 = G[5], r12 = G[6], t1 = G[7];
    const double r20 = G[8], r21 = G[9], r22 = G[10], t2 = G[11];
    if (threadIdx.x == 0) {
        double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
        for (int k = 0; k < 8; k++) {
            const double qx = (k & 1) ? -BEV_X0 : BEV_X0;
            const double qy = (k & 2) ? -BEV_Y0 : BEV_Y0;
            const double qz = (k & 4) ? BEV_Z0 + BEV_NZ * BEV_DZ : BEV_Z0;
            const double px = r00 * qx + r01 * qy + r02 * qz + t0;
            const double py = r10 * qx + r11 * qy + r12 * qz + t1;
            xmin = fmin(xmin, px);
            xmax = fmax(xmax, px);
            ymin = fmin(ymin, py);
            ymax = fmax(ymax, py);
        }
        /* tile range, clamped to the tile grid; empty when the footprint misses it */
        double fx0 = floor((xmin - 0.5 - tg.x0) / MAP_TILE), fx1 = floor((xmax + 0.5 - tg.x0) / MAP_TILE);
        double fy0 = floor((ymin - 0.5 - tg.y0) / MAP_TILE), fy1 = floor((ymax + 0.5 - tg.y0) / MAP_TILE);
        int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
        if (fx1 >= 0.0 && fx0 < (double)tg.ntx && fy1 >= 0.0 && fy0 < (double)tg.nty) {
            tx0 = (int)fmax(fx0, 0.0);
            tx1 = (int)fmin(fx1, (double)(tg.ntx - 1));
            ty0 = (int)fmax(fy0, 0.0);
            ty1 = (int)fmin(fy1, (double)(tg.nty - 1));
        }
        /* points visited: a bound on every cell's count */
        int64_t n = 0;
        for (int ty = ty0; ty <= ty1 && tx0 <= tx1; ty++)
            n += (int64_t)tg.start[(int64_t)ty * tg.ntx + tx1 + 1] - (int64_t)tg.start[(int64_t)ty * tg.ntx + tx0];
        s_rng[0] = tx0;
        s_rng[1] = tx1;
        s_rng[2] = ty0;
        s_rng[3] = ty1;
        s_n = n;
    }
    __syncthreads();
    const int tx0 = s_rng[0], tx1 = s_rng[1], ty0 = s_rng[2], ty1 = s_rng[3];
    const bool packed = s_n <= BEV_PACKED_MAX;
    uint16_t* o = X + (int64_t)b * ldx + col_off;
    for (int half = 0; half < (packed ? 1 : 2); half++) {
        bev_clear(hist);
        __syncthreads();
        for (int ty = ty0; ty <= ty1 && tx0 <= tx1; ty++) {
            const uint32_t pb = tg.start[(int64_t)ty * tg.ntx + tx0];
            const uint32_t pe = tg.start[(int64_t)ty * tg.ntx + tx1 + 1];
            for (uint32_t i = pb + threadIdx.x; i < pe; i += BEV_BLOCK) {
                const float4 p = tg.pts[i];
                const double dx = (double)p.x - t0, dy = (double)p.y - t1, dz = (double)p.z - t2;
                const double qx = r00 * dx + r10 * dy + r20 * dz;
                const double qy = r01 * dx + r11 * dy + r21 * dz;
                const double qz = r02 * dx + r12 * dy + r22 * dz;
                bev_add(hist, bev_index(qx, qy, qz), packed, half);
            }
        }
        __syncthreads();
        bev_emit(hist, packed, half, lut, o);
        __syncthreads();
    }
}

int launch_bev_crop(TileGridDev tg, const double* guesses, int32_t B, const uint16_t* lut, uint16_t* X, int64_t ldx,
                    int64_t col_off, hipStream_t stream) {
    if (B <= 0) return 0;
    hipLaunchKernelGGL(k_bev_crop, dim3(B), dim3(BEV_BLOCK), 0, stream, tg, guesses, lut, X, ldx, col_off);
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------ registration */

constexpr int WAVE = 64;
constexpr int REG_WORDS = BEV_CELLS / 32;        /* occupancy bits of one grid as 32-bit words */
constexpr int REG_ROWS = BEV_NZ * BEV_NY;        /* 64-cell rows (z, y) of a grid */
constexpr int REG_SIDE = 2 * REG_MAX_RADIUS + 1; /* shifts per axis at the largest radius */
constexpr int REG_UNROLL = 4;                    /* map points a thread loads at once */
constexpr int REG_TROWS = 32;                    /* tile rows indexed together */
/* height search: 5 cm bins over [-4, 12) m (bin floor((z + 4) * 20)), shifts of
 * -1.5 .. 1.5 m; each wave counts into its own copy of a histogram (the ground's
 * few bins take most points), the copies summed before the scores */
constexpr int ZH_BINS = 320, ZH_RANGE = 30, ZH_COPIES = REG_BLOCK / WAVE;
constexpr double ZH_Z0 = -4.0, ZH_BIN = 0.05, ZH_PER_M = 20.0;

/* height bin of z, or -1 outside (NaN included) */
__device__ inline int zh_bin(double z) {
    const double f = floor((z - ZH_Z0) * ZH_PER_M);
    return (f >= 0.0 && f < (double)ZH_BINS) ? (int)f : -1;
}

/* Offset of the vertex of the parabola through (-1, a), (0, b), (1, c) when it
 * opens downwards, else 0; with b the largest of the three it lies in [-0.5, 0.5]. */
__device__ inline double reg_parabola(double a, double b, double c) {
    const double d = a - 2.0 * b + c;
    return d < 0.0 ? 0.5 * (a - c) / d : 0.0;
}

/* Cosine and sine of a small angle (|d| <= 0.025 rad) from their Taylor series up
 * to d^8 / 8! and d^7 / 7!, term by term (truncation error below 1e-19). */
__device__ inline void reg_small_cos_sin(double d, double& c, double& s) {
    const double d2 = d * d;
    c = 1.0 - d2 / 2.0 + d2 * d2 / 24.0 - d2 * d2 * d2 / 720.0 + d2 * d2 * d2 * d2 / 40320.0;
    s = d - d * d2 / 6.0 + d * d2 * d2 / 120.0 - d * d2 * d2 * d2 / 5040.0;
}

/* One workgroup per (scan, guess) row.  The map points under the footprints of the
 * yaw-rotated grids are binned once into one occupancy bitmap per yaw (LDS): a
 * point's guess-frame position q = R^T (p - t) is turned by each yaw as
 * q_k = (qx + t_k qy, qy - t_k qx, qz), t_k = tan(yaw_k) (the rotation by -yaw_k
 * scaled by 1 / cos(yaw_k), at most 1.0014 for 3 degrees: 2 cm at the grid's
 * edge), so that its cell on x is floor((2 qx + 32) + t_k (2 qy)) and on y
 * floor((2 qy + 32) - t_k (2 qx)): two products and two sums per yaw.  A 64-cell row (z, y) of a bitmap
 * is one 64-bit word (bit x = cell x), so the score of a shift (dx, dy) is a sum
 * of popcounts of scan row AND shifted map row over the scan's non-empty rows.
 * A thread scores one (yaw, dy) over a quarter of the scan's rows for every dx at
 * once (the radius is a template parameter, so the shifts are constants), and the
 * quarters add into the score table.  The best candidate: highest score, then the
 * lowest rank (nearest the guess: |yaw index offset|, |dy|, |dx|, negative before
 * positive), packed with the candidate's index into one 64-bit key whose maximum
 * is taken over the block. */
template <int R>
__global__ __launch_bounds__(REG_BLOCK) void k_bev_register(TileGridDev tg, const uint16_t* __restrict__ sbev,
                                                            const float4* __restrict__ spts,
                                                            const int64_t* __restrict__ soff,
                                                            const int32_t* __restrict__ scnt,
                                                            const int32_t* __restrict__ idx, const double* gl,
                                                            const double* gw, RegisterDev rp, double* gl_out,
                                                            double* gw_out, int32_t* __restrict__ score_out) {
    constexpr int W = 2 * R + 1;
    __shared__ __align__(16) uint32_t occ[REG_MAX_YAWS * REG_WORDS];
    __shared__ uint64_t srow[REG_ROWS];
    __shared__ uint16_t rows[REG_ROWS];
    __shared__ uint32_t sc[REG_MAX_YAWS * REG_SIDE * REG_SIDE];
    __shared__ uint64_t wbest[REG_BLOCK / WAVE];
    __shared__ double s_g[12], s_c[REG_MAX_YAWS], s_s[REG_MAX_YAWS], s_t[REG_MAX_YAWS];
    __shared__ int s_rng[4], s_nrows;
    __shared__ uint32_t s_rbeg[REG_TROWS], s_rcnt[REG_TROWS], s_pre[REG_TROWS + 1];
    __shared__ uint32_t zh[2][ZH_COPIES][ZH_BINS]; /* heights: 0 scan, 1 map */
    __shared__ uint64_t zsc[2 * ZH_RANGE + 1];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const int K = rp.yaws, kc = (K - 1) / 2;
    if (tid < 12) s_g[tid] = gl[16 * (int64_t)b + tid];
    if (tid < K) {
        s_c[tid] = rp.c[tid];
        s_s[tid] = rp.s[tid];
        s_t[tid] = rp.t[tid];
    }
    for (int i = tid; i < K * REG_WORDS; i += REG_BLOCK) occ[i] = 0u;
    for (int i = tid; i < K * W * W; i += REG_BLOCK) sc[i] = 0u;
    const bool zs = rp.z_search != 0;
    for (int i = tid; i < 2 * ZH_COPIES * ZH_BINS; i += REG_BLOCK) (&zh[0][0][0])[i] = 0u;
    /* the scan's occupied cells: a thread reads 8 cells (16 bytes) of a row at a
     * time, all its loads in flight together, and writes their occupancy as byte
     * x / 8 of the row's word (little-endian: bit x of the word is cell x) */
    {
        const uint4* sg = (const uint4*)(sbev + (int64_t)idx[b] * BEV_CELLS);
        constexpr int PER = REG_ROWS * (BEV_NX / 8) / REG_BLOCK;
        uint4 v[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) v[u] = sg[tid + u * REG_BLOCK];
        uint8_t* bytes = (uint8_t*)srow;
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const uint32_t w[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
            uint32_t m8 = 0;
#pragma unroll
            for (int h = 0; h < 4; h++)
                m8 |= (((w[h] & 0xffffu) != 0u) ? 1u : 0u) << (2 * h) | ((w[h] >> 16) != 0u ? 1u : 0u) << (2 * h + 1);
            bytes[tid + u * REG_BLOCK] = (uint8_t)m8;
        }
    }
    __syncthreads();
    if (zs) {
        /* the scan's point heights, eight loads in flight per thread */
        const int s = idx[b];
        const float* sz = (const float*)(spts + soff[s]) + 2;
        const int n = scnt[s];
        uint32_t* h = zh[0][wave];
        for (int i0 = tid; i0 < n; i0 += 8 * REG_BLOCK) {
            float z[8];
#pragma unroll
            for (int u = 0; u < 8; u++) z[u] = sz[4 * (int64_t)min(i0 + u * REG_BLOCK, n - 1)];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int bz = zh_bin((double)z[u]);
                if (i0 + u * REG_BLOCK < n && bz >= 0) atomicAdd(&h[bz], 1u);
            }
        }
    }
    const double r00 = s_g[0], r01 = s_g[1], r02 = s_g[2], t0 = s_g[3];
    const double r10 = s_g[4], r11 = s_g[5], r12 = s_g[6], t1 = s_g[7];
    const double r20 = s_g[8], r21 = s_g[9], r22 = s_g[10], t2 = s_g[11];
    if (tid == 0) {
        /* tiles under the union of the yaw-rotated footprints (corners at both z
         * limits, mapped to the map frame), clamped to the tile grid */
        double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
        for (int k = 0; k < K; k++)
            for (int v = 0; v < 8; v++) {
                const double ax = (v & 1) ? -BEV_X0 : BEV_X0;
                const double ay = (v & 2) ? -BEV_Y0 : BEV_Y0;
                const double qz = (v & 4) ? BEV_Z0 + BEV_NZ * BEV_DZ : BEV_Z0;
                const double qx = s_c[k] * ax - s_s[k] * ay, qy = s_s[k] * ax + s_c[k] * ay;
                const double px = r00 * qx + r01 * qy + r02 * qz + t0;
                const double py = r10 * qx + r11 * qy + r12 * qz + t1;
                xmin = fmin(xmin, px);
                xmax = fmax(xmax, px);
                ymin = fmin(ymin, py);
                ymax = fmax(ymax, py);
            }
        double fx0 = floor((xmin - 0.5 - tg.x0) / MAP_TILE), fx1 = floor((xmax + 0.5 - tg.x0) / MAP_TILE);
        double fy0 = floor((ymin - 0.5 - tg.y0) / MAP_TILE), fy1 = floor((ymax + 0.5 - tg.y0) / MAP_TILE);
        int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
        if (fx1 >= 0.0 && fx0 < (double)tg.ntx && fy1 >= 0.0 && fy0 < (double)tg.nty) {
            tx0 = (int)fmax(fx0, 0.0);
            tx1 = (int)fmin(fx1, (double)(tg.ntx - 1));
            ty0 = (int)fmax(fy0, 0.0);
            ty1 = (int)fmin(fy1, (double)(tg.nty - 1));
        }
        s_rng[0] = tx0;
        s_rng[1] = tx1;
        s_rng[2] = ty0;
        s_rng[3] = ty1;
    }
    if (wave == 0) {
        /* the scan's non-empty rows, in order */
        int n = 0;
        for (int r0 = 0; r0 < REG_ROWS; r0 += WAVE) {
            const bool ne = srow[r0 + lane] != 0;
            const uint64_t bal = __ballot(ne);
            const int pos = n + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            if (ne) rows[pos] = (uint16_t)(r0 + lane);
            n += __popcll(bal);
        }
        if (lane == 0) s_nrows = n;
    }
    __syncthreads();
    const int tx0 = s_rng[0], tx1 = s_rng[1], ty0 = s_rng[2], ty1 = s_rng[3];
    double tk[REG_MAX_YAWS];
#pragma unroll
    for (int k = 0; k < REG_MAX_YAWS; k++) tk[k] = k < K ? s_t[k] : 0.0;
    /* a point with |qx| or |qy| above 16 (1 + max |t_k|) lies outside every yaw's
     * grid; half a metre more keeps the test clear of rounding */
    const double reach = -BEV_X0 * (1.0 + fmax(fabs(s_t[0]), fabs(s_t[K - 1]))) + 0.5;
    /* the points of up to REG_TROWS tile rows at a time as one index range (each
     * row's run of tiles tx0..tx1 is contiguous), so that every thread has points
     * and REG_UNROLL loads in flight however the points spread over the rows */
    for (int rb = ty0; rb <= ty1 && tx0 <= tx1; rb += REG_TROWS) {
        const int nr = min(REG_TROWS, ty1 - rb + 1);
        __syncthreads(); /* the previous rows' table is no longer read */
        if (tid < nr) {
            const int64_t row = (int64_t)(rb + tid) * tg.ntx;
            const uint32_t pb = tg.start[row + tx0];
            s_rbeg[tid] = pb;
            s_rcnt[tid] = tg.start[row + tx1 + 1] - pb;
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t acc = 0;
            for (int r = 0; r < nr; r++) {
                s_pre[r] = acc;
                acc += s_rcnt[r];
            }
            s_pre[nr] = acc;
        }
        __syncthreads();
        const uint32_t total = s_pre[nr];
        int r = 0;
        for (uint32_t i0 = tid; i0 < total; i0 += REG_UNROLL * REG_BLOCK) {
            /* REG_UNROLL points per thread, their loads issued first; the row of an
             * index only moves forward (indices past the end repeat the last) */
            float4 pp[REG_UNROLL];
#pragma unroll
            for (int u = 0; u < REG_UNROLL; u++) {
                const uint32_t gi = min(i0 + u * REG_BLOCK, total - 1);
                while (s_pre[r + 1] <= gi) r++;
                pp[u] = tg.pts[s_rbeg[r] + (gi - s_pre[r])];
            }
#pragma unroll
            for (int u = 0; u < REG_UNROLL; u++) {
                if (i0 + u * REG_BLOCK >= total) continue;
                const float4 p = pp[u];
                const double dx = (double)p.x - t0, dy = (double)p.y - t1, dz = (double)p.z - t2;
                const double qx = r00 * dx + r10 * dy + r20 * dz;
                const double qy = r01 * dx + r11 * dy + r21 * dz;
                const double qz = r02 * dx + r12 * dy + r22 * dz;
                if (!(fabs(qx) <= reach && fabs(qy) <= reach)) continue;
                if (zs && fabs(qx) < -BEV_X0 && fabs(qy) < -BEV_Y0) {
                    /* the map's heights under the grid's footprint (a yaw turn and
                     * a shift in the plane leave them unchanged) */
                    const int bz = zh_bin(qz);
                    if (bz >= 0) atomicAdd(&zh[1][wave][bz], 1u);
                }
                const double fz = floor((qz - BEV_Z0) / BEV_DZ);
                if (!(fz >= 0.0 && fz < (double)BEV_NZ)) continue;
                const int zoff = (int)fz * (BEV_NY * BEV_NX);
                /* a point that passed the z test is finite (a NaN or infinite
                 * coordinate makes qz NaN), so the cells convert exactly or
                 * saturate, and one unsigned compare per axis bounds them */
                const double ax = 2.0 * qx + 32.0, bx = 2.0 * qy, ay = 2.0 * qy + 32.0, by = 2.0 * qx;
#pragma unroll
                for (int k = 0; k < REG_MAX_YAWS; k++) {
                    if (k < K) {
                        const int ix = (int)floor(ax + tk[k] * bx);
                        const int iy = (int)floor(ay - tk[k] * by);
                        if ((unsigned)ix < (unsigned)BEV_NX && (unsigned)iy < (unsigned)BEV_NY) {
                            const int c = zoff + iy * BEV_NX + ix;
                            atomicOr(&occ[k * REG_WORDS + (c >> 5)], 1u << (c & 31));
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    if (zs) {
        /* the histograms' copies into copy 0 */
        for (int i = tid; i < 2 * ZH_BINS; i += REG_BLOCK) {
            uint32_t* h = &zh[i / ZH_BINS][0][i % ZH_BINS];
            uint32_t a = 0;
            for (int c = 0; c < ZH_COPIES; c++) a += h[c * ZH_BINS];
            h[0] = a;
        }
        __syncthreads();
    }
    /* scores of every (yaw, dy, dx): a thread takes one (yaw, dy) over every
     * `parts`-th non-empty scan row and counts all dx at once */
    const int ncand = K * W * W, nrows = s_nrows, nkd = K * W;
    const int parts = nkd >= REG_BLOCK ? 1 : REG_BLOCK / nkd;
    for (int it = tid; it < nkd * parts; it += REG_BLOCK) {
        const int kd = it % nkd, part = it / nkd;
        const int k = kd / W, dy = kd % W - R;
        const uint2* ok = (const uint2*)(occ + k * REG_WORDS);
        uint32_t cnt[W];
#pragma unroll
        for (int q = 0; q < W; q++) cnt[q] = 0u;
        for (int j = part; j < nrows; j += parts) {
            const int r = rows[j];
            const int y = (r & (BEV_NY - 1)) + dy;
            if (y < 0 || y >= BEV_NY) continue;
            const uint2 mw = ok[(r & ~(BEV_NY - 1)) + y];
            const uint64_t m = ((uint64_t)mw.y << 32) | (uint64_t)mw.x;
            const uint64_t sr = srow[r];
#pragma unroll
            for (int q = 0; q < W; q++) {
                const int dx = q - R;
                cnt[q] += (uint32_t)__popcll(sr & (dx >= 0 ? m >> dx : m << (-dx)));
            }
        }
        uint32_t* out = sc + k * W * W + (dy + R) * W;
#pragma unroll
        for (int q = 0; q < W; q++)
            if (cnt[q]) atomicAdd(&out[q], cnt[q]);
    }
    if (zs && tid >= REG_BLOCK - (2 * ZH_RANGE + 1)) {
        /* height scores sum_b scan(b) map(b + k), by the last threads */
        const int j = tid - (REG_BLOCK - (2 * ZH_RANGE + 1)), k = j - ZH_RANGE;
        uint64_t acc = 0;
        for (int q = max(0, -k); q < ZH_BINS - max(0, k); q++) acc += (uint64_t)zh[0][0][q] * zh[1][0][q + k];
        zsc[j] = acc;
    }
    __syncthreads();
    uint64_t best = 0;
    for (int n = tid; n < ncand; n += REG_BLOCK) {
        const int k = n / (W * W), rem = n % (W * W), dy = rem / W - R, dx = rem % W - R;
        const uint32_t rank =
            (((((uint32_t)abs(k - kc) * 2u + (k < kc ? 1u : 0u)) * 16u + (uint32_t)abs(dy)) * 2u + (dy < 0 ? 1u : 0u)) *
                 16u + (uint32_t)abs(dx)) * 2u + (dx < 0 ? 1u : 0u);
        const uint64_t key = ((uint64_t)sc[n] << 32) | ((uint64_t)(0x3fffu - rank) << 12) | (uint64_t)n;
        best = key > best ? key : best;
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(best, off);
        best = o > best ? o : best;
    }
    if (lane == 0) wbest[wave] = best;
    __syncthreads();
    if (tid != 0) return;
    for (int w = 1; w < REG_BLOCK / WAVE; w++) best = wbest[w] > best ? wbest[w] : best;
    const uint32_t v = (uint32_t)(best >> 32);
    const int n = (int)(best & 0xfffu);
    const int k = n / (W * W), rem = n % (W * W), dy = rem / W - R, dx = rem % W - R;
    double* outs[2] = {gl_out + 16 * (int64_t)b, gw_out + 16 * (int64_t)b};
    const double* ins[2] = {gl + 16 * (int64_t)b, gw + 16 * (int64_t)b};
    double G[2][16];
    for (int q = 0; q < 2; q++)
        for (int e = 0; e < 16; e++) G[q][e] = ins[q][e];
    if (v != 0) {
        const uint32_t* S = sc + k * W * W;
        const int i0 = (dy + R) * W + (dx + R);
        const double ox = abs(dx) < R ? reg_parabola((double)S[i0 - 1], (double)S[i0], (double)S[i0 + 1]) : 0.0;
        const double oy = abs(dy) < R ? reg_parabola((double)S[i0 - W], (double)S[i0], (double)S[i0 + W]) : 0.0;
        const double ck = s_c[k], sk = s_s[k];
        double c = ck, s = sk;
        if (rp.refine_yaw && k > 0 && k < K - 1) {
            const double o = reg_parabola((double)sc[i0 + (k - 1) * W * W], (double)S[i0],
                                          (double)sc[i0 + (k + 1) * W * W]);
            double cd, sd;
            reg_small_cos_sin(o * rp.yaw_step, cd, sd);
            c = ck * cd - sk * sd;
            s = sk * cd + ck * sd;
        }
        const double d0 = ((double)dx + ox) * 0.5, d1 = ((double)dy + oy) * 0.5;
        for (int q = 0; q < 2; q++) {
            double* g = G[q];
            for (int r = 0; r < 3; r++) {
                const double a = g[4 * r], bb = g[4 * r + 1];
                /* translation along the axes of the yaw-k frame, then the rotation */
                const double ak = a * ck + bb * sk, bk = bb * ck - a * sk;
                g[4 * r + 3] = g[4 * r + 3] + ak * d0 + bk * d1;
                g[4 * r] = a * c + bb * s;
                g[4 * r + 1] = bb * c - a * s;
            }
        }
    }
    if (zs) {
        /* the best height shift: highest score, then the smallest |k|, negative
         * first; refined by a parabola; applied along the guess's z axis (the yaw
         * turn leaves that column as it is) */
        int jb = ZH_RANGE;
        for (int j = 0; j < 2 * ZH_RANGE + 1; j++) {
            const int k = j - ZH_RANGE, kb = jb - ZH_RANGE;
            const int rk = 2 * abs(k) + (k < 0 ? 1 : 0), rb = 2 * abs(kb) + (kb < 0 ? 1 : 0);
            if (zsc[j] > zsc[jb] || (zsc[j] == zsc[jb] && rk < rb)) jb = j;
        }
        if (zsc[jb] != 0) {
            const double o = (jb > 0 && jb < 2 * ZH_RANGE)
                                 ? reg_parabola((double)zsc[jb - 1], (double)zsc[jb], (double)zsc[jb + 1])
                                 : 0.0;
            const double dz = ((double)(jb - ZH_RANGE) + o) * ZH_BIN;
            for (int q = 0; q < 2; q++)
                for (int r = 0; r < 3; r++) G[q][4 * r + 3] = G[q][4 * r + 3] + G[q][4 * r + 2] * dz;
        }
    }
    for (int q = 0; q < 2; q++)
        for (int e = 0; e < 16; e++) outs[q][e] = G[q][e];
    if (score_out) score_out[b] = (int32_t)v;
}

/* ------------------------------------------------------------------ leveling
 * One 1024-thread workgroup per registered guess, one thread per 1 m cell of
 * [-16, 16)^2 (cell (j, i) = (floor(y + 16), floor(x + 16)), index j * 32 + i).
 * The lowest scan point (sensor frame) and the lowest map point in the guess frame
 * (q = R^T (p - t), as the registration computes it) of each cell, heights in
 * [-8, 8) only, are kept in LDS as order-preserving 64-bit keys (atomicMin, issued
 * only by a point below the cell's current lowest: most points of the dense cells
 * near the sensor then skip the contended atomic).  Where
 * a cell has both, d = map - scan.  A bitonic sort gives the lower median of d;
 * the cells within 0.25 m of it enter the least-squares plane d = a x + b y + c
 * (cell centres x = i - 15.5, y = j - 15.5): nine sums, each by a butterfly over
 * a wave's 64 cells and then over the 16 waves in order, and Cramer's rule on one
 * lane.  A rotation of the guess by Rx(roll) Ry(pitch) lifts the map in its frame
 * by about pitch x - roll y, so the guess turns by (Rx(-b) Ry(a))^-1, both angles'
 * cosines and sines from the series of the registration's yaw refinement.  The
 * guess stays as it is with fewer than 32 cells in common or kept, a singular
 * system, or an angle beyond 0.035 rad. */

constexpr int LEV_N = 32, LEV_CELLS = LEV_N * LEV_N, LEV_MIN_CELLS = 32, LEV_SUMS = 9, LEV_UNROLL = 4;
constexpr double LEV_R = 16.0, LEV_Z0 = -8.0, LEV_Z1 = 8.0, LEV_TRIM = 0.25, LEV_MAX_ANGLE = 0.035;
constexpr unsigned long long LEV_EMPTY = ~0ull;
static_assert(LEV_CELLS == REG_BLOCK, "one cell per thread");

/* order-preserving key of a double: a < b <=> key(a) < key(b) (no NaN) */
__device__ inline unsigned long long lev_key(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ inline double lev_value(unsigned long long k) {
    const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}
/* cell of a guess-frame point, -1 outside the window or its height range (NaN too) */
__device__ inline int lev_cell(double x, double y, double z) {
    const double fi = floor(x + LEV_R), fj = floor(y + LEV_R);
    if (!(fi >= 0.0 && fi < (double)LEV_N && fj >= 0.0 && fj < (double)LEV_N && z >= LEV_Z0 && z < LEV_Z1))
        return -1;
    return (int)fj * LEV_N + (int)fi;
}

__global__ __launch_bounds__(REG_BLOCK) void k_bev_level(TileGridDev tg, const float4* __restrict__ spts,
                                                         const int64_t* __restrict__ soff,
                                                         const int32_t* __restrict__ scnt,
                                                         const int32_t* __restrict__ idx, double* gl, double* gw) {
    __shared__ unsigned long long zs[LEV_CELLS], zm[LEV_CELLS];
    __shared__ double dv[2][LEV_CELLS];
    __shared__ double wsum[REG_BLOCK / WAVE][LEV_SUMS];
    __shared__ double s_g[12];
    __shared__ int s_rng[4], s_n, s_keep;
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    if (tid < 12) s_g[tid] = gl[16 * (int64_t)b + tid];
    zs[tid] = LEV_EMPTY;
    zm[tid] = LEV_EMPTY;
    if (tid == 0) {
        s_n = 0;
        s_keep = 0;
    }
    __syncthreads();
    const double r00 = s_g[0], r01 = s_g[1], r02 = s_g[2], t0 = s_g[3];
    const double r10 = s_g[4], r11 = s_g[5], r12 = s_g[6], t1 = s_g[7];
    const double r20 = s_g[8], r21 = s_g[9], r22 = s_g[10], t2 = s_g[11];
    if (tid == 0) {
        /* tiles under the window's box [-16, 16]^2 x [-8, 8] in the map frame */
        double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
        for (int k = 0; k < 8; k++) {
            const double qx = (k & 1) ? LEV_R : -LEV_R;
            const double qy = (k & 2) ? LEV_R : -LEV_R;
            const double qz = (k & 4) ? LEV_Z1 : LEV_Z0;
            const double px = r00 * qx + r01 * qy + r02 * qz + t0;
            const double py = r10 * qx + r11 * qy + r12 * qz + t1;
            xmin = fmin(xmin, px);
            xmax = fmax(xmax, px);
            ymin = fmin(ymin, py);
            ymax = fmax(ymax, py);
        }
        double fx0 = floor((xmin - 0.5 - tg.x0) / MAP_TILE), fx1 = floor((xmax + 0.5 - tg.x0) / MAP_TILE);
        double fy0 = floor((ymin - 0.5 - tg.y0) / MAP_TILE), fy1 = floor((ymax + 0.5 - tg.y0) / MAP_TILE);
        int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
        if (fx1 >= 0.0 && fx0 < (double)tg.ntx && fy1 >= 0.0 && fy0 < (double)tg.nty) {
            tx0 = (int)fmax(fx0, 0.0);
            tx1 = (int)fmin(fx1, (double)(tg.ntx - 1));
            ty0 = (int)fmax(fy0, 0.0);
            ty1 = (int)fmin(fy1, (double)(tg.nty - 1));
        }
        s_rng[0] = tx0;
        s_rng[1] = tx1;
        s_rng[2] = ty0;
        s_rng[3] = ty1;
    }
    {
        /* LEV_UNROLL points in flight per thread */
        const int s = idx[b];
        const float4* sp = spts + soff[s];
        const int ns = scnt[s];
        for (int i0 = tid; i0 < ns; i0 += LEV_UNROLL * REG_BLOCK) {
            float4 p[LEV_UNROLL];
#pragma unroll
            for (int u = 0; u < LEV_UNROLL; u++)
                if (i0 + u * REG_BLOCK < ns) p[u] = sp[i0 + u * REG_BLOCK];
#pragma unroll
            for (int u = 0; u < LEV_UNROLL; u++) {
                if (i0 + u * REG_BLOCK >= ns) break;
                const int c = lev_cell((double)p[u].x, (double)p[u].y, (double)p[u].z);
                if (c < 0) continue;
                const unsigned long long key = lev_key((double)p[u].z);
                if (key < zs[c]) atomicMin(&zs[c], key);
            }
        }
    }
    __syncthreads();
    const int tx0 = s_rng[0], tx1 = s_rng[1], ty0 = s_rng[2], ty1 = s_rng[3];
    for (int ty = ty0; ty <= ty1 && tx0 <= tx1; ty++) {
        const uint32_t pb = tg.start[(int64_t)ty * tg.ntx + tx0];
        const uint32_t pe = tg.start[(int64_t)ty * tg.ntx + tx1 + 1];
        for (uint32_t i0 = pb + tid; i0 < pe; i0 += LEV_UNROLL * REG_BLOCK) {
            float4 p[LEV_UNROLL];
#pragma unroll
            for (int u = 0; u < LEV_UNROLL; u++)
                if (i0 + u * REG_BLOCK < pe) p[u] = tg.pts[i0 + u * REG_BLOCK];
#pragma unroll
            for (int u = 0; u < LEV_UNROLL; u++) {
                if (i0 + u * REG_BLOCK >= pe) break;
                const double dx = (double)p[u].x - t0, dy = (double)p[u].y - t1, dz = (double)p[u].z - t2;
                const double qx = r00 * dx + r10 * dy + r20 * dz;
                const double qy = r01 * dx + r11 * dy + r21 * dz;
                const double qz = r02 * dx + r12 * dy + r22 * dz;
                const int c = lev_cell(qx, qy, qz);
                if (c < 0) continue;
                const unsigned long long key = lev_key(qz);
                if (key < zm[c]) atomicMin(&zm[c], key);
            }
        }
    }
    __syncthreads();
    const bool both = zs[tid] != LEV_EMPTY && zm[tid] != LEV_EMPTY;
    const double d = both ? lev_value(zm[tid]) - lev_value(zs[tid]) : INFINITY;
    if (both) atomicAdd(&s_n, 1);
    __syncthreads();
    const int n = s_n;
    if (n < LEV_MIN_CELLS) return; /* uniform over the workgroup */
    /* ascending bitonic sort of the 1024 differences (+inf where a cell lacks one),
     * one value per thread: partners within a wave by lane shuffles, the rest
     * through two LDS buffers in turn; a thread keeps its own value on a tie */
    double v = d;
    int buf = 0;
    for (int k = 2; k <= LEV_CELLS; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            double o;
            if (j >= WAVE) {
                dv[buf][tid] = v;
                __syncthreads();
                o = dv[buf][tid ^ j];
                buf ^= 1;
            } else {
                o = __shfl_xor(v, j);
            }
            const bool take_min = ((tid & j) == 0) == ((tid & k) == 0);
            v = take_min ? ((o < v) ? o : v) : ((v < o) ? o : v);
        }
    }
    dv[buf][tid] = v;
    __syncthreads();
    const double med = dv[buf][(n - 1) / 2];
    const bool keep = both && fabs(d - med) < LEV_TRIM;
    if (keep) atomicAdd(&s_keep, 1);
    const double x = (double)(tid % LEV_N) - 15.5, y = (double)(tid / LEV_N) - 15.5;
    const double one = keep ? 1.0 : 0.0, dd = keep ? d : 0.0;
    double t[LEV_SUMS] = {one, one * x, one * y, one * x * x, one * x * y, one * y * y, dd, dd * x, dd * y};
#pragma unroll
    for (int q = 0; q < LEV_SUMS; q++) {
        for (int off = WAVE / 2; off > 0; off >>= 1) t[q] = t[q] + __shfl_xor(t[q], off);
        if (lane == 0) wsum[wave][q] = t[q];
    }
    __syncthreads();
    if (tid != 0 || s_keep < LEV_MIN_CELLS) return;
    double S[LEV_SUMS];
    for (int q = 0; q < LEV_SUMS; q++) {
        S[q] = wsum[0][q];
        for (int w = 1; w < REG_BLOCK / WAVE; w++) S[q] = S[q] + wsum[w][q];
    }
    const double sn = S[0], sx = S[1], sy = S[2], sxx = S[3], sxy = S[4], syy = S[5], sd = S[6], sxd = S[7],
                 syd = S[8];
    const double m0 = syy * sn - sy * sy;
    const double m1 = sxy * sn - sy * sx;
    const double m2 = sxy * sy - syy * sx;
    const double det = sxx * m0 - sxy * m1 + sx * m2;
    if (!(det > 0.0)) return;
    const double a = (sxd * m0 - sxy * (syd * sn - sy * sd) + sx * (syd * sy - syy * sd)) / det;
    const double bb = (sxx * (syd * sn - sy * sd) - sxd * m1 + sx * (sxy * sd - syd * sx)) / det;
    const double pitch = a, roll = -bb;
    if (!(fabs(pitch) <= LEV_MAX_ANGLE && fabs(roll) <= LEV_MAX_ANGLE)) return;
    double ct, st, cf, sf;
    reg_small_cos_sin(pitch, ct, st);
    reg_small_cos_sin(roll, cf, sf);
    /* M = Ry(-pitch) Rx(-roll) */
    const double M[3][3] = {{ct, st * sf, -(st * cf)}, {0.0, cf, sf}, {st, -(ct * sf), ct * cf}};
    double* outs[2] = {gl + 16 * (int64_t)b, gw + 16 * (int64_t)b};
    for (int q = 0; q < 2; q++) {
        double* g = outs[q];
        for (int r = 0; r < 3; r++) {
            const double a0 = g[4 * r], a1 = g[4 * r + 1], a2 = g[4 * r + 2];
            for (int k = 0; k < 3; k++) g[4 * r + k] = a0 * M[0][k] + a1 * M[1][k] + a2 * M[2][k];
        }
    }
}

int launch_bev_register(TileGridDev tg, const uint16_t* sbev, const float4* pts, const int64_t* offsets,
                        const int32_t* counts, const int32_t* idx, const double* gl, const double* gw, int32_t B,
                        const RegisterDev& rp, double* gl_out, double* gw_out, int32_t* score, hipStream_t stream) {
    if (B <= 0) return 0;
    if (rp.radius < 1 || rp.radius > REG_MAX_RADIUS || rp.yaws < 1 || rp.yaws > REG_MAX_YAWS) return (int)hipErrorInvalidValue;
#define REG_LAUNCH(r)                                                                                              \
    case r:                                                                                                        \
        hipLaunchKernelGGL(k_bev_register<r>, dim3(B), dim3(REG_BLOCK), 0, stream, tg, sbev, pts, offsets, counts,    \
                           idx, gl, gw, rp, gl_out, gw_out, score);                                                \
        break;
    switch (rp.radius) {
        REG_LAUNCH(1) REG_LAUNCH(2) REG_LAUNCH(3) REG_LAUNCH(4) REG_LAUNCH(5) REG_LAUNCH(6) REG_LAUNCH(7) REG_LAUNCH(8)
    }
#undef REG_LAUNCH
    if (rp.level)
        hipLaunchKernelGGL(k_bev_level, dim3(B), dim3(REG_BLOCK), 0, stream, tg, pts, offsets, counts, idx, gl_out,
                           gw_out);
    return (int)hipGetLastError();
}

/* dst row b = src row idx[b] (ncols u16 values, ncols a multiple of 8) */
__global__ void k_gather_rows(const uint16_t* __restrict__ src, const int32_t* __restrict__ idx, int64_t ncols,
                              uint16_t* __restrict__ dst, int64_t ldd) {
    const int b = blockIdx.y;
    const int64_t r = idx[b];
    const uint4* s = (const uint4*)(src + r * ncols);
    uint4* d = (uint4*)(dst + (int64_t)b * ldd);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ncols / 8; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

int launch_gather_rows(const uint16_t* src, const int32_t* idx, int32_t B, int64_t ncols, uint16_t* dst, int64_t ldd,
                       hipStream_t stream) {
    if (B <= 0) return 0;
    hipLaunchKernelGGL(k_gather_rows, dim3(8, B), dim3(256), 0, stream, src, idx, ncols, dst, ldd);
    return (int)hipGetLastError();
}

/* dst[c][r] = src[r][c]; rows and cols multiples of 64 */
__global__ __launch_bounds__(256) void k_transpose_u16(const uint16_t* __restrict__ src, int64_t rows, int64_t cols,
                                                       uint16_t* __restrict__ dst) {
    __shared__ uint16_t t[64][66];
    const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int k = ty; k < 64; k += 4) t[k][tx] = src[(r0 + k) * cols + c0 + tx];
    __syncthreads();
    for (int k = ty; k < 64; k += 4) dst[(c0 + k) * rows + r0 + tx] = t[tx][k];
}

int launch_transpose_u16(const uint16_t* src, int64_t rows, int64_t cols, uint16_t* dst, hipStream_t stream) {
    if (rows % 64 || cols % 64) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_transpose_u16, dim3((unsigned)(cols / 64), (unsigned)(rows / 64)), dim3(256), 0, stream, src,
                       rows, cols, dst);
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------ GEMM
 * C_z[M x N] = A[M x K_z] * B[N x K_z]^T with A and B bf16, row-major in K (lda,
 * ldb in elements), float32 accumulation; split z covers K range
 * [z K/splits, (z+1) K/splits) and writes its own M x N float32 slice of C.
 *
 * 256 threads = 4 waves in a 2 x 2 arrangement, each wave owns a 64 x 64 block
 * of C as 4 x 4 tiles of v_mfma_f32_16x16x32_bf16.  For that instruction lane l
 * supplies row (l mod 16) of A and column (l mod 16) of B over the 8 K values
 * 8 (l / 16) .. 8 (l / 16) + 7 of the 32-deep step, and receives
 * D[4 (l / 16) + r][l mod 16], r = 0..3.  A and B take their K values from the
 * same positions, so the product is independent of how the hardware orders the
 * K values inside one lane's 8.  K steps of 64 are staged through two LDS
 * buffers: the next step's global loads are in flight while the current one is
 * multiplied.  Rows are padded by 8 elements (16 bytes). */
constexpr int GPAD = 8;

__global__ __launch_bounds__(256) void k_gemm_nt(const uint16_t* __restrict__ A, int64_t lda,
                                                 const uint16_t* __restrict__ B, int64_t ldb, float* __restrict__ C,
                                                 int64_t ldc, int64_t split_stride, int kps) {
    __shared__ __attribute__((aligned(16))) uint16_t As[2][GEMM_BM][GEMM_BK + GPAD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][GEMM_BN][GEMM_BK + GPAD];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int64_t m0 = (int64_t)blockIdx.y * GEMM_BM, n0 = (int64_t)blockIdx.x * GEMM_BN;
    const int64_t k0 = (int64_t)blockIdx.z * kps;
    const int nk = kps / GEMM_BK;
    const uint16_t* Ab = A + m0 * lda + k0;
    const uint16_t* Bb = B + n0 * ldb + k0;
    uint4 ra[4], rb[4];
    /* tile chunk c = tid + 256 q (q = 0..3): row c / 8, columns 8 (c mod 8) .. +7 */
    const int lr = tid >> 3, lc = (tid & 7) * 8;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        ra[q] = *(const uint4*)(Ab + (int64_t)(lr + 32 * q) * lda + lc);
        rb[q] = *(const uint4*)(Bb + (int64_t)(lr + 32 * q) * ldb + lc);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        *(uint4*)&As[0][lr + 32 * q][lc] = ra[q];
        *(uint4*)&Bs[0][lr + 32 * q][lc] = rb[q];
    }
    __syncthreads();
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    for (int kt = 0; kt < nk; kt++) {
        const int buf = kt & 1;
        /* next K step (the last iteration reloads its own step: in bounds, unused) */
        const int64_t ko = (int64_t)(kt + 1 < nk ? kt + 1 : kt) * GEMM_BK;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            ra[q] = *(const uint4*)(Ab + (int64_t)(lr + 32 * q) * lda + ko + lc);
            rb[q] = *(const uint4*)(Bb + (int64_t)(lr + 32 * q) * ldb + ko + lc);
        }
#pragma unroll
        for (int kk = 0; kk < GEMM_BK; kk += 32) {
            bf16x8 a[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = *(const bf16x8*)&As[buf][wm * 64 + i * 16 + fr][kk + fk];
#pragma unroll
            for (int j = 0; j < 4; j++) bv[j] = *(const bf16x8*)&Bs[buf][wn * 64 + j * 16 + fr][kk + fk];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bv[j], acc[i][j], 0, 0, 0);
        }
        /* the other buffer was last read in the previous iteration, before its barrier */
#pragma unroll
        for (int q = 0; q < 4; q++) {
            *(uint4*)&As[buf ^ 1][lr + 32 * q][lc] = ra[q];
            *(uint4*)&Bs[buf ^ 1][lr + 32 * q][lc] = rb[q];
        }
        __syncthreads();
    }
    float* Cz = C + (int64_t)blockIdx.z * split_stride;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = m0 + wm * 64 + i * 16 + (lane >> 4) * 4 + r;
                const int64_t col = n0 + wn * 64 + j * 16 + fr;
                Cz[row * ldc + col] = acc[i][j][r];
            }
}

int launch_gemm_nt(const uint16_t* A, int64_t lda, const uint16_t* B, int64_t ldb, float* C, int64_t ldc, int M, int N,
                   int K, int splits, hipStream_t stream) {
    if (M <= 0 || N <= 0 || K <= 0 || splits <= 0 || M % GEMM_BM || N % GEMM_BN || K % (GEMM_BK * splits) ||
        lda % 8 || ldb % 8 || lda < K || ldb < K || ldc < N)
        return (int)hipErrorInvalidValue;
    const int kps = K / splits;
    hipLaunchKernelGGL(k_gemm_nt, dim3(N / GEMM_BN, M / GEMM_BM, splits), dim3(256), 0, stream, A, lda, B, ldb, C, ldc,
                       (int64_t)M * ldc, kps);
    return (int)hipGetLastError();
}

/* out[m][n] = act(sum_z P_z[m][n] + bias[n]) in bf16, and optionally its transpose */
__global__ void k_bias_act(const float* __restrict__ P, int splits, int M, int N, const float* __restrict__ bias,
                           int relu, uint16_t* __restrict__ out, uint16_t* __restrict__ outT) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)M * N) return;
    const int m = (int)(i / N), n = (int)(i % N);
    float v = 0.f;
    for (int z = 0; z < splits; z++) v += P[(int64_t)z * M * N + i];
    v += bias[n];
    if (relu && !(v > 0.f)) v = 0.f;
    const uint16_t h = f2bf(v);
    out[i] = h;
    if (outT) outT[(int64_t)n * M + m] = h;
}

int launch_bias_act(const float* P, int splits, int M, int N, const float* bias, int relu, uint16_t* out,
                    uint16_t* outT, hipStream_t stream) {
    const int64_t n = (int64_t)M * N;
    hipLaunchKernelGGL(k_bias_act, dim3(blocks_for(n, 256)), dim3(256), 0, stream, P, splits, M, N, bias, relu, out,
                       outT);
    return (int)hipGetLastError();
}

/* Output layer, one wave per sample: Y = H2 W3^T + b3.  With a target, also the
 * weighted squared error of the sample and dL/dY for the batch mean
 * L = (1/B) sum_b sum_o w_o (Y_bo - T_bo)^2. */
__global__ __launch_bounds__(256) void k_head_fwd(const uint16_t* __restrict__ H2, const uint16_t* __restrict__ W3,
                                                  const float* __restrict__ b3, int B, int F2, float* __restrict__ Y,
                                                  const float* __restrict__ target, const float* __restrict__ wts,
                                                  float* __restrict__ dY, float* __restrict__ loss_b) {
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b >= B) return;
    const uint16_t* h = H2 + (int64_t)b * F2;
    float y[NOUT];
#pragma unroll
    for (int o = 0; o < NOUT; o++) {
        float s = 0.f;
        for (int k = lane; k < F2; k += 64) s += bf2f(h[k]) * bf2f(W3[(int64_t)o * F2 + k]);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        y[o] = s + b3[o];
    }
    if (lane == 0) {
        float l = 0.f;
        for (int o = 0; o < NOUT; o++) {
            Y[(int64_t)b * NOUT + o] = y[o];
            if (target) {
                const float d = y[o] - target[(int64_t)b * NOUT + o];
                l += wts[o] * d * d;
                dY[(int64_t)b * NOUT + o] = 2.f * wts[o] * d / (float)B;
            }
        }
        if (target) loss_b[b] = l;
    }
}

int launch_head_fwd(const uint16_t* H2, const uint16_t* W3, const float* b3, int B, int F2, float* Y,
                    const float* target, const float* weights, float* dY, float* loss_b, hipStream_t stream) {
    hipLaunchKernelGGL(k_head_fwd, dim3(blocks_for(B, 4)), dim3(256), 0, stream, H2, W3, b3, B, F2, Y, target, weights,
                       dY, loss_b);
    return (int)hipGetLastError();
}

/* The two backward kernels below cut the (sample, unit) plane into tiles of
 * BWD_ROWS samples x 64 units, one tile per 256-thread workgroup: each of the four
 * waves takes BWD_ROWS / 4 consecutive samples of the tile, in order.  Sums over
 * samples are formed per wave, added over the waves in wave order and left per
 * tile in `part`; a second kernel adds the tiles in tile order.  Every sum
 * therefore has one fixed order.  The transposed bf16 outputs go through an LDS
 * tile so that both layouts are written in contiguous runs. */
constexpr int BWD_ROWS = 32;
constexpr int BWD_WAVES = 4;
constexpr int BWD_PER_WAVE = BWD_ROWS / BWD_WAVES;

int bwd_tiles(int B) { return (B + BWD_ROWS - 1) / BWD_ROWS; }

/* tile [64 units][BWD_ROWS samples] -> dst[unit][sample], row length B */
__device__ inline void store_tile_t(const uint16_t (*tile)[BWD_ROWS + 2], int n0, int b0, int F, int B,
                                    uint16_t* __restrict__ dst) {
    for (int e = threadIdx.x; e < 64 * BWD_ROWS; e += 256) {
        const int u = e / BWD_ROWS, r = e % BWD_ROWS;
        const int n = n0 + u, b = b0 + r;
        if (n < F && b < B) dst[(int64_t)n * B + b] = tile[u][r];
    }
}

/* Backward of the output layer and the second ReLU over one tile:
 * dZ2[b][n] = [H2[b][n] > 0] sum_o dY[b][o] W3[o][n] (bf16, also transposed into
 * dZ2T), and per tile the partial sums over its samples of dY[b][o] H2[b][n]
 * (-> dW3[o][n]), of dZ2[b][n] (-> db2[n]) and, in the first column of tiles,
 * of dY[b][o] (-> db3[o]).  part holds tiles x (NOUT + 1) x F2 values, then
 * tiles x NOUT. */
__global__ __launch_bounds__(256) void k_head_bwd(const float* __restrict__ dY, const uint16_t* __restrict__ H2,
                                                  const uint16_t* __restrict__ W3, int B, int F2,
                                                  uint16_t* __restrict__ dZ2, uint16_t* __restrict__ dZ2T,
                                                  float* __restrict__ part) {
    __shared__ uint16_t tile[64][BWD_ROWS + 2];
    __shared__ float sy[BWD_ROWS][NOUT];
    __shared__ float psum[BWD_WAVES][NOUT + 1][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n0 = blockIdx.x * 64, n = n0 + lane;
    const int b0 = blockIdx.y * BWD_ROWS;
    for (int e = threadIdx.x; e < BWD_ROWS * NOUT; e += 256) {
        const int b = b0 + e / NOUT;
        sy[e / NOUT][e % NOUT] = b < B ? dY[(int64_t)b * NOUT + e % NOUT] : 0.f;
    }
    float w3[NOUT], gw[NOUT];
#pragma unroll
    for (int o = 0; o < NOUT; o++) {
        w3[o] = n < F2 ? bf2f(W3[(int64_t)o * F2 + n]) : 0.f;
        gw[o] = 0.f;
    }
    __syncthreads();
    float gb = 0.f;
#pragma unroll
    for (int i = 0; i < BWD_PER_WAVE; i++) {
        const int r = wv * BWD_PER_WAVE + i, b = b0 + r;
        const bool in = n < F2 && b < B;
        const float h = in ? bf2f(H2[(int64_t)b * F2 + n]) : 0.f;
        float dz = 0.f;
#pragma unroll
        for (int o = 0; o < NOUT; o++) {
            const float d = sy[r][o];
            dz += d * w3[o];
            gw[o] += d * h;
        }
        if (!(h > 0.f)) dz = 0.f;
        const uint16_t zb = f2bf(dz);
        if (in) dZ2[(int64_t)b * F2 + n] = zb;
        tile[lane][r] = zb;
        gb += dz;
    }
#pragma unroll
    for (int o = 0; o < NOUT; o++) psum[wv][o][lane] = gw[o];
    psum[wv][NOUT][lane] = gb;
    __syncthreads();
    if (wv == 0 && n < F2) {
#pragma unroll
        for (int q = 0; q <= NOUT; q++)
            part[((int64_t)blockIdx.y * (NOUT + 1) + q) * F2 + n] =
                ((psum[0][q][lane] + psum[1][q][lane]) + psum[2][q][lane]) + psum[3][q][lane];
    }
    if (blockIdx.x == 0 && threadIdx.x < NOUT) {
        float s = 0.f;
        for (int r = 0; r < BWD_ROWS; r++) s += sy[r][threadIdx.x];
        part[(int64_t)gridDim.y * (NOUT + 1) * F2 + blockIdx.y * NOUT + threadIdx.x] = s;
    }
    store_tile_t(tile, n0, b0, F2, B, dZ2T);
}

/* dW3, db2 and db3 from the per-tile partials of k_head_bwd, tiles in order;
 * blockIdx.y picks the output (dW3 row q < NOUT, or db2) */
__global__ void k_head_bwd_sum(const float* __restrict__ part, int tiles, int F2, float* __restrict__ dW3,
                               float* __restrict__ db2, float* __restrict__ db3) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (n < F2) {
        float s = 0.f;
        for (int t = 0; t < tiles; t++) s += part[((int64_t)t * (NOUT + 1) + q) * F2 + n];
        if (q < NOUT) dW3[(int64_t)q * F2 + n] = s;
        else db2[n] = s;
    }
    if (blockIdx.x == 0 && q == 0 && threadIdx.x < NOUT) {
        const float* p3 = part + (int64_t)tiles * (NOUT + 1) * F2;
        float s = 0.f;
        for (int t = 0; t < tiles; t++) s += p3[t * NOUT + threadIdx.x];
        db3[threadIdx.x] = s;
    }
}

int launch_head_bwd(const float* dY, const uint16_t* H2, const uint16_t* W3, int B, int F2, float* dW3, float* db3,
                    uint16_t* dZ2, uint16_t* dZ2T, float* db2, float* part, hipStream_t stream) {
    const int tiles = bwd_tiles(B);
    hipLaunchKernelGGL(k_head_bwd, dim3(blocks_for(F2, 64), tiles), dim3(256), 0, stream, dY, H2, W3, B, F2, dZ2, dZ2T,
                       part);
    hipLaunchKernelGGL(k_head_bwd_sum, dim3(blocks_for(F2, 64), NOUT + 1), dim3(64), 0, stream, part, tiles, F2, dW3,
                       db2, db3);
    return (int)hipGetLastError();
}

/* Backward of a ReLU layer over one tile: dZ[b][n] = [H[b][n] > 0] dH[b][n],
 * written transposed in bf16, and per tile the partial sums over its samples of
 * dZ (-> db[n]); part holds tiles x F values. */
__global__ __launch_bounds__(256) void k_relu_bwd(const float* __restrict__ dH, const uint16_t* __restrict__ H, int B,
                                                  int F, uint16_t* __restrict__ dZT, float* __restrict__ part) {
    __shared__ uint16_t tile[64][BWD_ROWS + 2];
    __shared__ float psum[BWD_WAVES][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n0 = blockIdx.x * 64, n = n0 + lane;
    const int b0 = blockIdx.y * BWD_ROWS;
    float gb = 0.f;
#pragma unroll
    for (int i = 0; i < BWD_PER_WAVE; i++) {
        const int r = wv * BWD_PER_WAVE + i, b = b0 + r;
        float dz = 0.f;
        if (n < F && b < B) {
            const float h = bf2f(H[(int64_t)b * F + n]);
            dz = (h > 0.f) ? dH[(int64_t)b * F + n] : 0.f;
        }
        tile[lane][r] = f2bf(dz);
        gb += dz;
    }
    psum[wv][lane] = gb;
    __syncthreads();
    if (wv == 0 && n < F) part[(int64_t)blockIdx.y * F + n] = ((psum[0][lane] + psum[1][lane]) + psum[2][lane]) + psum[3][lane];
    store_tile_t(tile, n0, b0, F, B, dZT);
}

/* out[n] = sum over tiles t, in order, of part[t][n] */
__global__ void k_sum_tiles(const float* __restrict__ part, int tiles, int F, float* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= F) return;
    float s = 0.f;
    for (int t = 0; t < tiles; t++) s += part[(int64_t)t * F + n];
    out[n] = s;
}

int launch_relu_bwd(const float* dH, const uint16_t* H, int B, int F, uint16_t* dZT, float* db, float* part,
                    hipStream_t stream) {
    const int tiles = bwd_tiles(B);
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(F, 64), tiles), dim3(256), 0, stream, dH, H, B, F, dZT, part);
    hipLaunchKernelGGL(k_sum_tiles, dim3(blocks_for(F, 64)), dim3(64), 0, stream, part, tiles, F, db);
    return (int)hipGetLastError();
}

/* Adam (Kingma & Ba 2015) with decoupled weight decay wd (Loshchilov & Hutter
 * 2019) on float32 master weights; bc1 = 1 - beta1^t and bc2 = 1 - beta2^t; wb
 * (optional) receives the bf16 copy used by the GEMMs. */
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g,
                       int64_t n, float lr, float beta1, float beta2, float eps, float bc1, float bc2, float wd,
                       uint16_t* __restrict__ wb) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * gi;
    const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * ((mi / bc1) / (sqrtf(vi / bc2) + eps) + wd * w[i]);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

int launch_adam(float* w, float* m, float* v, const float* g, int64_t n, float lr, float beta1, float beta2, float eps,
                float bc1, float bc2, float wd, uint16_t* wb, hipStream_t stream) {
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n, 256)), dim3(256), 0, stream, w, m, v, g, n, lr, beta1, beta2, eps, bc1,
                       bc2, wd, wb);
    return (int)hipGetLastError();
}

__global__ void k_to_bf16(const float* __restrict__ x, int64_t n, uint16_t* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(x[i]);
}

int launch_to_bf16(const float* x, int64_t n, uint16_t* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n, 256)), dim3(256), 0, stream, x, n, out);
    return (int)hipGetLastError();
}

/* out[0] += x[0] + ... + x[n-1] in one fixed order: lane l of a 64-thread
 * workgroup sums x[l], x[l + 64], ... in turn, then the 64 lane sums are added
 * by a pairwise tree in LDS. */
__global__ __launch_bounds__(64) void k_sum_ordered(const float* __restrict__ x, int n, float* __restrict__ out) {
    __shared__ float t[64];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 64) s += x[i];
    t[threadIdx.x] = s;
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) t[threadIdx.x] = t[threadIdx.x] + t[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] += t[0];
}

int launch_sum_ordered(const float* x, int n, float* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_sum_ordered, dim3(1), dim3(64), 0, stream, x, n, out);
    return (int)hipGetLastError();
}


/* ------------------------------------------------------------------ guesses */

__global__ void k_guesses_local(const double* __restrict__ world, int32_t S, int32_t B, double ox, double oy,
                                double oz, double* __restrict__ local) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const double* G = world + 16 * (int64_t)(b < S ? b : 0);
    double* L = local + 16 * (int64_t)b;
    for (int k = 0; k < 16; k++) L[k] = G[k];
    L[3] = G[3] - ox;
    L[7] = G[7] - oy;
    L[11] = G[11] - oz;
}

int launch_guesses_local(const double* world, int32_t S, int32_t B, double ox, double oy, double oz, double* local,
                         hipStream_t stream) {
    hipLaunchKernelGGL(k_guesses_local, dim3(blocks_for(B, 64)), dim3(64), 0, stream, world, S, B, ox, oy, oz, local);
    return (int)hipGetLastError();
}

__global__ void k_refine(const float* __restrict__ Y, const double* __restrict__ guesses, int32_t S, CorrScale scale,
                         int32_t keep_rotation, double* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    double p6[NOUT], C[16], G[16], R[16];
    for (int k = 0; k < NOUT; k++) {
        double v = (double)Y[NOUT * (int64_t)s + k];
        if (!isfinite(v)) v = 0.0;
        p6[k] = v * scale.v[k];
    }
    pose_matrix(p6, C);
    for (int k = 0; k < 16; k++) G[k] = guesses[16 * (int64_t)s + k];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            double v = 0.0;
            for (int k = 0; k < 4; k++) v += G[r * 4 + k] * C[k * 4 + c];
            R[r * 4 + c] = v;
        }
    if (keep_rotation)
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) R[r * 4 + c] = G[r * 4 + c];
    for (int k = 0; k < 16; k++) out[16 * (int64_t)s + k] = R[k];
}

int launch_refine(const float* Y, const double* guesses, int32_t S, CorrScale scale, int32_t keep_rotation,
                  double* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_refine, dim3(blocks_for(S, 64)), dim3(64), 0, stream, Y, guesses, S, scale, keep_rotation,
                       out);
    return (int)hipGetLastError();
}

} // namespace ndt
