/*
 * ndt_learn_internal.h - device structures and kernel launchers of the learned
 * warm start (ndt_learn.hip, hipcc) shared with its host side (ndt_learn.cpp, g++).
 *
 * Features: two bird's-eye-view (BEV) occupancy grids of 64 x 64 x 8 cells of
 * 0.5 m x 0.5 m x 1.5 m, x and y in [-16, 16) m, z in [-2.5, 9.5) m, one of the
 * scan in its sensor frame and one of the map points around the guess, in the
 * guess frame.  A cell holds log1p(point count) in bfloat16.  Feature index =
 * grid * 32768 + (iz * 64 + iy) * 64 + ix (grid 0 = scan, 1 = map).
 *
 * Model: bf16 multilayer perceptron F0 -> F1 -> F2 -> 6 (default 65536 -> 1024 ->
 * 256 -> 6), ReLU after the two hidden layers, float32 accumulation, predicting
 * the pose correction of the guess (scaled per component).
 */
#ifndef NDT_LEARN_INTERNAL_H
#define NDT_LEARN_INTERNAL_H

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#else
#include <hip/hip_runtime_api.h>
#endif

namespace ndt {

constexpr int BEV_NX = 64, BEV_NY = 64, BEV_NZ = 8;
constexpr int BEV_CELLS = BEV_NX * BEV_NY * BEV_NZ; /* 32768 per grid */
constexpr double BEV_CELL = 0.5;
constexpr double BEV_X0 = -16.0, BEV_Y0 = -16.0, BEV_Z0 = -2.5, BEV_DZ = 1.5;
constexpr double MAP_TILE = 4.0; /* map points are bucketed in 4 m x 4 m tiles */
constexpr int NOUT = 6;

/* GEMM tile: C tiles of 128 x 128, K steps of 64 */
constexpr int GEMM_BM = 128, GEMM_BN = 128, GEMM_BK = 64;

/* map points bucketed by 2-D tile, row-major over (ty, tx) */
struct TileGridDev {
    const float4* pts;         /* map points sorted by tile */
    const uint32_t* start;     /* ntx * nty + 1 offsets */
    int ntx, nty;
    double x0, y0;             /* lower corner of tile (0, 0), map-local */
};

int launch_bev_scan(const float4* pts, const int64_t* offsets, const int32_t* counts, int32_t S,
                    const uint16_t* lut, uint16_t* out, hipStream_t stream);
int launch_bev_crop(TileGridDev tg, const double* guesses, int32_t B, const uint16_t* lut, uint16_t* X, int64_t ldx,
                    int64_t col_off, hipStream_t stream);
int launch_gather_rows(const uint16_t* src, const int32_t* idx, int32_t B, int64_t ncols, uint16_t* dst, int64_t ldd,
                       hipStream_t stream);
int launch_transpose_u16(const uint16_t* src, int64_t rows, int64_t cols, uint16_t* dst, hipStream_t stream);
int launch_gemm_nt(const uint16_t* A, int64_t lda, const uint16_t* B, int64_t ldb, float* C, int64_t ldc, int M, int N,
                   int K, int splits, hipStream_t stream);
int launch_bias_act(const float* P, int splits, int M, int N, const float* bias, int relu, uint16_t* out,
                    uint16_t* outT, hipStream_t stream);
int launch_head_fwd(const uint16_t* H2, const uint16_t* W3, const float* b3, int B, int F2, float* Y,
                    const float* target, const float* weights, float* dY, float* loss_b, hipStream_t stream);
/* sample tiles of the backward kernels; their partial-sum buffer `part` holds
 * bwd_tiles(B) * ((NOUT + 1) * F2 + NOUT) floats (head) or bwd_tiles(B) * F (ReLU) */
int bwd_tiles(int B);
int launch_head_bwd(const float* dY, const uint16_t* H2, const uint16_t* W3, int B, int F2, float* dW3, float* db3,
                    uint16_t* dZ2, uint16_t* dZ2T, float* db2, float* part, hipStream_t stream);
int launch_relu_bwd(const float* dH, const uint16_t* H, int B, int F, uint16_t* dZT, float* db, float* part,
                    hipStream_t stream);
int launch_adam(float* w, float* m, float* v, const float* g, int64_t n, float lr, float beta1, float beta2, float eps,
                float bc1, float bc2, float wd, uint16_t* wb, hipStream_t stream);
int launch_sum_ordered(const float* x, int n, float* out, hipStream_t stream);
int launch_to_bf16(const float* x, int64_t n, uint16_t* out, hipStream_t stream);

/* BEV registration of a guess (ndt_register_params): yaw k of the search is
 * (k - (yaws - 1) / 2) * yaw_step, its cosine, sine and tangent computed on the
 * host; `level` adds the leveling kernel after it */
constexpr int REG_MAX_RADIUS = 8, REG_MAX_YAWS = 11;
constexpr int REG_BLOCK = 1024;
struct RegisterDev {
    int32_t radius, yaws, refine_yaw, z_search, level, pad;
    double yaw_step;
    double c[REG_MAX_YAWS], s[REG_MAX_YAWS], t[REG_MAX_YAWS];
};
/* Registers row b (B rows): scan grid sbev[idx[b]] (bf16 bits, 0 = no point),
 * and with z_search or level the scan's points (pts from offsets[idx[b]],
 * counts[idx[b]] of them), against the map around the map-local guess gl[b];
 * writes the registered
 * map-local guess to gl_out[b], the registered world guess (from gw[b]) to
 * gw_out[b] and, when `score` is not null, the best score; with level, a second
 * kernel then levels both against the map around gl_out[b].  gl_out may be gl
 * and gw_out may be gw (each row is read before it is written). */
int launch_bev_register(TileGridDev tg, const uint16_t* sbev, const float4* pts, const int64_t* offsets,
                        const int32_t* counts, const int32_t* idx, const double* gl, const double* gw, int32_t B,
                        const RegisterDev& rp, double* gl_out, double* gw_out, int32_t* score, hipStream_t stream);

/* scale of each output of the model: the prediction of component k is the pose
 * correction's component k divided by scale[k] (x, y, z in metres; roll, pitch,
 * yaw in radians) */
struct CorrScale {
    double v[NOUT];
};
/* local[b] = world pose of guess min(b, S - 1 if b >= S: 0) minus the origin, for
 * b < B (rows past S repeat guess 0) */
int launch_guesses_local(const double* world, int32_t S, int32_t B, double ox, double oy, double oz, double* local,
                         hipStream_t stream);
/* refined guesses out[s] = guesses[s] * pose(Y[s] * scale) (R = Rx Ry Rz), a
 * non-finite prediction counting as 0; keep_rotation keeps the guess's rotation */
int launch_refine(const float* Y, const double* guesses, int32_t S, CorrScale scale, int32_t keep_rotation,
                  double* out, hipStream_t stream);

} // namespace ndt

#endif
