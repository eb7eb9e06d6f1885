/*
 * ndt_learn.h - C ABI of the learned warm start of the NDT engine (libndthip.so).
 *
 * A small learned pose proposer refines the initial guess before NDT runs.  The
 * reference has no counterpart: it aligns from the guess it is given
 * (src/CPU/ndt_mapping/kernel.cpp:1187-1206, where init_guess becomes the start
 * pose of computeTransformation).  NDT stays the judge: the final NDT score decides
 * between the learned start and plain NDT from the given guess (see
 * NDT_LEARNED_RACE and NDT_LEARNED_FALLBACK).
 *
 * Features: two bird's-eye-view occupancy grids of 64 x 64 x 8 cells
 * (0.5 m x 0.5 m x 1.5 m; x, y in [-16, 16) m, z in [-2.5, 9.5) m): the scan in its
 * sensor frame and the map points around the guess, expressed in the guess frame.
 * Cells hold log1p(count) in bfloat16 (counts saturate at 65535); feature index =
 * grid * 32768 + (iz * 64 + iy) * 64 + ix, grid 0 = scan, grid 1 = map.
 *
 * Model: bf16 MLP f0 -> f1 -> f2 -> 6 with ReLU after both hidden layers and
 * float32 accumulation (default 65536 -> 1024 -> 256 -> 6), trained with Adam on
 * the GPU.  It predicts the pose correction C = guess^-1 * target as
 * (x, y, z, roll, pitch, yaw) (R = Rx Ry Rz), each divided by a fixed scale
 * (3 m, 3 m, 1 m, 0.05 rad, 0.05 rad, 0.2 rad); the refined guess is guess * C.
 *
 * Poses are row-major 4x4 float64 in world coordinates (map-local + the origin
 * given to ndt_bevmap_build), as for ndt_align_scans.
 */
#ifndef NDT_LEARN_H
#define NDT_LEARN_H

#include <stdint.h>

#include "ndt_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ndt_bevmap ndt_bevmap;   /* map points bucketed for BEV crops */
typedef struct ndt_model ndt_model;     /* MLP weights, bf16 copies, Adam state */
typedef struct ndt_dataset ndt_dataset; /* training triples resident on the device */

enum { NDT_BEV_CELLS = 32768, NDT_FEATURES = 65536, NDT_MODEL_OUT = 6 };

/* Parameter tensors of a model, float32, row-major:
 * W1 [f1 x f0], b1 [f1], W2 [f2 x f1], b2 [f2], W3 [6 x f2], b3 [6]. */
enum { NDT_W1 = 0, NDT_B1 = 1, NDT_W2 = 2, NDT_B2 = 3, NDT_W3 = 4, NDT_B3 = 5 };

typedef struct ndt_train_params {
    int32_t epochs;        /* passes over the dataset */
    int32_t batch;         /* samples per step, a multiple of 128 */
    float lr;              /* Adam step size */
    float beta1, beta2, eps;
    uint64_t seed;         /* sample order */
    float weights[6];      /* loss weight per output component */
    float weight_decay;    /* decoupled weight decay of the weight matrices (AdamW) */
    int32_t cosine;        /* 1: step size follows a cosine from lr down to lr / 20 */
} ndt_train_params;

/* How ndt_align_scans_learned uses the refined guess:
 *   NDT_LEARNED_RACE     NDT runs from the given and from the refined guess in one
 *                        batch; the run with the better final score is kept (a
 *                        converged run beats one that did not converge), so the
 *                        result is never worse, by the NDT score, than plain NDT;
 *   NDT_LEARNED_FALLBACK NDT judges the refined guess: it runs from the refined guess
 *                        when the NDT score there is at least the given guess's, and
 *                        plain NDT from the given guess replaces that run where it
 *                        ended with a lower score than the given guess has;
 *                        otherwise plain NDT runs from the given guess.
 *                        All of it runs in one launch: each scan opens with one pass
 *                        at its given guess (which is also the first pass of a plain
 *                        run from there), one at the refined guess, and continues
 *                        with the chosen run while the other scans go on; each run
 *                        splits and sums its passes as a launch of its own would, so
 *                        every result equals, bit for bit, ndt_align_scans from one
 *                        of the two guesses (passes count every pass spent).
 * Either may be combined (bitwise or) with
 *   NDT_LEARNED_KEEP_ROTATION  the refined guess takes only the model's translation
 *                        and keeps the rotation of the guess the model corrects
 *                        (the given guess, or the registered one when the model
 *                        has a registration set). */
enum { NDT_LEARNED_RACE = 0, NDT_LEARNED_FALLBACK = 1, NDT_LEARNED_KEEP_ROTATION = 0x10 };

/* BEV registration, the stage before the model when the model has it set
 * (ndt_model_set_register): the occupancy of the scan grid (cells with a count)
 * is correlated with the occupancy of the map around the guess, for `yaws` yaw
 * angles (yaw_step apart, centred on the guess's yaw) and for every shift of
 * -radius .. radius cells on x and y.  The best (score = occupied cells in common)
 * gives the registered guess guess * Rz(yaw) * T(dx, dy, 0), the shifts refined by
 * a parabola through the neighbouring scores (and the yaw too with refine_yaw);
 * ties go to the candidate nearest the guess, and a guess with no cell in common
 * at any candidate stays as it is.  With z_search the height follows: histograms
 * of point heights in 5 cm bins over [-4, 12) m (bin floor((z + 4) * 20)), the
 * scan's in its sensor frame
 * and the map's under the grid's footprint in the guess frame, correlated over
 * shifts of -1.5 .. 1.5 m; the best shift, refined by a parabola, moves the
 * registered guess along its z axis.  With level the roll and pitch follow: per
 * 1 m cell of [-16, 16)^2 m the lowest scan point and the lowest map point in the
 * registered guess's frame (heights in [-8, 8) m), their differences within
 * 0.25 m of the median fitted by a plane, whose slopes give the guess's tilt
 * against the scan (at most 0.035 rad each way; at least 32 cells; otherwise the
 * guess keeps its rotation).  The model then sees the map around the registered
 * guess and corrects it. */
typedef struct ndt_register_params {
    int32_t radius;     /* cells (0.5 m) searched each way on x and y, 0..8; 0 = off */
    int32_t yaws;       /* yaw angles searched: odd, 1..11 */
    double yaw_step;    /* radians between them, 0 < yaw_step <= 0.05 */
    int32_t refine_yaw; /* 1: the yaw refined by a parabola through its neighbours */
    int32_t z_search;   /* 1: the height registered too */
    int32_t level;      /* 1: roll and pitch leveled against the map after the search */
    int32_t pad;        /* 0 */
} ndt_register_params;

enum { NDT_REGISTER_MAX_RADIUS = 8, NDT_REGISTER_MAX_YAWS = 11 };

/* radius 6 (+-3 m), 7 yaws 1 degree apart (+-3 degrees), refine_yaw 1, z_search 1,
 * level 1 */
void ndt_register_params_default(ndt_register_params* rp);

/* Buckets the map points (map-local coordinates, m x 4 floats) into 4 m x 4 m
 * tiles for the map grids; origin = world position of the map-local frame. */
int ndt_bevmap_build(ndt_ctx* ctx, const float* map_xyzw, int64_t m, const double origin[3], ndt_bevmap** out);
int ndt_bevmap_destroy(ndt_bevmap* bm);

/* Features of B (scan, guess) pairs into out (B x 65536 bf16 bit patterns, host):
 * row b = [BEV of scans[scan_index[b]] | map BEV around guesses[b]]. */
int ndt_features(ndt_ctx* ctx, const ndt_bevmap* bm, const ndt_scans* scans, const int32_t* scan_index,
                 const double* guesses, int32_t B, uint16_t* out);

/* Registered guesses of B (scan, guess) pairs (see ndt_register_params):
 * out_guesses[b] for scan scans[scan_index[b]] from guesses[b]; out_score
 * (optional) receives the best score.  radius 0 copies the guesses. */
int ndt_bev_register(ndt_ctx* ctx, const ndt_bevmap* bm, const ndt_scans* scans, const int32_t* scan_index,
                     const double* guesses, int32_t B, const ndt_register_params* rp, double* out_guesses,
                     int32_t* out_score);

/* Creates a model with He-normal hidden weights (seeded), small output weights
 * and zero biases.  f0, f1, f2 must be multiples of 128; features need f0 = 65536.
 * A model holds its batch workspace and the captured learned step: it must not be
 * used by two threads at once. */
int ndt_model_create(ndt_ctx* ctx, int32_t f0, int32_t f1, int32_t f2, uint64_t seed, ndt_model** out);
int ndt_model_destroy(ndt_model* model);
int ndt_model_dims(const ndt_model* model, int32_t dims[3]);
/* The BEV registration the model runs behind (radius 0: none, the default).  A
 * model trained on crops around registered guesses needs the same setting:
 * ndt_model_predict and ndt_align_scans_learned register each guess first and
 * refine the registered guess. */
int ndt_model_set_register(ndt_model* model, const ndt_register_params* rp);
int ndt_model_get_register(const ndt_model* model, ndt_register_params* rp);
/* Float32 weights to / from a file, with the registration setting (Adam state is
 * not stored). */
int ndt_model_save(const ndt_model* model, const char* path);
int ndt_model_load(ndt_ctx* ctx, const char* path, ndt_model** out);
/* Copies parameter tensor `which` (NDT_W1 ... NDT_B3) to / from host float32. */
int ndt_model_get(const ndt_model* model, int32_t which, float* out, int64_t n);
int ndt_model_set(ndt_model* model, int32_t which, const float* in, int64_t n);

/* Forward pass on host features (B x f0 bf16 bit patterns, B a multiple of 128):
 * y (B x 6) receives the scaled outputs. */
int ndt_model_forward(ndt_ctx* ctx, ndt_model* model, const uint16_t* x, int32_t B, float* y);
/* Forward and backward of the loss L = (1/B) sum_b sum_o w_o (y_bo - t_bo)^2 on host
 * features; stores the gradients (read them with ndt_model_grad) and the loss. */
int ndt_model_backward(ndt_ctx* ctx, ndt_model* model, const uint16_t* x, const float* target,
                       const float* weights, int32_t B, float* loss);
int ndt_model_grad(const ndt_model* model, int32_t which, float* out, int64_t n);
/* One Adam step with the stored gradients (weight_decay applies to W1, W2, W3). */
int ndt_model_adam_step(ndt_ctx* ctx, ndt_model* model, float lr, float beta1, float beta2, float eps,
                        float weight_decay);

void ndt_train_params_default(ndt_train_params* tp);

/* Training triples: scan scan_index[i] of `scans`, the guess guesses[i] and the
 * target pose targets[i] (an NDT-converged pose), n samples.  Scan grids are
 * computed once and kept on the device with the guesses and scaled corrections.
 * The dataset holds `bm` (training crops its map grids) until it is destroyed. */
int ndt_dataset_create(ndt_ctx* ctx, const ndt_bevmap* bm, const ndt_scans* scans, const int32_t* scan_index,
                       const double* guesses, const double* targets, int64_t n, ndt_dataset** out);
/* Appends triples that refer to another scan batch (same map), so that datasets
 * larger than one resident batch of scan points can be built batch by batch. */
int ndt_dataset_add(ndt_dataset* ds, const ndt_scans* scans, const int32_t* scan_index, const double* guesses,
                    const double* targets, int64_t n);
/* Number of samples and of scan grids held. */
int ndt_dataset_size(const ndt_dataset* ds, int64_t* samples, int64_t* scans);
int ndt_dataset_destroy(ndt_dataset* ds);

/* Trains on the dataset; epoch_loss (optional, `epochs` values) receives the mean
 * loss of each epoch.  Drops the last partial batch of an epoch. */
int ndt_model_train(ndt_ctx* ctx, ndt_model* model, const ndt_dataset* ds, const ndt_train_params* tp,
                    float* epoch_loss);

/* Refined guesses of S scans: out_guesses[s] = G * C(model output), G = guesses[s]
 * or, with a registration set on the model, the registered guesses[s]. */
int ndt_model_predict(ndt_ctx* ctx, ndt_model* model, const ndt_bevmap* bm, const ndt_scans* scans,
                      const double* guesses, double* out_guesses);

/* Batched NDT with the learned warm start under `policy` (NDT_LEARNED_*).
 * used_learned[s] (optional) = 1 where the result comes from the refined guess.
 * passes in the results count every derivative pass spent on the scan.
 * ndt_ctx_last_timing reports the NDT alignment kernel(s) of the call and
 * ndt_ctx_last_step_time the device time of the whole call (features, model and
 * NDT).  The fallback policy runs as one graph, captured at the first call for a
 * (model, BEV map, scans, grid, parameters, rotation choice) and replayed after. */
int ndt_align_scans_learned(ndt_ctx* ctx, const ndt_grid* grid, const ndt_bevmap* bm, ndt_model* model,
                            const ndt_scans* scans, const double* guesses, const ndt_params* params, int32_t policy,
                            ndt_result* results, int32_t* used_learned);

/* Device milliseconds of the whole last ndt_align_scans_learned call on this
 * context (for the other alignment calls, the alignment kernels' time). */
int ndt_ctx_last_step_time(ndt_ctx* ctx, double* device_ms);

#ifdef __cplusplus
}
#endif

#endif
