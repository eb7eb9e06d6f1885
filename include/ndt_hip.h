/*
 * ndt_hip.h - C ABI of the MI355X (gfx950) NDT scan-matching engine, libndthip.so.
 *
 * Drop-in boundary for the ndt_mapping kernel of the DAPHNE benchmark.  The
 * reference computes each testcase with
 *     CallbackResult partial_points_callback(PointCloud& scan, Matrix4f& guess, PointCloud& map)
 * (src/CPU/ndt_mapping/kernel.cpp:1405-1414, called from run() :1428); every entry
 * point below replaces one piece of that path and says which.
 *
 * Data layouts follow the reference (src/CPU/ndt_mapping/datatypes.h):
 *   points  - PointXYZI{float data[4]} (:11-13): float32 x, y, z, w/intensity, AoS;
 *   poses   - Matrix4f (:15-17): row-major 4x4 float32; this ABI also takes and
 *             returns row-major float64 poses (world = map-local + origin).
 * Ownership: the caller owns every host buffer (read-only during a call); the
 * library owns device memory behind opaque handles.  Nothing throws or exits
 * across the ABI: calls return an ndt_status and ndt_last_error() describes the
 * last failure of the calling thread.  A context owns one HIP stream; use one
 * context per host thread.
 *
 * Two semantics share the engine (ndt_params.mode):
 *   NDT_MODE_COMPAT - the reference's own arithmetic and quirks (dense-grid cell
 *                     indexing, quirky covariance and Hessian, two Newton steps of
 *                     0.1 x the Newton direction, translation left at the guess);
 *   NDT_MODE_NDT    - textbook Magnusson/PCL NDT (hashed voxels, eq. 6.13 Hessian,
 *                     live More-Thuente line search, true convergence flag).
 */
#ifndef NDT_HIP_H
#define NDT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ndt_status {
    NDT_OK = 0,
    NDT_ERR_INVALID = 1,     /* bad argument (null pointer, empty cloud, bad parameter, an
                                ndt-mode guess or pose with a non-finite entry) */
    NDT_ERR_HIP = 2,         /* HIP runtime error */
    NDT_ERR_NOMEM = 3,       /* device allocation failed (or would exceed the byte count in
                                the environment variable NDT_DEVICE_MEMORY_LIMIT) */
    NDT_ERR_TIMEOUT = 4,     /* a device-side wait exceeded its bound */
    NDT_ERR_UNSUPPORTED = 5, /* map extent or sizes beyond what the grid encodes */
} ndt_status;

enum { NDT_MODE_COMPAT = 0, NDT_MODE_NDT = 1 };
/* per-point derivative arithmetic */
enum { NDT_POINT_FP32 = 0, NDT_POINT_FP64 = 1 };
/* accumulation of score / gradient / Hessian terms within a thread */
enum { NDT_SUM_FP32 = 0, NDT_SUM_COMP32 = 1, NDT_SUM_FP64 = 2 };
/* scheduling of batched alignment: a work queue in which every scan advances on
 * its own (default), or rounds in which all active scans run one pass between
 * grid barriers; both give bit-identical results */
enum { NDT_SCHED_DEFAULT = 0, NDT_SCHED_ROUNDS = 1, NDT_SCHED_QUEUE = 2 };

/* NDT parameters.  The reference keeps them as members of ndt_mapping
 * (src/CPU/ndt_mapping/kernel.cpp:37-41,56-57); ndt_params_default() fills them. */
typedef struct ndt_params {
    int32_t mode;             /* NDT_MODE_COMPAT or NDT_MODE_NDT */
    float resolution;         /* voxel edge, metres (resolution_ :38) */
    double step_size;         /* maximum line-search step (step_size_ :40) */
    double trans_eps;         /* convergence threshold on the step (transformation_epsilon_ :56) */
    int32_t max_iterations;   /* Newton iterations (max_iterations_ :57) */
    double outlier_ratio;     /* (outlier_ratio_ :37) */
    int32_t min_points;       /* ndt: voxels need at least this many points */
    double min_eig_ratio;     /* ndt: eigenvalues clamped to >= ratio * largest */
    int32_t point_precision;  /* NDT_POINT_FP32 / NDT_POINT_FP64 (ndt) */
    int32_t sum_precision;    /* NDT_SUM_FP32 / NDT_SUM_COMP32 / NDT_SUM_FP64 (ndt) */
    double origin[3];         /* world position of the map-local frame (ndt) */
    int32_t block_size;       /* 0 = default (256); threads per workgroup (ndt align) */
    int32_t points_per_wg;    /* scan points per partial sum (ndt align); 0 = 2048 (batched
                                 throughput), 256 suits single-scan latency */
    int32_t scheduler;        /* NDT_SCHED_* (ndt align) */
    int32_t hessian_at_accept; /* ndt align: 0 (default) = every derivative pass also sums the
                                  Hessian, so the accepted step's Hessian is at hand; 1 = the
                                  line-search loop's trial passes sum score and gradient only and
                                  one more pass computes the Hessian at the accepted step, the
                                  reference's order (computeDerivatives(..., false) at
                                  kernel.cpp:1123, computeHessian at :1141-1142).  Same poses,
                                  more passes. */
    int32_t fine_pass;        /* ndt align: passes of a run from this one on cut a scan into
                                 chunks a quarter the size (at least 256 points), spreading the
                                 late passes of the slowest scans over more workgroups; 0 = 24.
                                 Part of the launch shape: results depend on it at the level of
                                 rounding, as on points_per_wg */
} ndt_params;

typedef struct ndt_result {
    float final_transformation[16];   /* row-major, as CallbackResult.final_transformation */
    double final_transformation_f64[16];
    double fitness_score;             /* CallbackResult.fitness_score; 0 unless computed */
    double score;                     /* NDT score at the final pose */
    double trans_probability;         /* score / scan points */
    int32_t iterations;               /* Newton iterations */
    int32_t converged;                /* CallbackResult.converged */
    int32_t passes;                   /* derivative passes over the scan */
    int32_t status;                   /* per-scan ndt_status */
} ndt_result;

typedef struct ndt_grid_info {
    int64_t map_points;
    int64_t voxels;          /* occupied voxels */
    int64_t valid_voxels;    /* voxels used for matching (ndt) */
    int64_t query_cells;     /* hashed lookup cells (ndt: occupied cells dilated by one) */
    int64_t candidates;      /* candidate entries of the lookup lists */
    int64_t hash_capacity;
    int64_t dense_cells;     /* compat: cells of the reference's dense grid; ndt: cells of the
                                dense query-cell index, 0 when the index is hashed */
    int32_t dims[3];         /* compat: dense grid dimensions */
    float min_corner[3];     /* compat: minVoxel */
    float max_corner[3];     /* compat: maxVoxel */
    double build_ms;         /* device time of the build */
} ndt_grid_info;

typedef struct ndt_ctx ndt_ctx;
typedef struct ndt_grid ndt_grid;
typedef struct ndt_scans ndt_scans;

/* Fills the defaults of a mode: compat = the reference's members (outlier 0.55,
 * resolution 1, step 0.1, epsilon 0.1, max_iterations 0); ndt = resolution 1,
 * step 0.1, epsilon 0.01, 30 iterations, outlier 0.55, 6 points per voxel,
 * eigenvalue ratio 0.01, fp64 point math, fp64 sums. */
void ndt_params_default(ndt_params* params, int32_t mode);

/* Message of the last failed call on this thread ("" if none). */
const char* ndt_last_error(void);

/* Library version string. */
const char* ndt_version(void);

/* Creates a context on HIP device `device` (one stream, device workspace).
 * Replaces the reference's per-process setup in ndt_mapping::init()
 * (src/CPU/ndt_mapping/kernel.cpp:345-377) for the compute side. */
int ndt_ctx_create(int32_t device, ndt_ctx** out);
/* Ends the caller's use of the context.  Grids, scan batches and learned-model
 * handles made on it keep it alive until they are destroyed, so they may be
 * destroyed in any order; it is not passed to other calls after this. */
int ndt_ctx_destroy(ndt_ctx* ctx);
/* Number of compute units and the device name. */
int ndt_ctx_device_info(ndt_ctx* ctx, int32_t* compute_units, char* name, int32_t name_len);

/* Builds the voxel grid of a map on the GPU.  Replaces initCompute()
 * (src/CPU/ndt_mapping/kernel.cpp:1302-1373; CUDA variant
 * src/Cuda/ndt_mapping/kernel.cu:1406-1451).  map_xyzw: m points, 4 floats each,
 * map-local coordinates (params->origin gives their world offset in ndt mode). */
int ndt_grid_build(ndt_ctx* ctx, const float* map_xyzw, int64_t m, const ndt_params* params,
                   ndt_grid** out);
int ndt_grid_info_get(const ndt_grid* grid, ndt_grid_info* info);
int ndt_grid_destroy(ndt_grid* grid);
/* Copies the occupied voxels to host memory (tests / diagnostics).
 * ndt: keys int64[3*k] (cell coordinates), n int32[k], mean f64[3*k], icov f64[9*k],
 *      valid int32[k], for the k = min(cap, voxels) voxels in key order (z, y, x).
 * compat: keys = dense cell index in keys[3*i], icov is the full 3x3 matrix.
 * Any output pointer may be null. */
int ndt_grid_voxels(const ndt_grid* grid, int64_t cap, int64_t* keys, int32_t* n, double* mean,
                    double* icov, int32_t* valid);

/* Aligns one scan against a grid.  Replaces partial_points_callback's
 * ndt_align / computeTransformation (src/CPU/ndt_mapping/kernel.cpp:1175-1272,
 * :1375-1392) for a grid that is already built.  guess: row-major float64 world
 * pose.  scan_xyzw: n points (x, y, z, w) in the sensor frame. */
int ndt_align(ndt_ctx* ctx, const ndt_grid* grid, const float* scan_xyzw, int64_t n,
              const double guess[16], const ndt_params* params, ndt_result* result);

/* Aligns S scans in one launch (ndt mode): per-scan Newton state stays on the
 * device, scans of uneven size share the GPU.  No counterpart in the reference,
 * which processes one testcase at a time (kernel.cpp:1425-1429). */
int ndt_align_batch(ndt_ctx* ctx, const ndt_grid* grid, const float* const* scans,
                    const int64_t* counts, int32_t S, const double* guesses,
                    const ndt_params* params, ndt_result* results);

/* Device-resident scan batches, so that timed runs start with inputs in HBM.
 * ndt_align_scans captures its submission into a graph held by the scans handle (and
 * pinned guess / result buffers with it) and replays it on later calls with the same
 * grid and parameters: a scans handle must not be used by two threads at once, even
 * through the const pointer (each thread uploads its own batch). */
int ndt_scans_upload(ndt_ctx* ctx, const float* const* scans, const int64_t* counts, int32_t S,
                     ndt_scans** out);
int ndt_scans_destroy(ndt_scans* scans);
int ndt_align_scans(ndt_ctx* ctx, const ndt_grid* grid, const ndt_scans* scans,
                    const double* guesses, const ndt_params* params, ndt_result* results);

/* The reference's per-testcase callback in compat mode: builds the grid of `map`
 * and aligns `scan` from `guess` with the reference's semantics.  Replaces
 * partial_points_callback (src/CPU/ndt_mapping/kernel.cpp:1405-1414). */
int ndt_compat_callback(ndt_ctx* ctx, const float* scan_xyzw, int64_t n, const float guess[16],
                        const float* map_xyzw, int64_t m, ndt_result* result);

/* One derivative pass: score, gradient (6) and Hessian (6x6, row-major) of the
 * scan at pose parameters p = (x, y, z, roll, pitch, yaw) (map-local frame).
 * Replaces computeDerivatives (src/CPU/ndt_mapping/kernel.cpp:651-705); in compat
 * mode the pose matrix is the float Matrix4f `compat_transform` (row-major) with
 * which the reference transforms the scan, p gives the angle tables. */
int ndt_derivatives(ndt_ctx* ctx, const ndt_grid* grid, const float* scan_xyzw, int64_t n,
                    const double p[6], const float compat_transform[16], const ndt_params* params,
                    double* score, double gradient[6], double hessian[36], int64_t* pairs);

/* Fitness score of S device-resident scans at world poses (row-major float64,
 * S x 16): mean squared distance from each transformed scan point to its nearest
 * map point, over the points whose nearest map point lies within max_range; the
 * largest double when no point qualifies.  This is PCL's
 * Registration::getFitnessScore, which the reference declares as
 * ndt_getFitnessScore (src/Cuda/ndt_mapping/kernel.cu:124,
 * src/OpenCl/ndt_mapping/kernel.cpp:212) but never implements.  max_range <= 0 or
 * infinite searches the cells within Chebyshev distance 16 of each point's cell
 * (points whose nearest map point lies beyond count as having none); a finite
 * max_range up to 15 resolutions is exact.  counts (optional) receives the number of points used.
 * Needs an ndt-mode grid. */
int ndt_fitness_score(ndt_ctx* ctx, const ndt_grid* grid, const ndt_scans* scans, const double* poses,
                      double max_range, double* fitness, int64_t* counts);

/* Autotuner of the batched align launch: times every candidate (block size 128,
 * 256, 512 x chunk 1024, 2048, 4096 points) on the given device-resident batch
 * and guesses, and writes the fastest into params->block_size and
 * params->points_per_wg; best_ms (optional) receives its kernel time.  With a
 * cache_path, the choice is stored per device, precision and batch shape (scan
 * count, total points to 3 significant digits) and later calls with the same key
 * read it back without timing.  The chunk size changes the rounding of the sums:
 * results depend on it as they depend on any explicit points_per_wg.  No
 * counterpart in the reference, which processes one testcase at a time. */
int ndt_autotune(ndt_ctx* ctx, const ndt_grid* grid, const ndt_scans* scans, const double* guesses,
                 ndt_params* params, const char* cache_path, double* best_ms);

/* ndt_align_scans for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference: it exists for those callers only.  Same arguments,
 * results and kernels as ndt_align_scans (with the work-queue scheduler, results
 * equal bit for bit), but every workspace buffer lives in a fixed device array of
 * 64 MiB per device, held from the library's first use of the device until the
 * process ends; the call allocates nothing.  A batch whose workspace (about
 * 2.9 KB per partial-sum slot at 30 iterations, plus 1.5 KB per scan) exceeds the
 * array returns NDT_ERR_UNSUPPORTED, as does scheduler NDT_SCHED_ROUNDS.  All
 * calls on one device share the array: calls that overlap on a device (from two
 * threads or contexts) race and must be serialised by the caller. */
int ndt_align_scans_noalloc(ndt_ctx* ctx, const ndt_grid* grid, const ndt_scans* scans,
                            const double* guesses, const ndt_params* params, ndt_result* results);

/* Timing of the last ndt_align / ndt_align_batch / ndt_align_scans call on this
 * context: device milliseconds of the alignment kernel(s), measured with HIP
 * events on the context's stream, and the number of kernel launches. */
int ndt_ctx_last_timing(ndt_ctx* ctx, double* align_kernel_ms, int32_t* launches, int64_t* pairs);

/* Phase profile of the last ndt alignment on this context, in microseconds as seen
 * by the first workgroup of the launch: building the list of active scans, point
 * compute + workgroup reductions, grid barriers, and the per-scan reductions and
 * Newton / line-search controllers. */
int ndt_ctx_last_profile(ndt_ctx* ctx, double phases_us[4]);

#ifdef __cplusplus
}
#endif

#endif
