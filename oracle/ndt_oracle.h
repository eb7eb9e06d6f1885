/*
 * ndt_oracle.h - CPU oracle for the NDT scan-matching engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libndthip.so and the DAPHNE harness) never links or calls it.
 *
 * Two semantics are restated here:
 *
 *  compat  - a line-faithful restatement of the reference CPU path,
 *            src/CPU/ndt_mapping/kernel.cpp (grid build :1302-1373, radius search
 *            :246-286, derivatives :438-560/:651-705, angle tables :707-825,
 *            solve :293-343, More-Thuente :425-436/:827-946/:1015-1144, matrix
 *            build :949-1013, Euler angles :1146-1173, driver :1175-1272/:1375-1414),
 *            including its quirks, with the same operation order and the same libm
 *            calls.  Compiled with -ffp-contract=off it is bit-exact with the
 *            reference compiled by g++ (pinned by tests/test_oracle_compat.py
 *            against fixtures written by the compiled reference).
 *
 *  ndt     - the textbook Magnusson/PCL NDT in float64 that the engine's "ndt" mode
 *            is measured against: hashed voxel grid (occupied voxels with at least
 *            min_points points, sample covariance, eigenvalues clamped to at least
 *            min_eig_ratio * largest), neighbours = voxels within one resolution of
 *            the point, Hessian per Magnusson (2009) eq. 6.13, live More-Thuente line
 *            search, translation update, R = Rx(roll) Ry(pitch) Rz(yaw), and a true
 *            convergence flag.  It shares with compat: the angle tables, point
 *            derivatives, score/gradient terms, solve and the More-Thuente helpers
 *            (those are pinned by the reference); the remainder is pinned by finite
 *            differences, closed forms and ground-truth recovery
 *            (tests/test_oracle_ndt.py).  A line-search trial at the step of the
 *            trial before it takes that pass's sums without a new pass (the pose is
 *            the same); repeat_passes = 1 makes the pass again, and both give the
 *            same poses bit for bit (tests/test_oracle_linesearch.py).
 */
#ifndef NDT_ORACLE_H
#define NDT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same layout as the reference's Voxel (src/CPU/ndt_mapping/datatypes.h:67-71). */
typedef struct orc_voxel {
    double icov[3][3];
    double mean[3];
    int32_t n;
} orc_voxel;

typedef struct orc_compat_grid {
    float minv[4], maxv[4];
    int32_t dims[3];
    int64_t ncells;
    orc_voxel* cells; /* dense, ncells entries, like target_cells_ */
} orc_compat_grid;

typedef struct orc_result {
    float final_f32[16];   /* row-major final transformation (float, as the reference) */
    double final_f64[16];  /* row-major final transformation in float64 (ndt) */
    double score;          /* NDT score at the final pose */
    double trans_probability; /* score / number of scan points */
    int32_t iterations;
    int32_t converged;
    int32_t passes;        /* derivative passes performed */
    int32_t pad;
} orc_result;

/* ---------------- shared helpers (pinned by the reference) ---------------- */
void orc_gauss_consts(double resolution, double outlier_ratio, double* d1, double* d2);
void orc_angle_tables(const double p[6], double jang[8][3], double hang[15][3]);
/* The tables of the textbook mode: orc_angle_tables with h_ang_d1_[2] negated
 * (the exact second derivative; see ndt_oracle.c). */
void orc_ndt_angle_tables(const double p[6], double jang[8][3], double hang[15][3]);
void orc_point_derivs(const double x[3], const double jang[8][3], const double hang[15][3],
                      double pg[3][6], double ph[18][6]);
void orc_solve6(double result[6], const double A[36], const double b[6]);
/* Textbook-mode solve: as orc_solve6 but the right-hand side is swapped with the rows. */
void orc_solve6_pivot(double result[6], const double A[36], const double b[6]);
int orc_mt_update_interval(double* a_l, double* f_l, double* g_l, double* a_u, double* f_u,
                           double* g_u, double a_t, double f_t, double g_t);
double orc_mt_trial_value(double a_l, double f_l, double g_l, double a_u, double f_u, double g_u,
                          double a_t, double f_t, double g_t);

/* More-Thuente line search (computeStepLengthMT, kernel.cpp:1015-1144) over a scalar
 * function.  phi(ctx, t, with_hessian, &dphi) returns phi(t) and writes phi'(t) for
 * a signed step t along the caller's direction; with_hessian tells an NDT caller
 * whether the evaluation is one whose Hessian is kept.  phi_0 = phi(0), d_phi_0 =
 * phi'(0).  The result: step length a >= 0 along sign * direction. */
typedef double (*orc_line_fn)(void* ctx, double t, int with_hessian, double* dphi);
enum {
    ORC_MT_ACCEPTED = 0,   /* sufficient decrease and curvature condition hold at a */
    ORC_MT_INTERVAL = 1,   /* the interval of uncertainty collapsed (updateIntervalMT) */
    ORC_MT_MAX_ITER = 2,   /* 10 loop iterations without an accepted step */
    ORC_MT_NO_DESCENT = 3  /* phi'(0) == 0: no step, no evaluation */
};
enum {
    ORC_MT_DEAD_LOOP = 1,         /* planted mistake for tests: the reference's start
                                     interval_converged = (step_max - step_min) > 0 (:1053) */
    ORC_MT_HESSIAN_AT_ACCEPT = 2  /* loop evaluations without the Hessian, one more with it at
                                     the returned step when the loop ran (:1123, :1141-1142) */
};
typedef struct orc_mt_out {
    double a;
    int32_t sign;            /* -1 when phi'(0) > 0 reversed the direction */
    int32_t evals;           /* evaluations of phi */
    int32_t loop_iterations; /* iterations of the More-Thuente loop */
    int32_t exit_reason;     /* ORC_MT_* */
} orc_mt_out;
int orc_mt_search(orc_line_fn phi, void* ctx, double phi_0, double d_phi_0, double step_init,
                  double step_max, double step_min, int32_t flags, orc_mt_out* out);

/* ---------------- compat (reference semantics) ---------------- */
int orc_compat_grid_build(const float* map4, int64_t m, float resolution, orc_compat_grid* g);
void orc_compat_grid_free(orc_compat_grid* g);
/* Returns the number of neighbours; out_cells receives dense cell indices in probe order. */
int orc_compat_radius_search(const orc_compat_grid* g, float resolution, const float pt[4],
                             double radius, int64_t out_cells[27], float out_dist[27]);
double orc_compat_update(double grad[6], double H[36], const double x_trans[3],
                         const double cinv[3][3], const double pg[3][6], const double ph[18][6],
                         double d1, double d2, int compute_hessian);
/* One derivative pass (computeDerivatives) at p over trans4 (transformed scan). */
double orc_compat_derivatives(const orc_compat_grid* g, float resolution, const float* scan4,
                              const float* trans4, int64_t n, const double p[6], double d1, double d2,
                              double grad[6], double H[36], int64_t* npairs);
void orc_compat_transform(const float* in4, float* out4, int64_t n, const float M[16]);
void orc_compat_build_matrix(float M[16], const double p[6]);
void orc_compat_euler(const float M[16], double out[3]);
/* partial_points_callback: grid build + alignment, reference parameters. */
int orc_compat_align(const float* map4, int64_t m, const float* scan4, int64_t n,
                     const float guess[16], orc_result* out);
/* Same, with the reference's OpenMP regions (src/OpenMP/ndt_mapping/kernel.cpp:390,1311,1341,1370). */
int orc_compat_align_omp(const float* map4, int64_t m, const float* scan4, int64_t n,
                         const float guess[16], orc_result* out);

/* ---------------- ndt (textbook float64) ---------------- */
typedef struct orc_ndt_params {
    double resolution;
    double step_size;
    double trans_eps;
    double outlier_ratio;
    int32_t max_iterations;
    int32_t min_points;
    double min_eig_ratio;
    int32_t hessian_at_accept; /* 1: the line search's trial passes skip the Hessian, which
                                  one more pass computes at the accepted step (the
                                  reference's order, kernel.cpp:1123,1141-1142); 0: every
                                  pass computes it (fewer passes, same poses) */
    int32_t repeat_passes;     /* 1: a line-search trial at the step of the trial before it
                                  (the trial value clamped to the same bound again) makes
                                  its pass again, as the reference's loop does; 0: it takes
                                  that pass's sums, which the pass would reproduce bit for
                                  bit (fewer passes, same poses) */
} orc_ndt_params;

void orc_ndt_params_default(orc_ndt_params* p);

typedef struct orc_ndt_voxel {
    int64_t key[3];
    double mean[3];
    double cov[9];
    double icov[9];
    int32_t n;
    int32_t valid;
} orc_ndt_voxel;

typedef struct orc_ndt_grid {
    double resolution;
    int64_t nvox;          /* occupied voxels (valid or not) */
    orc_ndt_voxel* vox;    /* sorted by key (z, y, x) */
    int64_t cap;           /* hash capacity (power of two) */
    int64_t* slot;         /* voxel index or -1 */
} orc_ndt_grid;

int orc_ndt_grid_build(const float* map4, int64_t m, const orc_ndt_params* prm, orc_ndt_grid* g);
void orc_ndt_grid_free(orc_ndt_grid* g);
int64_t orc_ndt_grid_find(const orc_ndt_grid* g, int64_t ix, int64_t iy, int64_t iz);
/* Symmetric 3x3 eigen-decomposition by cyclic Jacobi: evals ascending, evecs columns. */
void orc_sym3_eig(const double A[9], double evals[3], double evecs[9]);
void orc_ndt_pose_matrix(const double p[6], double M[16]);
void orc_ndt_euler(const double M[16], double out[3]);
double orc_ndt_derivatives(const orc_ndt_grid* g, const orc_ndt_params* prm, const float* scan4,
                           int64_t n, const double p[6], double grad[6], double H[36],
                           int compute_hessian, int64_t* npairs);
/* Score only (no derivatives) at pose p; used by finite-difference tests. */
double orc_ndt_score(const orc_ndt_grid* g, const orc_ndt_params* prm, const float* scan4, int64_t n,
                     const double p[6]);
int orc_ndt_align(const orc_ndt_grid* g, const orc_ndt_params* prm, const float* scan4, int64_t n,
                  const double guess[16], orc_result* out);
/* Aligns S scans spread across `threads` OpenMP threads (one scan per thread at a time). */
int orc_ndt_align_many(const orc_ndt_grid* g, const orc_ndt_params* prm, const float* const* scans,
                       const int64_t* counts, const double* guesses, int32_t S, int32_t threads,
                       orc_result* out);

#ifdef __cplusplus
}
#endif

#endif
