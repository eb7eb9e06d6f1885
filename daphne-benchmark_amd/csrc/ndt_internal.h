/*
 * ndt_internal.h - structures and float64 math shared by the gfx950 kernels
 * (ndt_kernels.hip, compiled by hipcc) and the host side of the C ABI
 * (ndt_api.cpp, compiled by g++ with -ffp-contract=off).
 *
 * NDT_HD marks functions compiled for both the host and the device.
 * Citations are to the reference tree (src/CPU/ndt_mapping/kernel.cpp unless
 * another file is named).
 */
#ifndef NDT_INTERNAL_H
#define NDT_INTERNAL_H

#include <stdint.h>
#include <math.h>

#include "ndt_hip.h"

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define NDT_HD __host__ __device__ inline
#else
#include <hip/hip_runtime_api.h>
#define NDT_HD inline
#endif

namespace ndt {

constexpr int NACC = 28;        /* score, gradient (6), Hessian upper triangle (21) */
constexpr int NACC_COMPAT = 43; /* score, gradient (6), full Hessian (36) */
constexpr int NRED = NACC + 1;  /* + pair count */
constexpr int NRED_COMPAT = NACC_COMPAT + 1;
constexpr uint64_t EMPTY_KEY = ~0ull;
constexpr int MAX_GROUP = 1024; /* workgroups cooperating on one scan */

/* ---------------------------------------------------------------- math */

NDT_HD double dmin_std(double a, double b) { return (b < a) ? b : a; } /* std::min */
NDT_HD double dmax_std(double a, double b) { return (a < b) ? b : a; } /* std::max */

NDT_HD double dot6(const double* a, const double* b) {
    /* dot_product6, kernel.cpp:417-420 */
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4] + a[5] * b[5];
}

/* Upper-triangle index of (i, j), i <= j, row-major over the 6x6 Hessian. */
NDT_HD int tri_index(int i, int j) { return i * 6 - (i * (i - 1)) / 2 + (j - i); }

NDT_HD void gauss_consts(double resolution, double outlier_ratio, double* d1, double* d2) {
    /* kernel.cpp:1181-1185 (Magnusson 2009 eq. 6.8) */
    double c1 = 10 * (1 - outlier_ratio);
    double c2 = outlier_ratio / pow(resolution, 3);
    double d3 = -log(c2);
    *d1 = -log(c1 + c2) - d3;
    *d2 = -2 * log((-log(c1 * exp(-0.5) + c2) - d3) / *d1);
}

/* Angular derivative tables (eq. 6.19 / 6.21), kernel.cpp:707-825, including the
 * small-angle rule |angle| < 1e-4 -> cos = 1, sin = 0.  j: a..h, h: a2 a3 b2 b3 c2
 * c3 d1 d2 d3 e1 e2 e3 f1 f2 f3. */
NDT_HD void angle_tables(const double* p, double j[8][3], double h[15][3]) {
    double cx, cy, cz, sx, sy, sz;
    if (fabs(p[3]) < 10e-5) { cx = 1.0; sx = 0.0; } else { cx = cos(p[3]); sx = sin(p[3]); }
    if (fabs(p[4]) < 10e-5) { cy = 1.0; sy = 0.0; } else { cy = cos(p[4]); sy = sin(p[4]); }
    if (fabs(p[5]) < 10e-5) { cz = 1.0; sz = 0.0; } else { cz = cos(p[5]); sz = sin(p[5]); }
    j[0][0] = (-sx * sz + cx * sy * cz); j[0][1] = (-sx * cz - cx * sy * sz); j[0][2] = (-cx * cy);
    j[1][0] = (cx * sz + sx * sy * cz);  j[1][1] = (cx * cz - sx * sy * sz);  j[1][2] = (-sx * cy);
    j[2][0] = (-sy * cz);                j[2][1] = sy * sz;                   j[2][2] = cy;
    j[3][0] = sx * cy * cz;              j[3][1] = (-sx * cy * sz);           j[3][2] = sx * sy;
    j[4][0] = (-cx * cy * cz);           j[4][1] = cx * cy * sz;              j[4][2] = (-cx * sy);
    j[5][0] = (-cy * sz);                j[5][1] = (-cy * cz);                j[5][2] = 0;
    j[6][0] = (cx * cz - sx * sy * sz);  j[6][1] = (-cx * sz - sx * sy * cz); j[6][2] = 0;
    j[7][0] = (sx * cz + cx * sy * sz);  j[7][1] = (cx * sy * cz - sx * sz);  j[7][2] = 0;
    h[0][0] = (-cx * sz - sx * sy * cz);  h[0][1] = (-cx * cz + sx * sy * sz); h[0][2] = sx * cy;
    h[1][0] = (-sx * sz + cx * sy * cz);  h[1][1] = (-cx * sy * sz - sx * cz); h[1][2] = (-cx * cy);
    h[2][0] = (cx * cy * cz);             h[2][1] = (-cx * cy * sz);           h[2][2] = (cx * sy);
    h[3][0] = (sx * cy * cz);             h[3][1] = (-sx * cy * sz);           h[3][2] = (sx * sy);
    h[4][0] = (-sx * cz - cx * sy * sz);  h[4][1] = (sx * sz - cx * sy * cz);  h[4][2] = 0;
    h[5][0] = (cx * cz - sx * sy * sz);   h[5][1] = (-sx * sy * cz - cx * sz); h[5][2] = 0;
    h[6][0] = (-cy * cz);                 h[6][1] = (cy * sz);                 h[6][2] = (sy);
    h[7][0] = (-sx * sy * cz);            h[7][1] = (sx * sy * sz);            h[7][2] = (sx * cy);
    h[8][0] = (cx * sy * cz);             h[8][1] = (-cx * sy * sz);           h[8][2] = (-cx * cy);
    h[9][0] = (sy * sz);                  h[9][1] = (sy * cz);                 h[9][2] = 0;
    h[10][0] = (-sx * cy * sz);           h[10][1] = (-sx * cy * cz);          h[10][2] = 0;
    h[11][0] = (cx * cy * sz);            h[11][1] = (cx * cy * cz);           h[11][2] = 0;
    h[12][0] = (-cy * cz);                h[12][1] = (cy * sz);                h[12][2] = 0;
    h[13][0] = (-cx * sz - sx * sy * cz); h[13][1] = (-cx * cz + sx * sy * sz); h[13][2] = 0;
    h[14][0] = (-sx * sz + cx * sy * cz); h[14][1] = (-cx * sy * sz - sx * cz); h[14][2] = 0;
}

/* 6x6 solve, kernel.cpp:293-343: Gaussian elimination with partial pivoting by
 * column maximum; an all-zero pivot column gets 0.001 on the diagonal. */
NDT_HD void solve6(double* result, const double* Ain, const double* bin) {
    double A[6][6], b[6];
    for (int i = 0; i < 6; i++) {
        b[i] = bin[i];
        for (int k = 0; k < 6; k++) A[i][k] = Ain[i * 6 + k];
    }
    for (int j = 0; j < 6; j++) {
        double mx = fabs(A[j][j]);
        int mi = j;
        for (int i = j + 1; i < 6; i++)
            if (fabs(A[i][j]) > mx) { mi = i; mx = fabs(A[i][j]); }
        if (mi != j)
            for (int i = 0; i < 6; i++) { double t = A[mi][i]; A[mi][i] = A[j][i]; A[j][i] = t; }
        if (mx == 0.0) A[j][j] = 0.001;
        for (int i = j + 1; i < 6; i++) {
            double pivot = A[i][j] / A[j][j];
            for (int k = 0; k < 6; k++) A[i][k] = A[i][k] - pivot * A[j][k];
            b[i] = b[i] - pivot * b[j];
        }
    }
    result[5] = b[5] / A[5][5];
    for (int i = 4; i >= 0; i--) {
        double sum = 0.0;
        for (int j = i + 1; j < 6; j++) sum = sum + A[i][j] * result[j];
        result[i] = (b[i] - sum) / A[i][i];
    }
}

NDT_HD double psi_mt(double a, double f_a, double f_0, double g_0, double mu) {
    return (f_a - f_0 - mu * g_0 * a); /* auxilaryFunction_PsiMT, kernel.cpp:433-436 */
}
NDT_HD double dpsi_mt(double g_a, double g_0, double mu) {
    return (g_a - mu * g_0); /* auxilaryFunction_dPsiMT, kernel.cpp:425-428 */
}

/* updateIntervalMT, kernel.cpp:827-864 (More & Thuente 1994). Returns 1 when the
 * interval has converged. */
NDT_HD int mt_update_interval(double& a_l, double& f_l, double& g_l, double& a_u, double& f_u, double& g_u,
                              double a_t, double f_t, double g_t) {
    if (f_t > f_l) {
        a_u = a_t; f_u = f_t; g_u = g_t;
        return 0;
    } else if (g_t * (a_l - a_t) > 0) {
        a_l = a_t; f_l = f_t; g_l = g_t;
        return 0;
    } else if (g_t * (a_l - a_t) < 0) {
        a_u = a_l; f_u = f_l; g_u = g_l;
        a_l = a_t; f_l = f_t; g_l = g_t;
        return 0;
    }
    return 1;
}

/* trialValueSelectionMT, kernel.cpp:867-946 (Sun & Yuan 2006 eqs. 2.4.2, 2.4.5,
 * 2.4.52, 2.4.56). */
NDT_HD double mt_trial_value(double a_l, double f_l, double g_l, double a_u, double f_u, double g_u,
                             double a_t, double f_t, double g_t) {
    if (f_t > f_l) {
        double z = 3 * (f_t - f_l) / (a_t - a_l) - g_t - g_l;
        double w = sqrt(z * z - g_t * g_l);
        double a_c = a_l + (a_t - a_l) * (w - g_l - z) / (g_t - g_l + 2 * w);
        double a_q = a_l - 0.5 * (a_l - a_t) * g_l / (g_l - (f_l - f_t) / (a_l - a_t));
        if (fabs(a_c - a_l) < fabs(a_q - a_l)) return a_c;
        return 0.5 * (a_q + a_c);
    } else if (g_t * g_l < 0) {
        double z = 3 * (f_t - f_l) / (a_t - a_l) - g_t - g_l;
        double w = sqrt(z * z - g_t * g_l);
        double a_c = a_l + (a_t - a_l) * (w - g_l - z) / (g_t - g_l + 2 * w);
        double a_s = a_l - (a_l - a_t) / (g_l - g_t) * g_l;
        if (fabs(a_c - a_t) >= fabs(a_s - a_t)) return a_c;
        return a_s;
    } else if (fabs(g_t) <= fabs(g_l)) {
        double z = 3 * (f_t - f_l) / (a_t - a_l) - g_t - g_l;
        double w = sqrt(z * z - g_t * g_l);
        double a_c = a_l + (a_t - a_l) * (w - g_l - z) / (g_t - g_l + 2 * w);
        double a_s = a_l - (a_l - a_t) / (g_l - g_t) * g_l;
        double a_t_next = (fabs(a_c - a_t) < fabs(a_s - a_t)) ? a_c : a_s;
        if (a_t > a_l) return dmin_std(a_t + 0.66 * (a_u - a_t), a_t_next);
        return dmax_std(a_t + 0.66 * (a_u - a_t), a_t_next);
    } else {
        double z = 3 * (f_t - f_u) / (a_t - a_u) - g_t - g_u;
        double w = sqrt(z * z - g_t * g_u);
        return (a_u + (a_t - a_u) * (w - g_u - z) / (g_t - g_u + 2 * w));
    }
}

/* ndt pose: T = [Rx(p3) Ry(p4) Rz(p5) | p0..2], the convention of angle_tables,
 * from the angles' cosines c[0..2] and sines s[0..2]. */
NDT_HD void pose_matrix_cs(const double* p, const double* c, const double* s, double* M) {
    const double cx = c[0], sx = s[0], cy = c[1], sy = s[1], cz = c[2], sz = s[2];
    M[0] = cy * cz;
    M[1] = -cy * sz;
    M[2] = sy;
    M[4] = cx * sz + sx * sy * cz;
    M[5] = cx * cz - sx * sy * sz;
    M[6] = -sx * cy;
    M[8] = sx * sz - cx * sy * cz;
    M[9] = sx * cz + cx * sy * sz;
    M[10] = cx * cy;
    M[3] = p[0];
    M[7] = p[1];
    M[11] = p[2];
    M[12] = M[13] = M[14] = 0.0;
    M[15] = 1.0;
}

NDT_HD void pose_matrix(const double* p, double* M) {
    const double c[3] = {cos(p[3]), cos(p[4]), cos(p[5])};
    const double s[3] = {sin(p[3]), sin(p[4]), sin(p[5])};
    pose_matrix_cs(p, c, s, M);
}

/* inverse of pose_matrix for |pitch| < pi/2 */
NDT_HD void pose_euler(const double* M, double* out) {
    double s = M[2];
    if (s > 1.0) s = 1.0;
    if (s < -1.0) s = -1.0;
    out[0] = atan2(-M[6], M[10]);
    out[1] = asin(s);
    out[2] = atan2(-M[1], M[0]);
}

/* ---------------------------------------------------------------- grid */

/* Hashed lookup table slot: a cell key and the [begin, begin + count) range of
 * its entry list. */
struct HashSlot {
    uint64_t key;
    uint32_t begin;
    uint32_t count;
};

/* Fibonacci hashing of a packed cell key. */
NDT_HD uint64_t hash_slot(uint64_t key, int log2cap) {
    return (key * 0x9E3779B97F4A7C15ull) >> (64 - log2cap);
}

/* Device view of an ndt-mode grid. */
/* Candidates: one entry per (query cell, valid voxel within reach), the entries
 * of a query cell contiguous in probe order, in two parallel arrays:
 *   head  float4 (mean - query cell corner as float32, voxel id bits): the range
 *         prefilter reads only these, 16 bytes per candidate;
 *   s32   2 x float4 (c00 c01 c02 c11)(c12 c22 0 0): float32 inverse covariance.
 * The float64 statistics are stored once per valid voxel, indexed by voxel id:
 *   vs64  5 x double2 (mx my)(mz c00)(c01 c02)(c11 c12)(c22 0): float64 mean and
 *         inverse covariance, 80 bytes per voxel (27 times less than per
 *         candidate, so that they stay in the caches). */
/* double2 entries per voxel in vs64 (80 bytes; padding records to 128 bytes, one
 * cache line each, made the alignment launch 1 % slower) */
constexpr int VS64_STRIDE = 5;
struct NdtGridDev {
    const uint2* dense;      /* query cell -> (begin, count), dense over the map's cell box, or null */
    int32_t dims[3];         /* cells per axis of the dense index */
    const HashSlot* slots;   /* query cell -> candidate range (when dense is null) */
    const float4* head;      /* candidates: relative mean, voxel id */
    const float4* s32;       /* candidates: float32 statistics (2 per entry) */
    const double2* vs64;     /* valid voxels: float64 statistics (5 per voxel) */
    int log2cap;
    int bits[3];             /* key bit widths x, y, z */
    int64_t origin_cell[3];  /* cell coordinates of key 0 */
    double origin_cell_d[3]; /* the same as float64 (exact: |cell| < 2^53) */
    float resolution;
    float inv_resolution;
    double resolution_d;
    double inv_resolution_d; /* 1 / resolution, exact when res_pow2 */
    int32_t res_pow2;        /* the resolution is a power of two */
};

/* Device view of the map points of an ndt grid, sorted by voxel, for nearest-point
 * queries (fitness score). */
struct FitGridDev {
    const HashSlot* slots;   /* occupied voxel -> range of sorted points */
    const float4* pts;       /* map points sorted by voxel key */
    int log2cap;
    int bits[3];
    int64_t origin_cell[3];
    double resolution;
};
constexpr int FIT_MAX_SHELLS = 16; /* nearest-point search radius cap, in cells */

/* Device view of a compat-mode grid (the reference's dense indexing, hashed). */
struct CompatGridDev {
    const HashSlot* slots;   /* dense cell index -> voxel id (begin) */
    const double* mean;      /* 3 per voxel */
    const double* icov;      /* 9 per voxel, full matrix */
    int log2cap;
    int dims[3];
    float minv[3];
    float maxv[3];
    float resolution;
};

/* ---------------------------------------------------------------- align */

/* PH_MT_HESS: one more pass at an accepted step whose line-search loop ran with
 * gradient-only passes (hess_at_accept) */
/* PH_PROBE: the pass at the given guess that opens a launch with refined starts
 * (its score decides the fallback) */
enum Phase : int32_t { PH_INIT = 0, PH_MT_FIRST = 1, PH_MT_LOOP = 2, PH_DONE = 3, PH_MT_HESS = 4, PH_PROBE = 5 };
/* runs of a scan in a launch with refined starts: the probe at the given guess,
 * NDT from the refined guess (when it scores at least as high there), then plain
 * NDT from the given guess (when it does not, or when the refined run ended with
 * a lower score than the given guess has) */
enum FbStage : int32_t { FB_NONE = 0, FB_PROBE = 1, FB_REFINED = 2, FB_FALLBACK = 3 };
/* the pass a scan in this phase runs needs no Hessian (with hess_at_accept) */
NDT_HD bool grad_only_pass(int32_t phase) { return phase == PH_MT_LOOP; }

/* Per-scan Newton / line-search state, resident in device memory for the whole
 * alignment.  Exchanged between workgroups as 64-bit words. */
struct ScanState {
    double p[6];       /* accepted pose parameters (map-local frame) */
    double x_t[6];     /* pose of the pass being evaluated */
    double dir[6];     /* normalised Newton direction */
    double score;
    double grad[6];
    double hess[21];   /* upper triangle */
    double phi_0, d_phi_0, a_l, f_l, g_l, a_u, f_u, g_u, a_t;
    double final_m[16];
    double pairs;      /* (point, voxel) pairs summed over all passes */
    double pp[32];     /* pass parameters of x_t (rotation, translation, axes, constants) */
    double alt[6];     /* launches with refined starts: the start not being run (the refined
                          guess during the probe, then the given guess) */
    double probe_score; /* score at the given guess (the probe pass) */
    double probe_d[27];  /* gradient and Hessian there: a run from the given guess starts
                            from them (the probe is its first pass) */
    int32_t fb_stage, used_alt; /* FbStage; 1 when the result is the refined run's */
    int32_t iter, converged;
    int32_t phase, done;
    int32_t step_iterations, open_interval;
    int32_t interval_converged, passes;
    int32_t status;
    int32_t run_base;  /* pass count at the start of the current run (chunking follows the
                          pass number within the run, so a run in a launch with refined
                          starts splits and sums its passes as a launch of its own does) */
};
/* pass number within the scan's current run */
NDT_HD uint32_t run_pass(const ScanState& st) { return (uint32_t)(st.passes - st.run_base); }
constexpr int STATE_WORDS = sizeof(ScanState) / 8;
static_assert(sizeof(ScanState) % 8 == 0, "state must be whole 64-bit words");

struct ScanDesc {
    int64_t offset;     /* first point in the packed point buffer */
    int64_t chunk_base; /* first partial-sum slot of this scan */
    int32_t n;          /* points */
    int32_t nchunks;    /* chunks of `chunk` points */
    int32_t chunk;      /* points per partial sum */
    int32_t fnchunks;   /* chunks of fine_chunk(chunk) points (passes from fine_pass on) */
    int64_t qbase;      /* queue position of the scan's first-pass chunks */
    int32_t fine_pass;  /* first pass of a run in fine chunks */
    int32_t pad;
};

/* Passes of a run from fine_pass on (default DEFAULT_FINE_PASS) cut a scan into
 * chunks a quarter the size (at least 256 points): the scans that need many passes
 * set a batch's finish time, and finer chunks spread each of their late passes
 * over more workgroups (on the 512-scan, 1 M-point-map batch from its given
 * guesses: 6.45 ms against 7.0 ms with one chunk size; passes from 20, 24, 32 or 36
 * on, chunks an eighth or half the size, or a further level of sixteenth-size
 * chunks from pass 36, 40 or 44 on, gave 6.45-7.0 ms; with starts a registration
 * put near the optimum, from pass 12 on, and pass 8 once leveled starts shortened
 * the runs further).  Once a line-search trial at the step
 * of the trial before it no longer made a pass (runs about 12 % shorter), pass 24
 * did best on that batch: 5.55 ms against 5.65 at pass 28, 5.63 at 22, 5.76 at 20
 * and 6.10 at 16.  The chunking is a function of the pass number only, so results
 * stay independent of scheduling. */
constexpr int32_t DEFAULT_FINE_PASS = 24;
#ifndef NDT_FINE_DIV
#define NDT_FINE_DIV 4
#endif
NDT_HD int32_t fine_chunk(int32_t chunk) {
    const int32_t q = chunk / NDT_FINE_DIV;
    return q >= 256 ? q : (chunk < 256 ? chunk : 256);
}
NDT_HD int32_t pass_chunk(const ScanDesc& d, uint32_t pass) {
    return pass >= (uint32_t)d.fine_pass ? fine_chunk(d.chunk) : d.chunk;
}
NDT_HD int32_t pass_nchunks(const ScanDesc& d, uint32_t pass) {
    return pass >= (uint32_t)d.fine_pass ? d.fnchunks : d.nchunks;
}

struct AlignParams {
    double step_size, trans_eps, d1, d2;
    double resolution_d;
    double origin[3];
    int32_t max_iterations;
    int32_t point_fp64;
    int32_t sum_mode;
    int32_t nscans;
    int32_t hess_at_accept; /* line-search loop passes without the Hessian (ndt_params) */
    int32_t pad;
    uint64_t wait_ticks; /* bound of one device-side wait, s_memrealtime ticks (100 MHz) */
};

/* grid-barrier state of one cooperative align launch: arrivals so far (grows by
 * the workgroup count per barrier) and the abort flag of a timed-out wait */
struct RoundSync {
    uint32_t count;
    uint32_t abort;
    uint32_t pad[62];
};

/* Work queue of one asynchronous align launch: items are (pass, scan, chunk)
 * triples, written once each at increasing positions; head = next position to
 * claim, tail = next position to fill, remaining = scans not finished. */
struct WorkQueue {
    uint32_t head;
    uint32_t tail;
    uint32_t remaining;
    uint32_t abort;
    uint32_t pad[60];
};
constexpr uint64_t ITEM_VALID = 1ull << 63;
constexpr int ITEM_CHUNK_BITS = 20;
constexpr int ITEM_SCAN_BITS = 20;
constexpr int ITEM_PASS_BITS = 23;
NDT_HD uint64_t make_item(uint32_t pass, uint32_t scan, uint32_t chunk) {
    return ITEM_VALID | ((uint64_t)pass << (ITEM_CHUNK_BITS + ITEM_SCAN_BITS)) | ((uint64_t)scan << ITEM_CHUNK_BITS) |
           (uint64_t)chunk;
}
/* upper bound of the derivative passes of one scan: the first pass, then per
 * Newton iteration at most 1 + 10 line-search evaluations and, with
 * hess_at_accept, one Hessian pass; a launch with refined starts adds the probe
 * and a second run */
NDT_HD int64_t max_passes(int32_t max_iterations) { return 1 + 12 * (int64_t)max_iterations; }
NDT_HD int64_t max_passes_fallback(int32_t max_iterations) { return 1 + 2 * max_passes(max_iterations); }

/* per-scan output written by the device */
struct ScanOut {
    double final_m[16];
    double score;
    double pairs;
    double t_prologue, t_compute, t_barrier, t_control; /* microseconds, workgroup 0 of the scan */
    int32_t iterations, converged, passes, status;
    int32_t used_alt, pad; /* launches with refined starts: 1 when the refined run was kept */
};

/* ---------------------------------------------------------------- launchers
 * (defined in ndt_kernels.hip; every launcher enqueues on `stream` and returns a
 * hipError_t value) */

int launch_bbox(const float4* pts, int64_t n, uint32_t* minmax_bits /* 6 */, hipStream_t stream);
int launch_keys_ndt(const float4* pts, int64_t n, double resolution, const int64_t* origin_cell,
                    const int* bits, uint64_t* keys, uint32_t* vals, hipStream_t stream);
int launch_keys_compat(const float4* pts, int64_t n, const float* minv, const int* dims, float resolution,
                       uint64_t* keys, uint32_t* vals, hipStream_t stream);
/* stable LSD radix sort of (key, value) pairs on the low `nbits` bits; the result
 * ends in (keys, vals); tmp buffers hold n entries each; scratch from
 * radix_sort_scratch_bytes */
size_t radix_sort_scratch_bytes(int64_t n);
int radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_tmp, uint32_t* vals_tmp, int64_t n,
                     int nbits, void* scratch, hipStream_t stream);
size_t scan_scratch_bytes(int64_t n);
/* exclusive prefix sum of n uint32 values; total (sum) written to *total_dev */
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev, void* scratch,
                       hipStream_t stream);
/* flags[i] = 1 where a run of equal (key >> shift) starts; keys >= limit are excluded */
int launch_run_heads(const uint64_t* keys, int64_t n, int shift, uint64_t limit, uint32_t* flags,
                     hipStream_t stream);
/* starts[rank] = i for each head i (rank = exclusive scan of flags) */
int launch_run_starts(const uint32_t* flags, const uint32_t* ranks, int64_t n, uint32_t* starts,
                      hipStream_t stream);
int launch_voxel_stats_ndt(const float4* pts, const uint32_t* sorted_idx, const uint64_t* sorted_keys,
                           const uint32_t* seg_start, int64_t nseg, int64_t nvalid_pts, int min_points,
                           double min_eig_ratio, const int* bits, double* mean, double* icov, double* cov,
                           int32_t* count, uint32_t* valid, uint64_t* vkey, hipStream_t stream);
int launch_compact_voxels(const uint32_t* valid, const uint32_t* rank, int64_t nvox, const double* mean,
                          const double* icov, const uint64_t* vkey, double* mean_out, double* icov_out,
                          float4* icov_f, uint64_t* vkey_out, hipStream_t stream);
int launch_dilate_emit(const uint64_t* vkey, int64_t nvalid, const int* bits, uint64_t* keys, uint32_t* vals,
                       hipStream_t stream);
int launch_hash_insert(const uint64_t* sorted_keys, const uint32_t* starts, int64_t nq, int64_t nentries,
                       int shift, HashSlot* slots, int log2cap, hipStream_t stream);
int launch_hash_insert_ids(const uint64_t* vkey, int64_t nvox, HashSlot* slots, int log2cap, hipStream_t stream);
int launch_cand_fill(const uint64_t* sorted_keys, const uint32_t* sorted_vals, int64_t nentries, const double* mean,
                     const double* icov, const int* bits, const int64_t* origin_cell, double res, float4* head,
                     float4* s32, hipStream_t stream);
int launch_voxel_s64(const double* mean, const double* icov, int64_t nvalid, double2* vs64, hipStream_t stream);
int launch_hash_clear(HashSlot* slots, int64_t cap, hipStream_t stream);
int launch_voxel_stats_compat(const float4* pts, const uint32_t* sorted_idx, const uint64_t* sorted_keys,
                              const uint32_t* seg_start, int64_t nseg, int64_t npts, double ncells,
                              double* mean, double* icov, int32_t* count, uint64_t* vkey, hipStream_t stream);

/* ndt alignment: init states from guesses, then the cooperative persistent kernel */
int launch_dense_fill(const uint64_t* sorted_keys, const uint32_t* starts, int64_t nq, int64_t nentries,
                      const int* bits, const int32_t* dims, uint2* dense, hipStream_t stream);
int launch_gather_points(const float4* pts, const uint32_t* idx, int64_t n, float4* out, hipStream_t stream);
int launch_fitness(FitGridDev fg, const float4* pts, const int64_t* offsets, const int32_t* counts, int32_t S,
                   const double* poses, double max_d2, int shells, int nblk, double* partials, hipStream_t stream);
int launch_init_queue(const ScanDesc* scans, int32_t S, uint64_t total_chunks, WorkQueue* q, uint64_t* items,
                      uint32_t* done_cnt, hipStream_t stream);
int launch_align_wq(NdtGridDev g, const float4* pts, const ScanDesc* scans, int32_t nwg, int block_size,
                    AlignParams prm, ScanState* states, WorkQueue* q, uint64_t* items, uint64_t cap, uint32_t* done_cnt,
                    double* partials, ScanOut* out, double* prof, int cooperative, hipStream_t stream);
/* writes n zero words (32-bit) from dst: a kernel, so that graphs can hold it */
int launch_zero_u32(void* dst, int64_t n, hipStream_t stream);
int align_wq_max_blocks(int block_size, int point_fp64, int sum_mode, int* blocks_per_cu);
/* alt (optional, S x 16): refined starts; then every scan opens with a probe pass at
 * its given guess, runs from the refined guess and falls back to the given guess */
int launch_init_states(const double* guesses, const double* alt, int32_t S, AlignParams prm, ScanState* states,
                       RoundSync* rs, hipStream_t stream);
int align_kernel_max_blocks(int block_size, int point_fp64, int sum_mode, int* blocks_per_cu);
constexpr int MAX_BATCH_SCANS = 2048;
/* Scan points per partial sum of the align kernels when ndt_params.points_per_wg
 * is 0.  A constant, so that a scan's result never depends on the batch it is
 * aligned in; 2048 maximises batched throughput (256 minimises the latency of a
 * single large scan, at a different rounding of its sums). */
constexpr int32_t ALIGN_CHUNK_POINTS = 2048;
NDT_HD int32_t align_chunk_points(int64_t /*n*/) { return ALIGN_CHUNK_POINTS; }
/* prof receives 5 doubles: phase microseconds (active list, compute, barriers,
 * control) of workgroup 0 and an abort flag */
int launch_align(NdtGridDev grid, const float4* pts, const ScanDesc* scans, int32_t S, int32_t nwg,
                 int block_size, AlignParams prm, ScanState* states, RoundSync* rs, double* partials, ScanOut* out,
                 double* prof, hipStream_t stream);

/* single derivative passes (tests and the compat path) */
/* partials: NRED (ndt) / NRED_COMPAT (compat) doubles per block, the last one the pair count */
int launch_deriv_ndt(NdtGridDev grid, const float4* pts, int64_t n, const double* pose_params, AlignParams prm,
                     double* partials, int nblocks, hipStream_t stream);
int launch_deriv_compat(CompatGridDev grid, const float4* scan, int64_t n, const float* M16, const double* jh,
                        double d1, double d2, double* partials, int nblocks, hipStream_t stream);
int launch_reduce_partials(const double* partials, int nblocks, int width, double* out, hipStream_t stream);

/* allocation-free alignment (ndt_align_scans_noalloc): a file-scope device array of
 * NOALLOC_SCRATCH_BYTES per device holds the whole workspace of a call */
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
/* device address of this device's workspace array (call with `device` current) */
int noalloc_scratch(int device, void** ptr);
/* chunk descriptors of S scans from their point counts (device memory) */
int launch_make_descs(const int64_t* counts, int32_t S, int32_t chunk, int32_t fine_pass, ScanDesc* desc,
                      hipStream_t stream);
/* ndt_result records from the outputs of a launch; prof[5] receives the pair total */
int launch_noalloc_results(const ScanOut* out, const ScanDesc* desc, int32_t S, double* prof, ndt_result* res,
                           hipStream_t stream);
static_assert(sizeof(ndt_result) == 232, "ndt_result layout shared by host and device");
static_assert(sizeof(ScanOut) % 4 == 0, "ScanOut is cleared as 32-bit words");

}  // namespace ndt

#endif
