/*
 * CPU oracle for the NDT scan-matching engine.  TEST INFRASTRUCTURE ONLY:
 * see ndt_oracle.h for who may use it and what pins each part.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  Contraction of
 * a*b+c into FMA must stay off: the reference is compiled without FMA and its
 * output is sensitive to the last bit of its sums.
 *
 * Citations are to the reference tree (src/CPU/ndt_mapping/kernel.cpp unless
 * another file is named).
 */
#include "ndt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.1415926535897932384626433832795 /* datatypes.h:74 */
#define ORC_SINGULAR_PIVOT 0.001                /* MAX_TRANSLATION_EPS, kernel.cpp:18,317-320 */

/* ======================================================================
 * shared helpers
 * ====================================================================== */

static inline double dot3(const double a[3], const double b[3]) {
    /* dot_product, kernel.cpp:409-412 */
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static inline double dot6(const double a[6], const double b[6]) {
    /* dot_product6, kernel.cpp:417-420 */
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4] + a[5] * b[5];
}

/* std::min / std::max as libstdc++ defines them (NaN handling included) */
static inline double std_min(double a, double b) { return (b < a) ? b : a; }
static inline double std_max(double a, double b) { return (a < b) ? b : a; }

void orc_gauss_consts(double resolution, double outlier_ratio, double* d1, double* d2) {
    /* kernel.cpp:1181-1185 (Magnusson 2009 eq. 6.8) */
    double c1 = 10 * (1 - outlier_ratio);
    double c2 = outlier_ratio / pow(resolution, 3);
    double d3 = -log(c2);
    *d1 = -log(c1 + c2) - d3;
    *d2 = -2 * log((-log(c1 * exp(-0.5) + c2) - d3) / *d1);
}

void orc_angle_tables(const double p[6], double j[8][3], double h[15][3]) {
    /* computeAngleDerivatives, kernel.cpp:707-825 (eq. 6.19 / 6.21) */
    double cx, cy, cz, sx, sy, sz;
    if (fabs(p[3]) < 10e-5) { cx = 1.0; sx = 0.0; } else { cx = cos(p[3]); sx = sin(p[3]); }
    if (fabs(p[4]) < 10e-5) { cy = 1.0; sy = 0.0; } else { cy = cos(p[4]); sy = sin(p[4]); }
    if (fabs(p[5]) < 10e-5) { cz = 1.0; sz = 0.0; } else { cz = cos(p[5]); sz = sin(p[5]); }
    /* a..h */
    j[0][0] = (-sx * sz + cx * sy * cz); j[0][1] = (-sx * cz - cx * sy * sz); j[0][2] = (-cx * cy);
    j[1][0] = (cx * sz + sx * sy * cz);  j[1][1] = (cx * cz - sx * sy * sz);  j[1][2] = (-sx * cy);
    j[2][0] = (-sy * cz);                j[2][1] = sy * sz;                   j[2][2] = cy;
    j[3][0] = sx * cy * cz;              j[3][1] = (-sx * cy * sz);           j[3][2] = sx * sy;
    j[4][0] = (-cx * cy * cz);           j[4][1] = cx * cy * sz;              j[4][2] = (-cx * sy);
    j[5][0] = (-cy * sz);                j[5][1] = (-cy * cz);                j[5][2] = 0;
    j[6][0] = (cx * cz - sx * sy * sz);  j[6][1] = (-cx * sz - sx * sy * cz); j[6][2] = 0;
    j[7][0] = (sx * cz + cx * sy * sz);  j[7][1] = (cx * sy * cz - sx * sz);  j[7][2] = 0;
    /* a2 a3 b2 b3 c2 c3 d1 d2 d3 e1 e2 e3 f1 f2 f3 */
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

void orc_ndt_angle_tables(const double p[6], double j[8][3], double h[15][3]) {
    /* The reference's tables, except h_ang_d1_[2] (kernel.cpp:786): the second
     * derivative of Rx Ry Rz x with respect to pitch has -sin(pitch) in its third
     * component, the reference (like the PCL code it derives from) stores +sin(pitch).
     * The textbook mode uses the exact derivative (pinned by finite differences). */
    orc_angle_tables(p, j, h);
    h[6][2] = -h[6][2];
}

void orc_point_derivs(const double x[3], const double j[8][3], const double h[15][3],
                      double pg[3][6], double ph[18][6]) {
    /* point_gradient_ initialisation kernel.cpp:1191-1195, computePointDerivatives :495-560 */
    memset(pg, 0, sizeof(double) * 18);
    memset(ph, 0, sizeof(double) * 108);
    pg[0][0] = 1.0; pg[1][1] = 1.0; pg[2][2] = 1.0;
    pg[1][3] = dot3(x, j[0]);
    pg[2][3] = dot3(x, j[1]);
    pg[0][4] = dot3(x, j[2]);
    pg[1][4] = dot3(x, j[3]);
    pg[2][4] = dot3(x, j[4]);
    pg[0][5] = dot3(x, j[5]);
    pg[1][5] = dot3(x, j[6]);
    pg[2][5] = dot3(x, j[7]);
    double a[3], b[3], c[3], d[3], e[3], f[3];
    a[0] = 0; a[1] = dot3(x, h[0]); a[2] = dot3(x, h[1]);
    b[0] = 0; b[1] = dot3(x, h[2]); b[2] = dot3(x, h[3]);
    c[0] = 0; c[1] = dot3(x, h[4]); c[2] = dot3(x, h[5]);
    d[0] = dot3(x, h[6]); d[1] = dot3(x, h[7]); d[2] = dot3(x, h[8]);
    e[0] = dot3(x, h[9]); e[1] = dot3(x, h[10]); e[2] = dot3(x, h[11]);
    f[0] = dot3(x, h[12]); f[1] = dot3(x, h[13]); f[2] = dot3(x, h[14]);
    for (int r = 0; r < 3; r++) {
        ph[9 + r][3] = a[r];  ph[12 + r][3] = b[r]; ph[15 + r][3] = c[r];
        ph[9 + r][4] = b[r];  ph[12 + r][4] = d[r]; ph[15 + r][4] = e[r];
        ph[9 + r][5] = c[r];  ph[12 + r][5] = e[r]; ph[15 + r][5] = f[r];
    }
}

void orc_solve6(double result[6], const double Ain[36], const double bin[6]) {
    /* solve, kernel.cpp:293-343: Gaussian elimination, partial pivoting by column
     * maximum, zero pivot replaced by 0.001, back substitution. */
    double A[6][6], b[6];
    memcpy(A, Ain, sizeof(A));
    memcpy(b, bin, sizeof(b));
    double pivot;
    for (int j = 0; j < 6; j++) {
        double max = fabs(A[j][j]);
        int mi = j;
        for (int i = j + 1; i < 6; i++)
            if (fabs(A[i][j]) > max) { mi = i; max = fabs(A[i][j]); }
        if (mi != j)
            for (int i = 0; i < 6; i++) { double t = A[mi][i]; A[mi][i] = A[j][i]; A[j][i] = t; }
        if (max == 0.0) A[j][j] = ORC_SINGULAR_PIVOT;
        for (int i = j + 1; i < 6; i++) {
            pivot = A[i][j] / A[j][j];
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

void orc_solve6_pivot(double result[6], const double Ain[36], const double bin[6]) {
    /* Gaussian elimination with partial pivoting that swaps the right-hand side
     * together with the rows of A (the textbook mode's solve).  The reference's
     * solve (kernel.cpp:306-313) swaps the rows of A only, which leaves b in the
     * wrong order whenever a pivot row moves. */
    double A[6][6], b[6];
    memcpy(A, Ain, sizeof(A));
    memcpy(b, bin, sizeof(b));
    for (int j = 0; j < 6; j++) {
        double max = fabs(A[j][j]);
        int mi = j;
        for (int i = j + 1; i < 6; i++)
            if (fabs(A[i][j]) > max) { mi = i; max = fabs(A[i][j]); }
        if (mi != j) {
            for (int i = 0; i < 6; i++) { double t = A[mi][i]; A[mi][i] = A[j][i]; A[j][i] = t; }
            double t = b[mi]; b[mi] = b[j]; b[j] = t;
        }
        if (max == 0.0) A[j][j] = ORC_SINGULAR_PIVOT;
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

static inline double psi_mt(double a, double f_a, double f_0, double g_0, double mu) {
    return (f_a - f_0 - mu * g_0 * a); /* kernel.cpp:433-436 */
}
static inline double dpsi_mt(double g_a, double g_0, double mu) {
    return (g_a - mu * g_0); /* kernel.cpp:425-428 */
}

int orc_mt_update_interval(double* a_l, double* f_l, double* g_l, double* a_u, double* f_u,
                           double* g_u, double a_t, double f_t, double g_t) {
    /* updateIntervalMT, kernel.cpp:827-864 (More & Thuente 1994, updating algorithm) */
    if (f_t > *f_l) {
        *a_u = a_t; *f_u = f_t; *g_u = g_t;
        return 0;
    } else if (g_t * (*a_l - a_t) > 0) {
        *a_l = a_t; *f_l = f_t; *g_l = g_t;
        return 0;
    } else if (g_t * (*a_l - a_t) < 0) {
        *a_u = *a_l; *f_u = *f_l; *g_u = *g_l;
        *a_l = a_t; *f_l = f_t; *g_l = g_t;
        return 0;
    }
    return 1;
}

double orc_mt_trial_value(double a_l, double f_l, double g_l, double a_u, double f_u, double g_u,
                          double a_t, double f_t, double g_t) {
    /* trialValueSelectionMT, kernel.cpp:867-946 (Sun & Yuan 2006 eqs. 2.4.2, 2.4.5,
     * 2.4.52, 2.4.56) */
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
        double a_t_next;
        if (fabs(a_c - a_t) < fabs(a_s - a_t)) a_t_next = a_c;
        else a_t_next = a_s;
        if (a_t > a_l) return std_min(a_t + 0.66 * (a_u - a_t), a_t_next);
        return std_max(a_t + 0.66 * (a_u - a_t), a_t_next);
    } else {
        double z = 3 * (f_t - f_u) / (a_t - a_u) - g_t - g_u;
        double w = sqrt(z * z - g_t * g_u);
        return (a_u + (a_t - a_u) * (w - g_u - z) / (g_t - g_u + 2 * w));
    }
}

/* ======================================================================
 * compat: reference semantics
 * ====================================================================== */

static inline int compat_linearize(const orc_compat_grid* g, float resolution, float x, float y, float z) {
    /* linearizeCoord / linearizeAddr, kernel.cpp:231-244 (truncation toward zero) */
    int idx_x = (int)((x - g->minv[0]) / resolution);
    int idx_y = (int)((y - g->minv[1]) / resolution);
    int idx_z = (int)((z - g->minv[2]) / resolution);
    return (idx_x + g->dims[0] * (idx_y + g->dims[1] * idx_z));
}

static void invert33(double m[3][3]) {
    /* invertMatrix, kernel.cpp:1277-1300: adjugate / determinant, no singularity guard */
    double t[3][3];
    double det = m[0][0] * (m[2][2] * m[1][1] - m[2][1] * m[1][2]) -
                 m[1][0] * (m[2][2] * m[0][1] - m[2][1] * m[0][2]) +
                 m[2][0] * (m[1][2] * m[0][1] - m[1][1] * m[0][2]);
    double invDet = 1.0 / det;
    t[0][0] = m[2][2] * m[1][1] - m[2][1] * m[1][2];
    t[0][1] = -(m[2][2] * m[0][1] - m[2][1] * m[0][2]);
    t[0][2] = m[1][2] * m[0][1] - m[1][1] * m[0][2];
    t[1][0] = -(m[2][2] * m[0][1] - m[2][0] * m[1][2]);
    t[1][1] = m[2][2] * m[0][0] - m[2][1] * m[0][2];
    t[1][2] = -(m[1][2] * m[0][0] - m[1][0] * m[0][2]);
    t[2][0] = m[2][1] * m[1][0] - m[2][0] * m[1][1];
    t[2][1] = -(m[2][1] * m[0][0] - m[2][0] * m[0][1]);
    t[2][2] = m[1][1] * m[0][0] - m[1][0] * m[0][1];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) m[r][c] = t[r][c] * invDet;
}

static int compat_grid_build_impl(const float* map4, int64_t m, float resolution, orc_compat_grid* g,
                                  int use_omp) {
    /* initCompute, kernel.cpp:1302-1373 */
    memset(g, 0, sizeof(*g));
    if (!map4 || m <= 0) return 1;
    float mn[3], mx[3];
    for (int e = 0; e < 3; e++) mn[e] = mx[e] = map4[e];
    if (use_omp) {
        /* src/OpenMP/ndt_mapping/kernel.cpp:1305-1334 (min/max are exact, any order) */
        float mn0 = mn[0], mn1 = mn[1], mn2 = mn[2], mx0 = mx[0], mx1 = mx[1], mx2 = mx[2];
#pragma omp parallel for reduction(min : mn0, mn1, mn2) reduction(max : mx0, mx1, mx2)
        for (int64_t i = 1; i < m; i++) {
            const float* p = map4 + 4 * i;
            mn0 = (p[0] < mn0) ? p[0] : mn0; mn1 = (p[1] < mn1) ? p[1] : mn1; mn2 = (p[2] < mn2) ? p[2] : mn2;
            mx0 = (p[0] > mx0) ? p[0] : mx0; mx1 = (p[1] > mx1) ? p[1] : mx1; mx2 = (p[2] > mx2) ? p[2] : mx2;
        }
        mn[0] = mn0; mn[1] = mn1; mn[2] = mn2; mx[0] = mx0; mx[1] = mx1; mx[2] = mx2;
    } else {
        for (int64_t i = 1; i < m; i++)
            for (int e = 0; e < 3; e++) {
                if (map4[4 * i + e] > mx[e]) mx[e] = map4[4 * i + e];
                if (map4[4 * i + e] < mn[e]) mn[e] = map4[4 * i + e];
            }
    }
    g->minv[3] = map4[3];
    g->maxv[3] = map4[3];
    for (int i = 0; i < 3; i++) {
        g->minv[i] = mn[i] - 0.01f;
        g->maxv[i] = mx[i] + 0.01f;
        g->dims[i] = (int)((g->maxv[i] - g->minv[i]) / resolution + 1);
    }
    int ncells = g->dims[0] * g->dims[1] * g->dims[2];
    if (ncells <= 0) return 1;
    g->ncells = ncells;
    g->cells = (orc_voxel*)malloc(sizeof(orc_voxel) * (size_t)ncells);
    if (!g->cells) return 1;
#pragma omp parallel for if (use_omp)
    for (int i = 0; i < ncells; i++) {
        orc_voxel* v = &g->cells[i];
        v->n = 0;
        v->mean[0] = v->mean[1] = v->mean[2] = 0;
        memset(v->icov, 0, sizeof(v->icov));
        v->icov[2][0] = 1.0; v->icov[1][1] = 1.0; v->icov[0][2] = 1.0; /* :1335-1337 */
    }
    for (int64_t i = 0; i < m; i++) {
        const float* p = map4 + 4 * i;
        int vi = compat_linearize(g, resolution, p[0], p[1], p[2]);
        orc_voxel* v = &g->cells[vi];
        v->mean[0] += p[0];
        v->mean[1] += p[1];
        v->mean[2] += p[2];
        v->n++;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) v->icov[r][c] += (double)(p[r] * p[c]); /* float product */
    }
    const double C = (double)(size_t)ncells;
#pragma omp parallel for if (use_omp)
    for (int i = 0; i < ncells; i++) {
        orc_voxel* v = &g->cells[i];
        double ps[3] = {v->mean[0], v->mean[1], v->mean[2]};
        v->mean[0] /= v->n;
        v->mean[1] /= v->n;
        v->mean[2] /= v->n;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                /* :1366-1369: normalised by the total cell count, not the point count */
                v->icov[r][c] = (v->icov[r][c] - 2 * (ps[r] * v->mean[c])) / C + v->mean[r] * v->mean[c];
                v->icov[r][c] *= (C - 1.0) / v->n;
            }
        invert33(v->icov);
    }
    return 0;
}

int orc_compat_grid_build(const float* map4, int64_t m, float resolution, orc_compat_grid* g) {
    return compat_grid_build_impl(map4, m, resolution, g, 0);
}

void orc_compat_grid_free(orc_compat_grid* g) {
    if (g) {
        free(g->cells);
        g->cells = NULL;
    }
}

int orc_compat_radius_search(const orc_compat_grid* g, float resolution, const float pt[4],
                             double radius, int64_t out_cells[27], float out_dist[27]) {
    /* voxelRadiusSearch, kernel.cpp:246-286: float-stepped 3x3x3 probe, z outermost,
     * bounds skip, keep the cell if the float distance to its mean is < radius
     * (NaN means of empty cells never pass). */
    int result = 0;
    float radiusFinal = (float)(radius + 0.001f);
    for (float z = (float)(pt[2] - radius); z <= pt[2] + radiusFinal; z += resolution)
        for (float y = (float)(pt[1] - radius); y <= pt[1] + radiusFinal; y += resolution)
            for (float x = (float)(pt[0] - radius); x <= pt[0] + radiusFinal; x += resolution) {
                if ((x < g->minv[0]) || (x > g->maxv[0]) || (y < g->minv[1]) || (y > g->maxv[1]) ||
                    (z < g->minv[2]) || (z > g->maxv[2]))
                    continue;
                int idx = compat_linearize(g, resolution, x, y, z);
                const double* c = g->cells[idx].mean;
                float dx = (float)(c[0] - pt[0]);
                float dy = (float)(c[1] - pt[1]);
                float dz = (float)(c[2] - pt[2]);
                float dist = (float)sqrt((double)(dx * dx + dy * dy + dz * dz));
                if (dist < radius) {
                    if (result < 27) {
                        out_cells[result] = idx;
                        out_dist[result] = dist;
                    }
                    result++;
                }
            }
    return result;
}

double orc_compat_update(double grad[6], double H[36], const double xt[3], const double ci[3][3],
                         const double pg[3][6], const double ph[18][6], double d1, double d2,
                         int compute_hessian) {
    /* updateDerivatives, kernel.cpp:438-493, including the non-textbook Hessian term
     * -d2 (x'C J_i)(x'C (J_j + H_ij)) + J_j C J_i (:476-488). */
    double xCx = ci[0][0] * xt[0] * xt[0] + ci[1][1] * xt[1] * xt[1] + ci[2][2] * xt[2] * xt[2] +
                 (ci[0][1] + ci[1][0]) * xt[0] * xt[1] + (ci[0][2] + ci[2][0]) * xt[0] * xt[2] +
                 (ci[1][2] + ci[2][1]) * xt[1] * xt[2];
    double e = exp(-d2 * (xCx) / 2);
    double score_inc = -d1 * e;
    e = d2 * e;
    if (e > 1 || e < 0 || e != e) return 0;
    e *= d1;
    double cov[3];
    for (int i = 0; i < 6; i++) {
        for (int r = 0; r < 3; r++) {
            cov[r] = 0;
            for (int c = 0; c < 3; c++) cov[r] += ci[r][c] * pg[c][i];
        }
        grad[i] += dot3(xt, cov) * e;
        if (compute_hessian) {
            for (int j = 0; j < 6; j++) {
                double colVec[3] = {pg[0][j], pg[1][j], pg[2][j]};
                double cvh[3] = {colVec[0] + ph[3 * i][j], colVec[1] + ph[3 * i + 1][j],
                                 colVec[2] + ph[3 * i + 2][j]};
                double mp[3];
                for (int r = 0; r < 3; r++) {
                    mp[r] = 0;
                    for (int c = 0; c < 3; c++) mp[r] += ci[r][c] * cvh[c];
                }
                H[i * 6 + j] += e * (-d2 * dot3(xt, cov) * dot3(xt, mp) + dot3(colVec, cov));
            }
        }
    }
    return score_inc;
}

double orc_compat_derivatives(const orc_compat_grid* g, float resolution, const float* scan4,
                              const float* trans4, int64_t n, const double p[6], double d1, double d2,
                              double grad[6], double H[36], int64_t* npairs) {
    /* computeDerivatives, kernel.cpp:651-705 */
    double jang[8][3], hang[15][3], pg[3][6], ph[18][6];
    memset(grad, 0, sizeof(double) * 6);
    memset(H, 0, sizeof(double) * 36);
    double score = 0.0;
    int64_t pairs = 0;
    orc_angle_tables(p, jang, hang);
    int64_t cells[27];
    float dist[27];
    for (int64_t idx = 0; idx < n; idx++) {
        const float* xtp = trans4 + 4 * idx;
        int k = orc_compat_radius_search(g, resolution, xtp, (double)resolution, cells, dist);
        if (k > 27) k = 27;
        for (int q = 0; q < k; q++) {
            const orc_voxel* v = &g->cells[cells[q]];
            double x[3] = {scan4[4 * idx], scan4[4 * idx + 1], scan4[4 * idx + 2]};
            double xt[3] = {xtp[0], xtp[1], xtp[2]};
            xt[0] -= v->mean[0];
            xt[1] -= v->mean[1];
            xt[2] -= v->mean[2];
            orc_point_derivs(x, jang, hang, pg, ph);
            score += orc_compat_update(grad, H, xt, v->icov, pg, ph, d1, d2, 1);
            pairs++;
        }
    }
    if (npairs) *npairs = pairs;
    return score;
}

void orc_compat_transform(const float* in4, float* out4, int64_t n, const float M[16]) {
    /* transformPointCloud, kernel.cpp:385-404 (float, w left undefined -> 0 here) */
    for (int64_t it = 0; it < n; it++) {
        float t[3];
        const float* p = in4 + 4 * it;
        for (int row = 0; row < 3; row++)
            t[row] = M[row * 4 + 0] * p[0] + M[row * 4 + 1] * p[1] + M[row * 4 + 2] * p[2] + M[row * 4 + 3];
        out4[4 * it + 0] = t[0];
        out4[4 * it + 1] = t[1];
        out4[4 * it + 2] = t[2];
        out4[4 * it + 3] = 0.0f;
    }
}

void orc_compat_build_matrix(float M[16], const double t[6]) {
    /* buildTransformationMatrix, kernel.cpp:949-1013: per-axis float quaternions with
     * the product components relabelled as written (:970-980); rotation only, the
     * translation column is left as it was (:1000-1004). */
    const float q_ha = (float)(0.5f * t[3]);
    const float q_w = (float)cos(q_ha);
    const float q_x = (float)sin(q_ha);
    const float q_y = 0.0f;
    const float q_z = 0.0f;
    const float q_ha2 = (float)(0.5f * t[4]);
    const float q_w2 = (float)cos(q_ha2);
    const float q_x2 = 0.0f;
    const float q_y2 = (float)sin(q_ha2);
    const float q_z2 = 0.0f;
    const float q_ha3 = (float)(0.5f * t[5]);
    const float q_w3 = (float)cos(q_ha3);
    const float q_x3 = 0.0f;
    const float q_y3 = 0.0f;
    const float q_z3 = (float)sin(q_ha3);
    const float r_x = q_w * q_w2 - q_x * q_x2 - q_y * q_y2 - q_z * q_z2;
    const float r_y = q_w * q_x2 + q_x * q_w2 + q_y * q_z2 - q_z * q_y2;
    const float r_z = q_w * q_y2 + q_y * q_w2 + q_z * q_x2 - q_x * q_z2;
    const float r_w = q_w * q_z2 + q_z * q_w2 + q_x * q_y2 - q_y * q_x2;
    const float r2_x = r_w * q_w3 - r_x * q_x3 - r_y * q_y3 - r_z * q_z3;
    const float r2_y = r_w * q_x3 + r_x * q_w3 + r_y * q_z3 - r_z * q_y3;
    const float r2_z = r_w * q_y3 + r_y * q_w3 + r_z * q_x3 - r_x * q_z3;
    const float r2_w = r_w * q_z3 + r_z * q_w3 + r_x * q_y3 - r_y * q_x3;
    const float tx = 2.0f * r2_x;
    const float ty = 2.0f * r2_y;
    const float tz = 2.0f * r2_z;
    const float twx = tx * r2_w;
    const float twy = ty * r2_w;
    const float twz = tz * r2_w;
    const float txx = tx * r2_x;
    const float txy = ty * r2_x;
    const float txz = tz * r2_x;
    const float tyy = ty * r2_y;
    const float tyz = tz * r2_y;
    const float tzz = tz * r2_z;
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

void orc_compat_euler(const float M[16], double result[3]) {
    /* eulerAngles, kernel.cpp:1146-1173 (Eigen eulerAngles(0,1,2) restated, result negated) */
    double res[3];
    res[0] = atan2((double)M[1 * 4 + 2], (double)M[2 * 4 + 2]);
    double n1 = M[0];
    double n2 = M[1];
    double c2 = sqrt(n1 * n1 + n2 * n2);
    if (res[0] > 0.0) {
        res[0] -= ORC_PI;
        res[1] = atan2((double)(-M[2]), -c2);
    } else {
        res[1] = atan2((double)(-M[2]), c2);
    }
    double s1 = sin(res[0]);
    double c1 = cos(res[0]);
    res[2] = atan2(s1 * M[8] - c1 * M[4], c1 * M[5] - s1 * M[9]);
    result[0] = -res[0];
    result[1] = -res[1];
    result[2] = -res[2];
}

typedef struct {
    const orc_compat_grid* g;
    float resolution;
    const float* scan4;
    float* trans4;
    int64_t n;
    double d1, d2;
    float final_m[16];
    int passes;
    int use_omp;
} compat_ctx;

static void compat_transform_ctx(compat_ctx* c, const float* in4, const float M[16]) {
    if (c->use_omp) {
#pragma omp parallel for
        for (int64_t it = 0; it < c->n; it++) orc_compat_transform(in4 + 4 * it, c->trans4 + 4 * it, 1, M);
    } else {
        orc_compat_transform(in4, c->trans4, c->n, M);
    }
}

static double compat_step_length(compat_ctx* c, const double x[6], double step_dir[6], double step_init,
                                 double step_max, double step_min, double* score, double grad[6],
                                 double H[36]) {
    /* computeStepLengthMT, kernel.cpp:1015-1144 with the reference's initialisation
     * interval_converged = (step_max - step_min) > 0 (:1053), so the loop below never
     * runs for the reference parameters; it is kept as written. */
    double phi_0 = -*score;
    double d_phi_0 = -(dot6(grad, step_dir));
    double x_t[6] = {0, 0, 0, 0, 0, 0};
    if (d_phi_0 >= 0) {
        if (d_phi_0 == 0) return 0;
        d_phi_0 *= -1;
        for (int i = 0; i < 6; i++) step_dir[i] *= -1;
    }
    int max_step_iterations = 10, step_iterations = 0;
    double mu = 1.e-4, nu = 0.9;
    double a_l = 0, a_u = 0;
    double f_l = psi_mt(a_l, phi_0, phi_0, d_phi_0, mu);
    double g_l = dpsi_mt(d_phi_0, d_phi_0, mu);
    double f_u = psi_mt(a_u, phi_0, phi_0, d_phi_0, mu);
    double g_u = dpsi_mt(d_phi_0, d_phi_0, mu);
    int interval_converged = (step_max - step_min) > 0, open_interval = 1;
    double a_t = step_init;
    a_t = std_min(a_t, step_max);
    a_t = std_max(a_t, step_min);
    for (int i = 0; i < 6; i++) x_t[i] = x[i] + step_dir[i] * a_t;
    orc_compat_build_matrix(c->final_m, x_t);
    compat_transform_ctx(c, c->scan4, c->final_m);
    *score = orc_compat_derivatives(c->g, c->resolution, c->scan4, c->trans4, c->n, x_t, c->d1, c->d2,
                                    grad, H, NULL);
    c->passes++;
    double phi_t = -*score;
    double d_phi_t = -(dot6(grad, step_dir));
    double psi_t = psi_mt(a_t, phi_t, phi_0, d_phi_0, mu);
    double d_psi_t = dpsi_mt(d_phi_t, d_phi_0, mu);
    while (!interval_converged && step_iterations < max_step_iterations &&
           !(psi_t <= 0 && d_phi_t <= -nu * d_phi_0)) {
        if (open_interval)
            a_t = orc_mt_trial_value(a_l, f_l, g_l, a_u, f_u, g_u, a_t, psi_t, d_psi_t);
        else
            a_t = orc_mt_trial_value(a_l, f_l, g_l, a_u, f_u, g_u, a_t, phi_t, d_phi_t);
        a_t = std_min(a_t, step_max);
        a_t = std_max(a_t, step_min);
        for (int r = 0; r < 6; r++) x_t[r] = x[r] + step_dir[r] * a_t;
        orc_compat_build_matrix(c->final_m, x_t);
        compat_transform_ctx(c, c->scan4, c->final_m);
        *score = orc_compat_derivatives(c->g, c->resolution, c->scan4, c->trans4, c->n, x_t, c->d1,
                                        c->d2, grad, H, NULL);
        c->passes++;
        phi_t = -*score;
        d_phi_t = -(dot6(grad, step_dir));
        psi_t = psi_mt(a_t, phi_t, phi_0, d_phi_0, mu);
        d_psi_t = dpsi_mt(d_phi_t, d_phi_0, mu);
        if (open_interval && (psi_t <= 0 && d_psi_t >= 0)) {
            open_interval = 0;
            f_l = f_l + phi_0 - mu * d_phi_0 * a_l;
            g_l = g_l + mu * d_phi_0;
            f_u = f_u + phi_0 - mu * d_phi_0 * a_u;
            g_u = g_u + mu * d_phi_0;
        }
        if (open_interval)
            interval_converged = orc_mt_update_interval(&a_l, &f_l, &g_l, &a_u, &f_u, &g_u, a_t, psi_t, d_psi_t);
        else
            interval_converged = orc_mt_update_interval(&a_l, &f_l, &g_l, &a_u, &f_u, &g_u, a_t, phi_t, d_phi_t);
        step_iterations++;
    }
    /* computeHessian at x_t when the loop ran (:1141-1142) recomputes the same Hessian
     * that the full derivative pass above already produced. */
    return a_t;
}

static int compat_align_impl(const float* map4, int64_t m, const float* scan4, int64_t n,
                             const float guess[16], orc_result* out, int use_omp) {
    /* partial_points_callback -> ndt_align -> computeTransformation,
     * kernel.cpp:1405-1414, :1375-1392, :1175-1272 with the reference parameters
     * (:37-57): outlier 0.55, resolution 1, step 0.1, epsilon 0.1, max_iterations 0. */
    memset(out, 0, sizeof(*out));
    if (!map4 || m <= 0 || !scan4 || n <= 0 || !guess) return 1;
    const float resolution = 1.0f;
    const double outlier_ratio = 0.55, step_size = 0.1, trans_eps = 0.1;
    const int max_iterations = 0;
    orc_compat_grid g;
    if (compat_grid_build_impl(map4, m, resolution, &g, use_omp)) return 1;
    compat_ctx c;
    memset(&c, 0, sizeof(c));
    c.g = &g;
    c.resolution = resolution;
    c.scan4 = scan4;
    c.n = n;
    c.use_omp = use_omp;
    c.trans4 = (float*)malloc(sizeof(float) * 4 * (size_t)n);
    if (!c.trans4) { orc_compat_grid_free(&g); return 1; }
    orc_gauss_consts(resolution, outlier_ratio, &c.d1, &c.d2);
    memcpy(c.final_m, guess, sizeof(float) * 16);
    compat_transform_ctx(&c, scan4, guess);
    double p[6], delta_p[6], grad[6], H[36];
    p[0] = c.final_m[3];
    p[1] = c.final_m[7];
    p[2] = c.final_m[11];
    double ea[3];
    orc_compat_euler(c.final_m, ea);
    p[3] = ea[0];
    p[4] = ea[1];
    p[5] = ea[2];
    double score = orc_compat_derivatives(&g, resolution, scan4, c.trans4, n, p, c.d1, c.d2, grad, H, NULL);
    c.passes = 1;
    int converged = 0, nr_iterations = 0;
    while (!converged) {
        double neg_grad[6] = {-grad[0], -grad[1], -grad[2], -grad[3], -grad[4], -grad[5]};
        orc_solve6(delta_p, H, neg_grad);
        double delta_p_norm = 1; /* :1231 overwrites the norm */
        for (int i = 0; i < 6; i++) delta_p[i] /= delta_p_norm;
        delta_p_norm = compat_step_length(&c, p, delta_p, delta_p_norm, step_size, trans_eps / 2, &score,
                                          grad, H);
        for (int i = 0; i < 6; i++) delta_p[i] *= delta_p_norm;
        for (int i = 0; i < 6; i++) p[i] = p[i] + delta_p[i];
        if (nr_iterations > max_iterations || (nr_iterations && (fabs(delta_p_norm) < trans_eps)))
            converged = 1;
        nr_iterations++;
    }
    memcpy(out->final_f32, c.final_m, sizeof(float) * 16);
    for (int i = 0; i < 16; i++) out->final_f64[i] = c.final_m[i];
    out->score = score;
    out->trans_probability = score / (double)n;
    out->iterations = nr_iterations;
    out->converged = converged;
    out->passes = c.passes;
    free(c.trans4);
    orc_compat_grid_free(&g);
    return 0;
}

int orc_compat_align(const float* map4, int64_t m, const float* scan4, int64_t n, const float guess[16],
                     orc_result* out) {
    return compat_align_impl(map4, m, scan4, n, guess, out, 0);
}

int orc_compat_align_omp(const float* map4, int64_t m, const float* scan4, int64_t n,
                         const float guess[16], orc_result* out) {
    return compat_align_impl(map4, m, scan4, n, guess, out, 1);
}

/* ======================================================================
 * ndt: textbook float64 NDT
 * ====================================================================== */

void orc_ndt_params_default(orc_ndt_params* p) {
    p->resolution = 1.0;
    p->step_size = 0.1;
    p->trans_eps = 0.01;
    p->outlier_ratio = 0.55;
    p->max_iterations = 30;
    p->min_points = 6;
    p->min_eig_ratio = 0.01;
    p->hessian_at_accept = 0;
    p->repeat_passes = 0;
}

void orc_sym3_eig(const double Ain[9], double evals[3], double evecs[9]) {
    /* cyclic Jacobi rotations on a symmetric 3x3 matrix (Golub & Van Loan, sec. 8.5) */
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
            if (d[idx[j]] < d[idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
    for (int i = 0; i < 3; i++) {
        evals[i] = d[idx[i]];
        for (int r = 0; r < 3; r++) evecs[r * 3 + i] = v[r][idx[i]];
    }
}

typedef struct {
    int64_t kx, ky, kz;
    int64_t idx;
} key_idx;

static int key_idx_cmp(const void* a, const void* b) {
    const key_idx* x = (const key_idx*)a;
    const key_idx* y = (const key_idx*)b;
    if (x->kz != y->kz) return x->kz < y->kz ? -1 : 1;
    if (x->ky != y->ky) return x->ky < y->ky ? -1 : 1;
    if (x->kx != y->kx) return x->kx < y->kx ? -1 : 1;
    if (x->idx != y->idx) return x->idx < y->idx ? -1 : 1;
    return 0;
}

static inline uint64_t ndt_hash3(int64_t x, int64_t y, int64_t z) {
    uint64_t h = (uint64_t)x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= (uint64_t)z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    return h ^ (h >> 29);
}

int64_t orc_ndt_grid_find(const orc_ndt_grid* g, int64_t ix, int64_t iy, int64_t iz) {
    uint64_t mask = (uint64_t)g->cap - 1;
    uint64_t s = ndt_hash3(ix, iy, iz) & mask;
    for (int64_t probe = 0; probe < g->cap; probe++) {
        int64_t v = g->slot[s];
        if (v < 0) return -1;
        const orc_ndt_voxel* vx = &g->vox[v];
        if (vx->key[0] == ix && vx->key[1] == iy && vx->key[2] == iz) return v;
        s = (s + 1) & mask;
    }
    return -1;
}

int orc_ndt_grid_build(const float* map4, int64_t m, const orc_ndt_params* prm, orc_ndt_grid* g) {
    /* Voxel statistics in the PCL VoxelGridCovariance convention, made numerically
     * robust: two-pass re-centred sums in point index order, sample covariance
     * (n - 1), at least min_points points, eigenvalues clamped to at least
     * min_eig_ratio * largest; inverse from the clamped eigen-decomposition. */
    memset(g, 0, sizeof(*g));
    if (!map4 || m <= 0 || !prm || prm->resolution <= 0) return 1;
    const double res = prm->resolution;
    g->resolution = res;
    key_idx* ki = (key_idx*)malloc(sizeof(key_idx) * (size_t)m);
    if (!ki) return 1;
    int64_t nk = 0;
    for (int64_t i = 0; i < m; i++) {
        const float* p = map4 + 4 * i;
        if (!isfinite(p[0]) || !isfinite(p[1]) || !isfinite(p[2])) continue;
        ki[nk].kx = (int64_t)floor((double)p[0] / res);
        ki[nk].ky = (int64_t)floor((double)p[1] / res);
        ki[nk].kz = (int64_t)floor((double)p[2] / res);
        ki[nk].idx = i;
        nk++;
    }
    qsort(ki, (size_t)nk, sizeof(key_idx), key_idx_cmp);
    int64_t nv = 0;
    for (int64_t i = 0; i < nk; i++)
        if (i == 0 || ki[i].kx != ki[i - 1].kx || ki[i].ky != ki[i - 1].ky || ki[i].kz != ki[i - 1].kz) nv++;
    g->vox = (orc_ndt_voxel*)calloc((size_t)(nv > 0 ? nv : 1), sizeof(orc_ndt_voxel));
    if (!g->vox) { free(ki); return 1; }
    g->nvox = nv;
    int64_t v = -1, start = 0;
    for (int64_t i = 0; i <= nk; i++) {
        int head = (i == nk) || i == 0 || ki[i].kx != ki[i - 1].kx || ki[i].ky != ki[i - 1].ky ||
                   ki[i].kz != ki[i - 1].kz;
        if (!head) continue;
        if (v >= 0) {
            orc_ndt_voxel* vx = &g->vox[v];
            int64_t cnt = i - start;
            vx->n = (int32_t)cnt;
            double s[3] = {0, 0, 0};
            for (int64_t k = start; k < i; k++) {
                const float* p = map4 + 4 * ki[k].idx;
                s[0] += p[0]; s[1] += p[1]; s[2] += p[2];
            }
            for (int c = 0; c < 3; c++) vx->mean[c] = s[c] / (double)cnt;
            double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int64_t k = start; k < i; k++) {
                const float* p = map4 + 4 * ki[k].idx;
                double d[3] = {p[0] - vx->mean[0], p[1] - vx->mean[1], p[2] - vx->mean[2]};
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) C[r * 3 + c] += d[r] * d[c];
            }
            vx->valid = 0;
            if (cnt >= prm->min_points && cnt >= 2) {
                for (int k = 0; k < 9; k++) C[k] /= (double)(cnt - 1);
                double ev[3], V[9];
                orc_sym3_eig(C, ev, V);
                if (ev[2] > 0 && isfinite(ev[2])) {
                    double lo = prm->min_eig_ratio * ev[2];
                    for (int k = 0; k < 3; k++)
                        if (ev[k] < lo) ev[k] = lo;
                    for (int r = 0; r < 3; r++)
                        for (int c = 0; c < 3; c++) {
                            double a = 0, b = 0;
                            for (int k = 0; k < 3; k++) {
                                a += V[r * 3 + k] * ev[k] * V[c * 3 + k];
                                b += V[r * 3 + k] * (1.0 / ev[k]) * V[c * 3 + k];
                            }
                            vx->cov[r * 3 + c] = a;
                            vx->icov[r * 3 + c] = b;
                        }
                    vx->valid = 1;
                }
            }
        }
        if (i < nk) {
            v++;
            start = i;
            g->vox[v].key[0] = ki[i].kx;
            g->vox[v].key[1] = ki[i].ky;
            g->vox[v].key[2] = ki[i].kz;
        }
    }
    free(ki);
    int64_t cap = 16;
    while (cap < 2 * nv) cap <<= 1;
    g->cap = cap;
    g->slot = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    if (!g->slot) { orc_ndt_grid_free(g); return 1; }
    for (int64_t s = 0; s < cap; s++) g->slot[s] = -1;
    for (int64_t i = 0; i < nv; i++) {
        uint64_t mask = (uint64_t)cap - 1;
        uint64_t s = ndt_hash3(g->vox[i].key[0], g->vox[i].key[1], g->vox[i].key[2]) & mask;
        while (g->slot[s] >= 0) s = (s + 1) & mask;
        g->slot[s] = i;
    }
    return 0;
}

void orc_ndt_grid_free(orc_ndt_grid* g) {
    if (!g) return;
    free(g->vox);
    free(g->slot);
    g->vox = NULL;
    g->slot = NULL;
}

void orc_ndt_pose_matrix(const double p[6], double M[16]) {
    /* T = [Rx(p3) Ry(p4) Rz(p5) | p0..2], the convention of the angle tables */
    double cx = cos(p[3]), sx = sin(p[3]);
    double cy = cos(p[4]), sy = sin(p[4]);
    double cz = cos(p[5]), sz = sin(p[5]);
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

void orc_ndt_euler(const double M[16], double out[3]) {
    /* inverse of orc_ndt_pose_matrix for |pitch| < pi/2 */
    double s = M[2];
    if (s > 1.0) s = 1.0;
    if (s < -1.0) s = -1.0;
    out[0] = atan2(-M[6], M[10]);
    out[1] = asin(s);
    out[2] = atan2(-M[1], M[0]);
}

typedef struct {
    const orc_ndt_grid* g;
    const orc_ndt_params* prm;
    const float* scan4;
    int64_t n;
    double d1, d2;
    int passes;
} ndt_ctx;

static double ndt_pass(const orc_ndt_grid* g, const orc_ndt_params* prm, double d1, double d2,
                       const float* scan4, int64_t n, const double p[6], double grad[6], double H[36],
                       int compute_derivs, int compute_hessian, int64_t* npairs) {
    double jang[8][3], hang[15][3], pg[3][6], ph[18][6], M[16];
    double Hu[36];
    memset(Hu, 0, sizeof(Hu));
    if (grad) memset(grad, 0, sizeof(double) * 6);
    orc_ndt_angle_tables(p, jang, hang);
    orc_ndt_pose_matrix(p, M);
    const double res = g->resolution;
    double score = 0.0;
    int64_t pairs = 0;
    for (int64_t i = 0; i < n; i++) {
        double x[3] = {scan4[4 * i], scan4[4 * i + 1], scan4[4 * i + 2]};
        double xt[3];
        for (int r = 0; r < 3; r++) xt[r] = M[r * 4 + 0] * x[0] + M[r * 4 + 1] * x[1] + M[r * 4 + 2] * x[2] + M[r * 4 + 3];
        int64_t cx = (int64_t)floor(xt[0] / res), cy = (int64_t)floor(xt[1] / res), cz = (int64_t)floor(xt[2] / res);
        int have_pd = 0;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int64_t v = orc_ndt_grid_find(g, cx + dx, cy + dy, cz + dz);
                    if (v < 0 || !g->vox[v].valid) continue;
                    const orc_ndt_voxel* vx = &g->vox[v];
                    double d[3] = {xt[0] - vx->mean[0], xt[1] - vx->mean[1], xt[2] - vx->mean[2]};
                    if (!(sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) < res)) continue;
                    pairs++;
                    const double* C = vx->icov;
                    double q[3];
                    for (int r = 0; r < 3; r++) q[r] = C[r * 3 + 0] * d[0] + C[r * 3 + 1] * d[1] + C[r * 3 + 2] * d[2];
                    double xCx = dot3(d, q);
                    double e = exp(-d2 * xCx / 2);
                    double e2 = d2 * e;
                    /* invalid-value guard of updateDerivatives (kernel.cpp:456-457): the
                     * pair contributes neither score nor derivatives */
                    if (e2 > 1 || e2 < 0 || e2 != e2) continue;
                    score += -d1 * e;
                    if (!compute_derivs) continue;
                    double f = e2 * d1;
                    if (!have_pd) {
                        orc_point_derivs(x, jang, hang, pg, ph);
                        have_pd = 1;
                    }
                    double cj[6][3], xcj[6];
                    for (int k = 0; k < 6; k++) {
                        for (int r = 0; r < 3; r++)
                            cj[k][r] = C[r * 3 + 0] * pg[0][k] + C[r * 3 + 1] * pg[1][k] + C[r * 3 + 2] * pg[2][k];
                        xcj[k] = dot3(d, cj[k]);
                        grad[k] += xcj[k] * f;
                    }
                    if (!compute_hessian) continue;
                    for (int a = 0; a < 6; a++)
                        for (int b = a; b < 6; b++) {
                            double hv[3] = {ph[3 * a][b], ph[3 * a + 1][b], ph[3 * a + 2][b]};
                            double ch = dot3(q, hv); /* x'^T C H_ab (C symmetric) */
                            double jj = pg[0][b] * cj[a][0] + pg[1][b] * cj[a][1] + pg[2][b] * cj[a][2];
                            Hu[a * 6 + b] += f * (-d2 * xcj[a] * xcj[b] + ch + jj);
                        }
                }
    }
    if (H) {
        for (int a = 0; a < 6; a++)
            for (int b = 0; b < 6; b++) H[a * 6 + b] = (b >= a) ? Hu[a * 6 + b] : Hu[b * 6 + a];
    }
    if (npairs) *npairs = pairs;
    return score;
}

double orc_ndt_derivatives(const orc_ndt_grid* g, const orc_ndt_params* prm, const float* scan4,
                           int64_t n, const double p[6], double grad[6], double H[36],
                           int compute_hessian, int64_t* npairs) {
    double d1, d2;
    orc_gauss_consts(prm->resolution, prm->outlier_ratio, &d1, &d2);
    return ndt_pass(g, prm, d1, d2, scan4, n, p, grad, H, 1, compute_hessian, npairs);
}

double orc_ndt_score(const orc_ndt_grid* g, const orc_ndt_params* prm, const float* scan4, int64_t n,
                     const double p[6]) {
    double d1, d2;
    orc_gauss_consts(prm->resolution, prm->outlier_ratio, &d1, &d2);
    return ndt_pass(g, prm, d1, d2, scan4, n, p, NULL, NULL, 0, 0, NULL);
}

int orc_mt_search(orc_line_fn phi, void* ctx, double phi_0, double d_phi_0, double step_init,
                  double step_max, double step_min, int32_t flags, orc_mt_out* out) {
    /* computeStepLengthMT (kernel.cpp:1015-1144) over a scalar function: phi(t) and
     * phi'(t) along the caller's direction.  An ascent direction is reversed
     * (kernel.cpp:1024-1037), so that the search runs on phi_s(a) = phi(sign a),
     * phi_s'(a) = sign phi'(sign a).  The More-Thuente loop is live: interval_converged
     * starts false (the reference starts it at (step_max - step_min) > 0, :1053,
     * which never iterates; ORC_MT_DEAD_LOOP restores that).  With
     * ORC_MT_HESSIAN_AT_ACCEPT the evaluations inside the loop ask for no Hessian
     * (:1123) and one more evaluation with the Hessian follows at the returned step
     * when the loop ran (:1141-1142). */
    memset(out, 0, sizeof(*out));
    out->sign = 1;
    if (d_phi_0 >= 0) {
        if (d_phi_0 == 0) {
            out->exit_reason = ORC_MT_NO_DESCENT;
            return 0;
        }
        d_phi_0 *= -1;
        out->sign = -1;
    }
    const double sg = (double)out->sign;
    const int hess_in_loop = !(flags & ORC_MT_HESSIAN_AT_ACCEPT);
    int max_step_iterations = 10, step_iterations = 0;
    double mu = 1.e-4, nu = 0.9;
    double a_l = 0, a_u = 0;
    double f_l = psi_mt(a_l, phi_0, phi_0, d_phi_0, mu);
    double g_l = dpsi_mt(d_phi_0, d_phi_0, mu);
    double f_u = psi_mt(a_u, phi_0, phi_0, d_phi_0, mu);
    double g_u = dpsi_mt(d_phi_0, d_phi_0, mu);
    int interval_converged = (flags & ORC_MT_DEAD_LOOP) ? ((step_max - step_min) > 0) : 0;
    int open_interval = 1;
    double a_t = step_init;
    a_t = std_min(a_t, step_max);
    a_t = std_max(a_t, step_min);
    double dphi;
    double phi_t = phi(ctx, sg * a_t, 1, &dphi);
    out->evals++;
    double d_phi_t = sg * dphi;
    double psi_t = psi_mt(a_t, phi_t, phi_0, d_phi_0, mu);
    double d_psi_t = dpsi_mt(d_phi_t, d_phi_0, mu);
    while (!interval_converged && step_iterations < max_step_iterations &&
           !(psi_t <= 0 && d_phi_t <= -nu * d_phi_0)) {
        if (open_interval)
            a_t = orc_mt_trial_value(a_l, f_l, g_l, a_u, f_u, g_u, a_t, psi_t, d_psi_t);
        else
            a_t = orc_mt_trial_value(a_l, f_l, g_l, a_u, f_u, g_u, a_t, phi_t, d_phi_t);
        a_t = std_min(a_t, step_max);
        a_t = std_max(a_t, step_min);
        phi_t = phi(ctx, sg * a_t, hess_in_loop, &dphi);
        out->evals++;
        d_phi_t = sg * dphi;
        psi_t = psi_mt(a_t, phi_t, phi_0, d_phi_0, mu);
        d_psi_t = dpsi_mt(d_phi_t, d_phi_0, mu);
        if (open_interval && (psi_t <= 0 && d_psi_t >= 0)) {
            open_interval = 0;
            f_l = f_l + phi_0 - mu * d_phi_0 * a_l;
            g_l = g_l + mu * d_phi_0;
            f_u = f_u + phi_0 - mu * d_phi_0 * a_u;
            g_u = g_u + mu * d_phi_0;
        }
        if (open_interval)
            interval_converged = orc_mt_update_interval(&a_l, &f_l, &g_l, &a_u, &f_u, &g_u, a_t, psi_t, d_psi_t);
        else
            interval_converged = orc_mt_update_interval(&a_l, &f_l, &g_l, &a_u, &f_u, &g_u, a_t, phi_t, d_phi_t);
        step_iterations++;
    }
    if (psi_t <= 0 && d_phi_t <= -nu * d_phi_0)
        out->exit_reason = ORC_MT_ACCEPTED;
    else if (interval_converged)
        out->exit_reason = ORC_MT_INTERVAL;
    else
        out->exit_reason = ORC_MT_MAX_ITER;
    if (step_iterations && !hess_in_loop) {
        (void)phi(ctx, sg * a_t, 1, &dphi);
        out->evals++;
    }
    out->a = a_t;
    out->loop_iterations = step_iterations;
    return 0;
}

typedef struct ndt_line {
    ndt_ctx* c;
    const double* x;
    const double* dir;
    double* x_t;
    double* score;
    double* grad;
    double* H;
    double last_t; /* line parameter of the previous evaluation (NaN before the first) */
    int last_h;    /* the pass whose sums are in the state computed the Hessian */
} ndt_line;

/* phi(t) = -score(x + t dir), phi'(t) = -grad . dir: one derivative pass at the trial
 * pose (kernel.cpp:1058-1066, :1114-1128), leaving score, gradient and (when asked
 * for) Hessian of the last trial in the driver's state.  An evaluation at the same
 * t as the one before it (the search clamps successive trials to the same bound,
 * step_min or step_max) is at the same pose, and its pass would reproduce the sums
 * still in the state bit for bit: unless prm->repeat_passes is set, or it asks for
 * a Hessian that pass did not compute, it takes them without a pass. */
static double ndt_line_eval(void* vctx, double t, int with_hessian, double* dphi) {
    ndt_line* L = (ndt_line*)vctx;
    if (t == L->last_t && (!with_hessian || L->last_h) && !L->c->prm->repeat_passes) {
        *dphi = -(dot6(L->grad, L->dir));
        return -*L->score;
    }
    for (int i = 0; i < 6; i++) L->x_t[i] = L->x[i] + L->dir[i] * t;
    *L->score = ndt_pass(L->c->g, L->c->prm, L->c->d1, L->c->d2, L->c->scan4, L->c->n, L->x_t, L->grad,
                         with_hessian ? L->H : NULL, 1, with_hessian, NULL);
    L->c->passes++;
    L->last_t = t;
    L->last_h = with_hessian;
    *dphi = -(dot6(L->grad, L->dir));
    return -*L->score;
}

int orc_ndt_align(const orc_ndt_grid* g, const orc_ndt_params* prm, const float* scan4, int64_t n,
                  const double guess[16], orc_result* out) {
    /* computeTransformation (kernel.cpp:1175-1272) in its textbook form: normalised
     * Newton direction, step length from the live line search (at most step_size),
     * p += step, converged when a step after the first is shorter than trans_eps. */
    memset(out, 0, sizeof(*out));
    if (!g || !prm || !scan4 || n <= 0 || !guess) return 1;
    ndt_ctx c;
    c.g = g;
    c.prm = prm;
    c.scan4 = scan4;
    c.n = n;
    orc_gauss_consts(prm->resolution, prm->outlier_ratio, &c.d1, &c.d2);
    double p[6], grad[6], H[36], ea[3];
    p[0] = guess[3];
    p[1] = guess[7];
    p[2] = guess[11];
    orc_ndt_euler(guess, ea);
    p[3] = ea[0];
    p[4] = ea[1];
    p[5] = ea[2];
    double score = ndt_pass(g, prm, c.d1, c.d2, scan4, n, p, grad, H, 1, 1, NULL);
    c.passes = 1;
    int iter = 0, converged = 0;
    while (iter < prm->max_iterations) {
        double neg_grad[6] = {-grad[0], -grad[1], -grad[2], -grad[3], -grad[4], -grad[5]};
        double delta_p[6], x_t[6];
        orc_solve6_pivot(delta_p, H, neg_grad);
        double norm = sqrt(delta_p[0] * delta_p[0] + delta_p[1] * delta_p[1] + delta_p[2] * delta_p[2] +
                           delta_p[3] * delta_p[3] + delta_p[4] * delta_p[4] + delta_p[5] * delta_p[5]);
        if (norm == 0 || norm != norm) {
            converged = (norm == norm);
            break;
        }
        for (int i = 0; i < 6; i++) delta_p[i] /= norm;
        /* computeStepLengthMT (kernel.cpp:1015-1144), phi(0) = -score, phi'(0) = -grad . dir */
        ndt_line L = {&c, p, delta_p, x_t, &score, grad, H, NAN, 0};
        orc_mt_out mt;
        orc_mt_search(ndt_line_eval, &L, -score, -(dot6(grad, delta_p)), norm, prm->step_size, prm->trans_eps / 2,
                      prm->hessian_at_accept ? ORC_MT_HESSIAN_AT_ACCEPT : 0, &mt);
        if (mt.sign < 0)
            for (int i = 0; i < 6; i++) delta_p[i] *= -1;
        const double a = mt.a;
        for (int i = 0; i < 6; i++) p[i] = p[i] + delta_p[i] * a;
        iter++;
        if (iter > 1 && fabs(a) < prm->trans_eps) {
            converged = 1;
            break;
        }
    }
    orc_ndt_pose_matrix(p, out->final_f64);
    for (int i = 0; i < 16; i++) out->final_f32[i] = (float)out->final_f64[i];
    out->score = score;
    out->trans_probability = score / (double)n;
    out->iterations = iter;
    out->converged = converged;
    out->passes = c.passes;
    return 0;
}

int orc_ndt_align_many(const orc_ndt_grid* g, const orc_ndt_params* prm, const float* const* scans,
                       const int64_t* counts, const double* guesses, int32_t S, int32_t threads,
                       orc_result* out) {
    int err = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(| : err)
#endif
    for (int32_t s = 0; s < S; s++) err |= orc_ndt_align(g, prm, scans[s], counts[s], guesses + 16 * s, &out[s]);
    (void)threads;
    return err;
}
