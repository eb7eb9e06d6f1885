/*
 * Seeded synthetic LiDAR scene generator (map, scans, ground truth, guesses).
 * Interface and scene model: include/ndt_scene.h.
 *
 * Determinism: only integer PRNG state and IEEE double arithmetic feed the
 * outputs, plus log/sqrt/cos/sin from libm, so the same build gives the same
 * points on every machine with the same C library.
 */
#include "ndt_scene.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SCENE_PI 3.14159265358979323846

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t* r, uint64_t seed, uint64_t stream) {
    uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull);
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* xoshiro256** */
static inline uint64_t rng_next(rng_t* r) {
    uint64_t* s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* uniform in [0, 1) with 53 random bits */
static inline double rng_u01(rng_t* r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static inline double rng_uniform(rng_t* r, double a, double b) { return a + (b - a) * rng_u01(r); }

/* standard normal, Box-Muller (one value per call, no caching: keeps the stream simple) */
static double rng_normal(rng_t* r) {
    double u1 = rng_u01(r);
    double u2 = rng_u01(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * SCENE_PI * u2);
}

/* uniform index in [0, n) (Lemire's multiply-shift, bias < 2^-64 * n, irrelevant here) */
static inline uint64_t rng_index(rng_t* r, uint64_t n) {
    return (uint64_t)(((unsigned __int128)rng_next(r) * n) >> 64);
}

void ndt_scene_cfg_default(ndt_scene_cfg* cfg, uint64_t seed) {
    memset(cfg, 0, sizeof(*cfg));
    cfg->seed = seed;
    cfg->extent_x = 100.0;
    cfg->extent_y = 100.0;
    cfg->n_boxes = 40;
    cfg->n_cylinders = 60;
    cfg->box_side_min = 2.0;
    cfg->box_side_max = 15.0;
    cfg->box_h_min = 3.0;
    cfg->box_h_max = 12.0;
    cfg->cyl_radius = 0.15;
    cfg->cyl_height = 4.0;
    cfg->ground_sigma = 0.02;
}

void ndt_scan_cfg_default(ndt_scan_cfg* cfg) {
    memset(cfg, 0, sizeof(*cfg));
    cfg->n_points = 2000;
    cfg->max_range = 40.0;
    cfg->noise_sigma = 0.01;
    cfg->sensor_height = 1.8;
    cfg->sensor_margin = 20.0;
    cfg->perturb_xyz[0] = cfg->perturb_xyz[1] = cfg->perturb_xyz[2] = 0.5;
    cfg->perturb_yaw_deg = 2.0;
    cfg->perturb_rollpitch_deg = 0.0;
}

void ndt_scene_pose(double x, double y, double z, double roll, double pitch, double yaw,
                    double out[16]) {
    double cx = cos(roll), sx = sin(roll);
    double cy = cos(pitch), sy = sin(pitch);
    double cz = cos(yaw), sz = sin(yaw);
    /* R = Rx * Ry * Rz */
    out[0] = cy * cz;
    out[1] = -cy * sz;
    out[2] = sy;
    out[4] = cx * sz + sx * sy * cz;
    out[5] = cx * cz - sx * sy * sz;
    out[6] = -sx * cy;
    out[8] = sx * sz - cx * sy * cz;
    out[9] = sx * cz + cx * sy * sz;
    out[10] = cx * cy;
    out[3] = x;
    out[7] = y;
    out[11] = z;
    out[12] = out[13] = out[14] = 0.0;
    out[15] = 1.0;
}

typedef struct {
    double cx, cy, sx, sy, h;
} box_t;

typedef struct {
    double cx, cy;
} cyl_t;

/* surface kinds */
enum { SURF_GROUND = 0, SURF_WALL_XN, SURF_WALL_XP, SURF_WALL_YN, SURF_WALL_YP, SURF_TOP, SURF_CYL };

typedef struct {
    int kind;
    int obj;
    double area;
} surf_t;

int ndt_scene_map(const ndt_scene_cfg* cfg, int64_t m, float* out) {
    if (!cfg || !out || m <= 0 || cfg->extent_x <= 0 || cfg->extent_y <= 0 || cfg->n_boxes < 0 ||
        cfg->n_cylinders < 0)
        return 1;
    rng_t r;
    rng_seed(&r, cfg->seed, 1);
    const double W = cfg->extent_x, H = cfg->extent_y;
    int nb = cfg->n_boxes, nc = cfg->n_cylinders;
    box_t* boxes = (box_t*)malloc(sizeof(box_t) * (nb > 0 ? nb : 1));
    cyl_t* cyls = (cyl_t*)malloc(sizeof(cyl_t) * (nc > 0 ? nc : 1));
    int ns = 1 + 5 * nb + nc;
    surf_t* surf = (surf_t*)malloc(sizeof(surf_t) * ns);
    double* cum = (double*)malloc(sizeof(double) * ns);
    if (!boxes || !cyls || !surf || !cum) {
        free(boxes); free(cyls); free(surf); free(cum);
        return 1;
    }
    for (int b = 0; b < nb; b++) {
        double sx = rng_uniform(&r, cfg->box_side_min, cfg->box_side_max);
        double sy = rng_uniform(&r, cfg->box_side_min, cfg->box_side_max);
        double h = rng_uniform(&r, cfg->box_h_min, cfg->box_h_max);
        double hx = sx * 0.5 < W * 0.5 ? W * 0.5 - sx * 0.5 : 0.0;
        double hy = sy * 0.5 < H * 0.5 ? H * 0.5 - sy * 0.5 : 0.0;
        boxes[b].cx = rng_uniform(&r, -hx, hx);
        boxes[b].cy = rng_uniform(&r, -hy, hy);
        boxes[b].sx = sx;
        boxes[b].sy = sy;
        boxes[b].h = h;
    }
    for (int c = 0; c < nc; c++) {
        double hx = W * 0.5 - cfg->cyl_radius, hy = H * 0.5 - cfg->cyl_radius;
        cyls[c].cx = rng_uniform(&r, -hx, hx);
        cyls[c].cy = rng_uniform(&r, -hy, hy);
    }
    int k = 0;
    surf[k].kind = SURF_GROUND; surf[k].obj = 0; surf[k].area = W * H; k++;
    for (int b = 0; b < nb; b++) {
        const box_t* B = &boxes[b];
        surf[k].kind = SURF_WALL_XN; surf[k].obj = b; surf[k].area = B->sy * B->h; k++;
        surf[k].kind = SURF_WALL_XP; surf[k].obj = b; surf[k].area = B->sy * B->h; k++;
        surf[k].kind = SURF_WALL_YN; surf[k].obj = b; surf[k].area = B->sx * B->h; k++;
        surf[k].kind = SURF_WALL_YP; surf[k].obj = b; surf[k].area = B->sx * B->h; k++;
        surf[k].kind = SURF_TOP; surf[k].obj = b; surf[k].area = B->sx * B->sy; k++;
    }
    for (int c = 0; c < nc; c++) {
        surf[k].kind = SURF_CYL; surf[k].obj = c;
        surf[k].area = 2.0 * SCENE_PI * cfg->cyl_radius * cfg->cyl_height;
        k++;
    }
    double total = 0.0;
    for (int i = 0; i < ns; i++) {
        total += surf[i].area;
        cum[i] = total;
    }
    for (int64_t i = 0; i < m; i++) {
        double u = rng_u01(&r) * total;
        int lo = 0, hi = ns - 1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (cum[mid] > u) hi = mid; else lo = mid + 1;
        }
        const surf_t* S = &surf[lo];
        double x = 0, y = 0, z = 0;
        if (S->kind == SURF_GROUND) {
            x = rng_uniform(&r, -W * 0.5, W * 0.5);
            y = rng_uniform(&r, -H * 0.5, H * 0.5);
            z = cfg->ground_sigma * rng_normal(&r);
        } else if (S->kind == SURF_CYL) {
            const cyl_t* C = &cyls[S->obj];
            double a = rng_uniform(&r, 0.0, 2.0 * SCENE_PI);
            x = C->cx + cfg->cyl_radius * cos(a);
            y = C->cy + cfg->cyl_radius * sin(a);
            z = rng_uniform(&r, 0.0, cfg->cyl_height);
        } else {
            const box_t* B = &boxes[S->obj];
            double x0 = B->cx - 0.5 * B->sx, x1 = B->cx + 0.5 * B->sx;
            double y0 = B->cy - 0.5 * B->sy, y1 = B->cy + 0.5 * B->sy;
            switch (S->kind) {
            case SURF_WALL_XN: x = x0; y = rng_uniform(&r, y0, y1); z = rng_uniform(&r, 0.0, B->h); break;
            case SURF_WALL_XP: x = x1; y = rng_uniform(&r, y0, y1); z = rng_uniform(&r, 0.0, B->h); break;
            case SURF_WALL_YN: y = y0; x = rng_uniform(&r, x0, x1); z = rng_uniform(&r, 0.0, B->h); break;
            case SURF_WALL_YP: y = y1; x = rng_uniform(&r, x0, x1); z = rng_uniform(&r, 0.0, B->h); break;
            default: x = rng_uniform(&r, x0, x1); y = rng_uniform(&r, y0, y1); z = B->h; break;
            }
        }
        out[4 * i + 0] = (float)x;
        out[4 * i + 1] = (float)y;
        out[4 * i + 2] = (float)z;
        out[4 * i + 3] = 1.0f;
    }
    free(boxes); free(cyls); free(surf); free(cum);
    return 0;
}

/* Map points bucketed on an 8 m grid in the x-y plane, indices increasing
 * within each bucket, so that the points within sensor range can be gathered
 * without visiting the whole map. */
typedef struct {
    double x0, y0, cell;
    int64_t nx, ny;
    int64_t* start; /* nx * ny + 1 */
    int64_t* idx;
} bucket_index;

static void bucket_free(bucket_index* b) {
    free(b->start);
    free(b->idx);
    b->start = b->idx = NULL;
}

static int64_t bucket_of(const bucket_index* b, double x, double y) {
    int64_t ix = (int64_t)floor((x - b->x0) / b->cell), iy = (int64_t)floor((y - b->y0) / b->cell);
    if (ix < 0) ix = 0;
    if (iy < 0) iy = 0;
    if (ix >= b->nx) ix = b->nx - 1;
    if (iy >= b->ny) iy = b->ny - 1;
    return iy * b->nx + ix;
}

static int bucket_build(const float* map, int64_t m, bucket_index* b) {
    memset(b, 0, sizeof(*b));
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
    for (int64_t i = 0; i < m; i++) {
        const double x = map[4 * i], y = map[4 * i + 1], z = map[4 * i + 2];
        if (!isfinite(x) || !isfinite(y) || !isfinite(z)) continue;
        if (x < lo[0]) lo[0] = x;
        if (x > hi[0]) hi[0] = x;
        if (y < lo[1]) lo[1] = y;
        if (y > hi[1]) hi[1] = y;
    }
    if (lo[0] > hi[0]) return 1;
    b->cell = 8.0;
    b->x0 = lo[0];
    b->y0 = lo[1];
    b->nx = (int64_t)floor((hi[0] - lo[0]) / b->cell) + 1;
    b->ny = (int64_t)floor((hi[1] - lo[1]) / b->cell) + 1;
    if (b->nx * b->ny > (1ll << 26)) return 1;
    b->start = (int64_t*)calloc((size_t)(b->nx * b->ny + 1), sizeof(int64_t));
    b->idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    if (!b->start || !b->idx) {
        bucket_free(b);
        return 1;
    }
    for (int64_t i = 0; i < m; i++) {
        const double x = map[4 * i], y = map[4 * i + 1], z = map[4 * i + 2];
        if (!isfinite(x) || !isfinite(y) || !isfinite(z)) continue;
        b->start[bucket_of(b, x, y) + 1]++;
    }
    for (int64_t k = 1; k <= b->nx * b->ny; k++) b->start[k] += b->start[k - 1];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(b->nx * b->ny));
    if (!fill) {
        bucket_free(b);
        return 1;
    }
    memcpy(fill, b->start, sizeof(int64_t) * (size_t)(b->nx * b->ny));
    for (int64_t i = 0; i < m; i++) {
        const double x = map[4 * i], y = map[4 * i + 1], z = map[4 * i + 2];
        if (!isfinite(x) || !isfinite(y) || !isfinite(z)) continue;
        b->idx[fill[bucket_of(b, x, y)]++] = i;
    }
    free(fill);
    return 0;
}

/* ascending sort of n non-negative indices below 2^48: LSD radix sort, 16-bit
 * digits, through tmp (n entries); digits that are zero everywhere are skipped */
static void sort_indices(int64_t* a, int64_t* tmp, int64_t n) {
    int64_t hi = 0;
    for (int64_t i = 0; i < n; i++) hi |= a[i];
    static const int B = 16;
    size_t cnt[1 << 16];
    for (int shift = 0; shift < 48 && (hi >> shift); shift += B) {
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < n; i++) cnt[(a[i] >> shift) & 0xFFFF]++;
        size_t sum = 0;
        for (int d = 0; d < (1 << B); d++) {
            size_t c = cnt[d];
            cnt[d] = sum;
            sum += c;
        }
        for (int64_t i = 0; i < n; i++) tmp[cnt[(a[i] >> shift) & 0xFFFF]++] = a[i];
        memcpy(a, tmp, sizeof(int64_t) * (size_t)n);
    }
}

/* Indices of the map points within range of (sx, sy, sz), in increasing order:
 * every point (bidx == NULL) or the points of the buckets the range box touches.
 * Both give the same list.  Returns the count, or -1 when out of memory. */
static int64_t gather_candidates(const float* map, int64_t m, const bucket_index* bidx, double sx, double sy,
                                 double sz, double range, int64_t** cand, int64_t* cap) {
    const double r2 = range * range;
    int64_t nc = 0;
#define NDT_SCENE_KEEP(i)                                                                                     \
    do {                                                                                                      \
        double dx = (double)map[4 * (i)] - sx, dy = (double)map[4 * (i) + 1] - sy,                            \
               dz = (double)map[4 * (i) + 2] - sz;                                                            \
        if (dx * dx + dy * dy + dz * dz <= r2) {                                                              \
            if (nc == *cap) {                                                                                 \
                int64_t* nb = (int64_t*)realloc(*cand, sizeof(int64_t) * (size_t)(*cap * 2));                 \
                if (!nb) return -1;                                                                           \
                *cand = nb;                                                                                   \
                *cap *= 2;                                                                                    \
            }                                                                                                 \
            (*cand)[nc++] = (i);                                                                              \
        }                                                                                                     \
    } while (0)
    if (!bidx) {
        for (int64_t i = 0; i < m; i++) NDT_SCENE_KEEP(i);
        return nc;
    }
    const int64_t b0 = bucket_of(bidx, sx - range, sy - range), b1 = bucket_of(bidx, sx + range, sy + range);
    const int64_t x0 = b0 % bidx->nx, y0 = b0 / bidx->nx, x1 = b1 % bidx->nx, y1 = b1 / bidx->nx;
    for (int64_t by = y0; by <= y1; by++)
        for (int64_t bx = x0; bx <= x1; bx++) {
            const int64_t k = by * bidx->nx + bx;
            for (int64_t j = bidx->start[k]; j < bidx->start[k + 1]; j++) NDT_SCENE_KEEP(bidx->idx[j]);
        }
#undef NDT_SCENE_KEEP
    int64_t* tmp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc > 0 ? nc : 1));
    if (!tmp) return -1;
    sort_indices(*cand, tmp, nc);
    free(tmp);
    return nc;
}

static int scan_impl(const ndt_scene_cfg* scene, const float* map, int64_t m, const bucket_index* bidx,
                     const ndt_scan_cfg* cfg, uint64_t scan_seed, float* out, double gt_pose[16],
                     double guess_pose[16]) {
    if (!scene || !map || m <= 0 || !cfg || cfg->n_points <= 0 || !out || !gt_pose || !guess_pose)
        return 1;
    rng_t r;
    rng_seed(&r, scene->seed ^ 0x5CA11AB1Eull, 1000 + scan_seed);
    const double W = scene->extent_x, H = scene->extent_y;
    double mx = W * 0.5 - cfg->sensor_margin, my = H * 0.5 - cfg->sensor_margin;
    if (mx < 0) mx = 0;
    if (my < 0) my = 0;
    double sx = rng_uniform(&r, -mx, mx);
    double sy = rng_uniform(&r, -my, my);
    double sz = cfg->sensor_height;
    double yaw = rng_uniform(&r, -SCENE_PI, SCENE_PI);
    double R[16];
    ndt_scene_pose(sx + scene->origin[0], sy + scene->origin[1], sz + scene->origin[2], 0.0, 0.0, yaw, R);
    memcpy(gt_pose, R, sizeof(double) * 16);

    /* candidate map points within range (local frame), in map order */
    int64_t cap = 1024;
    int64_t* cand = (int64_t*)malloc(sizeof(int64_t) * cap);
    if (!cand) return 1;
    int64_t nc = gather_candidates(map, m, bidx, sx, sy, sz, cfg->max_range, &cand, &cap);
    if (nc < 0) {
        free(cand);
        return 1;
    }
    if (nc == 0) {
        free(cand);
        return 2;
    }
    int64_t n = cfg->n_points;
    for (int64_t j = 0; j < n; j++) {
        int64_t idx;
        if (n <= nc) {
            /* partial Fisher-Yates: draw without replacement */
            int64_t pick = j + (int64_t)rng_index(&r, (uint64_t)(nc - j));
            int64_t t = cand[j]; cand[j] = cand[pick]; cand[pick] = t;
            idx = cand[j];
        } else {
            idx = cand[rng_index(&r, (uint64_t)nc)];
        }
        /* sensor frame: R^T (x - t) */
        double dx = (double)map[4 * idx] - sx, dy = (double)map[4 * idx + 1] - sy, dz = (double)map[4 * idx + 2] - sz;
        double px = R[0] * dx + R[4] * dy + R[8] * dz;
        double py = R[1] * dx + R[5] * dy + R[9] * dz;
        double pz = R[2] * dx + R[6] * dy + R[10] * dz;
        px += cfg->noise_sigma * rng_normal(&r);
        py += cfg->noise_sigma * rng_normal(&r);
        pz += cfg->noise_sigma * rng_normal(&r);
        out[4 * j + 0] = (float)px;
        out[4 * j + 1] = (float)py;
        out[4 * j + 2] = (float)pz;
        out[4 * j + 3] = 1.0f;
    }
    free(cand);
    const double deg = SCENE_PI / 180.0;
    double ex = rng_uniform(&r, -cfg->perturb_xyz[0], cfg->perturb_xyz[0]);
    double ey = rng_uniform(&r, -cfg->perturb_xyz[1], cfg->perturb_xyz[1]);
    double ez = rng_uniform(&r, -cfg->perturb_xyz[2], cfg->perturb_xyz[2]);
    double eyaw = rng_uniform(&r, -cfg->perturb_yaw_deg, cfg->perturb_yaw_deg) * deg;
    double eroll = rng_uniform(&r, -cfg->perturb_rollpitch_deg, cfg->perturb_rollpitch_deg) * deg;
    double epitch = rng_uniform(&r, -cfg->perturb_rollpitch_deg, cfg->perturb_rollpitch_deg) * deg;
    ndt_scene_pose(sx + scene->origin[0] + ex, sy + scene->origin[1] + ey, sz + scene->origin[2] + ez,
                   eroll, epitch, yaw + eyaw, guess_pose);
    return 0;
}

int ndt_scene_scan(const ndt_scene_cfg* scene, const float* map, int64_t m, const ndt_scan_cfg* cfg,
                   uint64_t scan_seed, float* out, double gt_pose[16], double guess_pose[16]) {
    return scan_impl(scene, map, m, NULL, cfg, scan_seed, out, gt_pose, guess_pose);
}

int ndt_scene_scans(const ndt_scene_cfg* scene, const float* map, int64_t m, const ndt_scan_cfg* cfgs,
                    const uint64_t* seeds, int32_t n, const int64_t* offsets, float* out, double* gt_poses,
                    double* guess_poses, int32_t threads, int32_t* status) {
    if (!scene || !map || m <= 0 || !cfgs || !seeds || n < 0 || !offsets || !out || !gt_poses || !guess_poses ||
        !status)
        return 1;
    bucket_index b;
    if (bucket_build(map, m, &b)) return 1;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int32_t k = 0; k < n; k++)
        status[k] = scan_impl(scene, map, m, &b, &cfgs[k], seeds[k], out + 4 * offsets[k], gt_poses + 16 * (int64_t)k,
                              guess_poses + 16 * (int64_t)k);
    (void)threads;
    bucket_free(&b);
    return 0;
}
