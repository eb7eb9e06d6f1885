/*
 * ndt_scene.h - seeded synthetic LiDAR scenes for the NDT scan-matching engine.
 *
 * The reference benchmark reads its inputs from data/ndt_input.dat
 * (src/CPU/ndt_mapping/kernel.cpp:143-173), a file that is not distributed with
 * the sources.  This generator produces maps, scans and perturbed initial guesses
 * of the same kind from a 64-bit seed, so every test and benchmark input can be
 * regenerated bit-for-bit on any machine that runs the same build.
 *
 * No counterpart in the reference: it stands in for the missing data file.
 *
 * Scene model: ground plane z = 0 with N(0, ground_sigma) height noise,
 * axis-aligned boxes (side lengths U[box_side_min, box_side_max], heights
 * U[box_h_min, box_h_max]) and vertical cylinders (radius cyl_radius, height
 * cyl_height), all inside [-extent_x/2, extent_x/2] x [-extent_y/2, extent_y/2].
 * Map points are sampled uniformly by surface area and stored as float32
 * (x, y, z, 1) relative to `origin` (a float64 world offset, so UTM-sized
 * coordinates keep float32 resolution).
 *
 * A scan is a random subset of the map points within max_range of a sensor
 * pose, mapped into the sensor frame through the inverse ground-truth pose, with
 * N(0, noise_sigma) noise per axis.  Poses are row-major 4x4 float64 matrices
 * mapping sensor coordinates to world coordinates (world = local + origin).
 */
#ifndef NDT_SCENE_H
#define NDT_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ndt_scene_cfg {
    uint64_t seed;
    double extent_x, extent_y;     /* scene footprint in metres */
    int32_t n_boxes, n_cylinders;
    double box_side_min, box_side_max;
    double box_h_min, box_h_max;
    double cyl_radius, cyl_height;
    double ground_sigma;           /* std-dev of ground height noise */
    double origin[3];              /* world offset of the local frame */
} ndt_scene_cfg;

typedef struct ndt_scan_cfg {
    int64_t n_points;              /* points per scan */
    double max_range;              /* sensor range in metres */
    double noise_sigma;            /* per-axis scan noise */
    double sensor_height;          /* sensor z above ground */
    double sensor_margin;          /* keep the sensor this far inside the scene */
    double perturb_xyz[3];         /* guess translation error U[-a, a] per axis */
    double perturb_yaw_deg;        /* guess yaw error U[-a, a] */
    double perturb_rollpitch_deg;  /* guess roll/pitch error U[-a, a] */
} ndt_scan_cfg;

/* Defaults of the 100 m x 100 m scene (40 boxes, 60 cylinders). */
void ndt_scene_cfg_default(ndt_scene_cfg* cfg, uint64_t seed);
/* Defaults of a 2 k point scan with +-0.5 m / +-2 deg guess errors. */
void ndt_scan_cfg_default(ndt_scan_cfg* cfg);

/* Writes m map points (x, y, z, 1) in the local frame into out_xyzw (4*m floats).
 * Returns 0 on success, non-zero on invalid arguments. */
int ndt_scene_map(const ndt_scene_cfg* cfg, int64_t m, float* out_xyzw);

/* Draws one scan from a map produced by ndt_scene_map.  scan_seed selects the
 * sensor pose, the point subset, the noise and the guess perturbation.
 * out_scan_xyzw receives cfg->n_points points (x, y, z, 1) in the sensor frame,
 * gt_pose and guess_pose receive row-major 4x4 float64 world poses.
 * Points are drawn without replacement while enough map points are in range,
 * with replacement otherwise.  Returns 0 on success, 1 on invalid arguments,
 * 2 when no map point is in range of the drawn sensor position. */
int ndt_scene_scan(const ndt_scene_cfg* scene, const float* map_xyzw, int64_t m,
                   const ndt_scan_cfg* cfg, uint64_t scan_seed,
                   float* out_scan_xyzw, double gt_pose[16], double guess_pose[16]);

/* Draws n scans on `threads` host threads (<= 0: all), scan k with cfgs[k] and
 * seeds[k], exactly as ndt_scene_scan would: its points go to
 * out_xyzw + 4 * offsets[k], its poses to gt_poses + 16 k and guess_poses + 16 k,
 * its ndt_scene_scan return code to status[k].  The map is bucketed once, so a
 * scan visits only the map points near its sensor.  Returns non-zero on invalid
 * arguments. */
int ndt_scene_scans(const ndt_scene_cfg* scene, const float* map_xyzw, int64_t m, const ndt_scan_cfg* cfgs,
                    const uint64_t* seeds, int32_t n, const int64_t* offsets, float* out_xyzw, double* gt_poses,
                    double* guess_poses, int32_t threads, int32_t* status);

/* Row-major 4x4 pose from translation (x, y, z) and Euler angles with
 * R = Rx(roll) * Ry(pitch) * Rz(yaw). */
void ndt_scene_pose(double x, double y, double z, double roll, double pitch, double yaw,
                    double out[16]);

#ifdef __cplusplus
}
#endif

#endif
