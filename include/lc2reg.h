/*
 * lc2reg.h -- C ABI of the registration layer above the LC2 engine: a whole YAML-driven
 * registration, and the pieces it is built from (optimizers, configuration I/O) for
 * callers and tests that reach them from C or through an FFI.
 * Statuses: 0 ok, negative on error (message in lc2reg_last_error()).
 */
#ifndef LC2REG_H
#define LC2REG_H

#include "lc2e.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* lc2reg_last_error(void);

/* Host-only geometry (no device needed), the same functions the engine and the facade run:
 * lc2reg_pair_geometry fills info->virt, fixed_center, shift, grid_ndim and src_ndim (the
 * other fields 0) for fixed/moving geometries as lc2e_pair_create does (virtual grid
 * interpolator.cu:320, :337-339; centre and CenterOverlaid shift :495-529);
 * lc2reg_pose_to_matrix is lc2e_pose_to_matrix on such an info (spatialMappers.cuh:579-850,
 * :977-1033); lc2reg_moving_to_fixed is GetMovingToFixedTransform (interpolator.cu:633-679,
 * :1382-1425): inv(VirtualToMoving) as 16 floats (3D sources) or the 2D 3x3 (9 floats),
 * row-major, bottom row forced. */
int lc2reg_pair_geometry(const lc2e_geom* fixed, const lc2e_geom* moving,
                         const float resampled_spacing[3], int dim, int center_overlaid,
                         lc2e_pair_info* out);
int lc2reg_pose_to_matrix(const lc2e_pair_info* info, const double* params, int n_params,
                          float out12[12]);
int lc2reg_moving_to_fixed(const lc2e_pair_info* info, const double* params, int n_params,
                           float* out);

/* One registration as the reference's CLI runs it for -m 2D | 3D | 2DTo3D -c <yaml>
 * (main.cu:728-827): Interpolator + Optimizers from the YAML, DIRECT stage and BOBYQA
 * rounds, GetOptimalTransform, ExportResults (writes <yaml stem>_result.<ext>), WriteOut
 * when `write_images` is set. Outputs (any may be NULL): the optimum parameter vector
 * (12 values for 3D / 2DTo3D, 6 for 2D), the last LC2, the evaluation count, the
 * moving-to-fixed matrix (16 or 9 floats, row-major). */
int lc2reg_register_yaml(const char* mode, const char* yaml_path, int write_images,
                         double* out_params, double* out_lc2, int* out_nevals, float* out_matrix);

/* lc2reg_register_yaml with the reference CLI's OpenIGTLink pose stream (main.cu:761-769,
 * interpolator.cu:359-388, :545-573), 3D only: igtl_port > 0 opens the server on that port
 * and waits for a client before the registration; every evaluated pose is then sent as a
 * TRANSFORM (lc2igtl.h). igtl_port 0: no server (lc2reg_register_yaml). */
int lc2reg_register_yaml_igtl(const char* mode, const char* yaml_path, int write_images,
                              int igtl_port, double* out_params, double* out_lc2, int* out_nevals,
                              float* out_matrix);

/* LC2 of one parameter vector for a YAML-configured pair (Interpolator*::GenerateMappers +
 * Interpolate + GetSimilarityMeasure with the YAML's PatchSize). */
int lc2reg_evaluate_yaml(const char* mode, const char* yaml_path, const double* params, int n_params,
                         double* out_lc2);

/* Batch objective: F[i] = f(X[i*d .. i*d+d-1]) for i < n. */
typedef void (*lc2reg_batch_fn)(int n, int d, const double* X, double* F, void* user);

/* The optimizers behind Optimizers* (see optim.hpp): DIRECT-L and the bound-constrained
 * quadratic-model trust region used for the BOBYQA rounds. */
int lc2reg_minimize_direct(int d, const double* lb, const double* ub, int maxeval,
                           lc2reg_batch_fn f, void* user, double* x_out, double* f_out,
                           int* nevals);
int lc2reg_minimize_qmtr(int d, const double* x0, const double* lb, const double* ub,
                         double xtol_rel, int maxeval, lc2reg_batch_fn f, void* user,
                         double* x_out, double* f_out, int* nevals);

/* The nlopt::opt-shaped optimizer facade (optim.hpp) that the Optimizers* classes use:
 * algorithm 0 = GN_DIRECT_L, 1 = GN_ORIG_DIRECT_L, 2 = LN_BOBYQA (optimizers.cu:298, :854,
 * :937). x0 is BOBYQA's start (ignored by DIRECT); step (NULL: NLopt's default initial
 * step) its initial step per variable; xtol_rel its set_xtol_rel; maxeval the evaluation
 * budget (required by DIRECT; <= 0: none for BOBYQA). Returns the NLopt result code
 * (1 success, 4 xtol reached, 5 maxeval reached; negative on failure). */
int lc2reg_minimize_nlopt(int algorithm, int d, const double* x0, const double* lb,
                          const double* ub, const double* step, double xtol_rel, int maxeval,
                          lc2reg_batch_fn f, void* user, double* x_out, double* f_out,
                          int* nevals);

/* The reference's LC2 classes in one call: LC2_3D_class(x, y, z, US, HR) (z > 1) or
 * LC2_2D_class(x, y, US, HR) (z == 1) on two images that already share a grid
 * (cu_lc2_3d.cuh:14-55, cu_lc2_2d.cuh:14-47), then GetSimilarityMetric(..., patch_size).
 * us and hr hold x*y*z floats, x fastest. Optional outputs (NULL to skip), x*y*z floats
 * each: GetHighResolutionGradientBuffer (|Sobel| of hr) and GetWeightedMapBuffer
 * (measure / w per voxel, 0 where w = 0). */
int lc2reg_lc2_class_metric(int x, int y, int z, const float* us, const float* hr, int patch_size,
                            double* out_lc2, float* out_gradient, float* out_weighted_map);

/* Parse a parameters*.yaml with the configuration reader and write it back out. */
int lc2reg_yaml_roundtrip(const char* in_path, const char* out_path);

#ifdef __cplusplus
}
#endif
#endif
