/*
 * LC2 oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C CPU restatement of the reference's LC2 similarity path, used as the
 * checker for the HIP engine. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the engine never does.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).
 *
 * Parity status per function (see the test that pins each):
 *   geometry (affine/rigid init, inverse)   pinned: closed forms + golden warp vectors
 *   warp 2D/3D                              pinned: golden vectors fixed.bin,
 *                                           CT_KidneyPhantom_virtual.bin (identity),
 *                                           moving.bin, US_KidneyPhantom_virtual.bin (rigid)
 *   sobel 2D/3D (both orders)               pinned: ramp closed forms + brute force
 *   lc2_exact (FP64)                        pinned: numpy lstsq brute force, known answers,
 *                                           invariances
 *   lc2_cg (FP32 CGLS replica)              pinned: converges to lc2_exact in FP64 mode;
 *                                           FP32 rounding path itself is parity unpinned
 *                                           (needs nvcc to run the reference)
 */
#ifndef LC2_ORACLE_H
#define LC2_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Grid geometry: x-fastest layout, idx = x + y*X + z*X*Y. */
typedef struct {
    int   dims[3];
    float spacing[3];
    float origin[3];
} lc2o_geom;

typedef struct {
    double sum_measure;
    double sum_w;
    double lc2;          /* sum_measure / sum_w, NaN/Inf -> 0 */
    long long n_valid;   /* voxels with w > 0 */
} lc2o_result;

/* ---- geometry (spatialMappers.cuh) ---- */
void lc2o_center(const lc2o_geom* g, int ndim, float out[3]);
void lc2o_virtual_dims(const lc2o_geom* fixed, const float resampled[3], int ndim, int out[3]);
void lc2o_affine_init_3d(const float rot_deg[3], const float trans[3], const float scale[3],
                         const float shear[3], const float center[3], const float shift[3],
                         float out16[16]);
void lc2o_affine_init_2d(float rot_deg, const float trans[2], const float scale[2], float shear,
                         const float center[2], const float shift[2], float out9[9]);
int  lc2o_inverse_4x4(const float m[16], float inv[16]);
int  lc2o_inverse_3x3(const float m[9], float inv[9]);

/* ---- warp (interpolator.cu) ---- */
void lc2o_warp_3d(const float* src, const lc2o_geom* src_geom, const lc2o_geom* virt,
                  const float m12[12], float* out);
void lc2o_warp_2d(const float* src, const lc2o_geom* src_geom, const lc2o_geom* virt,
                  const float m6[6], float* out);

/* ---- Sobel gradient magnitude (cu_lc2_3d.cu / cu_lc2_2d.cu) ---- */
void lc2o_sobel_3d_ref(const float* I, const int dims[3], float* G);
void lc2o_sobel_3d_sep(const float* I, const int dims[3], float* G);
void lc2o_sobel_2d_ref(const float* I, const int dims[2], float* G);
void lc2o_sobel_2d_sep(const float* I, const int dims[2], float* G);

/* ---- LC2 ---- */
/* ndim: 2 (patches (2r+1)^2 in the xy plane, dims[2] must be 1) or 3.
 * zlo/zhi restrict the OUTPUT voxels to z in [zlo, zhi) (2D: rows y in [zlo, zhi)).
 * w_map / m_map may be NULL. */
int lc2o_lc2_exact(const float* U, const float* I, const float* G, const int dims[3], int ndim,
                   int r, int zlo, int zhi, double* w_map, double* m_map, lc2o_result* out);
int lc2o_lc2_cg(const float* U, const float* I, const float* G, const int dims[3], int ndim,
                int r, int zlo, int zhi, int fp64_arith, float* w_map, float* m_map,
                lc2o_result* out);

/* Full single-pose evaluation in the reference's data flow: warp both images into the
 * virtual grid, Sobel of the high-resolution image, LC2.
 *  high_res_is_moving: 1 -> warped moving image is I, warped fixed image is U.
 *  src_ndim: 2 for a 2D moving image (bilinear), 3 for a 3D moving volume (trilinear).
 *  sobel_order: 0 = reference serial tap order, 1 = separable order (engine's). */
int lc2o_eval(const float* fixed, const lc2o_geom* fixed_geom, const float* moving,
              const lc2o_geom* moving_geom, const lc2o_geom* virt, int grid_ndim, int src_ndim,
              int high_res_is_moving, int r, const float m12[12], int solver_cg,
              int sobel_order, lc2o_result* out);

/* Normalized mutual information of the patchSize == 0 path
 * (LC2_3D_class::GetMutualInformationSimilarityMetric, cu_lc2_3d.cu:1196-1301, kernels
 * computeHistogram / computeJointHistogram :868-923, functors cu_lc2_helper.cuh:20-41):
 * bins from min/max of each image over the whole grid, index
 * roundf((v - min) * (bins - 1.0) / (max - min)), p = count / n (empty bins count as
 * p = 1), H = -sum p ln p; NMI = (H_U + H_I) / H_joint, joint index iU + binsU * iI.
 * hist_u / hist_i / hist_joint (optional) receive the integer histograms.
 * fp32_entropy: 1 sums p ln p in FP32 in bin order (the reference's element type; its
 * thrust::reduce order is unpinned), 0 sums in FP64. A constant image returns 0. */
double lc2o_nmi(const float* U, const float* I, long long n, int bins_u, int bins_i,
                int fp32_entropy, int* hist_u, int* hist_i, int* hist_joint);

int lc2o_num_threads(void);
/* Threads of the parallel loops that follow (n >= 1); returns the count in force. */
int lc2o_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
