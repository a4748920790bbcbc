/*
 * lc2e.h -- C ABI of the MI355X LC2 similarity engine.
 *
 * Plain C types only (no HIP or torch types): opaque handles, pointers and sizes.
 * Every entry point returns an int status: LC2E_OK (0) or LC2E_OK_DEGENERATE (1)
 * on success, a negative LC2E_ERR_* code on failure; lc2e_last_error() gives the
 * message of the last failure on the calling thread. No C++ exception crosses it.
 * Host memory passed in is caller-owned and only read/written during the call
 * (the *_async entry point excepted: its device buffers must stay alive until
 * the stream reaches the work). Device memory is engine-owned, resident until
 * the matching *_destroy. One engine serves one host thread at a time.
 *
 * The reference keeps this path behind C++ classes (Interpolator*, LC2_*_class)
 * and CUDA __constant__ state; each entry point below names the reference
 * interface it replaces (file:line, relative to the reference repository).
 */
#ifndef LC2E_H
#define LC2E_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LC2E_OK 0
#define LC2E_OK_DEGENERATE 1     /* LC2 was NaN/Inf (sum of weights 0) and is returned as 0.0 */
#define LC2E_ERR_INVALID (-1)    /* bad argument or handle */
#define LC2E_ERR_HIP (-2)        /* HIP runtime failure */
#define LC2E_ERR_BUDGET (-3)     /* allocation would exceed the engine's HBM budget */
#define LC2E_ERR_UNSUPPORTED (-4)
#define LC2E_ERR_SINGULAR (-5)   /* singular transform */

typedef struct lc2e_engine lc2e_engine;
typedef struct lc2e_volume lc2e_volume;
typedef struct lc2e_pair lc2e_pair;

typedef enum {
    LC2E_U8 = 0,
    LC2E_I16 = 1,
    LC2E_F16 = 2,
    LC2E_F32 = 3,
    LC2E_STORE_AUTO = 100 /* storage only: smallest lossless of F16 / F32 */
} lc2e_dtype;

/* x-fastest layout: idx = x + y*dims[0] + z*dims[0]*dims[1]; 2D images use dims[2] = 1,
 * a stack of 2D frames uses dims[2] = number of frames. */
typedef struct {
    int dims[3];
    float spacing[3];
    float origin[3];
} lc2e_geom;

typedef enum {
    LC2E_STAGE_WARP_I = 0,   /* high-resolution image on the virtual grid, float32 */
    LC2E_STAGE_WARP_US = 1,  /* ultrasound image on the virtual grid, float32 */
    LC2E_STAGE_GRAD = 2,     /* |Sobel(I)|, float32 */
    LC2E_STAGE_W = 3,        /* per-voxel weight w = std(US patch), float64 */
    LC2E_STAGE_MEASURE = 4   /* per-voxel measure, float64 */
} lc2e_stage;

/* Static description of a pair's virtual grid and the constants the pose ->
 * matrix conversion needs (Interpolator*::InitiateMappers, interpolator.cu:495-529). */
typedef struct {
    lc2e_geom virt;          /* origin = fixed origin, spacing = resampled, dims per ceil rule */
    float fixed_center[3];   /* rotation centre = fixed image centre */
    float shift[3];          /* c_moving - c_virtual if center_overlaid, else 0 */
    int grid_ndim;           /* 2 or 3 */
    int src_ndim;            /* 2 (bilinear 2D moving image) or 3 (trilinear volume) */
    int high_res_is_moving;
    int patch_radius;
    int n_frames;            /* planes of an independent-2D-frame stack (1 otherwise) */
    int per_frame_matrices;  /* 1: each pose carries n_frames matrices */
} lc2e_pair_info;

/* Last error message of the calling thread ("" if none). */
const char* lc2e_last_error(void);

/* Engine on one device with an HBM budget for everything it allocates (0 = no limit).
 * Replaces the reference's implicit default-device/default-stream state
 * (interpolator.cuh:27-41 global __constant__ mappers). */
int lc2e_create(int device, size_t hbm_budget_bytes, lc2e_engine** out);
void lc2e_destroy(lc2e_engine* e);
/* Bytes currently allocated by the engine and its budget. */
int lc2e_memory(lc2e_engine* e, size_t* used, size_t* budget);

/* Upload a resident fixed or moving image. Replaces Interpolator*'s thrust copies of the
 * volumes (interpolator.cu:276, :315; :1616-1627 for raw files). `storage` is LC2E_F16,
 * LC2E_F32 or LC2E_STORE_AUTO (U8 -> F16, F16 -> F16, I16 -> F16 when every value lies in
 * [-2048, 2048] (CT in HU; exact in half precision) else F32, F32 -> F32: all lossless).
 * The volume and a zeroed tail of X*Y + X + 64 voxels must fit 4 GiB (32-bit byte
 * offsets): about 2^31 voxels with F16 storage, 2^30 with F32 (LC2E_ERR_UNSUPPORTED
 * otherwise). Every voxel must be finite, and F32 data stored as F16 must lie within
 * +-65504 (LC2E_ERR_INVALID otherwise). */
int lc2e_volume_upload(lc2e_engine* e, const void* host, lc2e_dtype dtype, const lc2e_geom* geom,
                       lc2e_dtype storage, lc2e_volume** out);
void lc2e_volume_destroy(lc2e_volume* v);

/* A registration pair: virtual grid (interpolator.cu:320, :337-339), centre and shift
 * (InitiateMappers, interpolator.cu:495-529), LC2 buffers (LC2_*_class::PortedFromInterpolator,
 * cu_lc2_3d.cu:942-975) and the pose-independent warp of the fixed image.
 *  high_res_is_moving: HighResolutionModality "Moving" (1) or "Fixed" (0), interpolator.cu:590-605.
 *  dim: LC2 grid dimensionality; 2 with a 3D moving volume is the 2D->3D mode
 *       (Interpolator2DTo3D, interpolator.cu:1538-1786); a fixed image with dims[2] > 1 and
 *       dim 2 is a stack of independent 2D frames.
 *  patch_radius: 1..64. Radii up to 4 (3D) / 10 (2D) run the fused kernels; larger radii run
 *       a separable path (the reference's global-memory kernel for large patches,
 *       cu_lc2_3d.cu:769-827), whose buffers -- the moving image and its |Sobel| on the grid
 *       and two f64 box-sum buffers of 7 (moving high-resolution image) or 10 channels -- are
 *       allocated here, against the budget. */
int lc2e_pair_create(lc2e_engine* e, lc2e_volume* fixed, lc2e_volume* moving,
                     const float resampled_spacing[3], int high_res_is_moving,
                     int center_overlaid, int patch_radius, int dim, lc2e_pair** out);
void lc2e_pair_destroy(lc2e_pair* p);
int lc2e_pair_get_info(lc2e_pair* p, lc2e_pair_info* info);
/* Frame stacks only: 1 = every pose carries one matrix per frame (tracked sweeps). */
int lc2e_pair_set_per_frame(lc2e_pair* p, int per_frame_matrices);

/* Pose parameters -> VirtualToMoving 3x4 row-major float matrix of this pair
 * (Spatial3DMapper::AffineInit/RigidInit spatialMappers.cuh:579-850,
 *  Spatial2DMapper::AffineInit/RigidInit :977-1033).
 *  3D sources: params = [TX,TY,TZ,RX,RY,RZ,SX,SY,SZ,ShX,ShY,ShZ] (n = 12) or rigid (n = 6).
 *  2D sources: params = [TX,TY,R,SX,SY,Sh] (n = 6) or rigid (n = 3); the 2x3 matrix is
 *  embedded as [[a,b,0,tx],[c,d,0,ty],[0,0,1,0]]. */
int lc2e_pose_to_matrix(lc2e_pair* p, const double* params, int n_params, float out12[12]);

/* One pose: GenerateMappers -> Interpolate -> GetSimilarityMeasure
 * (interpolator.cu:542-631; LC2_*_class::GetSimilarityMetric cu_lc2_3d.cu:1119-1194,
 * cu_lc2_2d.cu:599-653). m12 = VirtualToMoving as 3x4 row-major (n_frames*12 floats for
 * per-frame stacks). LC2 NaN/Inf -> 0 with status LC2E_OK_DEGENERATE. */
int lc2e_eval(lc2e_pair* p, const float* m12, double* lc2);

/* n poses in one fused launch; no counterpart in the reference, whose global __constant__
 * transform allows one pose in flight (interpolator.cuh:27-41). Results are bit-identical
 * to n single evaluations. Optional outputs (NULL to skip): sum of measures, sum of w. */
int lc2e_eval_batch(lc2e_pair* p, int n, const float* mats, double* lc2, double* sum_measure,
                    double* sum_w);

/* lc2e_eval_batch without any allocation: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The launch's matrices, partial sums and results live in a fixed
 * device array of the library (16 MiB per device, held until the process ends) instead of
 * the pair's batch buffers; the call uses no hipMalloc/hipFree and no allocating library
 * routine. Same kernels, so the results are bit-identical to lc2e_eval_batch. A call needs
 * n*(24 + 48*mats_per_pose) + n*16*(partial sums per pose) bytes (each part rounded up to
 * 256); a larger call fails with LC2E_ERR_INVALID. Every call on a device shares that
 * device's array: concurrent calls would race on it, so the library serialises them per
 * device (a call waits for the one in progress). */
int lc2e_eval_batch_noalloc(lc2e_pair* p, int n, const float* mats, double* lc2,
                            double* sum_measure, double* sum_w);

/* Stream-ordered form for serving loops: d_mats and d_out are DEVICE pointers,
 * d_out receives 3 doubles per pose (sum_measure, sum_w, lc2). `stream` is a hipStream_t
 * (NULL = the engine's stream). Nothing is synchronised. The pair's partial-sum buffer is
 * shared by its launches: two launches of ONE pair must not overlap (enqueue them on one
 * stream, or wait for the first); launches of different pairs may run concurrently. A
 * batch larger than any before it on this pair grows that buffer first (hipFree and
 * hipMalloc, which wait for the device). */
int lc2e_eval_batch_async(lc2e_pair* p, int n, const float* d_mats, double* d_out, void* stream);
int lc2e_sync(lc2e_engine* e);

/* Normalized mutual information, the reference's patchSize == 0 metric
 * (LC2_3D_class::GetMutualInformationSimilarityMetric, cu_lc2_3d.cu:1196-1301, reached from
 * Interpolator3D::GetSimilarityMeasure, interpolator.cu:617-622; its callers maximise it,
 * optimizers.cu:722-724, :855-858): for each of n poses, (H_US + H_HR) / H_joint of the
 * two images on the virtual grid, with bins_us / bins_hr bins (2..256; the reference uses
 * 256 and 256) spread over each image's minimum..maximum and bin index
 * roundf((v - min) * (bins - 1) / (max - min)). Entropies are summed in FP64 (the
 * reference: FP32). nmi[n]; 0 when an image is constant. */
int lc2e_eval_mi_batch(lc2e_pair* p, int n, const float* mats, int bins_us, int bins_hr,
                       double* nmi);

/* Per-stage buffers for parity tests at one pose (the reference's buffer getters,
 * cu_lc2_3d.cuh:45-49, cu_lc2_2d.cuh:40-46). host receives virt voxels of float32
 * (WARP_I, WARP_US, GRAD) or float64 (W, MEASURE). */
int lc2e_debug_fetch(lc2e_pair* p, lc2e_stage stage, const float* m12, void* host);

/* Launch configuration of the 3D fused kernel: output planes per work-group (z-chunk).
 * Each launch of n poses uses, in order: the value forced by lc2e_pair_set_tuning
 * (0 = none), the value set or tuned for batches of exactly n poses
 * (lc2e_pair_set_tuning_n, lc2e_pair_autotune; 0 = none), or a model of the launch
 * (work-group rounds over the CUs times the planes each work-group marches).
 * get_tuning reports the forced value, else the one the last launch used.
 * No counterpart in the reference (its launch shape is fixed at 8x8x8). */
int lc2e_pair_set_tuning(lc2e_pair* p, int zchunk);
int lc2e_pair_set_tuning_n(lc2e_pair* p, int n, int zchunk);
int lc2e_pair_get_tuning(lc2e_pair* p, int* zchunk);
/* Autotuner: times the fused kernel over candidate z-chunks (HIP events, `iters` launches
 * each after one warm-up) for batches of n poses `mats`, records the fastest for batches
 * of n and returns it. A z-chunk forced with lc2e_pair_set_tuning stays in force (and
 * takes precedence) until it is cleared with 0. 2D pairs have nothing to tune and
 * return 1. No counterpart in the reference. */
int lc2e_pair_autotune(lc2e_pair* p, int n, const float* mats, int iters, int* best_zchunk,
                       double* best_ms);

/* Kernel timing with HIP events recorded around every fused-kernel launch on the engine's
 * stream. While timing is on, batches are launched eagerly even where graph replay is on
 * (the events bracket the kernel itself); replay resumes when it is turned off. enable = 1
 * resets the counters. read returns the summed fused-kernel time in ms and the number of
 * launches since the reset. */
int lc2e_pair_profile(lc2e_pair* p, int enable);
int lc2e_pair_profile_read(lc2e_pair* p, double* kernel_ms, long long* launches);

/* Record the batched evaluation of n poses (matrix upload, kernels, result download) in a
 * HIP graph on its second lc2e_eval_batch with n poses and replay it for later calls with
 * the same n (1 = on, 0 = off, the default). Graphs are dropped when the pair's buffers or
 * launch shape change. Results are identical either way. On MI355X with ROCm 7.2 a
 * replayed batch took 5-13 us longer than an eager one at 1 to 64 poses, so nothing in
 * this library turns it on by itself (the facade does with LC2E_GRAPHS=1). */
int lc2e_pair_use_graph(lc2e_pair* p, int enable);

#ifdef __cplusplus
}
#endif
#endif
