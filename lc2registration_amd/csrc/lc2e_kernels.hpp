// Internal interface between the engine (lc2e_engine.cpp) and the device code
// (lc2e_kernels.hip). Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace lc2e {

constexpr int kTile2X = 16;    // output tile of the 2D fused kernel: kTile2X x kTile2Y,
constexpr int kTile2Y = 16;    // one thread per output
constexpr int kThreads = kTile2X * kTile2Y;  // work-group size of the 2D fused kernel
constexpr int kTile3X = 16;    // output tile of the 3D fused kernel: 16 x 16 columns,
constexpr int kTile3Y = 16;    // one work-group of 256 threads (two per CU), one column each
// Resident volumes carry this many zeroed elements past their last voxel: the pipelined
// samplers read the voxel pair (x0, x0 + 1) of rows y0 and y0 + 1 in planes z0 and z0 + 1,
// and send samples outside the source to the first tail element. The farthest element
// read is X*Y + X + 1 past the last voxel's successor; 64 more keep the end aligned.
constexpr long long volume_pad_elems(long long X, long long Y) { return X * Y + X + 64; }
constexpr int kMaxR3 = 4;      // largest patch radius of the 3D fused kernel
constexpr int kZBlock = 8;     // 3D partial sums per block of 8 output planes; z-chunks are multiples
constexpr int kXcds = 8;       // MI355X: 8 XCDs, each with its own L2, 32 CUs each
constexpr int kMaxR2 = 10;     // largest patch radius of the 2D fused kernel
constexpr int kMaxRSep = 64;   // largest patch radius of the separable path (larger radii)
constexpr int kFrameRun = 16;  // frames per 2D work-group (frame stacks)
// The fused kernels' samplers form voxel offsets with 24-bit multiplies: a 3D source whose
// slices hold this many voxels or more is evaluated by the separable path instead.
constexpr long long kMaxFusedSlice = 1LL << 24;

struct VGeom {  // virtual grid
    int X, Y, Z;
    float sx, sy, sz, ox, oy, oz;
};

struct SGeom {  // a resident image
    int X, Y, Z;
    float sx, sy, sz, ox, oy, oz;
    // 1/spacing; used only when every spacing is a power of two, where (p - o) * (1/s)
    // equals (p - o) / s exactly
    float rsx, rsy, rsz;
    int pow2;
};

inline SGeom make_sgeom(const int d[3], const float s[3], const float o[3]) {
    SGeom g{d[0], d[1], d[2], s[0], s[1], s[2], o[0], o[1], o[2], 1.0f / s[0], 1.0f / s[1],
            1.0f / s[2], 1};
    for (int a = 0; a < 3; ++a) {
        int e;
        const float m = std::frexp(s[a], &e);
        if (m != 0.5f) g.pow2 = 0;
    }
    return g;
}

struct FusedArgs {
    VGeom v;
    SGeom s;                 // moving image
    const void* src;         // moving image data (float or half)
    const float* stat_img;   // fixed image resampled on the virtual grid
    const float* stat_grad;  // |Sobel| of stat_img when it is the high-resolution image
    const double* ustat_mean;  // static ultrasound: patch mean per voxel
    const double* ustat_cyy;   // static ultrasound: centred patch energy, < 0 where w = 0
    const float* mats;       // [pose][mats_per_pose][12] VirtualToMoving, 3x4 row-major
    double* partials;        // [pose][blocks_per_pose][2] (sum measure, sum w); 3D: the
                             // blocks are [z block of kZBlock planes][tile column][wave]
    double* wmap;            // optional per-voxel w (single pose), may be null
    double* mmap;            // optional per-voxel measure, may be null
    int r;
    int zchunk;              // 3D: output planes per work-group
    int ntx, nty, nzc;       // tiles along x, y and z chunks (3D) / frames (2D)
    int blocks_per_pose;     // partial sums per pose
    int tiles;               // work-groups per pose
    int mats_per_pose;
    int src_f16;             // moving image stored as half
    int npose;               // poses in this launch (set by launch_fused)
    int prun;                // 2D single images: poses per work-group (0 or 1: one)
    // separable path (patch radii beyond the fused kernels): the moving image and its
    // |Sobel| on the grid (V floats each) and two [channels][V] f64 box-sum buffers
    float* sep_img;
    float* sep_grad;
    double* sep_a;
    double* sep_b;
};

// Number of work-groups one pose needs (0 if the grid has no valid patch centre).
int fused_blocks_3d(const VGeom& v, int r, int zchunk, int* ntx, int* nty, int* nzc);
// Partial sums per pose of the 3D kernel (tile columns x blocks of kZBlock planes x the
// 4 waves of a work-group).
int fused_partials_3d(const VGeom& v, int r);
int fused_blocks_2d(const VGeom& v, int r, int* ntx, int* nty);

hipError_t launch_fused(const FusedArgs& a, int grid_ndim, int src_ndim, int high_res_is_moving,
                        int n_poses, hipStream_t st);
// Poses per work-group of the 2D kernel on a single image of a 2D source (1 otherwise):
// launch_fused runs several poses of one tile in one work-group while the launch keeps
// enough work-groups.
int pose_run(const VGeom& v, int grid_ndim, int src_ndim, int mats_per_pose, int tiles,
             int n_poses);
// Work-groups of the fused kernel a launch with these arguments would use that fit on one
// CU at once (hipOccupancyMaxActiveBlocksPerMultiprocessor).
hipError_t fused_resident_per_cu(const FusedArgs& a, int grid_ndim, int src_ndim, int movi,
                                 int* per_cu);
hipError_t launch_reduce(const double* partials, int blocks_per_pose, int n_poses, double* out,
                         hipStream_t st);
// Separable path for radii beyond the fused kernels (3D r > kMaxR3, 2D r > kMaxR2): per
// pose, warp (+ |Sobel|), box sums along each axis through sep_a / sep_b, combine; one
// partial per 256 output voxels (sep_blocks per pose, fixed positions).
int sep_blocks(const VGeom& v, int r, int grid_ndim);
hipError_t launch_sep(const FusedArgs& a, int grid_ndim, int src_ndim, int movi, int n_poses,
                      hipStream_t st);
constexpr int kSepChannels(bool movi) { return movi ? 7 : 10; }
// Device scratch of the allocation-free batch entry point (lc2e_eval_batch_noalloc): one
// file-scope device array of kScratchBytes per device, holding a launch's matrices,
// partial sums and results. scratch_address() returns its address on `device` (looked up
// once per device with that device current, then cached).
constexpr size_t kScratchBytes = size_t(16) << 20;
hipError_t scratch_address(int device, void** out);
// Fixed image -> virtual grid at identity (frames of a 2D stack one by one).
hipError_t launch_warp_static(const void* fixed, int fixed_f16, const SGeom& fg, const VGeom& v,
                              int grid_ndim, int src_ndim, float* out, hipStream_t st);
// Moving image -> virtual grid at one pose (debug stages).
hipError_t launch_warp_moving(const FusedArgs& a, int grid_ndim, int src_ndim, float* out,
                              hipStream_t st);
// Patch statistics of a static ultrasound grid: mean and centred energy per voxel
// (negative where the reference's weight is 0), FP64, fixed summation order.
hipError_t launch_ustats(const float* img, const VGeom& v, int grid_ndim, int r, double* mean,
                         double* cyy, hipStream_t st);
// |Sobel| of a virtual-grid image (3D, or per frame for 2D grids), separable order.
hipError_t launch_sobel(const float* img, const VGeom& v, int grid_ndim, float* out,
                        hipStream_t st);
// Normalized mutual information per pose over the virtual grid (the patchSize 0 metric):
// extremes (the static image's once, stat_ready = 0 computes them into ext_stat), the
// joint histogram [pose][bins_i][bins_u] and FP64 entropies into out[pose].
// ext_mov: 2 ints per pose; hist: n_poses * bins_u * bins_i counters.
constexpr int kMiMaxBins = 256;
hipError_t launch_mi(const FusedArgs& a, int grid_ndim, int src_ndim, int high_res_is_moving,
                     int n_poses, int bins_u, int bins_i, const float* stat, int* ext_stat,
                     int stat_ready, int* ext_mov, unsigned* hist, double* out, hipStream_t st);
// Host bytes -> resident storage (u8/i16/f16/f32 -> f16/f32).
hipError_t launch_convert(const void* d_in, int in_dtype, void* d_out, int out_f16, long long n,
                          hipStream_t st);

}  // namespace lc2e
