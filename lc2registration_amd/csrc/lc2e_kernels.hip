// Device code of the LC2 engine for gfx950 (MI355X).
//
// One fused kernel per pose batch computes, for every patch centre of the
// virtual grid, the reference's LC2 measure and weight
// (cu_lc2_3d.cu:181-618 / cu_lc2_2d.cu:88-480) and reduces them per work-group:
//   warp of the moving image (interpolator.cu:69-232) -> |Sobel| (cu_lc2_3d.cu:100-179)
//   -> FP64 patch moments -> closed-form least-squares projection -> measure, w.
// The reference's per-voxel FP32 conjugate gradient (cu_lc2_3d.cu:289-524) is
// replaced by the projection it converges to, computed from centred moments:
//   var_res * (N-1) = C_yy - s^T C^+ s,  s = (C_yI, C_yG),  C = [[C_II, C_IG], [C_IG, C_GG]].
//
// Numerics: the FP32 stages (coordinates, trilinear/bilinear lerps, Sobel) are
// written with explicit fmaf and the file is compiled with -ffp-contract=off, so
// they reproduce the oracle bit for bit. Patch moments are FP64 (products of two
// FP32 values are exact in FP64), which removes the cancellation an FP32
// sliding-window sum would suffer on intensities with small local variance.
//
// 3D structure (k_fused3d): a 256-thread work-group (two per CU) owns a 16x16 column
// of output voxels over `zchunk` planes and marches along z. Per input plane:
// warp the plane (+ Sobel halo) into LDS, separable Sobel of each 3x3 neighbourhood
// with a 3-plane register ring along z, FP64 products into LDS, x then y window sums in
// LDS (van Herk / Gil-Werman: no partial sum holds a value from outside its window), and
// a (2R+1)-slot register ring of plane sums and pair sums of planes gives the z window
// sum; complete patch sums are combined in FP64 and accumulated per thread. No
// intermediate volume touches HBM. Patch radii beyond the fused kernels run the
// separable path (k_sep_pass / k_sep_combine) through HBM.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>

#include "lc2e_kernels.hpp"

namespace lc2e {

namespace {

constexpr double kZeroThreshold = 1e-6;  // cu_lc2_helper.cuh:5
constexpr double kTauConst = 1e-10;      // centred energy below this fraction of raw -> constant
constexpr double kTauColl = 1e-10;       // Gram determinant below this fraction -> collinear

// Compile-time loop: calls f(integral_constant<int, K>) for K = B .. E-1.
template <int B, int E>
struct Unroll {
    template <class F>
    __device__ __forceinline__ static void run(F&& f) {
        f(std::integral_constant<int, B>{});
        Unroll<B + 1, E>::run(f);
    }
};
template <int E>
struct Unroll<E, E> {
    template <class F>
    __device__ __forceinline__ static void run(F&&) {}
};

template <typename T>
__device__ __forceinline__ float ld1(const T* p, int i) {
    if constexpr (sizeof(T) == 2) return __half2float(p[i]);
    else return p[i];
}

// The voxels at x0 and x0 + 1 of one row in a single load. Volumes are allocated with a
// zeroed tail, so x0 + 1 is readable even for the last voxel; when the sample lies on
// the integer x0 the reference reads x0 twice (ceil == floor), which is restored here.
template <typename T>
__device__ __forceinline__ void ld2(const T* p, int i, bool same, float& a, float& b) {
    if constexpr (sizeof(T) == 2) {
        __half2 h;
        __builtin_memcpy(&h, p + i, sizeof(h));
        a = __low2float(h);
        b = __high2float(h);
    } else {
        float2 f;
        __builtin_memcpy(&f, p + i, sizeof(f));
        a = f.x;
        b = f.y;
    }
    if (same) b = a;
}

// Voxel coordinates of a position in mm: (p - origin) / spacing with one rounding, as the
// reference's ToPxl computes it (spatialMappers.cuh:143-169). The fused kernels are
// instantiated per form, so their samplers carry no branch for it: P2 (every spacing a
// power of two) multiplies by the exact 1/s; otherwise the IEEE division. (A division by a
// constant as fmaf(a, RN(1/s), RN(a * zl)), exact for a checked spacing, took more time
// on the CT workload than the division it replaced, once its guard for tiny numerators was
// in: 12.05 vs 11.20 ms per launch, measured.)
template <bool P2>
__device__ __forceinline__ float to_voxel1(float p, float o, float sp, float rs) {
    if (P2) return (p - o) * rs;
    return (p - o) / sp;
}

template <bool P2>
__device__ __forceinline__ void to_voxel2(const SGeom& s, float px, float py, float& vx,
                                          float& vy) {
    vx = to_voxel1<P2>(px, s.ox, s.sx, s.rsx);
    vy = to_voxel1<P2>(py, s.oy, s.sy, s.rsy);
}

template <bool P2>
__device__ __forceinline__ void to_voxel(const SGeom& s, float px, float py, float pz, float& vx,
                                         float& vy, float& vz) {
    to_voxel2<P2>(s, px, py, vx, vy);
    vz = to_voxel1<P2>(pz, s.oz, s.sz, s.rsz);
}

// The same with the form chosen at run time (the one-off samplers).
__device__ __forceinline__ void to_voxel_rt(const SGeom& s, float px, float py, float pz,
                                            float& vx, float& vy, float& vz) {
    if (s.pow2) to_voxel<true>(s, px, py, pz, vx, vy, vz);
    else to_voxel<false>(s, px, py, pz, vx, vy, vz);
}
__device__ __forceinline__ void to_voxel2_rt(const SGeom& s, float px, float py, float& vx,
                                             float& vy) {
    if (s.pow2) to_voxel2<true>(s, px, py, vx, vy);
    else to_voxel2<false>(s, px, py, vx, vy);
}

// round(b - a) in f32 for the low / high halves of two half2 words: v_fma_mix_f32 with
// f16 sources b and -a and the f32 constant 1.0, i.e. fma(b, 1, -a) with one rounding.
__device__ __forceinline__ float hsub_lo(unsigned b, unsigned a) {
    float r;
    asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(b), "v"(a));
    return r;
}
__device__ __forceinline__ float hsub_hi(unsigned b, unsigned a) {
    float r;
    asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel:[1,0,1] op_sel_hi:[1,0,1]"
        : "=v"(r) : "v"(b), "v"(a));
    return r;
}
__device__ __forceinline__ float hlo(unsigned w) {
    __half2 h;
    __builtin_memcpy(&h, &w, sizeof(h));
    return __low2float(h);
}
__device__ __forceinline__ float hhi(unsigned w) {
    __half2 h;
    __builtin_memcpy(&h, &w, sizeof(h));
    return __high2float(h);
}

// Interpolate_3Dkernel sample (interpolator.cu:76-174): L -> M*L -> voxel -> trilinear
// z, y, x with floor/ceil corners; 0 outside [0, dim-1].
template <typename T>
__device__ __forceinline__ float sample3(const T* __restrict__ src, const SGeom& s, const float* m,
                                         float lx, float ly, float lz) {
    const float px = fmaf(m[2], lz, fmaf(m[1], ly, m[0] * lx)) + m[3];
    const float py = fmaf(m[6], lz, fmaf(m[5], ly, m[4] * lx)) + m[7];
    const float pz = fmaf(m[10], lz, fmaf(m[9], ly, m[8] * lx)) + m[11];
    float vx, vy, vz;
    to_voxel_rt(s, px, py, pz, vx, vy, vz);
    if (!(vx >= 0.0f && vx <= (float)(s.X - 1) && vy >= 0.0f && vy <= (float)(s.Y - 1) &&
          vz >= 0.0f && vz <= (float)(s.Z - 1)))
        return 0.0f;
    const float fx = floorf(vx), fy = floorf(vy), fz = floorf(vz);
    const int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
    const int y1 = (int)ceilf(vy), z1 = (int)ceilf(vz);
    const bool xs = (fx == vx);
    const float rx = vx - fx, ry = vy - fy, rz = vz - fz;
    const int sxy = s.X * s.Y;
    const int r00 = y0 * s.X + z0 * sxy, r01 = y0 * s.X + z1 * sxy;
    const int r10 = y1 * s.X + z0 * sxy, r11 = y1 * s.X + z1 * sxy;
    if constexpr (sizeof(T) == 2) {
        // half2 words of the two x neighbours; on an integer x (xs) rx == 0 selects x0
        const char* base = reinterpret_cast<const char*>(src);
        unsigned wa, wb, wc, wd;
        __builtin_memcpy(&wa, base + (size_t)(x0 + r00) * 2, 4);
        __builtin_memcpy(&wb, base + (size_t)(x0 + r01) * 2, 4);
        __builtin_memcpy(&wc, base + (size_t)(x0 + r10) * 2, 4);
        __builtin_memcpy(&wd, base + (size_t)(x0 + r11) * 2, 4);
        const float lu = fmaf(rz, hsub_lo(wb, wa), hlo(wa));
        const float ru = fmaf(rz, hsub_hi(wb, wa), hhi(wa));
        const float lb = fmaf(rz, hsub_lo(wd, wc), hlo(wc));
        const float rb = fmaf(rz, hsub_hi(wd, wc), hhi(wc));
        const float yl = fmaf(ry, lb - lu, lu);
        const float yr = fmaf(ry, rb - ru, ru);
        return fmaf(rx, yr - yl, yl);
    }
    float a0, a1, b0, b1;
    ld2(src, x0 + r00, xs, a0, a1);  // (x0, y0, z0), (x1, y0, z0)
    ld2(src, x0 + r01, xs, b0, b1);  // (x0, y0, z1), (x1, y0, z1)
    const float lu = fmaf(rz, b0 - a0, a0);
    const float ru = fmaf(rz, b1 - a1, a1);
    ld2(src, x0 + r10, xs, a0, a1);  // (x0, y1, z0), (x1, y1, z0)
    ld2(src, x0 + r11, xs, b0, b1);  // (x0, y1, z1), (x1, y1, z1)
    const float lb = fmaf(rz, b0 - a0, a0);
    const float rb = fmaf(rz, b1 - a1, a1);
    const float yl = fmaf(ry, lb - lu, lu);
    const float yr = fmaf(ry, rb - ru, ru);
    return fmaf(rx, yr - yl, yl);
}

// The same trilinear sample split in two for software pipelining: issue3() computes the
// voxel coordinates and starts the four paired loads, finish3() does the lerps once the
// data has arrived, so a whole z-step of work can run in between. Nothing but the loaded
// values and the fractions is carried across the step (no per-lane flags):
//  * a sample outside the source, or a position outside the grid (inside == false),
//    loads from the volume's zeroed tail, so its lerps give exactly 0;
//  * on an integer x the reference reads x0 twice (ceil == floor); here rx == 0 then, and
//    fmaf(0, yr - yl, yl) == yl for the finite values a resident volume holds (f16
//    storage rejects values beyond its range), up to the sign of a zero.
// Inside the volume the clamps change nothing: the result is the reference's.
// f32 sources keep the eight corner values; half sources keep the four loaded half2 words
// as they are, and the lerps read them through mixed-precision FMAs (exactly the f32
// arithmetic on the converted values, without separate conversions).
template <typename T>
struct Pending3 {
    float a0, a1, b0, b1, c0, c1, d0, d1;
    float rx, ry, rz;
};
template <>
struct Pending3<__half> {
    unsigned a, b, c, d;  // half2: x0 in the low half, x0 + 1 in the high half
    float rx, ry, rz;
};


template <typename T>
__device__ __forceinline__ void ld2_off(const T* __restrict__ base, unsigned idx, float& a,
                                        float& b) {
    // 32-bit element offset from a uniform base: one saddr + voffset load, no 64-bit math
    const char* p = reinterpret_cast<const char*>(base) + idx * (unsigned)sizeof(T);
    if constexpr (sizeof(T) == 2) {
        __half2 h;
        __builtin_memcpy(&h, p, sizeof(h));
        a = __low2float(h);
        b = __high2float(h);
    } else {
        float2 f;
        __builtin_memcpy(&f, p, sizeof(f));
        a = f.x;
        b = f.y;
    }
}

template <bool P2, typename T>
__device__ __forceinline__ void issue3v(const T* __restrict__ src, const SGeom& s, float px,
                                        float py, float pz, bool inside, Pending3<T>& p);

template <bool P2, typename T>
__device__ __forceinline__ void issue3(const T* __restrict__ src, const SGeom& s, const float* m,
                                       float lx, float ly, float lz, bool inside, Pending3<T>& p) {
    const float px = fmaf(m[2], lz, fmaf(m[1], ly, m[0] * lx)) + m[3];
    const float py = fmaf(m[6], lz, fmaf(m[5], ly, m[4] * lx)) + m[7];
    const float pz = fmaf(m[10], lz, fmaf(m[9], ly, m[8] * lx)) + m[11];
    issue3v<P2>(src, s, px, py, pz, inside, p);
}

// v clamped to [0, hi] in one instruction (fminf / fmaxf cost four: IEEE mode makes each
// quieten its inputs first). The coordinates here come from arithmetic, so a NaN is quiet,
// and v_med3_f32 then returns the minimum of the other two operands, 0.
__device__ __forceinline__ float clamp_med3(float v, float hi) {
    return __builtin_amdgcn_fmed3f(v, 0.0f, hi);
}

// The paired loads of the pipelined samplers always take the rows y0 + 1 and planes
// z0 + 1, where the reference takes ceil(v) (equal to v on an integer coordinate). On an
// integer coordinate the fraction is exactly 0 there, so fmaf(0, b - a, a) == a for the
// finite values a resident volume holds (uploads reject non-finite voxels), up to the sign
// of a zero: the reference's value. One row / plane past the last one lands in the zeroed
// tail (volume_pad_elems), as do all four corners of a sample outside the source. So a
// sample needs one 32-bit element offset: the four corners are two byte offsets (plane z0
// and z0 + 1) from two uniform bases (rows y0 and y0 + 1), with no per-corner arithmetic.
template <typename T>
__device__ __forceinline__ void corners_off(const T* __restrict__ src, const SGeom& s, unsigned e0,
                                            unsigned plane, unsigned& o0, unsigned& o1,
                                            const char*& b0, const char*& b1) {
    o0 = e0 * (unsigned)sizeof(T);
    o1 = o0 + plane * (unsigned)sizeof(T);
    b0 = reinterpret_cast<const char*>(src);
    b1 = b0 + (size_t)s.X * sizeof(T);
}

// issue3 from the source position (px, py, pz) in mm.
template <bool P2, typename T>
__device__ __forceinline__ void issue3v(const T* __restrict__ src, const SGeom& s, float px,
                                        float py, float pz, bool inside, Pending3<T>& p) {
    float vx, vy, vz;
    to_voxel<P2>(s, px, py, pz, vx, vy, vz);
    const float xm = (float)(s.X - 1), ym = (float)(s.Y - 1), zm = (float)(s.Z - 1);
    // clamped coordinates (one v_med3_f32 each; with a quiet-NaN input it returns the
    // smaller bound, 0); a coordinate lies in [0, dim - 1] exactly when the clamp leaves it
    // unchanged, which NaN never is
    const float cx = clamp_med3(vx, xm), cy = clamp_med3(vy, ym), cz = clamp_med3(vz, zm);
    const bool ok = inside && cx == vx && cy == vy && cz == vz;
    // on c >= 0: conversion truncates to floor(c), and v_fract_f32 gives c - floor(c) exactly
    p.rx = __builtin_amdgcn_fractf(cx);
    p.ry = __builtin_amdgcn_fractf(cy);
    p.rz = __builtin_amdgcn_fractf(cz);
    const unsigned X = (unsigned)s.X, sxy = (unsigned)(s.X * s.Y);
    const unsigned tail = sxy * (unsigned)s.Z;  // first element of the zeroed tail
    unsigned e0 = tail;
    if (ok) {
        const unsigned x0 = (unsigned)cx, y0 = (unsigned)cy, z0 = (unsigned)cz;
        // 24-bit multiply-adds (full rate; a 32-bit multiply is quarter rate): dimensions
        // are below 2^24, and so is X*Y for every source the fused kernels sample (pairs
        // with larger slices take the separable path)
        e0 = __umul24(z0, sxy) + (__umul24(y0, X) + x0);
    }
    unsigned o0, o1;
    const char *b0, *b1;
    corners_off(src, s, e0, sxy, o0, o1, b0, b1);
    if constexpr (sizeof(T) == 2) {
        __builtin_memcpy(&p.a, b0 + o0, 4);  // (y0, z0)
        __builtin_memcpy(&p.b, b0 + o1, 4);  // (y0, z0 + 1)
        __builtin_memcpy(&p.c, b1 + o0, 4);  // (y0 + 1, z0)
        __builtin_memcpy(&p.d, b1 + o1, 4);  // (y0 + 1, z0 + 1)
    } else {
        float2 f;
        __builtin_memcpy(&f, b0 + o0, 8);
        p.a0 = f.x, p.a1 = f.y;
        __builtin_memcpy(&f, b0 + o1, 8);
        p.b0 = f.x, p.b1 = f.y;
        __builtin_memcpy(&f, b1 + o0, 8);
        p.c0 = f.x, p.c1 = f.y;
        __builtin_memcpy(&f, b1 + o1, 8);
        p.d0 = f.x, p.d1 = f.y;
    }
}

// issue3 with the in-plane part of the source coordinates given: c = m*(lx, ly, 0) as
// issue3 forms it, so px = fmaf(m[2], lz, c[0]) + m[3] is issue3's px bit for bit.
template <bool P2, typename T>
__device__ __forceinline__ void issue3c(const T* __restrict__ src, const SGeom& s, const float* m,
                                        const float* c, float lz, bool inside, Pending3<T>& p) {
    const float px = fmaf(m[2], lz, c[0]) + m[3];
    const float py = fmaf(m[6], lz, c[1]) + m[7];
    const float pz = fmaf(m[10], lz, c[2]) + m[11];
    issue3v<P2>(src, s, px, py, pz, inside, p);
}

template <typename T>
__device__ __forceinline__ float finish3(const Pending3<T>& p) {
    float lu, ru, lb, rb;
    if constexpr (sizeof(T) == 2) {
        lu = fmaf(p.rz, hsub_lo(p.b, p.a), hlo(p.a));
        ru = fmaf(p.rz, hsub_hi(p.b, p.a), hhi(p.a));
        lb = fmaf(p.rz, hsub_lo(p.d, p.c), hlo(p.c));
        rb = fmaf(p.rz, hsub_hi(p.d, p.c), hhi(p.c));
    } else {
        lu = fmaf(p.rz, p.b0 - p.a0, p.a0);
        ru = fmaf(p.rz, p.b1 - p.a1, p.a1);
        lb = fmaf(p.rz, p.d0 - p.c0, p.c0);
        rb = fmaf(p.rz, p.d1 - p.c1, p.c1);
    }
    const float yl = fmaf(p.ry, lb - lu, lu);
    const float yr = fmaf(p.ry, rb - ru, ru);
    return fmaf(p.rx, yr - yl, yl);
}

// Interpolate_2Dkernel sample (interpolator.cu:186-230): bilinear, y then x. The 2x3
// matrix sits in the 3x4 layout as m[0], m[1], m[3] / m[4], m[5], m[7].
template <typename T>
__device__ __forceinline__ float sample2(const T* __restrict__ src, const SGeom& s, const float* m,
                                         float lx, float ly) {
    const float px = fmaf(m[1], ly, m[0] * lx) + m[3];
    const float py = fmaf(m[5], ly, m[4] * lx) + m[7];
    float vx, vy;
    to_voxel2_rt(s, px, py, vx, vy);
    if (!(vx >= 0.0f && vx <= (float)(s.X - 1) && vy >= 0.0f && vy <= (float)(s.Y - 1)))
        return 0.0f;
    const float fx = floorf(vx), fy = floorf(vy);
    const int x0 = (int)fx, y0 = (int)fy, y1 = (int)ceilf(vy);
    const bool xs = (fx == vx);
    const float ry = vy - fy;
    if constexpr (sizeof(T) == 2) {
        const char* base = reinterpret_cast<const char*>(src);
        unsigned wa, wb;
        __builtin_memcpy(&wa, base + (size_t)(x0 + y0 * s.X) * 2, 4);
        __builtin_memcpy(&wb, base + (size_t)(x0 + y1 * s.X) * 2, 4);
        const float xl = fmaf(ry, hsub_lo(wb, wa), hlo(wa));
        const float xr = fmaf(ry, hsub_hi(wb, wa), hhi(wa));
        return fmaf(vx - fx, xr - xl, xl);  // 0 * (xr - xl) on an integer x
    }
    float a0, a1, b0, b1;
    ld2(src, x0 + y0 * s.X, xs, a0, a1);
    ld2(src, x0 + y1 * s.X, xs, b0, b1);
    const float xl = fmaf(ry, b0 - a0, a0);
    const float xr = fmaf(ry, b1 - a1, a1);
    const float rx = vx - fx;
    return fmaf(rx, xr - xl, xl);
}

// sample2 split like issue3 / finish3: coordinates and the two paired loads now, the
// lerps later. Same flag-free rules: outside the source (or inside == false) the loads
// read the zeroed tail; on an integer x, rx == 0 selects x0.
template <typename T>
struct Pending2 {
    float a0, a1, b0, b1;
    float rx, ry;
};
template <>
struct Pending2<__half> {
    unsigned a, b;  // half2 words at (x0, y0) and (x0, y1)
    float rx, ry;
};

template <bool P2, typename T>
__device__ __forceinline__ void issue2(const T* __restrict__ src, const SGeom& s, const float* m,
                                       float lx, float ly, bool inside, Pending2<T>& p) {
    const float px = fmaf(m[1], ly, m[0] * lx) + m[3];
    const float py = fmaf(m[5], ly, m[4] * lx) + m[7];
    float vx, vy;
    to_voxel2<P2>(s, px, py, vx, vy);
    const float xm = (float)(s.X - 1), ym = (float)(s.Y - 1);
    const float cx = clamp_med3(vx, xm), cy = clamp_med3(vy, ym);
    const bool ok = inside && cx == vx && cy == vy;
    p.rx = __builtin_amdgcn_fractf(cx);  // as in issue3v
    p.ry = __builtin_amdgcn_fractf(cy);
    const unsigned X = (unsigned)s.X;
    const unsigned tail = X * (unsigned)s.Y * (unsigned)s.Z;  // first element of the zeroed tail
    // rows y0 and y0 + 1 (see corners_off): one offset from two uniform bases
    const unsigned o =
        (ok ? __umul24((unsigned)cy, X) + (unsigned)cx : tail) * (unsigned)sizeof(T);
    const char* b0 = reinterpret_cast<const char*>(src);
    const char* b1 = b0 + (size_t)s.X * sizeof(T);
    if constexpr (sizeof(T) == 2) {
        __builtin_memcpy(&p.a, b0 + o, 4);
        __builtin_memcpy(&p.b, b1 + o, 4);
    } else {
        float2 f;
        __builtin_memcpy(&f, b0 + o, 8);
        p.a0 = f.x, p.a1 = f.y;
        __builtin_memcpy(&f, b1 + o, 8);
        p.b0 = f.x, p.b1 = f.y;
    }
}

template <typename T>
__device__ __forceinline__ float finish2(const Pending2<T>& p) {
    float xl, xr;
    if constexpr (sizeof(T) == 2) {
        xl = fmaf(p.ry, hsub_lo(p.b, p.a), hlo(p.a));
        xr = fmaf(p.ry, hsub_hi(p.b, p.a), hhi(p.a));
    } else {
        xl = fmaf(p.ry, p.b0 - p.a0, p.a0);
        xr = fmaf(p.ry, p.b1 - p.a1, p.a1);
    }
    return fmaf(p.rx, xr - xl, xl);
}

// FP64 reciprocal and reciprocal square root: hardware estimate + two Newton steps
// (relative error a few ulp), instead of the long IEEE division / sqrt sequences.
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

__device__ __forceinline__ double rsq64(double x) {
    double r = __builtin_amdgcn_rsq(x);
    r = r * fma(-0.5 * x * r, r, 1.5);
    r = r * fma(-0.5 * x * r, r, 1.5);
    return r;
}

// s^T C^+ s for the two-column centred fit. Rank rule: a column whose centred energy is
// at most kTauConst of its raw energy is dropped; two columns with det <= kTauColl*a*b
// are collinear and U is projected on the leading singular direction of the centred,
// unit-normalised columns, I' + sign(rho) G' (of squared norm 2 + 2|rho|): least squares
// by truncated SVD with the cutoff sigma_min/sigma_max <= sqrt(kTauColl)/2.
__device__ __forceinline__ double explained_var(double a, double b, double c, double p, double q,
                                                double rawa, double rawb) {
    const bool ikeep = a > kTauConst * rawa;
    const bool gkeep = b > kTauConst * rawb;
    if (ikeep && gkeep) {
        const double det = a * b - c * c;
        if (det > kTauColl * a * b)
            return (b * p * p - 2.0 * c * p * q + a * q * q) * rcp64(det);
        // reciprocal square roots through the refined hardware estimates (a few ulp)
        const double ra = rsq64(a), rb = rsq64(b);
        const double t = p * ra + (c < 0.0 ? -1.0 : 1.0) * q * rb;
        return (t * t) * rcp64(2.0 + 2.0 * fabs(c) * ra * rb);
    }
    if (ikeep) return p * p * rcp64(a);
    if (gkeep) return q * q * rcp64(b);
    return 0.0;
}

// Weight and measure from the centred energy of U (cyy) and the explained part e
// (cu_lc2_3d.cu:258-284 weight, :596-611 measure): w = sqrt(cyy/(N-1)),
// measure = w - var_res / w with var_res = (cyy - e)/(N-1); 0 if not > 0.
__device__ __forceinline__ void weight_measure(double cyy, double e, double inv_n1, double& w,
                                               double& meas) {
    const double var_u = cyy * inv_n1;
    const double rw = rsq64(var_u);  // 1 / w
    w = var_u * rw;
    e = fmin(fmax(e, 0.0), cyy);
    meas = w - ((cyy - e) * inv_n1) * rw;
    if (!(meas > 0.0) || isinf(meas)) meas = 0.0;
}

// Channels I, G, U, II, GG, IG, UU, UI, UG and the count of U > 1e-6. false where w = 0.
__device__ __forceinline__ bool combine(const double* S, int cnt, int N, double& w_out,
                                        double& m_out) {
    if (cnt != N) return false;
    const double inv_n = 1.0 / (double)N, inv_n1 = 1.0 / (double)(N - 1);
    const double sy = S[2];
    const double cyy = S[6] - sy * sy * inv_n;
    if (!(cyy * inv_n1 > kZeroThreshold)) return false;
    const double si = S[0], sg = S[1];
    const double cii = S[3] - si * si * inv_n;
    const double cgg = S[4] - sg * sg * inv_n;
    const double cig = S[5] - si * sg * inv_n;
    const double cyi = S[7] - sy * si * inv_n;
    const double cyg = S[8] - sy * sg * inv_n;
    const double e = explained_var(cii, cgg, cig, cyi, cyg, S[3], S[4]);
    weight_measure(cyy, e, inv_n1, w_out, m_out);
    return true;
}

// Static ultrasound: channels I, G, II, GG, IG, UI, UG; mu = patch mean of U,
// cyy = its centred energy (< 0 where w = 0).
__device__ __forceinline__ bool combine_static_u(const double* S, double mu, double cyy, int N,
                                                 double& w_out, double& m_out) {
    if (!(cyy >= 0.0)) return false;
    const double inv_n = 1.0 / (double)N, inv_n1 = 1.0 / (double)(N - 1);
    const double si = S[0], sg = S[1];
    const double cii = S[2] - si * si * inv_n;
    const double cgg = S[3] - sg * sg * inv_n;
    const double cig = S[4] - si * sg * inv_n;
    const double cyi = S[5] - mu * si;
    const double cyg = S[6] - mu * sg;
    const double e = explained_var(cii, cgg, cig, cyi, cyg, S[2], S[3]);
    weight_measure(cyy, e, inv_n1, w_out, m_out);
    return true;
}

// Work-group -> (tile, pose) on a 1D grid of roundup(tiles, 8) x npose work-groups.
// Work-groups go to the 8 XCDs round-robin by linear id; this mapping keeps all poses of
// one tile on the same XCD and adjacent in time (32 poses fill that XCD's 32 CUs), so the
// tile's static data (ultrasound statistics, fixed image) and the overlapping source
// footprint are fetched into that XCD's L2 once and shared by every pose. Returns false
// for the padding work-groups (tile >= tiles), which exit at once.
__device__ __forceinline__ bool map_block(int tiles, int npose, int& tile, int& pose) {
    const int w = blockIdx.x;
    const int xcd = w % kXcds, slot = w / kXcds;
    pose = slot % npose;
    tile = (slot / npose) * kXcds + xcd;
    return tile < tiles;
}

template <int NTHR>
__device__ __forceinline__ void block_reduce_store(double am, double aw, double* s_red,
                                                   double* dst) {
    for (int off = 32; off > 0; off >>= 1) {
        am += __shfl_down(am, off, 64);
        aw += __shfl_down(aw, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        s_red[2 * wid] = am;
        s_red[2 * wid + 1] = aw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0, w = 0.0;
        for (int i = 0; i < NTHR / 64; ++i) {
            m += s_red[2 * i];
            w += s_red[2 * i + 1];
        }
        dst[0] = m;
        dst[1] = w;
    }
}

// Sums am and aw over the 64 lanes of a wave; lane 0 writes them to dst[0], dst[1].
__device__ __forceinline__ void wave_reduce_store(double am, double aw, double* dst) {
    for (int off = 32; off > 0; off >>= 1) {
        am += __shfl_down(am, off, 64);
        aw += __shfl_down(aw, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        dst[0] = am;
        dst[1] = aw;
    }
}

// LDS row strides of the products planes and of the x-pass output: 2 (mod 4) doubles.
// The x pass reads and writes each lane's row two doubles at a time (ds_read_b128 /
// ds_write2_b64), 8 lanes per LDS cycle; with a stride of 4k + 2 doubles the 8 lanes that
// walk down consecutive rows cover 8 distinct groups of 4 banks. (Odd strides suit
// single-double accesses and left 33% more conflict cycles here.)
__host__ __device__ constexpr int lds_stride(int n) { return n + ((6 - n % 4) % 4); }
// The 2D kernel's 16-wide tiles: odd strides. Its x pass has one run of 16 per row, and
// there odd strides give 1.5% more throughput than 2 (mod 4), by measurement.
__host__ __device__ constexpr int lds_stride_2d(int n) { return n | 1; }

// Products of one position into the channel planes of the LDS buffer (plane size np,
// position i). NC = 9 (+count): I, G, U, II, GG, IG, UU, UI, UG.
// NC = 7: I, G, II, GG, IG, UI, UG.
template <int NC>
__device__ __forceinline__ void store_products(double* s_prod, int* s_cnt, int np, int i, float I,
                                               float G, float U) {
    const double di = I, dg = G, du = U;
    if (NC == 9) {
        s_prod[0 * np + i] = di;
        s_prod[1 * np + i] = dg;
        s_prod[2 * np + i] = du;
        s_prod[3 * np + i] = di * di;
        s_prod[4 * np + i] = dg * dg;
        s_prod[5 * np + i] = di * dg;
        s_prod[6 * np + i] = du * du;
        s_prod[7 * np + i] = du * di;
        s_prod[8 * np + i] = du * dg;
        s_cnt[i] = (du > kZeroThreshold) ? 1 : 0;
    } else {
        s_prod[0 * np + i] = di;
        s_prod[1 * np + i] = dg;
        s_prod[2 * np + i] = di * di;
        s_prod[3 * np + i] = dg * dg;
        s_prod[4 * np + i] = di * dg;
        s_prod[5 * np + i] = du * di;
        s_prod[6 * np + i] = du * dg;
    }
}

// Window sums of width W over a run of NO outputs: out[o] = in[o] + ... + in[o + W - 1],
// o = 0..NO-1, inputs in[0 .. NO + W - 2] (elements `stride` apart). van Herk / Gil-Werman
// form: the run is cut into blocks of W inputs and a window is the suffix of its block
// plus the prefix of the next one, so a partial sum never holds a value from outside the
// window. (A sliding sum that adds the entering and subtracts the leaving value keeps the
// rounding residue of every large value that passed through it, which swamps windows of
// small variance further along the run.) Each input is read once.
template <int W, typename V, int NO = 16>
__device__ __forceinline__ void window_run(const V* in, int stride, V* out, int ostride) {
    constexpr int NB = (NO + W - 1) / W;  // blocks that hold outputs
    constexpr int NIN = NO - 1 + W;       // inputs of the run
    V raw[W];
#pragma unroll
    for (int j = 0; j < W; ++j) raw[j] = in[j * stride];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        V suf[W];  // suffix sums of block b
        suf[W - 1] = raw[W - 1];
#pragma unroll
        for (int j = W - 2; j >= 0; --j) suf[j] = raw[j] + suf[j + 1];
        V pre = 0;  // prefix of block b + 1: elements 0 .. j - 1 at output b * W + j
#pragma unroll
        for (int j = 0; j < W; ++j) {
            const int o = b * W + j;
            if (o < NO) out[o * ostride] = j == 0 ? suf[0] : suf[j] + pre;
            const int t = (b + 1) * W + j;
            if (t < NIN) {
                raw[j] = in[t * stride];
                pre = j == 0 ? raw[j] : pre + raw[j];
            }
        }
    }
}

// Box sums of width W along x then y, for NC FP64 channels and (CNT) one integer count
// channel, over a region of py rows (row stride ps >= TX + W - 1) into TX x TY outputs:
//   s_prod[c][py][ps] -> s_xp[c][py][TX + 1] -> s_out[c][TY][TX]   (s_out may alias s_prod).
// Rows (x pass) are cut into runs of 16 outputs and columns (y pass) into runs of 8, one
// task each (runs of 8 put twice as many threads to work in the y pass); neighbouring
// lanes take neighbouring rows (x pass) or columns (y pass).
// PRIO_X / PRIO_Y: run the x / y pass at raised wave priority. With several work-groups
// per CU out of phase, the LDS passes gain from going first (2D +3.2%, 3D +1.3% with
// three barriers per z-step, measured). The 3D kernel's half-step overlap keeps only the
// y pass raised (x pass raised too: 0.5% slower, measured).
// x pass of box_xy: s_prod[c][py][ps] -> s_xp[c][py][XS] (TX outputs per row).
template <int NC, bool CNT, int TX, int NTHR, int W, int XS, bool PRIO = false>
__device__ __forceinline__ void box_x(const double* s_prod, const int* s_cnt, double* s_xp,
                                      int* s_xcnt, int ps, int py, int tid) {
    static_assert(TX % 16 == 0, "tile edges are multiples of 16");
    static_assert(XS > TX, "x-pass rows hold TX outputs");
    constexpr int NT = NC + (CNT ? 1 : 0);
    constexpr int SX = TX / 16;
    const int np = ps * py;
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(2);
    for (int t = tid; t < NT * py * SX; t += NTHR) {
        const int row = t % py, rest = t / py;
        const int seg = rest % SX, c = rest / SX;
        if (c < NC)
            window_run<W>(s_prod + c * np + row * ps + seg * 16, 1,
                        s_xp + (c * py + row) * XS + seg * 16, 1);
        else
            window_run<W>(s_cnt + row * ps + seg * 16, 1, s_xcnt + row * XS + seg * 16, 1);
    }
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
}

// y pass of box_xy: s_xp[c][py][XS] -> s_out[c][TY][TX].
template <int NC, bool CNT, int TX, int TY, int NTHR, int W, int XS, bool PRIO = false>
__device__ __forceinline__ void box_y(const double* s_xp, const int* s_xcnt, double* s_out,
                                      int* s_ocnt, int py, int tid) {
    static_assert(TY % 16 == 0, "tile edges are multiples of 16");
    constexpr int NT = NC + (CNT ? 1 : 0);
    constexpr int SY8 = TY / 8;  // y runs of 8 outputs
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(2);
    for (int t = tid; t < NT * TX * SY8; t += NTHR) {
        const int col = t % TX, rest = t / TX;
        const int seg = rest % SY8, c = rest / SY8;
        if (c < NC)
            window_run<W, double, 8>(s_xp + (c * py + seg * 8) * XS + col, XS,
                                   s_out + c * TX * TY + seg * 8 * TX + col, TX);
        else
            window_run<W, int, 8>(s_xcnt + seg * 8 * XS + col, XS, s_ocnt + seg * 8 * TX + col, TX);
    }
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
}

template <int NC, bool CNT, int TX, int TY, int NTHR, int W, int XS, bool PRIO_X = false,
          bool PRIO_Y = false>
__device__ __forceinline__ void box_xy(const double* s_prod, const int* s_cnt, double* s_xp,
                                       int* s_xcnt, double* s_out, int* s_ocnt, int ps, int py,
                                       int tid) {
    box_x<NC, CNT, TX, NTHR, W, XS, PRIO_X>(s_prod, s_cnt, s_xp, s_xcnt, ps, py, tid);
    __syncthreads();
    box_y<NC, CNT, TX, TY, NTHR, W, XS, PRIO_Y>(s_xp, s_xcnt, s_out, s_ocnt, py, tid);
    __syncthreads();
}

// Row of output oy of the in-place box sums (box_xy_inplace) in the products planes.
__host__ __device__ constexpr int box_row(int oy, int w) { return oy < 8 ? oy : oy + w - 1; }

// Box sums in place (the 2D kernel), for 16x16 outputs: the x pass writes each row's 16
// outputs over the start of that row of s_prod, and the y pass writes each column run's 8
// outputs over rows that no other run reads: run 0 over rows 0..7, run 1 over rows
// W+7 .. W+14 (the last 8 of its own input rows). A task reads each of its inputs before it
// overwrites it, so output (ox, oy) of channel c ends up at
// s_prod[c * ps * py + box_row(oy, W) * ps + ox] with no second buffer: the LDS of the x
// pass's output and of the sums is saved, which lets more work-groups share a CU.
template <int NC, bool CNT, int NTHR, int W, bool PRIO_X = false, bool PRIO_Y = false>
__device__ __forceinline__ void box_xy_inplace(double* s_prod, int* s_cnt, int ps, int py,
                                               int tid) {
    constexpr int NT = NC + (CNT ? 1 : 0);
    const int np = ps * py;
    if constexpr (PRIO_X) __builtin_amdgcn_s_setprio(2);
    for (int t = tid; t < NT * py; t += NTHR) {
        const int row = t % py, c = t / py;
        if (c < NC) {
            double* r = s_prod + c * np + row * ps;
            window_run<W>(r, 1, r, 1);
        } else {
            int* r = s_cnt + row * ps;
            window_run<W>(r, 1, r, 1);
        }
    }
    if constexpr (PRIO_X) __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    if constexpr (PRIO_Y) __builtin_amdgcn_s_setprio(2);
    for (int t = tid; t < NT * 16 * 2; t += NTHR) {
        const int col = t % 16, rest = t / 16;
        const int seg = rest % 2, c = rest / 2;
        const int in_ofs = seg * 8 * ps + col, out_ofs = box_row(seg * 8, W) * ps + col;
        if (c < NC)
            window_run<W, double, 8>(s_prod + c * np + in_ofs, ps, s_prod + c * np + out_ofs, ps);
        else
            window_run<W, int, 8>(s_cnt + in_ofs, ps, s_cnt + out_ofs, ps);
    }
    if constexpr (PRIO_Y) __builtin_amdgcn_s_setprio(0);
    __syncthreads();
}

// box_xy_inplace for a patch radius known only at run time, 1 <= r <= RMAX.
template <int NC, bool CNT, int NTHR, int RMAX = kMaxR2, int RR = 1>
__device__ __forceinline__ void box_xy_inplace_r(int r, double* s_prod, int* s_cnt, int tid) {
    if (r == RR)
        box_xy_inplace<NC, CNT, NTHR, 2 * RR + 1, true, true>(s_prod, s_cnt,
                                                             lds_stride_2d(16 + 2 * RR), 16 + 2 * RR, tid);
    else if constexpr (RR < RMAX)
        box_xy_inplace_r<NC, CNT, NTHR, RMAX, RR + 1>(r, s_prod, s_cnt, tid);
}

// ---------------------------------------------------------------------------
// 3D fused kernel: a TX x TY column of output voxels over `zchunk` planes of one pose.
// MOVI: the moving image is the high-resolution I (warped + Sobel here), the fixed
// ultrasound U is static with precomputed patch statistics (7 channels). Otherwise U is
// warped here and I, |grad I| are static (9 channels + count).
// Global loads are software-pipelined one z-step ahead: the gathers of the next warped
// plane, the static image values of the next products plane and the statistics of the
// next output plane are issued early in a step and consumed in the following one.
// ---------------------------------------------------------------------------
template <int R, bool MOVI, typename T, int TX, int TY, bool P2>
__global__ __launch_bounds__(TX * TY, 1) void k_fused3d(FusedArgs a) {
    constexpr int NTHR = TX * TY;
    constexpr int W = 2 * R + 1;
    constexpr int PX = TX + 2 * R, PY = TY + 2 * R;  // patch-support region
    constexpr int NP = PX * PY;
    constexpr int PS = lds_stride(PX);               // LDS row stride of the products planes
    constexpr int XS = lds_stride(TX + 1);           // LDS row stride of the x-pass output
    constexpr int IXE = PX + 2, IYE = PY + 2;        // Sobel input region
    constexpr int NI = MOVI ? IXE * IYE : NP;        // warped positions per plane
    constexpr int SPT = (NI + NTHR - 1) / NTHR;      // warped samples per thread
    constexpr int PPT = (NP + NTHR - 1) / NTHR;      // products positions per thread
    constexpr int N3 = W * W * W;
    constexpr int LAG = MOVI ? 2 : 0;
    constexpr int NC = MOVI ? 7 : 9;
    constexpr bool CNT = !MOVI;
    // PIPE: steps overlap by half a step (two barriers per z-step instead of four; the
    // loop below says how). The plane sums then need their own LDS (s_sum) instead of
    // reusing s_prod, which the next step's products overwrite while the ring still
    // reads; that fits two work-groups per CU with 7 channels, not with 9. Not at R = 1,
    // where the longer-lived gather registers cost a wave per SIMD (2.64 -> 3.44 ms with
    // the gathers finished a phase early, measured).
    constexpr bool PIPE = MOVI && R >= 2;
    // PAIR: with the moving image warped here and two products slots per thread, a thread
    // takes two vertically adjacent positions of one column (rows ly0 and ly0 + 1, ly0
    // even) instead of positions tid and tid + NTHR: their 3x3 neighbourhoods share two
    // rows, so 4 rows of the warped plane are read and smoothed instead of 6, with the
    // same operations per position. The warped plane's LDS rows are then SIS = 32 + PX/2
    // floats apart: the PX lanes of a pair row read one row of it, the next PX lanes the
    // row two below (2 SIS = 64 + PX words on), so the groups of a wave fall into
    // distinct banks (32 or 64 banks).
    constexpr int NPAIR = PX * (PY / 2);
    constexpr bool PAIR = MOVI && PY % 2 == 0 && (NPAIR + NTHR - 1) / NTHR * 2 == PPT;
    constexpr int SIS = PAIR ? 32 + PX / 2 : IXE;  // row stride of the warped plane in LDS
    static_assert(SIS >= IXE, "warped rows fit their stride");

    __shared__ float s_in[PAIR ? IYE * SIS : NI];
    __shared__ double s_prod[NC * PY * PS];
    __shared__ int s_cnt[CNT ? PY * PS : 1];
    __shared__ double s_xp[NC * PY * XS];
    __shared__ int s_xcnt[CNT ? PY * XS : 1];
    __shared__ double s_sum[PIPE ? NC * TX * TY : 1];
    double* const s_out = PIPE ? s_sum : s_prod;

    int tile, pose;
    if (!map_block(a.tiles, a.npose, tile, pose)) return;  // whole work-group
    const int tid = threadIdx.x;
    const int tcol = tile % (a.ntx * a.nty), zc = tile / (a.ntx * a.nty);
    const int tx = tcol % a.ntx, ty = tcol / a.ntx;
    const int x0 = R + tx * TX, y0 = R + ty * TY;
    const int z0 = R + zc * a.zchunk;
    const int zend = min(z0 + a.zchunk, a.v.Z - R);
    const int nout = zend - z0;
    const int X = a.v.X, Y = a.v.Y, Z = a.v.Z;
    const T* __restrict__ src = static_cast<const T*>(a.src);
    float m[12];
    {
        const float* mp = a.mats + (size_t)pose * a.mats_per_pose * 12;
#pragma unroll
        for (int k = 0; k < 12; ++k) m[k] = mp[k];
    }
    // warped plane of step s: MOVI -> zi = z0-R-1+s (Sobel input), else zp = z0-R+s
    const int zw0 = MOVI ? z0 - R - 1 : z0 - R;
    const int wex = MOVI ? IXE : PX;     // warped region edge along x
    const int wofs = MOVI ? R + 1 : R;   // warped region starts at x0 - wofs, y0 - wofs
    const int nsteps = nout + 2 * R + LAG;

    Pending3<T> pend[SPT];
    // A wave whose 64 positions of slot j all lie past the region skips that slot (its
    // gathers and lerps, or its Sobel and products). The test is wave-uniform (scalar
    // branch).
    const int wbase = __builtin_amdgcn_readfirstlane(tid & ~63);
    auto wave_has = [&](int j, int n) { return (j + 1) * NTHR <= n || j * NTHR + wbase < n; };
    // The partial last gather slot goes to the last waves instead of the first (with 576
    // positions on 256 threads the third slot belongs to wave 3 alone): those are the
    // waves the x pass leaves idle (its 7 x PY row tasks fill the first waves).
    constexpr int GREM = NI - (SPT - 1) * NTHR;           // positions in the last slot
    constexpr int GSKIP = NTHR - (GREM + 63) / 64 * 64;   // threads before its first wave
    auto gpos = [&](int j) { return j < SPT - 1 ? tid + j * NTHR : j * NTHR + tid - GSKIP; };
    auto gwave_has = [&](int j) { return j < SPT - 1 || wbase >= GSKIP; };
    // the in-plane part of each warped position's source coordinates, m*(Lx, Ly, 0), is the
    // same for every plane: computed once, in the order issue3 uses
    float cxy[SPT][3];
    int sin_ofs[SPT];  // LDS offset of warped position gpos(j) in s_in
    unsigned in_xy = 0;
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int i = gpos(j);
        const int lx = i % wex, ly = i / wex;
        sin_ofs[j] = PAIR ? ly * SIS + lx : i;
        const int gx = x0 - wofs + lx, gy = y0 - wofs + ly;
        if (i >= j * NTHR && i < NI && gx >= 0 && gx < X && gy >= 0 && gy < Y) in_xy |= 1u << j;
        const float Lx = fmaf((float)gx, a.v.sx, a.v.ox), Ly = fmaf((float)gy, a.v.sy, a.v.oy);
        cxy[j][0] = fmaf(m[1], Ly, m[0] * Lx);
        cxy[j][1] = fmaf(m[5], Ly, m[4] * Lx);
        cxy[j][2] = fmaf(m[9], Ly, m[8] * Lx);
    }
    auto issue_plane = [&](int zi) {
        const bool zin = zi >= 0 && zi < Z;
        const float lz = fmaf((float)zi, a.v.sz, a.v.oz);
#pragma unroll
        for (int j = 0; j < SPT; ++j)
            if (gwave_has(j)) issue3c<P2>(src, a.s, m, cxy[j], lz, zin && ((in_xy >> j) & 1u), pend[j]);
    };
    // finish the gathers of the plane in flight into s_in. The finishing wave runs at
    // raised priority so that every wave reaches the next barrier sooner (+0.3%, measured).
    auto finish_plane = [&]() {
        __builtin_amdgcn_s_setprio(2);
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            if (!gwave_has(j)) continue;
            const int i = gpos(j);
            const float v = finish3(pend[j]);
            if (i >= j * NTHR && i < NI) s_in[sin_ofs[j]] = v;
        }
        __builtin_amdgcn_s_setprio(0);
    };
    // static image values (U if MOVI, else I and |grad I|) of the products positions
    float pre0[PPT], pre1[PPT];
    // products position (lx, ly) of slot j; false past the region
    auto ppos = [&](int j, int& lx, int& ly) {
        if (PAIR) {
            const int t = tid + (j >> 1) * NTHR;
            lx = t % PX;
            ly = 2 * (t / PX) + (j & 1);
            return t < NPAIR;
        }
        const int i = tid + j * NTHR;
        lx = i % PX;
        ly = i / PX;
        return i < NP;
    };
    auto issue_static = [&](int zp) {
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            int lx, ly;
            const bool pin = ppos(j, lx, ly);
            const int gx = x0 - R + lx, gy = y0 - R + ly;
            const bool in = pin && gx < X && gy < Y && zp >= 0 && zp < Z;
            const int vidx = (zp * Y + gy) * X + gx;
            pre0[j] = in ? a.stat_img[vidx] : 0.0f;
            if (!MOVI) pre1[j] = in ? a.stat_grad[vidx] : 0.0f;
        }
    };
    // static ultrasound statistics of this thread's output voxel
    const int ox = tid % TX, oy = tid / TX;
    const bool col_valid = (x0 + ox < X - R) && (y0 + oy < Y - R);
    double mu_pre = 0.0, cyy_pre = -1.0;
    auto issue_stats = [&](int zo) {
        if (MOVI) {
            const int vo = (zo * Y + (y0 + oy)) * X + (x0 + ox);
            const bool ok = col_valid && zo >= z0 && zo < zend;
            mu_pre = ok ? a.ustat_mean[vo] : 0.0;
            cyy_pre = ok ? a.ustat_cyy[vo] : -1.0;
        }
    };

    float tA[PPT][3], tBx[PPT][3], tBy[PPT][3], Iprev[PPT], Icur[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
#pragma unroll
        for (int k = 0; k < 3; ++k) { tA[j][k] = 0.f; tBx[j][k] = 0.f; tBy[j][k] = 0.f; }
        Iprev[j] = 0.f;
        Icur[j] = 0.f;
        pre0[j] = 0.f;
        pre1[j] = 0.f;
    }
    double ring[W][NC];
    int ringc[W];
    int SC = 0;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        ringc[k] = 0;
#pragma unroll
        for (int c = 0; c < NC; ++c) ring[k][c] = 0.0;
    }
    double acc_m = 0.0, acc_w = 0.0;
    const int cols = a.ntx * a.nty;

    issue_plane(zw0);
    issue_static(z0 - R);      // products plane of the first products step
    issue_stats(z0);           // first output plane
    if (PIPE) {
        finish_plane();
        __syncthreads();
    }

    // 2. in-plane Sobel of the 3x3 neighbourhood of the warped plane in s_in (rows: smooth
    // s = (l + 2c) + r and difference d = r - l along x; then A, Bx, By down the rows), a
    // 3-plane ring in registers along z, |grad| and, for q >= 0, the products of plane
    // z0 - R + q into s_prod
    auto sobel_products = [&](int q, const float* cur0, const float* cur1) {
        if constexpr (PAIR) {
#pragma unroll
            for (int jp = 0; jp < PPT / 2; ++jp) {
                if (!wave_has(jp, NPAIR)) continue;
                const int t = tid + jp * NTHR;
                if (t < NPAIR) {
                    const int lx = t % PX, ly0 = 2 * (t / PX);
                    // rows ly0 .. ly0 + 3: smooth s = (l + 2c) + r, difference d = r - l
                    const float* row = s_in + ly0 * SIS + lx;
                    float sm[4], df[4], cm[4];
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) {
                        const float l = row[rr * SIS], c = row[rr * SIS + 1], r = row[rr * SIS + 2];
                        sm[rr] = (l + 2.0f * c) + r;
                        df[rr] = r - l;
                        cm[rr] = c;
                    }
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int j = 2 * jp + e;
                        // A = (s0 + 2 s1) + s2, Bx = (d0 + 2 d1) + d2, By = s2 - s0 over the
                        // position's rows ly0 + e .. ly0 + e + 2
                        tA[j][0] = tA[j][1]; tA[j][1] = tA[j][2];
                        tBx[j][0] = tBx[j][1]; tBx[j][1] = tBx[j][2];
                        tBy[j][0] = tBy[j][1]; tBy[j][1] = tBy[j][2];
                        tA[j][2] = (sm[e] + 2.0f * sm[e + 1]) + sm[e + 2];
                        tBx[j][2] = (df[e] + 2.0f * df[e + 1]) + df[e + 2];
                        tBy[j][2] = sm[e + 2] - sm[e];
                        Iprev[j] = Icur[j];
                        Icur[j] = cm[e + 1];
                        if (q >= 0) {
                            const float gxv = (tBx[j][0] + 2.0f * tBx[j][1]) + tBx[j][2];
                            const float gyv = (tBy[j][0] + 2.0f * tBy[j][1]) + tBy[j][2];
                            const float gzv = tA[j][2] - tA[j][0];
                            const float G = sqrtf(fmaf(gzv, gzv, fmaf(gyv, gyv, gxv * gxv)));
                            store_products<NC>(s_prod, s_cnt, PY * PS, (ly0 + e) * PS + lx,
                                               Iprev[j], G, cur0[j]);
                        }
                    }
                }
            }
            return;
        }
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            if (!wave_has(j, NP)) continue;
            const int i = tid + j * NTHR;
            if (i < NP) {
                const int lx = i % PX, ly = i / PX;
                if (MOVI) {
                    // rows in order, accumulating A = (s0 + 2 s1) + s2,
                    // Bx = (d0 + 2 d1) + d2 and By = s2 - s0 as they arrive
                    const float* row = s_in + ly * IXE + lx;
                    float l = row[0], c = row[1], r = row[2];
                    const float s0 = (l + 2.0f * c) + r;
                    float a = s0, bx = r - l;
                    row += IXE;
                    l = row[0], c = row[1], r = row[2];
                    a = a + 2.0f * ((l + 2.0f * c) + r);
                    bx = bx + 2.0f * (r - l);
                    const float cc = c;
                    row += IXE;
                    l = row[0], c = row[1], r = row[2];
                    const float s2 = (l + 2.0f * c) + r;
                    tA[j][0] = tA[j][1]; tA[j][1] = tA[j][2];
                    tBx[j][0] = tBx[j][1]; tBx[j][1] = tBx[j][2];
                    tBy[j][0] = tBy[j][1]; tBy[j][1] = tBy[j][2];
                    tA[j][2] = a + s2;
                    tBx[j][2] = bx + (r - l);
                    tBy[j][2] = s2 - s0;
                    Iprev[j] = Icur[j];
                    Icur[j] = cc;
                }
                if (q >= 0) {
                    float I, G, U;
                    if (MOVI) {
                        const float gxv = (tBx[j][0] + 2.0f * tBx[j][1]) + tBx[j][2];
                        const float gyv = (tBy[j][0] + 2.0f * tBy[j][1]) + tBy[j][2];
                        const float gzv = tA[j][2] - tA[j][0];
                        G = sqrtf(fmaf(gzv, gzv, fmaf(gyv, gyv, gxv * gxv)));
                        I = Iprev[j];
                        U = cur0[j];
                    } else {
                        U = s_in[i];
                        I = cur0[j];
                        G = cur1[j];
                    }
                    store_products<NC>(s_prod, s_cnt, PY * PS, ly * PS + lx, I, G, U);
                }
            }
        }
    };
    // 4. z ring of plane sums and combine, for the step of products plane q (q >= 0) whose
    // plane sums are in s_out, ring slot SLOT. The patch sums add the W plane sums in z
    // order (no running sum), so a voxel's value does not depend on where its work-group's
    // z-chunk starts; integer counts can run.
    auto ring_combine = [&](auto slot_c, int q) {
        constexpr int slot = decltype(slot_c)::value;
        const int zp = z0 - R + q;
        if (CNT) {
            const int cc = s_cnt[tid];
            SC += cc - ringc[slot];
            ringc[slot] = cc;
        }
        // pair sums: slot (q-1) turns from plane q-1 into Q(q-1) = P(q-1) + P(q), slot q
        // takes P(q); a window of W = 2R+1 planes ending at q is then the R pair sums
        // Q(q-2R), Q(q-2R+2), ..., Q(q-2) plus P(q): R+1 adds per channel, not 2R
        constexpr int prev = (slot + W - 1) % W;
        double Pn[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            Pn[c] = s_out[c * TX * TY + tid];
            ring[prev][c] = ring[prev][c] + Pn[c];
            ring[slot][c] = Pn[c];
        }
        if (q >= 2 * R) {
            const double mu = mu_pre, cyy = cyy_pre;
            if (q + 1 < nout + 2 * R) issue_stats(zp - R + 1);
            double S[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                double t = ring[(slot + 1) % W][c];  // Q(q-2R)
#pragma unroll
                for (int j = 1; j < R; ++j) t += ring[(slot + 1 + 2 * j) % W][c];
                S[c] = t + Pn[c];
            }
            if (col_valid) {
                double w, meas;
                bool ok;
                if (MOVI) ok = combine_static_u(S, mu, cyy, N3, w, meas);
                else ok = combine(S, SC, N3, w, meas);
                if (ok) {
                    acc_m += meas;
                    acc_w += w;
                    if (a.wmap) {
                        const int vo = ((zp - R) * Y + (y0 + oy)) * X + (x0 + ox);
                        a.wmap[vo] = w;
                        a.mmap[vo] = meas;
                    }
                }
            }
            // Partial sums per block of kZBlock output planes, tile column and wave, at
            // fixed positions in z: every z-chunk (a multiple of kZBlock) and every batch
            // size then adds the same numbers in the same order. One partial per wave needs
            // no barrier (a work-group sum would take one every kZBlock steps).
            const int zo = zp - R;  // this step's output plane (block-uniform)
            if ((zo - R + 1) % kZBlock == 0 || zo + 1 == zend) {
                constexpr int NW = NTHR / 64;
                double* dst = a.partials + (((size_t)pose * a.blocks_per_pose +
                                             ((size_t)((zo - R) / kZBlock) * cols + tcol) * NW +
                                             (tid >> 6)) * 2);
                wave_reduce_store(acc_m, acc_w, dst);
                acc_m = 0.0;
                acc_w = 0.0;
            }
        }
    };

    if constexpr (PIPE) {
        // Two barriers per z-step: step s's Sobel and products share a phase with step
        // s - 1's y pass, and step s's x pass with step s - 1's ring and combine and the
        // finishing of plane s + 1's gathers. Each phase writes only LDS that the other
        // phase reads, so one barrier closes it: phase 1 writes s_prod and s_sum and reads
        // s_in and s_xp, phase 2 writes s_xp and s_in and reads s_prod and s_sum.
        for (int s0 = 0; s0 <= nsteps; s0 += W) {
            // W consecutive steps with compile-time ring slots, so the z ring stays in VGPRs.
            Unroll<0, W>::run([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                const int s = s0 + k;
                if (s > nsteps) return;  // block-uniform; step nsteps only drains s - 1
                const int q = s - LAG;
                const bool here = s < nsteps;
                const bool more = s + 1 < nsteps;
                const bool prev = s >= 1 && q >= 1;  // step s - 1 made products
                // phase 1
                if (here) {
                    float cur0[PPT], cur1[PPT];
#pragma unroll
                    for (int j = 0; j < PPT; ++j) {
                        cur0[j] = pre0[j];
                        cur1[j] = pre1[j];
                    }
                    if (q >= 0 && q + 1 < nout + 2 * R) issue_static(z0 - R + q + 1);
                    sobel_products(q, cur0, cur1);
                    // the next plane's gathers fly during the y pass, the barrier and the
                    // x pass
                    if (more) issue_plane(zw0 + s + 1);
                }
                if (prev) box_y<NC, CNT, TX, TY, NTHR, W, XS, true>(s_xp, s_xcnt, s_out, s_cnt, PY, tid);
                __syncthreads();
                // phase 2
                if (here && q >= 0)
                    box_x<NC, CNT, TX, NTHR, W, XS, false>(s_prod, s_cnt, s_xp, s_xcnt, PS, PY, tid);
                if (more) finish_plane();
                // the ring and combine at raised priority: 0.7% faster (measured); the x pass
                // stays at normal priority
                __builtin_amdgcn_s_setprio(1);
                if (prev) ring_combine(std::integral_constant<int, (k - 1 - LAG + 2 * W) % W>{}, q - 1);
                __builtin_amdgcn_s_setprio(0);
                __syncthreads();
            });
        }
        return;
    }

    for (int s0 = 0; s0 < nsteps; s0 += W) {
        // W consecutive steps with compile-time ring slots (k), so the z ring stays in VGPRs.
        Unroll<0, W>::run([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const int s = s0 + k;
            if (s >= nsteps) return;  // block-uniform
            const int q = s - LAG;
            const int zp = z0 - R + q;
            // 1. finish the plane issued last step
            finish_plane();
            float cur0[PPT], cur1[PPT];
#pragma unroll
            for (int j = 0; j < PPT; ++j) {
                cur0[j] = pre0[j];
                cur1[j] = pre1[j];
            }
            if (q >= 0 && q + 1 < nout + 2 * R) issue_static(zp + 1);
            __syncthreads();
            sobel_products(q, cur0, cur1);
            __syncthreads();  // s_in reads done; products visible
            // the next plane's gathers fly during the box sums and the combine
            if (s + 1 < nsteps) issue_plane(zw0 + s + 1);
            if (q < 0) return;  // block-uniform: Sobel ring warm-up only
            // 3. in-plane box sums
            box_xy<NC, CNT, TX, TY, NTHR, W, XS, true, true>(s_prod, s_cnt, s_xp, s_xcnt, s_out,
                                                            s_cnt, PS, PY, tid);
            ring_combine(std::integral_constant<int, (k - LAG + W) % W>{}, q);
        });
    }
}

// ---------------------------------------------------------------------------
// 2D fused kernel: one 16x16 output tile of a run of `zchunk` (kFrameRun) consecutive
// frames of one pose (a single image is one frame). Dynamic LDS sized for the runtime patch
// radius. Frame f + 1's gathers (and its static image values and ultrasound statistics)
// are issued while frame f's box sums and combine run, as in the 3D kernel's z-steps. Each
// thread adds its pixel's measure and weight over the run's frames in frame order, and the
// work-group reduces them once into the partial of (run, tile). The run length does not
// depend on the batch, so neither do the results.
// ---------------------------------------------------------------------------
template <bool MOVI, bool SRC3D, typename T, int RMAX, bool P2>
// MOVI at r <= 3: at most 128 VGPRs, four work-groups per CU (the f32-source variant
// would take 129 and drop to three)
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(MOVI && RMAX <= 3 ? 4 : 1)))
void k_fused2d(FusedArgs a) {
    extern __shared__ double s_dyn[];
    static_assert(kTile2X == 16 && kTile2Y == 16, "box_xy_inplace sums 16x16 tiles");
    constexpr int NC = MOVI ? 7 : 9;
    constexpr bool CNT = !MOVI;
    // per-thread gathers / products positions at the largest radius of this variant
    // (r <= RMAX): the pipelined gathers are held in registers
    constexpr int SPT = ((kTile2X + 2 * RMAX + 2) * (kTile2Y + 2 * RMAX + 2) + kThreads - 1) / kThreads;
    constexpr int PPT = ((kTile2X + 2 * RMAX) * (kTile2Y + 2 * RMAX) + kThreads - 1) / kThreads;
    const int R = a.r, W = 2 * R + 1, PX = kTile2X + 2 * R, PY = kTile2Y + 2 * R, NP = PX * PY;
    const int IXE = PX + 2, IYE = PY + 2;
    // MOVI: warped rows 32 + PX/2 floats apart in LDS, so the paired Sobel reads of a wave
    // (PX lanes per row group, groups two rows apart) fall into distinct banks
    const int SIS = MOVI ? 32 + PX / 2 : IXE;
    const int PS = lds_stride_2d(PX);                         // products row stride
    const int wex = MOVI ? IXE : PX;                          // warped region edge
    const int wofs = MOVI ? R + 1 : R;                        // it starts at x0 - wofs
    const int NI = MOVI ? IXE * IYE : NP;
    // The box sums run in place in the products planes (box_xy_inplace).
    double* s_prod = s_dyn;                                   // NC*PY*PS
    double* s_red = s_prod + NC * PY * PS;                    // 8
    float* s_in = reinterpret_cast<float*>(s_red + 8);        // NI
    int* s_cnt = reinterpret_cast<int*>(s_in + NI);           // PY*PS

    // A work-group takes one tile of either a run of frames of one pose (frame stacks, 3D
    // sources) or, with a.prun > 1 (single images of a 2D source), a run of a.prun poses:
    // their gathers pipeline like a run's frames, and the tile's static values are loaded
    // once. Each pose still gets one partial per tile, so the results do not depend on the
    // run length. (With a 3D source prun is the constant 1: the frame stacks' code is not
    // touched by the pose runs.)
    const int prun = (!SRC3D && a.prun > 1) ? a.prun : 1;
    int tile, prun_idx;
    if (!map_block(a.tiles, (a.npose + prun - 1) / prun, tile, prun_idx)) return;  // whole work-group
    const int p0 = prun_idx * prun, p1 = min(p0 + prun, a.npose);
    const int tid = threadIdx.x;
    const int cols = a.ntx * a.nty;
    const int tcol = tile % cols, run = tile / cols;
    const int f0 = run * a.zchunk, f1 = min(f0 + a.zchunk, a.v.Z);
    const int nit = prun > 1 ? p1 - p0 : f1 - f0;  // iterations: poses or frames
    const int tx = tcol % a.ntx, ty = tcol / a.ntx;
    const int x0 = R + tx * kTile2X, y0 = R + ty * kTile2Y;
    const int X = a.v.X, Y = a.v.Y;
    const float lz = a.v.oz;
    const T* __restrict__ src = static_cast<const T*>(a.src);

    // Per-thread positions, the same for every frame (the radius is a run-time value, so
    // the divisions are done once here, not per frame): virtual-grid coordinates of the
    // warped positions, and the LDS / image offsets of the products positions.
    float gLx[SPT], gLy[SPT];
    int g_sin[SPT];     // LDS offset of warped position j in s_in
    unsigned g_in = 0;  // bit j: warped position j lies in the grid
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int i = tid + j * kThreads;
        const int lx = i % wex, ly = i / wex;
        g_sin[j] = MOVI ? ly * SIS + lx : i;
        const int gx = x0 - wofs + lx, gy = y0 - wofs + ly;
        if (i < NI && gx >= 0 && gx < X && gy >= 0 && gy < Y) g_in |= 1u << j;
        gLx[j] = fmaf((float)gx, a.v.sx, a.v.ox);
        gLy[j] = fmaf((float)gy, a.v.sy, a.v.oy);
    }
    // MOVI: slots 2k and 2k + 1 are the vertically adjacent products positions (lx, ly0)
    // and (lx, ly0 + 1), ly0 even, of pair t = tid + k * kThreads: their 3x3 Sobel
    // neighbourhoods share two rows, so 4 rows are read and smoothed instead of 6, with the
    // same operations per position. (PPT is even at every RMAX; PY = 16 + 2R is even.)
    static_assert(PPT % 2 == 0, "products slots come in pairs");
    const int NPAIR = PX * (PY / 2);
    int p_in_ofs[PPT], p_prod_ofs[PPT], p_img_ofs[PPT];  // s_in, s_prod, image (frame 0)
    unsigned p_in = 0;  // bit j: products position j holds an image value
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int t = tid + (j >> 1) * kThreads;
        const int i = MOVI ? t : tid + j * kThreads;
        const int lx = MOVI ? t % PX : i % PX;
        const int ly = MOVI ? 2 * (t / PX) + (j & 1) : i / PX;
        const int gx = x0 - R + lx, gy = y0 - R + ly;
        if ((MOVI ? t < NPAIR : i < NP) && gx < X && gy < Y) p_in |= 1u << j;
        p_in_ofs[j] = MOVI ? ly * SIS + lx : i;
        p_prod_ofs[j] = ly * PS + lx;
        p_img_ofs[j] = gy * X + gx;
    }

    using Pend = typename std::conditional<SRC3D, Pending3<T>, Pending2<T>>::type;
    Pend pend[SPT];
    // (No wave-uniform skip of empty slots here, unlike the 3D kernel: in this kernel it
    // cost 6% on the tracked sweep, 29.09 -> 30.93 ms per launch, measured.)
    // the matrix of iteration it: pose p0 + it (pose runs) or frame f0 + it of pose p0
    auto issue_frame = [&](int it) {
        float m[12];
        const float* mp = prun > 1 ? a.mats + (size_t)(p0 + it) * a.mats_per_pose * 12
                                   : a.mats + ((size_t)p0 * a.mats_per_pose +
                                               (a.mats_per_pose > 1 ? f0 + it : 0)) * 12;
#pragma unroll
        for (int k = 0; k < 12; ++k) m[k] = mp[k];
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            if (j * kThreads >= NI) break;  // uniform
            const bool inside = (g_in >> j) & 1u;
            if constexpr (SRC3D) issue3<P2>(src, a.s, m, gLx[j], gLy[j], lz, inside, pend[j]);
            else issue2<P2>(src, a.s, m, gLx[j], gLy[j], inside, pend[j]);
        }
    };
    // static image values of the products positions (U if MOVI, else I and |grad I|)
    float pre0[PPT], pre1[PPT];
    auto issue_static = [&](int frame) {
        const int plane = frame * X * Y;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            // uniform: slot j holds pair j / 2 (MOVI) or position tid + j * kThreads
            if (MOVI ? (j >> 1) * kThreads >= NPAIR : j * kThreads >= NP) break;
            const bool in = (p_in >> j) & 1u;
            const int vidx = plane + p_img_ofs[j];
            pre0[j] = in ? a.stat_img[vidx] : 0.0f;
            if (!MOVI) pre1[j] = in ? a.stat_grad[vidx] : 0.0f;
        }
    };
    // static ultrasound statistics of this thread's output pixel
    const int ox = tid % kTile2X, oy = tid / kTile2X;
    const bool out_ok = (x0 + ox < X - R) && (y0 + oy < Y - R);
    double mu_pre = 0.0, cyy_pre = -1.0;
    auto issue_stats = [&](int frame) {
        if (MOVI && out_ok) {
            const int vo = frame * X * Y + (y0 + oy) * X + (x0 + ox);
            mu_pre = a.ustat_mean[vo];
            cyy_pre = a.ustat_cyy[vo];
        }
    };

    double acc_m = 0.0, acc_w = 0.0;
    issue_frame(0);
    issue_static(f0);
    issue_stats(f0);
    for (int it = 0; it < nit; ++it) {
        const int f = prun > 1 ? f0 : f0 + it;  // frame of this iteration
        // 1. the warped tile of iteration it (raised wave priority, as in the 3D kernel)
        __builtin_amdgcn_s_setprio(2);
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            if (j * kThreads >= NI) break;  // uniform
            const int i = tid + j * kThreads;
            float v;
            if constexpr (SRC3D) v = finish3(pend[j]);
            else v = finish2(pend[j]);
            if (i < NI) s_in[g_sin[j]] = v;
        }
        __builtin_amdgcn_s_setprio(0);
        float cur0[PPT], cur1[PPT];
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            cur0[j] = pre0[j];
            cur1[j] = pre1[j];
        }
        const double mu = mu_pre, cyy = cyy_pre;
        if (prun == 1 && f + 1 < f1) {  // the next frame's static values (pose runs: the same)
            issue_static(f + 1);
            issue_stats(f + 1);
        }
        __syncthreads();
        // 2. |Sobel| and the products
        if constexpr (MOVI) {
#pragma unroll
            for (int jp = 0; jp < PPT / 2; ++jp) {
                if (jp * kThreads >= NPAIR) break;  // uniform
                if (tid + jp * kThreads < NPAIR) {
                    // 3x3 Sobel read from the warped tile: rows ly0 .. ly0 + 3, smooth
                    // s = (l + 2c) + r and difference d = r - l along x
                    float sv[4], dv[4], cv[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const float* row = s_in + p_in_ofs[2 * jp] + k * SIS;
                        const float l = row[0], c = row[1], r = row[2];
                        dv[k] = r - l;
                        sv[k] = (l + 2.0f * c) + r;
                        cv[k] = c;
                    }
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const float gxv = (dv[e] + 2.0f * dv[e + 1]) + dv[e + 2];
                        const float gyv = sv[e + 2] - sv[e];
                        const float G = sqrtf(fmaf(gyv, gyv, gxv * gxv));
                        store_products<NC>(s_prod, s_cnt, PY * PS, p_prod_ofs[2 * jp + e],
                                           cv[e + 1], G, cur0[2 * jp + e]);
                    }
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < PPT; ++j) {
                if (j * kThreads >= NP) break;  // uniform
                if (tid + j * kThreads < NP)
                    store_products<NC>(s_prod, s_cnt, PY * PS, p_prod_ofs[j], cur0[j], cur1[j],
                                       s_in[p_in_ofs[j]]);
            }
        }
        __syncthreads();  // s_in reads done; products visible
        // the next iteration's gathers fly during the box sums and the combine
        if (it + 1 < nit) issue_frame(it + 1);
        // 3. box sums
        box_xy_inplace_r<NC, CNT, kThreads, RMAX>(R, s_prod, s_cnt, tid);
        // 4. combine; each thread adds its pixel's values over the run's frames in order
        if (out_ok) {
            const int so = box_row(oy, W) * PS + ox;  // this pixel's sums
            double S[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) S[c] = s_prod[c * PY * PS + so];
            double w, meas;
            bool ok;
            if (MOVI) ok = combine_static_u(S, mu, cyy, W * W, w, meas);
            else ok = combine(S, s_cnt[so], W * W, w, meas);
            if (ok) {
                acc_m += meas;
                acc_w += w;
                if (a.wmap) {
                    const int vo = f * X * Y + (y0 + oy) * X + (x0 + ox);
                    a.wmap[vo] = w;
                    a.mmap[vo] = meas;
                }
            }
        }
        if (prun > 1) {  // block-uniform: this pose's partial of the tile
            double* dst = a.partials + ((size_t)(p0 + it) * a.blocks_per_pose + (size_t)run * cols + tcol) * 2;
            block_reduce_store<kThreads>(acc_m, acc_w, s_red, dst);
            acc_m = 0.0;
            acc_w = 0.0;
        }
    }
    if (prun == 1) {
        double* dst = a.partials + ((size_t)p0 * a.blocks_per_pose + (size_t)run * cols + tcol) * 2;
        block_reduce_store<kThreads>(acc_m, acc_w, s_red, dst);
    }
}

// ---------------------------------------------------------------------------
// Separable path for patch radii beyond the fused kernels (3D r > kMaxR3, 2D r > kMaxR2):
// the reference switches to a global-memory kernel there (cu_lc2_3d.cu:769-827, chosen at
// :1133-1177). Here, per pose: the moving image is warped (and, if it is the high-resolution
// one, its |Sobel| taken) into two f32 grids, then the patch sums of every moment channel are
// box sums along x, y (and z) through two f64 buffers, each a van Herk / Gil-Werman pass
// (exact in the sense of the fused kernels: no partial sum holds a value from outside its
// window), and a last kernel combines them per voxel and reduces per work-group.
// Channels: moving high-resolution image (7): I, G, II, GG, IG, UI, UG; moving ultrasound
// (10): I, G, U, II, GG, IG, UU, UI, UG and the count of U > 1e-6 (exact as a double).
// ---------------------------------------------------------------------------
template <bool MOVI>
__device__ __forceinline__ double sep_product(int c, float I, float G, float U) {
    const double di = I, dg = G, du = U;
    if (MOVI) {
        switch (c) {
            case 0: return di;
            case 1: return dg;
            case 2: return di * di;
            case 3: return dg * dg;
            case 4: return di * dg;
            case 5: return du * di;
            default: return du * dg;
        }
    }
    switch (c) {
        case 0: return di;
        case 1: return dg;
        case 2: return du;
        case 3: return di * di;
        case 4: return dg * dg;
        case 5: return di * dg;
        case 6: return du * du;
        case 7: return du * di;
        case 8: return du * dg;
        default: return (du > kZeroThreshold) ? 1.0 : 0.0;
    }
}

// Window sums of width W along one line of n inputs: out(s) = in(s) + ... + in(s + W - 1)
// for s < n - W + 1, van Herk form with the output itself holding the block suffix sums
// between the two passes (no per-thread storage of W values).
template <class In, class Out>
__device__ __forceinline__ void vh_line(int n, int W, In in, Out out) {
    const int nw = n - W + 1;
    for (int b = 0; b < nw; b += W) {  // suffix sums of each block, from its end
        double acc = 0.0;
        for (int i = min(b + W, n) - 1; i >= b; --i) {
            acc += in(i);
            if (i < nw) out(i) = acc;
        }
    }
    for (int b = 0; b < nw; b += W) {  // plus the prefix of the next block
        double pre = 0.0;
        for (int j = 0; j < W && b + j < nw; ++j) {
            if (j > 0) out(b + j) = out(b + j) + pre;
            if (b + W + j < n) pre += in(b + W + j);
        }
    }
}

// One box pass. axis 0 reads the products of the I, G, U grids (U: ultrasound on the grid,
// I: high-resolution image, G: its |Sobel|); axes 1 and 2 read `src`. Window sums of lines
// along `axis` go to `dst` at the index of the window start. Buffers are [channel][V].
template <bool MOVI>
__global__ __launch_bounds__(256) void k_sep_pass(const float* __restrict__ I,
                                                  const float* __restrict__ G,
                                                  const float* __restrict__ U,
                                                  const double* __restrict__ src,
                                                  double* __restrict__ dst, VGeom v, int W,
                                                  int axis) {
    constexpr int NC = MOVI ? 7 : 10;
    const long long V = (long long)v.X * v.Y * v.Z;
    const int nX = axis == 0 ? 1 : v.X - W + 1;  // x positions of the lines (window starts)
    const int nY = axis == 2 ? v.Y - W + 1 : (axis == 1 ? 1 : v.Y);
    const long long lines = axis == 0 ? (long long)v.Y * v.Z
                          : axis == 1 ? (long long)nX * v.Z
                                      : (long long)nX * nY;
    const long long tasks = lines * NC;
    for (long long t = blockIdx.x * 256LL + threadIdx.x; t < tasks; t += (long long)gridDim.x * 256) {
        const int c = (int)(t / lines);
        const long long l = t % lines;
        long long base;  // index of the line's first element
        long long stride;
        int n;
        if (axis == 0) {
            base = l * v.X;
            stride = 1;
            n = v.X;
        } else if (axis == 1) {
            const int x = (int)(l % nX), z = (int)(l / nX);
            base = (long long)z * v.X * v.Y + x;
            stride = v.X;
            n = v.Y;
        } else {
            const int x = (int)(l % nX), y = (int)(l / nX);
            base = (long long)y * v.X + x;
            stride = (long long)v.X * v.Y;
            n = v.Z;
        }
        double* o = dst + (long long)c * V + base;
        if (axis == 0) {
            vh_line(n, W, [&](int i) { return sep_product<MOVI>(c, I[base + i], G[base + i], U[base + i]); },
                    [&](int s) -> double& { return o[s]; });
        } else {
            const double* in = src + (long long)c * V + base;
            vh_line(n, W, [&](int i) { return in[i * stride]; },
                    [&](int s) -> double& { return o[s * stride]; });
        }
    }
}

// Per output voxel: the patch sums (at the window-start index of `S`), weight and measure;
// per work-group sums into the pose's partials (fixed positions, fixed order).
template <bool MOVI>
__global__ __launch_bounds__(256) void k_sep_combine(const double* __restrict__ S,
                                                     const double* __restrict__ ustat_mean,
                                                     const double* __restrict__ ustat_cyy,
                                                     VGeom v, int r, int grid3, double* partials,
                                                     double* wmap, double* mmap) {
    __shared__ double s_red[8];
    constexpr int NC = MOVI ? 7 : 10;
    const long long V = (long long)v.X * v.Y * v.Z;
    const int W = 2 * r + 1;
    const int ex = v.X - 2 * r, ey = v.Y - 2 * r, ez = grid3 ? v.Z - 2 * r : v.Z;
    const int N = grid3 ? W * W * W : W * W;
    const long long nout = (long long)ex * ey * ez;
    const long long o = blockIdx.x * 256LL + threadIdx.x;
    double am = 0.0, aw = 0.0;
    if (o < nout) {
        const int sx = (int)(o % ex), sy = (int)((o / ex) % ey), sz = (int)(o / ((long long)ex * ey));
        const long long si = ((long long)sz * v.Y + sy) * v.X + sx;           // window start
        const long long vo = ((long long)(sz + (grid3 ? r : 0)) * v.Y + (sy + r)) * v.X + (sx + r);
        double Sc[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) Sc[c] = S[(long long)c * V + si];
        double w, meas;
        bool ok;
        if (MOVI) ok = combine_static_u(Sc, ustat_mean[vo], ustat_cyy[vo], N, w, meas);
        else ok = combine(Sc, (int)Sc[9], N, w, meas);
        if (ok) {
            am = meas;
            aw = w;
            if (wmap) {
                wmap[vo] = w;
                mmap[vo] = meas;
            }
        }
    }
    block_reduce_store<256>(am, aw, s_red, partials + (size_t)blockIdx.x * 2);
}

// Per-pose fixed-order reduction of the work-group partials -> (sum m, sum w, LC2)
// (GetSimilarityMetric, cu_lc2_3d.cu:1181-1191; sums in FP64, NaN/Inf -> 0).
__global__ __launch_bounds__(256) void k_reduce(const double* __restrict__ partials, int nb,
                                                double* __restrict__ out) {
    __shared__ double sm[256], sw[256];
    const int pose = blockIdx.x;
    const double2* p = reinterpret_cast<const double2*>(partials) + (size_t)pose * nb;
    double m = 0.0, w = 0.0;
    // each thread adds partials tid, tid + 256, ... in that order; four loads are in
    // flight at a time (the order of the additions, and so the result, is unchanged)
    int i = threadIdx.x;
    for (; i + 3 * 256 < nb; i += 4 * 256) {
        const double2 a0 = p[i], a1 = p[i + 256], a2 = p[i + 512], a3 = p[i + 768];
        m += a0.x; w += a0.y;
        m += a1.x; w += a1.y;
        m += a2.x; w += a2.y;
        m += a3.x; w += a3.y;
    }
    for (; i < nb; i += 256) {
        const double2 a = p[i];
        m += a.x;
        w += a.y;
    }
    sm[threadIdx.x] = m;
    sw[threadIdx.x] = w;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sm[threadIdx.x] += sm[threadIdx.x + h];
            sw[threadIdx.x] += sw[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double M = sm[0], Wt = sw[0];
        double l = M / Wt;
        if (isnan(l) || isinf(l)) l = 0.0;
        out[pose * 3 + 0] = M;
        out[pose * 3 + 1] = Wt;
        out[pose * 3 + 2] = l;
    }
}

template <typename T>
__global__ void k_warp_static(const T* fixed, SGeom fg, VGeom v, int mode, float* out) {
    // mode 0: 3D grid, trilinear; 1: 2D grid, bilinear per frame (frames = v.Z)
    const long long n = (long long)v.X * v.Y * v.Z;
    const float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % v.X), y = (int)((i / v.X) % v.Y), z = (int)(i / ((long long)v.X * v.Y));
        const float lx = fmaf((float)x, v.sx, v.ox), ly = fmaf((float)y, v.sy, v.oy);
        float val;
        if (mode == 0)
            val = sample3(fixed, fg, ident, lx, ly, fmaf((float)z, v.sz, v.oz));
        else
            val = sample2(fixed + (size_t)z * fg.X * fg.Y, fg, ident, lx, ly);
        out[i] = val;
    }
}

template <typename T>
__global__ void k_warp_moving(FusedArgs a, int grid3, int src3, float* out) {
    const VGeom v = a.v;
    const T* src = static_cast<const T*>(a.src);
    const long long n = (long long)v.X * v.Y * v.Z;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % v.X), y = (int)((i / v.X) % v.Y), z = (int)(i / ((long long)v.X * v.Y));
        const float* mp = a.mats + (size_t)(grid3 ? 0 : (a.mats_per_pose > 1 ? z : 0)) * 12;
        float m[12];
        for (int k = 0; k < 12; ++k) m[k] = mp[k];
        const float lx = fmaf((float)x, v.sx, v.ox), ly = fmaf((float)y, v.sy, v.oy);
        const float lz = grid3 ? fmaf((float)z, v.sz, v.oz) : v.oz;
        out[i] = src3 ? sample3(src, a.s, m, lx, ly, lz) : sample2(src, a.s, m, lx, ly);
    }
}

// ---------------------------------------------------------------------------
// Normalized mutual information, the reference's patchSize == 0 metric
// (GetMutualInformationSimilarityMetric, cu_lc2_3d.cu:1196-1301; histogram kernels
// :868-923): NMI = (H_U + H_I) / H_joint over the whole virtual grid, with bins spread
// between each image's minimum and maximum.
// ---------------------------------------------------------------------------

// Floats as ints whose signed order is the float order (for atomic min / max).
__device__ __forceinline__ int ford(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float fdec(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

// The moving image warped at virtual voxel i of pose `pose` (as k_warp_moving).
template <typename T>
__device__ __forceinline__ float warped_at(const FusedArgs& a, int grid3, int src3, int pose,
                                           long long i) {
    const VGeom& v = a.v;
    const int x = (int)(i % v.X), y = (int)((i / v.X) % v.Y), z = (int)(i / ((long long)v.X * v.Y));
    const float* mp = a.mats + ((size_t)pose * a.mats_per_pose +
                                (grid3 ? 0 : (a.mats_per_pose > 1 ? z : 0))) * 12;
    float m[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) m[k] = mp[k];
    const float lx = fmaf((float)x, v.sx, v.ox), ly = fmaf((float)y, v.sy, v.oy);
    const float lz = grid3 ? fmaf((float)z, v.sz, v.oz) : v.oz;
    const T* src = static_cast<const T*>(a.src);
    if (src3) return sample3(src, a.s, m, lx, ly, lz);
    return sample2(src, a.s, m, lx, ly);
}

template <int NTHR>
__device__ __forceinline__ void block_minmax(float mn, float mx, int* s_red, int* dst) {
    for (int off = 32; off > 0; off >>= 1) {
        mn = fminf(mn, __shfl_down(mn, off, 64));
        mx = fmaxf(mx, __shfl_down(mx, off, 64));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        s_red[2 * wid] = ford(mn);
        s_red[2 * wid + 1] = ford(mx);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = s_red[0], b = s_red[1];
        for (int w = 1; w < NTHR / 64; ++w) {
            a = min(a, s_red[2 * w]);
            b = max(b, s_red[2 * w + 1]);
        }
        atomicMin(&dst[0], a);
        atomicMax(&dst[1], b);
    }
}

// Extremes of the warped moving image per pose (blockIdx.y): ext[pose] = {min, max}.
template <typename T>
__global__ __launch_bounds__(256) void k_mi_extrema(FusedArgs a, int grid3, int src3, int* ext) {
    __shared__ int s_red[8];
    const int pose = blockIdx.y;
    const long long n = (long long)a.v.X * a.v.Y * a.v.Z;
    float mn = INFINITY, mx = -INFINITY;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const float v = warped_at<T>(a, grid3, src3, pose, i);
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    block_minmax<256>(mn, mx, s_red, ext + 2 * pose);
}

// Extremes of a resident image.
__global__ __launch_bounds__(256) void k_mi_extrema_img(const float* img, long long n, int* ext) {
    __shared__ int s_red[8];
    float mn = INFINITY, mx = -INFINITY;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        mn = fminf(mn, img[i]);
        mx = fmaxf(mx, img[i]);
    }
    block_minmax<256>(mn, mx, s_red, ext);
}

__global__ void k_mi_init(int* ext, int n_pairs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_pairs) {
        ext[2 * i] = 0x7fffffff;
        ext[2 * i + 1] = (int)0x80000000;
    }
}

// cu_lc2_3d.cu:882-887: bin of v; the scaling is formed in double as there, roundf()
// rounds its float conversion.
__device__ __forceinline__ int mi_bin(float v, float mn, float mx, int bins) {
    const double t = (double)(v - mn) * (bins - 1.0) / (double)(mx - mn);
    return (int)roundf((float)t);
}

// Joint histogram [pose][bin_i][bin_u] of (U, I) over chunks of at most 65535 voxels per
// work-group: privatised in LDS as 16-bit counters packed two to a word (128 KiB for
// 256 x 256 bins), then the non-zero bins are added to the pose's histogram.
template <typename T>
__global__ __launch_bounds__(512) void k_mi_joint(FusedArgs a, int grid3, int src3,
                                                  const float* stat, int stat_is_u,
                                                  const int* ext_mov, const int* ext_stat,
                                                  int bins_u, int bins_i, int chunk,
                                                  unsigned* hist) {
    extern __shared__ unsigned s_h[];
    const int pose = blockIdx.y;
    const int nb = bins_u * bins_i, nw = (nb + 1) / 2;
    for (int w = threadIdx.x; w < nw; w += 512) s_h[w] = 0u;
    const float mnm = fdec(ext_mov[2 * pose]), mxm = fdec(ext_mov[2 * pose + 1]);
    const float mns = fdec(ext_stat[0]), mxs = fdec(ext_stat[1]);
    const float mnu = stat_is_u ? mns : mnm, mxu = stat_is_u ? mxs : mxm;
    const float mni = stat_is_u ? mnm : mns, mxi = stat_is_u ? mxm : mxs;
    __syncthreads();
    const long long n = (long long)a.v.X * a.v.Y * a.v.Z;
    const long long i0 = (long long)blockIdx.x * chunk;
    const long long i1 = min(n, i0 + chunk);
    for (long long i = i0 + threadIdx.x; i < i1; i += 512) {
        const float vm = warped_at<T>(a, grid3, src3, pose, i);
        const float vs = stat[i];
        const int bu = mi_bin(stat_is_u ? vs : vm, mnu, mxu, bins_u);
        const int bi = mi_bin(stat_is_u ? vm : vs, mni, mxi, bins_i);
        if (bu >= 0 && bu < bins_u && bi >= 0 && bi < bins_i) {
            const int j = bu + bins_u * bi;
            atomicAdd(&s_h[j >> 1], (j & 1) ? 65536u : 1u);
        }
    }
    __syncthreads();
    unsigned* h = hist + (size_t)pose * nb;
    for (int w = threadIdx.x; w < nw; w += 512) {
        const unsigned c = s_h[w];
        if (c & 0xffffu) atomicAdd(&h[2 * w], c & 0xffffu);
        if ((c >> 16) && 2 * w + 1 < nb) atomicAdd(&h[2 * w + 1], c >> 16);
    }
}

// Entropies in FP64 from the joint histogram (its marginals are the single-image
// histograms: every finite value falls into a bin) and NMI per pose (blockIdx.x);
// 0 when an image is constant.
__global__ __launch_bounds__(256) void k_mi_entropy(const unsigned* hist, const int* ext_mov,
                                                    const int* ext_stat, int bins_u, int bins_i,
                                                    double n, double* out) {
    __shared__ double s_red[3][256];
    const int pose = blockIdx.x;
    const unsigned* h = hist + (size_t)pose * bins_u * bins_i;
    double hj = 0.0, hu = 0.0, hi = 0.0;
    for (int j = threadIdx.x; j < bins_u * bins_i; j += 256)
        if (h[j]) {
            const double p = h[j] / n;
            hj -= p * log(p);
        }
    for (int u = threadIdx.x; u < bins_u; u += 256) {
        unsigned long long c = 0;
        for (int i = 0; i < bins_i; ++i) c += h[u + bins_u * i];
        if (c) {
            const double p = c / n;
            hu -= p * log(p);
        }
    }
    for (int i = threadIdx.x; i < bins_i; i += 256) {
        unsigned long long c = 0;
        for (int u = 0; u < bins_u; ++u) c += h[u + bins_u * i];
        if (c) {
            const double p = c / n;
            hi -= p * log(p);
        }
    }
    s_red[0][threadIdx.x] = hj;
    s_red[1][threadIdx.x] = hu;
    s_red[2][threadIdx.x] = hi;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k)
            for (int q = 0; q < 3; ++q) s_red[q][threadIdx.x] += s_red[q][threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const bool flat = !(fdec(ext_mov[2 * pose + 1]) > fdec(ext_mov[2 * pose])) ||
                          !(fdec(ext_stat[1]) > fdec(ext_stat[0]));
        const double HJ = s_red[0][0], HU = s_red[1][0], HI = s_red[2][0];
        out[pose] = (flat || !(HJ > 0.0)) ? 0.0 : (HU + HI) / HJ;
    }
}

// Static ultrasound patch statistics (weight half of cu_lc2_3d.cu:181-287 /
// cu_lc2_2d.cu:88-180) per voxel in FP64: count of U > 1e-6, sum, sum of squares in a
// fixed x-fastest order; mean and centred energy, or -1 where w = 0.
__global__ void k_ustats(const float* img, VGeom v, int grid3, int r, double* mean, double* cyy) {
    const long long n = (long long)v.X * v.Y * v.Z;
    const int w = 2 * r + 1;
    const int N = grid3 ? w * w * w : w * w;
    const int rz = grid3 ? r : 0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % v.X), y = (int)((i / v.X) % v.Y), z = (int)(i / ((long long)v.X * v.Y));
        double m = 0.0, c = -1.0;
        const bool inside = x >= r && y >= r && x + r < v.X && y + r < v.Y &&
                            (!grid3 || (z >= r && z + r < v.Z));
        if (inside) {
            int cnt = 0;
            double s1 = 0.0, s2 = 0.0;
            for (int dz = -rz; dz <= rz; ++dz)
                for (int dy = -r; dy <= r; ++dy) {
                    const float* row = img + ((long long)(z + dz) * v.Y + (y + dy)) * v.X + x;
                    for (int dx = -r; dx <= r; ++dx) {
                        const double u = row[dx];
                        cnt += (u > kZeroThreshold) ? 1 : 0;
                        s1 += u;
                        s2 += u * u;
                    }
                }
            if (cnt == N) {
                const double cc = s2 - s1 * s1 * (1.0 / (double)N);
                if (cc * (1.0 / (double)(N - 1)) > kZeroThreshold) {
                    m = s1 * (1.0 / (double)N);
                    c = cc;
                }
            }
        }
        mean[i] = m;
        cyy[i] = c;
    }
}

__device__ __forceinline__ float at_img(const float* img, const VGeom& v, int x, int y, int z) {
    if (x < 0 || y < 0 || x >= v.X || y >= v.Y) return 0.0f;
    return img[(z * v.Y + y) * v.X + x];
}

// Separable Sobel triplet (A, Bx, By) of plane z at (x, y), zero padding.
__device__ __forceinline__ void sobel_triplet(const float* img, const VGeom& v, int x, int y,
                                              int z, float& A, float& Bx, float& By) {
    float s[3], d[3];
    for (int k = 0; k < 3; ++k) {
        const int yy = y + k - 1;
        const float l = at_img(img, v, x - 1, yy, z), c = at_img(img, v, x, yy, z),
                    r = at_img(img, v, x + 1, yy, z);
        d[k] = r - l;
        s[k] = (l + 2.0f * c) + r;
    }
    A = (s[0] + 2.0f * s[1]) + s[2];
    Bx = (d[0] + 2.0f * d[1]) + d[2];
    By = s[2] - s[0];
}

__global__ void k_sobel(const float* img, VGeom v, int grid3, float* out) {
    const long long n = (long long)v.X * v.Y * v.Z;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % v.X), y = (int)((i / v.X) % v.Y), z = (int)(i / ((long long)v.X * v.Y));
        float A, Bx, By;
        sobel_triplet(img, v, x, y, z, A, Bx, By);
        float gx, gy, gz;
        if (grid3) {
            float Am = 0.f, Bxm = 0.f, Bym = 0.f, Ap = 0.f, Bxp = 0.f, Byp = 0.f;
            if (z > 0) sobel_triplet(img, v, x, y, z - 1, Am, Bxm, Bym);
            if (z + 1 < v.Z) sobel_triplet(img, v, x, y, z + 1, Ap, Bxp, Byp);
            gx = (Bxm + 2.0f * Bx) + Bxp;
            gy = (Bym + 2.0f * By) + Byp;
            gz = Ap - Am;
        } else {
            gx = Bx;
            gy = By;
            gz = 0.0f;
        }
        out[i] = sqrtf(fmaf(gz, gz, fmaf(gy, gy, gx * gx)));
    }
}

__global__ void k_convert(const void* in, int in_dtype, void* out, int out_f16, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        float v;
        switch (in_dtype) {
            case 0: v = (float)static_cast<const unsigned char*>(in)[i]; break;
            case 1: v = (float)static_cast<const short*>(in)[i]; break;
            case 2: v = __half2float(static_cast<const __half*>(in)[i]); break;
            default: v = static_cast<const float*>(in)[i]; break;
        }
        if (out_f16) static_cast<__half*>(out)[i] = __float2half_rn(v);
        else static_cast<float*>(out)[i] = v;
    }
}

// The dynamic LDS layout of k_fused2d: nc FP64 products planes (7 for a moving
// high-resolution image, 9 plus an integer count plane for a moving ultrasound), the
// reduction scratch and the warped tile.
constexpr size_t lds_bytes_2d(int r, bool movi) {
    const size_t nc = movi ? 7 : 9;
    const size_t PX = kTile2X + 2 * r, PY = kTile2Y + 2 * r, IXE = PX + 2, IYE = PY + 2;
    const size_t PS = lds_stride_2d((int)PX);
    const size_t NI = movi ? (32 + PX / 2) * IYE : PX * PY;
    return sizeof(double) * (nc * PY * PS + 8) + sizeof(float) * NI +
           (movi ? 0 : sizeof(int) * PY * PS);
}
static_assert(lds_bytes_2d(kMaxR2, false) <= 160 * 1024, "2D fused kernel exceeds the 160 KiB of LDS");

// The launch helpers below either launch (occ == nullptr) or report, in *occ, how many
// work-groups of the kernel they would launch fit on one CU at once.
template <int R, typename T, bool P2>
hipError_t launch3(const FusedArgs& a, bool movi, dim3 grid, hipStream_t st, int* occ) {
    constexpr int nt = kTile3X * kTile3Y;
    if (occ) {
        return movi ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          occ, k_fused3d<R, true, T, kTile3X, kTile3Y, P2>, nt, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          occ, k_fused3d<R, false, T, kTile3X, kTile3Y, P2>, nt, 0);
    }
    if (movi) hipLaunchKernelGGL((k_fused3d<R, true, T, kTile3X, kTile3Y, P2>), grid, dim3(nt), 0, st, a);
    else hipLaunchKernelGGL((k_fused3d<R, false, T, kTile3X, kTile3Y, P2>), grid, dim3(nt), 0, st, a);
    return hipGetLastError();
}

template <typename T, bool P2>
hipError_t launch3r(const FusedArgs& a, bool movi, dim3 grid, hipStream_t st, int* occ) {
    switch (a.r) {
        case 1: return launch3<1, T, P2>(a, movi, grid, st, occ);
        case 2: return launch3<2, T, P2>(a, movi, grid, st, occ);
        case 3: return launch3<3, T, P2>(a, movi, grid, st, occ);
        case 4: return launch3<4, T, P2>(a, movi, grid, st, occ);
        default: return hipErrorInvalidValue;
    }
}

// The fused kernel for the moving image's storage type and spacing form (SGeom::pow2).
hipError_t launch3d(const FusedArgs& a, bool movi, dim3 grid, hipStream_t st, int* occ) {
    if (a.s.pow2)
        return a.src_f16 ? launch3r<__half, true>(a, movi, grid, st, occ)
                         : launch3r<float, true>(a, movi, grid, st, occ);
    return a.src_f16 ? launch3r<__half, false>(a, movi, grid, st, occ)
                     : launch3r<float, false>(a, movi, grid, st, occ);
}

template <bool MOVI, bool SRC3D, typename T, int RMAX, bool P2>
hipError_t launch2r(const FusedArgs& a, dim3 grid, hipStream_t st, int* occ) {
    const size_t lds = lds_bytes_2d(a.r, MOVI);
    // allow this variant's largest radius, once per process (client threads and a serving
    // worker may get here at the same time)
    static std::once_flag once;
    static hipError_t set_err = hipSuccess;
    std::call_once(once, [] {
        set_err = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fused2d<MOVI, SRC3D, T, RMAX, P2>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds_bytes_2d(RMAX, MOVI));
    });
    if (set_err != hipSuccess) return set_err;
    if (occ) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_fused2d<MOVI, SRC3D, T, RMAX, P2>,
                                                                  kThreads, lds);
    hipLaunchKernelGGL((k_fused2d<MOVI, SRC3D, T, RMAX, P2>), grid, dim3(kThreads), lds, st, a);
    return hipGetLastError();
}

// Radius classes of the 2D kernel: fewer registers for the common small patches.
template <bool MOVI, bool SRC3D, typename T, bool P2>
hipError_t launch2(const FusedArgs& a, dim3 grid, hipStream_t st, int* occ) {
    if (a.r <= 3) return launch2r<MOVI, SRC3D, T, 3, P2>(a, grid, st, occ);
    if (a.r <= 6) return launch2r<MOVI, SRC3D, T, 6, P2>(a, grid, st, occ);
    return launch2r<MOVI, SRC3D, T, kMaxR2, P2>(a, grid, st, occ);
}

template <typename T, bool P2>
hipError_t launch2t(const FusedArgs& a, bool movi, int src_ndim, dim3 grid, hipStream_t st, int* occ) {
    if (movi)
        return src_ndim == 3 ? launch2<true, true, T, P2>(a, grid, st, occ)
                             : launch2<true, false, T, P2>(a, grid, st, occ);
    return src_ndim == 3 ? launch2<false, true, T, P2>(a, grid, st, occ)
                         : launch2<false, false, T, P2>(a, grid, st, occ);
}

hipError_t launch2d(const FusedArgs& a, bool movi, int src_ndim, dim3 grid, hipStream_t st,
                    int* occ) {
    if (a.s.pow2)
        return a.src_f16 ? launch2t<__half, true>(a, movi, src_ndim, grid, st, occ)
                         : launch2t<float, true>(a, movi, src_ndim, grid, st, occ);
    return a.src_f16 ? launch2t<__half, false>(a, movi, src_ndim, grid, st, occ)
                     : launch2t<float, false>(a, movi, src_ndim, grid, st, occ);
}

}  // namespace

int fused_blocks_3d(const VGeom& v, int r, int zchunk, int* ntx, int* nty, int* nzc) {
    const int ex = v.X - 2 * r, ey = v.Y - 2 * r, ez = v.Z - 2 * r;
    if (ex <= 0 || ey <= 0 || ez <= 0 || zchunk <= 0) { *ntx = *nty = *nzc = 0; return 0; }
    *ntx = (ex + kTile3X - 1) / kTile3X;
    *nty = (ey + kTile3Y - 1) / kTile3Y;
    *nzc = (ez + zchunk - 1) / zchunk;
    return *ntx * *nty * *nzc;
}

int fused_partials_3d(const VGeom& v, int r) {
    int ntx, nty, nzc;
    if (fused_blocks_3d(v, r, kZBlock, &ntx, &nty, &nzc) <= 0) return 0;
    return ntx * nty * nzc * (kTile3X * kTile3Y / 64);  // one per wave
}

int fused_blocks_2d(const VGeom& v, int r, int* ntx, int* nty) {
    const int ex = v.X - 2 * r, ey = v.Y - 2 * r;
    if (ex <= 0 || ey <= 0) { *ntx = *nty = 0; return 0; }
    *ntx = (ex + kTile2X - 1) / kTile2X;
    *nty = (ey + kTile2Y - 1) / kTile2Y;
    return *ntx * *nty * v.Z;
}

int pose_run(const VGeom& v, int grid_ndim, int src_ndim, int mats_per_pose, int tiles,
             int n_poses) {
    // runs of poses per work-group while at least about two work-groups per CU slot remain
    constexpr int kMaxPoseRun = 8;
    constexpr long long kPoseRunWGs = 2048;
    if (grid_ndim != 2 || src_ndim != 2 || v.Z != 1 || mats_per_pose != 1) return 1;
    const long long tiles8 = (tiles + kXcds - 1) / kXcds * (long long)kXcds;
    return (int)std::max(1LL, std::min<long long>(kMaxPoseRun, tiles8 * n_poses / kPoseRunWGs));
}

hipError_t launch_fused(const FusedArgs& a, int grid_ndim, int src_ndim, int movi, int n_poses,
                        hipStream_t st) {
    if (a.tiles == 0 || n_poses <= 0) return hipSuccess;
    if (grid_ndim == 3 && (a.zchunk <= 0 || a.zchunk % kZBlock != 0)) return hipErrorInvalidValue;
    if (grid_ndim == 2 && a.zchunk <= 0) return hipErrorInvalidValue;
    FusedArgs b = a;
    b.npose = n_poses;
    const long long tiles8 = (a.tiles + kXcds - 1) / kXcds * (long long)kXcds;
    b.prun = pose_run(a.v, grid_ndim, src_ndim, a.mats_per_pose, a.tiles, n_poses);
    const long long runs = (n_poses + b.prun - 1) / b.prun;
    if (tiles8 * runs > 0x7fffffffLL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(tiles8 * runs));
    if (grid_ndim == 3) return launch3d(b, movi, grid, st, nullptr);
    if (b.r < 1 || b.r > kMaxR2) return hipErrorInvalidValue;
    return launch2d(b, movi, src_ndim, grid, st, nullptr);
}

hipError_t fused_resident_per_cu(const FusedArgs& a, int grid_ndim, int src_ndim, int movi,
                                 int* per_cu) {
    *per_cu = 0;
    if (grid_ndim == 3) return launch3d(a, movi, dim3(1), nullptr, per_cu);
    if (a.r < 1 || a.r > kMaxR2) return hipErrorInvalidValue;
    return launch2d(a, movi, src_ndim, dim3(1), nullptr, per_cu);
}

static dim3 grid_for(long long n) {
    long long b = (n + 255) / 256;
    if (b > 65535LL * 16) b = 65535LL * 16;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

int sep_blocks(const VGeom& v, int r, int grid_ndim) {
    const long long ex = v.X - 2 * r, ey = v.Y - 2 * r, ez = grid_ndim == 3 ? v.Z - 2 * r : v.Z;
    if (ex <= 0 || ey <= 0 || ez <= 0) return 0;
    return (int)((ex * ey * ez + 255) / 256);
}

hipError_t launch_sep(const FusedArgs& a, int grid_ndim, int src_ndim, int movi, int n_poses,
                      hipStream_t st) {
    const VGeom& v = a.v;
    const int W = 2 * a.r + 1;
    const int nb = sep_blocks(v, a.r, grid_ndim);
    if (nb == 0 || n_poses <= 0) return hipSuccess;
    const long long V = (long long)v.X * v.Y * v.Z;
    const int nc = movi ? 7 : 10;
    for (int pose = 0; pose < n_poses; ++pose) {
        FusedArgs b = a;
        b.mats = a.mats + (size_t)pose * a.mats_per_pose * 12;
        // the moving image on the grid; its |Sobel| when it is the high-resolution image
        hipError_t e = launch_warp_moving(b, grid_ndim, src_ndim, a.sep_img, st);
        if (e != hipSuccess) return e;
        if (movi && (e = launch_sobel(a.sep_img, v, grid_ndim, a.sep_grad, st)) != hipSuccess) return e;
        const float* I = movi ? a.sep_img : a.stat_img;
        const float* G = movi ? a.sep_grad : a.stat_grad;
        const float* U = movi ? a.stat_img : a.sep_img;
        const int passes = grid_ndim == 3 ? 3 : 2;
        for (int axis = 0; axis < passes; ++axis) {
            const long long lines = axis == 0 ? (long long)v.Y * v.Z
                                  : axis == 1 ? (long long)(v.X - W + 1) * v.Z
                                              : (long long)(v.X - W + 1) * (v.Y - W + 1);
            const dim3 grid = grid_for(lines * nc);
            const double* src = axis == 1 ? a.sep_a : a.sep_b;
            double* dst = axis == 1 ? a.sep_b : a.sep_a;
            if (movi)
                hipLaunchKernelGGL(k_sep_pass<true>, grid, dim3(256), 0, st, I, G, U, src, dst, v, W, axis);
            else
                hipLaunchKernelGGL(k_sep_pass<false>, grid, dim3(256), 0, st, I, G, U, src, dst, v, W, axis);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        const double* S = passes == 3 ? a.sep_a : a.sep_b;
        double* part = a.partials + (size_t)pose * a.blocks_per_pose * 2;
        if (movi)
            hipLaunchKernelGGL(k_sep_combine<true>, dim3(nb), dim3(256), 0, st, S, a.ustat_mean,
                               a.ustat_cyy, v, a.r, grid_ndim == 3 ? 1 : 0, part, a.wmap, a.mmap);
        else
            hipLaunchKernelGGL(k_sep_combine<false>, dim3(nb), dim3(256), 0, st, S, a.ustat_mean,
                               a.ustat_cyy, v, a.r, grid_ndim == 3 ? 1 : 0, part, a.wmap, a.mmap);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    (void)V;
    return hipSuccess;
}

hipError_t launch_reduce(const double* partials, int blocks_per_pose, int n_poses, double* out,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_reduce, dim3(n_poses), dim3(256), 0, st, partials, blocks_per_pose, out);
    return hipGetLastError();
}


hipError_t launch_warp_static(const void* fixed, int fixed_f16, const SGeom& fg, const VGeom& v,
                              int grid_ndim, int src_ndim, float* out, hipStream_t st) {
    const long long n = (long long)v.X * v.Y * v.Z;
    // A 2D grid over a 3D fixed "volume" of one slice (2D->3D) samples trilinearly with
    // Z = 1, which equals the bilinear per-frame path bit for bit.
    const int mode = (grid_ndim == 3) ? 0 : 1;
    (void)src_ndim;
    if (fixed_f16)
        hipLaunchKernelGGL(k_warp_static<__half>, grid_for(n), dim3(256), 0, st,
                           static_cast<const __half*>(fixed), fg, v, mode, out);
    else
        hipLaunchKernelGGL(k_warp_static<float>, grid_for(n), dim3(256), 0, st,
                           static_cast<const float*>(fixed), fg, v, mode, out);
    return hipGetLastError();
}

hipError_t launch_warp_moving(const FusedArgs& a, int grid_ndim, int src_ndim, float* out,
                              hipStream_t st) {
    const long long n = (long long)a.v.X * a.v.Y * a.v.Z;
    if (a.src_f16)
        hipLaunchKernelGGL(k_warp_moving<__half>, grid_for(n), dim3(256), 0, st, a,
                           grid_ndim == 3 ? 1 : 0, src_ndim == 3 ? 1 : 0, out);
    else
        hipLaunchKernelGGL(k_warp_moving<float>, grid_for(n), dim3(256), 0, st, a,
                           grid_ndim == 3 ? 1 : 0, src_ndim == 3 ? 1 : 0, out);
    return hipGetLastError();
}

hipError_t launch_ustats(const float* img, const VGeom& v, int grid_ndim, int r, double* mean,
                         double* cyy, hipStream_t st) {
    const long long n = (long long)v.X * v.Y * v.Z;
    hipLaunchKernelGGL(k_ustats, grid_for(n), dim3(256), 0, st, img, v, grid_ndim == 3 ? 1 : 0, r,
                       mean, cyy);
    return hipGetLastError();
}

hipError_t launch_sobel(const float* img, const VGeom& v, int grid_ndim, float* out,
                        hipStream_t st) {
    const long long n = (long long)v.X * v.Y * v.Z;
    hipLaunchKernelGGL(k_sobel, grid_for(n), dim3(256), 0, st, img, v, grid_ndim == 3 ? 1 : 0,
                       out);
    return hipGetLastError();
}

hipError_t launch_mi(const FusedArgs& a, int grid_ndim, int src_ndim, int movi, int n_poses,
                     int bins_u, int bins_i, const float* stat, int* ext_stat, int stat_ready,
                     int* ext_mov, unsigned* hist, double* out, hipStream_t st) {
    if (bins_u < 2 || bins_i < 2 || bins_u > kMiMaxBins || bins_i > kMiMaxBins || n_poses <= 0)
        return hipErrorInvalidValue;
    const long long n = (long long)a.v.X * a.v.Y * a.v.Z;
    const int g3 = grid_ndim == 3 ? 1 : 0, s3 = src_ndim == 3 ? 1 : 0;
    const int nb = bins_u * bins_i;
    hipLaunchKernelGGL(k_mi_init, dim3((n_poses + 1 + 255) / 256), dim3(256), 0, st, ext_mov,
                       n_poses);
    if (!stat_ready) {
        hipLaunchKernelGGL(k_mi_init, dim3(1), dim3(256), 0, st, ext_stat, 1);
        hipLaunchKernelGGL(k_mi_extrema_img, grid_for(n), dim3(256), 0, st, stat, n, ext_stat);
    }
    const dim3 ge((unsigned)std::min<long long>((n + 255) / 256, 1024), (unsigned)n_poses);
    if (a.src_f16)
        hipLaunchKernelGGL(k_mi_extrema<__half>, ge, dim3(256), 0, st, a, g3, s3, ext_mov);
    else
        hipLaunchKernelGGL(k_mi_extrema<float>, ge, dim3(256), 0, st, a, g3, s3, ext_mov);
    hipError_t e = hipMemsetAsync(hist, 0, (size_t)n_poses * nb * sizeof(unsigned), st);
    if (e != hipSuccess) return e;
    const int chunk = 32768;  // voxels per work-group, below the 16-bit counter limit
    const dim3 gj((unsigned)((n + chunk - 1) / chunk), (unsigned)n_poses);
    const size_t lds = (size_t)((nb + 1) / 2) * sizeof(unsigned);
    static std::once_flag once;
    static hipError_t set_err = hipSuccess;
    std::call_once(once, [] {
        for (const void* f : {reinterpret_cast<const void*>(&k_mi_joint<float>),
                              reinterpret_cast<const void*>(&k_mi_joint<__half>)}) {
            set_err = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)((kMiMaxBins * kMiMaxBins / 2) * sizeof(unsigned)));
            if (set_err != hipSuccess) return;
        }
    });
    if (set_err != hipSuccess) return set_err;
    const int stat_is_u = movi ? 1 : 0;
    if (a.src_f16)
        hipLaunchKernelGGL(k_mi_joint<__half>, gj, dim3(512), lds, st, a, g3, s3, stat, stat_is_u,
                           ext_mov, ext_stat, bins_u, bins_i, chunk, hist);
    else
        hipLaunchKernelGGL(k_mi_joint<float>, gj, dim3(512), lds, st, a, g3, s3, stat, stat_is_u,
                           ext_mov, ext_stat, bins_u, bins_i, chunk, hist);
    hipLaunchKernelGGL(k_mi_entropy, dim3(n_poses), dim3(256), 0, st, hist, ext_mov, ext_stat,
                       bins_u, bins_i, (double)n, out);
    return hipGetLastError();
}

// The scratch array of lc2e_eval_batch_noalloc. In host code its name is only a stand-in
// for the symbol: kernels get the device address that scratch_address() looks up.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

hipError_t scratch_address(int device, void** out) {
    static std::mutex mu;
    static std::map<int, void*> addr;  // device -> address of its copy of the array
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(device);
    if (it != addr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    addr[device] = ptr;
    *out = ptr;
    return hipSuccess;
}

hipError_t launch_convert(const void* d_in, int in_dtype, void* d_out, int out_f16, long long n,
                          hipStream_t st) {
    hipLaunchKernelGGL(k_convert, grid_for(n), dim3(256), 0, st, d_in, in_dtype, d_out, out_f16, n);
    return hipGetLastError();
}

}  // namespace lc2e
