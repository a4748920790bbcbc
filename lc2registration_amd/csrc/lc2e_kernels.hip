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
// 3D structure (k_fused3d): a 256-thread work-group owns a 16x16 column of
// output voxels over `zchunk` planes and marches along z. Per input plane:
// warp the plane (+ Sobel halo) into LDS, separable Sobel row pass in LDS,
// column pass in registers (3-plane ring per position), FP64 products into LDS,
// sliding x then y box sums in LDS, and a (2R+1)-plane ring of plane sums in
// registers gives the z box sum; complete patch sums are combined in FP64 and
// accumulated per thread. No intermediate volume touches HBM.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "lc2e_kernels.hpp"

namespace lc2e {

namespace {

constexpr int NCH = 9;  // I, G, U, II, GG, IG, UU, UI, UG
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

__device__ __forceinline__ float load_src(const void* p, int f16, int i) {
    if (f16) return __half2float(static_cast<const __half*>(p)[i]);
    return static_cast<const float*>(p)[i];
}

// Interpolate_3Dkernel sample (interpolator.cu:76-174): L -> M*L -> voxel -> trilinear
// z, y, x with floor/ceil corners; 0 outside [0, dim-1].
__device__ __forceinline__ float sample3(const void* src, int f16, const SGeom& s, const float* m,
                                         float lx, float ly, float lz) {
    const float px = fmaf(m[2], lz, fmaf(m[1], ly, m[0] * lx)) + m[3];
    const float py = fmaf(m[6], lz, fmaf(m[5], ly, m[4] * lx)) + m[7];
    const float pz = fmaf(m[10], lz, fmaf(m[9], ly, m[8] * lx)) + m[11];
    const float vx = (px - s.ox) / s.sx;
    const float vy = (py - s.oy) / s.sy;
    const float vz = (pz - s.oz) / s.sz;
    if (!(vx >= 0.0f && vx <= (float)(s.X - 1) && vy >= 0.0f && vy <= (float)(s.Y - 1) &&
          vz >= 0.0f && vz <= (float)(s.Z - 1)))
        return 0.0f;
    const int x0 = (int)floorf(vx), x1 = (int)ceilf(vx);
    const int y0 = (int)floorf(vy), y1 = (int)ceilf(vy);
    const int z0 = (int)floorf(vz), z1 = (int)ceilf(vz);
    const float rx = vx - (float)x0, ry = vy - (float)y0, rz = vz - (float)z0;
    const int sxy = s.X * s.Y;
    const int r00 = y0 * s.X + z0 * sxy, r01 = y0 * s.X + z1 * sxy;
    const int r10 = y1 * s.X + z0 * sxy, r11 = y1 * s.X + z1 * sxy;
    float a, b;
    a = load_src(src, f16, x0 + r00); b = load_src(src, f16, x0 + r01);
    const float lu = fmaf(rz, b - a, a);
    a = load_src(src, f16, x1 + r00); b = load_src(src, f16, x1 + r01);
    const float ru = fmaf(rz, b - a, a);
    a = load_src(src, f16, x1 + r10); b = load_src(src, f16, x1 + r11);
    const float rb = fmaf(rz, b - a, a);
    a = load_src(src, f16, x0 + r10); b = load_src(src, f16, x0 + r11);
    const float lb = fmaf(rz, b - a, a);
    const float yl = fmaf(ry, lb - lu, lu);
    const float yr = fmaf(ry, rb - ru, ru);
    return fmaf(rx, yr - yl, yl);
}

// Interpolate_2Dkernel sample (interpolator.cu:186-230): bilinear, y then x. The 2x3
// matrix sits in the 3x4 layout as m[0], m[1], m[3] / m[4], m[5], m[7].
__device__ __forceinline__ float sample2(const void* src, int f16, const SGeom& s, const float* m,
                                         float lx, float ly) {
    const float px = fmaf(m[1], ly, m[0] * lx) + m[3];
    const float py = fmaf(m[5], ly, m[4] * lx) + m[7];
    const float vx = (px - s.ox) / s.sx;
    const float vy = (py - s.oy) / s.sy;
    if (!(vx >= 0.0f && vx <= (float)(s.X - 1) && vy >= 0.0f && vy <= (float)(s.Y - 1)))
        return 0.0f;
    const int x0 = (int)floorf(vx), x1 = (int)ceilf(vx);
    const int y0 = (int)floorf(vy), y1 = (int)ceilf(vy);
    const float ry = vy - floorf(vy);
    float a = load_src(src, f16, x0 + y0 * s.X), b = load_src(src, f16, x0 + y1 * s.X);
    const float xl = fmaf(ry, b - a, a);
    a = load_src(src, f16, x1 + y0 * s.X);
    b = load_src(src, f16, x1 + y1 * s.X);
    const float xr = fmaf(ry, b - a, a);
    const float rx = vx - floorf(vx);
    return fmaf(rx, xr - xl, xl);
}

// s^T C^+ s for the two-column centred fit with the rank rule shared with the oracle.
__device__ __forceinline__ double explained_var(double a, double b, double c, double p, double q,
                                                double rawa, double rawb) {
    const bool iconst = !(a > kTauConst * rawa);
    const bool gconst = !(b > kTauConst * rawb);
    if (iconst) { a = 0.0; c = 0.0; p = 0.0; }
    if (gconst) { b = 0.0; c = 0.0; q = 0.0; }
    const double det = a * b - c * c;
    if (a > 0.0 && b > 0.0 && det > kTauColl * a * b)
        return (b * p * p - 2.0 * c * p * q + a * q * q) / det;
    const double tr = a + b;
    if (tr > 0.0) {
        const double u = sqrt(a) * p + (c < 0.0 ? -1.0 : 1.0) * sqrt(b) * q;
        return (u * u) / (tr * tr);
    }
    return 0.0;
}

// Patch sums -> (w, measure) with the reference's thresholds
// (cu_lc2_3d.cu:258-284 weight, :596-611 measure). Returns false if w = 0.
__device__ __forceinline__ bool combine(const double* S, int cnt, int N, double& w_out,
                                        double& m_out) {
    if (cnt != N) return false;
    const double n = (double)N;
    const double sy = S[2];
    const double cyy = S[6] - sy * sy / n;
    const double var_u = cyy / (n - 1.0);
    if (!(var_u > kZeroThreshold)) return false;
    const double w = sqrt(var_u);
    const double si = S[0], sg = S[1];
    const double cii = S[3] - si * si / n;
    const double cgg = S[4] - sg * sg / n;
    const double cig = S[5] - si * sg / n;
    const double cyi = S[7] - sy * si / n;
    const double cyg = S[8] - sy * sg / n;
    double e = explained_var(cii, cgg, cig, cyi, cyg, S[3], S[4]);
    e = fmin(fmax(e, 0.0), cyy);
    const double var_res = (cyy - e) / (n - 1.0);
    double meas = w - var_res / w;
    if (!(meas > 0.0) || isinf(meas)) meas = 0.0;
    w_out = w;
    m_out = meas;
    return true;
}

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
        for (int i = 0; i < kThreads / 64; ++i) {
            m += s_red[2 * i];
            w += s_red[2 * i + 1];
        }
        dst[0] = m;
        dst[1] = w;
    }
}

// Products of one position into the channel planes of the LDS buffer.
__device__ __forceinline__ void store_products(double* s_prod, int* s_cnt, int np, int i, float I,
                                               float G, float U) {
    const double di = I, dg = G, du = U;
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
}

// Sliding box sums along x (rows of width px -> kTile outputs) then along y.
// s_prod[c][py*px] -> s_xp[c][py*kTile] -> s_out[c][kTile*kTile] (s_out may alias s_prod).
__device__ __forceinline__ void box_xy(const double* s_prod, const int* s_cnt, double* s_xp,
                                       int* s_xcnt, double* s_out, int* s_ocnt, int px, int w,
                                       int tid) {
    const int np = px * px;
    // x pass: task = (channel, row)
    for (int t = tid; t < (NCH + 1) * px; t += kThreads) {
        const int c = t / px, row = t - c * px;
        if (c < NCH) {
            const double* in = s_prod + c * np + row * px;
            double* out = s_xp + c * kTile * px + row * kTile;
            double acc = 0.0;
            for (int k = 0; k < w; ++k) acc += in[k];
            out[0] = acc;
            for (int xo = 1; xo < kTile; ++xo) {
                acc += in[xo + w - 1] - in[xo - 1];
                out[xo] = acc;
            }
        } else {
            const int* in = s_cnt + row * px;
            int* out = s_xcnt + row * kTile;
            int acc = 0;
            for (int k = 0; k < w; ++k) acc += in[k];
            out[0] = acc;
            for (int xo = 1; xo < kTile; ++xo) {
                acc += in[xo + w - 1] - in[xo - 1];
                out[xo] = acc;
            }
        }
    }
    __syncthreads();
    // y pass: task = (channel, column)
    for (int t = tid; t < (NCH + 1) * kTile; t += kThreads) {
        const int c = t / kTile, col = t - c * kTile;
        if (c < NCH) {
            const double* in = s_xp + c * kTile * px + col;
            double* out = s_out + c * kTile * kTile + col;
            double acc = 0.0;
            for (int k = 0; k < w; ++k) acc += in[k * kTile];
            out[0] = acc;
            for (int yo = 1; yo < kTile; ++yo) {
                acc += in[(yo + w - 1) * kTile] - in[(yo - 1) * kTile];
                out[yo * kTile] = acc;
            }
        } else {
            const int* in = s_xcnt + col;
            int* out = s_ocnt + col;
            int acc = 0;
            for (int k = 0; k < w; ++k) acc += in[k * kTile];
            out[0] = acc;
            for (int yo = 1; yo < kTile; ++yo) {
                acc += in[(yo + w - 1) * kTile] - in[(yo - 1) * kTile];
                out[yo * kTile] = acc;
            }
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// 3D fused kernel
// ---------------------------------------------------------------------------
template <int R, bool MOVI>
__global__ __launch_bounds__(kThreads, 2) void k_fused3d(FusedArgs a) {
    constexpr int W = 2 * R + 1;
    constexpr int PX = kTile + 2 * R;  // patch-support region edge
    constexpr int NP = PX * PX;
    constexpr int IXE = PX + 2;        // Sobel input region edge
    constexpr int NI = MOVI ? IXE * IXE : NP;
    constexpr int PPT = (NP + kThreads - 1) / kThreads;
    constexpr int N3 = W * W * W;
    constexpr int LAG = MOVI ? 2 : 0;

    __shared__ float s_in[NI];
    __shared__ float s_s[MOVI ? PX * IXE : 1];
    __shared__ float s_d[MOVI ? PX * IXE : 1];
    __shared__ double s_prod[NCH * NP];
    __shared__ int s_cnt[NP];
    __shared__ double s_xp[NCH * kTile * PX];
    __shared__ int s_xcnt[kTile * PX];
    __shared__ double s_red[2 * kThreads / 64];

    const int tid = threadIdx.x;
    const int tx = blockIdx.x % a.ntx, ty = blockIdx.x / a.ntx;
    const int x0 = R + tx * kTile, y0 = R + ty * kTile;
    const int z0 = R + blockIdx.y * a.zchunk;
    const int zend = min(z0 + a.zchunk, a.v.Z - R);
    const int nout = zend - z0;
    const int pose = blockIdx.z;
    const int X = a.v.X, Y = a.v.Y, Z = a.v.Z;
    float m[12];
    {
        const float* mp = a.mats + (size_t)pose * a.mats_per_pose * 12;
#pragma unroll
        for (int k = 0; k < 12; ++k) m[k] = mp[k];
    }

    float tA[PPT][3], tBx[PPT][3], tBy[PPT][3], Iprev[PPT], Icur[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
#pragma unroll
        for (int k = 0; k < 3; ++k) { tA[j][k] = 0.f; tBx[j][k] = 0.f; tBy[j][k] = 0.f; }
        Iprev[j] = 0.f;
        Icur[j] = 0.f;
    }
    double ring[W][NCH];
    int ringc[W];
    double S[NCH];
    int SC = 0;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        ringc[k] = 0;
#pragma unroll
        for (int c = 0; c < NCH; ++c) ring[k][c] = 0.0;
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) S[c] = 0.0;

    const int ox = tid % kTile, oy = tid / kTile;
    const bool col_valid = (x0 + ox < X - R) && (y0 + oy < Y - R);
    double acc_m = 0.0, acc_w = 0.0;

    const int nsteps = nout + 2 * R + LAG;
    for (int s0 = 0; s0 < nsteps; s0 += W) {
        // W consecutive steps with compile-time ring slots (k), so the z ring stays in VGPRs.
        Unroll<0, W>::run([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const int s = s0 + k;
            if (s >= nsteps) return;  // block-uniform
            // 1. warp one plane into LDS
            if (MOVI) {
                const int zi = z0 - R - 1 + s;
                const bool zin = zi >= 0 && zi < Z;
                const float lz = fmaf((float)zi, a.v.sz, a.v.oz);
                for (int i = tid; i < NI; i += kThreads) {
                    const int lx = i % IXE, ly = i / IXE;
                    const int gx = x0 - R - 1 + lx, gy = y0 - R - 1 + ly;
                    float val = 0.0f;
                    if (zin && gx >= 0 && gx < X && gy >= 0 && gy < Y)
                        val = sample3(a.src, a.src_f16, a.s, m, fmaf((float)gx, a.v.sx, a.v.ox),
                                      fmaf((float)gy, a.v.sy, a.v.oy), lz);
                    s_in[i] = val;
                }
            } else {
                const int zp = z0 - R + s;
                const float lz = fmaf((float)zp, a.v.sz, a.v.oz);
                for (int i = tid; i < NI; i += kThreads) {
                    const int lx = i % PX, ly = i / PX;
                    const int gx = x0 - R + lx, gy = y0 - R + ly;
                    float val = 0.0f;
                    if (gx < X && gy < Y)
                        val = sample3(a.src, a.src_f16, a.s, m, fmaf((float)gx, a.v.sx, a.v.ox),
                                      fmaf((float)gy, a.v.sy, a.v.oy), lz);
                    s_in[i] = val;
                }
            }
            __syncthreads();
            // 2. Sobel row pass (smooth s and difference d along x)
            if (MOVI) {
                for (int i = tid; i < PX * IXE; i += kThreads) {
                    const int lx = i % PX, ly = i / PX;
                    const float* row = s_in + ly * IXE + lx;
                    const float l = row[0], c = row[1], r = row[2];
                    s_d[i] = r - l;
                    s_s[i] = (l + 2.0f * c) + r;
                }
                __syncthreads();
            }
            // 3. column pass into the 3-plane ring, |grad| and products of plane zp
            const int q = s - LAG;
            const int zp = z0 - R + q;
#pragma unroll
            for (int j = 0; j < PPT; ++j) {
                const int i = tid + j * kThreads;
                if (i < NP) {
                    const int lx = i % PX, ly = i / PX;
                    if (MOVI) {
                        const float s0v = s_s[ly * PX + lx], s1v = s_s[(ly + 1) * PX + lx],
                                    s2v = s_s[(ly + 2) * PX + lx];
                        const float d0v = s_d[ly * PX + lx], d1v = s_d[(ly + 1) * PX + lx],
                                    d2v = s_d[(ly + 2) * PX + lx];
                        tA[j][0] = tA[j][1]; tA[j][1] = tA[j][2];
                        tBx[j][0] = tBx[j][1]; tBx[j][1] = tBx[j][2];
                        tBy[j][0] = tBy[j][1]; tBy[j][1] = tBy[j][2];
                        tA[j][2] = (s0v + 2.0f * s1v) + s2v;
                        tBx[j][2] = (d0v + 2.0f * d1v) + d2v;
                        tBy[j][2] = s2v - s0v;
                        Iprev[j] = Icur[j];
                        Icur[j] = s_in[(ly + 1) * IXE + lx + 1];
                    }
                    if (q >= 0) {
                        const int gx = x0 - R + lx, gy = y0 - R + ly;
                        const bool in = gx < X && gy < Y;
                        const int vidx = (zp * Y + gy) * X + gx;
                        float I, G, U;
                        if (MOVI) {
                            const float gxv = (tBx[j][0] + 2.0f * tBx[j][1]) + tBx[j][2];
                            const float gyv = (tBy[j][0] + 2.0f * tBy[j][1]) + tBy[j][2];
                            const float gzv = tA[j][2] - tA[j][0];
                            G = sqrtf(fmaf(gzv, gzv, fmaf(gyv, gyv, gxv * gxv)));
                            I = Iprev[j];
                            U = in ? a.stat_img[vidx] : 0.0f;
                        } else {
                            U = s_in[i];
                            I = in ? a.stat_img[vidx] : 0.0f;
                            G = in ? a.stat_grad[vidx] : 0.0f;
                        }
                        store_products(s_prod, s_cnt, NP, i, I, G, U);
                    }
                }
            }
            __syncthreads();  // s_in / s_s / s_d reads done; products visible
            if (q < 0) return;  // block-uniform: Sobel ring warm-up only
            // 4. in-plane box sums
            box_xy(s_prod, s_cnt, s_xp, s_xcnt, s_prod, s_cnt, PX, W, tid);
            // 5. z ring of plane sums, running patch sums, combine
            constexpr int slot = (k - LAG + W) % W;
            const int cc = s_cnt[tid];
            SC += cc - ringc[slot];
            ringc[slot] = cc;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const double v = s_prod[c * kTile * kTile + tid];
                S[c] += v - ring[slot][c];
                ring[slot][c] = v;
            }
            if (q >= 2 * R && col_valid) {
                double w, meas;
                if (combine(S, SC, N3, w, meas)) {
                    acc_m += meas;
                    acc_w += w;
                    if (a.wmap) {
                        const int vo = ((zp - R) * Y + (y0 + oy)) * X + (x0 + ox);
                        a.wmap[vo] = w;
                        a.mmap[vo] = meas;
                    }
                }
            }
        });
    }
    double* dst = a.partials + ((size_t)pose * a.blocks_per_pose +
                                (size_t)blockIdx.y * (a.ntx * a.nty) + blockIdx.x) * 2;
    block_reduce_store(acc_m, acc_w, s_red, dst);
}

// ---------------------------------------------------------------------------
// 2D fused kernel: one 16x16 output tile of one frame (plane) of one pose.
// Dynamic LDS sized for the runtime patch radius.
// ---------------------------------------------------------------------------
template <bool MOVI, bool SRC3D>
__global__ __launch_bounds__(kThreads) void k_fused2d(FusedArgs a) {
    extern __shared__ double s_dyn[];
    const int R = a.r, W = 2 * R + 1, PX = kTile + 2 * R, NP = PX * PX, IXE = PX + 2;
    const int NI = MOVI ? IXE * IXE : NP;
    double* s_prod = s_dyn;                                   // NCH*NP
    double* s_xp = s_prod + NCH * NP;                         // NCH*kTile*PX
    double* s_red = s_xp + NCH * kTile * PX;                  // 8
    float* s_in = reinterpret_cast<float*>(s_red + 8);        // NI
    float* s_s = s_in + NI;                                   // PX*IXE
    float* s_d = s_s + PX * IXE;                              // PX*IXE
    int* s_cnt = reinterpret_cast<int*>(s_d + PX * IXE);      // NP
    int* s_xcnt = s_cnt + NP;                                 // kTile*PX

    const int tid = threadIdx.x;
    const int tx = blockIdx.x % a.ntx, ty = blockIdx.x / a.ntx;
    const int x0 = R + tx * kTile, y0 = R + ty * kTile;
    const int frame = blockIdx.y;
    const int pose = blockIdx.z;
    const int X = a.v.X, Y = a.v.Y;
    const int plane = frame * X * Y;
    float m[12];
    {
        const float* mp =
            a.mats + ((size_t)pose * a.mats_per_pose + (a.mats_per_pose > 1 ? frame : 0)) * 12;
#pragma unroll
        for (int k = 0; k < 12; ++k) m[k] = mp[k];
    }
    const float lz = a.v.oz;
    const void* src = a.src;
    // 1. warp
    if (MOVI) {
        for (int i = tid; i < NI; i += kThreads) {
            const int lx = i % IXE, ly = i / IXE;
            const int gx = x0 - R - 1 + lx, gy = y0 - R - 1 + ly;
            float val = 0.0f;
            if (gx >= 0 && gx < X && gy >= 0 && gy < Y) {
                const float Lx = fmaf((float)gx, a.v.sx, a.v.ox), Ly = fmaf((float)gy, a.v.sy, a.v.oy);
                val = SRC3D ? sample3(src, a.src_f16, a.s, m, Lx, Ly, lz)
                            : sample2(src, a.src_f16, a.s, m, Lx, Ly);
            }
            s_in[i] = val;
        }
    } else {
        for (int i = tid; i < NI; i += kThreads) {
            const int lx = i % PX, ly = i / PX;
            const int gx = x0 - R + lx, gy = y0 - R + ly;
            float val = 0.0f;
            if (gx < X && gy < Y) {
                const float Lx = fmaf((float)gx, a.v.sx, a.v.ox), Ly = fmaf((float)gy, a.v.sy, a.v.oy);
                val = SRC3D ? sample3(src, a.src_f16, a.s, m, Lx, Ly, lz)
                            : sample2(src, a.src_f16, a.s, m, Lx, Ly);
            }
            s_in[i] = val;
        }
    }
    __syncthreads();
    if (MOVI) {
        for (int i = tid; i < PX * IXE; i += kThreads) {
            const int lx = i % PX, ly = i / PX;
            const float* row = s_in + ly * IXE + lx;
            const float l = row[0], c = row[1], r = row[2];
            s_d[i] = r - l;
            s_s[i] = (l + 2.0f * c) + r;
        }
        __syncthreads();
    }
    for (int i = tid; i < NP; i += kThreads) {
        const int lx = i % PX, ly = i / PX;
        const int gx = x0 - R + lx, gy = y0 - R + ly;
        const bool in = gx < X && gy < Y;
        const int vidx = plane + gy * X + gx;
        float I, G, U;
        if (MOVI) {
            const float s0v = s_s[ly * PX + lx], s2v = s_s[(ly + 2) * PX + lx];
            const float d0v = s_d[ly * PX + lx], d1v = s_d[(ly + 1) * PX + lx],
                        d2v = s_d[(ly + 2) * PX + lx];
            const float gxv = (d0v + 2.0f * d1v) + d2v;
            const float gyv = s2v - s0v;
            G = sqrtf(fmaf(gyv, gyv, gxv * gxv));
            I = s_in[(ly + 1) * IXE + lx + 1];
            U = in ? a.stat_img[vidx] : 0.0f;
        } else {
            U = s_in[i];
            I = in ? a.stat_img[vidx] : 0.0f;
            G = in ? a.stat_grad[vidx] : 0.0f;
        }
        store_products(s_prod, s_cnt, NP, i, I, G, U);
    }
    __syncthreads();
    box_xy(s_prod, s_cnt, s_xp, s_xcnt, s_prod, s_cnt, PX, W, tid);
    const int ox = tid % kTile, oy = tid / kTile;
    double acc_m = 0.0, acc_w = 0.0;
    if ((x0 + ox < X - R) && (y0 + oy < Y - R)) {
        double S[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) S[c] = s_prod[c * kTile * kTile + tid];
        double w, meas;
        if (combine(S, s_cnt[tid], W * W, w, meas)) {
            acc_m = meas;
            acc_w = w;
            if (a.wmap) {
                const int vo = plane + (y0 + oy) * X + (x0 + ox);
                a.wmap[vo] = w;
                a.mmap[vo] = meas;
            }
        }
    }
    double* dst = a.partials + ((size_t)pose * a.blocks_per_pose +
                                (size_t)frame * (a.ntx * a.nty) + blockIdx.x) * 2;
    block_reduce_store(acc_m, acc_w, s_red, dst);
}

// Per-pose fixed-order reduction of the work-group partials -> (sum m, sum w, LC2)
// (GetSimilarityMetric, cu_lc2_3d.cu:1181-1191; sums in FP64, NaN/Inf -> 0).
__global__ __launch_bounds__(256) void k_reduce(const double* __restrict__ partials, int nb,
                                                double* __restrict__ out) {
    __shared__ double sm[256], sw[256];
    const int pose = blockIdx.x;
    const double* p = partials + (size_t)pose * nb * 2;
    double m = 0.0, w = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) {
        m += p[2 * i];
        w += p[2 * i + 1];
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

__global__ void k_warp_static(const void* fixed, int f16, SGeom fg, VGeom v, int mode,
                              float* out) {
    // mode 0: 3D grid, trilinear; 1: 2D grid, bilinear per frame (frames = v.Z)
    const long long n = (long long)v.X * v.Y * v.Z;
    const float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % v.X), y = (int)((i / v.X) % v.Y), z = (int)(i / ((long long)v.X * v.Y));
        const float lx = fmaf((float)x, v.sx, v.ox), ly = fmaf((float)y, v.sy, v.oy);
        float val;
        if (mode == 0) {
            val = sample3(fixed, f16, fg, ident, lx, ly, fmaf((float)z, v.sz, v.oz));
        } else {
            const int off = z * fg.X * fg.Y;
            const void* fp = f16 ? (const void*)(static_cast<const __half*>(fixed) + off)
                                 : (const void*)(static_cast<const float*>(fixed) + off);
            val = sample2(fp, f16, fg, ident, lx, ly);
        }
        out[i] = val;
    }
}

__global__ void k_warp_moving(FusedArgs a, int grid3, int src3, float* out) {
    const VGeom v = a.v;
    const long long n = (long long)v.X * v.Y * v.Z;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % v.X), y = (int)((i / v.X) % v.Y), z = (int)(i / ((long long)v.X * v.Y));
        const float* mp = a.mats + (size_t)(grid3 ? 0 : (a.mats_per_pose > 1 ? z : 0)) * 12;
        float m[12];
        for (int k = 0; k < 12; ++k) m[k] = mp[k];
        const float lx = fmaf((float)x, v.sx, v.ox), ly = fmaf((float)y, v.sy, v.oy);
        const float lz = grid3 ? fmaf((float)z, v.sz, v.oz) : v.oz;
        out[i] = src3 ? sample3(a.src, a.src_f16, a.s, m, lx, ly, lz)
                      : sample2(a.src, a.src_f16, a.s, m, lx, ly);
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

size_t lds_bytes_2d(int r) {
    const size_t PX = kTile + 2 * r, NP = PX * PX, IXE = PX + 2;
    return sizeof(double) * (NCH * NP + NCH * kTile * PX + 8) +
           sizeof(float) * (IXE * IXE + 2 * PX * IXE) + sizeof(int) * (NP + kTile * PX);
}

template <int R>
hipError_t launch3(const FusedArgs& a, bool movi, dim3 grid, hipStream_t st) {
    if (movi) hipLaunchKernelGGL((k_fused3d<R, true>), grid, dim3(kThreads), 0, st, a);
    else hipLaunchKernelGGL((k_fused3d<R, false>), grid, dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

template <bool MOVI, bool SRC3D>
hipError_t launch2(const FusedArgs& a, dim3 grid, hipStream_t st) {
    const size_t lds = lds_bytes_2d(a.r);
    static bool configured = false;  // allow the largest radius once; smaller ones fit under it
    if (!configured) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fused2d<MOVI, SRC3D>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds_bytes_2d(kMaxR2));
        if (e != hipSuccess) return e;
        configured = true;
    }
    hipLaunchKernelGGL((k_fused2d<MOVI, SRC3D>), grid, dim3(kThreads), lds, st, a);
    return hipGetLastError();
}

}  // namespace

int fused_blocks_3d(const VGeom& v, int r, int zchunk, int* ntx, int* nty, int* nzc) {
    const int ex = v.X - 2 * r, ey = v.Y - 2 * r, ez = v.Z - 2 * r;
    if (ex <= 0 || ey <= 0 || ez <= 0 || zchunk <= 0) { *ntx = *nty = *nzc = 0; return 0; }
    *ntx = (ex + kTile - 1) / kTile;
    *nty = (ey + kTile - 1) / kTile;
    *nzc = (ez + zchunk - 1) / zchunk;
    return *ntx * *nty * *nzc;
}

int fused_blocks_2d(const VGeom& v, int r, int* ntx, int* nty) {
    const int ex = v.X - 2 * r, ey = v.Y - 2 * r;
    if (ex <= 0 || ey <= 0) { *ntx = *nty = 0; return 0; }
    *ntx = (ex + kTile - 1) / kTile;
    *nty = (ey + kTile - 1) / kTile;
    return *ntx * *nty * v.Z;
}

hipError_t launch_fused(const FusedArgs& a, int grid_ndim, int src_ndim, int movi, int n_poses,
                        hipStream_t st) {
    if (a.blocks_per_pose == 0) return hipSuccess;
    if (grid_ndim == 3) {
        dim3 grid(a.ntx * a.nty, a.nzc, n_poses);
        switch (a.r) {
            case 1: return launch3<1>(a, movi, grid, st);
            case 2: return launch3<2>(a, movi, grid, st);
            case 3: return launch3<3>(a, movi, grid, st);
            case 4: return launch3<4>(a, movi, grid, st);
            default: return hipErrorInvalidValue;
        }
    }
    if (a.r < 1 || a.r > kMaxR2) return hipErrorInvalidValue;
    dim3 grid(a.ntx * a.nty, a.v.Z, n_poses);
    if (movi) return src_ndim == 3 ? launch2<true, true>(a, grid, st) : launch2<true, false>(a, grid, st);
    return src_ndim == 3 ? launch2<false, true>(a, grid, st) : launch2<false, false>(a, grid, st);
}

hipError_t launch_reduce(const double* partials, int blocks_per_pose, int n_poses, double* out,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_reduce, dim3(n_poses), dim3(256), 0, st, partials, blocks_per_pose, out);
    return hipGetLastError();
}

static dim3 grid_for(long long n) {
    long long b = (n + 255) / 256;
    if (b > 65535LL * 16) b = 65535LL * 16;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

hipError_t launch_warp_static(const void* fixed, int fixed_f16, const SGeom& fg, const VGeom& v,
                              int grid_ndim, int src_ndim, float* out, hipStream_t st) {
    const long long n = (long long)v.X * v.Y * v.Z;
    // A 2D grid over a 3D fixed "volume" of one slice (2D->3D) samples trilinearly with
    // Z = 1, which equals the bilinear per-frame path bit for bit.
    const int mode = (grid_ndim == 3) ? 0 : 1;
    (void)src_ndim;
    hipLaunchKernelGGL(k_warp_static, grid_for(n), dim3(256), 0, st, fixed, fixed_f16, fg, v, mode,
                       out);
    return hipGetLastError();
}

hipError_t launch_warp_moving(const FusedArgs& a, int grid_ndim, int src_ndim, float* out,
                              hipStream_t st) {
    const long long n = (long long)a.v.X * a.v.Y * a.v.Z;
    hipLaunchKernelGGL(k_warp_moving, grid_for(n), dim3(256), 0, st, a, grid_ndim == 3 ? 1 : 0,
                       src_ndim == 3 ? 1 : 0, out);
    return hipGetLastError();
}

hipError_t launch_sobel(const float* img, const VGeom& v, int grid_ndim, float* out,
                        hipStream_t st) {
    const long long n = (long long)v.X * v.Y * v.Z;
    hipLaunchKernelGGL(k_sobel, grid_for(n), dim3(256), 0, st, img, v, grid_ndim == 3 ? 1 : 0,
                       out);
    return hipGetLastError();
}

hipError_t launch_convert(const void* d_in, int in_dtype, void* d_out, int out_f16, long long n,
                          hipStream_t st) {
    hipLaunchKernelGGL(k_convert, grid_for(n), dim3(256), 0, st, d_in, in_dtype, d_out, out_f16, n);
    return hipGetLastError();
}

}  // namespace lc2e
