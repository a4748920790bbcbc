// Host geometry of the engine: virtual grid, rotation centre and shift, pose -> matrix.
//   virtual grid            interpolator.cu:320, :337-339 (3D), :1079-1080 (2D), :1603-1605 (2D->3D)
//   centre / shift          interpolator.cu:495-529, :1198-1234, :1635-1669
//   pose -> VirtualToMoving spatialMappers.cuh:579-850 (3D), :977-1033 (2D)
//   inverse / output matrix interpolator.cu:633-679, spatialMappers.cuh:315-445, :946-975
#pragma once

#include <cmath>

#include "lc2e.h"

namespace lc2e {

struct PairGeometry {
    lc2e_geom virt;
    float fixed_center[3];
    float shift[3];
};

// PhysioSpatial*Mapper::SetCenter (spatialMappers.cuh:107-111): origin + spacing*dim/2, FP32.
inline void image_center(const lc2e_geom& g, float c[3]) {
    for (int a = 0; a < 3; ++a) c[a] = g.origin[a] + g.spacing[a] * (float)g.dims[a] / 2.0f;
}

inline PairGeometry make_pair_geometry(const lc2e_geom& fixed, const lc2e_geom& moving,
                                       const float resampled[3], int dim, int src_ndim,
                                       bool center_overlaid) {
    PairGeometry pg{};
    for (int a = 0; a < 3; ++a) pg.virt.origin[a] = fixed.origin[a];
    for (int a = 0; a < 3; ++a) {
        const bool resampled_axis = a < 2 || dim == 3 || (src_ndim == 3 && resampled[a] > 0.0f);
        pg.virt.spacing[a] = resampled_axis ? resampled[a] : fixed.spacing[a];
    }
    for (int a = 0; a < 3; ++a) {
        if (a < 2 || dim == 3) {
            const float ext = (float)fixed.dims[a] * fixed.spacing[a];
            pg.virt.dims[a] = (int)std::ceil((double)(ext / pg.virt.spacing[a]));
        } else {
            pg.virt.dims[a] = fixed.dims[2];  // 2D grid: one plane per fixed frame
        }
    }
    float cm[3], cv[3];
    image_center(fixed, pg.fixed_center);
    image_center(moving, cm);
    image_center(pg.virt, cv);
    for (int a = 0; a < 3; ++a) pg.shift[a] = center_overlaid ? cm[a] - cv[a] : 0.0f;
    return pg;
}

inline void mat3_mul(const double a[9], const double b[9], double c[9]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a[i * 3 + k] * b[k * 3 + j];
            c[i * 3 + j] = s;
        }
}

// VirtualToMoving M(x) = A (x - C + shift) + C + T, A = Rz Ry Rx Sh diag(S) (3D) or
// R Sh diag(S) (2D), degrees. The reference casts its double parameters to float first
// (optimizers.cu:677-694); the matrix is composed in FP64 and rounded once.
// The FP64 linear part A (3x3 row-major) and translation t of that matrix, before rounding.
inline void affine_3d_fp64(const float T[3], const float Rdeg[3], const float Sc[3],
                           const float Sh[3], const float C[3], const float shift[3],
                           double A[9], double t[3]) {
    const double d2r = 3.14159265358979323846 / 180.0;
    const double cx = std::cos(Rdeg[0] * d2r), sx = std::sin(Rdeg[0] * d2r);
    const double cy = std::cos(Rdeg[1] * d2r), sy = std::sin(Rdeg[1] * d2r);
    const double cz = std::cos(Rdeg[2] * d2r), sz = std::sin(Rdeg[2] * d2r);
    const double rx[9] = {1, 0, 0, 0, cx, -sx, 0, sx, cx};
    const double ry[9] = {cy, 0, sy, 0, 1, 0, -sy, 0, cy};
    const double rz[9] = {cz, -sz, 0, sz, cz, 0, 0, 0, 1};
    const double sh[9] = {1, Sh[0], Sh[1], 0, 1, Sh[2], 0, 0, 1};
    const double sc[9] = {Sc[0], 0, 0, 0, Sc[1], 0, 0, 0, Sc[2]};
    double t0[9], t1[9], t2[9];
    mat3_mul(rz, ry, t0);
    mat3_mul(t0, rx, t1);
    mat3_mul(t1, sh, t2);
    mat3_mul(t2, sc, A);
    for (int i = 0; i < 3; ++i) {
        t[i] = (double)C[i] + (double)T[i];
        for (int k = 0; k < 3; ++k) t[i] += A[i * 3 + k] * ((double)shift[k] - (double)C[k]);
    }
}

inline void affine_matrix_3d(const float T[3], const float Rdeg[3], const float Sc[3],
                             const float Sh[3], const float C[3], const float shift[3],
                             float out12[12]) {
    double A[9], t[3];
    affine_3d_fp64(T, Rdeg, Sc, Sh, C, shift, A, t);
    for (int i = 0; i < 3; ++i) {
        for (int k = 0; k < 3; ++k) out12[i * 4 + k] = (float)A[i * 3 + k];
        out12[i * 4 + 3] = (float)t[i];
    }
}

inline void affine_matrix_2d(const float T[2], float Rdeg, const float Sc[2], float Sh,
                             const float C[2], const float shift[2], float out12[12]) {
    const double d2r = 3.14159265358979323846 / 180.0;
    const double c = std::cos(Rdeg * d2r), s = std::sin(Rdeg * d2r);
    const double A[4] = {Sc[0] * c, Sc[1] * (c * Sh - s), Sc[0] * s, Sc[1] * (s * Sh + c)};
    double t[2];
    for (int i = 0; i < 2; ++i) {
        t[i] = (double)C[i] + (double)T[i];
        for (int k = 0; k < 2; ++k) t[i] += A[i * 2 + k] * ((double)shift[k] - (double)C[k]);
    }
    const float m[12] = {(float)A[0], (float)A[1], 0.0f, (float)t[0],
                         (float)A[2], (float)A[3], 0.0f, (float)t[1],
                         0.0f,        0.0f,        1.0f, 0.0f};
    for (int i = 0; i < 12; ++i) out12[i] = m[i];
}

// params -> 3x4. Returns 0 or LC2E_ERR_INVALID.
inline int pose_to_matrix(const lc2e_pair_info& info, const double* p, int n, float out12[12]) {
    if (info.src_ndim == 3) {
        float T[3], R[3], S[3] = {1, 1, 1}, H[3] = {0, 0, 0};
        if (n != 6 && n != 12) return LC2E_ERR_INVALID;
        for (int a = 0; a < 3; ++a) {
            T[a] = (float)p[a];
            R[a] = (float)p[3 + a];
            if (n == 12) {
                S[a] = (float)p[6 + a];
                H[a] = (float)p[9 + a];
            }
        }
        affine_matrix_3d(T, R, S, H, info.fixed_center, info.shift, out12);
        return 0;
    }
    if (n != 3 && n != 6) return LC2E_ERR_INVALID;
    const float T[2] = {(float)p[0], (float)p[1]};
    const float S[2] = {n == 6 ? (float)p[3] : 1.0f, n == 6 ? (float)p[4] : 1.0f};
    const float H = n == 6 ? (float)p[5] : 0.0f;
    affine_matrix_2d(T, (float)p[2], S, H, info.fixed_center, info.shift, out12);
    return 0;
}

// Inverse of the affine 3x4 (bottom row [0,0,0,1]) in FP64, written as a 4x4 row-major.
// Spatial3DMapper::GetMyselfInversed (spatialMappers.cuh:315-445). Returns false if singular.
inline bool invert_affine(const float m12[12], float out16[16]) {
    const double a = m12[0], b = m12[1], c = m12[2], d = m12[4], e = m12[5], f = m12[6],
                 g = m12[8], h = m12[9], k = m12[10];
    const double c00 = e * k - f * h, c01 = -(d * k - f * g), c02 = d * h - e * g;
    const double det = a * c00 + b * c01 + c * c02;
    if (det == 0.0 || !std::isfinite(det)) return false;
    const double id = 1.0 / det;
    const double Ai[9] = {c00 * id, -(b * k - c * h) * id, (b * f - c * e) * id,
                          c01 * id, (a * k - c * g) * id,  -(a * f - c * d) * id,
                          c02 * id, -(a * h - b * g) * id, (a * e - b * d) * id};
    for (int i = 0; i < 3; ++i) {
        double t = 0.0;
        for (int j = 0; j < 3; ++j) t -= Ai[i * 3 + j] * (double)m12[j * 4 + 3];
        for (int j = 0; j < 3; ++j) out16[i * 4 + j] = (float)Ai[i * 3 + j];
        out16[i * 4 + 3] = (float)t;
    }
    out16[12] = out16[13] = out16[14] = 0.0f;
    out16[15] = 1.0f;
    return true;
}

// GetMovingToFixedTransform (interpolator.cu:633-679 3D, :1382-1425 2D, :1729-1775 2D->3D):
// VirtualToFixed (the identity) times inv(VirtualToMoving), bottom row forced to [0,0,0,1];
// out is 4x4 row-major (out_dim 4) or the 2D 3x3 (out_dim 3). Returns false if singular.
inline bool moving_to_fixed(const float m12[12], int out_dim, float* out) {
    float inv[16];
    if (!invert_affine(m12, inv)) return false;
    if (out_dim == 4) {
        for (int i = 0; i < 16; ++i) out[i] = inv[i];
    } else {
        const float o[9] = {inv[0], inv[1], inv[3], inv[4], inv[5], inv[7], 0, 0, 1};
        for (int i = 0; i < 9; ++i) out[i] = o[i];
    }
    return true;
}

}  // namespace lc2e
