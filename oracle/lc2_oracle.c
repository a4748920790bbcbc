/*
 * LC2 oracle -- TEST INFRASTRUCTURE ONLY (see lc2_oracle.h).
 *
 * CPU restatement of the reference's LC2 evaluation. Compiled with
 * -ffp-contract=off; every fused multiply-add the reference's nvcc build would
 * form is written out as fmaf() so the FP32 stages are reproducible bit for bit.
 */
#include "lc2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LC2O_ZERO_THRESHOLD 1e-6   /* cu_lc2_helper.cuh:5 */
/* Rank handling of the exact least-squares projection (the reference's CGLS
 * converges to this projection; it has no explicit rule of its own). A column
 * whose centred energy is below TAU_CONST of its raw energy is constant within
 * rounding and dropped; a pair whose Gram determinant is below TAU_COLL*a*b is
 * collinear and U is projected on the leading singular direction of the two centred,
 * unit-normalised columns. That is least squares by truncated SVD on those columns
 * with the cutoff sigma_min/sigma_max <= sqrt(TAU_COLL)/2 (numpy lstsq, rcond 5e-6). */
#define LC2O_TAU_CONST 1e-10
#define LC2O_TAU_COLL 1e-10

int lc2o_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Threads of the parallel loops that follow (n >= 1); returns the count now in force. */
int lc2o_set_num_threads(int n) {
#ifdef _OPENMP
    if (n >= 1) omp_set_num_threads(n);
#endif
    return lc2o_num_threads();
}

/* ------------------------------------------------------------------------ */
/* Geometry                                                                  */
/* ------------------------------------------------------------------------ */

/* PhysioSpatial3DMapper::SetCenter, spatialMappers.cuh:107-111:
 * centre = origin + spacing * dim / 2 (FP32). */
void lc2o_center(const lc2o_geom* g, int ndim, float out[3]) {
    for (int a = 0; a < 3; ++a) {
        if (a < ndim)
            out[a] = g->origin[a] + g->spacing[a] * (float)g->dims[a] / 2.0f;
        else
            out[a] = 0.0f;
    }
}

/* Virtual grid dims, interpolator.cu:337-339: ceil(dim_f * spacing_f / spacing_v). */
void lc2o_virtual_dims(const lc2o_geom* fixed, const float resampled[3], int ndim, int out[3]) {
    for (int a = 0; a < 3; ++a) {
        if (a < ndim) {
            float ext = (float)fixed->dims[a] * fixed->spacing[a];
            out[a] = (int)ceil((double)(ext / resampled[a]));
        } else {
            out[a] = 1;
        }
    }
}

/* 3x3 product helper in double: c = a*b */
static void mm3(const double a[9], const double b[9], double c[9]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a[i * 3 + k] * b[k * 3 + j];
            c[i * 3 + j] = s;
        }
}

/* Spatial3DMapper::AffineInit, spatialMappers.cuh:637-850 (matrix entries
 * :800-840, elementary-matrix order in the derivation :642-797):
 *   M(x) = A (x - C + shift) + C + T,   A = Rz Ry Rx Sh diag(S),
 *   Sh = [[1, shx, shy], [0, 1, shz], [0, 0, 1]], angles in degrees.
 * RigidInit (:579-624) is the special case S = 1, Sh = 0.
 * Evaluated in FP64 and rounded to FP32 once. */
void lc2o_affine_init_3d(const float rot_deg[3], const float trans[3], const float scale[3],
                         const float shear[3], const float center[3], const float shift[3],
                         float out16[16]) {
    const double d2r = 3.14159265358979323846 / 180.0;
    double cx = cos(rot_deg[0] * d2r), sx = sin(rot_deg[0] * d2r);
    double cy = cos(rot_deg[1] * d2r), sy = sin(rot_deg[1] * d2r);
    double cz = cos(rot_deg[2] * d2r), sz = sin(rot_deg[2] * d2r);
    double Rx[9] = {1, 0, 0, 0, cx, -sx, 0, sx, cx};
    double Ry[9] = {cy, 0, sy, 0, 1, 0, -sy, 0, cy};
    double Rz[9] = {cz, -sz, 0, sz, cz, 0, 0, 0, 1};
    double Sh[9] = {1, shear[0], shear[1], 0, 1, shear[2], 0, 0, 1};
    double S[9] = {scale[0], 0, 0, 0, scale[1], 0, 0, 0, scale[2]};
    double t1[9], t2[9], t3[9], A[9];
    mm3(Rz, Ry, t1);
    mm3(t1, Rx, t2);
    mm3(t2, Sh, t3);
    mm3(t3, S, A);
    for (int i = 0; i < 3; ++i) {
        double t = (double)center[i] + (double)trans[i];
        for (int k = 0; k < 3; ++k) t += A[i * 3 + k] * ((double)shift[k] - (double)center[k]);
        for (int k = 0; k < 3; ++k) out16[i * 4 + k] = (float)A[i * 3 + k];
        out16[i * 4 + 3] = (float)t;
    }
    out16[12] = 0.0f; out16[13] = 0.0f; out16[14] = 0.0f; out16[15] = 1.0f;
}

/* Spatial2DMapper::AffineInit, spatialMappers.cuh:999-1033:
 *   A = R(theta) [[1, sh], [0, 1]] diag(sx, sy),  M(x) = A (x - C + shift) + C + T.
 * RigidInit (:977-997) is S = 1, sh = 0. FP64, rounded once. */
void lc2o_affine_init_2d(float rot_deg, const float trans[2], const float scale[2], float shear,
                         const float center[2], const float shift[2], float out9[9]) {
    const double d2r = 3.14159265358979323846 / 180.0;
    double c = cos(rot_deg * d2r), s = sin(rot_deg * d2r);
    double A[4];
    A[0] = scale[0] * c;
    A[1] = scale[1] * (c * shear - s);
    A[2] = scale[0] * s;
    A[3] = scale[1] * (s * shear + c);
    for (int i = 0; i < 2; ++i) {
        double t = (double)center[i] + (double)trans[i];
        for (int k = 0; k < 2; ++k) t += A[i * 2 + k] * ((double)shift[k] - (double)center[k]);
        out9[i * 3 + 0] = (float)A[i * 2 + 0];
        out9[i * 3 + 1] = (float)A[i * 2 + 1];
        out9[i * 3 + 2] = (float)t;
    }
    out9[6] = 0.0f; out9[7] = 0.0f; out9[8] = 1.0f;
}

/* Spatial3DMapper::GetMyselfInversed, spatialMappers.cuh:315-445: full inverse,
 * then the bottom row forced to [0,0,0,1] (:439-442). Done here for the affine
 * case as [A^-1, -A^-1 t] in FP64. Returns 0 if singular. */
int lc2o_inverse_4x4(const float m[16], float inv[16]) {
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9],
           k = m[10];
    double co00 = e * k - f * h, co01 = -(d * k - f * g), co02 = d * h - e * g;
    double det = a * co00 + b * co01 + c * co02;
    if (det == 0.0) return 0;
    double id = 1.0 / det;
    double Ai[9];
    Ai[0] = co00 * id;
    Ai[1] = -(b * k - c * h) * id;
    Ai[2] = (b * f - c * e) * id;
    Ai[3] = co01 * id;
    Ai[4] = (a * k - c * g) * id;
    Ai[5] = -(a * f - c * d) * id;
    Ai[6] = co02 * id;
    Ai[7] = -(a * h - b * g) * id;
    Ai[8] = (a * e - b * d) * id;
    for (int i = 0; i < 3; ++i) {
        double t = 0.0;
        for (int j = 0; j < 3; ++j) t -= Ai[i * 3 + j] * (double)m[j * 4 + 3];
        for (int j = 0; j < 3; ++j) inv[i * 4 + j] = (float)Ai[i * 3 + j];
        inv[i * 4 + 3] = (float)t;
    }
    inv[12] = 0.0f; inv[13] = 0.0f; inv[14] = 0.0f; inv[15] = 1.0f;
    return 1;
}

/* Spatial2DMapper::GetMyselfInversed, spatialMappers.cuh:946-975 (bottom row forced). */
int lc2o_inverse_3x3(const float m[9], float inv[9]) {
    double a = m[0], b = m[1], d = m[3], e = m[4];
    double det = a * e - b * d;
    if (det == 0.0) return 0;
    double id = 1.0 / det;
    double Ai[4] = {e * id, -b * id, -d * id, a * id};
    for (int i = 0; i < 2; ++i) {
        double t = -(Ai[i * 2 + 0] * (double)m[2] + Ai[i * 2 + 1] * (double)m[5]);
        inv[i * 3 + 0] = (float)Ai[i * 2 + 0];
        inv[i * 3 + 1] = (float)Ai[i * 2 + 1];
        inv[i * 3 + 2] = (float)t;
    }
    inv[6] = 0.0f; inv[7] = 0.0f; inv[8] = 1.0f;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Warp                                                                      */
/* ------------------------------------------------------------------------ */

/* Interpolate_3Dkernel, interpolator.cu:69-179 (one of its two identical branches).
 *  L   = ToLoc(p)        spatialMappers.cuh:76-82  (p*spacing + origin, one fma)
 *  L_m = M * L           Spatial3DMapper::operator*, spatialMappers.cuh:274-295
 *                        (accumulated 0 + m0*x + m1*y + m2*z + m3, fma-chained)
 *  v   = ToPxl(L_m)      spatialMappers.cuh:84-90  ((L - origin) / spacing)
 *  sample iff 0 <= v <= dim-1 on every axis, trilinear z -> y -> x with floor/ceil
 *  corner indices (interpolator.cu:139-174), else 0. */
void lc2o_warp_3d(const float* src, const lc2o_geom* sg, const lc2o_geom* vg, const float m[12],
                  float* out) {
    const int VX = vg->dims[0], VY = vg->dims[1], VZ = vg->dims[2];
    const int SX = sg->dims[0], SY = sg->dims[1], SZ = sg->dims[2];
    const long long sxy = (long long)SX * SY;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < VZ; ++z) {
        for (int y = 0; y < VY; ++y) {
            for (int x = 0; x < VX; ++x) {
                float lx = fmaf((float)x, vg->spacing[0], vg->origin[0]);
                float ly = fmaf((float)y, vg->spacing[1], vg->origin[1]);
                float lz = fmaf((float)z, vg->spacing[2], vg->origin[2]);
                float px = fmaf(m[2], lz, fmaf(m[1], ly, m[0] * lx)) + m[3];
                float py = fmaf(m[6], lz, fmaf(m[5], ly, m[4] * lx)) + m[7];
                float pz = fmaf(m[10], lz, fmaf(m[9], ly, m[8] * lx)) + m[11];
                float vx = (px - sg->origin[0]) / sg->spacing[0];
                float vy = (py - sg->origin[1]) / sg->spacing[1];
                float vz = (pz - sg->origin[2]) / sg->spacing[2];
                float val = 0.0f;
                if (vx >= 0.0f && vx <= (float)(SX - 1) && vy >= 0.0f && vy <= (float)(SY - 1) &&
                    vz >= 0.0f && vz <= (float)(SZ - 1)) {
                    int x0 = (int)floorf(vx), x1 = (int)ceilf(vx);
                    int y0 = (int)floorf(vy), y1 = (int)ceilf(vy);
                    int z0 = (int)floorf(vz), z1 = (int)ceilf(vz);
                    float rx = vx - (float)x0, ry = vy - (float)y0, rz = vz - (float)z0;
#define S3(X, Y, Z) src[(long long)(X) + (long long)(Y) * SX + (long long)(Z) * sxy]
                    float a, b;
                    a = S3(x0, y0, z0); b = S3(x0, y0, z1);
                    float lu = fmaf(rz, b - a, a);
                    a = S3(x1, y0, z0); b = S3(x1, y0, z1);
                    float ru = fmaf(rz, b - a, a);
                    a = S3(x1, y1, z0); b = S3(x1, y1, z1);
                    float rb = fmaf(rz, b - a, a);
                    a = S3(x0, y1, z0); b = S3(x0, y1, z1);
                    float lb = fmaf(rz, b - a, a);
#undef S3
                    float yl = fmaf(ry, lb - lu, lu);
                    float yr = fmaf(ry, rb - ru, ru);
                    val = fmaf(rx, yr - yl, yl);
                }
                out[(long long)x + (long long)y * VX + (long long)z * VX * VY] = val;
            }
        }
    }
}

/* Interpolate_2Dkernel, interpolator.cu:181-232 (and the reference's own CPU
 * restatement, interpolator.cu:1273-1342): bilinear, y first then x. */
void lc2o_warp_2d(const float* src, const lc2o_geom* sg, const lc2o_geom* vg, const float m[6],
                  float* out) {
    const int VX = vg->dims[0], VY = vg->dims[1];
    const int SX = sg->dims[0], SY = sg->dims[1];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < VY; ++y) {
        for (int x = 0; x < VX; ++x) {
            float lx = fmaf((float)x, vg->spacing[0], vg->origin[0]);
            float ly = fmaf((float)y, vg->spacing[1], vg->origin[1]);
            float px = fmaf(m[1], ly, m[0] * lx) + m[2];
            float py = fmaf(m[4], ly, m[3] * lx) + m[5];
            float vx = (px - sg->origin[0]) / sg->spacing[0];
            float vy = (py - sg->origin[1]) / sg->spacing[1];
            float val = 0.0f;
            if (vx >= 0.0f && vx <= (float)(SX - 1) && vy >= 0.0f && vy <= (float)(SY - 1)) {
                int x0 = (int)floorf(vx), x1 = (int)ceilf(vx);
                int y0 = (int)floorf(vy), y1 = (int)ceilf(vy);
                float ry = vy - floorf(vy);
                float a = src[x0 + (long long)y0 * SX], b = src[x0 + (long long)y1 * SX];
                float xl = fmaf(ry, b - a, a);
                a = src[x1 + (long long)y0 * SX]; b = src[x1 + (long long)y1 * SX];
                float xr = fmaf(ry, b - a, a);
                float rx = vx - floorf(vx);
                val = fmaf(rx, xr - xl, xl);
            }
            out[x + (long long)y * VX] = val;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Sobel gradient magnitude                                                  */
/* ------------------------------------------------------------------------ */

static inline float at3(const float* I, int X, int Y, int Z, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= X || y >= Y || z >= Z) return 0.0f;
    return I[(long long)x + (long long)y * X + (long long)z * X * Y];
}

/* gradient3DSobelKernel, cu_lc2_3d.cu:100-179 with taps from PrepareGradientFilter
 * :977-1005 (S_x[ix,iy,iz] = hp[ix] h[iy] h[iz], h = [1,2,1], hp = [-1,0,1]),
 * correlation form (tap k <-> offset k-1), zero padding outside the grid, taps
 * accumulated in the kernel's z, y, x loop order, |g| = sqrt(gx^2+gy^2+gz^2). */
void lc2o_sobel_3d_ref(const float* I, const int dims[3], float* G) {
    static const float h[3] = {1.0f, 2.0f, 1.0f};
    static const float hp[3] = {-1.0f, 0.0f, 1.0f};
    float kx[27], ky[27], kz[27];
    for (int iz = 0; iz < 3; ++iz)
        for (int iy = 0; iy < 3; ++iy)
            for (int ix = 0; ix < 3; ++ix) {
                int k = ix + iy * 3 + iz * 9;
                kx[k] = hp[ix] * h[iy] * h[iz];
                ky[k] = h[ix] * hp[iy] * h[iz];
                kz[k] = h[ix] * h[iy] * hp[iz];
            }
    const int X = dims[0], Y = dims[1], Z = dims[2];
#pragma omp parallel for schedule(static)
    for (int z = 0; z < Z; ++z)
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X; ++x) {
                float gx = 0.0f, gy = 0.0f, gz = 0.0f;
                for (int iz = 0; iz < 3; ++iz)
                    for (int iy = 0; iy < 3; ++iy)
                        for (int ix = 0; ix < 3; ++ix) {
                            int k = ix + iy * 3 + iz * 9;
                            float v = at3(I, X, Y, Z, x + ix - 1, y + iy - 1, z + iz - 1);
                            gx = fmaf(kx[k], v, gx);
                            gy = fmaf(ky[k], v, gy);
                            gz = fmaf(kz[k], v, gz);
                        }
                G[(long long)x + (long long)y * X + (long long)z * X * Y] =
                    sqrtf(fmaf(gz, gz, fmaf(gy, gy, gx * gx)));
            }
}

/* Same operator evaluated separably (the engine's order): per plane
 *   d = I[x+1]-I[x-1], s = (I[x-1]+2I[x])+I[x+1]           (along x)
 *   A = (s[y-1]+2s[y])+s[y+1], Bx = (d[y-1]+2d[y])+d[y+1], By = s[y+1]-s[y-1]
 * then gx = (Bx[z-1]+2Bx[z])+Bx[z+1], gy likewise from By, gz = A[z+1]-A[z-1].
 * Same taps and padding as lc2o_sobel_3d_ref; only FP32 rounding of the sums differs. */
void lc2o_sobel_3d_sep(const float* I, const int dims[3], float* G) {
    const int X = dims[0], Y = dims[1], Z = dims[2];
    const long long P = (long long)X * Y;
    float* A = (float*)malloc(sizeof(float) * P * Z);
    float* Bx = (float*)malloc(sizeof(float) * P * Z);
    float* By = (float*)malloc(sizeof(float) * P * Z);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < Z; ++z)
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X; ++x) {
                float s[3], d[3];
                for (int k = 0; k < 3; ++k) {
                    int yy = y + k - 1;
                    float l = at3(I, X, Y, Z, x - 1, yy, z);
                    float c = at3(I, X, Y, Z, x, yy, z);
                    float r = at3(I, X, Y, Z, x + 1, yy, z);
                    d[k] = r - l;
                    s[k] = (l + 2.0f * c) + r;
                }
                long long i = (long long)x + (long long)y * X + (long long)z * P;
                A[i] = (s[0] + 2.0f * s[1]) + s[2];
                Bx[i] = (d[0] + 2.0f * d[1]) + d[2];
                By[i] = s[2] - s[0];
            }
#pragma omp parallel for schedule(static)
    for (int z = 0; z < Z; ++z)
        for (long long p = 0; p < P; ++p) {
            long long i = p + (long long)z * P;
            float bxm = z > 0 ? Bx[i - P] : 0.0f, bxp = z + 1 < Z ? Bx[i + P] : 0.0f;
            float bym = z > 0 ? By[i - P] : 0.0f, byp = z + 1 < Z ? By[i + P] : 0.0f;
            float am = z > 0 ? A[i - P] : 0.0f, ap = z + 1 < Z ? A[i + P] : 0.0f;
            float gx = (bxm + 2.0f * Bx[i]) + bxp;
            float gy = (bym + 2.0f * By[i]) + byp;
            float gz = ap - am;
            G[i] = sqrtf(fmaf(gz, gz, fmaf(gy, gy, gx * gx)));
        }
    free(A);
    free(Bx);
    free(By);
}

/* gradient2DSobelKernel, cu_lc2_2d.cu:25-86; taps cu_lc2_2d.cu:548-578
 * (S_x[ix,iy] = hp[ix] h[iy], S_y = h[ix] hp[iy]), zero padding, row-major tap order. */
void lc2o_sobel_2d_ref(const float* I, const int dims[2], float* G) {
    static const float h[3] = {1.0f, 2.0f, 1.0f};
    static const float hp[3] = {-1.0f, 0.0f, 1.0f};
    const int X = dims[0], Y = dims[1];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < Y; ++y)
        for (int x = 0; x < X; ++x) {
            float gx = 0.0f, gy = 0.0f;
            for (int iy = 0; iy < 3; ++iy)
                for (int ix = 0; ix < 3; ++ix) {
                    float v = at3(I, X, Y, 1, x + ix - 1, y + iy - 1, 0);
                    gx = fmaf(hp[ix] * h[iy], v, gx);
                    gy = fmaf(h[ix] * hp[iy], v, gy);
                }
            G[x + (long long)y * X] = sqrtf(fmaf(gy, gy, gx * gx));
        }
}

/* Separable order of the 2D operator (engine's): gx = Bx, gy = By as in the 3D case. */
void lc2o_sobel_2d_sep(const float* I, const int dims[2], float* G) {
    const int X = dims[0], Y = dims[1];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < Y; ++y)
        for (int x = 0; x < X; ++x) {
            float s[3], d[3];
            for (int k = 0; k < 3; ++k) {
                int yy = y + k - 1;
                float l = at3(I, X, Y, 1, x - 1, yy, 0);
                float c = at3(I, X, Y, 1, x, yy, 0);
                float r = at3(I, X, Y, 1, x + 1, yy, 0);
                d[k] = r - l;
                s[k] = (l + 2.0f * c) + r;
            }
            float gx = (d[0] + 2.0f * d[1]) + d[2];
            float gy = s[2] - s[0];
            G[x + (long long)y * X] = sqrtf(fmaf(gy, gy, gx * gx));
        }
}

/* ------------------------------------------------------------------------ */
/* LC2                                                                       */
/* ------------------------------------------------------------------------ */

/* Explained variance s^T S^+ s of the centred two-column fit, with the rank rule
 * documented at LC2O_TAU_*. a = C_II, b = C_GG, c = C_IG, p = C_yI, q = C_yG;
 * rawa / rawb are the uncentred energies used as the scale of "zero". */
static double explained_var(double a, double b, double c, double p, double q, double rawa,
                            double rawb) {
    int ikeep = a > LC2O_TAU_CONST * rawa;
    int gkeep = b > LC2O_TAU_CONST * rawb;
    if (ikeep && gkeep) {
        double det = a * b - c * c;
        if (det > LC2O_TAU_COLL * a * b)
            return (b * p * p - 2.0 * c * p * q + a * q * q) / det;
        /* collinear: unit columns I', G' with correlation rho = c / sqrt(ab); the leading
         * singular direction is I' + sign(rho) G', of squared norm 2 + 2|rho| */
        double ra = 1.0 / sqrt(a), rb = 1.0 / sqrt(b);
        double t = p * ra + (c < 0.0 ? -1.0 : 1.0) * q * rb;
        return (t * t) / (2.0 + 2.0 * fabs(c) * ra * rb);
    }
    if (ikeep) return p * p / a;
    if (gkeep) return q * q / b;
    return 0.0;
}

/* The exact (FP64) LC2 -- the fixed point of the reference's CGLS -- for the
 * semantics of calculate{3,2}DWeightingFactorOnSharedMemory (cu_lc2_3d.cu:181-287,
 * cu_lc2_2d.cu:88-180), calculate{3,2}DWeightedSimilarityOnSharedMemory
 * (cu_lc2_3d.cu:526-618, cu_lc2_2d.cu:392-480) and GetSimilarityMetric
 * (cu_lc2_3d.cu:1119-1194, cu_lc2_2d.cu:599-653):
 *   patch P(v) = (2r+1)^D cube, U zero-padded outside the grid;
 *   w = 0 unless all N values of U > 1e-6 and var_U = sum (U-m)^2/(N-1) > 1e-6,
 *   else w = sqrt(var_U);
 *   fit U ~ a I + b G + c by least squares over P(v), res = U - fit,
 *   var_res = sum (res - mean res)^2 / (N-1);
 *   measure = w - var_res / w, 0 if NaN/Inf/<= 0;
 *   LC2 = sum measure / sum w (NaN/Inf -> 0).
 * Moments are taken two-pass (centred) in FP64 per patch. */
int lc2o_lc2_exact(const float* U, const float* I, const float* G, const int dims[3], int ndim,
                   int r, int zlo, int zhi, double* w_map, double* m_map, lc2o_result* out) {
    const int X = dims[0], Y = dims[1], Z = dims[2];
    if (ndim != 2 && ndim != 3) return -1;
    if (ndim == 2 && Z != 1) return -1;
    if (r < 1) return -1;
    const int d = 2 * r + 1;
    const int N = ndim == 3 ? d * d * d : d * d;
    const long long XY = (long long)X * Y;
    const int outer = ndim == 3 ? Z : Y;
    if (zlo < 0) zlo = 0;
    if (zhi > outer) zhi = outer;
    if (w_map) memset(w_map, 0, sizeof(double) * XY * Z);
    if (m_map) memset(m_map, 0, sizeof(double) * XY * Z);
    /* per-slab partial sums, added in slab order afterwards: the result does not depend on
     * the thread count or schedule */
    const int nslab = zhi > zlo ? zhi - zlo : 0;
    double* psm = (double*)calloc(nslab > 0 ? (size_t)nslab : 1, sizeof(double));
    double* psw = (double*)calloc(nslab > 0 ? (size_t)nslab : 1, sizeof(double));
    long long* pnv = (long long*)calloc(nslab > 0 ? (size_t)nslab : 1, sizeof(long long));
    if (!psm || !psw || !pnv) {
        free(psm);
        free(psw);
        free(pnv);
        return -1;
    }
    const int rz = ndim == 3 ? r : 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int o = zlo; o < zhi; ++o) {
        double sm = 0.0, sw = 0.0;
        long long nv = 0;
        int z0 = ndim == 3 ? o : 0;
        int ylo = ndim == 3 ? 0 : o, yhi = ndim == 3 ? Y : o + 1;
        for (int y = ylo; y < yhi; ++y)
            for (int x = 0; x < X; ++x) {
                /* border patches include zero padding -> not all U > threshold */
                if (x < r || y < r || x + r >= X || y + r >= Y) continue;
                if (ndim == 3 && (z0 < r || z0 + r >= Z)) continue;
                int cnt = 0;
                double su = 0.0, si = 0.0, sg = 0.0;
                for (int dz = -rz; dz <= rz; ++dz)
                    for (int dy = -r; dy <= r; ++dy)
                        for (int dx = -r; dx <= r; ++dx) {
                            long long i = (long long)(x + dx) + (long long)(y + dy) * X +
                                          (long long)(z0 + dz) * XY;
                            if ((double)U[i] > LC2O_ZERO_THRESHOLD) ++cnt;
                            su += U[i];
                            si += I[i];
                            sg += G[i];
                        }
                if (cnt != N) continue;
                double mu = su / N, mi = si / N, mg = sg / N;
                double cyy = 0, cii = 0, cgg = 0, cig = 0, cyi = 0, cyg = 0, rii = 0, rgg = 0;
                for (int dz = -rz; dz <= rz; ++dz)
                    for (int dy = -r; dy <= r; ++dy)
                        for (int dx = -r; dx <= r; ++dx) {
                            long long i = (long long)(x + dx) + (long long)(y + dy) * X +
                                          (long long)(z0 + dz) * XY;
                            double yu = (double)U[i] - mu, ii = (double)I[i] - mi,
                                   gg = (double)G[i] - mg;
                            cyy += yu * yu;
                            cii += ii * ii;
                            cgg += gg * gg;
                            cig += ii * gg;
                            cyi += yu * ii;
                            cyg += yu * gg;
                            rii += (double)I[i] * (double)I[i];
                            rgg += (double)G[i] * (double)G[i];
                        }
                double var_u = cyy / (N - 1);
                if (!(var_u > LC2O_ZERO_THRESHOLD)) continue;
                double w = sqrt(var_u);
                double ex = explained_var(cii, cgg, cig, cyi, cyg, rii, rgg);
                if (ex < 0.0) ex = 0.0;
                if (ex > cyy) ex = cyy;
                double var_res = (cyy - ex) / (N - 1);
                double meas = w - var_res / w;
                if (!(meas > 0.0) || isinf(meas)) meas = 0.0;
                long long vi = (long long)x + (long long)y * X + (long long)z0 * XY;
                if (w_map) w_map[vi] = w;
                if (m_map) m_map[vi] = meas;
                sm += meas;
                sw += w;
                nv += 1;
            }
        psm[o - zlo] = sm;
        psw[o - zlo] = sw;
        pnv[o - zlo] = nv;
    }
    double sm = 0.0, sw = 0.0;
    long long nv = 0;
    for (int k = 0; k < nslab; ++k) {
        sm += psm[k];
        sw += psw[k];
        nv += pnv[k];
    }
    free(psm);
    free(psw);
    free(pnv);
    out->sum_measure = sm;
    out->sum_w = sw;
    out->n_valid = nv;
    double l = sm / sw;
    out->lc2 = (isnan(l) || isinf(l)) ? 0.0 : l;
    return 0;
}

/* Pairwise FP32 sum: a stand-in for thrust::reduce's tree order
 * (cu_lc2_3d.cu:1181-1182); the exact thrust order is not reproducible here. */
static float pairwise_sum_f(const float* v, long long n) {
    if (n <= 64) {
        float s = 0.0f;
        for (long long i = 0; i < n; ++i) s += v[i];
        return s;
    }
    long long h = n / 2;
    return pairwise_sum_f(v, h) + pairwise_sum_f(v + h, n - h);
}

/* The reference's own arithmetic: FP32 weights (cu_lc2_3d.cu:226-284), FP32
 * CGLS with x0 = 0, at most 20 iterations, stop when ||C^T r|| < 1e-4 (3D,
 * cu_lc2_3d.cu:346) or 1e-6 (2D, cu_lc2_2d.cu:233), residual variance with
 * ddof 1 (cu_lc2_3d.cu:289-524), measure = w - var/w (cu_lc2_3d.cu:596-611),
 * FP32 reductions. fp64_arith = 1 runs the same recurrence in double. */
#define CG_BODY(T, SQRT)                                                                     \
    {                                                                                        \
        T xm0 = 0, xm1 = 0, xm2 = 0;                                                         \
        T g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0, g5 = 0;                                    \
        for (int i = 0; i < N; ++i) {                                                        \
            T res = (T)py[i] - (xm0 * (T)pc1[i] + xm1 * (T)pc2[i] + xm2 * (T)1);             \
            g0 += (T)pc1[i] * res;                                                           \
            g1 += (T)pc2[i] * res;                                                           \
            g2 += (T)1 * res;                                                                \
        }                                                                                    \
        T beta = 0, p0 = 0, p1 = 0, p2 = 0, alpha = 0;                                       \
        for (int it = 1; it <= 20; ++it) {                                                   \
            if (SQRT(g0 * g0 + g1 * g1 + g2 * g2) < (T)tol) break;                          \
            if (it > 1) beta = -(g0 * g0 + g1 * g1 + g2 * g2) / (g3 * g3 + g4 * g4 + g5 * g5); \
            if (it == 1) { p0 = g0; p1 = g1; p2 = g2; }                                      \
            else { p0 = g0 - beta * p0; p1 = g1 - beta * p1; p2 = g2 - beta * p2; }          \
            T nsq = 0;                                                                       \
            for (int i = 0; i < N; ++i) {                                                    \
                T cp = (T)pc1[i] * p0 + (T)pc2[i] * p1 + (T)1 * p2;                          \
                nsq += cp * cp;                                                              \
            }                                                                                \
            alpha = (g0 * g0 + g1 * g1 + g2 * g2) / nsq;                                     \
            g3 = g0; g4 = g1; g5 = g2;                                                       \
            g0 = 0; g1 = 0; g2 = 0;                                                          \
            for (int i = 0; i < N; ++i) {                                                    \
                T res = (T)py[i] - (xm0 * (T)pc1[i] + xm1 * (T)pc2[i] + xm2 * (T)1);         \
                T cp = (T)pc1[i] * p0 + (T)pc2[i] * p1 + (T)1 * p2;                          \
                g0 += (T)pc1[i] * (res - alpha * cp);                                        \
                g1 += (T)pc2[i] * (res - alpha * cp);                                        \
                g2 += (T)1 * (res - alpha * cp);                                             \
            }                                                                                \
            xm0 = xm0 + alpha * p0; xm1 = xm1 + alpha * p1; xm2 = xm2 + alpha * p2;          \
        }                                                                                    \
        T mres = 0;                                                                          \
        for (int i = 0; i < N; ++i)                                                          \
            mres += (T)py[i] - (xm0 * (T)pc1[i] + xm1 * (T)pc2[i] + xm2 * (T)1);             \
        mres /= (T)N;                                                                        \
        T var = 0;                                                                           \
        for (int i = 0; i < N; ++i) {                                                        \
            T res = (T)py[i] - (xm0 * (T)pc1[i] + xm1 * (T)pc2[i] + xm2 * (T)1);             \
            var += (mres - res) * (mres - res);                                              \
        }                                                                                    \
        var_out = (double)(var / (T)(N - 1));                                                \
    }

int lc2o_lc2_cg(const float* U, const float* I, const float* G, const int dims[3], int ndim,
                int r, int zlo, int zhi, int fp64_arith, float* w_map, float* m_map,
                lc2o_result* out) {
    const int X = dims[0], Y = dims[1], Z = dims[2];
    if (ndim != 2 && ndim != 3) return -1;
    if (ndim == 2 && Z != 1) return -1;
    if (r < 1) return -1;
    const int d = 2 * r + 1;
    const int N = ndim == 3 ? d * d * d : d * d;
    const long long XY = (long long)X * Y, V = XY * Z;
    const double tol = ndim == 3 ? 1e-4 : 1e-6;
    const int outer = ndim == 3 ? Z : Y;
    if (zlo < 0) zlo = 0;
    if (zhi > outer) zhi = outer;
    const int rz = ndim == 3 ? r : 0;
    float* wm = w_map ? w_map : (float*)calloc((size_t)V, sizeof(float));
    float* mm = m_map ? m_map : (float*)calloc((size_t)V, sizeof(float));
    if (w_map) memset(w_map, 0, sizeof(float) * V);
    if (m_map) memset(m_map, 0, sizeof(float) * V);
    long long nv = 0;
#pragma omp parallel reduction(+ : nv)
    {
        float* py = (float*)malloc(sizeof(float) * N);
        float* pc1 = (float*)malloc(sizeof(float) * N);
        float* pc2 = (float*)malloc(sizeof(float) * N);
#pragma omp for schedule(dynamic, 1)
        for (int o = zlo; o < zhi; ++o) {
            int z0 = ndim == 3 ? o : 0;
            int ylo = ndim == 3 ? 0 : o, yhi = ndim == 3 ? Y : o + 1;
            for (int y = ylo; y < yhi; ++y)
                for (int x = 0; x < X; ++x) {
                    if (x < r || y < r || x + r >= X || y + r >= Y) continue;
                    if (ndim == 3 && (z0 < r || z0 + r >= Z)) continue;
                    int n = 0, cnt = 0;
                    float sum = 0.0f;
                    for (int dz = -rz; dz <= rz; ++dz)
                        for (int dy = -r; dy <= r; ++dy)
                            for (int dx = -r; dx <= r; ++dx) {
                                long long i = (long long)(x + dx) + (long long)(y + dy) * X +
                                              (long long)(z0 + dz) * XY;
                                py[n] = U[i];
                                pc1[n] = I[i];
                                pc2[n] = G[i];
                                if ((double)U[i] > LC2O_ZERO_THRESHOLD) {
                                    ++cnt;
                                    sum += U[i];
                                }
                                ++n;
                            }
                    if (cnt != N) continue;
                    float mean = sum / (float)cnt, var = 0.0f;
                    for (int i = 0; i < N; ++i) var = fmaf(py[i] - mean, py[i] - mean, var);
                    var /= (float)(cnt - 1);
                    if (!((double)var > LC2O_ZERO_THRESHOLD)) continue;
                    float w = sqrtf(var);
                    double var_out = 0.0;
                    if (fp64_arith) {
                        CG_BODY(double, sqrt)
                    } else {
                        CG_BODY(float, sqrtf)
                    }
                    float meas = w - (float)var_out / w;
                    if (isnan(meas) || isinf(meas) || !(meas > 0.0f)) meas = 0.0f;
                    long long vi = (long long)x + (long long)y * X + (long long)z0 * XY;
                    wm[vi] = w;
                    mm[vi] = meas;
                    nv += 1;
                }
        }
        free(py);
        free(pc1);
        free(pc2);
    }
    double sm = (double)pairwise_sum_f(mm, V), sw = (double)pairwise_sum_f(wm, V);
    out->sum_measure = sm;
    out->sum_w = sw;
    out->n_valid = nv;
    double l = sm / sw;
    out->lc2 = (isnan(l) || isinf(l)) ? 0.0 : l;
    if (!w_map) free(wm);
    if (!m_map) free(mm);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* One evaluation in the reference's data flow                               */
/* ------------------------------------------------------------------------ */

/* Interpolator*::Interpolate + GetSimilarityMeasure (interpolator.cu:583-631,
 * :1266-1380, :1688-1727): both images are warped into the virtual grid (fixed
 * at identity, moving by m12); HighResolutionModality picks which warped image
 * is the high-resolution I (interpolator.cu:590-605); Sobel of I; LC2. */
int lc2o_eval(const float* fixed, const lc2o_geom* fg, const float* moving, const lc2o_geom* mg,
              const lc2o_geom* vg, int grid_ndim, int src_ndim, int high_res_is_moving, int r,
              const float m12[12], int solver_cg, int sobel_order, lc2o_result* out) {
    const long long V = (long long)vg->dims[0] * vg->dims[1] * vg->dims[2];
    float* wf = (float*)malloc(sizeof(float) * V);
    float* wmv = (float*)malloc(sizeof(float) * V);
    float* g = (float*)malloc(sizeof(float) * V);
    if (src_ndim == 3) {
        const float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        lc2o_warp_3d(fixed, fg, vg, ident, wf);
        lc2o_warp_3d(moving, mg, vg, m12, wmv);
    } else {
        const float ident[6] = {1, 0, 0, 0, 1, 0};
        lc2o_warp_2d(fixed, fg, vg, ident, wf);
        lc2o_warp_2d(moving, mg, vg, m12, wmv);
    }
    const float* Iimg = high_res_is_moving ? wmv : wf;
    const float* Uimg = high_res_is_moving ? wf : wmv;
    if (grid_ndim == 3) {
        if (sobel_order == 0) lc2o_sobel_3d_ref(Iimg, vg->dims, g);
        else lc2o_sobel_3d_sep(Iimg, vg->dims, g);
    } else {
        if (sobel_order == 0) lc2o_sobel_2d_ref(Iimg, vg->dims, g);
        else lc2o_sobel_2d_sep(Iimg, vg->dims, g);
    }
    int outer = grid_ndim == 3 ? vg->dims[2] : vg->dims[1];
    int rc;
    if (solver_cg)
        rc = lc2o_lc2_cg(Uimg, Iimg, g, vg->dims, grid_ndim, r, 0, outer, solver_cg == 2, NULL,
                         NULL, out);
    else
        rc = lc2o_lc2_exact(Uimg, Iimg, g, vg->dims, grid_ndim, r, 0, outer, NULL, NULL, out);
    free(wf);
    free(wmv);
    free(g);
    return rc;
}

/* cu_lc2_3d.cu:873-923: bin of v. The product with (bins - 1.0) is formed in double, as
 * the reference's expression does, and roundf() rounds its float conversion. */
static int nmi_bin(float v, float mn, float mx, int bins) {
    const double t = (double)(v - mn) * (bins - 1.0) / (double)(mx - mn);
    return (int)roundf((float)t);
}

static double nmi_entropy(const int* h, int nb, long long n, int fp32) {
    /* cu_lc2_helper.cuh:20-41: p = count / n, an empty bin counts as p = 1 (0 contribution) */
    if (fp32) {
        float acc = 0.0f;
        for (int b = 0; b < nb; ++b) {
            const float p = h[b] == 0 ? 1.0f : (float)h[b] / (float)n;
            acc += p * logf(p);
        }
        return -(double)acc;
    }
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) {
        if (h[b] == 0) continue;
        const double p = (double)h[b] / (double)n;
        acc += p * log(p);
    }
    return -acc;
}

double lc2o_nmi(const float* U, const float* I, long long n, int bins_u, int bins_i,
                int fp32_entropy, int* hist_u, int* hist_i, int* hist_joint) {
    /* cu_lc2_3d.cu:1196-1201: extremes over the whole grid */
    float mnu = U[0], mxu = U[0], mni = I[0], mxi = I[0];
    for (long long k = 1; k < n; ++k) {
        if (U[k] < mnu) mnu = U[k];
        if (U[k] > mxu) mxu = U[k];
        if (I[k] < mni) mni = I[k];
        if (I[k] > mxi) mxi = I[k];
    }
    /* a constant image has no bin width (the reference divides by zero): degenerate */
    if (!(mxu > mnu) || !(mxi > mni)) return 0.0;
    int* hu = (int*)calloc((size_t)bins_u, sizeof(int));
    int* hi = (int*)calloc((size_t)bins_i, sizeof(int));
    int* hj = (int*)calloc((size_t)bins_u * bins_i, sizeof(int));
    for (long long k = 0; k < n; ++k) {
        const int bu = nmi_bin(U[k], mnu, mxu, bins_u);
        const int bi = nmi_bin(I[k], mni, mxi, bins_i);
        const int oku = bu >= 0 && bu <= bins_u - 1, oki = bi >= 0 && bi <= bins_i - 1;
        if (oku) hu[bu]++;
        if (oki) hi[bi]++;
        if (oku && oki) hj[bu + bins_u * bi]++;
    }
    const double h_u = nmi_entropy(hu, bins_u, n, fp32_entropy);
    const double h_i = nmi_entropy(hi, bins_i, n, fp32_entropy);
    const double h_j = nmi_entropy(hj, bins_u * bins_i, n, fp32_entropy);
    if (hist_u) memcpy(hist_u, hu, sizeof(int) * bins_u);
    if (hist_i) memcpy(hist_i, hi, sizeof(int) * bins_i);
    if (hist_joint) memcpy(hist_joint, hj, sizeof(int) * (size_t)bins_u * bins_i);
    free(hu);
    free(hi);
    free(hj);
    if (fp32_entropy) return (double)(((float)h_u + (float)h_i) / (float)h_j);
    return (h_u + h_i) / h_j;
}
