// Device math shared by the HIP kernels: density-normalised real spherical harmonics,
// Slater Fourier-Bessel transforms and the P_lm frame rotation.
#pragma once
#include <hip/hip_runtime.h>

#include "hc_tables.h"

namespace discamb_amd {

__device__ __constant__ double c_density_norm[25];
__device__ __constant__ double c_gcoef[400];
__device__ __constant__ int c_goff[5][13];
// Projection quadrature for the P_lm rotation: 5 Gauss-Legendre nodes in cos(theta) x 10 equally
// spaced azimuths integrate products of two harmonics with l <= 4 exactly.  c_proj_dir holds
// the 50 unit vectors, c_proj_w[p][i] = w_p d_i(u_p) / int d_i^2.
__device__ __constant__ double c_proj_dir[50][3];
__device__ __constant__ double c_proj_w[50][25];

template <typename T>
__device__ __forceinline__ void ylm_density(T x, T y, T z, T* d)
{
    T xx = x * x, yy = y * y, zz = z * z;
    d[0] = T(1);
    d[1] = y; d[2] = z; d[3] = x;
    d[4] = x * y; d[5] = y * z; d[6] = T(2) * zz - xx - yy; d[7] = x * z; d[8] = xx - yy;
    d[9] = (T(3) * xx - yy) * y; d[10] = x * y * z; d[11] = y * (T(4) * zz - xx - yy);
    d[12] = z * (T(2) * zz - T(3) * xx - T(3) * yy); d[13] = x * (T(4) * zz - xx - yy);
    d[14] = (xx - yy) * z; d[15] = (xx - T(3) * yy) * x;
    d[16] = x * y * (xx - yy); d[17] = (T(3) * xx - yy) * y * z; d[18] = x * y * (T(7) * zz - T(1));
    d[19] = y * z * (T(7) * zz - T(3)); d[20] = T(35) * zz * zz - T(30) * zz + T(3);
    d[21] = x * z * (T(7) * zz - T(3)); d[22] = (xx - yy) * (T(7) * zz - T(1));
    d[23] = (xx - T(3) * yy) * x * z; d[24] = xx * (xx - T(3) * yy) - yy * (T(3) * xx - yy);
#pragma unroll
    for (int i = 0; i < 25; i++) d[i] *= T(c_density_norm[i]);
}

// Gradients (d/dx, d/dy, d/dz) of the 25 density-normalised harmonic polynomials of ylm_density,
// taken as polynomials in (x, y, z).
__device__ __forceinline__ void ylm_density_grad(double x, double y, double z, double (*g)[3])
{
    const double xx = x * x, yy = y * y, zz = z * z;
    const double q[25][3] = {
        {0, 0, 0},
        {0, 1, 0}, {0, 0, 1}, {1, 0, 0},
        {y, x, 0}, {0, z, y}, {-2 * x, -2 * y, 4 * z}, {z, 0, x}, {2 * x, -2 * y, 0},
        {6 * x * y, 3 * xx - 3 * yy, 0}, {y * z, x * z, x * y}, {-2 * x * y, 4 * zz - xx - 3 * yy, 8 * y * z},
        {-6 * x * z, -6 * y * z, 6 * zz - 3 * xx - 3 * yy}, {4 * zz - 3 * xx - yy, -2 * x * y, 8 * x * z},
        {2 * x * z, -2 * y * z, xx - yy}, {3 * xx - 3 * yy, -6 * x * y, 0},
        {3 * xx * y - y * yy, x * xx - 3 * x * yy, 0}, {6 * x * y * z, 3 * xx * z - 3 * yy * z, 3 * xx * y - y * yy},
        {y * (7 * zz - 1), x * (7 * zz - 1), 14 * x * y * z}, {0, 7 * z * zz - 3 * z, 21 * y * zz - 3 * y},
        {0, 0, 140 * z * zz - 60 * z}, {7 * z * zz - 3 * z, 0, 21 * x * zz - 3 * x},
        {2 * x * (7 * zz - 1), -2 * y * (7 * zz - 1), 14 * z * (xx - yy)},
        {3 * xx * z - 3 * yy * z, -6 * x * y * z, x * xx - 3 * x * yy},
        {4 * x * xx - 12 * x * yy, -12 * xx * y + 4 * y * yy, 0}};
    for (int i = 0; i < 25; i++)
        for (int c = 0; c < 3; c++) g[i][c] = q[i][c] * c_density_norm[i];
}

// int_0^inf r^n exp(-z r) j_l(k r) dr with k = 2 pi h (FP64)
__device__ __forceinline__ double g_function(int l, int n, double h, double z)
{
    const double k = 6.283185307179586476925 * h;
    const int off = c_goff[l][n];
    const double d = k * k + z * z;
    // homogeneous numerator sum_i c_i k^i z^(n-1-i), nested in k with rising powers of z
    double acc = c_gcoef[off + n - 1], zp = 1.0;
    for (int i = n - 2; i >= 0; i--) {
        zp *= z;
        acc = acc * k + c_gcoef[off + i] * zp;
    }
    const double num = acc;
    double dn = 1.0;
    for (int i = 0; i < n; i++) dn *= d;
    return num / dn;
}

// g(l, n, h, z) and its derivative with respect to h (K = 2 pi h):
// d/dK [P / d^n] = (P_K d - 2 n K P) / d^(n+1)
__device__ __forceinline__ double g_function_dh(int l, int n, double h, double z, double* dg_dh)
{
    const double k = 6.283185307179586476925 * h;
    const int off = c_goff[l][n];
    const double d = k * k + z * z;
    double acc = c_gcoef[off + n - 1], dacc = (n - 1) * c_gcoef[off + n - 1], zp = 1.0;
    for (int i = n - 2; i >= 0; i--) {
        zp *= z;
        acc = acc * k + c_gcoef[off + i] * zp;
        if (i >= 1) dacc = dacc * k + i * c_gcoef[off + i] * zp;
    }
    // dacc holds sum_i i c_i k^(i-1) z^(n-1-i) once the loop has consumed i = 1
    double dn = 1.0;
    for (int i = 0; i < n; i++) dn *= d;
    const double g = acc / dn;
    const double dgdk = (dacc * d - 2.0 * n * k * acc) / (dn * d);
    *dg_dh = 6.283185307179586476925 * dgdk;
    return g;
}

}  // namespace discamb_amd
