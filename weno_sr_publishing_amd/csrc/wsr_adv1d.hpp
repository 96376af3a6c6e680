// wsr_adv1d.hpp -- face reconstruction of the 1-D advection path (torch semantics
// of the reference's training / GSTE code; paths relative to the reference checkout):
//   features: code/snapshot/teacherfree_lab_weno5/weno5_core.py:186-201 (scale (log10+16)/16)
//             followed by the V9 remap code/snapshot/teacherfree_lab_weno5_v9_exact_fvm/v9_model.py:16-28
//   MLP:      weno5_core.py:43-55 (h = x @ W + b, swish, 6 tanh(z/6), softmax)
//   weights:  weno5_core.py:220-239 (omega_from_ratio / classical_omega), plateau -> d (:133-137)
#pragma once

#include "wsr_math.hpp"

namespace wsr {

static inline float host_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;
    u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    float y;
    memcpy(&y, &u, 4);
    return y;
}

// weno5_core.py:133-137 plateau_mask
// The feature arithmetic is evaluated unfused, operation by operation as torch
// does, so that exact ties (a max of equal second differences, a zero
// difference under abs) fall the same way and the adjoint picks the same
// subgradient as autograd.
__device__ __forceinline__ bool plateau5(const double q[5]) {
#pragma clang fp contract(off)
    double d20 = q[0] - 2.0 * q[1] + q[2];
    double d21 = q[1] - 2.0 * q[2] + q[3];
    double d22 = q[2] - 2.0 * q[3] + q[4];
    const double c1312 = 13.0 / 12.0;
    double delta0 = c1312 * fabs(d20) + 0.25 * fabs(q[0] - 4.0 * q[1] + 3.0 * q[2]);
    double delta1 = c1312 * fabs(d21) + 0.25 * fabs(q[1] - q[3]);
    double delta2 = c1312 * fabs(d22) + 0.25 * fabs(3.0 * q[2] - 4.0 * q[3] + q[4]);
    double delta_max = dmax(dmax(delta0, delta1), delta2);
    double q_scale = dmax(dmax(dmax(dmax(fabs(q[0]), fabs(q[1])), fabs(q[2])), dmax(fabs(q[3]), fabs(q[4]))), 1.0);
    return delta_max <= 1.0e-13 * q_scale;
}

// plateau test and features of one stencil; false = plateau (features not formed)
__device__ __forceinline__ bool features_1d(const double q[5], double f[5]) {
#pragma clang fp contract(off)
    double d20 = q[0] - 2.0 * q[1] + q[2];
    double d21 = q[1] - 2.0 * q[2] + q[3];
    double d22 = q[2] - 2.0 * q[3] + q[4];
    const double c1312 = 13.0 / 12.0;
    double delta0 = c1312 * fabs(d20) + 0.25 * fabs(q[0] - 4.0 * q[1] + 3.0 * q[2]);
    double delta1 = c1312 * fabs(d21) + 0.25 * fabs(q[1] - q[3]);
    double delta2 = c1312 * fabs(d22) + 0.25 * fabs(3.0 * q[2] - 4.0 * q[3] + q[4]);
    double delta_max = dmax(dmax(delta0, delta1), delta2);
    double q_scale = dmax(dmax(dmax(dmax(fabs(q[0]), fabs(q[1])), fabs(q[2])), dmax(fabs(q[3]), fabs(q[4]))), 1.0);
    if (delta_max <= 1.0e-13 * q_scale) return false;
    const double eps = 1.0e-15;
    double g0 = fabs(d20) / (fabs(q[1] - q[0]) + fabs(q[2] - q[1]) + eps);
    double g1 = fabs(d21) / (fabs(q[2] - q[1]) + fabs(q[3] - q[2]) + eps);
    double g2 = fabs(d22) / (fabs(q[3] - q[2]) + fabs(q[4] - q[3]) + eps);
    double inv_delta_max = 1.0 / dmax(delta_max, 1.0e-15);
    double relative = dmax(delta_max / q_scale, 1.0e-30);
    double old_scale = dmin(dmax((log10(relative) + 16.0) / 16.0, 0.0), 1.0);
    f[0] = delta0 * inv_delta_max;
    f[1] = delta1 * inv_delta_max;
    f[2] = delta2 * inv_delta_max;
    f[3] = dmin(dmax(dmax(dmax(g0, g1), g2), 0.0), 1.0);
    f[4] = dmin(dmax((16.0 * old_scale - 16.0 + 4.0) / 4.0, 0.0), 1.0);
    return true;
}

// parameters read through a device pointer (training kernels); P.p[i] works for
// both this view and the by-value Mlp64 / Mlp32 blocks
struct MlpDev64 { const double *__restrict__ p; };
struct MlpDev32 { const float *__restrict__ p; };

template <class MP>
__device__ __forceinline__ void mlp_torch64(const double x[5], const MP &P, double r[3]) {
    double h1[10], h2[6], h3[6];
#pragma unroll
    for (int o = 0; o < 10; ++o) {
        double z = x[0] * P.p[OFF_W1 + o];
#pragma unroll
        for (int k = 1; k < 5; ++k) z = z + x[k] * P.p[OFF_W1 + k * 10 + o];
        z = z + P.p[OFF_B1 + o];
        h1[o] = z * (1.0 / (1.0 + exp(-z)));
    }
#pragma unroll
    for (int o = 0; o < 6; ++o) {
        double z = h1[0] * P.p[OFF_W2 + o];
#pragma unroll
        for (int k = 1; k < 10; ++k) z = z + h1[k] * P.p[OFF_W2 + k * 6 + o];
        z = z + P.p[OFF_B2 + o];
        h2[o] = z * (1.0 / (1.0 + exp(-z)));
    }
#pragma unroll
    for (int o = 0; o < 6; ++o) {
        double z = h2[0] * P.p[OFF_W3 + o];
#pragma unroll
        for (int k = 1; k < 6; ++k) z = z + h2[k] * P.p[OFF_W3 + k * 6 + o];
        z = z + P.p[OFF_B3 + o];
        h3[o] = z * (1.0 / (1.0 + exp(-z)));
    }
    double b[3];
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        double z = h3[0] * P.p[OFF_W4 + o];
#pragma unroll
        for (int k = 1; k < 6; ++k) z = z + h3[k] * P.p[OFF_W4 + k * 3 + o];
        b[o] = 6.0 * tanh((z + P.p[OFF_B4 + o]) / 6.0);
    }
    double m = dmax(dmax(b[0], b[1]), b[2]);
    double e0 = exp(b[0] - m), e1 = exp(b[1] - m), e2 = exp(b[2] - m);
    double s = e0 + e1 + e2;
    r[0] = e0 / s; r[1] = e1 / s; r[2] = e2 / s;
}

template <bool BF16, class MP>
__device__ __forceinline__ void mlp_torch32(const float xin[5], const MP &P, float r[3]) {
    float x[5], h1[10], h2[6], h3[6];
#pragma unroll
    for (int k = 0; k < 5; ++k) x[k] = BF16 ? round_bf16(xin[k]) : xin[k];
#pragma unroll
    for (int o = 0; o < 10; ++o) {
        float z = x[0] * P.p[OFF_W1 + o];
#pragma unroll
        for (int k = 1; k < 5; ++k) z = fmaf(x[k], P.p[OFF_W1 + k * 10 + o], z);
        z = z + P.p[OFF_B1 + o];
        h1[o] = BF16 ? round_bf16(swish32(z)) : swish32(z);
    }
#pragma unroll
    for (int o = 0; o < 6; ++o) {
        float z = h1[0] * P.p[OFF_W2 + o];
#pragma unroll
        for (int k = 1; k < 10; ++k) z = fmaf(h1[k], P.p[OFF_W2 + k * 6 + o], z);
        z = z + P.p[OFF_B2 + o];
        h2[o] = BF16 ? round_bf16(swish32(z)) : swish32(z);
    }
#pragma unroll
    for (int o = 0; o < 6; ++o) {
        float z = h2[0] * P.p[OFF_W3 + o];
#pragma unroll
        for (int k = 1; k < 6; ++k) z = fmaf(h2[k], P.p[OFF_W3 + k * 6 + o], z);
        z = z + P.p[OFF_B3 + o];
        h3[o] = BF16 ? round_bf16(swish32(z)) : swish32(z);
    }
    float b[3];
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        float z = h3[0] * P.p[OFF_W4 + o];
#pragma unroll
        for (int k = 1; k < 6; ++k) z = fmaf(h3[k], P.p[OFF_W4 + k * 3 + o], z);
        b[o] = 6.0f * tanhf((z + P.p[OFF_B4 + o]) / 6.0f);
    }
    float m = fmax_(fmax_(b[0], b[1]), b[2]);
    float e0 = exp_fast(b[0] - m), e1 = exp_fast(b[1] - m), e2 = exp_fast(b[2] - m);
    float inv = rcp_fast(e0 + e1 + e2);
    r[0] = e0 * inv; r[1] = e1 * inv; r[2] = e2 * inv;
}

template <int SCHEME, class MLP>
__device__ __forceinline__ void sym_ratios_1d(const double f[5], const MLP &P, double r[3]) {
    double d[3], e[3];
    if constexpr (SCHEME == SCHEME_SR64) {
        double fr[5] = {f[2], f[1], f[0], f[3], f[4]};
        mlp_torch64(f, P, d);
        mlp_torch64(fr, P, e);
    } else {
        float xd[5] = {(float)f[0], (float)f[1], (float)f[2], (float)f[3], (float)f[4]};
        float xr[5] = {xd[2], xd[1], xd[0], xd[3], xd[4]};
        float rd[3], re[3];
        mlp_torch32<SCHEME == SCHEME_SRBF16>(xd, P, rd);
        mlp_torch32<SCHEME == SCHEME_SRBF16>(xr, P, re);
        for (int k = 0; k < 3; ++k) { d[k] = rd[k]; e[k] = re[k]; }
    }
    r[0] = 0.5 * (d[0] + e[2]);
    r[1] = 0.5 * (d[1] + e[1]);
    r[2] = 0.5 * (d[2] + e[0]);
}

// face value x_{i+1/2} from the stencil q (lr1: cells i-2..i+2; else cells i-1..i+3, lr=2 of cell i+1)
template <int SCHEME, class MLP>
__device__ __forceinline__ double face_1d(const double q[5], bool lr1, const MLP &P) {
    double s[3], d[3], w[3];
    candidates_normal(q, lr1, s);
    d_normal(lr1, d);
    if constexpr (SCHEME == SCHEME_JS) {
        if (plateau5(q)) {
            w[0] = d[0]; w[1] = d[1]; w[2] = d[2];
        } else {
            double b[3];
            js_beta(q, b);
            double a0 = d[0] / ((b[0] + 1.0e-6) * (b[0] + 1.0e-6));
            double a1 = d[1] / ((b[1] + 1.0e-6) * (b[1] + 1.0e-6));
            double a2 = d[2] / ((b[2] + 1.0e-6) * (b[2] + 1.0e-6));
            double sum = a0 + a1 + a2;
            w[0] = a0 / sum; w[1] = a1 / sum; w[2] = a2 / sum;
        }
    } else {
        double f[5];
        if (!features_1d(q, f)) {
            w[0] = d[0]; w[1] = d[1]; w[2] = d[2];
        } else {
            double r[3];
            sym_ratios_1d<SCHEME>(f, P, r);
            double a0 = d[0] / ((3.0 * r[0] + 1.0e-12) * (3.0 * r[0] + 1.0e-12));
            double a1 = d[1] / ((3.0 * r[1] + 1.0e-12) * (3.0 * r[1] + 1.0e-12));
            double a2 = d[2] / ((3.0 * r[2] + 1.0e-12) * (3.0 * r[2] + 1.0e-12));
            double sum = a0 + a1 + a2;
            w[0] = a0 / sum; w[1] = a1 / sum; w[2] = a2 / sum;
        }
    }
    return w[0] * s[0] + w[1] * s[1] + w[2] * s[2];
}

// upwind face value at x_{i+1/2} of a periodic row (rk3_advection.py:82-104)
template <int SCHEME, class MLP>
__device__ __forceinline__ double face_signed(const double *row, int n, int i, double a, const MLP &P) {
    double q[5];
    const int base = a > 0.0 ? i - 2 : i - 1;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        int k = base + m;
        k = k < 0 ? k + n : (k >= n ? k - n : k);
        q[m] = row[k];
    }
    return face_1d<SCHEME>(q, a > 0.0, P);
}

}  // namespace wsr
