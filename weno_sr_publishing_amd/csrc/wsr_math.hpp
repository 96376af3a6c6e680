// wsr_math.hpp -- device arithmetic of the WENO5 finite-volume step (gfx950).
//
// Every function restates a reference formula (paths relative to the reference
// checkout; H5 = code/snapshot/warp_weno5_helpers.py, C64 = code/snapshot/
// teacherfree_lab_weno5_v20_distance_balanced/weno5_rk3_diff_v20_deploy.py,
// C32 = code/snapshot/teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/
// warp_mlp_f32/weno5_rk3_diff_mlp_f32.py, AJS = code/snapshot/for_paper_results/
// solvers/weno5_hllc.py, WZ = code/snapshot/weno_z_borges_p2_results/
// weno5_helpers_z.py).  The expressions keep the reference's structure; the
// compiler is free to fuse multiply-adds, so results agree with an unfused
// evaluation to rounding, not bitwise.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace wsr {

enum { SCHEME_JS = 0, SCHEME_SR64 = 1, SCHEME_SR32 = 2, SCHEME_SRBF16 = 3, SCHEME_Z = 4 };
enum { RIEMANN_HLLC = 0, RIEMANN_FORCE = 1 };
enum { HEAD_NL1 = 0, HEAD_NL2 = 1, HEAD_GL1 = 2, HEAD_GL2 = 3 };

constexpr int MLP5_COUNT = 189;
// parameter block: w1[5][10] b1[10] w2[10][6] b2[6] w3[6][6] b3[6] w4[6][3] b4[3]
constexpr int OFF_W1 = 0, OFF_B1 = 50, OFF_W2 = 60, OFF_B2 = 120, OFF_W3 = 126, OFF_B3 = 162,
              OFF_W4 = 168, OFF_B4 = 186;

struct Mlp64 { double p[MLP5_COUNT]; };
struct Mlp32 { float p[MLP5_COUNT]; };
struct MlpNone { int unused; };

// by-value parameter blocks (1-D kernels)
template <int SCHEME> struct MlpFor { using type = MlpNone; };
template <> struct MlpFor<SCHEME_SR64> { using type = Mlp64; };
template <> struct MlpFor<SCHEME_SR32> { using type = Mlp32; };
template <> struct MlpFor<SCHEME_SRBF16> { using type = Mlp32; };

// Device-resident parameter block of the 2-D sweeps, neuron-major: for every
// layer and output neuron o the run [b_o, w_0o, ..., w_(K-1)o], so that one
// neuron's weights are contiguous and arrive through scalar loads.
constexpr int DEV_L1 = 0, DEV_L2 = 60, DEV_L3 = 126, DEV_L4 = 168;
template <class T> struct MlpK { const T *p; };
template <int SCHEME> struct MlpSweep { using type = MlpNone; };
template <> struct MlpSweep<SCHEME_SR64> { using type = MlpK<double>; };
template <> struct MlpSweep<SCHEME_SR32> { using type = MlpK<float>; };
template <> struct MlpSweep<SCHEME_SRBF16> { using type = MlpK<float>; };

// reference layout -> neuron-major (host or device)
template <class T>
__host__ __device__ inline void pack_neuron_major(const T *ref, T *dev) {
    const int K[4] = {5, 10, 6, 6}, N[4] = {10, 6, 6, 3};
    const int W[4] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4}, B[4] = {OFF_B1, OFF_B2, OFF_B3, OFF_B4};
    int o = 0;
    for (int l = 0; l < 4; ++l)
        for (int n = 0; n < N[l]; ++n) {
            dev[o++] = ref[B[l] + n];
            for (int k = 0; k < K[l]; ++k) dev[o++] = ref[W[l] + k * N[l] + n];
        }
}

// FP32 closure with its activation constants folded into the packed parameters
// (host side, at upload; BF16 keeps the plain parameters).  With L = log2(e) the
// hidden layers compute z' = -L z (layer 1 weights and biases and the biases of
// layers 2 and 3 scaled by -L), so that the swish is h' = z' / (1 + 2^z') = -L swish(z)
// with no multiply, and the output layer computes u = (L / 3) z4 (weights scaled by
// -1/3, which undoes the -L of h', biases by L / 3), the exponent of e^(z4 / 3) that
// the bounded head 6 tanh(z4 / 6) = 6 - 12 / (1 + e^(z4 / 3)) needs.
constexpr double LOG2E = 1.4426950408889634;
inline void fold_fp32_closure(float *dev) {
    const int start[4] = {DEV_L1, DEV_L2, DEV_L3, DEV_L4}, K[4] = {5, 10, 6, 6}, N[4] = {10, 6, 6, 3};
    for (int l = 0; l < 4; ++l)
        for (int n = 0; n < N[l]; ++n) {
            float *run = dev + start[l] + n * (K[l] + 1);   // [bias, w_0 .. w_(K-1)]
            const double wscale = l == 0 ? -LOG2E : (l == 3 ? -1.0 / 3.0 : 1.0);
            const double bscale = l == 3 ? LOG2E / 3.0 : -LOG2E;
            run[0] = (float)((double)run[0] * bscale);
            for (int k = 1; k <= K[l]; ++k) run[k] = (float)((double)run[k] * wscale);
        }
}

// Parameters are read through the constant address space (scalar loads into
// SGPRs, one SGPR operand per FMA).  The empty asm makes the pointer opaque
// before every layer so that the loads stay next to their use instead of being
// hoisted to the kernel entry (which would exceed the SGPR file and spill).
template <class T> using kconst_ptr = const T __attribute__((address_space(4))) *;
template <class T>
__device__ __forceinline__ kconst_ptr<T> kparams(const T *p) {
    kconst_ptr<T> k = (kconst_ptr<T>)p;
    asm volatile("" : "+s"(k));
    return k;
}

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ float fmax_(float a, float b) { return a > b ? a : b; }

// ---------------------------------------------------------------- Euler algebra
// H5:246-252 con_to_pri
__device__ __forceinline__ void con_to_pri(const double q[4], double gamma, double w[4]) {
    double rho = q[0];
    double u = q[1] / rho;
    double v = q[2] / rho;
    w[0] = rho; w[1] = u; w[2] = v;
    w[3] = (gamma - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v));
}

// H5:264-285 pri_to_flux
template <int DIR>
__device__ __forceinline__ void pri_to_flux(const double w[4], double gamma, double f[4]) {
    double rho = w[0], u = w[1], v = w[2], p = w[3];
    double e = 0.5 * rho * (u * u + v * v) + p / (gamma - 1.0);
    if (DIR == 1) {
        f[0] = rho * u; f[1] = rho * u * u + p; f[2] = rho * u * v; f[3] = u * (p + e);
    } else {
        f[0] = rho * v; f[1] = rho * u * v; f[2] = rho * v * v + p; f[3] = v * (p + e);
    }
}

__device__ __forceinline__ double floor_signed(double d, double tiny) {
    if (fabs(d) < tiny) return d < 0.0 ? -tiny : tiny;
    return d;
}

// AJS:23-115 hllc_flux (Davis bounds, floors 1e-16)
template <int DIR>
__device__ void hllc_flux(const double ul0[4], const double ur0[4], double gamma, double out[4]) {
    const double tiny = 1.0e-16;
    double wl[4], wr[4];
    con_to_pri(ul0, gamma, wl);
    con_to_pri(ur0, gamma, wr);
    double rho_l = dmax(wl[0], tiny), rho_r = dmax(wr[0], tiny);
    double p_l = dmax(wl[3], tiny), p_r = dmax(wr[3], tiny);
    double a_l = sqrt(gamma * p_l / rho_l);
    double a_r = sqrt(gamma * p_r / rho_r);
    double un_l = DIR == 1 ? wl[1] : wl[2];
    double un_r = DIR == 1 ? wr[1] : wr[2];
    double s_l = dmin(un_l - a_l, un_r - a_r);
    double s_r = dmax(un_l + a_l, un_r + a_r);
    double f_l[4], f_r[4];
    pri_to_flux<DIR>(wl, gamma, f_l);
    pri_to_flux<DIR>(wr, gamma, f_r);
    if (0.0 <= s_l) { for (int c = 0; c < 4; ++c) out[c] = f_l[c]; return; }
    if (0.0 >= s_r) { for (int c = 0; c < 4; ++c) out[c] = f_r[c]; return; }
    double denom = floor_signed(rho_l * (s_l - un_l) - rho_r * (s_r - un_r), tiny);
    double s_star = (p_r - p_l + rho_l * un_l * (s_l - un_l) - rho_r * un_r * (s_r - un_r)) / denom;
    const bool right = s_star < 0.0;
    double u_side[4], w_side[4], f_side[4];
    for (int c = 0; c < 4; ++c) {
        u_side[c] = right ? ur0[c] : ul0[c];
        w_side[c] = right ? wr[c] : wl[c];
        f_side[c] = right ? f_r[c] : f_l[c];
    }
    double s_side = right ? s_r : s_l;
    double rho = dmax(w_side[0], tiny);
    double u = w_side[1], v = w_side[2];
    double p = dmax(w_side[3], tiny);
    double un = DIR == 1 ? u : v;
    double denom_star = floor_signed(s_side - s_star, tiny);
    double factor = rho * (s_side - un) / denom_star;
    double u_star[4];
    u_star[0] = factor;
    u_star[1] = DIR == 1 ? factor * s_star : factor * u;
    u_star[2] = DIR == 1 ? factor * v : factor * s_star;
    double denom_side = floor_signed(s_side - un, tiny);
    u_star[3] = factor * ((u_side[3] / rho) + (s_star - un) * (s_star + p / (rho * denom_side)));
    for (int c = 0; c < 4; ++c) out[c] = f_side[c] + s_side * (u_star[c] - u_side[c]);
}

// H5:401-435 force_flux, c = dt / h
template <int DIR>
__device__ void force_flux(const double ul0[4], const double ur0[4], double c, double gamma, double out[4]) {
    double wl0[4], wr0[4], fl[4], fr[4], ri_u[4], wri[4], ri[4];
    con_to_pri(ul0, gamma, wl0);
    con_to_pri(ur0, gamma, wr0);
    pri_to_flux<DIR>(wl0, gamma, fl);
    pri_to_flux<DIR>(wr0, gamma, fr);
    for (int k = 0; k < 4; ++k) ri_u[k] = 0.5 * (ul0[k] + ur0[k]) - 0.5 * c * (fr[k] - fl[k]);
    con_to_pri(ri_u, gamma, wri);
    pri_to_flux<DIR>(wri, gamma, ri);
    double dx_dt = 1.0 / c;
    for (int k = 0; k < 4; ++k) {
        double lf = 0.5 * (fl[k] + fr[k]) - 0.5 * dx_dt * (ur0[k] - ul0[k]);
        out[k] = 0.5 * (lf + ri[k]);
    }
}

// ------------------------------------------- Roe average + eigen-structure (H5:598-748)
struct Eigen {
    double u, v, a, h;   // Jacobian values of the Roe state (H5:631-642)
};

__device__ __forceinline__ Eigen roe_eigen(const double L[4], const double R[4], double gamma) {
    const double tiny = 1.0e-15;
    double gm1 = gamma - 1.0;
    // H5:598-628 roe_average_state
    double rho_l = dmax(L[0], tiny);
    double u_l = L[1] / rho_l, v_l = L[2] / rho_l, e_l = L[3];
    double p_l = dmax(gm1 * (e_l - 0.5 * rho_l * (u_l * u_l + v_l * v_l)), tiny);
    double h_l = (e_l + p_l) / rho_l;
    double rho_r = dmax(R[0], tiny);
    double u_r = R[1] / rho_r, v_r = R[2] / rho_r, e_r = R[3];
    double p_r = dmax(gm1 * (e_r - 0.5 * rho_r * (u_r * u_r + v_r * v_r)), tiny);
    double h_r = (e_r + p_r) / rho_r;
    double sqrt_l = sqrt(rho_l), sqrt_r = sqrt(rho_r);
    double inv = 1.0 / (sqrt_l + sqrt_r);
    double u_roe = (sqrt_l * u_l + sqrt_r * u_r) * inv;
    double v_roe = (sqrt_l * v_l + sqrt_r * v_r) * inv;
    double h_roe = (sqrt_l * h_l + sqrt_r * h_r) * inv;
    double rho_roe = sqrt_l * sqrt_r;
    double q2_roe = u_roe * u_roe + v_roe * v_roe;
    double e_roe = (rho_roe * h_roe + 0.5 * gm1 * rho_roe * q2_roe) / gamma;
    double m1 = rho_roe * u_roe, m2 = rho_roe * v_roe;
    // H5:631-642 jacobian_state_values of the conserved Roe state
    double rho = dmax(rho_roe, tiny);
    Eigen E;
    E.u = m1 / rho;
    E.v = m2 / rho;
    double q2 = E.u * E.u + E.v * E.v;
    double p = dmax(gm1 * (e_roe - 0.5 * rho * q2), tiny);
    E.a = sqrt(gamma * p / rho);
    E.h = 0.5 * q2 + E.a * E.a / gm1;
    return E;
}

// Row k of the left eigenvector matrix (H5:645-679 x, H5:699-731 y), already
// multiplied by scale = (gamma-1)/(2 a^2):  c_k = (l0 q0 + l1 q1 + l2 q2 + l3 q3) * scale.
template <int DIR>
__device__ __forceinline__ void left_row(int k, const Eigen &E, double gamma, double l[4], double &scale) {
    double gm1 = gamma - 1.0;
    double inv_gm1 = 1.0 / gm1;
    double a = E.a, h = E.h, a2 = a * a;
    double tn = DIR == 1 ? E.v : E.u;   // tangential velocity
    double nn = DIR == 1 ? E.u : E.v;   // normal velocity
    scale = gm1 / (2.0 * a2);
    // rows in the normal/tangential frame; columns (q0, q_normal, q_tangent, q3)
    double r0 = h + a * (nn - a) * inv_gm1, rn0 = -(nn + a * inv_gm1), rt0 = -tn;
    double r1 = -2.0 * h + 4.0 * a2 * inv_gm1, rn1 = 2.0 * nn, rt1 = 2.0 * tn;
    double r2 = -2.0 * tn * a2 * inv_gm1, rt2 = 2.0 * a2 * inv_gm1;
    double r3 = h - a * (nn + a) * inv_gm1, rn3 = -nn + a * inv_gm1, rt3 = -tn;
    double c0 = k == 0 ? r0 : k == 1 ? r1 : k == 2 ? r2 : r3;
    double cn = k == 0 ? rn0 : k == 1 ? rn1 : k == 2 ? 0.0 : rn3;
    double ct = k == 0 ? rt0 : k == 1 ? rt1 : k == 2 ? rt2 : rt3;
    double c3 = k == 0 ? 1.0 : k == 1 ? -2.0 : k == 2 ? 0.0 : 1.0;
    l[0] = c0;
    l[1] = DIR == 1 ? cn : ct;
    l[2] = DIR == 1 ? ct : cn;
    l[3] = c3;
}

// Row k of the right eigenvector matrix (H5:682-696 x, H5:734-748 y).
template <int DIR>
__device__ __forceinline__ void right_row(int k, const Eigen &E, double r[4]) {
    double a = E.a, h = E.h;
    double tn = DIR == 1 ? E.v : E.u;
    double nn = DIR == 1 ? E.u : E.v;
    double q2 = E.u * E.u + E.v * E.v;
    // component k of the conserved vector: 0 rho, normal momentum, tangential momentum, 3 energy
    const int kn = DIR == 1 ? 1 : 2;
    if (k == 0) { r[0] = 1.0; r[1] = 1.0; r[2] = 0.0; r[3] = 1.0; }
    else if (k == kn) { r[0] = nn - a; r[1] = nn; r[2] = 0.0; r[3] = nn + a; }
    else if (k == 3) { r[0] = h - nn * a; r[1] = 0.5 * q2; r[2] = tn; r[3] = h + nn * a; }
    else { r[0] = tn; r[1] = tn; r[2] = 1.0; r[3] = tn; }
}

// ------------------------------------------------------ candidate stencils
// H5:438-446 stencil_2d (face) and H5:798-810 stencil_gauss (Gauss points,
// the reference's 15-digit constants).
__device__ __forceinline__ void candidates_normal(const double q[5], bool lr1, double s[3]) {
    const double t13 = 1.0 / 3.0, t76 = 7.0 / 6.0, t116 = 11.0 / 6.0, t16 = 1.0 / 6.0, t56 = 5.0 / 6.0;
    if (lr1) {
        s[0] = t13 * q[0] - t76 * q[1] + t116 * q[2];
        s[1] = (-t16) * q[1] + t56 * q[2] + t13 * q[3];
        s[2] = t13 * q[2] + t56 * q[3] - t16 * q[4];
    } else {
        s[0] = (-t16) * q[0] + t56 * q[1] + t13 * q[2];
        s[1] = t13 * q[1] + t56 * q[2] - t16 * q[3];
        s[2] = t116 * q[2] - t76 * q[3] + t13 * q[4];
    }
}

__device__ __forceinline__ void candidates_gauss(const double q[5], double s1[3], double s2[3]) {
    s1[0] = -0.144337567297407 * q[0] + 0.577350269189626 * q[1] + 0.566987298107780 * q[2];
    s1[1] = 0.144337567297406 * q[1] + q[2] - 0.144337567297407 * q[3];
    s1[2] = 1.433012701892220 * q[2] - 0.577350269189626 * q[3] + 0.144337567297406 * q[4];
    s2[0] = 0.144337567297406 * q[0] - 0.577350269189626 * q[1] + 1.433012701892220 * q[2];
    s2[1] = -0.144337567297407 * q[1] + q[2] + 0.144337567297406 * q[3];
    s2[2] = 0.566987298107781 * q[2] + 0.577350269189625 * q[3] - 0.144337567297406 * q[4];
}

// linear weights (C64:136-152)
__device__ __forceinline__ void d_normal(bool lr1, double d[3]) {
    d[0] = lr1 ? 0.1 : 0.3; d[1] = 0.6; d[2] = lr1 ? 0.3 : 0.1;
}
__device__ __forceinline__ void d_gauss(double d1[3], double d2[3]) {
    const double root3 = 1.7320508075688772;  // sqrt(3.0), correctly rounded
    double lo = (210.0 - root3) / 1080.0, hi = (210.0 + root3) / 1080.0;
    d1[0] = hi; d1[1] = 11.0 / 18.0; d1[2] = lo;
    d2[0] = lo; d2[1] = 11.0 / 18.0; d2[2] = hi;
}

// ------------------------------------------------------ WENO5-JS / -Z weights
// H5:545-570 Jiang-Shu smoothness indicators
__device__ __forceinline__ void js_beta(const double q[5], double b[3]) {
    double t2 = q[2] - 2.0 * q[3] + q[4], e2 = 3.0 * q[2] - 4.0 * q[3] + q[4];
    double t1 = q[1] - 2.0 * q[2] + q[3], e1 = q[1] - q[3];
    double t0 = q[0] - 2.0 * q[1] + q[2], e0 = q[0] - 4.0 * q[1] + 3.0 * q[2];
    const double c1312 = 13.0 / 12.0;
    b[2] = c1312 * t2 * t2 + 0.25 * e2 * e2;
    b[1] = c1312 * t1 * t1 + 0.25 * e1 * e1;
    b[0] = c1312 * t0 * t0 + 0.25 * e0 * e0;
}

// H5:556-570: alpha_k = d_k / (1e-6 + beta_k)^2, normalised
__device__ __forceinline__ double js_combine(const double s[3], const double d[3], const double inv[3]) {
    double a0 = d[0] * inv[0] * inv[0], a1 = d[1] * inv[1] * inv[1], a2 = d[2] * inv[2] * inv[2];
    double sum = a0 + a1 + a2;
    return (a0 / sum) * s[0] + (a1 / sum) * s[1] + (a2 / sum) * s[2];
}

// WZ:558-570: alpha_k = d_k (1 + (tau5 / (beta_k + h^2))^2)
__device__ __forceinline__ double z_combine(const double s[3], const double d[3], const double r2[3]) {
    double a0 = d[0] * (1.0 + r2[0]), a1 = d[1] * (1.0 + r2[1]), a2 = d[2] * (1.0 + r2[2]);
    double sum = a0 + a1 + a2;
    return (a0 / sum) * s[0] + (a1 / sum) * s[1] + (a2 / sum) * s[2];
}

// ------------------------------------------------------------ SR features
// C64:77-133 raw sensors, features (scale clip((log10 rel + 4)/4, 0, 1)) and the
// plateau rule.  Returns true on a plateau (delta_max <= 1e-13 q_scale), in
// which case f is not computed.  The reflected feature vector of the reversed
// stencil is the permutation (f2, f1, f0, f3, f4) of this one.
__device__ __forceinline__ bool sr_features(const double q[5], double f[5]) {
    double d20 = q[0] - 2.0 * q[1] + q[2];
    double d21 = q[1] - 2.0 * q[2] + q[3];
    double d22 = q[2] - 2.0 * q[3] + q[4];
    const double c1312 = 13.0 / 12.0;
    double delta0 = c1312 * fabs(d20) + 0.25 * fabs(q[0] - 4.0 * q[1] + 3.0 * q[2]);
    double delta1 = c1312 * fabs(d21) + 0.25 * fabs(q[1] - q[3]);
    double delta2 = c1312 * fabs(d22) + 0.25 * fabs(3.0 * q[2] - 4.0 * q[3] + q[4]);
    double delta_max = dmax(dmax(delta0, delta1), delta2);
    double q_scale = dmax(dmax(dmax(fabs(q[0]), fabs(q[1])), dmax(fabs(q[2]), fabs(q[3]))), fabs(q[4]));
    q_scale = dmax(q_scale, 1.0);
    if (delta_max <= 1.0e-13 * q_scale) return true;
    const double eps = 1.0e-15;
    double gamma0 = fabs(d20) / (fabs(q[1] - q[0]) + fabs(q[2] - q[1]) + eps);
    double gamma1 = fabs(d21) / (fabs(q[2] - q[1]) + fabs(q[3] - q[2]) + eps);
    double gamma2 = fabs(d22) / (fabs(q[3] - q[2]) + fabs(q[4] - q[3]) + eps);
    double gamma_s = dmin(1.0, dmax(dmax(gamma0, gamma1), gamma2));
    double inv_delta_max = 1.0 / dmax(delta_max, eps);
    double relative = dmax(delta_max / q_scale, 1.0e-30);
    double scale = (log(relative) / 2.302585092994046 + 4.0) / 4.0;  // log(10.0)
    scale = dmin(1.0, dmax(0.0, scale));
    f[0] = delta0 * inv_delta_max;
    f[1] = delta1 * inv_delta_max;
    f[2] = delta2 * inv_delta_max;
    f[3] = gamma_s;
    f[4] = scale;
    return false;
}

// ----------------------------------------------------------------- the MLP
// C64:155-207 (FP64) / C32:152-202 (FP32): 5 -> 10 -> 6 -> 6 -> 3, swish hidden
// activations, r = softmax(6 tanh(z / 6)).
__device__ __forceinline__ double swish64(double x) { return x / (1.0 + exp(-x)); }

// Two raw evaluations (direct and reflected input) side by side: every weight
// read feeds two independent FMA chains.
template <int K, int N, class T, class ACT>
__device__ __forceinline__ void dense_pair(const T *xa, const T *xb, const T *w, int base, T *ha, T *hb, ACT act) {
    kconst_ptr<T> c = kparams(w + base);
#pragma unroll
    for (int o = 0; o < N; ++o) {
        if constexpr (sizeof(T) == 8) c = kparams(w + base + o * (K + 1));   // FP64: one neuron at a time
        const int b = sizeof(T) == 8 ? 0 : o * (K + 1);
        T za = c[b], zb = c[b];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            za = za + xa[k] * c[b + 1 + k];
            zb = zb + xb[k] * c[b + 1 + k];
        }
        ha[o] = act(za);
        hb[o] = act(zb);
    }
}

__device__ __forceinline__ void softmax_head64(const double z[3], double r[3]) {
    double b0 = 6.0 * tanh(z[0] / 6.0), b1 = 6.0 * tanh(z[1] / 6.0), b2 = 6.0 * tanh(z[2] / 6.0);
    double bm = dmax(dmax(b0, b1), b2);
    double e0 = exp(b0 - bm), e1 = exp(b1 - bm), e2 = exp(b2 - bm);
    double inv = 1.0 / (e0 + e1 + e2);
    r[0] = e0 * inv; r[1] = e1 * inv; r[2] = e2 * inv;
}

__device__ __forceinline__ void mlp_pair64(const double xa[5], const double xb[5], const double *w, double ra[3],
                                           double rb[3]) {
    double a1[10], b1[10], a2[6], b2[6], a3[6], b3[6], za[3], zb[3];
    auto sw = [](double z) { return swish64(z); };
    auto id = [](double z) { return z; };
    dense_pair<5, 10>(xa, xb, w, DEV_L1, a1, b1, sw);
    dense_pair<10, 6>(a1, b1, w, DEV_L2, a2, b2, sw);
    dense_pair<6, 6>(a2, b2, w, DEV_L3, a3, b3, sw);
    dense_pair<6, 3>(a3, b3, w, DEV_L4, za, zb, id);
    softmax_head64(za, ra);
    softmax_head64(zb, rb);
}

// FP32 closure on the VALU: v_exp_f32 / v_rcp_f32 for the swish, the bounded
// logits and the softmax.
__device__ __forceinline__ float exp_fast(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float swish32(float x) { return x * rcp_fast(1.0f + exp_fast(-x)); }
// tanh(y) = sign(y) (1 - e) / (1 + e), e = exp(-2|y|): absolute error ~1e-7
__device__ __forceinline__ float tanh_fast(float y) {
    float e = exp_fast(-2.0f * fabsf(y));
    return copysignf((1.0f - e) * rcp_fast(1.0f + e), y);
}

__device__ __forceinline__ float round_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;
    u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    return __uint_as_float(u);
}

// The FP32 closure on folded parameters (fold_fp32_closure): swish h' = z' / (1 + 2^z'),
// head e_k = 2^(6L - 12L / (1 + 2^u_k)) = e^(6 tanh(z_k / 6)) in (e^-6, e^6), softmax
// without the max shift (the bounded logits cannot overflow).
__device__ __forceinline__ float swish_folded(float zp) { return zp * rcp_fast(1.0f + __builtin_amdgcn_exp2f(zp)); }

__device__ __forceinline__ void softmax_head_folded(const float u[3], float r[3]) {
    const float c6 = (float)(6.0 * LOG2E), c12 = (float)(12.0 * LOG2E);
    float e[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float q = rcp_fast(1.0f + __builtin_amdgcn_exp2f(u[k]));
        e[k] = __builtin_amdgcn_exp2f(fmaf(-c12, q, c6));
    }
    float inv = rcp_fast(e[0] + e[1] + e[2]);
    r[0] = e[0] * inv; r[1] = e[1] * inv; r[2] = e[2] * inv;
}

// First layer of the reflected pair: the reflected input (x2, x1, x0, x3, x4) shares
// x1, x3 and x4 with the direct one, so each neuron sums the bias and those three
// terms once and adds the two remaining products per input (7 FMAs for the pair
// instead of 10).
template <class ACT>
__device__ __forceinline__ void dense_pair_reflected5(const float x[5], const float *w, float *ha, float *hb, ACT act) {
#pragma unroll
    for (int o = 0; o < 10; ++o) {
        kconst_ptr<float> c = kparams(w + DEV_L1 + o * 6);   // one neuron's run at a time
        const int b = 0;
        float shared = c[b] + x[1] * c[b + 2];
        shared = shared + x[3] * c[b + 4];
        shared = shared + x[4] * c[b + 5];
        float za = shared + x[0] * c[b + 1];
        za = za + x[2] * c[b + 3];
        float zb = shared + x[2] * c[b + 1];
        zb = zb + x[0] * c[b + 3];
        ha[o] = act(za);
        hb[o] = act(zb);
    }
}

// ra = M(x), rb = M(P5 x) with P5 x = (x2, x1, x0, x3, x4).
__device__ __forceinline__ void mlp_pair32_folded(const float x[5], const float *w, float ra[3], float rb[3]) {
    float a1[10], b1[10], a2[6], b2[6], a3[6], b3[6], za[3], zb[3];
    auto sw = [](float z) { return swish_folded(z); };
    auto id = [](float z) { return z; };
    dense_pair_reflected5(x, w, a1, b1, sw);
    dense_pair<10, 6>(a1, b1, w, DEV_L2, a2, b2, sw);
    dense_pair<6, 6>(a2, b2, w, DEV_L3, a3, b3, sw);
    dense_pair<6, 3>(a3, b3, w, DEV_L4, za, zb, id);
    softmax_head_folded(za, ra);
    softmax_head_folded(zb, rb);
}

// The folded FP32 closure of the 2-D sweep with its parameters in LDS (one copy per
// workgroup): every neuron's run [b, w_0 .. w_(K-1)] padded to whole 16-byte groups
// (strides 8 / 12 / 8 / 8 floats), read as ds_read_b128 broadcasts into VGPRs through an
// opaque per-neuron address (so that the reads stay next to their use and are not
// hoisted out of the sweep's loops).  Against scalar loads per neuron (an s_waitcnt each
// and SGPR-spill reloads) the window's SR32 sweep runs 3-4 % faster.
constexpr int LDS_L1 = 0, LDS_L2 = 80, LDS_L3 = 152, LDS_L4 = 200, LDS_MLP32 = 224;
using lds_cfloat = const float __attribute__((address_space(3)));
struct MlpL { lds_cfloat *p; };
// padded LDS slot -> index into the neuron-major block, or -1 for padding
__device__ __forceinline__ int lds_mlp32_source(int t) {
    int base, stride, K, start;
    if (t < LDS_L2) { base = LDS_L1; stride = 8; K = 5; start = DEV_L1; }
    else if (t < LDS_L3) { base = LDS_L2; stride = 12; K = 10; start = DEV_L2; }
    else if (t < LDS_L4) { base = LDS_L3; stride = 8; K = 6; start = DEV_L3; }
    else { base = LDS_L4; stride = 8; K = 6; start = DEV_L4; }
    const int n = (t - base) / stride, sl = (t - base) % stride;
    return sl <= K ? start + n * (K + 1) + sl : -1;
}
template <int STRIDE>
__device__ __forceinline__ void lds_run(lds_cfloat *p, float c[STRIDE]) {
    asm volatile("" : "+s"(p));
#pragma unroll
    for (int v = 0; v < STRIDE / 4; ++v) {
        typedef float v4f __attribute__((ext_vector_type(4)));
        v4f q = *(const v4f __attribute__((address_space(3))) *)(p + 4 * v);
        c[4 * v] = q.x; c[4 * v + 1] = q.y; c[4 * v + 2] = q.z; c[4 * v + 3] = q.w;
    }
}
template <int K, int N, int STRIDE, class ACT>
__device__ __forceinline__ void dense_pair_lds(const float *xa, const float *xb, lds_cfloat *w, float *ha, float *hb, ACT act) {
#pragma unroll
    for (int o = 0; o < N; ++o) {
        float c[STRIDE];
        lds_run<STRIDE>(w + o * STRIDE, c);
        float za = c[0], zb = c[0];
#pragma unroll
        for (int k = 0; k < K; ++k) { za = za + xa[k] * c[1 + k]; zb = zb + xb[k] * c[1 + k]; }
        ha[o] = act(za); hb[o] = act(zb);
    }
}
__device__ __forceinline__ void mlp_pair32_folded(const float x[5], lds_cfloat *w, float ra[3], float rb[3]) {
    float a1[10], b1[10], a2[6], b2[6], a3[6], b3[6], za[3], zb[3];
    auto sw = [](float z) { return swish_folded(z); };
    auto id = [](float z) { return z; };
#pragma unroll
    for (int o = 0; o < 10; ++o) {
        float c[8];
        lds_run<8>(w + LDS_L1 + o * 8, c);
        float shared = c[0] + x[1] * c[2];
        shared = shared + x[3] * c[4];
        shared = shared + x[4] * c[5];
        float za1 = shared + x[0] * c[1];
        za1 = za1 + x[2] * c[3];
        float zb1 = shared + x[2] * c[1];
        zb1 = zb1 + x[0] * c[3];
        a1[o] = sw(za1); b1[o] = sw(zb1);
    }
    dense_pair_lds<10, 6, 12>(a1, b1, w + LDS_L2, a2, b2, sw);
    dense_pair_lds<6, 6, 8>(a2, b2, w + LDS_L3, a3, b3, sw);
    dense_pair_lds<6, 3, 8>(a3, b3, w + LDS_L4, za, zb, id);
    softmax_head_folded(za, ra);
    softmax_head_folded(zb, rb);
}

// ------------------------------------------------ BF16 closure on packed BF16 dot products
// Weights in bfloat16, two per 32-bit word; the layer inputs (the features and the
// hidden activations) rounded to bfloat16 (round to nearest even) and packed in pairs by
// v_cvt_pk_bf16_f32; each neuron is its FP32 bias plus v_dot2c_f32_bf16 steps (two
// exact bf16 x bf16 products accumulated in FP32 per instruction); activations, the
// bounded head and the softmax in FP32.  Device block, per layer and neuron:
// [bias (FP32 bits), (w_0, w_1), (w_2, w_3), ...] with the last pair zero-padded
// for odd K: 4 / 6 / 4 / 4 words per neuron of the 5-10-6-6-3 layers.
constexpr int BF_L1 = 0, BF_L2 = 40, BF_L3 = 76, BF_L4 = 100, BF_WORDS = 112;

// host: reference-layout FP32 parameters -> the packed BF16 block (weights rounded to
// nearest even, biases kept in FP32)
inline void pack_bf16_closure(const float *ref, uint32_t *out) {
    const int K[4] = {5, 10, 6, 6}, N[4] = {10, 6, 6, 3};
    const int W[4] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4}, B[4] = {OFF_B1, OFF_B2, OFF_B3, OFF_B4};
    auto bits = [](float x) { uint32_t u; memcpy(&u, &x, 4); return u; };
    auto bf = [&](float x) -> uint32_t {
        uint32_t u = bits(x);
        if ((u & 0x7f800000u) == 0x7f800000u) return u >> 16;
        return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    };
    int o = 0;
    for (int l = 0; l < 4; ++l)
        for (int n = 0; n < N[l]; ++n) {
            out[o++] = bits(ref[B[l] + n]);
            for (int k = 0; k < K[l]; k += 2) {
                uint32_t lo = bf(ref[W[l] + k * N[l] + n]);
                uint32_t hi = k + 1 < K[l] ? bf(ref[W[l] + (k + 1) * N[l] + n]) : 0u;
                out[o++] = lo | (hi << 16);
            }
        }
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bf16x2 pk_bf16(float lo, float hi) { return bf16x2{(__bf16)lo, (__bf16)hi}; }
__device__ __forceinline__ float dot2_bf16(uint32_t w, bf16x2 x, float acc) {
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, w), x, acc, false);
}

// NP packed input pairs -> N neurons (pre-activation) of one evaluation pair; one
// scalar base per layer, the neurons' runs at constant offsets from it
template <int NP, int N, int STRIDE>
__device__ __forceinline__ void dense_bf16_pair(const bf16x2 *xa, const bf16x2 *xb, const float *wf, int base, float *za,
                                                float *zb) {
    kconst_ptr<uint32_t> c = kparams((const uint32_t *)wf + base);
#pragma unroll
    for (int o = 0; o < N; ++o) {
        float a = __builtin_bit_cast(float, c[o * STRIDE]), b = a;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            a = dot2_bf16(c[o * STRIDE + 1 + k], xa[k], a);
            b = dot2_bf16(c[o * STRIDE + 1 + k], xb[k], b);
        }
        za[o] = a;
        zb[o] = b;
    }
}

// ra = M(xa), rb = M(xb) for the BF16 closure (inputs rounded here).  The FP32 inputs
// pass through an empty asm so that each is rounded from its FP32 value, as the
// closure defines it (without it the compiler would fold an FP64 -> FP32 -> BF16
// chain into one direct rounding).  The head is softmax(6 tanh(z / 6)) in the form
// e_k = 2^(6L - 12L / (1 + 2^(z_k L / 3))), L = log2 e, without a max shift (the
// bounded logits cannot overflow).
__device__ __forceinline__ void mlp_pair_bf16(const float xa_in[5], const float xb_in[5], const float *w, float ra[3],
                                              float rb[3]) {
    float xa[5], xb[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        xa[k] = xa_in[k];
        xb[k] = xb_in[k];
        asm volatile("" : "+v"(xa[k]));
        asm volatile("" : "+v"(xb[k]));
    }
    bf16x2 pa[5], pb[5];
    pa[0] = pk_bf16(xa[0], xa[1]); pa[1] = pk_bf16(xa[2], xa[3]); pa[2] = pk_bf16(xa[4], 0.0f);
    pb[0] = pk_bf16(xb[0], xb[1]); pb[1] = pk_bf16(xb[2], xb[3]); pb[2] = pk_bf16(xb[4], 0.0f);
    float za[10], zb[10];
    dense_bf16_pair<3, 10, 4>(pa, pb, w, BF_L1, za, zb);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        pa[k] = pk_bf16(swish32(za[2 * k]), swish32(za[2 * k + 1]));
        pb[k] = pk_bf16(swish32(zb[2 * k]), swish32(zb[2 * k + 1]));
    }
    dense_bf16_pair<5, 6, 6>(pa, pb, w, BF_L2, za, zb);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        pa[k] = pk_bf16(swish32(za[2 * k]), swish32(za[2 * k + 1]));
        pb[k] = pk_bf16(swish32(zb[2 * k]), swish32(zb[2 * k + 1]));
    }
    dense_bf16_pair<3, 6, 4>(pa, pb, w, BF_L3, za, zb);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        pa[k] = pk_bf16(swish32(za[2 * k]), swish32(za[2 * k + 1]));
        pb[k] = pk_bf16(swish32(zb[2 * k]), swish32(zb[2 * k + 1]));
    }
    dense_bf16_pair<3, 3, 4>(pa, pb, w, BF_L4, za, zb);
    const float l3 = (float)(LOG2E / 3.0);
    float ua[3] = {za[0] * l3, za[1] * l3, za[2] * l3}, ub[3] = {zb[0] * l3, zb[1] * l3, zb[2] * l3};
    softmax_head_folded(ua, ra);
    softmax_head_folded(ub, rb);
}

// SR features for the FP32 closure: the plateau rule in FP64 as sr_features (the branch
// is exact), then the five features with FP32 reciprocals and one v_log_f32 from the
// FP64 differences (the reference computes them in FP64 and casts them to FP32;
// the difference is an FP32 rounding of each feature).
__device__ __forceinline__ bool sr_features32(const double q[5], float f[5]) {
    double d20 = q[0] - 2.0 * q[1] + q[2];
    double d21 = q[1] - 2.0 * q[2] + q[3];
    double d22 = q[2] - 2.0 * q[3] + q[4];
    const double c1312 = 13.0 / 12.0;
    double delta0 = c1312 * fabs(d20) + 0.25 * fabs(q[0] - 4.0 * q[1] + 3.0 * q[2]);
    double delta1 = c1312 * fabs(d21) + 0.25 * fabs(q[1] - q[3]);
    double delta2 = c1312 * fabs(d22) + 0.25 * fabs(3.0 * q[2] - 4.0 * q[3] + q[4]);
    double delta_max = dmax(dmax(delta0, delta1), delta2);
    double q_scale = dmax(dmax(dmax(fabs(q[0]), fabs(q[1])), dmax(fabs(q[2]), fabs(q[3]))), fabs(q[4]));
    q_scale = dmax(q_scale, 1.0);
    if (delta_max <= 1.0e-13 * q_scale) return true;
    // not a plateau: delta_max > 1e-13 q_scale >= 1e-13, above the reference's 1e-15 floors
    const double eps = 1.0e-15;
    float g0 = (float)fabs(d20) * rcp_fast((float)(fabs(q[1] - q[0]) + fabs(q[2] - q[1]) + eps));
    float g1 = (float)fabs(d21) * rcp_fast((float)(fabs(q[2] - q[1]) + fabs(q[3] - q[2]) + eps));
    float g2 = (float)fabs(d22) * rcp_fast((float)(fabs(q[3] - q[2]) + fabs(q[4] - q[3]) + eps));
    float inv_dm = rcp_fast((float)delta_max);
    float rel = (float)delta_max * rcp_fast((float)q_scale);
    // (log10(rel) + 4) / 4 = log2(rel) log10(2) / 4 + 1, clipped to [0, 1]
    float scale = fmaf(__builtin_amdgcn_logf(rel), (float)(0.30102999566398120 / 4.0), 1.0f);
    f[0] = (float)delta0 * inv_dm;
    f[1] = (float)delta1 * inv_dm;
    f[2] = (float)delta2 * inv_dm;
    f[3] = fminf(1.0f, fmax_(fmax_(g0, g1), g2));
    f[4] = fminf(1.0f, fmax_(0.0f, scale));
    return false;
}

// 1 / x in FP64 from v_rcp_f64 and two Newton steps (the FP32 closure's weights)
__device__ __forceinline__ double rcp_newton(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// FP32 closure: reflection-averaged ratios from the FP32 features and folded parameters
// A wave raises its issue priority while it evaluates the closure: the waves with
// closure work are the ones the sweep's workgroup barriers wait for (2-D SR32 sweep 3 %
// faster, 3-D passes unchanged).
template <class MLP>
__device__ __forceinline__ void sr_ratios32(const float f[5], const MLP &P, double br[3]) {
    float rd[3], rr[3];
    __builtin_amdgcn_s_setprio(3);
    mlp_pair32_folded(f, P.p, rd, rr);
    __builtin_amdgcn_s_setprio(0);
    br[0] = 0.5 * ((double)rd[0] + (double)rr[2]);
    br[1] = 0.5 * ((double)rd[1] + (double)rr[1]);
    br[2] = 0.5 * ((double)rd[2] + (double)rr[0]);
}

// Reflection-averaged ratios 0.5 [M(x) + P3 M(P5 x)] in FP64 (C64:225-237, C32:225-237).
template <int SCHEME, class MLP>
__device__ __forceinline__ void sr_ratios(const double f[5], const MLP &P, double br[3]) {
    double direct[3], refl[3];
    static_assert(SCHEME == SCHEME_SR64 || SCHEME == SCHEME_SRBF16, "the FP32 closure runs sr_ratios32");
    if constexpr (SCHEME == SCHEME_SR64) {
        double xr[5] = {f[2], f[1], f[0], f[3], f[4]};
        mlp_pair64(f, xr, P.p, direct, refl);
    } else {
        float xd[5] = {(float)f[0], (float)f[1], (float)f[2], (float)f[3], (float)f[4]};
        float xr[5] = {xd[2], xd[1], xd[0], xd[3], xd[4]};
        float rd[3], rr[3];
        mlp_pair_bf16(xd, xr, P.p, rd, rr);
        direct[0] = rd[0]; direct[1] = rd[1]; direct[2] = rd[2];
        refl[0] = rr[0]; refl[1] = rr[1]; refl[2] = rr[2];
    }
    br[0] = 0.5 * (direct[0] + refl[2]);
    br[1] = 0.5 * (direct[1] + refl[1]);
    br[2] = 0.5 * (direct[2] + refl[0]);
}

// C64:239-279: beta = 3 r, alpha = d / (beta + 1e-12)^2, normalised (+ optional ENO cutoff).
// FAST (the FP32 closure) takes the reciprocals from rcp_newton instead of an IEEE division.
template <bool FAST = false>
__device__ __forceinline__ void sr_inv_sq(const double br[3], double ib2[3]) {
    for (int k = 0; k < 3; ++k) {
        double x = 3.0 * br[k] + 1.0e-12;
        double ib = FAST ? rcp_newton(x) : 1.0 / x;
        ib2[k] = ib * ib;
    }
}

template <bool FAST = false>
__device__ __forceinline__ double sr_combine(const double s[3], const double d[3], const double ib2[3],
                                             bool plateau, int eno_cutoff) {
    double w0, w1, w2;
    if (plateau) {
        w0 = d[0]; w1 = d[1]; w2 = d[2];
    } else {
        double a0 = d[0] * ib2[0], a1 = d[1] * ib2[1], a2 = d[2] * ib2[2];
        double asum = dmax(a0 + a1 + a2, 1.0e-300);
        double inv = FAST ? rcp_newton(asum) : 1.0 / asum;
        w0 = a0 * inv; w1 = a1 * inv; w2 = a2 * inv;
        if (eno_cutoff) {
            const double cutoff = 4.0e-7;
            double p0 = w0 <= cutoff ? 0.0 : 1.0, p1 = w1 <= cutoff ? 0.0 : 1.0, p2 = w2 <= cutoff ? 0.0 : 1.0;
            double ic = 1.0 / dmax(p0 * w0 + p1 * w1 + p2 * w2, 1.0e-300);
            w0 = p0 * w0 * ic; w1 = p1 * w1 * ic; w2 = p2 * w2 * ic;
        }
    }
    return w0 * s[0] + w1 * s[1] + w2 * s[2];
}

// ------------------------------------------------ scalar reconstructions
// Normal face value (lr1: right face x_{i+1/2}; otherwise left face x_{i-1/2}).
template <int SCHEME, class MLP>
__device__ __forceinline__ double recon_normal(const double q[5], bool lr1, const MLP &P, int eno, double h,
                                               bool &closure_evaluated) {
    double s[3], d[3];
    closure_evaluated = false;
    candidates_normal(q, lr1, s);
    d_normal(lr1, d);
    if constexpr (SCHEME == SCHEME_JS) {
        double b[3], inv[3];
        js_beta(q, b);
        for (int k = 0; k < 3; ++k) inv[k] = 1.0 / (1.0e-6 + b[k]);
        return js_combine(s, d, inv);
    } else if constexpr (SCHEME == SCHEME_Z) {
        double b[3], r2[3];
        js_beta(q, b);
        double tau5 = fabs(b[0] - b[2]), eps_z = h * h;
        for (int k = 0; k < 3; ++k) { double r = tau5 / (b[k] + eps_z); r2[k] = r * r; }
        return z_combine(s, d, r2);
    } else if constexpr (SCHEME == SCHEME_SR32) {
        float f[5];
        double ib2[3] = {0.0, 0.0, 0.0};
        bool plateau = sr_features32(q, f);
        closure_evaluated = !plateau;
        if (!plateau) {
            double br[3];
            sr_ratios32(f, P, br);
            sr_inv_sq<true>(br, ib2);
        }
        return sr_combine<true>(s, d, ib2, plateau, eno);
    } else {
        double f[5], ib2[3] = {0.0, 0.0, 0.0};
        bool plateau = sr_features(q, f);
        closure_evaluated = !plateau;
        if (!plateau) {
            double br[3];
            sr_ratios<SCHEME>(f, P, br);
            sr_inv_sq(br, ib2);
        }
        return sr_combine(s, d, ib2, plateau, eno);
    }
}

// Both transverse Gauss-point values of one stencil; the closure (beta or MLP
// ratios) is evaluated once and shared by the two points.
template <int SCHEME, class MLP>
__device__ __forceinline__ void recon_gauss(const double q[5], const MLP &P, int eno, double h,
                                            double &g1, double &g2, bool &closure_evaluated) {
    double s1[3], s2[3], d1[3], d2[3];
    closure_evaluated = false;
    candidates_gauss(q, s1, s2);
    d_gauss(d1, d2);
    if constexpr (SCHEME == SCHEME_JS) {
        double b[3], inv[3];
        js_beta(q, b);
        for (int k = 0; k < 3; ++k) inv[k] = 1.0 / (1.0e-6 + b[k]);
        g1 = js_combine(s1, d1, inv);
        g2 = js_combine(s2, d2, inv);
    } else if constexpr (SCHEME == SCHEME_Z) {
        double b[3], r2[3];
        js_beta(q, b);
        double tau5 = fabs(b[0] - b[2]), eps_z = h * h;
        for (int k = 0; k < 3; ++k) { double r = tau5 / (b[k] + eps_z); r2[k] = r * r; }
        g1 = z_combine(s1, d1, r2);
        g2 = z_combine(s2, d2, r2);
    } else if constexpr (SCHEME == SCHEME_SR32) {
        float f[5];
        double ib2[3] = {0.0, 0.0, 0.0};
        bool plateau = sr_features32(q, f);
        closure_evaluated = !plateau;
        if (!plateau) {
            double br[3];
            sr_ratios32(f, P, br);
            sr_inv_sq<true>(br, ib2);
        }
        g1 = sr_combine<true>(s1, d1, ib2, plateau, eno);
        g2 = sr_combine<true>(s2, d2, ib2, plateau, eno);
    } else {
        double f[5], ib2[3] = {0.0, 0.0, 0.0};
        bool plateau = sr_features(q, f);
        closure_evaluated = !plateau;
        if (!plateau) {
            double br[3];
            sr_ratios<SCHEME>(f, P, br);
            sr_inv_sq(br, ib2);
        }
        g1 = sr_combine(s1, d1, ib2, plateau, eno);
        g2 = sr_combine(s2, d2, ib2, plateau, eno);
    }
}

}  // namespace wsr
