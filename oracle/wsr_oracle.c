/*
 * wsr_oracle.c -- CPU oracle for the two-dimensional WENO5 finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C, FP64 restatement of the
 * reference's WENO5-JS / WENO5-SR (FP64, FP32 and emulated-BF16 MLP closure) /
 * WENO5-Z characteristic reconstruction, transverse Gauss reconstruction,
 * HLLC / FORCE flux, SSPRK3 stage update, boundary kernels and CFL rule.  It is
 * used by the test-suite as the parity checker, by the smoke entry point and by
 * the benchmark's CPU-baseline leg.  The product path (the HIP library) never
 * links or calls it.
 *
 * It deliberately follows the reference's structure literally (one "thread"
 * per cell, Roe average and Jacobian values recomputed inside every projection,
 * reflected features recomputed from the reversed stencil, the MLP evaluated
 * separately for every head) and keeps every expression in the reference's
 * evaluation order.  Build with -ffp-contract=off so that no multiply-add is
 * fused (see Makefile).
 *
 * Reference files (paths relative to the reference checkout):
 *   H5  = code/snapshot/warp_weno5_helpers.py
 *   C64 = code/snapshot/teacherfree_lab_weno5_v20_distance_balanced/weno5_rk3_diff_v20_deploy.py
 *   C32 = code/snapshot/teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/warp_mlp_f32/weno5_rk3_diff_mlp_f32.py
 *   CB  = code/snapshot/weno5_rk3_diff.py
 *   AJS = code/snapshot/for_paper_results/solvers/weno5_hllc.py
 *   A64 = code/snapshot/teacherfree_lab_weno5_v20_distance_balanced/weno5_hllc_refsym.py
 *   A32 = code/snapshot/for_paper_results/solvers/weno5_hllc_mixed.py
 *   WZ  = code/snapshot/weno_z_borges_p2_results/weno5_helpers_z.py
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* enums shared with the Python wrapper                                      */
/* ------------------------------------------------------------------------ */
enum { SCHEME_JS = 0, SCHEME_SR64 = 1, SCHEME_SR32 = 2, SCHEME_SRBF16 = 3, SCHEME_Z = 4 };
enum { RIEMANN_HLLC = 0, RIEMANN_FORCE = 1 };
enum { BC_TRANSMISSIVE = 0, BC_PERIODIC = 1, BC_DOUBLE_MACH = 2 };
enum { HEAD_NORMAL_LR1 = 0, HEAD_NORMAL_LR2 = 1, HEAD_GAUSS_LR1 = 2, HEAD_GAUSS_LR2 = 3 };

#define MLP5_COUNT 189

typedef struct { double v[4]; } vec4;

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline float fmaxf_(float a, float b) { return a > b ? a : b; }

static inline vec4 v4(double a, double b, double c, double d) {
    vec4 r; r.v[0] = a; r.v[1] = b; r.v[2] = c; r.v[3] = d; return r;
}

/* ------------------------------------------------------------------------ */
/* Euler algebra (H5:245-285)                                                */
/* ------------------------------------------------------------------------ */
/* H5:246-252 con_to_pri */
static vec4 con_to_pri(vec4 q, double gamma) {
    double rho = q.v[0];
    double u = q.v[1] / rho;
    double v = q.v[2] / rho;
    double p = (gamma - 1.0) * (q.v[3] - 0.5 * rho * (u * u + v * v));
    return v4(rho, u, v, p);
}

/* H5:255-261 pri_to_con */
static vec4 pri_to_con(vec4 w, double gamma) {
    double rho = w.v[0], u = w.v[1], v = w.v[2], p = w.v[3];
    double e = 0.5 * rho * (u * u + v * v) + p / (gamma - 1.0);
    return v4(rho, rho * u, rho * v, e);
}

/* H5:264-285 pri_to_flux */
static vec4 pri_to_flux(vec4 w, int direction, double gamma) {
    double rho = w.v[0], u = w.v[1], v = w.v[2], p = w.v[3];
    double e = 0.5 * rho * (u * u + v * v) + p / (gamma - 1.0);
    if (direction == 1)
        return v4(rho * u, rho * u * u + p, rho * u * v, u * (p + e));
    return v4(rho * v, rho * u * v, rho * v * v + p, v * (p + e));
}

/* H5:401-435 force_flux: FORCE = 1/2 (Lax-Friedrichs + Richtmyer), c = dt/h */
static vec4 force_flux(vec4 ul0, vec4 ur0, int direction, double c, double gamma) {
    vec4 wl0 = con_to_pri(ul0, gamma);
    vec4 wr0 = con_to_pri(ur0, gamma);
    vec4 fl = pri_to_flux(wl0, direction, gamma);
    vec4 fr = pri_to_flux(wr0, direction, gamma);
    vec4 ri_u;
    for (int k = 0; k < 4; ++k)
        ri_u.v[k] = 0.5 * (ul0.v[k] + ur0.v[k]) - 0.5 * c * (fr.v[k] - fl.v[k]);
    vec4 ri = pri_to_flux(con_to_pri(ri_u, gamma), direction, gamma);
    double dx_dt = 1.0 / c;
    vec4 lf, out;
    for (int k = 0; k < 4; ++k)
        lf.v[k] = 0.5 * (fl.v[k] + fr.v[k]) - 0.5 * dx_dt * (ur0.v[k] - ul0.v[k]);
    for (int k = 0; k < 4; ++k) out.v[k] = 0.5 * (lf.v[k] + ri.v[k]);
    return out;
}

/* AJS:23-115 hllc_flux: Davis wave-speed bounds, sign-preserving floors */
static double floor_signed(double d, double tiny) {
    if (fabs(d) < tiny) return d < 0.0 ? -tiny : tiny;
    return d;
}

static vec4 hllc_flux(vec4 ul0, vec4 ur0, int direction, double gamma) {
    const double tiny = 1.0e-16;
    vec4 wl = con_to_pri(ul0, gamma);
    vec4 wr = con_to_pri(ur0, gamma);
    double rho_l = dmax(wl.v[0], tiny);
    double rho_r = dmax(wr.v[0], tiny);
    double p_l = dmax(wl.v[3], tiny);
    double p_r = dmax(wr.v[3], tiny);
    double a_l = sqrt(gamma * p_l / rho_l);
    double a_r = sqrt(gamma * p_r / rho_r);
    double un_l = wl.v[1], un_r = wr.v[1];
    if (direction == 2) { un_l = wl.v[2]; un_r = wr.v[2]; }
    double s_l = dmin(un_l - a_l, un_r - a_r);
    double s_r = dmax(un_l + a_l, un_r + a_r);
    vec4 f_l = pri_to_flux(wl, direction, gamma);
    vec4 f_r = pri_to_flux(wr, direction, gamma);
    if (0.0 <= s_l) return f_l;
    if (0.0 >= s_r) return f_r;
    double denom = floor_signed(rho_l * (s_l - un_l) - rho_r * (s_r - un_r), tiny);
    double s_star = (p_r - p_l + rho_l * un_l * (s_l - un_l) - rho_r * un_r * (s_r - un_r)) / denom;
    vec4 u_side = ul0, w_side = wl, f_side = f_l;
    double s_side = s_l;
    if (s_star < 0.0) { u_side = ur0; w_side = wr; f_side = f_r; s_side = s_r; }
    double rho = dmax(w_side.v[0], tiny);
    double u = w_side.v[1], v = w_side.v[2];
    double p = dmax(w_side.v[3], tiny);
    double un = direction == 2 ? v : u;
    double denom_star = floor_signed(s_side - s_star, tiny);
    double factor = rho * (s_side - un) / denom_star;
    double rho_star = factor;
    double mx_star = factor * s_star;
    double my_star = factor * v;
    if (direction == 2) { mx_star = factor * u; my_star = factor * s_star; }
    double denom_side = floor_signed(s_side - un, tiny);
    double e_star = factor * ((u_side.v[3] / rho) + (s_star - un) * (s_star + p / (rho * denom_side)));
    vec4 u_star = v4(rho_star, mx_star, my_star, e_star);
    vec4 out;
    for (int k = 0; k < 4; ++k) out.v[k] = f_side.v[k] + s_side * (u_star.v[k] - u_side.v[k]);
    return out;
}

/* ------------------------------------------------------------------------ */
/* Roe average and characteristic projections (H5:598-748)                   */
/* ------------------------------------------------------------------------ */
/* H5:598-628 roe_average_state */
static vec4 roe_average_state(vec4 left, vec4 right, double gamma) {
    const double tiny = 1.0e-15;
    double gm1 = gamma - 1.0;
    double rho_l = dmax(left.v[0], tiny);
    double u_l = left.v[1] / rho_l;
    double v_l = left.v[2] / rho_l;
    double e_l = left.v[3];
    double p_l = dmax(gm1 * (e_l - 0.5 * rho_l * (u_l * u_l + v_l * v_l)), tiny);
    double h_l = (e_l + p_l) / rho_l;
    double rho_r = dmax(right.v[0], tiny);
    double u_r = right.v[1] / rho_r;
    double v_r = right.v[2] / rho_r;
    double e_r = right.v[3];
    double p_r = dmax(gm1 * (e_r - 0.5 * rho_r * (u_r * u_r + v_r * v_r)), tiny);
    double h_r = (e_r + p_r) / rho_r;
    double sqrt_l = sqrt(rho_l);
    double sqrt_r = sqrt(rho_r);
    double inv = 1.0 / (sqrt_l + sqrt_r);
    double u_roe = (sqrt_l * u_l + sqrt_r * u_r) * inv;
    double v_roe = (sqrt_l * v_l + sqrt_r * v_r) * inv;
    double h_roe = (sqrt_l * h_l + sqrt_r * h_r) * inv;
    double rho_roe = sqrt_l * sqrt_r;
    double q2_roe = u_roe * u_roe + v_roe * v_roe;
    double e_roe = (rho_roe * h_roe + 0.5 * gm1 * rho_roe * q2_roe) / gamma;
    return v4(rho_roe, rho_roe * u_roe, rho_roe * v_roe, e_roe);
}

/* H5:631-642 jacobian_state_values -> (u, v, a, h) */
static vec4 jacobian_state_values(vec4 roe, double gamma) {
    const double tiny = 1.0e-15;
    double rho = dmax(roe.v[0], tiny);
    double u = roe.v[1] / rho;
    double v = roe.v[2] / rho;
    double gm1 = gamma - 1.0;
    double q2 = u * u + v * v;
    double p = dmax(gm1 * (roe.v[3] - 0.5 * rho * q2), tiny);
    double a = sqrt(gamma * p / rho);
    double h = 0.5 * q2 + a * a / gm1;
    return v4(u, v, a, h);
}

/* H5:645-679 con_to_char_x */
static vec4 con_to_char_x(vec4 q, vec4 roe, double gamma) {
    vec4 val = jacobian_state_values(roe, gamma);
    double u = val.v[0], v = val.v[1], a = val.v[2], h = val.v[3];
    double gm1 = gamma - 1.0;
    double a2 = a * a;
    double scale = gm1 / (2.0 * a2);
    double c0 = ((h + a * (u - a) / gm1) * q.v[0] - (u + a / gm1) * q.v[1] - v * q.v[2] + q.v[3]) * scale;
    double c1 = ((-2.0 * h + 4.0 * a2 / gm1) * q.v[0] + 2.0 * u * q.v[1] + 2.0 * v * q.v[2] - 2.0 * q.v[3]) * scale;
    double c2 = (-2.0 * v * a2 / gm1 * q.v[0] + 2.0 * a2 / gm1 * q.v[2]) * scale;
    double c3 = ((h - a * (u + a) / gm1) * q.v[0] + (-u + a / gm1) * q.v[1] - v * q.v[2] + q.v[3]) * scale;
    return v4(c0, c1, c2, c3);
}

/* H5:682-696 char_to_con_x */
static vec4 char_to_con_x(vec4 c, vec4 roe, double gamma) {
    vec4 val = jacobian_state_values(roe, gamma);
    double u = val.v[0], v = val.v[1], a = val.v[2], h = val.v[3];
    double q2 = u * u + v * v;
    return v4(c.v[0] + c.v[1] + c.v[3],
              (u - a) * c.v[0] + u * c.v[1] + (u + a) * c.v[3],
              v * c.v[0] + v * c.v[1] + c.v[2] + v * c.v[3],
              (h - u * a) * c.v[0] + 0.5 * q2 * c.v[1] + v * c.v[2] + (h + u * a) * c.v[3]);
}

/* H5:699-731 con_to_char_y */
static vec4 con_to_char_y(vec4 q, vec4 roe, double gamma) {
    vec4 val = jacobian_state_values(roe, gamma);
    double tangent = val.v[0], normal = val.v[1], a = val.v[2], h = val.v[3];
    double gm1 = gamma - 1.0;
    double a2 = a * a;
    double scale = gm1 / (2.0 * a2);
    double c0 = ((h + a * (normal - a) / gm1) * q.v[0] - tangent * q.v[1] - (normal + a / gm1) * q.v[2] + q.v[3]) * scale;
    double c1 = ((-2.0 * h + 4.0 * a2 / gm1) * q.v[0] + 2.0 * tangent * q.v[1] + 2.0 * normal * q.v[2] - 2.0 * q.v[3]) * scale;
    double c2 = (-2.0 * tangent * a2 / gm1 * q.v[0] + 2.0 * a2 / gm1 * q.v[1]) * scale;
    double c3 = ((h - a * (normal + a) / gm1) * q.v[0] - tangent * q.v[1] + (-normal + a / gm1) * q.v[2] + q.v[3]) * scale;
    return v4(c0, c1, c2, c3);
}

/* H5:734-748 char_to_con_y */
static vec4 char_to_con_y(vec4 c, vec4 roe, double gamma) {
    vec4 val = jacobian_state_values(roe, gamma);
    double tangent = val.v[0], normal = val.v[1], a = val.v[2], h = val.v[3];
    double q2 = tangent * tangent + normal * normal;
    return v4(c.v[0] + c.v[1] + c.v[3],
              tangent * c.v[0] + tangent * c.v[1] + c.v[2] + tangent * c.v[3],
              (normal - a) * c.v[0] + normal * c.v[1] + (normal + a) * c.v[3],
              (h - normal * a) * c.v[0] + 0.5 * q2 * c.v[1] + tangent * c.v[2] + (h + normal * a) * c.v[3]);
}

static vec4 con_to_char(vec4 q, vec4 roe, int direction, double gamma) {
    return direction == 1 ? con_to_char_x(q, roe, gamma) : con_to_char_y(q, roe, gamma);
}
static vec4 char_to_con(vec4 c, vec4 roe, int direction, double gamma) {
    return direction == 1 ? char_to_con_x(c, roe, gamma) : char_to_con_y(c, roe, gamma);
}

/* ------------------------------------------------------------------------ */
/* Candidate stencils (H5:438-446 face, H5:798-810 Gauss with the reference's */
/* 15-digit constants)                                                        */
/* ------------------------------------------------------------------------ */
static double stencil_2d(double u1, double u2, double u3, int location) {
    if (location == 1) return (1.0 / 3.0) * u1 - (7.0 / 6.0) * u2 + (11.0 / 6.0) * u3;
    if (location == 2) return (-1.0 / 6.0) * u1 + (5.0 / 6.0) * u2 + (1.0 / 3.0) * u3;
    if (location == 3) return (1.0 / 3.0) * u1 + (5.0 / 6.0) * u2 - (1.0 / 6.0) * u3;
    return (11.0 / 6.0) * u1 - (7.0 / 6.0) * u2 + (1.0 / 3.0) * u3;
}

static double stencil_gauss(double u1, double u2, double u3, int location) {
    if (location == 1) return -0.144337567297407 * u1 + 0.577350269189626 * u2 + 0.566987298107780 * u3;
    if (location == 2) return 0.144337567297406 * u1 + u2 - 0.144337567297407 * u3;
    if (location == 3) return 1.433012701892220 * u1 - 0.577350269189626 * u2 + 0.144337567297406 * u3;
    if (location == 4) return 0.144337567297406 * u1 - 0.577350269189626 * u2 + 1.433012701892220 * u3;
    if (location == 5) return -0.144337567297407 * u1 + u2 + 0.144337567297406 * u3;
    return 0.566987298107781 * u1 + 0.577350269189625 * u2 - 0.144337567297406 * u3;
}

/* candidate values and linear weights of one head */
static void candidates(const double q[5], int head, double s[3]) {
    switch (head) {
    case HEAD_NORMAL_LR1:
        s[0] = stencil_2d(q[0], q[1], q[2], 1); s[1] = stencil_2d(q[1], q[2], q[3], 2);
        s[2] = stencil_2d(q[2], q[3], q[4], 3); break;
    case HEAD_NORMAL_LR2:
        s[0] = stencil_2d(q[0], q[1], q[2], 2); s[1] = stencil_2d(q[1], q[2], q[3], 3);
        s[2] = stencil_2d(q[2], q[3], q[4], 4); break;
    case HEAD_GAUSS_LR1:
        s[0] = stencil_gauss(q[0], q[1], q[2], 1); s[1] = stencil_gauss(q[1], q[2], q[3], 2);
        s[2] = stencil_gauss(q[2], q[3], q[4], 3); break;
    default:
        s[0] = stencil_gauss(q[0], q[1], q[2], 4); s[1] = stencil_gauss(q[1], q[2], q[3], 5);
        s[2] = stencil_gauss(q[2], q[3], q[4], 6); break;
    }
}

/* C64:136-152 optimal_weights; H5:527-537 and H5:829-841 use the same d */
static void optimal_weights(int head, double c[3]) {
    double root3 = sqrt(3.0);
    switch (head) {
    case HEAD_NORMAL_LR1: c[0] = 0.1; c[1] = 0.6; c[2] = 0.3; break;
    case HEAD_NORMAL_LR2: c[0] = 0.3; c[1] = 0.6; c[2] = 0.1; break;
    case HEAD_GAUSS_LR1:
        c[0] = (210.0 + root3) / 1080.0; c[1] = 11.0 / 18.0; c[2] = (210.0 - root3) / 1080.0; break;
    default:
        c[0] = (210.0 - root3) / 1080.0; c[1] = 11.0 / 18.0; c[2] = (210.0 + root3) / 1080.0; break;
    }
}

/* ------------------------------------------------------------------------ */
/* WENO5-JS scalar reconstruction (H5:511-570 face, H5:813-873 Gauss)        */
/* ------------------------------------------------------------------------ */
static void js_beta(const double q[5], double beta[3]) {
    double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4];
    beta[2] = (13.0 / 12.0) * (q2 - 2.0 * q3 + q4) * (q2 - 2.0 * q3 + q4)
            + 0.25 * (3.0 * q2 - 4.0 * q3 + q4) * (3.0 * q2 - 4.0 * q3 + q4);
    beta[1] = (13.0 / 12.0) * (q1 - 2.0 * q2 + q3) * (q1 - 2.0 * q2 + q3)
            + 0.25 * (q1 - q3) * (q1 - q3);
    beta[0] = (13.0 / 12.0) * (q0 - 2.0 * q1 + q2) * (q0 - 2.0 * q1 + q2)
            + 0.25 * (q0 - 4.0 * q1 + 3.0 * q2) * (q0 - 4.0 * q1 + 3.0 * q2);
}

/* H5:241-242 safe_rcp */
static double safe_rcp(double x) { return 1.0 / (1.0e-6 + x); }

static double weno5_js_value(const double q[5], int head) {
    double s[3], d[3], beta[3];
    candidates(q, head, s);
    optimal_weights(head, d);
    js_beta(q, beta);
    double inv0 = safe_rcp(beta[0]), inv1 = safe_rcp(beta[1]), inv2 = safe_rcp(beta[2]);
    double alpha0 = d[0] * inv0 * inv0;
    double alpha1 = d[1] * inv1 * inv1;
    double alpha2 = d[2] * inv2 * inv2;
    double alpha_sum = alpha0 + alpha1 + alpha2;
    double w0 = alpha0 / alpha_sum, w1 = alpha1 / alpha_sum, w2 = alpha2 / alpha_sum;
    return w0 * s[0] + w1 * s[1] + w2 * s[2];
}

/* WZ:511-570 (face) and WZ:818-880 (Gauss): WENO5-Z, tau5 = |b0-b2|, eps = h^2, p = 2 */
static double weno5_z_value(const double q[5], int head, double h) {
    double s[3], d[3], beta[3];
    candidates(q, head, s);
    optimal_weights(head, d);
    js_beta(q, beta);
    double tau5 = fabs(beta[0] - beta[2]);
    double eps_z = h * h;
    double ratio0 = tau5 / (beta[0] + eps_z);
    double ratio1 = tau5 / (beta[1] + eps_z);
    double ratio2 = tau5 / (beta[2] + eps_z);
    double alpha0 = d[0] * (1.0 + ratio0 * ratio0);
    double alpha1 = d[1] * (1.0 + ratio1 * ratio1);
    double alpha2 = d[2] * (1.0 + ratio2 * ratio2);
    double alpha_sum = alpha0 + alpha1 + alpha2;
    double w0 = alpha0 / alpha_sum, w1 = alpha1 / alpha_sum, w2 = alpha2 / alpha_sum;
    return w0 * s[0] + w1 * s[1] + w2 * s[2];
}

/* ------------------------------------------------------------------------ */
/* WENO5-SR closure (C64:77-281, C32:74-283)                                 */
/* ------------------------------------------------------------------------ */
/* C64:77-94 weno5_raw_sensors -> (delta0, delta1, delta2, gamma_s) */
static void raw_sensors(double q0, double q1, double q2, double q3, double q4, double out[4]) {
    double d20 = q0 - 2.0 * q1 + q2;
    double d21 = q1 - 2.0 * q2 + q3;
    double d22 = q2 - 2.0 * q3 + q4;
    double delta0 = (13.0 / 12.0) * fabs(d20) + 0.25 * fabs(q0 - 4.0 * q1 + 3.0 * q2);
    double delta1 = (13.0 / 12.0) * fabs(d21) + 0.25 * fabs(q1 - q3);
    double delta2 = (13.0 / 12.0) * fabs(d22) + 0.25 * fabs(3.0 * q2 - 4.0 * q3 + q4);
    double eps = 1.0e-15;
    double gamma0 = fabs(d20) / (fabs(q1 - q0) + fabs(q2 - q1) + eps);
    double gamma1 = fabs(d21) / (fabs(q2 - q1) + fabs(q3 - q2) + eps);
    double gamma2 = fabs(d22) / (fabs(q3 - q2) + fabs(q4 - q3) + eps);
    double gamma_s = dmin(1.0, dmax(dmax(gamma0, gamma1), gamma2));
    out[0] = delta0; out[1] = delta1; out[2] = delta2; out[3] = gamma_s;
}

static double stencil_scale(double q0, double q1, double q2, double q3, double q4) {
    double q_scale = dmax(dmax(dmax(fabs(q0), fabs(q1)), dmax(fabs(q2), fabs(q3))), fabs(q4));
    return dmax(q_scale, 1.0);
}

/* C64:97-122 weno5_nn_features (scale feature clip((log10(rel)+4)/4, 0, 1)) */
static void nn_features(double q0, double q1, double q2, double q3, double q4, double f[5]) {
    double raw[4];
    raw_sensors(q0, q1, q2, q3, q4, raw);
    double eps = 1.0e-15;
    double delta_max = dmax(dmax(raw[0], raw[1]), raw[2]);
    double inv_delta_max = 1.0 / dmax(delta_max, eps);
    double q_scale = stencil_scale(q0, q1, q2, q3, q4);
    double relative_scale = dmax(delta_max / q_scale, 1.0e-30);
    double log10_relative_scale = log(relative_scale) / log(10.0);
    double scale_feature = (log10_relative_scale + 4.0) / 4.0;
    scale_feature = dmin(1.0, dmax(0.0, scale_feature));
    f[0] = raw[0] * inv_delta_max;
    f[1] = raw[1] * inv_delta_max;
    f[2] = raw[2] * inv_delta_max;
    f[3] = raw[3];
    f[4] = scale_feature;
}

/* C64:125-133 plateau_detected */
static int plateau_detected(double q0, double q1, double q2, double q3, double q4) {
    double raw[4];
    raw_sensors(q0, q1, q2, q3, q4, raw);
    double delta_max = dmax(dmax(raw[0], raw[1]), raw[2]);
    double q_scale = stencil_scale(q0, q1, q2, q3, q4);
    return delta_max <= 1.0e-13 * q_scale;
}

/* MLP parameter block: w1[5][10] b1[10] w2[10][6] b2[6] w3[6][6] b3[6] w4[6][3] b4[3] */
enum { OFF_W1 = 0, OFF_B1 = 50, OFF_W2 = 60, OFF_B2 = 120, OFF_W3 = 126, OFF_B3 = 162,
       OFF_W4 = 168, OFF_B4 = 186 };

/* C64:155-207 mlp_badness_ratios, FP64 */
static double swish64(double x) { return x / (1.0 + exp(-x)); }

static void mlp_ratios_f64(const double x[5], const double *p, double r[3]) {
    double h1[10], h2[6], h3[6];
    for (int o = 0; o < 10; ++o) {
        double z = p[OFF_B1 + o];
        for (int k = 0; k < 5; ++k) z = z + x[k] * p[OFF_W1 + k * 10 + o];
        h1[o] = swish64(z);
    }
    for (int o = 0; o < 6; ++o) {
        double z = p[OFF_B2 + o];
        for (int k = 0; k < 10; ++k) z = z + h1[k] * p[OFF_W2 + k * 6 + o];
        h2[o] = swish64(z);
    }
    for (int o = 0; o < 6; ++o) {
        double z = p[OFF_B3 + o];
        for (int k = 0; k < 6; ++k) z = z + h2[k] * p[OFF_W3 + k * 6 + o];
        h3[o] = swish64(z);
    }
    double rb0 = p[OFF_B4 + 0], rb1 = p[OFF_B4 + 1], rb2 = p[OFF_B4 + 2];
    for (int k = 0; k < 6; ++k) {
        rb0 = rb0 + h3[k] * p[OFF_W4 + k * 3 + 0];
        rb1 = rb1 + h3[k] * p[OFF_W4 + k * 3 + 1];
        rb2 = rb2 + h3[k] * p[OFF_W4 + k * 3 + 2];
    }
    double cmax = 6.0;
    double b0 = cmax * tanh(rb0 / cmax), b1 = cmax * tanh(rb1 / cmax), b2 = cmax * tanh(rb2 / cmax);
    double bmax = dmax(dmax(b0, b1), b2);
    double e0 = exp(b0 - bmax), e1 = exp(b1 - bmax), e2 = exp(b2 - bmax);
    double inv_sum = 1.0 / (e0 + e1 + e2);
    r[0] = e0 * inv_sum; r[1] = e1 * inv_sum; r[2] = e2 * inv_sum;
}

/* C32:152-202 mlp_badness_ratios, FP32 parameters and activations */
static float swish32(float x) { return x / (1.0f + expf(-x)); }

static void mlp_ratios_f32(const float x[5], const float *p, float r[3]) {
    float h1[10], h2[6], h3[6];
    for (int o = 0; o < 10; ++o) {
        float z = p[OFF_B1 + o];
        for (int k = 0; k < 5; ++k) z = z + x[k] * p[OFF_W1 + k * 10 + o];
        h1[o] = swish32(z);
    }
    for (int o = 0; o < 6; ++o) {
        float z = p[OFF_B2 + o];
        for (int k = 0; k < 10; ++k) z = z + h1[k] * p[OFF_W2 + k * 6 + o];
        h2[o] = swish32(z);
    }
    for (int o = 0; o < 6; ++o) {
        float z = p[OFF_B3 + o];
        for (int k = 0; k < 6; ++k) z = z + h2[k] * p[OFF_W3 + k * 6 + o];
        h3[o] = swish32(z);
    }
    float rb0 = p[OFF_B4 + 0], rb1 = p[OFF_B4 + 1], rb2 = p[OFF_B4 + 2];
    for (int k = 0; k < 6; ++k) {
        rb0 = rb0 + h3[k] * p[OFF_W4 + k * 3 + 0];
        rb1 = rb1 + h3[k] * p[OFF_W4 + k * 3 + 1];
        rb2 = rb2 + h3[k] * p[OFF_W4 + k * 3 + 2];
    }
    float cmax = 6.0f;
    float b0 = cmax * tanhf(rb0 / cmax), b1 = cmax * tanhf(rb1 / cmax), b2 = cmax * tanhf(rb2 / cmax);
    float bmax = fmaxf_(fmaxf_(b0, b1), b2);
    float e0 = expf(b0 - bmax), e1 = expf(b1 - bmax), e2 = expf(b2 - bmax);
    float inv_sum = 1.0f / (e0 + e1 + e2);
    r[0] = e0 * inv_sum; r[1] = e1 * inv_sum; r[2] = e2 * inv_sum;
}

/* Emulated BF16 closure (no counterpart in the reference): parameters and
 * layer inputs are rounded to bfloat16 (round-to-nearest-even), each affine
 * layer accumulates the exact bf16*bf16 products in FP32, and activations,
 * tanh and softmax run in FP32. */
static float round_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x; /* inf / nan unchanged */
    uint32_t lsb = (u >> 16) & 1u;
    u = (u + 0x7fffu + lsb) & 0xffff0000u;
    float y;
    memcpy(&y, &u, 4);
    return y;
}

static void mlp_ratios_bf16(const float x[5], const float *p, float r[3]) {
    float xin[10], h[10];
    for (int k = 0; k < 5; ++k) xin[k] = round_bf16(x[k]);
    for (int o = 0; o < 10; ++o) {
        float z = p[OFF_B1 + o];
        for (int k = 0; k < 5; ++k) z = z + xin[k] * round_bf16(p[OFF_W1 + k * 10 + o]);
        h[o] = round_bf16(swish32(z));
    }
    for (int k = 0; k < 10; ++k) xin[k] = h[k];
    for (int o = 0; o < 6; ++o) {
        float z = p[OFF_B2 + o];
        for (int k = 0; k < 10; ++k) z = z + xin[k] * round_bf16(p[OFF_W2 + k * 6 + o]);
        h[o] = round_bf16(swish32(z));
    }
    for (int k = 0; k < 6; ++k) xin[k] = h[k];
    for (int o = 0; o < 6; ++o) {
        float z = p[OFF_B3 + o];
        for (int k = 0; k < 6; ++k) z = z + xin[k] * round_bf16(p[OFF_W3 + k * 6 + o]);
        h[o] = round_bf16(swish32(z));
    }
    float rb0 = p[OFF_B4 + 0], rb1 = p[OFF_B4 + 1], rb2 = p[OFF_B4 + 2];
    for (int k = 0; k < 6; ++k) {
        rb0 = rb0 + h[k] * round_bf16(p[OFF_W4 + k * 3 + 0]);
        rb1 = rb1 + h[k] * round_bf16(p[OFF_W4 + k * 3 + 1]);
        rb2 = rb2 + h[k] * round_bf16(p[OFF_W4 + k * 3 + 2]);
    }
    float cmax = 6.0f;
    float b0 = cmax * tanhf(rb0 / cmax), b1 = cmax * tanhf(rb1 / cmax), b2 = cmax * tanhf(rb2 / cmax);
    float bmax = fmaxf_(fmaxf_(b0, b1), b2);
    float e0 = expf(b0 - bmax), e1 = expf(b1 - bmax), e2 = expf(b2 - bmax);
    float inv_sum = 1.0f / (e0 + e1 + e2);
    r[0] = e0 * inv_sum; r[1] = e1 * inv_sum; r[2] = e2 * inv_sum;
}

/* raw ratios M(x) in FP64 for the given closure precision (C64:155 / C32:152) */
static void raw_ratios(const double f[5], int scheme, const void *mlp, double r[3]) {
    if (scheme == SCHEME_SR64) {
        mlp_ratios_f64(f, (const double *)mlp, r);
        return;
    }
    float xf[5], rf[3];
    for (int k = 0; k < 5; ++k) xf[k] = (float)f[k];
    if (scheme == SCHEME_SR32) mlp_ratios_f32(xf, (const float *)mlp, rf);
    else mlp_ratios_bf16(xf, (const float *)mlp, rf);
    r[0] = (double)rf[0]; r[1] = (double)rf[1]; r[2] = (double)rf[2];
}

/* C64:209-281 / C32:206-283 mlp_weights: reflection average 0.5[M(x) + P3 M(P5 x)] */
static void mlp_weights(const double q[5], int head, int scheme, const void *mlp, int eno_cutoff,
                        double w[3]) {
    double c[3];
    optimal_weights(head, c);
    if (plateau_detected(q[0], q[1], q[2], q[3], q[4])) {
        w[0] = c[0]; w[1] = c[1]; w[2] = c[2];
        return;
    }
    double fd[5], fr[5], direct[3], reflected[3];
    nn_features(q[0], q[1], q[2], q[3], q[4], fd);
    nn_features(q[4], q[3], q[2], q[1], q[0], fr);
    raw_ratios(fd, scheme, mlp, direct);
    raw_ratios(fr, scheme, mlp, reflected);
    double br0 = 0.5 * (direct[0] + reflected[2]);
    double br1 = 0.5 * (direct[1] + reflected[1]);
    double br2 = 0.5 * (direct[2] + reflected[0]);
    double eps_ratio = 1.0e-12;
    double ratio_scale = 3.0;
    double sb0 = ratio_scale * br0, sb1 = ratio_scale * br1, sb2 = ratio_scale * br2;
    double ib0 = 1.0 / (sb0 + eps_ratio), ib1 = 1.0 / (sb1 + eps_ratio), ib2 = 1.0 / (sb2 + eps_ratio);
    double ib0s = ib0 * ib0, ib1s = ib1 * ib1, ib2s = ib2 * ib2;
    double alpha0 = c[0] * ib0s, alpha1 = c[1] * ib1s, alpha2 = c[2] * ib2s;
    double alpha_sum = alpha0 + alpha1 + alpha2;
    double inv_alpha_sum = 1.0 / dmax(alpha_sum, 1.0e-300);
    double w0 = alpha0 * inv_alpha_sum, w1 = alpha1 * inv_alpha_sum, w2 = alpha2 * inv_alpha_sum;
    if (eno_cutoff) {
        double psi0 = 1.0, psi1 = 1.0, psi2 = 1.0, cutoff = 4.0e-7;
        if (w0 <= cutoff) psi0 = 0.0;
        if (w1 <= cutoff) psi1 = 0.0;
        if (w2 <= cutoff) psi2 = 0.0;
        double cutoff_sum = psi0 * w0 + psi1 * w1 + psi2 * w2;
        double inv_cutoff_sum = 1.0 / dmax(cutoff_sum, 1.0e-300);
        w0 = psi0 * w0 * inv_cutoff_sum;
        w1 = psi1 * w1 * inv_cutoff_sum;
        w2 = psi2 * w2 * inv_cutoff_sum;
    }
    w[0] = w0; w[1] = w1; w[2] = w2;
}

/* C64:299-334 weno5_lr_value_mlp / C64:337-372 weno5_gauss_value_mlp */
static double weno5_sr_value(const double q[5], int head, int scheme, const void *mlp, int eno_cutoff) {
    double s[3], w[3];
    candidates(q, head, s);
    mlp_weights(q, head, scheme, mlp, eno_cutoff, w);
    return w[0] * s[0] + w[1] * s[1] + w[2] * s[2];
}

/* one scalar reconstruction for any closure; h is the WENO-Z grid spacing */
static double scalar_value(const double q[5], int head, int scheme, const void *mlp, int eno_cutoff,
                           double h) {
    if (scheme == SCHEME_JS) return weno5_js_value(q, head);
    if (scheme == SCHEME_Z) return weno5_z_value(q, head, h);
    return weno5_sr_value(q, head, scheme, mlp, eno_cutoff);
}

/* ------------------------------------------------------------------------ */
/* padded-array helpers: layout [j][i][comp], shape (ny+2g, nx+2g, 4)        */
/* ------------------------------------------------------------------------ */
#define IDX(j, i, c, nxt) ((((size_t)(j)) * (size_t)(nxt) + (size_t)(i)) * 4u + (size_t)(c))

static vec4 load4(const double *a, int j, int i, int nxt) {
    const double *p = a + IDX(j, i, 0, nxt);
    return v4(p[0], p[1], p[2], p[3]);
}
static void store4(double *a, int j, int i, int nxt, vec4 v) {
    double *p = a + IDX(j, i, 0, nxt);
    p[0] = v.v[0]; p[1] = v.v[1]; p[2] = v.v[2]; p[3] = v.v[3];
}

/* ------------------------------------------------------------------------ */
/* boundary kernels                                                         */
/* ------------------------------------------------------------------------ */
/* C64:391-409 copy_transmissive_boundary_kernel, C64:371-388 periodic */
static void copy_boundary(const double *src, double *dst, int nx, int ny, int gc, int periodic) {
    int nxt = nx + 2 * gc, nyt = ny + 2 * gc;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nyt; ++j) {
        for (int i = 0; i < nxt; ++i) {
            int si = i, sj = j;
            if (periodic) {
                if (i < gc) si = i + nx; else if (i >= nx + gc) si = i - nx;
                if (j < gc) sj = j + ny; else if (j >= ny + gc) sj = j - ny;
            } else {
                if (i < gc) si = gc; else if (i >= nx + gc) si = nx + gc - 1;
                if (j < gc) sj = gc; else if (j >= ny + gc) sj = ny + gc - 1;
            }
            store4(dst, j, i, nxt, load4(src, sj, si, nxt));
        }
    }
}

/* A64:25-54 double_mach_exact_conserved */
static vec4 double_mach_exact_conserved(double x, double y, double t, double gamma) {
    double root3 = sqrt(3.0);
    double x0 = 1.0 / 6.0;
    double shock_x = x0 + y / root3 + 20.0 * t / root3;
    if (x < shock_x) return pri_to_con(v4(8.0, 8.25 * root3 * 0.5, -8.25 * 0.5, 116.5), gamma);
    return pri_to_con(v4(1.4, 0.0, 0.0, 1.0), gamma);
}

/* A64:57-124 apply_double_mach_boundary_kernel (in place, after a transmissive copy) */
static void apply_double_mach(double *u, int nx, int ny, int gc, double dx, double dy, double time,
                              double gamma) {
    int nxt = nx + 2 * gc, nyt = ny + 2 * gc;
    for (int j = 0; j < nyt; ++j) {
        for (int i = 0; i < nxt; ++i) {
            int inside_x = i >= gc && i < nx + gc;
            int inside_y = j >= gc && j < ny + gc;
            if (inside_x && inside_y) continue;
            double x = ((double)(i - gc) + 0.5) * dx;
            double y = ((double)(j - gc) + 0.5) * dy;
            double x0 = 1.0 / 6.0;
            vec4 q = v4(0, 0, 0, 0);
            int use_exact = 0, use_reflect = 0, si = i, sj = j;
            if (j >= ny + gc) {
                q = double_mach_exact_conserved(x, y, time, gamma); use_exact = 1;
            } else if (j < gc) {
                if (x < x0) {
                    q = double_mach_exact_conserved(x, 0.0, time, gamma); use_exact = 1;
                } else {
                    use_reflect = 1;
                    sj = 2 * gc - 1 - j;
                    if (i < gc) si = gc; else if (i >= nx + gc) si = nx + gc - 1;
                }
            } else if (i < gc) {
                q = double_mach_exact_conserved(x, y, time, gamma); use_exact = 1;
            } else if (i >= nx + gc) {
                si = nx + gc - 1; sj = j;
            }
            if (use_exact) {
                store4(u, j, i, nxt, q);
            } else if (use_reflect) {
                vec4 s = load4(u, sj, si, nxt);
                s.v[2] = -s.v[2];
                store4(u, j, i, nxt, s);
            } else {
                store4(u, j, i, nxt, load4(u, sj, si, nxt));
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* normal (characteristic) reconstruction                                    */
/* H5:765-795 weno5_lr_value_characteristic, H5:948-979 / H5:1047-1078 (JS), */
/* C64:480-596 (SR).  Thread (j, i) covers the stencil u[j, i..i+4] (x) or    */
/* u[j..j+4, i] (y); lr=2 (left face, Roe(q1,q2)) -> *_l, lr=1 (right face,   */
/* Roe(q2,q3)) -> *_r.                                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    int scheme, eno_cutoff;
    const void *mlp;
    double gamma;
} closure_t;

static vec4 char_face(const vec4 qs[5], int lr, int direction, const closure_t *cl, double h) {
    vec4 left = qs[2], right = qs[3];
    if (lr == 2) { left = qs[1]; right = qs[2]; }
    vec4 roe = roe_average_state(left, right, cl->gamma);
    vec4 c[5];
    for (int m = 0; m < 5; ++m) c[m] = con_to_char(qs[m], roe, direction, cl->gamma);
    int head = lr == 1 ? HEAD_NORMAL_LR1 : HEAD_NORMAL_LR2;
    vec4 face;
    for (int k = 0; k < 4; ++k) {
        double q[5] = {c[0].v[k], c[1].v[k], c[2].v[k], c[3].v[k], c[4].v[k]};
        face.v[k] = scalar_value(q, head, cl->scheme, cl->mlp, cl->eno_cutoff, h);
    }
    return char_to_con(face, roe, direction, cl->gamma);
}

static void normal_stage(const double *bc, double *out_l, double *out_r, int nx, int ny, int gc,
                         int direction, const closure_t *cl, double h) {
    int nxt = nx + 2 * gc;
    int jn = direction == 1 ? ny + 6 : ny + 2;
    int in = direction == 1 ? nx + 2 : nx + 6;
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < jn; ++j) {
        for (int i = 0; i < in; ++i) {
            vec4 qs[5];
            for (int m = 0; m < 5; ++m)
                qs[m] = direction == 1 ? load4(bc, j, i + m, nxt) : load4(bc, j + m, i, nxt);
            store4(out_l, j, i, nxt, char_face(qs, 2, direction, cl, h));
            store4(out_r, j, i, nxt, char_face(qs, 1, direction, cl, h));
        }
    }
}

/* ------------------------------------------------------------------------ */
/* transverse Gauss reconstruction + Riemann flux                            */
/* AJS:226-285 (JS), A32:21-108 / A64:127-266 (SR), CB:572-712 (FORCE),       */
/* WZ-based weno5_hllc_z_min.py:170-245 (Z).                                  */
/* ------------------------------------------------------------------------ */
static void flux_stage(const double *t_l, const double *t_r, double *flux, int nx, int ny, int gc,
                       int direction, int loca, double dt_h, int riemann, const closure_t *cl,
                       double h_transverse) {
    int nxt = nx + 2 * gc;
    int jn = direction == 1 ? ny : ny + 1;
    int in = direction == 1 ? nx + 1 : nx;
    int head = loca == 1 ? HEAD_GAUSS_LR1 : HEAD_GAUSS_LR2;
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < jn; ++j) {
        for (int i = 0; i < in; ++i) {
            vec4 ul, ur;
            for (int comp = 0; comp < 4; ++comp) {
                double qr[5], ql[5];
                for (int m = 0; m < 5; ++m) {
                    if (direction == 1) {
                        qr[m] = t_l[IDX(j + 1 + m, i + 1, comp, nxt)];
                        ql[m] = t_r[IDX(j + 1 + m, i, comp, nxt)];
                    } else {
                        qr[m] = t_l[IDX(j + 1, i + 1 + m, comp, nxt)];
                        ql[m] = t_r[IDX(j, i + 1 + m, comp, nxt)];
                    }
                }
                ur.v[comp] = scalar_value(qr, head, cl->scheme, cl->mlp, cl->eno_cutoff, h_transverse);
                ul.v[comp] = scalar_value(ql, head, cl->scheme, cl->mlp, cl->eno_cutoff, h_transverse);
            }
            vec4 f = riemann == RIEMANN_FORCE ? force_flux(ul, ur, direction, dt_h, cl->gamma)
                                              : hllc_flux(ul, ur, direction, cl->gamma);
            for (int comp = 0; comp < 4; ++comp) flux[IDX(j, i, comp, nxt)] = f.v[comp] * dt_h;
        }
    }
}

/* CB:1003-1032 / C64:1028 update_rk3_out_kernel */
static void update_rk3(const double *u_stage, const double *u0, double *u_out, const double *fx1,
                       const double *fx2, const double *fy1, const double *fy2, int nx, int ny,
                       int gc, int rk) {
    int nxt = nx + 2 * gc;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            int jj = j + gc, ii = i + gc;
            for (int comp = 0; comp < 4; ++comp) {
                double fx_r = fx1[IDX(j, i + 1, comp, nxt)] + fx2[IDX(j, i + 1, comp, nxt)];
                double fx_l = fx1[IDX(j, i, comp, nxt)] + fx2[IDX(j, i, comp, nxt)];
                double fy_t = fy1[IDX(j + 1, i, comp, nxt)] + fy2[IDX(j + 1, i, comp, nxt)];
                double fy_b = fy1[IDX(j, i, comp, nxt)] + fy2[IDX(j, i, comp, nxt)];
                double rhs = u_stage[IDX(jj, ii, comp, nxt)] - 0.5 * (fx_r - fx_l) - 0.5 * (fy_t - fy_b);
                double out;
                if (rk == 1) out = rhs;
                else if (rk == 2) out = 0.75 * u0[IDX(jj, ii, comp, nxt)] + 0.25 * rhs;
                else out = (1.0 / 3.0) * u0[IDX(jj, ii, comp, nxt)] + (2.0 / 3.0) * rhs;
                u_out[IDX(jj, ii, comp, nxt)] = out;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* public entry points (ctypes)                                             */
/* ------------------------------------------------------------------------ */
size_t oracle_euler2d_work_doubles(int nx, int ny, int gc) {
    return (size_t)9 * (size_t)(nx + 2 * gc) * (size_t)(ny + 2 * gc) * 4u;
}

/* A32:121-168 / A64:298-435 / AJS:390-454 launch_stage.  work must hold
 * oracle_euler2d_work_doubles() doubles (bc, x_l, x_r, y_l, y_r, fx1, fx2, fy1, fy2). */
int oracle_euler2d_stage(const double *u_stage, const double *u0, double *u_out, int nx, int ny,
                         int gc, double dx, double dy, double gamma, double dt, int rk, int scheme,
                         int riemann, int boundary, double boundary_time, int eno_cutoff,
                         const void *mlp, double *work) {
    if (gc != 3 || nx < 1 || ny < 1) return -1;
    if (scheme < SCHEME_JS || scheme > SCHEME_Z) return -2;
    if (scheme != SCHEME_JS && scheme != SCHEME_Z && mlp == NULL) return -3;
    size_t n = (size_t)(nx + 2 * gc) * (size_t)(ny + 2 * gc) * 4u;
    double *bc = work, *x_l = work + n, *x_r = work + 2 * n, *y_l = work + 3 * n, *y_r = work + 4 * n;
    double *fx1 = work + 5 * n, *fx2 = work + 6 * n, *fy1 = work + 7 * n, *fy2 = work + 8 * n;
    memset(work, 0, 9 * n * sizeof(double));
    if (boundary == BC_PERIODIC) {
        copy_boundary(u_stage, bc, nx, ny, gc, 1);
    } else if (boundary == BC_TRANSMISSIVE || boundary == BC_DOUBLE_MACH) {
        copy_boundary(u_stage, bc, nx, ny, gc, 0);
        if (boundary == BC_DOUBLE_MACH) apply_double_mach(bc, nx, ny, gc, dx, dy, boundary_time, gamma);
    } else {
        return -4;
    }
    closure_t cl = {scheme, eno_cutoff, mlp, gamma};
    double dt_dx = dt / dx, dt_dy = dt / dy;
    normal_stage(bc, x_l, x_r, nx, ny, gc, 1, &cl, dx);
    flux_stage(x_l, x_r, fx1, nx, ny, gc, 1, 1, dt_dx, riemann, &cl, dy);
    flux_stage(x_l, x_r, fx2, nx, ny, gc, 1, 2, dt_dx, riemann, &cl, dy);
    normal_stage(bc, y_l, y_r, nx, ny, gc, 2, &cl, dy);
    flux_stage(y_l, y_r, fy1, nx, ny, gc, 2, 1, dt_dy, riemann, &cl, dx);
    flux_stage(y_l, y_r, fy2, nx, ny, gc, 2, 2, dt_dy, riemann, &cl, dx);
    update_rk3(bc, u0, u_out, fx1, fx2, fy1, fy2, nx, ny, gc, rk);
    return 0;
}

/* H5:876-897 compute_max_speed_kernel + np.max (NaN propagates like np.max) */
double oracle_euler2d_max_speed(const double *u, int nx, int ny, int gc, double gamma) {
    int nxt = nx + 2 * gc;
    double best = 0.0;
    int saw_nan = 0;
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            int jj = j + gc, ii = i + gc;
            double rho = dmax(u[IDX(jj, ii, 0, nxt)], 1.0e-15);
            double vel_x = u[IDX(jj, ii, 1, nxt)] / rho;
            double vel_y = u[IDX(jj, ii, 2, nxt)] / rho;
            double pressure = (gamma - 1.0) * (u[IDX(jj, ii, 3, nxt)] - 0.5 * rho * (vel_x * vel_x + vel_y * vel_y));
            pressure = dmax(pressure, 1.0e-15);
            double sound = sqrt(gamma * pressure / rho);
            double speed = dmax(fabs(vel_x) + sound, fabs(vel_y) + sound);
            if (speed != speed) saw_nan = 1;
            else if (speed > best) best = speed;
        }
    }
    return saw_nan ? NAN : best;
}

/* H5:189-207 compute_dt_from_warp_array */
double oracle_euler2d_dt(const double *u, int nx, int ny, int gc, double dx, double dy, double gamma,
                         double cfl) {
    double max_speed = oracle_euler2d_max_speed(u, nx, ny, gc, gamma);
    if (max_speed < 1.0e-15) return 1.0e10;
    return cfl * dmin(dx, dy) / max_speed;
}

/* A64:462-544 / A32:185-212 run_to_time without reporting.  u is advanced in place;
 * tmp must hold 3 padded states, work oracle_euler2d_work_doubles().  Returns the
 * number of steps; *t_out receives the final time and dts[] the step sizes. */
int oracle_euler2d_run(double *u, int nx, int ny, int gc, double dx, double dy, double gamma,
                       double cfl, double t_end, int max_steps, int scheme, int riemann,
                       int boundary, int eno_cutoff, const void *mlp, double *tmp, double *work,
                       double *dts, double *t_out) {
    size_t n = (size_t)(nx + 2 * gc) * (size_t)(ny + 2 * gc) * 4u;
    double *u1 = tmp, *u2 = tmp + n, *u3 = tmp + 2 * n;
    double t = 0.0;
    int step = 0;
    while (t < t_end - 1.0e-14 && step < max_steps) {
        double dt = dmin(oracle_euler2d_dt(u, nx, ny, gc, dx, dy, gamma, cfl), t_end - t);
        if (!isfinite(dt) || dt <= 0.0) break;
        int rc = oracle_euler2d_stage(u, u, u1, nx, ny, gc, dx, dy, gamma, dt, 1, scheme, riemann,
                                      boundary, t, eno_cutoff, mlp, work);
        if (rc) return rc;
        oracle_euler2d_stage(u1, u, u2, nx, ny, gc, dx, dy, gamma, dt, 2, scheme, riemann, boundary,
                             t + dt, eno_cutoff, mlp, work);
        oracle_euler2d_stage(u2, u, u3, nx, ny, gc, dx, dy, gamma, dt, 3, scheme, riemann, boundary,
                             t + 0.5 * dt, eno_cutoff, mlp, work);
        /* the reference swaps u0/u3; only the interior of u3 is defined */
        for (int j = gc; j < ny + gc; ++j)
            memcpy(u + IDX(j, gc, 0, nx + 2 * gc), u3 + IDX(j, gc, 0, nx + 2 * gc), (size_t)nx * 4u * sizeof(double));
        t += dt;
        if (dts) dts[step] = dt;
        step += 1;
    }
    if (t_out) *t_out = t;
    return step;
}

/* ------------------------------------------------------------------------ */
/* unit-level entry points used by the tests                                */
/* ------------------------------------------------------------------------ */
void oracle_hllc(const double *ul, const double *ur, int direction, double gamma, double *out) {
    vec4 f = hllc_flux(v4(ul[0], ul[1], ul[2], ul[3]), v4(ur[0], ur[1], ur[2], ur[3]), direction, gamma);
    memcpy(out, f.v, sizeof f.v);
}

void oracle_force(const double *ul, const double *ur, int direction, double c, double gamma, double *out) {
    vec4 f = force_flux(v4(ul[0], ul[1], ul[2], ul[3]), v4(ur[0], ur[1], ur[2], ur[3]), direction, c, gamma);
    memcpy(out, f.v, sizeof f.v);
}

void oracle_physical_flux(const double *q, int direction, double gamma, double *out) {
    vec4 f = pri_to_flux(con_to_pri(v4(q[0], q[1], q[2], q[3]), gamma), direction, gamma);
    memcpy(out, f.v, sizeof f.v);
}

void oracle_roe(const double *l, const double *r, double gamma, double *out) {
    vec4 f = roe_average_state(v4(l[0], l[1], l[2], l[3]), v4(r[0], r[1], r[2], r[3]), gamma);
    memcpy(out, f.v, sizeof f.v);
}

void oracle_char_roundtrip(const double *q, const double *roe, int direction, double gamma,
                           double *c_out, double *q_out) {
    vec4 rv = v4(roe[0], roe[1], roe[2], roe[3]);
    vec4 c = con_to_char(v4(q[0], q[1], q[2], q[3]), rv, direction, gamma);
    vec4 b = char_to_con(c, rv, direction, gamma);
    memcpy(c_out, c.v, sizeof c.v);
    memcpy(q_out, b.v, sizeof b.v);
}

void oracle_candidates(const double *q, int head, double *s) { candidates(q, head, s); }

double oracle_scalar_value(const double *q, int head, int scheme, const void *mlp, int eno_cutoff,
                           double h) {
    return scalar_value(q, head, scheme, mlp, eno_cutoff, h);
}

void oracle_features(const double *q, double *f) { nn_features(q[0], q[1], q[2], q[3], q[4], f); }

int oracle_plateau(const double *q) { return plateau_detected(q[0], q[1], q[2], q[3], q[4]); }

/* raw (one-orientation) MLP ratios for a batch of feature vectors (n x 5) */
void oracle_mlp_raw(const double *features, int n, int scheme, const void *mlp, double *out) {
    for (int b = 0; b < n; ++b) raw_ratios(features + 5 * b, scheme, mlp, out + 3 * b);
}

void oracle_mlp_weights(const double *q, int head, int scheme, const void *mlp, int eno_cutoff,
                        double *w) {
    mlp_weights(q, head, scheme, mlp, eno_cutoff, w);
}

float oracle_round_bf16(float x) { return round_bf16(x); }

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ======================================================================== */
/* 1-D periodic scalar advection (torch path of the reference)              */
/*   ADV = code/snapshot/teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/rk3_advection.py */
/*   W5  = code/snapshot/teacherfree_lab_weno5/weno5_core.py                 */
/*   AP  = code/snapshot/teacherfree_lab_weno5/apost_advect.py               */
/*   V9  = code/snapshot/teacherfree_lab_weno5_v9_exact_fvm/v9_model.py      */
/*   V12 = code/snapshot/teacherfree_lab_weno5_v12_reflection_sym/v12_model.py */
/*   M32 = code/snapshot/teacherfree_lab_weno5_v20_distance_balanced_mlp_f32/v20_mlp_f32_model.py */
/* ======================================================================== */

/* W5:186-201 weno5_features (scale channel (log10+16)/16), followed by the
 * model's V9:16-28 remap to (log10+4)/4 as the V12 / M32 models do */
static void features_1d(const double q[5], double f[5]) {
    double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4];
    double d20 = q0 - 2.0 * q1 + q2;
    double d21 = q1 - 2.0 * q2 + q3;
    double d22 = q2 - 2.0 * q3 + q4;
    double delta0 = (13.0 / 12.0) * fabs(d20) + 0.25 * fabs(q0 - 4.0 * q1 + 3.0 * q2);
    double delta1 = (13.0 / 12.0) * fabs(d21) + 0.25 * fabs(q1 - q3);
    double delta2 = (13.0 / 12.0) * fabs(d22) + 0.25 * fabs(3.0 * q2 - 4.0 * q3 + q4);
    double delta_max = dmax(dmax(delta0, delta1), delta2);
    double inv_delta_max = 1.0 / dmax(delta_max, 1.0e-15);
    const double eps = 1.0e-15;
    double g0 = fabs(d20) / (fabs(q1 - q0) + fabs(q2 - q1) + eps);
    double g1 = fabs(d21) / (fabs(q2 - q1) + fabs(q3 - q2) + eps);
    double g2 = fabs(d22) / (fabs(q3 - q2) + fabs(q4 - q3) + eps);
    double gamma_s = dmin(dmax(dmax(dmax(g0, g1), g2), 0.0), 1.0);
    double q_scale = dmax(dmax(dmax(dmax(fabs(q0), fabs(q1)), fabs(q2)), dmax(fabs(q3), fabs(q4))), 1.0);
    double relative = dmax(delta_max / q_scale, 1.0e-30);
    double old_scale = dmin(dmax((log10(relative) + 16.0) / 16.0, 0.0), 1.0);
    double log10_rel = 16.0 * old_scale - 16.0;
    double new_scale = dmin(dmax((log10_rel - (-4.0)) / 4.0, 0.0), 1.0);
    f[0] = delta0 * inv_delta_max;
    f[1] = delta1 * inv_delta_max;
    f[2] = delta2 * inv_delta_max;
    f[3] = gamma_s;
    f[4] = new_scale;
}

/* W5:43-55 SharedBadnessMLP.forward: h = x @ w + b (the bias after the product) */
static void mlp_ratios_torch64(const double x[5], const double *p, double r[3]) {
    double h1[10], h2[6], h3[6], raw[3];
    for (int o = 0; o < 10; ++o) {
        double z = 0.0;
        for (int k = 0; k < 5; ++k) z = z + x[k] * p[OFF_W1 + k * 10 + o];
        z = z + p[OFF_B1 + o];
        h1[o] = z * (1.0 / (1.0 + exp(-z)));
    }
    for (int o = 0; o < 6; ++o) {
        double z = 0.0;
        for (int k = 0; k < 10; ++k) z = z + h1[k] * p[OFF_W2 + k * 6 + o];
        z = z + p[OFF_B2 + o];
        h2[o] = z * (1.0 / (1.0 + exp(-z)));
    }
    for (int o = 0; o < 6; ++o) {
        double z = 0.0;
        for (int k = 0; k < 6; ++k) z = z + h2[k] * p[OFF_W3 + k * 6 + o];
        z = z + p[OFF_B3 + o];
        h3[o] = z * (1.0 / (1.0 + exp(-z)));
    }
    for (int o = 0; o < 3; ++o) {
        double z = 0.0;
        for (int k = 0; k < 6; ++k) z = z + h3[k] * p[OFF_W4 + k * 3 + o];
        raw[o] = 6.0 * tanh((z + p[OFF_B4 + o]) / 6.0);
    }
    double m = dmax(dmax(raw[0], raw[1]), raw[2]);
    double e0 = exp(raw[0] - m), e1 = exp(raw[1] - m), e2 = exp(raw[2] - m);
    double s = e0 + e1 + e2;
    r[0] = e0 / s; r[1] = e1 / s; r[2] = e2 / s;
}

/* M32 raw_forward: the same in FP32 after the cast of the remapped features */
static void mlp_ratios_torch32(const float x[5], const float *p, float r[3]) {
    float h1[10], h2[6], h3[6], raw[3];
    for (int o = 0; o < 10; ++o) {
        float z = 0.0f;
        for (int k = 0; k < 5; ++k) z = z + x[k] * p[OFF_W1 + k * 10 + o];
        z = z + p[OFF_B1 + o];
        h1[o] = z * (1.0f / (1.0f + expf(-z)));
    }
    for (int o = 0; o < 6; ++o) {
        float z = 0.0f;
        for (int k = 0; k < 10; ++k) z = z + h1[k] * p[OFF_W2 + k * 6 + o];
        z = z + p[OFF_B2 + o];
        h2[o] = z * (1.0f / (1.0f + expf(-z)));
    }
    for (int o = 0; o < 6; ++o) {
        float z = 0.0f;
        for (int k = 0; k < 6; ++k) z = z + h2[k] * p[OFF_W3 + k * 6 + o];
        z = z + p[OFF_B3 + o];
        h3[o] = z * (1.0f / (1.0f + expf(-z)));
    }
    for (int o = 0; o < 3; ++o) {
        float z = 0.0f;
        for (int k = 0; k < 6; ++k) z = z + h3[k] * p[OFF_W4 + k * 3 + o];
        raw[o] = 6.0f * tanhf((z + p[OFF_B4 + o]) / 6.0f);
    }
    float m = fmaxf_(fmaxf_(raw[0], raw[1]), raw[2]);
    float e0 = expf(raw[0] - m), e1 = expf(raw[1] - m), e2 = expf(raw[2] - m);
    float s = e0 + e1 + e2;
    r[0] = e0 / s; r[1] = e1 / s; r[2] = e2 / s;
}

/* V12:46-50 / M32 forward: 0.5 [M(x) + P3 M(P5 x)] */
static void sym_ratios_1d(const double f[5], int scheme, const void *mlp, double r[3]) {
    double fr[5] = {f[2], f[1], f[0], f[3], f[4]};
    double d[3], e[3];
    if (scheme == SCHEME_SR64) {
        mlp_ratios_torch64(f, (const double *)mlp, d);
        mlp_ratios_torch64(fr, (const double *)mlp, e);
    } else {
        float xf[5], xr[5], rd[3], re[3];
        for (int k = 0; k < 5; ++k) { xf[k] = (float)f[k]; xr[k] = (float)fr[k]; }
        if (scheme == SCHEME_SR32) {
            mlp_ratios_torch32(xf, (const float *)mlp, rd);
            mlp_ratios_torch32(xr, (const float *)mlp, re);
        } else {
            mlp_ratios_bf16(xf, (const float *)mlp, rd);
            mlp_ratios_bf16(xr, (const float *)mlp, re);
        }
        for (int k = 0; k < 3; ++k) { d[k] = rd[k]; e[k] = re[k]; }
    }
    r[0] = 0.5 * (d[0] + e[2]);
    r[1] = 0.5 * (d[1] + e[1]);
    r[2] = 0.5 * (d[2] + e[0]);
}

/* W5:133-137 plateau_mask */
static int plateau_1d(const double q[5]) {
    double raw[4];
    raw_sensors(q[0], q[1], q[2], q[3], q[4], raw);
    double delta_max = dmax(dmax(raw[0], raw[1]), raw[2]);
    double q_scale = dmax(dmax(dmax(dmax(fabs(q[0]), fabs(q[1])), fabs(q[2])), dmax(fabs(q[3]), fabs(q[4]))), 1.0);
    return delta_max <= 1.0e-13 * q_scale;
}

/* one face value of the 1-D scheme: head 0 (lr=1, x_{i+1/2} of cell i) or 1 (lr=2) */
static double face_1d(const double q[5], int lr, int scheme, const void *mlp) {
    int head = lr == 1 ? HEAD_NORMAL_LR1 : HEAD_NORMAL_LR2;
    double s[3], d[3], w[3];
    candidates(q, head, s);
    optimal_weights(head, d);
    if (plateau_1d(q)) {
        w[0] = d[0]; w[1] = d[1]; w[2] = d[2];
    } else if (scheme == SCHEME_JS) {
        /* W5:235-239 classical_omega: d / (beta + 1e-6)^2, normalised */
        double b[3];
        js_beta(q, b);
        double a0 = d[0] / ((b[0] + 1.0e-6) * (b[0] + 1.0e-6));
        double a1 = d[1] / ((b[1] + 1.0e-6) * (b[1] + 1.0e-6));
        double a2 = d[2] / ((b[2] + 1.0e-6) * (b[2] + 1.0e-6));
        double sum = a0 + a1 + a2;
        w[0] = a0 / sum; w[1] = a1 / sum; w[2] = a2 / sum;
    } else {
        /* W5:220-224 omega_from_ratio: beta = 3 r, alpha = d / (beta + 1e-12)^2 */
        double f[5], r[3];
        features_1d(q, f);
        sym_ratios_1d(f, scheme, mlp, r);
        double a0 = d[0] / ((3.0 * r[0] + 1.0e-12) * (3.0 * r[0] + 1.0e-12));
        double a1 = d[1] / ((3.0 * r[1] + 1.0e-12) * (3.0 * r[1] + 1.0e-12));
        double a2 = d[2] / ((3.0 * r[2] + 1.0e-12) * (3.0 * r[2] + 1.0e-12));
        double sum = a0 + a1 + a2;
        w[0] = a0 / sum; w[1] = a1 / sum; w[2] = a2 / sum;
    }
    return w[0] * s[0] + w[1] * s[1] + w[2] * s[2];
}

/* ADV:82-115 rhs_signed for one trajectory; face i is x_{i+1/2} */
static void rhs_1d(const double *u, int n, double a, double dxinv, int scheme, const void *mlp,
                   double *face, double *rhs) {
    for (int i = 0; i < n; ++i) {
        double q[5];
        if (a > 0.0) {
            for (int m = 0; m < 5; ++m) q[m] = u[((i - 2 + m) % n + n) % n];
            face[i] = a * face_1d(q, 1, scheme, mlp);
        } else {
            for (int m = 0; m < 5; ++m) q[m] = u[((i - 1 + m) % n + n) % n];
            face[i] = a * face_1d(q, 2, scheme, mlp);
        }
    }
    for (int i = 0; i < n; ++i) rhs[i] = -(face[i] - face[(i - 1 + n) % n]) * dxinv;
}

/* ADV:118-131 ssprk3_step_signed (AP:41-45 for a = +1) */
static void ssprk3_1d(double *u, int n, double a, double dt, double dxinv, int scheme, const void *mlp,
                      double *scratch) {
    double *face = scratch, *rhs = scratch + n, *u1 = scratch + 2 * n, *u2 = scratch + 3 * n;
    rhs_1d(u, n, a, dxinv, scheme, mlp, face, rhs);
    for (int i = 0; i < n; ++i) u1[i] = u[i] + dt * rhs[i];
    rhs_1d(u1, n, a, dxinv, scheme, mlp, face, rhs);
    for (int i = 0; i < n; ++i) u2[i] = 0.75 * u[i] + 0.25 * (u1[i] + dt * rhs[i]);
    rhs_1d(u2, n, a, dxinv, scheme, mlp, face, rhs);
    for (int i = 0; i < n; ++i) u[i] = u[i] / 3.0 + (2.0 / 3.0) * (u2[i] + dt * rhs[i]);
}

/* batch of trajectories u (batch x n, advanced in place) with velocities a[b] in {-1,+1};
 * dts[nsteps] is the step schedule; dxinv = 1/dx.  Threads over the batch. */
int oracle_adv1d_run(double *u, int batch, int n, const double *a, const double *dts, int nsteps,
                     double dxinv, int scheme, const void *mlp) {
    if (n < 5 || batch < 1) return -1;
    if (scheme != SCHEME_JS && scheme != SCHEME_SR64 && scheme != SCHEME_SR32 && scheme != SCHEME_SRBF16)
        return -2;
    int rc = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < batch; ++b) {
        double *scratch = (double *)malloc(sizeof(double) * 4 * (size_t)n);
        if (!scratch) { rc = -3; continue; }
        for (int s = 0; s < nsteps; ++s) ssprk3_1d(u + (size_t)b * n, n, a[b], dts[s], dxinv, scheme, mlp, scratch);
        free(scratch);
    }
    return rc;
}

double oracle_face_1d(const double *q, int lr, int scheme, const void *mlp) { return face_1d(q, lr, scheme, mlp); }

void oracle_features_1d(const double *q, double *f) { features_1d(q, f); }
