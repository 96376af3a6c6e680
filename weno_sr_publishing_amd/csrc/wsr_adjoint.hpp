// wsr_adjoint.hpp -- reverse mode (vector-Jacobian product) of one 1-D WENO5-SR
// face reconstruction: d(face)/d(stencil) and d(face)/d(MLP parameters).
//
// It differentiates exactly the forward of wsr_adv1d.hpp (torch semantics of
// code/snapshot/teacherfree_lab_weno5/weno5_core.py:43-55,133-239 with the
// remap of teacherfree_lab_weno5_v9_exact_fvm/v9_model.py:16-28) and follows
// torch autograd's conventions at non-smooth points, so that it reproduces the
// reference's gradients: d|x|/dx = sign(x) with sign(0) = 0; maximum(a, b) gives
// half of the gradient to each argument on a tie; clamp passes the gradient on
// the closed interval [min, max]; max over |q| sends it to the first maximiser;
// plateau stencils (linear weights) carry no gradient into the closure.
#pragma once

#include "wsr_adv1d.hpp"

namespace wsr {

// one raw MLP evaluation with the values its reverse pass needs
template <typename T>
struct MlpTape {
    T x[5], z1[10], s1[10], z2[6], s2[6], z3[6], s3[6], t4[3], p[3];
};

// parameter-gradient row of one evaluation: inputs x, h1, h2, h3, 1 and output adjoints
// z1bar, z2bar, z3bar, z4bar.  d(loss)/d(param) = sum over rows of row[in] * row[out].
constexpr int ROW = 53;
constexpr int R_X = 0, R_H1 = 5, R_H2 = 15, R_H3 = 21, R_ONE = 27, R_Z1 = 28, R_Z2 = 38, R_Z3 = 44, R_Z4 = 50;

__device__ __forceinline__ void param_row_slots(int p, int &in, int &out) {
    if (p < 50) { in = R_X + p / 10; out = R_Z1 + p % 10; }
    else if (p < 60) { in = R_ONE; out = R_Z1 + (p - 50); }
    else if (p < 120) { in = R_H1 + (p - 60) / 6; out = R_Z2 + (p - 60) % 6; }
    else if (p < 126) { in = R_ONE; out = R_Z2 + (p - 120); }
    else if (p < 162) { in = R_H2 + (p - 126) / 6; out = R_Z3 + (p - 126) % 6; }
    else if (p < 168) { in = R_ONE; out = R_Z3 + (p - 162); }
    else if (p < 186) { in = R_H3 + (p - 168) / 3; out = R_Z4 + (p - 168) % 3; }
    else { in = R_ONE; out = R_Z4 + (p - 186); }
}

__device__ __forceinline__ double sigmoid_t(double z) { return 1.0 / (1.0 + exp(-z)); }
__device__ __forceinline__ float sigmoid_t(float z) { return rcp_fast(1.0f + exp_fast(-z)); }
__device__ __forceinline__ double exp_t(double z) { return exp(z); }
__device__ __forceinline__ float exp_t(float z) { return exp_fast(z); }
__device__ __forceinline__ double tanh_t(double z) { return tanh(z); }
__device__ __forceinline__ float tanh_t(float z) { return tanhf(z); }
// softmax normalisation exactly as the forward rollout computes it
__device__ __forceinline__ double div_t(double a, double b) { return a / b; }
__device__ __forceinline__ float div_t(float a, float b) { return a * rcp_fast(b); }

template <typename T, class MP>
__device__ __forceinline__ void mlp_forward_tape(const T xin[5], const MP &P, MlpTape<T> &tp) {
    for (int k = 0; k < 5; ++k) tp.x[k] = xin[k];
    T h1[10], h2[6], h3[6];
#pragma unroll
    for (int o = 0; o < 10; ++o) {
        T z = tp.x[0] * P.p[OFF_W1 + o];
#pragma unroll
        for (int k = 1; k < 5; ++k) z = z + tp.x[k] * P.p[OFF_W1 + k * 10 + o];
        z = z + P.p[OFF_B1 + o];
        T s = sigmoid_t(z);
        tp.z1[o] = z; tp.s1[o] = s; h1[o] = z * s;
    }
#pragma unroll
    for (int o = 0; o < 6; ++o) {
        T z = h1[0] * P.p[OFF_W2 + o];
#pragma unroll
        for (int k = 1; k < 10; ++k) z = z + h1[k] * P.p[OFF_W2 + k * 6 + o];
        z = z + P.p[OFF_B2 + o];
        T s = sigmoid_t(z);
        tp.z2[o] = z; tp.s2[o] = s; h2[o] = z * s;
    }
#pragma unroll
    for (int o = 0; o < 6; ++o) {
        T z = h2[0] * P.p[OFF_W3 + o];
#pragma unroll
        for (int k = 1; k < 6; ++k) z = z + h2[k] * P.p[OFF_W3 + k * 6 + o];
        z = z + P.p[OFF_B3 + o];
        T s = sigmoid_t(z);
        tp.z3[o] = z; tp.s3[o] = s; h3[o] = z * s;
    }
    T y[3];
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        T z = h3[0] * P.p[OFF_W4 + o];
#pragma unroll
        for (int k = 1; k < 6; ++k) z = z + h3[k] * P.p[OFF_W4 + k * 3 + o];
        T t = tanh_t((z + P.p[OFF_B4 + o]) / T(6));
        tp.t4[o] = t;
        y[o] = T(6) * t;
    }
    T m = y[0] > y[1] ? y[0] : y[1];
    m = m > y[2] ? m : y[2];
    T e0 = exp_t(y[0] - m), e1 = exp_t(y[1] - m), e2 = exp_t(y[2] - m);
    T sum = e0 + e1 + e2;
    tp.p[0] = div_t(e0, sum); tp.p[1] = div_t(e1, sum); tp.p[2] = div_t(e2, sum);
}

// reverse pass of one evaluation: pbar -> xbar, and its parameter-gradient row
template <typename T, class MP>
__device__ __forceinline__ void mlp_backward(const MlpTape<T> &tp, const T pbar[3], const MP &P, T xbar[5], T *row) {
    // softmax and the bounded logits y = 6 tanh(z / 6)
    T dot = pbar[0] * tp.p[0] + pbar[1] * tp.p[1] + pbar[2] * tp.p[2];
    T z4b[3];
    for (int o = 0; o < 3; ++o) z4b[o] = tp.p[o] * (pbar[o] - dot) * (T(1) - tp.t4[o] * tp.t4[o]);
    T h3[6], h2[6], h1[10];
    for (int k = 0; k < 6; ++k) h3[k] = tp.z3[k] * tp.s3[k];
    for (int k = 0; k < 6; ++k) h2[k] = tp.z2[k] * tp.s2[k];
    for (int k = 0; k < 10; ++k) h1[k] = tp.z1[k] * tp.s1[k];
    // swish'(z) = s (1 + z (1 - s))
    T z3b[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        T hb = P.p[OFF_W4 + k * 3 + 0] * z4b[0] + P.p[OFF_W4 + k * 3 + 1] * z4b[1] + P.p[OFF_W4 + k * 3 + 2] * z4b[2];
        z3b[k] = hb * tp.s3[k] * (T(1) + tp.z3[k] * (T(1) - tp.s3[k]));
    }
    T z2b[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        T hb = T(0);
#pragma unroll
        for (int o = 0; o < 6; ++o) hb = hb + P.p[OFF_W3 + k * 6 + o] * z3b[o];
        z2b[k] = hb * tp.s2[k] * (T(1) + tp.z2[k] * (T(1) - tp.s2[k]));
    }
    T z1b[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        T hb = T(0);
#pragma unroll
        for (int o = 0; o < 6; ++o) hb = hb + P.p[OFF_W2 + k * 6 + o] * z2b[o];
        z1b[k] = hb * tp.s1[k] * (T(1) + tp.z1[k] * (T(1) - tp.s1[k]));
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        T xb = T(0);
#pragma unroll
        for (int o = 0; o < 10; ++o) xb = xb + P.p[OFF_W1 + k * 10 + o] * z1b[o];
        xbar[k] = xb;
    }
    for (int k = 0; k < 5; ++k) row[R_X + k] = tp.x[k];
    for (int k = 0; k < 10; ++k) row[R_H1 + k] = h1[k];
    for (int k = 0; k < 6; ++k) row[R_H2 + k] = h2[k];
    for (int k = 0; k < 6; ++k) row[R_H3 + k] = h3[k];
    row[R_ONE] = T(1);
    for (int k = 0; k < 10; ++k) row[R_Z1 + k] = z1b[k];
    for (int k = 0; k < 6; ++k) row[R_Z2 + k] = z2b[k];
    for (int k = 0; k < 6; ++k) row[R_Z3 + k] = z3b[k];
    for (int k = 0; k < 3; ++k) row[R_Z4 + k] = z4b[k];
}

__device__ __forceinline__ double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// torch.maximum(a, b) backward
__device__ __forceinline__ void max_bwd(double a, double b, double g, double &ga, double &gb) {
    if (a > b) { ga += g; }
    else if (b > a) { gb += g; }
    else { ga += 0.5 * g; gb += 0.5 * g; }
}

// reverse pass of features_1d (non-plateau stencil): fbar (5) -> qbar (5, accumulated)
__device__ void features_1d_backward(const double q[5], const double fbar[5], double qbar[5]) {
#pragma clang fp contract(off)
    const double c1312 = 13.0 / 12.0, eps = 1.0e-15;
    double d2[3] = {q[0] - 2.0 * q[1] + q[2], q[1] - 2.0 * q[2] + q[3], q[2] - 2.0 * q[3] + q[4]};
    double e[3] = {q[0] - 4.0 * q[1] + 3.0 * q[2], q[1] - q[3], 3.0 * q[2] - 4.0 * q[3] + q[4]};
    double delta[3];
    for (int k = 0; k < 3; ++k) delta[k] = c1312 * fabs(d2[k]) + 0.25 * fabs(e[k]);
    double m01 = dmax(delta[0], delta[1]);
    double dmx = dmax(m01, delta[2]);
    double dcl = dmax(dmx, 1.0e-15);
    double inv = 1.0 / dcl;
    double dl[4] = {q[1] - q[0], q[2] - q[1], q[3] - q[2], q[4] - q[3]};
    double D[3], g[3];
    for (int k = 0; k < 3; ++k) {
        D[k] = fabs(dl[k]) + fabs(dl[k + 1]) + eps;
        g[k] = fabs(d2[k]) / D[k];
    }
    double g01 = dmax(g[0], g[1]);
    double gm = dmax(g01, g[2]);
    // q_scale = clamp(max |q|, min 1): first maximiser
    int am = 0;
    double qa = fabs(q[0]);
    for (int k = 1; k < 5; ++k)
        if (fabs(q[k]) > qa) { qa = fabs(q[k]); am = k; }
    double qs = dmax(qa, 1.0);
    double ratio = dmx / qs;
    double rel = dmax(ratio, 1.0e-30);
    double old_in = (log10(rel) + 16.0) / 16.0;
    double old = dmin(dmax(old_in, 0.0), 1.0);
    double new_in = (16.0 * old - 16.0 + 4.0) / 4.0;

    double db[3] = {0.0, 0.0, 0.0}, d2b[3] = {0.0, 0.0, 0.0}, eb[3] = {0.0, 0.0, 0.0}, dlb[4] = {0.0, 0.0, 0.0, 0.0};
    double dmxb = 0.0, qab = 0.0;
    // f4: clamp(new_in) <- old (clamp) <- log10(rel) <- clamp(ratio, 1e-30) <- dmx / qs
    double nb = (new_in >= 0.0 && new_in <= 1.0) ? fbar[4] : 0.0;
    double ob = 4.0 * nb;
    double oib = (old_in >= 0.0 && old_in <= 1.0) ? ob : 0.0;
    double relb = (oib / 16.0) / (rel * 2.302585092994046);
    double ratiob = ratio >= 1.0e-30 ? relb : 0.0;
    dmxb += ratiob / qs;
    double qsb = -ratiob * dmx / (qs * qs);
    if (qa >= 1.0) qab += qsb;
    // f3: clamp(max(max(g0, g1), g2), 0, 1)
    double gmb = (gm >= 0.0 && gm <= 1.0) ? fbar[3] : 0.0;
    double g01b = 0.0, gb[3] = {0.0, 0.0, 0.0};
    max_bwd(g01, g[2], gmb, g01b, gb[2]);
    max_bwd(g[0], g[1], g01b, gb[0], gb[1]);
    for (int k = 0; k < 3; ++k) {
        d2b[k] += gb[k] * sgn(d2[k]) / D[k];
        double Db = -gb[k] * g[k] / D[k];
        dlb[k] += Db * sgn(dl[k]);
        dlb[k + 1] += Db * sgn(dl[k + 1]);
    }
    // f0..f2 = delta_k / clamp(delta_max, 1e-15)
    double invb = 0.0;
    for (int k = 0; k < 3; ++k) {
        db[k] += fbar[k] * inv;
        invb += fbar[k] * delta[k];
    }
    double dclb = -invb * inv * inv;
    if (dmx >= 1.0e-15) dmxb += dclb;
    double m01b = 0.0;
    max_bwd(m01, delta[2], dmxb, m01b, db[2]);
    max_bwd(delta[0], delta[1], m01b, db[0], db[1]);
    for (int k = 0; k < 3; ++k) {
        d2b[k] += db[k] * c1312 * sgn(d2[k]);
        eb[k] += db[k] * 0.25 * sgn(e[k]);
    }
    // back to the stencil
    qbar[0] += d2b[0] + eb[0] - dlb[0];
    qbar[1] += -2.0 * d2b[0] + d2b[1] - 4.0 * eb[0] + eb[1] + dlb[0] - dlb[1];
    qbar[2] += d2b[0] - 2.0 * d2b[1] + d2b[2] + 3.0 * eb[0] + 3.0 * eb[2] + dlb[1] - dlb[2];
    qbar[3] += d2b[1] - 2.0 * d2b[2] - eb[1] - 4.0 * eb[2] + dlb[2] - dlb[3];
    qbar[4] += d2b[2] + eb[2] + dlb[3];
    qbar[am] += qab * sgn(q[am]);
}

// candidate coefficients (face x_{i+1/2} of cell i for lr1, else x_{i-1/2} of cell i)
__device__ __forceinline__ void candidates_transpose(bool lr1, const double sb[3], double qbar[5]) {
    const double t13 = 1.0 / 3.0, t76 = 7.0 / 6.0, t116 = 11.0 / 6.0, t16 = 1.0 / 6.0, t56 = 5.0 / 6.0;
    if (lr1) {
        qbar[0] += t13 * sb[0];
        qbar[1] += -t76 * sb[0] - t16 * sb[1];
        qbar[2] += t116 * sb[0] + t56 * sb[1] + t13 * sb[2];
        qbar[3] += t13 * sb[1] + t56 * sb[2];
        qbar[4] += -t16 * sb[2];
    } else {
        qbar[0] += -t16 * sb[0];
        qbar[1] += t56 * sb[0] + t13 * sb[1];
        qbar[2] += t13 * sb[0] + t56 * sb[1] + t116 * sb[2];
        qbar[3] += -t16 * sb[1] - t76 * sb[2];
        qbar[4] += t13 * sb[2];
    }
}

// VJP of one SR face value: returns the forward value, accumulates qbar (5) and
// writes the two parameter-gradient rows (direct, reflected); false if the
// stencil is a plateau (no closure, rows unused).
template <int SCHEME, typename T, class MP>
__device__ bool face_vjp(const double q[5], bool lr1, double vbar, const MP &P, double qbar[5], T *row_d,
                         T *row_r, double &value) {
    double s[3], d[3];
    candidates_normal(q, lr1, s);
    d_normal(lr1, d);
    for (int m = 0; m < 5; ++m) qbar[m] = 0.0;
    double f[5];
    if (!features_1d(q, f)) {
        value = d[0] * s[0] + d[1] * s[1] + d[2] * s[2];
        double sb[3] = {vbar * d[0], vbar * d[1], vbar * d[2]};
        candidates_transpose(lr1, sb, qbar);
        return false;
    }
    T xd[5] = {(T)f[0], (T)f[1], (T)f[2], (T)f[3], (T)f[4]};
    T xr[5] = {xd[2], xd[1], xd[0], xd[3], xd[4]};
    MlpTape<T> td, tr;
    mlp_forward_tape(xd, P, td);
    mlp_forward_tape(xr, P, tr);
    double r[3], a[3], w[3];
    for (int k = 0; k < 3; ++k) r[k] = 0.5 * ((double)td.p[k] + (double)tr.p[2 - k]);
    for (int k = 0; k < 3; ++k) {
        double den = 3.0 * r[k] + 1.0e-12;
        a[k] = d[k] / (den * den);
    }
    double S = a[0] + a[1] + a[2];
    for (int k = 0; k < 3; ++k) w[k] = a[k] / S;
    value = w[0] * s[0] + w[1] * s[1] + w[2] * s[2];
    // reverse
    double sb[3] = {vbar * w[0], vbar * w[1], vbar * w[2]};
    candidates_transpose(lr1, sb, qbar);
    double wb[3] = {vbar * s[0], vbar * s[1], vbar * s[2]};
    double dot = wb[0] * w[0] + wb[1] * w[1] + wb[2] * w[2];
    double rb[3];
    for (int k = 0; k < 3; ++k) {
        double ab = (wb[k] - dot) / S;
        rb[k] = ab * (-6.0 * a[k] / (3.0 * r[k] + 1.0e-12));
    }
    T pbd[3] = {(T)(0.5 * rb[0]), (T)(0.5 * rb[1]), (T)(0.5 * rb[2])};
    T pbr[3] = {(T)(0.5 * rb[2]), (T)(0.5 * rb[1]), (T)(0.5 * rb[0])};
    T xbd[5], xbr[5];
    mlp_backward(td, pbd, P, xbd, row_d);
    mlp_backward(tr, pbr, P, xbr, row_r);
    double fb[5] = {(double)xbd[0] + (double)xbr[2], (double)xbd[1] + (double)xbr[1], (double)xbd[2] + (double)xbr[0],
                    (double)xbd[3] + (double)xbr[3], (double)xbd[4] + (double)xbr[4]};
    features_1d_backward(q, fb, qbar);
    return true;
}

}  // namespace wsr
