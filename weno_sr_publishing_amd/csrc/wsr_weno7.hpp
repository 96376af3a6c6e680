// wsr_weno7.hpp -- device arithmetic of the WENO7 step (gfx950): candidate
// polynomials, smoothness indicators, the 6-24-16-16-4 closure and the EVILIN
// state.
//
// Reference files (paths relative to the reference checkout):
//   W7  = code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/weno7_core.py
//   WH7 = code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/
//         warp_weno7_ader4_helpers_classical_only.py
//   EXT = code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/
//         warp_weno7_external_overwrite_fullall.py
//   BP  = code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/beta_provider.py
//   Z7  = code/snapshot/weno_z_borges_p2_results/weno7_helpers_z.py
#pragma once

#include "wsr_math.hpp"

namespace wsr {

constexpr int MLP7_COUNT = 908;
// reference layout: w1[6][24] b1[24] w2[24][16] b2[16] w3[16][16] b3[16] w4[16][4] b4[4]
constexpr int O7_W1 = 0, O7_B1 = 144, O7_W2 = 168, O7_B2 = 552, O7_W3 = 568, O7_B3 = 824, O7_W4 = 840,
              O7_B4 = 904;
// device layout, neuron-major runs [b_o, w_0o, ..., w_(K-1)o]
constexpr int D7_L1 = 0, D7_L2 = 168, D7_L3 = 568, D7_L4 = 840;

template <int SCHEME> struct Mlp7Sweep { using type = MlpNone; };
template <> struct Mlp7Sweep<SCHEME_SR64> { using type = MlpK<double>; };
template <> struct Mlp7Sweep<SCHEME_SR32> { using type = MlpK<float>; };

template <class T>
__host__ __device__ inline void pack7_neuron_major(const T *ref, T *dev) {
    const int K[4] = {6, 24, 16, 16}, N[4] = {24, 16, 16, 4};
    const int W[4] = {O7_W1, O7_W2, O7_W3, O7_W4}, B[4] = {O7_B1, O7_B2, O7_B3, O7_B4};
    int o = 0;
    for (int l = 0; l < 4; ++l)
        for (int n = 0; n < N[l]; ++n) {
            dev[o++] = ref[B[l] + n];
            for (int k = 0; k < K[l]; ++k) dev[o++] = ref[W[l] + k * N[l] + n];
        }
}

// W7:28-128 candidate coefficients and optimal weights of the four heads
// (1: x_{i+1/2}, 2: x_{i-1/2}, 3: Gauss point -sqrt(3)/6, 4: +sqrt(3)/6), as the
// double values of the reference's expressions.  WH7's Gauss location 1..4
// (loca 1) are head 4, locations 5..8 (loca 2) head 3.
template <int HEAD>
__device__ __forceinline__ void candidates7(const double q[7], double s[4]) {
    constexpr double C[4][4][4] = {
        {{-0.25, 1.0833333333333333, -1.9166666666666667, 2.0833333333333335},
         {0.08333333333333333, -0.4166666666666667, 1.0833333333333333, 0.25},
         {-0.08333333333333333, 0.5833333333333334, 0.5833333333333334, -0.08333333333333333},
         {0.25, 1.0833333333333333, -0.4166666666666667, 0.08333333333333333}},
        {{0.08333333333333333, -0.4166666666666667, 1.0833333333333333, 0.25},
         {-0.08333333333333333, 0.5833333333333334, 0.5833333333333334, -0.08333333333333333},
         {0.25, 1.0833333333333333, -0.4166666666666667, 0.08333333333333333},
         {2.0833333333333335, -1.9166666666666667, 1.0833333333333333, -0.25}},
        {{0.08820629112619283, -0.4089564406759849, 0.8419691425682042, 0.4787810069815879},
         {-0.056131276171213614, 0.31273139581104725, 0.8316061714863592, -0.08820629112619283},
         {0.08820629112619283, 1.1683938285136408, -0.31273139581104725, 0.056131276171213614},
         {1.5212189930184121, -0.8419691425682042, 0.4089564406759849, -0.08820629112619283}},
        {{-0.08820629112619283, 0.4089564406759849, -0.8419691425682042, 1.5212189930184121},
         {0.056131276171213614, -0.31273139581104725, 1.1683938285136408, 0.08820629112619283},
         {-0.08820629112619283, 0.8316061714863592, 0.31273139581104725, -0.056131276171213614},
         {0.4787810069815879, 0.8419691425682042, -0.4089564406759849, 0.08820629112619283}}};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        s[k] = q[k] * C[HEAD - 1][k][0] + q[k + 1] * C[HEAD - 1][k][1] + q[k + 2] * C[HEAD - 1][k][2] +
               q[k + 3] * C[HEAD - 1][k][3];
}

template <int HEAD>
__device__ __forceinline__ void optimal7(double d[4]) {
    constexpr double D[4][4] = {
        {0.02857142857142857, 0.34285714285714286, 0.5142857142857142, 0.11428571428571428},
        {0.11428571428571428, 0.5142857142857142, 0.34285714285714286, 0.02857142857142857},
        {0.06756615284017824, 0.4381942580088793, 0.42771483290021167, 0.06652475625073086},
        {0.06652475625073086, 0.42771483290021167, 0.4381942580088793, 0.06756615284017824}};
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = D[HEAD - 1][k];
}

// W7:139-163 the twelve difference forms of the four sub-stencils
__device__ __forceinline__ void delta_forms7(const double q[7], double f[4][3]) {
    f[0][0] = -2.0 * q[0] + 9.0 * q[1] - 18.0 * q[2] + 11.0 * q[3];
    f[0][1] = -q[0] + 4.0 * q[1] - 5.0 * q[2] + 2.0 * q[3];
    f[0][2] = -q[0] + 3.0 * q[1] - 3.0 * q[2] + q[3];
    f[1][0] = q[1] - 6.0 * q[2] + 3.0 * q[3] + 2.0 * q[4];
    f[1][1] = q[2] - 2.0 * q[3] + q[4];
    f[1][2] = -q[1] + 3.0 * q[2] - 3.0 * q[3] + q[4];
    f[2][0] = -2.0 * q[2] - 3.0 * q[3] + 6.0 * q[4] - q[5];
    f[2][1] = q[2] - 2.0 * q[3] + q[4];
    f[2][2] = -q[2] + 3.0 * q[3] - 3.0 * q[4] + q[5];
    f[3][0] = -11.0 * q[3] + 18.0 * q[4] - 9.0 * q[5] + 2.0 * q[6];
    f[3][1] = 2.0 * q[3] - 5.0 * q[4] + 4.0 * q[5] - q[6];
    f[3][2] = -q[3] + 3.0 * q[4] - 3.0 * q[5] + q[6];
}

// WH7:557-564 weno7_beta (expanded integer quadratic forms, 240 x Jiang-Shu)
__device__ __forceinline__ void beta7_poly(const double q[7], double b[4]) {
    const double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5], q6 = q[6];
    b[3] = 2107.0 * q3 * q3 - 9402.0 * q4 * q3 + 7042.0 * q3 * q5 - 1854.0 * q3 * q6 - 17246.0 * q4 * q5 -
           3882.0 * q5 * q6 + 11003.0 * q4 * q4 + 7043.0 * q5 * q5 + 547.0 * q6 * q6 + 4642.0 * q4 * q6;
    b[2] = 267.0 * q5 * q5 - 1642.0 * q5 * q4 + 1602.0 * q5 * q3 - 494.0 * q5 * q2 - 2522.0 * q3 * q2 +
           3443.0 * q3 * q3 + 547.0 * q2 * q2 - 5966.0 * q4 * q3 + 2843.0 * q4 * q4 + 1922.0 * q4 * q2;
    b[1] = 267.0 * q1 * q1 - 494.0 * q4 * q1 - 1642.0 * q2 * q1 + 1602.0 * q3 * q1 - 5966.0 * q3 * q2 +
           3443.0 * q3 * q3 - 2522.0 * q4 * q3 + 2843.0 * q2 * q2 + 547.0 * q4 * q4 + 1922.0 * q4 * q2;
    b[0] = 2107.0 * q3 * q3 + 7042.0 * q3 * q1 - 1854.0 * q3 * q0 - 9402.0 * q3 * q2 - 17246.0 * q2 * q1 +
           4642.0 * q2 * q0 + 547.0 * q0 * q0 - 3882.0 * q1 * q0 + 11003.0 * q2 * q2 + 7043.0 * q1 * q1;
}

// Z7:558-588 the same indicators as non-negative sums of squares
__device__ __forceinline__ void beta7_sos(const double q[7], double b[4]) {
    double f[4][3];
    delta_forms7(q, f);
    const double c0 = 20.0 / 3.0, c1 = 260.0, c2 = 781.0 / 3.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) b[k] = c0 * f[k][0] * f[k][0] + c1 * f[k][1] * f[k][1] + c2 * f[k][2] * f[k][2];
}

// W7:166-221 plateau test and the six closure inputs (torch semantics: scale
// channel (log10 + 16) / 16 without a remap); false = plateau
__device__ __forceinline__ bool features7(const double q[7], double f[6]) {
    double fm[4][3];
    delta_forms7(q, fm);
    const double c0 = 1.0 / 36.0, c1 = 13.0 / 12.0, c2 = 781.0 / 720.0;
    double delta[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) delta[k] = fabs(fm[k][0]) * c0 + fabs(fm[k][1]) * c1 + fabs(fm[k][2]) * c2;
    double dm = dmax(dmax(dmax(delta[0], delta[1]), delta[2]), delta[3]);
    double qs = fabs(q[0]);
#pragma unroll
    for (int m = 1; m < 7; ++m) qs = dmax(qs, fabs(q[m]));
    qs = dmax(qs, 1.0);
    if (dm <= 1.0e-13 * qs) return false;
    double inv = 1.0 / dmax(dm, 1.0e-15);
#pragma unroll
    for (int k = 0; k < 4; ++k) f[k] = delta[k] * inv;
    double g = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        double d2 = q[i] - 2.0 * q[i + 1] + q[i + 2];
        double den = fabs(q[i + 1] - q[i]) + fabs(q[i + 2] - q[i + 1]) + 1.0e-15;
        double r = fabs(d2) / den;
        g = i == 0 ? r : dmax(g, r);
    }
    f[4] = dmin(dmax(g, 0.0), 1.0);
    double rel = dmax(dm / qs, 1.0e-30);
    f[5] = dmin(dmax((log10(rel) + 16.0) / 16.0, 0.0), 1.0);
    return true;
}

// features7 for the FP32 closure: the plateau test in FP64, the six inputs in FP32
// (v_rcp_f32 quotients, v_log_f32 for the scale channel), as the 2-D FP32 closure
// computes its features; false = plateau (f untouched)
__device__ __forceinline__ bool features7_32(const double q[7], float f[6]) {
    double fm[4][3];
    delta_forms7(q, fm);
    const double c0 = 1.0 / 36.0, c1 = 13.0 / 12.0, c2 = 781.0 / 720.0;
    double delta[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) delta[k] = fabs(fm[k][0]) * c0 + fabs(fm[k][1]) * c1 + fabs(fm[k][2]) * c2;
    double dm = dmax(dmax(dmax(delta[0], delta[1]), delta[2]), delta[3]);
    double qs = fabs(q[0]);
#pragma unroll
    for (int m = 1; m < 7; ++m) qs = dmax(qs, fabs(q[m]));
    qs = dmax(qs, 1.0);
    if (dm <= 1.0e-13 * qs) return false;
    // not a plateau: dm > 1e-13 qs >= 1e-13, above the 1e-15 and 1e-30 floors
    const float inv = rcp_fast((float)dm);
#pragma unroll
    for (int k = 0; k < 4; ++k) f[k] = (float)delta[k] * inv;
    float g = 0.0f;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        double d2 = q[i] - 2.0 * q[i + 1] + q[i + 2];
        double den = fabs(q[i + 1] - q[i]) + fabs(q[i + 2] - q[i + 1]) + 1.0e-15;
        float r = (float)fabs(d2) * rcp_fast((float)den);
        g = i == 0 ? r : fmax_(g, r);
    }
    f[4] = fminf(g, 1.0f);
    const float rel = (float)dm * rcp_fast((float)qs);
    // (log10(rel) + 16) / 16 = log2(rel) log10(2) / 16 + 1, clipped to [0, 1]
    const float scale = fmaf(__builtin_amdgcn_logf(rel), (float)(0.30102999566398120 / 16.0), 1.0f);
    f[5] = fminf(1.0f, fmax_(0.0f, scale));
    return true;
}

// ----------------------------------------------------------- the closure
// W7:301-311: three swish layers, softmax(6 tanh(z / 6)) over 4 outputs; the
// reflection average 0.5 [M(x) + P4 M(P6 x)] with P6 reversing the deltas.
__device__ __forceinline__ void softmax4_64(const double z[4], double r[4]) {
    double b[4];
#pragma unroll
    for (int o = 0; o < 4; ++o) b[o] = 6.0 * tanh(z[o] / 6.0);
    double m = dmax(dmax(dmax(b[0], b[1]), b[2]), b[3]);
    double e[4], s = 0.0;
#pragma unroll
    for (int o = 0; o < 4; ++o) { e[o] = exp(b[o] - m); s = s + e[o]; }
    double inv = 1.0 / s;
#pragma unroll
    for (int o = 0; o < 4; ++o) r[o] = e[o] * inv;
}

__device__ __forceinline__ void softmax4_32(const float z[4], float r[4]) {
    const float sixth = 1.0f / 6.0f;
    float b[4];
#pragma unroll
    for (int o = 0; o < 4; ++o) b[o] = 6.0f * tanh_fast(z[o] * sixth);
    float m = fmax_(fmax_(fmax_(b[0], b[1]), b[2]), b[3]);
    float e[4], s = 0.0f;
#pragma unroll
    for (int o = 0; o < 4; ++o) { e[o] = exp_fast(b[o] - m); s = s + e[o]; }
    float inv = rcp_fast(s);
#pragma unroll
    for (int o = 0; o < 4; ++o) r[o] = e[o] * inv;
}

// one dense layer for a pair of inputs, one neuron's weights per scalar fetch
template <int K, int N, class T, class ACT>
__device__ __forceinline__ void dense7_pair(const T *xa, const T *xb, const T *w, int base, T *ha, T *hb, ACT act) {
#pragma unroll
    for (int o = 0; o < N; ++o) {
        kconst_ptr<T> c = kparams(w + base + o * (K + 1));
        T za = c[0], zb = c[0];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            za = za + xa[k] * c[1 + k];
            zb = zb + xb[k] * c[1 + k];
        }
        ha[o] = act(za);
        hb[o] = act(zb);
    }
}

// ratios of the reflection-averaged closure for the six inputs f (FP64 or FP32)
template <int SCHEME>
__device__ __forceinline__ void sym_ratios7(const double f[6], const typename Mlp7Sweep<SCHEME>::type &P,
                                            double r[4]) {
    if constexpr (SCHEME == SCHEME_SR64) {
        double xa[6] = {f[0], f[1], f[2], f[3], f[4], f[5]};
        double xb[6] = {f[3], f[2], f[1], f[0], f[4], f[5]};
        double a1[24], b1[24], a2[16], b2[16], a3[16], b3[16], za[4], zb[4], ra[4], rb[4];
        auto sw = [](double z) { return swish64(z); };
        auto id = [](double z) { return z; };
        dense7_pair<6, 24>(xa, xb, P.p, D7_L1, a1, b1, sw);
        dense7_pair<24, 16>(a1, b1, P.p, D7_L2, a2, b2, sw);
        dense7_pair<16, 16>(a2, b2, P.p, D7_L3, a3, b3, sw);
        dense7_pair<16, 4>(a3, b3, P.p, D7_L4, za, zb, id);
        softmax4_64(za, ra);
        softmax4_64(zb, rb);
#pragma unroll
        for (int k = 0; k < 4; ++k) r[k] = 0.5 * (ra[k] + rb[3 - k]);
    } else {
        float xa[6] = {(float)f[0], (float)f[1], (float)f[2], (float)f[3], (float)f[4], (float)f[5]};
        float xb[6] = {xa[3], xa[2], xa[1], xa[0], xa[4], xa[5]};
        float a1[24], b1[24], a2[16], b2[16], a3[16], b3[16], za[4], zb[4], ra[4], rb[4];
        auto sw = [](float z) { return swish32(z); };
        auto id = [](float z) { return z; };
        dense7_pair<6, 24>(xa, xb, P.p, D7_L1, a1, b1, sw);
        dense7_pair<24, 16>(a1, b1, P.p, D7_L2, a2, b2, sw);
        dense7_pair<16, 16>(a2, b2, P.p, D7_L3, a3, b3, sw);
        dense7_pair<16, 4>(a3, b3, P.p, D7_L4, za, zb, id);
        softmax4_32(za, ra);
        softmax4_32(zb, rb);
#pragma unroll
        for (int k = 0; k < 4; ++k) r[k] = 0.5 * ((double)ra[k] + (double)rb[3 - k]);
    }
}

// BP:27-33 beta = 4 r, ones on a plateau.  `reversed` evaluates the mirrored
// stencil and mirrors the result (BP:185-200, the lr = 2 side); with the
// reflection-averaged closure this equals the direct evaluation up to rounding.
// Returns false for a plateau.  With `fallback`, a non-finite or out-of-range
// ratio (outside [0, 1]) sets `invalid`.
template <int SCHEME>
__device__ __forceinline__ bool closure_beta7(const double v[7], bool reversed,
                                              const typename Mlp7Sweep<SCHEME>::type &P, double beta[4],
                                              bool &invalid) {
    double q[7], f[6], r[4];
#pragma unroll
    for (int m = 0; m < 7; ++m) q[m] = reversed ? v[6 - m] : v[m];
    invalid = false;
    if (!features7(q, f)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) beta[k] = 1.0;
        return false;
    }
    sym_ratios7<SCHEME>(f, P, r);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double rk = reversed ? r[3 - k] : r[k];
        invalid = invalid || !(rk >= 0.0 && rk <= 1.0);
        beta[k] = 4.0 * rk;
    }
    return true;
}

// ------------------------------------ the FP32 closure on the matrix cores
// The fused sweep evaluates the FP32 6-24-16-16-4 closure of a whole wave with
// v_mfma_f32_4x4x1f32 (16 blocks of 4 lanes): the B operand is each lane's own input
// (so a lane's stencil stays in its lane), the A operand four neurons' weights, read from
// one register for 16 instructions through the A broadcast (CBSZ 4, ABID = entry mod 16),
// and the 4 x 4 result of block b leaves neuron i of lane 4b + j's stencil in register i
// of lane 4b + j.  Every layer is laid out as groups of four neurons; the direct and the
// reflected evaluation are separate groups (reflected layer-1 weights permuted, so both
// read the direct features), and each group sums its bias first (B = 1) and then its
// inputs in ascending order, the order of the VALU closure.  Hidden activations and the
// head use the folded constants of fold7_fp32 (the swish z' / (1 + 2^z'), the bounded
// head 2^(6L - 12L / (1 + 2^u))).  Measured on MI355X against the VALU pair: 1.48x
// (9.6k against 14.2k SIMD cycles per wave).
typedef float v4f32 __attribute__((ext_vector_type(4)));
constexpr int W7M_L1 = 0, W7M_L2 = 84, W7M_L3 = 284, W7M_L4 = 420;
constexpr int W7M_ENTRIES = 454;                       // A entries (4 floats each)
constexpr int W7M_VGPRS = (W7M_ENTRIES + 15) / 16;     // 16 entries per A register
constexpr int W7M_WORDS = W7M_VGPRS * 64;              // table floats: register v, lane l at v * 64 + l

// host: the activation constants folded into the neuron-major FP32 block (as
// fold_fp32_closure for the 5-10-6-6-3 closure): layer-1 weights and all hidden biases
// scaled by -log2(e), output weights by -1/3 and output biases by log2(e)/3
inline void fold7_fp32(float *dev) {
    const int start[4] = {D7_L1, D7_L2, D7_L3, D7_L4}, K[4] = {6, 24, 16, 16}, N[4] = {24, 16, 16, 4};
    for (int l = 0; l < 4; ++l)
        for (int n = 0; n < N[l]; ++n) {
            float *run = dev + start[l] + n * (K[l] + 1);
            const double wscale = l == 0 ? -LOG2E : (l == 3 ? -1.0 / 3.0 : 1.0);
            const double bscale = l == 3 ? LOG2E / 3.0 : -LOG2E;
            run[0] = (float)((double)run[0] * bscale);
            for (int k = 1; k <= K[l]; ++k) run[k] = (float)((double)run[k] * wscale);
        }
}

// host: the A table from the folded neuron-major block.  Entry m of four floats (neuron
// p of its group) sits at T[m * 4 + p]; register v of lane l holds T[v * 64 + l].
// Layer 1: groups 0..5 direct neurons, 6..11 reflected (input k reads feature P6(k));
// layers 2, 3: groups 0..3 direct, 4..7 reflected; layer 4: group 0 direct, 1 reflected.
inline void mfma7_table(const float *dev, float *T) {
    for (int i = 0; i < W7M_WORDS; ++i) T[i] = 0.0f;
    const int reflect[6] = {3, 2, 1, 0, 4, 5};
    for (int G = 0; G < 12; ++G)
        for (int p = 0; p < 4; ++p) {
            const int n = 4 * G + p, o = n % 24;
            const bool refl = n >= 24;
            T[(W7M_L1 + G * 7) * 4 + p] = dev[D7_L1 + o * 7];
            for (int k = 0; k < 6; ++k)
                T[(W7M_L1 + G * 7 + 1 + k) * 4 + p] = dev[D7_L1 + o * 7 + 1 + (refl ? reflect[k] : k)];
        }
    auto hidden = [&](int m0, int start, int K) {
        for (int G = 0; G < 8; ++G)
            for (int p = 0; p < 4; ++p) {
                const int o = (4 * G + p) % 16;
                T[(m0 + G * (K + 1)) * 4 + p] = dev[start + o * (K + 1)];
                for (int k = 0; k < K; ++k) T[(m0 + G * (K + 1) + 1 + k) * 4 + p] = dev[start + o * (K + 1) + 1 + k];
            }
    };
    hidden(W7M_L2, D7_L2, 24);
    hidden(W7M_L3, D7_L3, 16);
    for (int G = 0; G < 2; ++G)
        for (int p = 0; p < 4; ++p) {
            T[(W7M_L4 + G * 17) * 4 + p] = dev[D7_L4 + p * 17];
            for (int k = 0; k < 16; ++k) T[(W7M_L4 + G * 17 + 1 + k) * 4 + p] = dev[D7_L4 + p * 17 + 1 + k];
        }
}

// A registers of the closure: the table in LDS, register v of this lane
struct Mfma7A {
    const float *lds;
    int lane;
    __device__ __forceinline__ float operator()(int v) const { return lds[v * 64 + lane]; }
};

template <int M>
__device__ __forceinline__ v4f32 mfma7(float a, float b, v4f32 c) {
    return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 4, M & 15, 0);
}
// one input step I of every group (independent accumulators back to back)
template <int M0, int K, int NG, bool SPLIT, int I, int G>
__device__ __forceinline__ void mfma7_kstep(const Mfma7A &wa, const float *x, v4f32 *acc) {
    if constexpr (G < NG) {
        constexpr int M = M0 + G * (K + 1) + 1 + I;
        acc[G] = mfma7<M>(wa(M >> 4), x[(SPLIT && G >= NG / 2 ? K : 0) + I], acc[G]);
        mfma7_kstep<M0, K, NG, SPLIT, I, G + 1>(wa, x, acc);
    }
}
template <int M0, int K, int NG, bool SPLIT, int I>
__device__ __forceinline__ void mfma7_kloop(const Mfma7A &wa, const float *x, v4f32 *acc) {
    if constexpr (I < K) {
        mfma7_kstep<M0, K, NG, SPLIT, I, 0>(wa, x, acc);
        mfma7_kloop<M0, K, NG, SPLIT, I + 1>(wa, x, acc);
    }
}
template <int M0, int K, int NG, int G>
__device__ __forceinline__ void mfma7_bias(const Mfma7A &wa, v4f32 *acc) {
    if constexpr (G < NG) {
        constexpr int M = M0 + G * (K + 1);
        const v4f32 z = {0.0f, 0.0f, 0.0f, 0.0f};
        acc[G] = mfma7<M>(wa(M >> 4), 1.0f, z);
        mfma7_bias<M0, K, NG, G + 1>(wa, acc);
    }
}
// NG groups of four neurons over K inputs (SPLIT: the second half of the groups reads
// inputs K..2K-1), swish on the hidden layers
template <int M0, int K, int NG, bool SPLIT, bool ACT>
__device__ __forceinline__ void mfma7_layer(const Mfma7A &wa, const float *x, float *h) {
    v4f32 acc[NG];
    mfma7_bias<M0, K, NG, 0>(wa, acc);
    mfma7_kloop<M0, K, NG, SPLIT, 0>(wa, x, acc);
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int p = 0; p < 4; ++p) h[4 * g + p] = ACT ? swish_folded(acc[g][p]) : acc[g][p];
}

__device__ __forceinline__ void head4_folded(const float u[4], float r[4]) {
    const float c6 = (float)(6.0 * LOG2E), c12 = (float)(12.0 * LOG2E);
    float e[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float q = rcp_fast(1.0f + __builtin_amdgcn_exp2f(u[k]));
        e[k] = __builtin_amdgcn_exp2f(fmaf(-c12, q, c6));
    }
    float inv = rcp_fast((e[0] + e[1]) + (e[2] + e[3]));
#pragma unroll
    for (int k = 0; k < 4; ++k) r[k] = e[k] * inv;
}

// ra = M(x), rb = M(P6 x) for every lane of the wave (call with all 64 lanes active)
__device__ __forceinline__ void closure7_mfma(const float x[6], const Mfma7A &wa, float ra[4], float rb[4]) {
    float h1[48], h2[32], h3[32], z4[8];
    mfma7_layer<W7M_L1, 6, 12, false, true>(wa, x, h1);
    mfma7_layer<W7M_L2, 24, 8, true, true>(wa, h1, h2);
    mfma7_layer<W7M_L3, 16, 8, true, true>(wa, h2, h3);
    mfma7_layer<W7M_L4, 16, 2, true, false>(wa, h3, z4);
    head4_folded(z4, ra);
    head4_folded(z4 + 4, rb);
}

// closure_beta7 for the FP32 closure of a whole wave (every lane calls it with all 64
// lanes active; lanes on a plateau get ones).  The matrix cores run when any lane of the
// wave needs the closure.
__device__ __forceinline__ bool closure_beta7_wave(const double v[7], bool reversed, const Mfma7A &wa,
                                                   double beta[4], bool &invalid) {
    double q[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) q[m] = reversed ? v[6 - m] : v[m];
    float f[6];
    const bool active = features7_32(q, f);
    float x[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) x[k] = active ? f[k] : 0.0f;
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    if (__ballot(active) != 0ull) {   // wave-uniform branch: EXEC stays full
        float ra[4], rb[4];
        closure7_mfma(x, wa, ra, rb);
#pragma unroll
        for (int k = 0; k < 4; ++k) r[k] = 0.5 * ((double)ra[k] + (double)rb[3 - k]);
    }
    invalid = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double rk = reversed ? r[3 - k] : r[k];
        invalid = invalid || (active && !(rk >= 0.0 && rk <= 1.0));
        beta[k] = active ? 4.0 * rk : 1.0;
    }
    return active;
}

// EXT:64-79 weights from beta: a = d / (beta + 1e-12)^2, normalised (+ ENO cutoff)
__device__ __forceinline__ double combine_beta7(const double s[4], const double d[4], const double beta[4], int eno) {
    double a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double ib = 1.0 / (beta[k] + 1.0e-12);
        a[k] = d[k] * ib * ib;
    }
    double inv = 1.0 / dmax(a[0] + a[1] + a[2] + a[3], 1.0e-300);
    double w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = a[k] * inv;
    if (eno) {
        double p[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) p[k] = w[k] <= 4.0e-7 ? 0.0 : 1.0;
        double ic = 1.0 / dmax(p[0] * w[0] + p[1] * w[1] + p[2] * w[2] + p[3] * w[3], 1.0e-300);
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] = p[k] * w[k] * ic;
    }
    return w[0] * s[0] + w[1] * s[1] + w[2] * s[2] + w[3] * s[3];
}

// The FP32 closure's weights: a = d / (beta + 1e-12)^2 with the reciprocal from v_rcp_f64
// and two Newton steps (within an ulp of the quotient), then combine_beta7's
// normalisation and ENO cutoff from the same a
__device__ __forceinline__ void weights_a7_fast(const double d[4], const double beta[4], double a[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double ib = rcp_newton(beta[k] + 1.0e-12);
        a[k] = d[k] * ib * ib;
    }
}

__device__ __forceinline__ double combine_a7_fast(const double s[4], const double a[4], int eno) {
    double inv = rcp_newton(dmax(a[0] + a[1] + a[2] + a[3], 1.0e-300));
    double w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = a[k] * inv;
    if (eno) {
        double p[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) p[k] = w[k] <= 4.0e-7 ? 0.0 : 1.0;
        double ic = rcp_newton(dmax(p[0] * w[0] + p[1] * w[1] + p[2] * w[2] + p[3] * w[3], 1.0e-300));
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] = p[k] * w[k] * ic;
    }
    return w[0] * s[0] + w[1] * s[1] + w[2] * s[2] + w[3] * s[3];
}

// WH7:741-776 (JS): a = d / (1e-6 + beta)^2 with the expanded indicators
__device__ __forceinline__ double combine_js7(const double s[4], const double d[4], const double b[4]) {
    double a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double r = 1.0 / (1.0e-6 + b[k]);
        a[k] = d[k] * r * r;
    }
    double asum = a[0] + a[1] + a[2] + a[3];
    return (a[0] * s[0] + a[1] * s[1] + a[2] * s[2] + a[3] * s[3]) / asum;
}

// Z7:780-809 (WENO-Z, p = 3, tau_7 of Castro-Costa-Don, eps = 240 h^3)
__device__ __forceinline__ double combine_z7(const double s[4], const double d[4], const double b[4], double h) {
    double tau7 = fabs(-b[0] - 3.0 * b[1] + 3.0 * b[2] + b[3]);
    double eps_z = 240.0 * h * h * h;
    double a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double r = tau7 / (b[k] + eps_z);
        a[k] = d[k] * (1.0 + r * r * r);
    }
    double asum = a[0] + a[1] + a[2] + a[3];
    return (a[0] * s[0] + a[1] * s[1] + a[2] * s[2] + a[3] * s[3]) / asum;
}

// Guard of the neural closure (WSR_FLAG_Z_FALLBACK).  The candidate the closure
// weights most, k* = argmax d_k / (beta_k + 1e-12)^2, must not be one that the WENO7-Z
// smoothness indicators bz mark as crossing a discontinuity:
// bz[k*] > Z7_GUARD_RATIO min_k bz_k + 240 h^3 (the WENO7-Z epsilon) rejects the closure
// output, and so does a ratio that is not finite or outside [0, 1] (`invalid`).  A state
// (the four components of one face side, or of one Gauss point of a side) with a
// rejected component is reconstructed with the WENO7-Z weights in all four components:
// the kernels decide per group of four lanes.
constexpr double Z7_GUARD_RATIO = 50.0;

// z7_guard_trips from the unnormalised weights a of weights_a7_fast (the FP32 closure)
__device__ __forceinline__ bool z7_guard_trips_a(const double a[4], const double bz[4], double h, bool invalid) {
    double best = -1.0, bk = bz[0];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool up = a[k] > best;
        best = up ? a[k] : best;
        bk = up ? bz[k] : bk;
    }
    double bmin = dmin(dmin(bz[0], bz[1]), dmin(bz[2], bz[3]));
    return invalid || bk > Z7_GUARD_RATIO * bmin + 240.0 * h * h * h;
}

__device__ __forceinline__ bool z7_guard_trips(const double d[4], const double beta[4], const double bz[4],
                                               double h, bool invalid) {
    // the dominant candidate's indicator is carried by selects (an index into bz would
    // put bz in scratch memory)
    double best = -1.0, bk = bz[0];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double ib = 1.0 / (beta[k] + 1.0e-12);
        double a = d[k] * ib * ib;
        const bool up = a > best;
        best = up ? a : best;
        bk = up ? bz[k] : bk;
    }
    double bmin = dmin(dmin(bz[0], bz[1]), dmin(bz[2], bz[3]));
    return invalid || bk > Z7_GUARD_RATIO * bmin + 240.0 * h * h * h;
}

// WENO7-Z value of one component at head HEAD (the fallback)
template <int HEAD>
__device__ __forceinline__ double z7_value(const double q[7], double h) {
    double s[4], d[4], bz[4];
    candidates7<HEAD>(q, s);
    optimal7<HEAD>(d);
    beta7_sos(q, bz);
    return combine_z7(s, d, bz, h);
}

// Normal face value of one characteristic component: lr 1 (HEAD 1) or lr 2 (HEAD 2).
// With z_fallback, `trip` reports whether the guard rejects the closure's value (the
// value returned is still the closure's).
template <int SCHEME, int HEAD>
__device__ __forceinline__ double recon7_normal(const double q[7], const typename Mlp7Sweep<SCHEME>::type &P,
                                                int eno, double h, bool z_fallback, bool &closure, bool &trip) {
    double s[4], d[4], b[4];
    candidates7<HEAD>(q, s);
    optimal7<HEAD>(d);
    closure = false;
    trip = false;
    if constexpr (SCHEME == SCHEME_JS) {
        beta7_poly(q, b);
        return combine_js7(s, d, b);
    } else if constexpr (SCHEME == SCHEME_Z) {
        beta7_sos(q, b);
        return combine_z7(s, d, b, h);
    } else {
        bool invalid;
        closure = closure_beta7<SCHEME>(q, HEAD == 2, P, b, invalid);
        if (z_fallback && closure) {
            double bz[4];
            beta7_sos(q, bz);
            trip = z7_guard_trips(d, b, bz, h, invalid);
        }
        return combine_beta7(s, d, b, eno);
    }
}

// recon7_normal for the FP32 closure on the matrix cores: side s of the lane (0: lr 1,
// head 1; 1: lr 2, head 2, the mirrored stencil) chosen by selects, so that every lane
// reaches the closure together.
__device__ __forceinline__ double recon7_normal_mfma(const double q[7], int s, const Mfma7A &wa, int eno, double h,
                                                     bool z_fallback, bool &closure, bool &trip) {
    double s1[4], s2[4], d1[4], d2[4], sv[4], d[4], b[4];
    candidates7<1>(q, s1);
    candidates7<2>(q, s2);
    optimal7<1>(d1);
    optimal7<2>(d2);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        sv[k] = s ? s2[k] : s1[k];
        d[k] = s ? d2[k] : d1[k];
    }
    bool invalid;
    closure = closure_beta7_wave(q, s == 1, wa, b, invalid);
    double a[4];
    weights_a7_fast(d, b, a);
    trip = false;
    if (z_fallback && closure) {
        double bz[4];
        beta7_sos(q, bz);
        trip = z7_guard_trips_a(a, bz, h, invalid);
    }
    return combine_a7_fast(sv, a, eno);
}

// recon7_gauss for the FP32 closure on the matrix cores (every lane of the wave calls it)
__device__ __forceinline__ void recon7_gauss_mfma(const double q[7], const Mfma7A &wa, int eno, double h,
                                                  bool z_fallback, double &g1, double &g2, bool &closure,
                                                  bool &trip1, bool &trip2) {
    double s1[4], s2[4], d1[4], d2[4], b[4];
    candidates7<4>(q, s1);
    candidates7<3>(q, s2);
    optimal7<4>(d1);
    optimal7<3>(d2);
    bool invalid;
    closure = closure_beta7_wave(q, false, wa, b, invalid);
    double a1[4], a2[4];   // weights_a7_fast for both points from one reciprocal
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double ib = rcp_newton(b[k] + 1.0e-12);
        a1[k] = d1[k] * ib * ib;
        a2[k] = d2[k] * ib * ib;
    }
    trip1 = false;
    trip2 = false;
    if (z_fallback && closure) {
        double bz[4];
        beta7_sos(q, bz);
        trip1 = z7_guard_trips_a(a1, bz, h, invalid);
        trip2 = z7_guard_trips_a(a2, bz, h, invalid);
    }
    g1 = combine_a7_fast(s1, a1, eno);
    g2 = combine_a7_fast(s2, a2, eno);
}

// Both transverse Gauss-point values (loca 1 = head 4, loca 2 = head 3) of one
// conservative component; one closure evaluation serves both points (the
// reference evaluates the mirrored stencil for loca 2, equal up to rounding).
// With z_fallback the guard judges each Gauss point on its own (its optimal weights
// pick the dominant candidate): trip1 / trip2 (the values are still the closure's).
template <int SCHEME>
__device__ __forceinline__ void recon7_gauss(const double q[7], const typename Mlp7Sweep<SCHEME>::type &P, int eno,
                                             double h, bool z_fallback, double &g1, double &g2, bool &closure,
                                             bool &trip1, bool &trip2) {
    double s1[4], s2[4], d1[4], d2[4], b[4];
    candidates7<4>(q, s1);
    candidates7<3>(q, s2);
    optimal7<4>(d1);
    optimal7<3>(d2);
    closure = false;
    trip1 = false;
    trip2 = false;
    if constexpr (SCHEME == SCHEME_JS) {
        beta7_poly(q, b);
        g1 = combine_js7(s1, d1, b);
        g2 = combine_js7(s2, d2, b);
    } else if constexpr (SCHEME == SCHEME_Z) {
        beta7_sos(q, b);
        g1 = combine_z7(s1, d1, b, h);
        g2 = combine_z7(s2, d2, b, h);
    } else {
        bool invalid;
        closure = closure_beta7<SCHEME>(q, false, P, b, invalid);
        if (z_fallback && closure) {
            double bz[4];
            beta7_sos(q, bz);
            trip1 = z7_guard_trips(d1, b, bz, h, invalid);
            trip2 = z7_guard_trips(d2, b, bz, h, invalid);
        }
        g1 = combine_beta7(s1, d1, b, eno);
        g2 = combine_beta7(s2, d2, b, eno);
    }
}

// WH7:256-342 evilin_state_2d; the flux is f(evilin state) (WH7:429-433)
template <int DIR>
__device__ void evilin_flux(const double ul0[4], const double ur0[4], double c, double gamma, double out[4]) {
    const double tiny = 1.0e-15;
    double wl0[4], wr0[4];
    con_to_pri(ul0, gamma, wl0);
    con_to_pri(ur0, gamma, wr0);
    const int n = DIR == 1 ? 1 : 2;
    double a_l = sqrt(gamma * wl0[3] / wl0[0]), a_r = sqrt(gamma * wr0[3] / wr0[0]);
    double smax = dmax(dmax(fabs(wl0[n]) + a_l, fabs(wr0[n]) + a_r), tiny);
    double dt_dx = c, dx_dt = 1.0 / dt_dx, local_cfl = c * smax;
    double fl[4], fr[4], qlw[4], wlw[4], flw[4], fgf[4], ulh[4], urh[4];
    pri_to_flux<DIR>(wl0, gamma, fl);
    pri_to_flux<DIR>(wr0, gamma, fr);
#pragma unroll
    for (int k = 0; k < 4; ++k) qlw[k] = 0.5 * (ul0[k] + ur0[k]) - 0.5 * dt_dx * (fr[k] - fl[k]);
    con_to_pri(qlw, gamma, wlw);
    pri_to_flux<DIR>(wlw, gamma, flw);
    double theta = (1.0 - local_cfl) / (1.0 + local_cfl);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double ffo = 0.25 * (fl[k] + 2.0 * flw[k] + fr[k] - dx_dt * (ur0[k] - ul0[k]));
        fgf[k] = theta * flw[k] + (1.0 - theta) * ffo;
        ulh[k] = ul0[k] - dt_dx * (fgf[k] - fl[k]);
        urh[k] = ur0[k] - dt_dx * (fr[k] - fgf[k]);
    }
    double wl[4], wr[4];
    con_to_pri(ulh, gamma, wl);
    con_to_pri(urh, gamma, wr);
    double rho_bar = dmax(0.5 * (wl[0] + wr[0]), tiny);
    double p_bar = dmax(0.5 * (wl[3] + wr[3]), tiny);
    double a_bar = dmax(sqrt(gamma * p_bar / rho_bar), tiny);
    double c1 = rho_bar * a_bar, c2 = rho_bar / a_bar;
    double un_le = wl[n], un_re = wr[n];
    double u_star = 0.5 * (un_le + un_re) - 0.5 * (wr[3] - wl[3]) / c1;
    double p_star = 0.5 * (wl[3] + wr[3]) - 0.5 * (un_re - un_le) * c1;
    double rho_star_l = dmax(wl[0] + (un_le - u_star) * c2, tiny);
    double rho_star_r = dmax(wr[0] + (u_star - un_re) * c2, tiny);
    p_star = dmax(p_star, tiny);
    double u_bar = 0.5 * (un_le + un_re);
    double w[4];
    if (u_bar - a_bar >= 0.0) {
        w[0] = wl[0]; w[1] = wl[1]; w[2] = wl[2]; w[3] = wl[3];
    } else if (u_bar >= 0.0) {
        w[0] = rho_star_l; w[1] = wl[1]; w[2] = wl[2]; w[3] = p_star; w[n] = u_star;
    } else if (u_bar + a_bar >= 0.0) {
        w[0] = rho_star_r; w[1] = wr[1]; w[2] = wr[2]; w[3] = p_star; w[n] = u_star;
    } else {
        w[0] = wr[0]; w[1] = wr[1]; w[2] = wr[2]; w[3] = wr[3];
    }
    // pri_to_con then con_to_pri (WH7:432) before the physical flux
    double e = 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2]) + w[3] / (gamma - 1.0);
    double qf[4] = {w[0], w[0] * w[1], w[0] * w[2], e}, wf[4];
    con_to_pri(qf, gamma, wf);
    pri_to_flux<DIR>(wf, gamma, out);
}

}  // namespace wsr
