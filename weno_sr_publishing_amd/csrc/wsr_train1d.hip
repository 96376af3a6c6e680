// wsr_train1d.hip -- closure training kernels of the 1-D advection path (gfx950).
//
// The reference trains the MLP with torch autograd through checkpointed
// 16-step chunks of the SSPRK3 rollout
// (code/snapshot/teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/fast_losses.py:412-648,
// rk3_advection.py:118-131).  Here the gradient is a hand-written adjoint:
//   k_train_forward : the rollout with the parameters read from device memory,
//                     storing the state at the start of every step (checkpoints);
//   k_train_adjoint : one workgroup per trajectory walks the steps backwards,
//                     recomputes the two intermediate stages of each step from
//                     its checkpoint, and propagates the adjoint through the three
//                     stages: per face the VJP of the reconstruction (wsr_adjoint.hpp)
//                     gives the stencil adjoints (gathered per cell, no atomics)
//                     and two parameter-gradient rows per MLP evaluation; the rows
//                     of a pass are reduced on the matrix cores (16x16x4 MFMA outer
//                     products, FP32 rows for the FP32 closure, FP64 for FP64) and
//                     the per-pass tiles accumulated in FP64.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/wsr.h"
#include "wsr_adjoint.hpp"

namespace wsr {

constexpr int T1_BLOCK = 256;
constexpr int T1_MAX_N = 1024;

template <int SCHEME> struct DevView;
template <> struct DevView<SCHEME_SR64> { using type = MlpDev64; using T = double; static constexpr int WAVES = 2; };
template <> struct DevView<SCHEME_SR32> { using type = MlpDev32; using T = float; static constexpr int WAVES = 4; };

template <int SCHEME>
__global__ __launch_bounds__(T1_BLOCK) void k_train_forward(double *__restrict__ u, const double *__restrict__ vel,
                                                            const double *__restrict__ dts, int nsteps, int n,
                                                            double dxinv, double *__restrict__ ckpt,
                                                            typename DevView<SCHEME>::type P) {
    extern __shared__ double lds[];
    double *s0 = lds, *s1 = lds + n, *fl = lds + 2 * n;
    const int b = blockIdx.x;
    const double a = vel[b];
    double *ub = u + (size_t)b * n;
    for (int i = threadIdx.x; i < n; i += T1_BLOCK) s0[i] = ub[i];
    __syncthreads();
    for (int step = 0; step < nsteps; ++step) {
        const double dt = dts[step];
        double *c = ckpt + ((size_t)b * (nsteps + 1) + step) * n;
        for (int i = threadIdx.x; i < n; i += T1_BLOCK) c[i] = s0[i];
        for (int stage = 0; stage < 3; ++stage) {
            const double *src = stage == 0 ? s0 : s1;
            for (int i = threadIdx.x; i < n; i += T1_BLOCK) fl[i] = a * face_signed<SCHEME>(src, n, i, a, P);
            __syncthreads();
            for (int i = threadIdx.x; i < n; i += T1_BLOCK) {
                double rhs = -(fl[i] - fl[i == 0 ? n - 1 : i - 1]) * dxinv;
                double v;
                if (stage == 0) v = s0[i] + dt * rhs;
                else if (stage == 1) v = 0.75 * s0[i] + 0.25 * (s1[i] + dt * rhs);
                else v = s0[i] / 3.0 + (2.0 / 3.0) * (s1[i] + dt * rhs);
                if (stage < 2) s1[i] = v;
                else s0[i] = v;
            }
            __syncthreads();
        }
    }
    double *c = ckpt + ((size_t)b * (nsteps + 1) + nsteps) * n;
    for (int i = threadIdx.x; i < n; i += T1_BLOCK) {
        ub[i] = s0[i];
        c[i] = s0[i];
    }
}

// dst[i] = a * face(src)_i for all faces (forward recompute inside the adjoint)
template <int SCHEME, class MP>
__device__ __forceinline__ void faces_into(const double *src, double *dst, int n, double a, const MP &P, int nthreads) {
    for (int i = threadIdx.x; i < n; i += nthreads) dst[i] = a * face_signed<SCHEME>(src, n, i, a, P);
}

// Parameter gradients as matrix-core outer products.  d(loss)/d(param) is a sum
// over gradient rows (one per MLP evaluation) of row[in] * row[out]; grouped by
// layer these are the products  IN^T OUT  of the rows' input activations and
// output adjoints, accumulated by 16 x 16 x 4 MFMA tiles over 4 rows at a time:
//   tile 0: i 0..5  = (x, 1) x j 0..9   = z1bar   (layer 1 weights, bias)
//           i 6..12 = (h2, 1) x j 10..15 = z3bar  (layer 3)
//   tile 1: i 0..10 = (h1, 1) x j 0..5  = z2bar   (layer 2)
//   tile 2: i 0..6  = (h3, 1) x j 0..2  = z4bar   (layer 4)
// (the other entries of tile 0 are cross terms and are never read).
constexpr int NTILE = 3;

__device__ __forceinline__ int tile_in_slot(int t, int i) {
    if (t == 0) {
        if (i < 5) return R_X + i;
        if (i == 5) return R_ONE;
        if (i < 12) return R_H2 + (i - 6);
        return i == 12 ? R_ONE : -1;
    }
    if (t == 1) return i < 10 ? R_H1 + i : (i == 10 ? R_ONE : -1);
    return i < 6 ? R_H3 + i : (i == 6 ? R_ONE : -1);
}
__device__ __forceinline__ int tile_out_slot(int t, int j) {
    if (t == 0) return j < 10 ? R_Z1 + j : R_Z3 + (j - 10);
    if (t == 1) return j < 6 ? R_Z2 + j : -1;
    return j < 3 ? R_Z4 + j : -1;
}
// (tile, i, j) of parameter p (inverse of param_row_slots)
__device__ __forceinline__ void param_tile(int p, int &t, int &i, int &j) {
    if (p < 50) { t = 0; i = p / 10; j = p % 10; }
    else if (p < 60) { t = 0; i = 5; j = p - 50; }
    else if (p < 120) { t = 1; i = (p - 60) / 6; j = (p - 60) % 6; }
    else if (p < 126) { t = 1; i = 10; j = p - 120; }
    else if (p < 162) { t = 0; i = 6 + (p - 126) / 6; j = 10 + (p - 126) % 6; }
    else if (p < 168) { t = 0; i = 12; j = 10 + (p - 162); }
    else if (p < 186) { t = 2; i = (p - 168) / 3; j = (p - 168) % 3; }
    else { t = 2; i = 6; j = p - 186; }
}

typedef float mf4 __attribute__((ext_vector_type(4)));
typedef double md4 __attribute__((ext_vector_type(4)));

template <typename T> struct Tile;
template <> struct Tile<float> {
    using V = mf4;
    static __device__ __forceinline__ V mfma(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // accumulator register r of lane l holds D[4 (l / 16) + r][l % 16]
    static __device__ __forceinline__ int row_of(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <> struct Tile<double> {
    using V = md4;
    static __device__ __forceinline__ V mfma(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // accumulator register r of lane l holds D[(l / 16) + 4 r][l % 16]
    static __device__ __forceinline__ int row_of(int lane, int r) { return (lane >> 4) + 4 * r; }
};

// VJP of R(X)_j = -(F_j - F_{j-1}) dxinv with cotangent g: target += R_X^T g and the
// parameter-gradient rows of the pass folded into acc (FP64, lane layout of Tile<T>).
template <int SCHEME, int NT, typename T, class MP>
__device__ void vjp_rhs(const double *X, const double *g, double *target, double *adjq, T *rows, int *used, int n,
                        double a, double dxinv, const MP &P, double (&acc)[NTILE][4]) {
    const int off = a > 0.0 ? 2 : 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // operand slots of this lane: A[i][k] = row_k[in(i)], B[k][j] = row_k[out(j)], i = j = lane % 16
    int sin[NTILE], sout[NTILE];
#pragma unroll
    for (int t = 0; t < NTILE; ++t) { sin[t] = tile_in_slot(t, lane & 15); sout[t] = tile_out_slot(t, lane & 15); }
    for (int base = 0; base < n; base += NT) {
        const int i = base + threadIdx.x;
        if (i < n) {
            double vbar = a * ((g[i + 1 == n ? 0 : i + 1] - g[i]) * dxinv);
            double q[5], qb[5], value;
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                int k = i - off + m;
                k = k < 0 ? k + n : (k >= n ? k - n : k);
                q[m] = X[k];
            }
            T *rd = rows + (size_t)(2 * threadIdx.x) * ROW;
            bool closure = face_vjp<SCHEME>(q, a > 0.0, vbar, P, qb, rd, rd + ROW, value);
            used[threadIdx.x] = closure ? 1 : 0;
#pragma unroll
            for (int m = 0; m < 5; ++m) adjq[i * 5 + m] = qb[m];
        } else {
            used[threadIdx.x] = 0;
        }
        __syncthreads();
        // this wave's 64 faces = 128 rows, 4 rows per MFMA step (row k of a step: lane / 16)
        typename Tile<T>::V d[NTILE];
#pragma unroll
        for (int t = 0; t < NTILE; ++t) d[t] = typename Tile<T>::V{T(0), T(0), T(0), T(0)};
        for (int kc = 0; kc < 32; ++kc) {
            const int r = wave * 128 + kc * 4 + (lane >> 4);
            const T *row = rows + (size_t)r * ROW;
            const bool u = used[r >> 1] != 0;
#pragma unroll
            for (int t = 0; t < NTILE; ++t) {
                T av = (u && sin[t] >= 0) ? row[sin[t]] : T(0);
                T bv = (u && sout[t] >= 0) ? row[sout[t]] : T(0);
                d[t] = Tile<T>::mfma(av, bv, d[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < NTILE; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[t][q] += (double)d[t][q];
        __syncthreads();
    }
    // gather the stencil adjoints per cell: cell j is entry m of face j + off - m
    for (int j = threadIdx.x; j < n; j += NT) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            int f = j + off - m;
            f = f < 0 ? f + n : (f >= n ? f - n : f);
            s += adjq[f * 5 + m];
        }
        target[j] += s;
    }
    __syncthreads();
}

template <int SCHEME>
__global__ __launch_bounds__(T1_BLOCK) void k_train_adjoint(const double *__restrict__ ckpt,
                                                            const double *__restrict__ lam_T,
                                                            const double *__restrict__ vel,
                                                            const double *__restrict__ dts, int nsteps, int n,
                                                            double dxinv, typename DevView<SCHEME>::type P,
                                                            double *__restrict__ grad, double *__restrict__ lam0) {
    using T = typename DevView<SCHEME>::T;
    constexpr int NT = 64 * DevView<SCHEME>::WAVES;
    extern __shared__ double lds[];
    double *u = lds, *u1 = u + n, *u2 = u1 + n, *lam = u2 + n, *lu = lam + n, *lx = lu + n, *g = lx + n;
    double *fl = g + n, *adjq = fl + n;
    int *used = (int *)(adjq + 5 * n);
    T *rows = (T *)(used + NT);
    const int b = blockIdx.x;
    const double a = vel[b];
    static_assert(NT % 64 == 0 && NT / 64 * 128 == 2 * NT, "one 128-row MFMA pass per wave");
    double acc[NTILE][4];
    for (int t = 0; t < NTILE; ++t)
        for (int q = 0; q < 4; ++q) acc[t][q] = 0.0;
    for (int j = threadIdx.x; j < n; j += NT) lam[j] = lam_T[(size_t)b * n + j];
    __syncthreads();
    for (int step = nsteps - 1; step >= 0; --step) {
        const double dt = dts[step];
        const double *c = ckpt + ((size_t)b * (nsteps + 1) + step) * n;
        for (int j = threadIdx.x; j < n; j += NT) u[j] = c[j];
        __syncthreads();
        // recompute the stage states u1, u2 of this step
        faces_into<SCHEME>(u, fl, n, a, P, NT);
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += NT) u1[j] = u[j] + dt * (-(fl[j] - fl[j == 0 ? n - 1 : j - 1]) * dxinv);
        __syncthreads();
        faces_into<SCHEME>(u1, fl, n, a, P, NT);
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += NT)
            u2[j] = 0.75 * u[j] + 0.25 * (u1[j] + dt * (-(fl[j] - fl[j == 0 ? n - 1 : j - 1]) * dxinv));
        // stage 3: u' = u/3 + 2/3 (u2 + dt R(u2))
        for (int j = threadIdx.x; j < n; j += NT) {
            lu[j] = lam[j] / 3.0;
            lx[j] = (2.0 / 3.0) * lam[j];
            g[j] = (2.0 / 3.0) * dt * lam[j];
        }
        __syncthreads();
        vjp_rhs<SCHEME, NT>(u2, g, lx, adjq, rows, used, n, a, dxinv, P, acc);
        // stage 2: u2 = 3/4 u + 1/4 (u1 + dt R(u1))
        for (int j = threadIdx.x; j < n; j += NT) {
            lu[j] += 0.75 * lx[j];
            g[j] = 0.25 * dt * lx[j];
            lx[j] = 0.25 * lx[j];
        }
        __syncthreads();
        vjp_rhs<SCHEME, NT>(u1, g, lx, adjq, rows, used, n, a, dxinv, P, acc);
        // stage 1: u1 = u + dt R(u)
        for (int j = threadIdx.x; j < n; j += NT) {
            lu[j] += lx[j];
            g[j] = dt * lx[j];
        }
        __syncthreads();
        vjp_rhs<SCHEME, NT>(u, g, lu, adjq, rows, used, n, a, dxinv, P, acc);
        for (int j = threadIdx.x; j < n; j += NT) lam[j] = lu[j];
        __syncthreads();
    }
    // sum the waves' tiles: red[wave][t][i][j] in the (now free) row buffer
    double *red = (double *)rows;
    {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int t = 0; t < NTILE; ++t)
            for (int q = 0; q < 4; ++q)
                red[((wave * NTILE + t) * 16 + Tile<T>::row_of(lane, q)) * 16 + (lane & 15)] = acc[t][q];
    }
    __syncthreads();
    for (int p = threadIdx.x; p < MLP5_COUNT; p += NT) {
        int t, i, j;
        param_tile(p, t, i, j);
        double sum = 0.0;
        for (int w = 0; w < NT / 64; ++w) sum += red[((w * NTILE + t) * 16 + i) * 16 + j];
        grad[(size_t)b * MLP5_COUNT + p] = sum;
    }
    if (lam0)
        for (int j = threadIdx.x; j < n; j += NT) lam0[(size_t)b * n + j] = lam[j];
}

template <int SCHEME>
static size_t adjoint_lds(int n) {
    using T = typename DevView<SCHEME>::T;
    constexpr int NT = 64 * DevView<SCHEME>::WAVES;
    return (size_t)13 * n * sizeof(double) + NT * sizeof(int) + (size_t)2 * NT * ROW * sizeof(T);
}

}  // namespace wsr

using namespace wsr;

extern "C" {

int wsr_adv1d_train_forward(const wsr_adv1d_config *cfg, double *u, const double *velocities, const double *dts,
                            int32_t n_steps, double *checkpoints, const void *params_device, void *stream) {
    if (!cfg || !u || !velocities || !dts || !checkpoints || !params_device || n_steps < 1) return WSR_E_ARG;
    if (cfg->batch < 1 || cfg->n < 5 || cfg->n > T1_MAX_N) return WSR_E_ARG;
    hipStream_t st = (hipStream_t)stream;
    size_t lds = 3 * (size_t)cfg->n * sizeof(double);
    if (cfg->scheme == WSR_SCHEME_SR64) {
        MlpDev64 P{(const double *)params_device};
        hipLaunchKernelGGL(k_train_forward<SCHEME_SR64>, dim3(cfg->batch), dim3(T1_BLOCK), lds, st, u, velocities, dts,
                           n_steps, cfg->n, cfg->dxinv, checkpoints, P);
    } else if (cfg->scheme == WSR_SCHEME_SR32) {
        MlpDev32 P{(const float *)params_device};
        hipLaunchKernelGGL(k_train_forward<SCHEME_SR32>, dim3(cfg->batch), dim3(T1_BLOCK), lds, st, u, velocities, dts,
                           n_steps, cfg->n, cfg->dxinv, checkpoints, P);
    } else {
        return WSR_E_ARG;
    }
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_adv1d_train_adjoint(const wsr_adv1d_config *cfg, const double *checkpoints, const double *lam_final,
                            const double *velocities, const double *dts, int32_t n_steps, const void *params_device,
                            double *grad, double *lam_initial, void *stream) {
    if (!cfg || !checkpoints || !lam_final || !velocities || !dts || !params_device || !grad || n_steps < 1)
        return WSR_E_ARG;
    if (cfg->batch < 1 || cfg->n < 5 || cfg->n > T1_MAX_N) return WSR_E_ARG;
    hipStream_t st = (hipStream_t)stream;
    if (cfg->scheme == WSR_SCHEME_SR64) {
        MlpDev64 P{(const double *)params_device};
        size_t lds = adjoint_lds<SCHEME_SR64>(cfg->n);
        if (lds > 160 * 1024) return WSR_E_ARG;
        if (hipFuncSetAttribute((const void *)k_train_adjoint<SCHEME_SR64>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return WSR_E_HIP;
        hipLaunchKernelGGL(k_train_adjoint<SCHEME_SR64>, dim3(cfg->batch), dim3(64 * DevView<SCHEME_SR64>::WAVES), lds,
                           st, checkpoints, lam_final, velocities, dts, n_steps, cfg->n, cfg->dxinv, P, grad,
                           lam_initial);
    } else if (cfg->scheme == WSR_SCHEME_SR32) {
        MlpDev32 P{(const float *)params_device};
        size_t lds = adjoint_lds<SCHEME_SR32>(cfg->n);
        if (lds > 160 * 1024) return WSR_E_ARG;
        if (hipFuncSetAttribute((const void *)k_train_adjoint<SCHEME_SR32>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return WSR_E_HIP;
        hipLaunchKernelGGL(k_train_adjoint<SCHEME_SR32>, dim3(cfg->batch), dim3(64 * DevView<SCHEME_SR32>::WAVES), lds,
                           st, checkpoints, lam_final, velocities, dts, n_steps, cfg->n, cfg->dxinv, P, grad,
                           lam_initial);
    } else {
        return WSR_E_ARG;
    }
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

}  // extern "C"
