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
//                     of a pass are reduced in LDS by one thread per parameter.
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

// VJP of R(X)_j = -(F_j - F_{j-1}) dxinv with cotangent g: target += R_X^T g, acc += R_theta^T g.
// parameters owned by one thread of an NT-thread workgroup: p = tid, tid + NT, ...
constexpr int OWN = 2;

template <int SCHEME, int NT, typename T, class MP>
__device__ void vjp_rhs(const double *X, const double *g, double *target, double *adjq, T *rows, int *used, int n,
                        double a, double dxinv, const MP &P, const int (&my_in)[OWN], const int (&my_out)[OWN],
                        double (&acc)[OWN]) {
    const int off = a > 0.0 ? 2 : 1;
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
        // each thread sums the rows of this pass for the parameters it owns
#pragma unroll
        for (int o = 0; o < OWN; ++o) {
            if (my_in[o] < 0) continue;
            double s = 0.0;
            for (int r = 0; r < NT; ++r) {
                if (!used[r]) continue;
                const T *rd = rows + (size_t)(2 * r) * ROW;
                s += (double)rd[my_in[o]] * (double)rd[my_out[o]] + (double)rd[ROW + my_in[o]] * (double)rd[ROW + my_out[o]];
            }
            acc[o] += s;
        }
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
    static_assert(OWN * NT >= MLP5_COUNT, "every parameter needs an owner");
    int my_in[OWN], my_out[OWN];
    double acc[OWN];
    for (int o = 0; o < OWN; ++o) {
        const int p = threadIdx.x + o * NT;
        my_in[o] = my_out[o] = -1;
        acc[o] = 0.0;
        if (p < MLP5_COUNT) param_row_slots(p, my_in[o], my_out[o]);
    }
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
        vjp_rhs<SCHEME, NT>(u2, g, lx, adjq, rows, used, n, a, dxinv, P, my_in, my_out, acc);
        // stage 2: u2 = 3/4 u + 1/4 (u1 + dt R(u1))
        for (int j = threadIdx.x; j < n; j += NT) {
            lu[j] += 0.75 * lx[j];
            g[j] = 0.25 * dt * lx[j];
            lx[j] = 0.25 * lx[j];
        }
        __syncthreads();
        vjp_rhs<SCHEME, NT>(u1, g, lx, adjq, rows, used, n, a, dxinv, P, my_in, my_out, acc);
        // stage 1: u1 = u + dt R(u)
        for (int j = threadIdx.x; j < n; j += NT) {
            lu[j] += lx[j];
            g[j] = dt * lx[j];
        }
        __syncthreads();
        vjp_rhs<SCHEME, NT>(u, g, lu, adjq, rows, used, n, a, dxinv, P, my_in, my_out, acc);
        for (int j = threadIdx.x; j < n; j += NT) lam[j] = lu[j];
        __syncthreads();
    }
    for (int o = 0; o < OWN; ++o) {
        const int p = threadIdx.x + o * NT;
        if (p < MLP5_COUNT) grad[(size_t)b * MLP5_COUNT + p] = acc[o];
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
