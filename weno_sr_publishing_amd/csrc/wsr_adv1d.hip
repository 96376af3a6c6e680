// wsr_adv1d.hip -- 1-D periodic scalar advection with the WENO5 closures (gfx950).
//
// Batched trajectories q_t + a q_x = 0, a in {-1, +1}, advanced with the signed
// SSPRK3 step of the reference's training/GSTE path (paths relative to the
// reference checkout):
//   code/snapshot/teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/rk3_advection.py:82-131
//   code/snapshot/teacherfree_lab_weno5/apost_advect.py:16-45, weno5_core.py:133-239
//   code/snapshot/teacherfree_lab_weno5_v9_exact_fvm/v9_model.py:16-28 (scale remap)
// One workgroup owns one trajectory for the whole rollout: the state, the two
// stage states and the face fluxes live in LDS, every face is one lane, and the
// kernel loops over all steps without returning to the host.  Optionally the
// state at the start of every step is written out (checkpoints for the adjoint).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/wsr.h"
#include "wsr_adv1d.hpp"

namespace wsr {

constexpr int A1_BLOCK = 256;
constexpr int A1_MAX_N = 2048;

template <int SCHEME>
__global__ __launch_bounds__(A1_BLOCK) void k_adv1d_rollout(double *__restrict__ u, const double *__restrict__ vel,
                                                            const double *__restrict__ dts, int nsteps, int n,
                                                            double dxinv, double *__restrict__ ckpt,
                                                            typename MlpFor<SCHEME>::type mlp) {
    extern __shared__ double lds[];
    double *s0 = lds;            // u^n
    double *s1 = lds + n;        // stage state
    double *fl = lds + 2 * n;    // face fluxes a * q_{i+1/2}
    const int b = blockIdx.x;
    const double a = vel[b];
    double *ub = u + (size_t)b * n;
    for (int i = threadIdx.x; i < n; i += A1_BLOCK) s0[i] = ub[i];
    __syncthreads();
    for (int step = 0; step < nsteps; ++step) {
        const double dt = dts[step];
        if (ckpt) {
            double *c = ckpt + ((size_t)b * (nsteps + 1) + step) * n;
            for (int i = threadIdx.x; i < n; i += A1_BLOCK) c[i] = s0[i];
        }
        for (int stage = 0; stage < 3; ++stage) {
            const double *src = stage == 0 ? s0 : s1;
            for (int i = threadIdx.x; i < n; i += A1_BLOCK) fl[i] = a * face_signed<SCHEME>(src, n, i, a, mlp);
            __syncthreads();
            for (int i = threadIdx.x; i < n; i += A1_BLOCK) {
                double rhs = -(fl[i] - fl[i == 0 ? n - 1 : i - 1]) * dxinv;
                double v;
                if (stage == 0) v = s0[i] + dt * rhs;
                else if (stage == 1) v = 0.75 * s0[i] + 0.25 * (s1[i] + dt * rhs);
                else v = s0[i] / 3.0 + (2.0 / 3.0) * (s1[i] + dt * rhs);
                if (stage < 2) s1[i] = v;   // each lane rewrites only its own cells
                else s0[i] = v;
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += A1_BLOCK) ub[i] = s0[i];
    if (ckpt) {
        double *c = ckpt + ((size_t)b * (nsteps + 1) + nsteps) * n;
        for (int i = threadIdx.x; i < n; i += A1_BLOCK) c[i] = s0[i];
    }
}

template <int SCHEME>
static void launch_rollout(const wsr_adv1d_config *cfg, double *u, const double *vel, const double *dts,
                           int nsteps, double *ckpt, hipStream_t st) {
    using MLP = typename MlpFor<SCHEME>::type;
    MLP mlp;
    memset(&mlp, 0, sizeof mlp);
    if constexpr (SCHEME == SCHEME_SR64 || SCHEME == SCHEME_SR32) memcpy(mlp.p, cfg->mlp, sizeof mlp.p);
    if constexpr (SCHEME == SCHEME_SRBF16) {
        const float *src = (const float *)cfg->mlp;
        for (int i = 0; i < MLP5_COUNT; ++i) mlp.p[i] = host_bf16(src[i]);
    }
    size_t lds = 3 * (size_t)cfg->n * sizeof(double);
    hipLaunchKernelGGL((k_adv1d_rollout<SCHEME>), dim3(cfg->batch), dim3(A1_BLOCK), lds, st, u, vel, dts, nsteps,
                       cfg->n, cfg->dxinv, ckpt, mlp);
}

}  // namespace wsr

using namespace wsr;

extern "C" int wsr_adv1d_rollout(const wsr_adv1d_config *cfg, double *u, const double *velocities,
                                 const double *dts, int32_t n_steps, double *checkpoints, void *stream) {
    if (!cfg || !u || !velocities || (!dts && n_steps > 0) || n_steps < 0) return WSR_E_ARG;
    if (cfg->batch < 1 || cfg->batch > (1 << 24) || cfg->n < 5 || cfg->n > A1_MAX_N) return WSR_E_ARG;
    if (cfg->scheme != WSR_SCHEME_JS && cfg->scheme != WSR_SCHEME_SR64 && cfg->scheme != WSR_SCHEME_SR32 &&
        cfg->scheme != WSR_SCHEME_SRBF16)
        return WSR_E_ARG;
    if (cfg->scheme != WSR_SCHEME_JS && !cfg->mlp) return WSR_E_MLP;
    hipStream_t st = (hipStream_t)stream;
    switch (cfg->scheme) {
        case WSR_SCHEME_JS: launch_rollout<SCHEME_JS>(cfg, u, velocities, dts, n_steps, checkpoints, st); break;
        case WSR_SCHEME_SR64: launch_rollout<SCHEME_SR64>(cfg, u, velocities, dts, n_steps, checkpoints, st); break;
        case WSR_SCHEME_SR32: launch_rollout<SCHEME_SR32>(cfg, u, velocities, dts, n_steps, checkpoints, st); break;
        default: launch_rollout<SCHEME_SRBF16>(cfg, u, velocities, dts, n_steps, checkpoints, st); break;
    }
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}
