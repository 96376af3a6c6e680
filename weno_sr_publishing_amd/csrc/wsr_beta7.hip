// wsr_beta7.hip -- the beta-provider protocol of the WENO7-SR point step (gfx950).
//
// The fused sweep of wsr_euler7.hip evaluates the 6-24-16-16-4 closure inside the
// reconstruction; no array of smoothness indicators ever exists there.  The
// reference's own interface is different: a beta provider object fills four arrays
// of smoothness indicators (beta = 4 r, ones on a plateau), and separate kernels
// reconstruct from whatever those arrays hold, so that a caller can plug in any
// provider.  This file restores that interface:
//   * fill kernels: the closure of this library as a provider (FP64 or FP32 MLP),
//     writing the reference layouts;
//   * consumers: normal face values from external beta (both sides), transverse
//     Gauss-point values from external cross beta with the Riemann flux of one Gauss
//     location, the right-hand side from the four flux arrays, and the Shu RK4 stage
//     combinations.
// One thread per array entry (or per face): a protocol path for pluggable providers
// and for inspection, not the production step.
//
// Layouts (FP64, C order), as the reference allocates them:
//   beta_x (ny+8, nx+2, 32), beta_y (ny+2, nx+8, 32): entry side*16 + comp*4 + k
//   beta_cross_x (ny, nx+2, 256), beta_cross_y (ny+2, nx, 256):
//     entry group*128 + side*16 + comp*4 + k (group 0 from x_r / y_r, 1 from x_l / y_l)
//   face values x_l, x_r, y_l, y_r: padded (ny+8, nx+8, 4), entries [j, i] with
//     i < nx+2 (x) / j < ny+2 (y)
//   fluxes fx (ny, nx+1, 4), fy (ny+1, nx, 4); rhs and stages padded (ny+8, nx+8, 4)
//
// Reference (paths relative to the reference checkout):
//   BETA = code/snapshot/for_paper_results/solvers/weno7_warp_beta.py (fill kernels, :266-560)
//   MLP7 = code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/point_rk4_mlp.py
//          (protocol :22-56, stage / flux kernels :61-186, compute_rhs_mlp :194-273)
//   EXT  = code/snapshot/weno7_external_clean/warp_weno7_external_overwrite_fullall.py (:20-240)
//   PRK4 = code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/point_rk4.py
//          (rhs_from_flux :290-312, Shu stages :315-403)
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/wsr.h"
#include "wsr_loop.hpp"
#include "wsr_weno7.hpp"

namespace wsr {
namespace b7 {

constexpr int GC = 4;
constexpr int BLOCK = 256;
constexpr int NSTRIDE = 32;    // normal beta entries per face
constexpr int XSTRIDE = 256;   // cross beta entries per face

struct Dims {
    int nx, ny, nxt;
    double gamma;
};

__device__ __forceinline__ void ld4(const double *a, int nxt, int j, int i, double v[4]) {
    const double2 *p = reinterpret_cast<const double2 *>(a + ((size_t)j * nxt + i) * 4);
    const double2 x = p[0], y = p[1];
    v[0] = x.x; v[1] = x.y; v[2] = y.x; v[3] = y.y;
}

__device__ __forceinline__ void st4(double *a, size_t rec, const double v[4]) {
    double2 *p = reinterpret_cast<double2 *>(a + rec * 4);
    p[0] = make_double2(v[0], v[1]);
    p[1] = make_double2(v[2], v[3]);
}

// Characteristic values of component k over the 7-cell window q (cells m = 0..6) with
// the Roe state of the face the side belongs to: lr 1 (right face of the centre cell)
// averages cells 3 and 4, lr 2 (left face) cells 2 and 3 (EXT:173-211, BETA:347-350).
template <int DIR>
__device__ __forceinline__ Eigen side_eigen(const double q[7][4], int lr, double gamma) {
    return lr == 1 ? roe_eigen(q[3], q[4], gamma) : roe_eigen(q[2], q[3], gamma);
}

// ------------------------------------------------------------ providers
// BETA:318-426 fill_normal_x_kernel / fill_normal_y_kernel: one thread per (face
// entry, side, component); the face entry [j, i] holds the window of cells
// i..i+6 of row j (x) or rows j..j+6 of column i (y).
template <int DIR, int SCHEME>
__global__ __launch_bounds__(BLOCK) void k_fill_normal(const double *__restrict__ u, double *__restrict__ beta, Dims d,
                                                       typename Mlp7Sweep<SCHEME>::type mlp) {
    const int rows = DIR == 1 ? d.ny + 8 : d.ny + 2, cols = DIR == 1 ? d.nx + 2 : d.nx + 8;
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= (long)rows * cols * 8) return;
    const int lane = (int)(t & 7);
    const long cell = t >> 3;
    const int j = (int)(cell / cols), i = (int)(cell % cols);
    const int side = lane >> 2, comp = lane & 3, lr = side + 1;
    double q[7][4];
#pragma unroll
    for (int m = 0; m < 7; ++m) ld4(u, d.nxt, DIR == 1 ? j : j + m, DIR == 1 ? i + m : i, q[m]);
    Eigen E = side_eigen<DIR>(q, lr, d.gamma);
    double l[4], scale;
    left_row<DIR>(comp, E, d.gamma, l, scale);
    double c[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) c[m] = (l[0] * q[m][0] + l[1] * q[m][1] + l[2] * q[m][2] + l[3] * q[m][3]) * scale;
    double b[4];
    bool invalid;
    closure_beta7<SCHEME>(c, lr == 2, mlp, b, invalid);
    double *o = beta + ((size_t)j * cols + i) * NSTRIDE + side * 16 + comp * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = b[k];
}

// BETA:445-560 fill_cross_x_kernel / fill_cross_y_kernel: (face entry, group, side,
// component); group 0 reads the right-face array, group 1 the left-face array, over
// the seven transverse entries rows j+1..j+7 of column i (x) / columns i+1..i+7 of
// row j (y), conservative components.
template <int DIR, int SCHEME>
__global__ __launch_bounds__(BLOCK) void k_fill_cross(const double *__restrict__ right, const double *__restrict__ left,
                                                      double *__restrict__ beta, Dims d,
                                                      typename Mlp7Sweep<SCHEME>::type mlp) {
    const int rows = DIR == 1 ? d.ny : d.ny + 2, cols = DIR == 1 ? d.nx + 2 : d.nx;
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= (long)rows * cols * 16) return;
    const int lane = (int)(t & 15);
    const long cell = t >> 4;
    const int j = (int)(cell / cols), i = (int)(cell % cols);
    const int group = lane >> 3, side = (lane >> 2) & 1, comp = lane & 3;
    const double *src = group == 0 ? right : left;
    double v[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) {
        const int jj = DIR == 1 ? j + 1 + m : j, ii = DIR == 1 ? i : i + 1 + m;
        v[m] = src[((size_t)jj * d.nxt + ii) * 4 + comp];
    }
    double b[4];
    bool invalid;
    closure_beta7<SCHEME>(v, side == 1, mlp, b, invalid);
    double *o = beta + ((size_t)j * cols + i) * XSTRIDE + group * 128 + side * 16 + comp * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = b[k];
}

// ------------------------------------------------------------ consumers
// EXT:173-211 + MLP7:61-109: both normal face values of the window [j, i] from
// external beta (lr 1 -> right, lr 2 -> left), characteristic reconstruction with
// the Roe state of each side's face.
template <int DIR>
__global__ __launch_bounds__(BLOCK) void k_stage_ext(const double *__restrict__ u, const double *__restrict__ beta,
                                                     double *__restrict__ left, double *__restrict__ right, Dims d,
                                                     int eno) {
    const int rows = DIR == 1 ? d.ny + 8 : d.ny + 2, cols = DIR == 1 ? d.nx + 2 : d.nx + 8;
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= (long)rows * cols) return;
    const int j = (int)(t / cols), i = (int)(t % cols);
    double q[7][4];
#pragma unroll
    for (int m = 0; m < 7; ++m) ld4(u, d.nxt, DIR == 1 ? j : j + m, DIR == 1 ? i + m : i, q[m]);
    const double *bt = beta + ((size_t)j * cols + i) * NSTRIDE;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int lr = side + 1;
        Eigen E = side_eigen<DIR>(q, lr, d.gamma);
        double cf[4];
#pragma unroll
        for (int comp = 0; comp < 4; ++comp) {
            double l[4], scale;
            left_row<DIR>(comp, E, d.gamma, l, scale);
            double c[7];
#pragma unroll
            for (int m = 0; m < 7; ++m)
                c[m] = (l[0] * q[m][0] + l[1] * q[m][1] + l[2] * q[m][2] + l[3] * q[m][3]) * scale;
            double s[4], w[4], b[4];
            if (lr == 1) { candidates7<1>(c, s); optimal7<1>(w); }
            else { candidates7<2>(c, s); optimal7<2>(w); }
#pragma unroll
            for (int k = 0; k < 4; ++k) b[k] = bt[side * 16 + comp * 4 + k];
            cf[comp] = combine_beta7(s, w, b, eno);
        }
        double out[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double r[4];
            right_row<DIR>(k, E, r);
            out[k] = r[0] * cf[0] + r[1] * cf[1] + r[2] * cf[2] + r[3] * cf[3];
        }
        st4(lr == 1 ? right : left, (size_t)j * d.nxt + i, out);
    }
}

// Gauss-point value (loca 1: head 4, side-0 beta; loca 2: head 3, side-1 beta) of the
// four conservative components over seven transverse entries (EXT:128-171,213-240).
__device__ __forceinline__ void gauss_state(const double q[7][4], const double *bt, int loca, int eno, double out[4]) {
    const int side = loca == 1 ? 0 : 1;
#pragma unroll
    for (int comp = 0; comp < 4; ++comp) {
        double v[7], s[4], w[4], b[4];
#pragma unroll
        for (int m = 0; m < 7; ++m) v[m] = q[m][comp];
        if (loca == 1) { candidates7<4>(v, s); optimal7<4>(w); }
        else { candidates7<3>(v, s); optimal7<3>(w); }
#pragma unroll
        for (int k = 0; k < 4; ++k) b[k] = bt[side * 16 + comp * 4 + k];
        out[comp] = combine_beta7(s, w, b, eno);
    }
}

// MLP7:111-186 compute_x_flux_point_mlp_kernel / compute_y_flux_point_mlp_kernel:
// the flux of one Gauss location at face [j, i] (x: j < ny, i < nx+1; y: j < ny+1,
// i < nx) from the right-face entries of the cell before and the left-face entries of
// the cell after; reverse swaps the states (the downwind operator).
template <int DIR, int RIEMANN>
__global__ __launch_bounds__(BLOCK) void k_flux_ext(const double *__restrict__ left, const double *__restrict__ right,
                                                    const double *__restrict__ beta, double *__restrict__ flux, Dims d,
                                                    int loca, double dt_h, int reverse, int eno) {
    const int rows = DIR == 1 ? d.ny : d.ny + 1, cols = DIR == 1 ? d.nx + 1 : d.nx;
    const int bcols = DIR == 1 ? d.nx + 2 : d.nx;   // columns of the cross beta array
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= (long)rows * cols) return;
    const int j = (int)(t / cols), i = (int)(t % cols);
    double ql[7][4], qr[7][4];
#pragma unroll
    for (int m = 0; m < 7; ++m) {
        if (DIR == 1) {
            ld4(right, d.nxt, j + 1 + m, i, ql[m]);
            ld4(left, d.nxt, j + 1 + m, i + 1, qr[m]);
        } else {
            ld4(right, d.nxt, j, i + 1 + m, ql[m]);
            ld4(left, d.nxt, j + 1, i + 1 + m, qr[m]);
        }
    }
    const double *btl = beta + ((size_t)j * bcols + i) * XSTRIDE;
    const double *btr = DIR == 1 ? beta + ((size_t)j * bcols + i + 1) * XSTRIDE + 128
                                 : beta + ((size_t)(j + 1) * bcols + i) * XSTRIDE + 128;
    double sl[4], sr[4], f[4];
    gauss_state(ql, btl, loca, eno, sl);
    gauss_state(qr, btr, loca, eno, sr);
    const double *a = reverse ? sr : sl, *b = reverse ? sl : sr;
    double ua[4] = {a[0], a[1], a[2], a[3]}, ub[4] = {b[0], b[1], b[2], b[3]};
    if constexpr (RIEMANN == WSR_RIEMANN_HLLC) hllc_flux<DIR>(ua, ub, d.gamma, f);
    else if constexpr (RIEMANN == WSR_RIEMANN_FORCE) force_flux<DIR>(ua, ub, dt_h, d.gamma, f);
    else evilin_flux<DIR>(ua, ub, dt_h, d.gamma, f);
    st4(flux, (size_t)j * cols + i, f);
}

// PRK4:290-312 rhs_from_flux_kernel (interior cells; ghost entries untouched)
__global__ __launch_bounds__(BLOCK) void k_rhs_from_flux(double *__restrict__ rhs, const double *__restrict__ fx1,
                                                         const double *__restrict__ fx2, const double *__restrict__ fy1,
                                                         const double *__restrict__ fy2, Dims d, double inv_dx,
                                                         double inv_dy) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= d.nx * d.ny) return;
    const int j = idx / d.nx, i = idx - j * d.nx;
    double o[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double fxl = fx1[((size_t)j * (d.nx + 1) + i) * 4 + c] + fx2[((size_t)j * (d.nx + 1) + i) * 4 + c];
        double fxr = fx1[((size_t)j * (d.nx + 1) + i + 1) * 4 + c] + fx2[((size_t)j * (d.nx + 1) + i + 1) * 4 + c];
        double fyb = fy1[((size_t)j * d.nx + i) * 4 + c] + fy2[((size_t)j * d.nx + i) * 4 + c];
        double fyt = fy1[((size_t)(j + 1) * d.nx + i) * 4 + c] + fy2[((size_t)(j + 1) * d.nx + i) * 4 + c];
        o[c] = -0.5 * (fxr - fxl) * inv_dx - 0.5 * (fyt - fyb) * inv_dy;
    }
    st4(rhs, (size_t)(j + GC) * d.nxt + (i + GC), o);
}

// PRK4:315-403 Shu RK4 stage combinations over every padded entry (stage 1: u1,
// 2: u2, 3: u3, 4: the new state), the reference's term order
struct ShuArgs {
    const double *in[7];
    double *out;
};

template <int STAGE>
__global__ __launch_bounds__(BLOCK) void k_shu(ShuArgs A, long n, double dt) {
    const long e = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= n) return;
    double r;
    if constexpr (STAGE == 1) {
        r = A.in[0][e] + 0.5 * dt * A.in[1][e];
    } else if constexpr (STAGE == 2) {
        r = 649.0 / 1600.0 * A.in[0][e] - 10890423.0 / 25193600.0 * dt * A.in[1][e] + 951.0 / 1600.0 * A.in[2][e] +
            5000.0 / 7873.0 * dt * A.in[3][e];
    } else if constexpr (STAGE == 3) {
        r = 53989.0 / 2500000.0 * A.in[0][e] - 102261.0 / 5000000.0 * dt * A.in[1][e] +
            4806213.0 / 20000000.0 * A.in[2][e] - 5121.0 / 20000.0 * dt * A.in[3][e] +
            23619.0 / 32000.0 * A.in[4][e] + 7873.0 / 10000.0 * dt * A.in[5][e];
    } else {
        r = 1.0 / 5.0 * A.in[0][e] + 1.0 / 10.0 * dt * A.in[1][e] + 6127.0 / 30000.0 * A.in[2][e] +
            1.0 / 6.0 * dt * A.in[3][e] + 7873.0 / 30000.0 * A.in[4][e] + 1.0 / 3.0 * A.in[5][e] +
            1.0 / 6.0 * dt * A.in[6][e];
    }
    A.out[e] = r;
}

// closure parameters: neuron-major block in the caller's workspace, uploaded by value
constexpr int PCHUNK = 448;
struct ParamChunk { uint32_t w[PCHUNK]; };

__global__ void k_put_params(ParamChunk src, uint32_t *dst, int nwords) {
    for (int i = threadIdx.x; i < nwords; i += blockDim.x) dst[i] = src.w[i];
}

static void upload_params(const wsr_euler7_config *cfg, void *ws, hipStream_t st) {
    uint32_t words[MLP7_COUNT * 2];
    int nwords;
    if (cfg->scheme == WSR_SCHEME_SR64) {
        double dev[MLP7_COUNT];
        pack7_neuron_major((const double *)cfg->mlp, dev);
        memcpy(words, dev, sizeof dev);
        nwords = MLP7_COUNT * 2;
    } else {
        float dev[MLP7_COUNT];
        pack7_neuron_major((const float *)cfg->mlp, dev);
        memcpy(words, dev, sizeof dev);
        nwords = MLP7_COUNT;
    }
    for (int o = 0; o < nwords; o += PCHUNK) {
        ParamChunk c;
        memset(&c, 0, sizeof c);
        const int n = nwords - o < PCHUNK ? nwords - o : PCHUNK;
        memcpy(c.w, words + o, (size_t)n * sizeof(uint32_t));
        hipLaunchKernelGGL(k_put_params, dim3(1), dim3(256), 0, st, c, (uint32_t *)ws + o, n);
    }
}

static int check_cfg(const wsr_euler7_config *cfg, bool need_mlp) {
    if (!cfg) return WSR_E_ARG;
    if (cfg->ghost != GC || cfg->nx < 1 || cfg->ny < 1 || cfg->nx > (1 << 15) || cfg->ny > (1 << 15))
        return WSR_E_ARG;
    if (need_mlp) {
        if (cfg->scheme != WSR_SCHEME_SR64 && cfg->scheme != WSR_SCHEME_SR32) return WSR_E_ARG;
        if (!cfg->mlp) return WSR_E_MLP;
    }
    return WSR_OK;
}

static Dims dims(const wsr_euler7_config *cfg) {
    Dims d;
    d.nx = cfg->nx; d.ny = cfg->ny; d.nxt = cfg->nx + 2 * GC; d.gamma = cfg->gamma;
    return d;
}

static dim3 blocks_for(long n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

constexpr size_t PARAM_BYTES = MLP7_COUNT * sizeof(double);

}  // namespace b7
}  // namespace wsr

using namespace wsr;
using namespace wsr::b7;

extern "C" {

size_t wsr_beta7_workspace_bytes(void) { return PARAM_BYTES; }

int wsr_beta7_fill_normal(const wsr_euler7_config *cfg, const double *u, double *beta, int32_t dir, void *workspace,
                          size_t workspace_bytes, void *stream) {
    int rc = check_cfg(cfg, true);
    if (rc) return rc;
    if (!u || !beta || !workspace || (dir != 1 && dir != 2)) return WSR_E_ARG;
    if (!all_aligned16(u, beta, workspace)) return WSR_E_ARG;
    if (workspace_bytes < PARAM_BYTES) return WSR_E_WORKSPACE;
    hipStream_t st = (hipStream_t)stream;
    upload_params(cfg, workspace, st);
    Dims d = dims(cfg);
    const long n = (long)(dir == 1 ? (d.ny + 8) * (long)(d.nx + 2) : (d.ny + 2) * (long)(d.nx + 8)) * 8;
    if (cfg->scheme == WSR_SCHEME_SR64) {
        MlpK<double> m{(const double *)workspace};
        if (dir == 1) hipLaunchKernelGGL((k_fill_normal<1, SCHEME_SR64>), blocks_for(n), dim3(BLOCK), 0, st, u, beta, d, m);
        else hipLaunchKernelGGL((k_fill_normal<2, SCHEME_SR64>), blocks_for(n), dim3(BLOCK), 0, st, u, beta, d, m);
    } else {
        MlpK<float> m{(const float *)workspace};
        if (dir == 1) hipLaunchKernelGGL((k_fill_normal<1, SCHEME_SR32>), blocks_for(n), dim3(BLOCK), 0, st, u, beta, d, m);
        else hipLaunchKernelGGL((k_fill_normal<2, SCHEME_SR32>), blocks_for(n), dim3(BLOCK), 0, st, u, beta, d, m);
    }
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_beta7_fill_cross(const wsr_euler7_config *cfg, const double *right, const double *left, double *beta,
                         int32_t dir, void *workspace, size_t workspace_bytes, void *stream) {
    int rc = check_cfg(cfg, true);
    if (rc) return rc;
    if (!right || !left || !beta || !workspace || (dir != 1 && dir != 2)) return WSR_E_ARG;
    if (!all_aligned16(right, left, beta, workspace)) return WSR_E_ARG;
    if (workspace_bytes < PARAM_BYTES) return WSR_E_WORKSPACE;
    hipStream_t st = (hipStream_t)stream;
    upload_params(cfg, workspace, st);
    Dims d = dims(cfg);
    const long n = (long)(dir == 1 ? d.ny * (long)(d.nx + 2) : (d.ny + 2) * (long)d.nx) * 16;
    if (cfg->scheme == WSR_SCHEME_SR64) {
        MlpK<double> m{(const double *)workspace};
        if (dir == 1) hipLaunchKernelGGL((k_fill_cross<1, SCHEME_SR64>), blocks_for(n), dim3(BLOCK), 0, st, right, left, beta, d, m);
        else hipLaunchKernelGGL((k_fill_cross<2, SCHEME_SR64>), blocks_for(n), dim3(BLOCK), 0, st, right, left, beta, d, m);
    } else {
        MlpK<float> m{(const float *)workspace};
        if (dir == 1) hipLaunchKernelGGL((k_fill_cross<1, SCHEME_SR32>), blocks_for(n), dim3(BLOCK), 0, st, right, left, beta, d, m);
        else hipLaunchKernelGGL((k_fill_cross<2, SCHEME_SR32>), blocks_for(n), dim3(BLOCK), 0, st, right, left, beta, d, m);
    }
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_euler7_stage_ext(const wsr_euler7_config *cfg, const double *u, const double *beta, double *left,
                         double *right, int32_t dir, void *stream) {
    int rc = check_cfg(cfg, false);
    if (rc) return rc;
    if (!u || !beta || !left || !right || (dir != 1 && dir != 2)) return WSR_E_ARG;
    if (!all_aligned16(u, beta, left, right)) return WSR_E_ARG;
    hipStream_t st = (hipStream_t)stream;
    Dims d = dims(cfg);
    const long n = dir == 1 ? (d.ny + 8) * (long)(d.nx + 2) : (d.ny + 2) * (long)(d.nx + 8);
    if (dir == 1) hipLaunchKernelGGL(k_stage_ext<1>, blocks_for(n), dim3(BLOCK), 0, st, u, beta, left, right, d, cfg->eno_cutoff);
    else hipLaunchKernelGGL(k_stage_ext<2>, blocks_for(n), dim3(BLOCK), 0, st, u, beta, left, right, d, cfg->eno_cutoff);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_euler7_flux_ext(const wsr_euler7_config *cfg, const double *left, const double *right,
                        const double *beta_cross, double *flux, int32_t dir, int32_t loca, double dt,
                        int32_t reverse_upwind, void *stream) {
    int rc = check_cfg(cfg, false);
    if (rc) return rc;
    if (!left || !right || !beta_cross || !flux || (dir != 1 && dir != 2) || (loca != 1 && loca != 2))
        return WSR_E_ARG;
    if (cfg->riemann != WSR_RIEMANN_HLLC && cfg->riemann != WSR_RIEMANN_FORCE && cfg->riemann != WSR_RIEMANN_EVILIN)
        return WSR_E_ARG;
    if (!all_aligned16(left, right, beta_cross, flux)) return WSR_E_ARG;
    hipStream_t st = (hipStream_t)stream;
    Dims d = dims(cfg);
    const long n = dir == 1 ? d.ny * (long)(d.nx + 1) : (d.ny + 1) * (long)d.nx;
    const double dt_h = dt / (dir == 1 ? cfg->dx : cfg->dy);
    const int rev = reverse_upwind ? 1 : 0, eno = cfg->eno_cutoff;
#define WSR_FLUX_CASE(R)                                                                                           \
    if (dir == 1)                                                                                                  \
        hipLaunchKernelGGL((k_flux_ext<1, R>), blocks_for(n), dim3(BLOCK), 0, st, left, right, beta_cross, flux, d, \
                           (int)loca, dt_h, rev, eno);                                                             \
    else                                                                                                           \
        hipLaunchKernelGGL((k_flux_ext<2, R>), blocks_for(n), dim3(BLOCK), 0, st, left, right, beta_cross, flux, d, \
                           (int)loca, dt_h, rev, eno);
    if (cfg->riemann == WSR_RIEMANN_HLLC) { WSR_FLUX_CASE(WSR_RIEMANN_HLLC) }
    else if (cfg->riemann == WSR_RIEMANN_FORCE) { WSR_FLUX_CASE(WSR_RIEMANN_FORCE) }
    else { WSR_FLUX_CASE(WSR_RIEMANN_EVILIN) }
#undef WSR_FLUX_CASE
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_euler7_rhs_from_flux(const wsr_euler7_config *cfg, double *rhs, const double *fx1, const double *fx2,
                             const double *fy1, const double *fy2, void *stream) {
    int rc = check_cfg(cfg, false);
    if (rc) return rc;
    if (!rhs || !fx1 || !fx2 || !fy1 || !fy2) return WSR_E_ARG;
    if (!all_aligned16(rhs, fx1, fx2, fy1, fy2)) return WSR_E_ARG;
    hipStream_t st = (hipStream_t)stream;
    Dims d = dims(cfg);
    hipLaunchKernelGGL(k_rhs_from_flux, blocks_for((long)d.nx * d.ny), dim3(BLOCK), 0, st, rhs, fx1, fx2, fy1, fy2, d,
                       1.0 / cfg->dx, 1.0 / cfg->dy);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_euler7_shu_combine(const wsr_euler7_config *cfg, int32_t stage, double dt, const double *const *inputs,
                           double *out, void *stream) {
    int rc = check_cfg(cfg, false);
    if (rc) return rc;
    static const int n_in[5] = {0, 2, 4, 6, 7};
    if (stage < 1 || stage > 4 || !inputs || !out) return WSR_E_ARG;
    ShuArgs A;
    memset(&A, 0, sizeof A);
    for (int k = 0; k < n_in[stage]; ++k) {
        if (!inputs[k]) return WSR_E_ARG;
        A.in[k] = inputs[k];
    }
    A.out = out;
    hipStream_t st = (hipStream_t)stream;
    const long n = (long)(cfg->nx + 2 * GC) * (cfg->ny + 2 * GC) * 4;
    switch (stage) {
        case 1: hipLaunchKernelGGL(k_shu<1>, blocks_for(n), dim3(BLOCK), 0, st, A, n, dt); break;
        case 2: hipLaunchKernelGGL(k_shu<2>, blocks_for(n), dim3(BLOCK), 0, st, A, n, dt); break;
        case 3: hipLaunchKernelGGL(k_shu<3>, blocks_for(n), dim3(BLOCK), 0, st, A, n, dt); break;
        default: hipLaunchKernelGGL(k_shu<4>, blocks_for(n), dim3(BLOCK), 0, st, A, n, dt); break;
    }
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

}  // extern "C"
