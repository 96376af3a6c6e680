// wsr_init.hip -- initial states built on the device.
//
// Each kernel computes one padded cell's Gauss-Legendre cell average with the same
// node order, the same expression order and no fused multiply-add (the whole file is
// compiled with contraction off), so that the piecewise-constant problems reproduce
// the host builders bit for bit:
//   * four-quadrant Riemann problems: for_paper_results/run_quadrant.py:64-106
//     (averages on the ghost-4 grid of the unit square, interior re-padded
//     periodically as run_weno5_timing.py:86-104 / run_vortex.py:34-35 do);
//   * Sod strip: teacherfree_lab_weno5/warp_sod_validation.py:48-62,176-210;
//   * double-Mach reflection: run_double_mach_compare.py:43-88;
//   * periodic isentropic vortex: weno7_point_rk4_shu/point_rk4.py:81-117 (exp and pow:
//     equal to the host builder to the last bits of those functions);
//   * 3-D shock-bubble: warp_weno5_3d_rk3/kernels.py:730-799 (Ma = 3 in SI units) and
//     the unit-spacing benchmark case.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/wsr.h"

namespace wsr {
namespace init {

constexpr int BLOCK = 256;

struct Quad {
    double x[16], w[16];
    int n;
};

struct Prim4 { double rho, u, v, p; };

__device__ __forceinline__ void con4(Prim4 s, double gamma, double q[4]) {
    q[0] = s.rho;
    q[1] = s.rho * s.u;
    q[2] = s.rho * s.v;
    q[3] = s.p / (gamma - 1.0) + 0.5 * s.rho * (s.u * s.u + s.v * s.v);
}

__device__ __forceinline__ void store4(double *dst, size_t idx, const double q[4]) {
    double *o = dst + idx * 4;
    o[0] = q[0]; o[1] = q[1]; o[2] = q[2]; o[3] = q[3];
}

// ---------------------------------------------------------------- quadrant
struct QuadrantArgs {
    Prim4 states[4];   // 1 = upper right, 2 = upper left, 3 = lower left, 4 = lower right
    int nx, ny, ghost, wrap;
    double x_min, y_min, dx, dy, gamma;
    Quad quad;
};

__global__ __launch_bounds__(BLOCK) void k_quadrant(double *__restrict__ u, QuadrantArgs A) {
    const int nxt = A.nx + 2 * A.ghost, nyt = A.ny + 2 * A.ghost;
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= nxt * nyt) return;
    const int jp = idx / nxt, ip = idx - jp * nxt;
    int i = ip - A.ghost, j = jp - A.ghost;
    if (A.wrap) {   // periodic re-pad of the interior: padded (jp, ip) copies interior cell (j, i)
        i %= A.nx;
        j %= A.ny;
        i += i < 0 ? A.nx : 0;
        j += j < 0 ? A.ny : 0;
    }
    const double xc = A.x_min + ((double)i + 0.5) * A.dx;
    const double yc = A.y_min + ((double)j + 0.5) * A.dy;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int a = 0; a < A.quad.n; ++a) {
        const double x = xc + 0.5 * A.dx * A.quad.x[a];
        for (int b = 0; b < A.quad.n; ++b) {
            const double y = yc + 0.5 * A.dy * A.quad.x[b];
            const int region = x >= 0.5 ? (y >= 0.5 ? 0 : 3) : (y >= 0.5 ? 1 : 2);
            double q[4];
            con4(A.states[region], A.gamma, q);
            const double c = 0.25 * A.quad.w[a] * A.quad.w[b];
            for (int k = 0; k < 4; ++k) acc[k] = acc[k] + c * q[k];
        }
    }
    store4(u, idx, acc);
}

// ---------------------------------------------------------------- Sod strip
struct SodArgs {
    int nx, ny, ghost, axis;   // axis 1 = x, 2 = y
    double origin, h, gamma;
    Quad quad;
};

__global__ __launch_bounds__(BLOCK) void k_sod(double *__restrict__ u, SodArgs A) {
    const int nxt = A.nx + 2 * A.ghost, nyt = A.ny + 2 * A.ghost;
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= nxt * nyt) return;
    const int jp = idx / nxt, ip = idx - jp * nxt;
    const int n = A.axis == 2 ? jp : ip;
    const double centre = A.origin + ((double)(n - A.ghost) + 0.5) * A.h;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int a = 0; a < A.quad.n; ++a) {
        const double coord = centre + 0.5 * A.h * A.quad.x[a];
        const bool left = coord < 0.0;
        const double rho = left ? 1.0 : 0.125, vel = 0.0, p = left ? 1.0 : 0.1;
        const double energy = 0.5 * rho * vel * vel + p / (A.gamma - 1.0);
        double q[4] = {rho, A.axis == 2 ? 0.0 : rho * vel, A.axis == 2 ? rho * vel : 0.0, energy};
        for (int k = 0; k < 4; ++k) acc[k] = acc[k] + A.quad.w[a] * q[k];
    }
    for (int k = 0; k < 4; ++k) acc[k] = 0.5 * acc[k];
    store4(u, idx, acc);
}

// ---------------------------------------------------------------- double Mach
struct DoubleMachArgs {
    int nx, ny, ghost;
    double dx, dy, gamma;
    Quad quad;
};

__global__ __launch_bounds__(BLOCK) void k_double_mach(double *__restrict__ u, DoubleMachArgs A) {
    const int nxt = A.nx + 2 * A.ghost, nyt = A.ny + 2 * A.ghost;
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= nxt * nyt) return;
    const int jp = idx / nxt, ip = idx - jp * nxt;
    const double xc = ((double)(ip - A.ghost) + 0.5) * A.dx;
    const double yc = ((double)(jp - A.ghost) + 0.5) * A.dy;
    const double root3 = sqrt(3.0);
    const Prim4 post = {8.0, 8.25 * root3 * 0.5, -8.25 * 0.5, 116.5}, pre = {1.4, 0.0, 0.0, 1.0};
    double qpost[4], qpre[4];
    con4(post, A.gamma, qpost);
    con4(pre, A.gamma, qpre);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int a = 0; a < A.quad.n; ++a) {
        const double x = xc + 0.5 * A.dx * A.quad.x[a];
        for (int b = 0; b < A.quad.n; ++b) {
            const double y = yc + 0.5 * A.dy * A.quad.x[b];
            const double shock_x = 1.0 / 6.0 + y / root3 + 20.0 * 0.0 / root3;
            const double *q = x < shock_x ? qpost : qpre;
            const double c = A.quad.w[a] * A.quad.w[b];
            for (int k = 0; k < 4; ++k) acc[k] = acc[k] + c * q[k];
        }
    }
    for (int k = 0; k < 4; ++k) acc[k] = 0.25 * acc[k];
    store4(u, idx, acc);
}

// ---------------------------------------------------------------- isentropic vortex
struct VortexArgs {
    int nx, ny, ghost;
    double t, x_min, y_min, dx, dy, gamma;
    Quad quad;
};

__global__ __launch_bounds__(BLOCK) void k_vortex(double *__restrict__ u, VortexArgs A) {
    const int nxt = A.nx + 2 * A.ghost, nyt = A.ny + 2 * A.ghost;
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= nxt * nyt) return;
    const int jp = idx / nxt, ip = idx - jp * nxt;
    const double xc = A.x_min + ((double)(ip - A.ghost) + 0.5) * A.dx;
    const double yc = A.y_min + ((double)(jp - A.ghost) + 0.5) * A.dy;
    const double beta = 5.0, length = 20.0, pi = 3.141592653589793, g = A.gamma;
    const double coef = -((g - 1.0) * beta * beta) / (8.0 * g * pi * pi);
    const double twopi_beta = beta / (2.0 * pi);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int a = 0; a < A.quad.n; ++a) {
        const double x = xc + 0.5 * A.dx * A.quad.x[a];
        for (int b = 0; b < A.quad.n; ++b) {
            const double y = yc + 0.5 * A.dy * A.quad.x[b];
            double ddx = x - 1.0 * A.t, ddy = y - 1.0 * A.t;
            ddx = ddx - length * rint(ddx / length);   // numpy round: half to even
            ddy = ddy - length * rint(ddy / length);
            const double r2 = ddx * ddx + ddy * ddy;
            const double d_t = coef * exp(1.0 - r2);
            Prim4 s;
            s.rho = 1.0 + (pow(1.0 + d_t, 1.0 / (g - 1.0)) - 1.0);
            s.p = 1.0 + (pow(1.0 + d_t, g / (g - 1.0)) - 1.0);
            const double factor = twopi_beta * exp(0.5 * (1.0 - r2));
            s.u = 1.0 - factor * ddy;
            s.v = 1.0 + factor * ddx;
            double q[4];
            con4(s, g, q);
            const double c = 0.25 * A.quad.w[a] * A.quad.w[b];
            for (int k = 0; k < 4; ++k) acc[k] = acc[k] + c * q[k];
        }
    }
    store4(u, idx, acc);
}

// ---------------------------------------------------------------- 3-D shock-bubble
struct Bubble3Args {
    int nx, ny, nz, ghost, kind;   // kind 0: the unit-spacing benchmark case, 1: Ma = 3 in SI units
    double dx, dy, dz, gamma;
    double rho_post, u_post, p_post;
    double cx, cy, cz, radius, rho_bubble;
    Quad quad;
};

__global__ __launch_bounds__(BLOCK) void k_shockbubble(double *__restrict__ q, Bubble3Args A) {
    const long long nxt = A.nx + 2 * A.ghost, nyt = A.ny + 2 * A.ghost, nzt = A.nz + 2 * A.ghost;
    const long long idx = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= nxt * nyt * nzt) return;
    const long long kp = idx / (nxt * nyt), rem = idx - kp * nxt * nyt;
    const long long jp = rem / nxt, ip = rem - jp * nxt;
    const double xc = ((double)(ip - A.ghost) + 0.5) * A.dx;
    const double yc = ((double)(jp - A.ghost) + 0.5) * A.dy;
    const double zc = ((double)(kp - A.ghost) + 0.5) * A.dz;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int gi = 0; gi < A.quad.n; ++gi) {
        const double x = xc + 0.5 * A.dx * A.quad.x[gi];
        for (int gj = 0; gj < A.quad.n; ++gj) {
            const double y = yc + 0.5 * A.dy * A.quad.x[gj];
            for (int gk = 0; gk < A.quad.n; ++gk) {
                const double z = zc + 0.5 * A.dz * A.quad.x[gk];
                const double w = A.quad.w[gi] * A.quad.w[gj] * A.quad.w[gk];
                double rho, u, p, energy;
                if (A.kind == 0) {
                    const bool shocked = x < 4.0;
                    const double bx = x - A.cx, by = y - A.cy, bz = z - A.cz;
                    const bool bubble = bx * bx + by * by + bz * bz <= A.radius * A.radius;
                    rho = shocked ? A.rho_post : (bubble ? A.rho_bubble : 1.0);
                    u = shocked ? A.u_post : 0.0;
                    p = shocked ? A.p_post : 1.0;
                    energy = 0.5 * rho * u * u + p / (A.gamma - 1.0);
                } else {
                    const double rho_air = 1.29, p_air = 101325.0;
                    const bool shocked = x < 0.005;
                    rho = shocked ? A.rho_post : rho_air;
                    u = shocked ? 736.911 : 0.0;
                    p = shocked ? A.p_post : p_air;
                    const double bx = x - 0.035, by = y - 0.0445, bz = z - 0.0445;
                    if (bx * bx + by * by + bz * bz <= 0.025 * 0.025) {
                        rho = A.rho_bubble;
                        u = 0.0;
                        p = p_air;
                    }
                    energy = 0.5 * rho * (u * u) + p / (A.gamma - 1.0);
                }
                const double c[5] = {rho, rho * u, 0.0, 0.0, energy};
                for (int k = 0; k < 5; ++k) acc[k] = acc[k] + w * c[k];
            }
        }
    }
    double *o = q + idx * 5;
    for (int k = 0; k < 5; ++k) o[k] = 0.125 * acc[k];
}

static int make_quad(const double *nodes, const double *weights, int n, Quad &quad) {
    if (!nodes || !weights || n < 1 || n > 16) return WSR_E_ARG;
    memset(&quad, 0, sizeof quad);
    for (int i = 0; i < n; ++i) {
        quad.x[i] = nodes[i];
        quad.w[i] = weights[i];
    }
    quad.n = n;
    return WSR_OK;
}

static unsigned blocks(long long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

}  // namespace init
}  // namespace wsr

using namespace wsr::init;

extern "C" {

int wsr_init_quadrant(double *u, int32_t nx, int32_t ny, int32_t ghost, int32_t wrap, const double *states,
                      double x_min, double x_max, double y_min, double y_max, double gamma, const double *nodes,
                      const double *weights, int32_t nq, void *stream) {
    if (!u || !states || nx < 1 || ny < 1 || ghost < 0) return WSR_E_ARG;
    QuadrantArgs A;
    memset(&A, 0, sizeof A);
    if (make_quad(nodes, weights, nq, A.quad)) return WSR_E_ARG;
    for (int r = 0; r < 4; ++r) A.states[r] = {states[4 * r], states[4 * r + 1], states[4 * r + 2], states[4 * r + 3]};
    A.nx = nx; A.ny = ny; A.ghost = ghost; A.wrap = wrap ? 1 : 0;
    A.x_min = x_min; A.y_min = y_min;
    A.dx = (x_max - x_min) / nx; A.dy = (y_max - y_min) / ny; A.gamma = gamma;
    const long long n = (long long)(nx + 2 * ghost) * (ny + 2 * ghost);
    hipLaunchKernelGGL(k_quadrant, dim3(blocks(n)), dim3(BLOCK), 0, (hipStream_t)stream, u, A);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_init_sod(double *u, int32_t nx, int32_t ny, int32_t ghost, int32_t axis, double x_length, double y_length,
                 double gamma, const double *nodes, const double *weights, int32_t nq, void *stream) {
    if (!u || nx < 1 || ny < 1 || ghost < 0 || (axis != 1 && axis != 2)) return WSR_E_ARG;
    SodArgs A;
    memset(&A, 0, sizeof A);
    if (make_quad(nodes, weights, nq, A.quad)) return WSR_E_ARG;
    A.nx = nx; A.ny = ny; A.ghost = ghost; A.axis = axis; A.gamma = gamma;
    A.origin = axis == 2 ? -0.5 * y_length : -0.5 * x_length;
    A.h = axis == 2 ? y_length / ny : x_length / nx;
    const long long n = (long long)(nx + 2 * ghost) * (ny + 2 * ghost);
    hipLaunchKernelGGL(k_sod, dim3(blocks(n)), dim3(BLOCK), 0, (hipStream_t)stream, u, A);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_init_double_mach(double *u, int32_t nx, int32_t ny, int32_t ghost, double x_length, double y_length,
                         double gamma, const double *nodes, const double *weights, int32_t nq, void *stream) {
    if (!u || nx < 1 || ny < 1 || ghost < 0) return WSR_E_ARG;
    DoubleMachArgs A;
    memset(&A, 0, sizeof A);
    if (make_quad(nodes, weights, nq, A.quad)) return WSR_E_ARG;
    A.nx = nx; A.ny = ny; A.ghost = ghost; A.gamma = gamma;
    A.dx = x_length / nx; A.dy = y_length / ny;
    const long long n = (long long)(nx + 2 * ghost) * (ny + 2 * ghost);
    hipLaunchKernelGGL(k_double_mach, dim3(blocks(n)), dim3(BLOCK), 0, (hipStream_t)stream, u, A);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_init_vortex(double *u, int32_t nx, int32_t ny, int32_t ghost, double t, double x_min, double x_max,
                    double y_min, double y_max, double gamma, const double *nodes, const double *weights, int32_t nq,
                    void *stream) {
    if (!u || nx < 1 || ny < 1 || ghost < 0) return WSR_E_ARG;
    VortexArgs A;
    memset(&A, 0, sizeof A);
    if (make_quad(nodes, weights, nq, A.quad)) return WSR_E_ARG;
    A.nx = nx; A.ny = ny; A.ghost = ghost; A.t = t; A.gamma = gamma;
    A.x_min = x_min; A.y_min = y_min;
    A.dx = (x_max - x_min) / (double)nx; A.dy = (y_max - y_min) / (double)ny;
    const long long n = (long long)(nx + 2 * ghost) * (ny + 2 * ghost);
    hipLaunchKernelGGL(k_vortex, dim3(blocks(n)), dim3(BLOCK), 0, (hipStream_t)stream, u, A);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_init_shockbubble3d(double *q, int32_t nx, int32_t ny, int32_t nz, int32_t ghost, int32_t kind,
                           const double *lengths, const double *shock, double density_ratio, double gamma,
                           const double *nodes, const double *weights, int32_t nq, void *stream) {
    if (!q || !lengths || !shock || nx < 1 || ny < 1 || nz < 1 || ghost < 0 || (kind != 0 && kind != 1))
        return WSR_E_ARG;
    Bubble3Args A;
    memset(&A, 0, sizeof A);
    if (make_quad(nodes, weights, nq, A.quad)) return WSR_E_ARG;
    A.nx = nx; A.ny = ny; A.nz = nz; A.ghost = ghost; A.kind = kind; A.gamma = gamma;
    A.dx = lengths[0] / nx; A.dy = lengths[1] / ny; A.dz = lengths[2] / nz;
    A.rho_post = shock[0]; A.u_post = shock[1]; A.p_post = shock[2];
    A.rho_bubble = density_ratio;
    A.radius = 0.25 * ny;
    A.cx = 0.5 * ny + 8.0; A.cy = 0.5 * ny; A.cz = 0.5 * nz;
    const long long n = (long long)(nx + 2 * ghost) * (ny + 2 * ghost) * (nz + 2 * ghost);
    hipLaunchKernelGGL(k_shockbubble, dim3(blocks(n)), dim3(BLOCK), 0, (hipStream_t)stream, q, A);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

}  // extern "C"
