// wsr_euler7.hip -- MI355X (gfx950) kernels and C ABI of the 2-D WENO7 step
// with Shu's fourth-order TVD Runge-Kutta method (downwind operator L~).
//
// One right-hand-side evaluation is a boundary fill (ghost 4) and one sweep
// launch that covers the x and the y tiles.  A sweep tile computes, per face and transverse line, the Roe
// average, the characteristic projection of the 8-cell window, both WENO7 face
// values (lr 1 / lr 2) and their back-projection into LDS (phase 1), then the
// 7-point transverse Gauss reconstruction of the eight face-value components
// (one closure for both Gauss points) and the Riemann fluxes (phase 2).  L and
// L~ share every reconstruction -- L~ only swaps the states entering the
// Riemann solver -- so a sweep that serves both writes F(ul, ur) and F(ur, ul)
// from the same face values: a Shu RK4 step costs four reconstructions
// (L,L~ of u; L,L~ of u1; L of u2; L of u3) instead of six.
//
// A step is: per stage [boundary, x+y sweep, stage combine]; the stage
// combines accumulate the RK4 sums in the reference's evaluation order.
//
// Reference (paths relative to the reference checkout):
//   RHS and stages: code/snapshot/weno7_point_rk4_shu/point_rk4.py:154-404
//   SR closure:     code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/
//                   point_rk4_mlp.py:55-170, warp_weno7_external_overwrite_fullall.py:64-230,
//                   beta_provider.py:12-219
//   double Mach:    code/snapshot/for_paper_results/solvers/weno7_double_mach.py:31-127,354-470
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/wsr.h"
#include "wsr_loop.hpp"
#include "wsr_weno7.hpp"

namespace wsr {
namespace w7 {

#ifndef WSR7_SPLIT
#define WSR7_SPLIT 2
#endif
#ifndef WSR7_WAVES
#define WSR7_WAVES 3
#endif
constexpr int GC = 4;
constexpr int BLOCK = 256;
constexpr int TN = 4;              // faces per tile along the sweep axis
constexpr int TT = 26;             // transverse flux lines per tile (phase 1 covers TT + 6)
constexpr int NT1 = TT + 6;
constexpr int NN = TN + 7;         // stencil cells per transverse line
constexpr int SPLIT = WSR7_SPLIT;  // threads per phase-1 face-line (2: twice the waves, half the passes)
constexpr int SBLOCK = NT1 * TN * SPLIT;
static_assert(SBLOCK % 64 == 0, "whole waves");
constexpr int N1 = NT1 * TN;
static_assert(N1 % (SBLOCK / 8) == 0, "phase 1b passes cover the face-lines exactly");
// waves per SIMD the register allocation aims for (the FP64 closure keeps the default)
template <int SCHEME> struct Sweep7Waves { static constexpr int value = SCHEME == SCHEME_SR64 ? 1 : WSR7_WAVES; };
constexpr int N2 = TT * TN;
constexpr int CH = SBLOCK / 2;
constexpr int FV_SIZE = NT1 * TN * 8;
constexpr int P1_SIZE = NN * NT1 * 4 + N1 * 4;
constexpr int P2_SIZE = CH * 16 + CH * 16;
constexpr int REG_SIZE = P1_SIZE > P2_SIZE ? P1_SIZE : P2_SIZE;

// ---------------------------------------------------------------- boundary
// point_rk4.py boundary kernels (ghost 4): transmissive clamp / periodic wrap;
// weno7_double_mach.py:63-127 on top of the transmissive copy.
__global__ __launch_bounds__(BLOCK) void k7_boundary(const double *__restrict__ src, double *__restrict__ dst, Grid g,
                                                     int boundary, double time_host, const Ctl *ctl,
                                                     double time_frac) {
    double time = time_host;
    if (ctl) {
        if (ctl->done) return;
        time = ctl->t + time_frac * ctl->dt;
    }
    int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= g.nxt * g.nyt) return;
    int j = idx / g.nxt, i = idx - j * g.nxt;
    bool inside = i >= GC && i < g.nx + GC && j >= GC && j < g.ny + GC;
    int si = i, sj = j;
    if (boundary == WSR_BC_PERIODIC) {
        if (i < GC) si = i + g.nx; else if (i >= g.nx + GC) si = i - g.nx;
        if (j < GC) sj = j + g.ny; else if (j >= g.ny + GC) sj = j - g.ny;
    } else {
        if (i < GC) si = GC; else if (i >= g.nx + GC) si = g.nx + GC - 1;
        if (j < GC) sj = GC; else if (j >= g.ny + GC) sj = g.ny + GC - 1;
    }
    const double2 *s2 = reinterpret_cast<const double2 *>(src + ((size_t)sj * g.nxt + si) * 4);
    double2 a = s2[0], b = s2[1];
    double q[4] = {a.x, a.y, b.x, b.y};
    if (boundary == WSR_BC_DOUBLE_MACH && !inside) {
        double x = ((double)(i - GC) + 0.5) * g.dx;
        double y = ((double)(j - GC) + 0.5) * g.dy;
        const double x0 = 1.0 / 6.0;
        if (j >= g.ny + GC) {
            dm_exact(x, y, time, g.gamma, q);
        } else if (j < GC) {
            if (x < x0) {
                dm_exact(x, 0.0, time, g.gamma, q);
            } else {
                int rj = 2 * GC - 1 - j;
                int ri = i < GC ? GC : (i >= g.nx + GC ? g.nx + GC - 1 : i);
                const double *r = src + ((size_t)rj * g.nxt + ri) * 4;
                q[0] = r[0]; q[1] = r[1]; q[2] = -r[2]; q[3] = r[3];
            }
        } else if (i < GC) {
            dm_exact(x, y, time, g.gamma, q);
        }
    }
    double2 *d2 = reinterpret_cast<double2 *>(dst + (size_t)idx * 4);
    d2[0] = make_double2(q[0], q[1]);
    d2[1] = make_double2(q[2], q[3]);
}

// The ghost cells of u in place (the reference's boundary kernels write the state they
// are given): only ghost entries are written, and they read interior cells only.
__global__ __launch_bounds__(BLOCK) void k7_ghosts_inplace(double *u, Grid g, int boundary, double time) {
    int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= g.nxt * g.nyt) return;
    int j = idx / g.nxt, i = idx - j * g.nxt;
    if (i >= GC && i < g.nx + GC && j >= GC && j < g.ny + GC) return;
    int si = i, sj = j;
    if (boundary == WSR_BC_PERIODIC) {
        if (i < GC) si = i + g.nx; else if (i >= g.nx + GC) si = i - g.nx;
        if (j < GC) sj = j + g.ny; else if (j >= g.ny + GC) sj = j - g.ny;
    } else {
        if (i < GC) si = GC; else if (i >= g.nx + GC) si = g.nx + GC - 1;
        if (j < GC) sj = GC; else if (j >= g.ny + GC) sj = g.ny + GC - 1;
    }
    const double *s = u + ((size_t)sj * g.nxt + si) * 4;
    double q[4] = {s[0], s[1], s[2], s[3]};
    if (boundary == WSR_BC_DOUBLE_MACH) {
        double x = ((double)(i - GC) + 0.5) * g.dx;
        double y = ((double)(j - GC) + 0.5) * g.dy;
        const double x0 = 1.0 / 6.0;
        if (j >= g.ny + GC) {
            dm_exact(x, y, time, g.gamma, q);
        } else if (j < GC) {
            if (x < x0) {
                dm_exact(x, 0.0, time, g.gamma, q);
            } else {
                int rj = 2 * GC - 1 - j;
                int ri = i < GC ? GC : (i >= g.nx + GC ? g.nx + GC - 1 : i);
                const double *r = u + ((size_t)rj * g.nxt + ri) * 4;
                q[0] = r[0]; q[1] = r[1]; q[2] = -r[2]; q[3] = r[3];
            }
        } else if (i < GC) {
            dm_exact(x, y, time, g.gamma, q);
        }
    }
    double *d = u + (size_t)idx * 4;
    d[0] = q[0]; d[1] = q[1]; d[2] = q[2]; d[3] = q[3];
}

// ------------------------------------------------------------------ sweep
struct SweepArgs7 {
    Grid g;
    double dt_host;
    const Ctl *ctl;
    unsigned long long *counters;
    unsigned long long *fallbacks;   // counter of values that fell back to WENO7-Z (with z_fallback)
    int eno;
    int z_fallback;
    int unit_gauss_h;   // WSR_FLAG_Z7_UNIT_GAUSS_H: transverse weights with h = 1
};

template <int DIR>
__device__ __forceinline__ void item_map(int it, int ntau, int &fl, int &tau) {
    if (DIR == 1) { fl = it % TN; tau = it / TN; }
    else { tau = it % ntau; fl = it / ntau; }
}

template <int DIR, int RIEMANN>
__device__ __forceinline__ void riemann7(const double ul[4], const double ur[4], double dt_h, double gamma,
                                         double f[4]) {
    if constexpr (RIEMANN == WSR_RIEMANN_HLLC) hllc_flux<DIR>(ul, ur, gamma, f);
    else if constexpr (RIEMANN == WSR_RIEMANN_FORCE) force_flux<DIR>(ul, ur, dt_h, gamma, f);
    else evilin_flux<DIR>(ul, ur, dt_h, gamma, f);
}

// F[face] = F(g1) + F(g2) with the states (left, right); with BOTH also
// Ft[face] = F(right, left) at both points (the downwind operator).
// One 4 x 26 face tile of the DIR sweep (tile coordinates bx along the sweep axis, by
// across it); smem is the block's LDS region of FV_SIZE + REG_SIZE doubles.
template <int DIR, int SCHEME, int RIEMANN, bool BOTH>
__device__ __forceinline__ void sweep7_tile(const double *__restrict__ bc, double *__restrict__ F,
                                            double *__restrict__ Ft, const SweepArgs7 &A,
                                            const typename Mlp7Sweep<SCHEME>::type &mlp, double dt, double *smem,
                                            float *sA, int bx, int by) {
    const Grid &g = A.g;
    const double h_n = DIR == 1 ? g.dx : g.dy;
    const double h_t = A.unit_gauss_h ? 1.0 : (DIR == 1 ? g.dy : g.dx);
    const double dt_h = dt / h_n;
    const int nfaces = (DIR == 1 ? g.nx : g.ny) + 1;
    const int ntrans = DIR == 1 ? g.ny : g.nx;
    const int f0 = bx * TN;
    const int t0 = by * TT;
    const bool zf = A.z_fallback != 0;

    double *FV = smem;
    double *T = smem + FV_SIZE;
    double *EIG = T + NN * NT1 * 4;
    double *GV = smem + FV_SIZE;
    double *FL = GV + CH * 16;

    // tile load: cell n of line tau is padded (t0 + 1 + tau, f0 + n) (sweep axis last)
    for (int e = threadIdx.x; e < NN * NT1; e += SBLOCK) {
        int n, tau, prow, pcol;
        if (DIR == 1) { n = e % NN; tau = e / NN; prow = t0 + 1 + tau; pcol = f0 + n; }
        else { tau = e % NT1; n = e / NT1; prow = f0 + n; pcol = t0 + 1 + tau; }
        prow = prow < g.nyt ? prow : g.nyt - 1;
        pcol = pcol < g.nxt ? pcol : g.nxt - 1;
        const double2 *s2 = reinterpret_cast<const double2 *>(bc + ((size_t)prow * g.nxt + pcol) * 4);
        double2 a = s2[0], b = s2[1];
        double *d = T + (n * NT1 + tau) * 4;
        d[0] = a.x; d[1] = a.y; d[2] = b.x; d[3] = b.y;
    }
    // the FP32 closure's matrix-core weights (A registers) into LDS
    if constexpr (SCHEME == SCHEME_SR32)
        for (int e = threadIdx.x; e < W7M_WORDS; e += SBLOCK) sA[e] = mlp.p[e];
    __syncthreads();
    const Mfma7A wa{sA, (int)(threadIdx.x & 63)};

    // phase 1a: Roe average of the two cells next to each face (cells n = fl+3, fl+4)
    if (threadIdx.x < N1) {
        int fl, tau;
        item_map<DIR>(threadIdx.x, NT1, fl, tau);
        Eigen E = roe_eigen(T + ((fl + 3) * NT1 + tau) * 4, T + ((fl + 4) * NT1 + tau) * 4, g.gamma);
        double *e = EIG + threadIdx.x * 4;
        e[0] = E.u; e[1] = E.v; e[2] = E.a; e[3] = E.h;
    }
    __syncthreads();

    const int lane = threadIdx.x;
    const int gi = lane & 7;
    const int s = gi >> 2;    // 0: left state of the face (lr 1 of the cell before), 1: right state (lr 2)
    const int k = gi & 3;
    const int slot = lane >> 3;
    LaneCounts cnt;
    unsigned nfell = 0;

    // phase 1b: characteristic projection, WENO7 face value, back-projection
    for (int base = 0; base < N1; base += SBLOCK / 8) {
        const int it = base + slot;
        int fl, tau;
        item_map<DIR>(it, NT1, fl, tau);
        const double *e = EIG + it * 4;
        Eigen E;
        E.u = e[0]; E.v = e[1]; E.a = e[2]; E.h = e[3];
        double l[4], scale;
        left_row<DIR>(k, E, g.gamma, l, scale);
        double c[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) {
            const double *q = T + ((fl + s + m) * NT1 + tau) * 4;
            c[m] = (l[0] * q[0] + l[1] * q[1] + l[2] * q[2] + l[3] * q[3]) * scale;
        }
        bool closure, trip;
        double val;
        if constexpr (SCHEME == SCHEME_SR32) val = recon7_normal_mfma(c, s, wa, A.eno, h_n, zf, closure, trip);
        else val = s == 0 ? recon7_normal<SCHEME, 1>(c, mlp, A.eno, h_n, zf, closure, trip)
                          : recon7_normal<SCHEME, 2>(c, mlp, A.eno, h_n, zf, closure, trip);
        if (A.counters) cnt.add(0, 4, true, closure);
        if (zf) {
            // the four lanes of one side hold the four characteristic components of one state
            const unsigned long long tb = __ballot(trip);
            if ((tb >> (lane & 60)) & 0xFull) {   // the group's bits in the wave mask
                val = s == 0 ? z7_value<1>(c, h_n) : z7_value<2>(c, h_n);
                // count each used state once: faces of the grid, and the face-lines this
                // tile owns (line L = t0 - 3 + tau in [-3, ntrans + 3); a halo line belongs
                // to the tile whose core holds L clamped to the grid)
                const int L = t0 - 3 + tau;
                const int Lc = L < 0 ? 0 : (L > ntrans - 1 ? ntrans - 1 : L);
                const bool owned = (f0 + fl) < nfaces && L < ntrans + 3 && Lc >= t0 && Lc < t0 + TT;
                nfell += (owned && k == 0) ? 1u : 0u;
            }
        }
        const int qb = lane & ~3;
        double cf0 = __shfl(val, qb + 0), cf1 = __shfl(val, qb + 1);
        double cf2 = __shfl(val, qb + 2), cf3 = __shfl(val, qb + 3);
        double r[4];
        right_row<DIR>(k, E, r);
        FV[(tau * TN + fl) * 8 + gi] = r[0] * cf0 + r[1] * cf1 + r[2] * cf2 + r[3] * cf3;
    }
    __syncthreads();

    // phase 2: transverse Gauss reconstruction, then the Riemann fluxes
    for (int c0 = 0; c0 < N2; c0 += CH) {
        for (int base = 0; base < CH; base += SBLOCK / 8) {
            const int ci = base + slot;
            const int it = c0 + ci;
            // items past the tile's N2 faces (whole waves of the last pass at 4 x 26)
            // have no face: no reconstruction, no flux.  The skip is decided per wave (a
            // uniform branch keeps EXEC full for the matrix-core closure); a wave with some
            // items past the faces (none at 4 x 26) reconstructs a valid item in their lanes
            // and drops the results.
            const bool valid = it < N2;
            if (__ballot(valid) == 0ull) continue;
            int fl, tl;
            item_map<DIR>(valid ? it : N2 - 1, TT, fl, tl);
            double q[7];
#pragma unroll
            for (int m = 0; m < 7; ++m) q[m] = FV[((tl + m) * TN + fl) * 8 + gi];
            double g1, g2;
            bool closure, trip1, trip2;
            if constexpr (SCHEME == SCHEME_SR32) recon7_gauss_mfma(q, wa, A.eno, h_t, zf, g1, g2, closure, trip1, trip2);
            else recon7_gauss<SCHEME>(q, mlp, A.eno, h_t, zf, g1, g2, closure, trip1, trip2);
            if (A.counters) cnt.add(2, 5, valid && (f0 + fl) < nfaces && (t0 + tl) < ntrans, valid && closure);
            if (zf) {
                // lanes of a group of four hold one side's four conservative components
                // (items past the tile skip whole groups of eight lanes)
                const unsigned long long b1 = __ballot(trip1), b2 = __ballot(trip2);
                const bool z1 = (b1 >> (lane & 60)) & 0xFull, z2 = (b2 >> (lane & 60)) & 0xFull;
                if (z1) g1 = z7_value<4>(q, h_t);
                if (z2) g2 = z7_value<3>(q, h_t);
                // states of faces inside the grid (the tile's edge lines past it are not counted)
                if (valid && k == 0 && (f0 + fl) < nfaces && (t0 + tl) < ntrans)
                    nfell += (z1 ? 1u : 0u) + (z2 ? 1u : 0u);
            }
            if (valid) {
                GV[ci * 16 + gi] = g1;
                GV[ci * 16 + 8 + gi] = g2;
            }
        }
        __syncthreads();
        if (c0 + (int)(threadIdx.x >> 1) < N2) {
            const int ci = threadIdx.x >> 1, gp = threadIdx.x & 1;
            const double *v = GV + ci * 16 + gp * 8;
            double ul[4] = {v[0], v[1], v[2], v[3]};
            double ur[4] = {v[4], v[5], v[6], v[7]};
            double f[4];
            riemann7<DIR, RIEMANN>(ul, ur, dt_h, g.gamma, f);
            double *o = FL + ci * 16 + gp * 4;
            o[0] = f[0]; o[1] = f[1]; o[2] = f[2]; o[3] = f[3];
            if constexpr (BOTH) {
                riemann7<DIR, RIEMANN>(ur, ul, dt_h, g.gamma, f);
                o += 8;
                o[0] = f[0]; o[1] = f[1]; o[2] = f[2]; o[3] = f[3];
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < CH * 4; e += SBLOCK) {
            const int ci = e >> 2, kk = e & 3;
            const int it = c0 + ci;
            if (it >= N2) continue;
            int fl, tl;
            item_map<DIR>(it, TT, fl, tl);
            const int face = f0 + fl, tt = t0 + tl;
            if (face < nfaces && tt < ntrans) {
                size_t o = DIR == 1 ? ((size_t)tt * (g.nx + 1) + face) * 4 + kk : ((size_t)face * g.nx + tt) * 4 + kk;
                F[o] = FL[ci * 16 + kk] + FL[ci * 16 + 4 + kk];
                if constexpr (BOTH) Ft[o] = FL[ci * 16 + 8 + kk] + FL[ci * 16 + 12 + kk];
            }
        }
        __syncthreads();
    }
    if (A.counters) cnt.flush(A.counters);
    if (A.fallbacks) {
        for (int off = 32; off > 0; off >>= 1) nfell += __shfl_down(nfell, off);
        if ((threadIdx.x & 63) == 0 && nfell) atomicAdd(A.fallbacks, (unsigned long long)nfell);
    }
}

// Both sweeps of a right-hand side in one launch: workgroups [0, gx*tx) take the x
// tiles (gx along x, tx across), the rest the y tiles; they read the same
// boundary-filled state and write separate flux arrays.
template <int SCHEME, int RIEMANN, bool BOTH>
__global__ __launch_bounds__(SBLOCK) __attribute__((amdgpu_waves_per_eu(Sweep7Waves<SCHEME>::value))) void k7_sweep_xy(const double *__restrict__ bc, double *__restrict__ Fx,
                                                      double *__restrict__ Fxt, double *__restrict__ Fy,
                                                      double *__restrict__ Fyt, SweepArgs7 A,
                                                      typename Mlp7Sweep<SCHEME>::type mlp) {
    double dt = A.dt_host;
    if (A.ctl) {
        if (A.ctl->done) return;
        dt = A.ctl->dt;
    }
    __shared__ double smem[FV_SIZE + REG_SIZE];
    __shared__ float sA[SCHEME == SCHEME_SR32 ? W7M_WORDS : 1];
    const int gx = (A.g.nx + 1 + TN - 1) / TN, tx = (A.g.ny + TT - 1) / TT;
    const int gy = (A.g.ny + 1 + TN - 1) / TN;
    const int id = blockIdx.x;
    if (id < gx * tx) {
        sweep7_tile<1, SCHEME, RIEMANN, BOTH>(bc, Fx, Fxt, A, mlp, dt, smem, sA, id % gx, id / gx);
    } else {
        const int iy = id - gx * tx;
        sweep7_tile<2, SCHEME, RIEMANN, BOTH>(bc, Fy, Fyt, A, mlp, dt, smem, sA, iy % gy, iy / gy);
    }
}

// ------------------------------------------------------------------ stages
// rhs = -0.5 (Fx_{i+1} - Fx_i) / dx - 0.5 (Fy_{j+1} - Fy_j) / dy (point_rk4.py:286-302)
__device__ __forceinline__ double rhs_at(const double *Fx, const double *Fy, const Grid &g, int j, int i, int c,
                                         double inv_dx, double inv_dy) {
    double fxl = Fx[((size_t)j * (g.nx + 1) + i) * 4 + c];
    double fxr = Fx[((size_t)j * (g.nx + 1) + i + 1) * 4 + c];
    double fyb = Fy[((size_t)j * g.nx + i) * 4 + c];
    double fyt = Fy[((size_t)(j + 1) * g.nx + i) * 4 + c];
    return -0.5 * (fxr - fxl) * inv_dx - 0.5 * (fyt - fyb) * inv_dy;
}

struct StageArrays {
    double *u, *u1, *u2, *u3, *rt0, *acc;
    const double *Fx, *Fy, *Fxt, *Fyt;
    double *dts;
};

// Shu RK4 (point_rk4.py:314-404), four combines:
//  A (L u, L~ u):   u1 = u + dt/2 r0; acc = u/5 + dt/10 r0; rt0 kept
//  B (L u1, L~ u1): u2 = 649/1600 u - c dt rt0 + 951/1600 u1 + 5000/7873 dt r1;
//                   acc = ((acc + 6127/30000 u1) + dt/6 r1) + 7873/30000 u2;
//                   u3 = ((53989/2500000 u - .. dt rt0) + .. u1) - 5121/20000 dt rt1
//  C (L u2):        u3 = (u3 + 23619/32000 u2) + 7873/10000 dt r2; acc = acc + u3/3
//  D (L u3):        u = acc + dt/6 r3 (and the next step's max speed)
template <int STAGE>
__global__ __launch_bounds__(BLOCK) void k7_stage(StageArrays S, Grid g, double dt_host, Ctl *ctl, int fold_speed) {
    double dt = dt_host;
    if (ctl) {
        if (ctl->done) return;
        dt = ctl->dt;
    }
    int idx = blockIdx.x * BLOCK + threadIdx.x;
    const bool active = idx < g.nx * g.ny;
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    if (active) {
        int j = idx / g.nx, i = idx - j * g.nx;
        size_t c = ((size_t)(j + GC) * g.nxt + (i + GC)) * 4;
        const double inv_dx = 1.0 / g.dx, inv_dy = 1.0 / g.dy;
        // every input is loaded (as 16-byte halves of the 32-byte cell and face
        // records) before the first store: the stage arrays are distinct, but the
        // compiler cannot know it and would otherwise wait out each component's
        // stores before the next component's loads.  The multiply-adds are spelled
        // out so that the rounding does not depend on the compiler's choice of the
        // product to fuse (a sum of two products fuses the second, a difference the
        // first).
        auto ld4 = [](const double *p, double v[4]) {
            const double2 a = reinterpret_cast<const double2 *>(p)[0], b = reinterpret_cast<const double2 *>(p)[1];
            v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
        };
        auto st4 = [](double *p, const double v[4]) {
            reinterpret_cast<double2 *>(p)[0] = make_double2(v[0], v[1]);
            reinterpret_cast<double2 *>(p)[1] = make_double2(v[2], v[3]);
        };
        // rhs_at for the four components of cell (j, i)
        auto rhs4 = [&](const double *Fx, const double *Fy, double out[4]) {
            double xl[4], xr[4], yb[4], yt[4];
            ld4(Fx + ((size_t)j * (g.nx + 1) + i) * 4, xl);
            ld4(Fx + ((size_t)j * (g.nx + 1) + i + 1) * 4, xr);
            ld4(Fy + ((size_t)j * g.nx + i) * 4, yb);
            ld4(Fy + ((size_t)(j + 1) * g.nx + i) * 4, yt);
#pragma unroll
            for (int k = 0; k < 4; ++k) out[k] = fma(-0.5 * (xr[k] - xl[k]), inv_dx, -0.5 * (yt[k] - yb[k]) * inv_dy);
        };
        double r[4];
        rhs4(S.Fx, S.Fy, r);
        if constexpr (STAGE == 0) {
            double rt[4], u[4], u1[4], acc[4];
            rhs4(S.Fxt, S.Fyt, rt);
            ld4(S.u + c, u);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                u1[k] = fma(0.5 * dt, r[k], u[k]);
                acc[k] = fma(1.0 / 10.0 * dt, r[k], 1.0 / 5.0 * u[k]);
            }
            st4(S.u1 + c, u1);
            st4(S.acc + c, acc);
            st4(S.rt0 + c, rt);
        } else if constexpr (STAGE == 1) {
            double rt1[4], u[4], u1[4], rt0[4], acc[4], u2[4], u3[4];
            rhs4(S.Fxt, S.Fyt, rt1);
            ld4(S.u + c, u);
            ld4(S.u1 + c, u1);
            ld4(S.rt0 + c, rt0);
            ld4(S.acc + c, acc);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                u2[k] = fma(5000.0 / 7873.0 * dt, r[k],
                            fma(951.0 / 1600.0, u1[k], fma(649.0 / 1600.0, u[k], -(10890423.0 / 25193600.0 * dt * rt0[k]))));
                acc[k] = fma(7873.0 / 30000.0, u2[k], fma(1.0 / 6.0 * dt, r[k], fma(6127.0 / 30000.0, u1[k], acc[k])));
                u3[k] = fma(-(5121.0 / 20000.0 * dt), rt1[k],
                            fma(4806213.0 / 20000000.0, u1[k], fma(53989.0 / 2500000.0, u[k], -(102261.0 / 5000000.0 * dt * rt0[k]))));
            }
            st4(S.u2 + c, u2);
            st4(S.acc + c, acc);
            st4(S.u3 + c, u3);
        } else if constexpr (STAGE == 2) {
            double u3[4], u2[4], acc[4];
            ld4(S.u3 + c, u3);
            ld4(S.u2 + c, u2);
            ld4(S.acc + c, acc);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                u3[k] = fma(7873.0 / 10000.0 * dt, r[k], fma(23619.0 / 32000.0, u2[k], u3[k]));
                acc[k] = fma(1.0 / 3.0, u3[k], acc[k]);
            }
            st4(S.u3 + c, u3);
            st4(S.acc + c, acc);
        } else {
            double acc[4];
            ld4(S.acc + c, acc);
#pragma unroll
            for (int k = 0; k < 4; ++k) q[k] = fma(1.0 / 6.0 * dt, r[k], acc[k]);
            st4(S.u + c, q);
        }
    }
    if constexpr (STAGE == 3) {
        if (fold_speed) {
            reduce_speed<BLOCK>(active ? cell_speed(q, g.gamma) : 0.0, active, ctl);
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                double d = ctl->dt;
                S.dts[ctl->step] = d;
                ctl->t = ctl->t + d;
                ctl->dt_last = d;
                ctl->step = ctl->step + 1;
            }
        }
    }
}

// the interior of L(u) (and L~(u)) into padded arrays with zero ghost cells
__global__ __launch_bounds__(BLOCK) void k7_rhs_out(const double *__restrict__ Fx, const double *__restrict__ Fy,
                                                    const double *__restrict__ Fxt, const double *__restrict__ Fyt,
                                                    double *rhs, double *rhs_t, Grid g) {
    int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= g.nxt * g.nyt) return;
    int j = idx / g.nxt - GC, i = idx % g.nxt - GC;
    bool inside = i >= 0 && i < g.nx && j >= 0 && j < g.ny;
    const double inv_dx = 1.0 / g.dx, inv_dy = 1.0 / g.dy;
    for (int comp = 0; comp < 4; ++comp) {
        rhs[(size_t)idx * 4 + comp] = inside ? rhs_at(Fx, Fy, g, j, i, comp, inv_dx, inv_dy) : 0.0;
        if (rhs_t) rhs_t[(size_t)idx * 4 + comp] = inside ? rhs_at(Fxt, Fyt, g, j, i, comp, inv_dx, inv_dy) : 0.0;
    }
}

__global__ __launch_bounds__(BLOCK) void k7_max_speed(const double *__restrict__ u, Grid g, Ctl *ctl) {
    int idx = blockIdx.x * BLOCK + threadIdx.x;
    const bool active = idx < g.nx * g.ny;
    double sp = 0.0;
    if (active) {
        int j = idx / g.nx, i = idx - j * g.nx;
        const double *c = u + ((size_t)(j + GC) * g.nxt + (i + GC)) * 4;
        double q[4] = {c[0], c[1], c[2], c[3]};
        sp = cell_speed(q, g.gamma);
    }
    reduce_speed<BLOCK>(sp, active, ctl);
}

// counters cover one wsr_euler7_rhs / wsr_euler7_step call (a kernel, not a memset:
// the call may be captured into a graph)
__global__ void k7_zero_counters(Ctl *ctl) {
    if (threadIdx.x < NCOUNT) ctl->counters[threadIdx.x] = 0ull;
}

__global__ void k7_reset_speed(Ctl *ctl) {
    ctl->speed_bits = 0ull;
    ctl->nan_flag = 0;
}

__global__ void k7_loop_init(Ctl *ctl, double t_end, double cfl_h, int max_steps) {
    ctl->t = 0.0; ctl->dt = 0.0; ctl->t_end = t_end; ctl->cfl_h = cfl_h; ctl->dt_last = 0.0;
    ctl->speed_bits = 0ull; ctl->step = 0; ctl->done = 0; ctl->nan_flag = 0; ctl->max_steps = max_steps;
    for (int i = 0; i < NCOUNT; ++i) ctl->counters[i] = 0ull;
}

// dt of the next step.  mode 0: point_rk4.py:557-563 (loop while t < t_end,
// dt clipped to t_end - t, stop when dt <= 0); mode 1: weno7_double_mach.py:
// 453-460 (loop while t < t_end - 1e-14, dt = min(dt_cfl, t_end - t), stop on
// a non-finite or non-positive dt).  A non-finite max speed ends both loops.
__global__ void k7_dt(Ctl *ctl, int mode) {
    if (ctl->done) return;
    double t = ctl->t;
    bool go = mode == 0 ? (t < ctl->t_end) : (t < ctl->t_end - 1.0e-14);
    if (!go || ctl->step >= ctl->max_steps) { ctl->done = 1; return; }
    double dt;
    if (ctl->nan_flag) {
        dt = __longlong_as_double(0x7ff8000000000000ll);
    } else {
        double max_speed = __longlong_as_double((long long)ctl->speed_bits);
        dt = max_speed < 1.0e-15 ? 1.0e10 : ctl->cfl_h / max_speed;
    }
    if (mode == 0) {
        if (t + dt > ctl->t_end) dt = ctl->t_end - t;
    } else {
        double rem = ctl->t_end - t;
        dt = rem < dt ? rem : dt;
    }
    if (!isfinite(dt) || dt <= 0.0) { ctl->done = 1; return; }
    ctl->dt = dt;
    ctl->speed_bits = 0ull;
    ctl->nan_flag = 0;
}

// parameters travel by value in launches of at most PCHUNK words (kernel
// arguments stay well under the 4 KiB argument limit)
constexpr int PCHUNK = 448;
struct ParamChunk { uint32_t w[PCHUNK]; };

__global__ void k7_put_params(ParamChunk src, uint32_t *dst, int nwords) {
    for (int i = threadIdx.x; i < nwords; i += blockDim.x) dst[i] = src.w[i];
}

// ----------------------------------------------------------- host side
struct Workspace7 {
    Ctl *ctl;
    void *mlp;
    double *bc, *Fx, *Fy, *Fxt, *Fyt, *u1, *u2, *u3, *rt0, *acc;
};

static size_t ws7_layout(int nx, int ny, Workspace7 *w, void *base) {
    size_t np = (size_t)(nx + 2 * GC) * (ny + 2 * GC) * 4 * sizeof(double);
    size_t nfx = (size_t)ny * (nx + 1) * 4 * sizeof(double);
    size_t nfy = (size_t)(ny + 1) * nx * 4 * sizeof(double);
    size_t off = 0;
    char *b = (char *)base;
    auto take = [&](size_t bytes) { char *p = b + off; off += align256(bytes); return (void *)p; };
    void *ctl = take(sizeof(Ctl));
    // FP64 closure: the neuron-major block; FP32: the matrix-core A table
    void *mlp = take(MLP7_COUNT * sizeof(double) > W7M_WORDS * sizeof(float) ? MLP7_COUNT * sizeof(double)
                                                                             : W7M_WORDS * sizeof(float));
    void *bc = take(np), *fx = take(nfx), *fy = take(nfy), *fxt = take(nfx), *fyt = take(nfy);
    void *u1 = take(np), *u2 = take(np), *u3 = take(np), *rt0 = take(np), *acc = take(np);
    if (w && base) {
        w->ctl = (Ctl *)ctl; w->mlp = mlp; w->bc = (double *)bc;
        w->Fx = (double *)fx; w->Fy = (double *)fy; w->Fxt = (double *)fxt; w->Fyt = (double *)fyt;
        w->u1 = (double *)u1; w->u2 = (double *)u2; w->u3 = (double *)u3; w->rt0 = (double *)rt0;
        w->acc = (double *)acc;
    }
    return off;
}

static int check7(const wsr_euler7_config *cfg) {
    if (!cfg) return WSR_E_ARG;
    if (cfg->ghost != GC || cfg->nx < 1 || cfg->ny < 1 || cfg->nx > (1 << 15) || cfg->ny > (1 << 15))
        return WSR_E_ARG;
    if (cfg->scheme != WSR_SCHEME_JS && cfg->scheme != WSR_SCHEME_SR64 && cfg->scheme != WSR_SCHEME_SR32 &&
        cfg->scheme != WSR_SCHEME_Z)
        return WSR_E_ARG;
    if (cfg->riemann != WSR_RIEMANN_HLLC && cfg->riemann != WSR_RIEMANN_FORCE && cfg->riemann != WSR_RIEMANN_EVILIN)
        return WSR_E_ARG;
    if (cfg->boundary < WSR_BC_TRANSMISSIVE || cfg->boundary > WSR_BC_DOUBLE_MACH) return WSR_E_ARG;
    if (cfg->loop_mode != WSR_LOOP7_POINT && cfg->loop_mode != WSR_LOOP7_DOUBLE_MACH) return WSR_E_ARG;
    bool sr = cfg->scheme == WSR_SCHEME_SR64 || cfg->scheme == WSR_SCHEME_SR32;
    if (sr && !cfg->mlp) return WSR_E_MLP;
    return WSR_OK;
}

static Grid grid7(const wsr_euler7_config *cfg) {
    Grid g;
    g.nx = cfg->nx; g.ny = cfg->ny; g.nxt = cfg->nx + 2 * GC; g.nyt = cfg->ny + 2 * GC;
    g.dx = cfg->dx; g.dy = cfg->dy; g.gamma = cfg->gamma;
    return g;
}

static void enqueue_params7(const wsr_euler7_config *cfg, const Workspace7 &w, hipStream_t st) {
    if (cfg->scheme != WSR_SCHEME_SR64 && cfg->scheme != WSR_SCHEME_SR32) return;
    static_assert(W7M_WORDS <= MLP7_COUNT * 2 + 64, "the word buffer holds either parameter block");
    uint32_t words[MLP7_COUNT * 2 + 64];
    int nwords;
    if (cfg->scheme == WSR_SCHEME_SR64) {
        double dev[MLP7_COUNT];
        pack7_neuron_major((const double *)cfg->mlp, dev);
        memcpy(words, dev, sizeof dev);
        nwords = MLP7_COUNT * 2;
    } else {
        // the FP32 closure runs on the matrix cores: folded parameters as the A table
        float dev[MLP7_COUNT], table[W7M_WORDS];
        pack7_neuron_major((const float *)cfg->mlp, dev);
        fold7_fp32(dev);
        mfma7_table(dev, table);
        memcpy(words, table, sizeof table);
        nwords = W7M_WORDS;
    }
    for (int o = 0; o < nwords; o += PCHUNK) {
        ParamChunk c;
        memset(&c, 0, sizeof c);
        const int n = nwords - o < PCHUNK ? nwords - o : PCHUNK;
        memcpy(c.w, words + o, (size_t)n * sizeof(uint32_t));
        hipLaunchKernelGGL(k7_put_params, dim3(1), dim3(256), 0, st, c, (uint32_t *)w.mlp + o, n);
    }
}

template <int SCHEME, int RIEMANN>
static void launch7_t(const double *bc, double *Fx, double *Fxt, double *Fy, double *Fyt, const SweepArgs7 &A,
                      const void *mlp_dev, hipStream_t st) {
    typename Mlp7Sweep<SCHEME>::type mlp;
    memset(&mlp, 0, sizeof mlp);
    if constexpr (SCHEME == SCHEME_SR64) mlp.p = (const double *)mlp_dev;
    else if constexpr (SCHEME == SCHEME_SR32) mlp.p = (const float *)mlp_dev;
    const int gx = (A.g.nx + 1 + TN - 1) / TN, gy = (A.g.ny + 1 + TN - 1) / TN;
    const int tx = (A.g.ny + TT - 1) / TT, ty = (A.g.nx + TT - 1) / TT;
    dim3 grid(gx * tx + gy * ty);
    if (Fxt)
        hipLaunchKernelGGL((k7_sweep_xy<SCHEME, RIEMANN, true>), grid, dim3(SBLOCK), 0, st, bc, Fx, Fxt, Fy, Fyt, A,
                           mlp);
    else
        hipLaunchKernelGGL((k7_sweep_xy<SCHEME, RIEMANN, false>), grid, dim3(SBLOCK), 0, st, bc, Fx, Fxt, Fy, Fyt, A,
                           mlp);
}

// x and y sweeps of one right-hand side (with BOTH: also the downwind fluxes Fxt, Fyt)
static void launch7(const wsr_euler7_config *cfg, const double *bc, double *Fx, double *Fxt, double *Fy, double *Fyt,
                    const SweepArgs7 &A, const void *mlp_dev, hipStream_t st) {
#define WSR7_CASE(S)                                                                                               \
    case S:                                                                                                        \
        if (cfg->riemann == WSR_RIEMANN_FORCE) launch7_t<S, WSR_RIEMANN_FORCE>(bc, Fx, Fxt, Fy, Fyt, A, mlp_dev, st); \
        else if (cfg->riemann == WSR_RIEMANN_EVILIN)                                                               \
            launch7_t<S, WSR_RIEMANN_EVILIN>(bc, Fx, Fxt, Fy, Fyt, A, mlp_dev, st);                               \
        else launch7_t<S, WSR_RIEMANN_HLLC>(bc, Fx, Fxt, Fy, Fyt, A, mlp_dev, st);                                \
        break;
    switch (cfg->scheme) {
        WSR7_CASE(SCHEME_JS)
        WSR7_CASE(SCHEME_SR64)
        WSR7_CASE(SCHEME_SR32)
        WSR7_CASE(SCHEME_Z)
    }
#undef WSR7_CASE
}

// boundary fill of src at time (host value, or ctl->t + frac * dt) and both sweeps
// optional HIP-event pairs around the sweep launches (kind 0: L and L~ sweep, 1: L only)
constexpr int TIMER7_MAX_STEPS = 32;
struct Timer7 {
    hipEvent_t ev[2 * 8 * TIMER7_MAX_STEPS];
    int kind[8 * TIMER7_MAX_STEPS];
    int used = 0, cap = 0;
    void begin(hipStream_t st) {
        if (used < cap) (void)hipEventRecord(ev[2 * used], st);
    }
    void end(int k, hipStream_t st) {
        if (used < cap) {
            (void)hipEventRecord(ev[2 * used + 1], st);
            kind[used] = k;
            used += 1;
        }
    }
};

static void enqueue_rhs7(const wsr_euler7_config *cfg, const Workspace7 &w, const double *src, bool both,
                         double dt, double time, const Ctl *ctl, double frac, hipStream_t st, Timer7 *tm = nullptr) {
    Grid g = grid7(cfg);
    const int np = g.nxt * g.nyt;
    hipLaunchKernelGGL(k7_boundary, dim3((np + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, src, w.bc, g,
                       cfg->boundary, time, ctl, frac);
    SweepArgs7 A;
    A.g = g; A.dt_host = dt; A.ctl = ctl; A.eno = cfg->eno_cutoff;
    A.z_fallback = (cfg->flags & WSR_FLAG_Z_FALLBACK) ? 1 : 0;
    A.unit_gauss_h = (cfg->flags & WSR_FLAG_Z7_UNIT_GAUSS_H) ? 1 : 0;
    A.counters = (cfg->flags & WSR_FLAG_COUNT_CLOSURE) ? w.ctl->counters : nullptr;
    A.fallbacks = A.z_fallback ? w.ctl->counters + 6 : nullptr;
    if (tm) tm->begin(st);
    launch7(cfg, w.bc, w.Fx, both ? w.Fxt : nullptr, w.Fy, both ? w.Fyt : nullptr, A, w.mlp, st);
    if (tm) tm->end(both ? 0 : 1, st);
}

// one Shu RK4 step of u (host dt/time, or the device-resident ones when ctl != nullptr)
static void enqueue_step7(const wsr_euler7_config *cfg, const Workspace7 &w, double *u, double dt, double time,
                          Ctl *ctl, double *dts, hipStream_t st, Timer7 *tm = nullptr) {
    Grid g = grid7(cfg);
    const int nc = g.nx * g.ny;
    dim3 grid((nc + BLOCK - 1) / BLOCK);
    StageArrays S;
    S.u = u; S.u1 = w.u1; S.u2 = w.u2; S.u3 = w.u3; S.rt0 = w.rt0; S.acc = w.acc;
    S.Fx = w.Fx; S.Fy = w.Fy; S.Fxt = w.Fxt; S.Fyt = w.Fyt; S.dts = dts;
    enqueue_rhs7(cfg, w, u, true, dt, time, ctl, 0.0, st, tm);
    hipLaunchKernelGGL(k7_stage<0>, grid, dim3(BLOCK), 0, st, S, g, dt, ctl, 0);
    enqueue_rhs7(cfg, w, w.u1, true, dt, time + 0.5 * dt, ctl, 0.5, st, tm);
    hipLaunchKernelGGL(k7_stage<1>, grid, dim3(BLOCK), 0, st, S, g, dt, ctl, 0);
    enqueue_rhs7(cfg, w, w.u2, false, dt, time + 0.5 * dt, ctl, 0.5, st, tm);
    hipLaunchKernelGGL(k7_stage<2>, grid, dim3(BLOCK), 0, st, S, g, dt, ctl, 0);
    enqueue_rhs7(cfg, w, w.u3, false, dt, time + dt, ctl, 1.0, st, tm);
    hipLaunchKernelGGL(k7_stage<3>, grid, dim3(BLOCK), 0, st, S, g, dt, ctl, ctl ? 1 : 0);
}

}  // namespace w7
}  // namespace wsr

using namespace wsr;
using namespace wsr::w7;

extern "C" {

size_t wsr_euler7_workspace_bytes(const wsr_euler7_config *cfg) {
    if (check7(cfg)) return 0;
    return ws7_layout(cfg->nx, cfg->ny, nullptr, nullptr);
}

int wsr_euler7_rhs(const wsr_euler7_config *cfg, const double *u, double *rhs, double *rhs_reverse, double dt,
                   double time, void *workspace, size_t workspace_bytes, void *stream) {
    int rc = check7(cfg);
    if (rc) return rc;
    if (!u || !rhs || !workspace) return WSR_E_ARG;
    if (!all_aligned16(u, rhs, rhs_reverse, workspace)) return WSR_E_ARG;
    if (workspace_bytes < ws7_layout(cfg->nx, cfg->ny, nullptr, nullptr)) return WSR_E_WORKSPACE;
    Workspace7 w;
    ws7_layout(cfg->nx, cfg->ny, &w, workspace);
    hipStream_t st = (hipStream_t)stream;
    Grid g = grid7(cfg);
    enqueue_params7(cfg, w, st);
    hipLaunchKernelGGL(k7_zero_counters, dim3(1), dim3(64), 0, st, w.ctl);
    enqueue_rhs7(cfg, w, u, rhs_reverse != nullptr, dt, time, nullptr, 0.0, st);
    const int np = g.nxt * g.nyt;
    hipLaunchKernelGGL(k7_rhs_out, dim3((np + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, (const double *)w.Fx,
                       (const double *)w.Fy, (const double *)w.Fxt, (const double *)w.Fyt, rhs, rhs_reverse, g);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_euler7_step(const wsr_euler7_config *cfg, double *u, double dt, double time, void *workspace,
                    size_t workspace_bytes, void *stream) {
    int rc = check7(cfg);
    if (rc) return rc;
    if (!u || !workspace) return WSR_E_ARG;
    if (!all_aligned16(u, workspace)) return WSR_E_ARG;
    if (workspace_bytes < ws7_layout(cfg->nx, cfg->ny, nullptr, nullptr)) return WSR_E_WORKSPACE;
    Workspace7 w;
    ws7_layout(cfg->nx, cfg->ny, &w, workspace);
    hipStream_t st = (hipStream_t)stream;
    enqueue_params7(cfg, w, st);
    hipLaunchKernelGGL(k7_zero_counters, dim3(1), dim3(64), 0, st, w.ctl);
    enqueue_step7(cfg, w, u, dt, time, nullptr, nullptr, st);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_euler7_loop_init(const wsr_euler7_config *cfg, const double *u, int32_t max_steps, void *workspace,
                         size_t workspace_bytes, void *stream) {
    int rc = check7(cfg);
    if (rc) return rc;
    if (!u || !workspace || max_steps < 0) return WSR_E_ARG;
    if (!all_aligned16(u, workspace)) return WSR_E_ARG;
    if (workspace_bytes < ws7_layout(cfg->nx, cfg->ny, nullptr, nullptr)) return WSR_E_WORKSPACE;
    Workspace7 w;
    ws7_layout(cfg->nx, cfg->ny, &w, workspace);
    hipStream_t st = (hipStream_t)stream;
    Grid g = grid7(cfg);
    double h = cfg->dx < cfg->dy ? cfg->dx : cfg->dy;
    hipLaunchKernelGGL(k7_loop_init, dim3(1), dim3(1), 0, st, w.ctl, cfg->t_end, cfg->cfl * h, max_steps);
    const int nc = g.nx * g.ny;
    hipLaunchKernelGGL(k7_max_speed, dim3((nc + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, u, g, w.ctl);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_euler7_loop_steps(const wsr_euler7_config *cfg, double *u, double *dts, int32_t n_steps, void *workspace,
                          size_t workspace_bytes, void *stream) {
    int rc = check7(cfg);
    if (rc) return rc;
    if (!u || !dts || !workspace || n_steps < 0) return WSR_E_ARG;
    if (!all_aligned16(u, workspace)) return WSR_E_ARG;
    if (workspace_bytes < ws7_layout(cfg->nx, cfg->ny, nullptr, nullptr)) return WSR_E_WORKSPACE;
    Workspace7 w;
    ws7_layout(cfg->nx, cfg->ny, &w, workspace);
    hipStream_t st = (hipStream_t)stream;
    if (n_steps > 0) enqueue_params7(cfg, w, st);
    for (int n = 0; n < n_steps; ++n) {
        hipLaunchKernelGGL(k7_dt, dim3(1), dim3(1), 0, st, w.ctl, cfg->loop_mode);
        enqueue_step7(cfg, w, u, 0.0, 0.0, w.ctl, dts, st);
    }
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_euler7_loop_steps_timed(const wsr_euler7_config *cfg, double *u, double *dts, int32_t n_steps,
                                void *workspace, size_t workspace_bytes, void *stream, double *sweep_ms,
                                int32_t *launches) {
    int rc = check7(cfg);
    if (rc) return rc;
    if (!u || !dts || !workspace || n_steps < 0 || n_steps > TIMER7_MAX_STEPS || !sweep_ms || !launches)
        return WSR_E_ARG;
    if (!all_aligned16(u, workspace)) return WSR_E_ARG;
    if (workspace_bytes < ws7_layout(cfg->nx, cfg->ny, nullptr, nullptr)) return WSR_E_WORKSPACE;
    Workspace7 w;
    ws7_layout(cfg->nx, cfg->ny, &w, workspace);
    hipStream_t st = (hipStream_t)stream;
    Timer7 *tm = new Timer7();
    tm->cap = 4 * n_steps;
    for (int i = 0; i < 2 * tm->cap; ++i)
        if (hipEventCreate(&tm->ev[i]) != hipSuccess) {
            while (i-- > 0) (void)hipEventDestroy(tm->ev[i]);
            delete tm;
            return WSR_E_HIP;
        }
    int32_t step0 = 0, step1 = 0;
    if (!loop_step_count(w.ctl, st, &step0)) rc = WSR_E_HIP;
    if (n_steps > 0) enqueue_params7(cfg, w, st);
    for (int n = 0; n < n_steps; ++n) {
        hipLaunchKernelGGL(k7_dt, dim3(1), dim3(1), 0, st, w.ctl, cfg->loop_mode);
        enqueue_step7(cfg, w, u, 0.0, 0.0, w.ctl, dts, st, tm);
    }
    if (rc == WSR_OK) rc = hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
    if (!loop_step_count(w.ctl, st, &step1)) rc = WSR_E_HIP;
    // only the launches of the steps that ran: four sweep launches per step, in order
    int ran = step1 - step0;
    ran = ran < 0 ? 0 : (ran > n_steps ? n_steps : ran);
    for (int k = 0; k < 2; ++k) { sweep_ms[k] = 0.0; launches[k] = 0; }
    for (int i = 0; i < tm->used && i < 4 * ran; ++i) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, tm->ev[2 * i], tm->ev[2 * i + 1]) == hipSuccess) {
            sweep_ms[tm->kind[i]] += ms;
            launches[tm->kind[i]] += 1;
        }
    }
    for (int i = 0; i < 2 * tm->cap; ++i) (void)hipEventDestroy(tm->ev[i]);
    delete tm;
    return rc;
}

int wsr_euler7_boundary(const wsr_euler7_config *cfg, double *u, double time, void *stream) {
    int rc = check7(cfg);
    if (rc) return rc;
    if (!u) return WSR_E_ARG;
    Grid g = grid7(cfg);
    const int np = g.nxt * g.nyt;
    hipLaunchKernelGGL(k7_ghosts_inplace, dim3((np + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, (hipStream_t)stream, u, g,
                       cfg->boundary, time);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_euler7_max_speed(const wsr_euler7_config *cfg, const double *u, void *workspace, size_t workspace_bytes,
                         double *max_speed_out, void *stream) {
    int rc = check7(cfg);
    if (rc) return rc;
    if (!u || !workspace || !max_speed_out) return WSR_E_ARG;
    if (!all_aligned16(u, workspace)) return WSR_E_ARG;
    // the reduction uses the control block at the start of the workspace only
    static_assert(sizeof(Ctl) <= WSR_SPEED_WORKSPACE_BYTES, "control block fits the speed workspace");
    if (workspace_bytes < WSR_SPEED_WORKSPACE_BYTES) return WSR_E_WORKSPACE;
    Ctl *ctl = (Ctl *)workspace;
    hipStream_t st = (hipStream_t)stream;
    Grid g = grid7(cfg);
    hipLaunchKernelGGL(k7_reset_speed, dim3(1), dim3(1), 0, st, ctl);
    const int nc = g.nx * g.ny;
    hipLaunchKernelGGL(k7_max_speed, dim3((nc + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, u, g, ctl);
    Ctl host;
    if (hipMemcpyAsync(&host, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st) != hipSuccess) return WSR_E_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return WSR_E_HIP;
    double m;
    memcpy(&m, &host.speed_bits, sizeof m);
    *max_speed_out = host.nan_flag ? NAN : m;
    return WSR_OK;
}

}  // extern "C"
