// wsr_euler3.hip -- MI355X (gfx950) kernels and C ABI of the 3-D WENO5 step
// (five conserved variables, ghost 3, SSPRK3).
//
// One SSPRK3 stage is eleven launches:
//   k3_boundary         : transmissive ghost fill of q in place (clamped copy of
//                         the nearest interior cell; the x, then y, then z
//                         copies of the reference give the same values).
//   per direction d = x, y, z:
//     k3_normal<d>      : characteristic normal reconstruction at every
//                         stencil start of the sweep box; blockIdx.y selects
//                         the face side (LR = 2 -> NL, LR = 1 -> NR) so a wave
//                         never diverges between the two paths.  LR = 1 runs
//                         on the stencil round-tripped through the LR = 2
//                         eigenvectors, R (L q), as the reference does.
//     k3_trans1<d>      : first transverse Gauss pass, one lane per (side,
//                         component, position); the closure is evaluated once
//                         for both Gauss points (-> S[first_lr][side]).
//     k3_flux<d>        : second transverse Gauss pass fused with the Riemann
//                         solves; four lanes per face for SR (first_lr x side),
//                         each reconstructing its side at both second_lr points
//                         and solving one of the four Riemann problems, two for
//                         JS (first_lr), then the face sum ((F1 + F2) + F3) + F4
//                         in the reference's flag order, times dt / h.
//   k3_update           : q <- SSPRK3 combination of q0 and q - (dFx+dFy+dFz)/4;
//                         in the device time loop the last stage also reduces
//                         the next step's max speed.
// Intermediate face arrays are component planes (SoA) over the padded grid,
// so every pass reads and writes contiguous x runs.
//
// Reference (paths relative to the reference checkout, directory
// code/snapshot/warp_weno5_3d_rk3/):
//   stage / flux sweeps: solver.py:165-301; advance: solver.py:303-329
//   normal reconstruction: kernels.py:570-588, 870-942; eigenvectors kernels.py:272-498
//   transverse passes: kernels.py:941-1014; EVILIN kernels.py:122-219; flux kernels.py:1016-1090
//   RK stages kernels.py:1092-1163; boundary kernels.py:802-852; max speed kernels.py:855-867
//   SR closure: mlp_kernels.py, mlp_kernels_f32.py (same closure as the 2-D path)
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/wsr.h"
#include "wsr_math.hpp"
#include "wsr_loop.hpp"

namespace wsr {
namespace e3 {

constexpr int BLOCK = 256;
constexpr int GC = 3;
#ifndef WSR3_WAVES_NORMAL
#define WSR3_WAVES_NORMAL 1
#endif
#ifndef WSR3_WAVES_FLUX
#define WSR3_WAVES_FLUX 5
#endif
// waves per SIMD the register allocation of the normal / flux passes aims for
// (FP32-closure flux lanes fit 5; the others keep the compiler's default)
template <int SCHEME> struct WavesNormal { static constexpr int value = SCHEME == SCHEME_SR64 ? 1 : WSR3_WAVES_NORMAL; };
template <int SCHEME> struct WavesFlux {
    static constexpr int value = (SCHEME == SCHEME_SR32 || SCHEME == SCHEME_SRBF16) ? WSR3_WAVES_FLUX : 1;
};

struct Grid3 {
    int nx, ny, nz, nxt, nyt, nzt;
    long long np;          // padded cells
    double dx, dy, dz, gamma;
};

struct V5 { double v[5]; };

// position of physical cell (k, j, i) in the padded grid
__device__ __forceinline__ long long pidx(const Grid3 &g, int k, int j, int i) {
    return ((long long)k * g.nyt + j) * g.nxt + i;
}

template <int DIR>
__device__ __forceinline__ long long axis_stride(const Grid3 &g) {
    return DIR == 1 ? 1ll : (DIR == 2 ? (long long)g.nxt : (long long)g.nxt * g.nyt);
}
// the first / second transverse axis of a sweep (solver.py:165-283: x faces
// y then z, y faces x then z, z faces y then x)
template <int DIR> struct Trans { static constexpr int T1 = DIR == 2 ? 1 : 2, T2 = DIR == 3 ? 1 : 3; };

// Workgroups are dealt to the eight XCDs in turn (workgroup w to XCD w mod 8), each with
// its own L2.  For 8-wide tiles (rows of 64 bytes) xcd_tile gives each XCD runs of two
// consecutive tiles, x neighbours whose rows share 128-byte lines: within every 16
// workgroups, workgroup w takes tile (w mod 8) * 2 + (w / 8) mod 2.  The runs stay
// interleaved across the XCDs, so spatially clustered closure work stays spread; the
// last partial group keeps the identity.  SR32 shock-bubble flux reads 1.12 -> 0.73 GB
// per launch (normal passes 14-23 % less), the same time.
template <int TILE_X>
__device__ __forceinline__ int xcd_tile(int w, int nw) {
    constexpr int R = TILE_X < 16 ? 2 : 1;
    if (R <= 1) return w;
    constexpr int G = 8 * R;
    if (w >= nw / G * G) return w;
    const int r = w % G;
    return w - r + (r % 8) * R + r / 8;
}

// Work item of a pass over a box ex x ey x ez (x fastest).  SAX = 1 walks x rows.
// SAX = 2 / 3 gives each workgroup a tile of TILE_X x positions times NPB / TILE_X lines
// along y / z, so that the 5-point stencils of one workgroup (along that axis) overlap
// in the same cache lines; in plain x-row order the neighbouring stencil lines belong
// to workgroups dispatched to other XCDs, whose L2s each fetch them (4-5x the
// algorithmic reads).  A wave of a tile is also a compact patch of the grid, which
// matters where a closure runs: fewer waves straddle a shock or an interface.
// active is false for the items of a partial tile (their coordinates are clamped
// into the box, so that their loads stay in bounds).
template <int SAX, int NPB, int TILE_X = 16>
__device__ __forceinline__ void tile_item(int f, int ex, int ey, int ez, int &i, int &j, int &k, bool &active) {
    if constexpr (SAX == 1) {
        const long long item = (long long)blockIdx.x * NPB + f;
        active = item < (long long)ex * ey * ez;
        const long long it = active ? item : (long long)ex * ey * ez - 1;
        i = (int)(it % ex);
        const long long r = it / ex;
        j = (int)(r % ey);
        k = (int)(r / ey);
    } else {
        static_assert(NPB % TILE_X == 0, "a tile holds whole x rows");
        constexpr int TS = NPB / TILE_X;
        const int es = SAX == 2 ? ey : ez;
        const int nbx = (ex + TILE_X - 1) / TILE_X, nbs = (es + TS - 1) / TS;
        const int b = xcd_tile<TILE_X>((int)blockIdx.x, (int)gridDim.x);
        const int bx = b % nbx, r = b / nbx;
        const int bs = r % nbs, other = r / nbs;
        i = bx * TILE_X + f % TILE_X;
        const int s = bs * TS + f / TILE_X;
        active = i < ex && s < es;
        i = i < ex ? i : ex - 1;
        const int sc = s < es ? s : es - 1;
        if (SAX == 2) { j = sc; k = other; }
        else { k = sc; j = other; }
    }
}
// Tile axis and width per pass (PASS 0 normal, 1 first transverse, 2 flux).  The
// closure passes tile along the stencil axis, or along y when the stencil runs along
// x, 8 wide for the normal and flux passes (8 x 8 / 8 x 2 patches per wave) and 16 wide
// for the first transverse pass.  Measured on the 224 x 88 x 88 shock-bubble (SR32):
// x rows everywhere 8.78 ms/step, 16-wide tiles along y / z 8.03, also for the x
// passes 7.55, these widths 7.28; 4-wide tiles were slower.  WENO-JS, whose light
// reconstructions leave these passes waiting on memory, keeps x rows for its normal
// and first passes (tiled: 8 % slower) and 16-wide tiles for its flux.
template <int SCHEME, int SAX> struct TileAxis { static constexpr int value = SCHEME == SCHEME_JS ? 1 : (SAX == 1 ? 2 : SAX); };
template <int SCHEME, int SAX> struct FluxAxis { static constexpr int value = SCHEME == SCHEME_JS ? SAX : (SAX == 1 ? 2 : SAX); };
template <int SCHEME, int PASS> struct TileWidth { static constexpr int value = SCHEME == SCHEME_JS ? 16 : (PASS == 1 ? 16 : 8); };

// workgroups of a tile_item pass
template <int SAX, int NPB, int TILE_X = 16>
static inline unsigned tile_blocks(int ex, int ey, int ez) {
    if (SAX == 1) return (unsigned)(((long long)ex * ey * ez + NPB - 1) / NPB);
    constexpr int TS = NPB / TILE_X;
    const int es = SAX == 2 ? ey : ez, eo = SAX == 2 ? ez : ey;
    return (unsigned)(((ex + TILE_X - 1) / TILE_X) * ((es + TS - 1) / TS) * eo);
}

// ------------------------------------------------------------ gas dynamics
__device__ __forceinline__ V5 con_to_pri5(const V5 &c, double gamma) {
    V5 p;
    p.v[0] = c.v[0];
    p.v[1] = c.v[1] / c.v[0];
    p.v[2] = c.v[2] / c.v[0];
    p.v[3] = c.v[3] / c.v[0];
    p.v[4] = (gamma - 1.0) * (c.v[4] - 0.5 * c.v[0] * (p.v[1] * p.v[1] + p.v[2] * p.v[2] + p.v[3] * p.v[3]));
    return p;
}
__device__ __forceinline__ V5 pri_to_con5(const V5 &p, double gamma) {
    V5 c;
    c.v[0] = p.v[0];
    c.v[1] = p.v[0] * p.v[1];
    c.v[2] = p.v[0] * p.v[2];
    c.v[3] = p.v[0] * p.v[3];
    c.v[4] = p.v[4] / (gamma - 1.0) + 0.5 * p.v[0] * (p.v[1] * p.v[1] + p.v[2] * p.v[2] + p.v[3] * p.v[3]);
    return c;
}
// kernels.py:58-97 physical flux of a primitive state
template <int DIR>
__device__ __forceinline__ V5 flux5(const V5 &p, double gamma) {
    double rho = p.v[0], u = p.v[1], v = p.v[2], w = p.v[3], pr = p.v[4];
    double e = 0.5 * rho * (u * u + v * v + w * w) + pr / (gamma - 1.0);
    V5 f;
    if (DIR == 1) {
        f.v[0] = rho * u; f.v[1] = rho * u * u + pr; f.v[2] = rho * u * v; f.v[3] = rho * u * w; f.v[4] = u * (pr + e);
    } else if (DIR == 2) {
        f.v[0] = rho * v; f.v[1] = rho * u * v; f.v[2] = rho * v * v + pr; f.v[3] = rho * v * w; f.v[4] = v * (pr + e);
    } else {
        f.v[0] = rho * w; f.v[1] = rho * u * w; f.v[2] = rho * v * w; f.v[3] = rho * w * w + pr; f.v[4] = w * (pr + e);
    }
    return f;
}

__device__ __forceinline__ void floor_rho_p(V5 &w, double eps) {
    w.v[0] = dmax(w.v[0], eps);
    w.v[4] = dmax(w.v[4], eps);
}

// kernels.py:122-219 evilin_state for the face with states (a | b) (a on the
// low side); the flux is f(evilin state) (kernels.py:1016-1090)
template <int DIR>
__device__ V5 evilin_flux5(const V5 &ul0, const V5 &ur0, double dt_dx, double gamma) {
    const double eps = 1.0e-15;
    V5 wl0 = con_to_pri5(ul0, gamma), wr0 = con_to_pri5(ur0, gamma);
    floor_rho_p(wl0, eps);
    floor_rho_p(wr0, eps);
    double unl0 = wl0.v[DIR], unr0 = wr0.v[DIR];
    double al0 = sqrt(gamma * wl0.v[4] / wl0.v[0]), ar0 = sqrt(gamma * wr0.v[4] / wr0.v[0]);
    double smax = dmax(dmax(fabs(unl0) + al0, fabs(unr0) + ar0), eps);
    double dx_dt = 1.0 / dt_dx, courant = smax * dt_dx;
    V5 fl = flux5<DIR>(wl0, gamma), fr = flux5<DIR>(wr0, gamma), u_lw, f_gf, ulh, urh;
#pragma unroll
    for (int c = 0; c < 5; ++c) u_lw.v[c] = 0.5 * (ul0.v[c] + ur0.v[c]) - 0.5 * dt_dx * (fr.v[c] - fl.v[c]);
    V5 w_lw = con_to_pri5(u_lw, gamma);
    floor_rho_p(w_lw, eps);
    V5 f_lw = flux5<DIR>(w_lw, gamma);
    double theta = (1.0 - courant) / (1.0 + courant);
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        double f_fo = 0.25 * (fl.v[c] + 2.0 * f_lw.v[c] + fr.v[c] - dx_dt * (ur0.v[c] - ul0.v[c]));
        f_gf.v[c] = theta * f_lw.v[c] + (1.0 - theta) * f_fo;
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        ulh.v[c] = ul0.v[c] - dt_dx * (f_gf.v[c] - fl.v[c]);
        urh.v[c] = ur0.v[c] - dt_dx * (fr.v[c] - f_gf.v[c]);
    }
    V5 wl = con_to_pri5(ulh, gamma), wr = con_to_pri5(urh, gamma);
    floor_rho_p(wl, eps);
    floor_rho_p(wr, eps);
    double rho_bar = dmax(0.5 * (wl.v[0] + wr.v[0]), eps), p_bar = dmax(0.5 * (wl.v[4] + wr.v[4]), eps);
    double a_bar = dmax(sqrt(gamma * p_bar / rho_bar), eps);
    double c1 = rho_bar * a_bar, c2 = rho_bar / a_bar;
    double unle = wl.v[DIR], unre = wr.v[DIR];
    double u_star = 0.5 * (unle + unre) - 0.5 * (wr.v[4] - wl.v[4]) / c1;
    double p_star = 0.5 * (wl.v[4] + wr.v[4]) - 0.5 * (unre - unle) * c1;
    double rho_star_l = dmax(wl.v[0] + (unle - u_star) * c2, eps);
    double rho_star_r = dmax(wr.v[0] + (u_star - unre) * c2, eps);
    p_star = dmax(p_star, eps);
    double u_bar = 0.5 * (unle + unre);
    V5 wf = wr;
    if (u_bar - a_bar >= 0.0) {
        wf = wl;
    } else if (u_bar >= 0.0) {
        wf = wl; wf.v[0] = rho_star_l; wf.v[DIR] = u_star; wf.v[4] = p_star;
    } else if (u_bar + a_bar >= 0.0) {
        wf = wr; wf.v[0] = rho_star_r; wf.v[DIR] = u_star; wf.v[4] = p_star;
    }
    floor_rho_p(wf, eps);
    // flux of the conserved round trip of the state, as the reference evaluates it
    return flux5<DIR>(con_to_pri5(pri_to_con5(wf, gamma), gamma), gamma);
}

// HLLC for five variables (the 2-D weno5_hllc.py:23-115 formula with both
// tangential velocities carried through the star state; no 3-D counterpart)
template <int DIR>
__device__ V5 hllc_flux5(const V5 &ul0, const V5 &ur0, double gamma) {
    const double tiny = 1.0e-16;
    V5 wl = con_to_pri5(ul0, gamma), wr = con_to_pri5(ur0, gamma);
    double rho_l = dmax(wl.v[0], tiny), rho_r = dmax(wr.v[0], tiny);
    double p_l = dmax(wl.v[4], tiny), p_r = dmax(wr.v[4], tiny);
    double a_l = sqrt(gamma * p_l / rho_l), a_r = sqrt(gamma * p_r / rho_r);
    double un_l = wl.v[DIR], un_r = wr.v[DIR];
    double s_l = dmin(un_l - a_l, un_r - a_r), s_r = dmax(un_l + a_l, un_r + a_r);
    V5 f_l = flux5<DIR>(wl, gamma), f_r = flux5<DIR>(wr, gamma);
    if (0.0 <= s_l) return f_l;
    if (0.0 >= s_r) return f_r;
    double denom = floor_signed(rho_l * (s_l - un_l) - rho_r * (s_r - un_r), tiny);
    double s_star = (p_r - p_l + rho_l * un_l * (s_l - un_l) - rho_r * un_r * (s_r - un_r)) / denom;
    const bool right = s_star < 0.0;
    const V5 &u_side = right ? ur0 : ul0;
    const V5 &w_side = right ? wr : wl;
    const V5 &f_side = right ? f_r : f_l;
    double s_side = right ? s_r : s_l;
    double rho = dmax(w_side.v[0], tiny), p = dmax(w_side.v[4], tiny), un = w_side.v[DIR];
    double factor = rho * (s_side - un) / floor_signed(s_side - s_star, tiny);
    V5 us;
    us.v[0] = factor;
    us.v[1] = factor * w_side.v[1];
    us.v[2] = factor * w_side.v[2];
    us.v[3] = factor * w_side.v[3];
    us.v[DIR] = factor * s_star;
    us.v[4] = factor * ((u_side.v[4] / rho) + (s_star - un) * (s_star + p / (rho * floor_signed(s_side - un, tiny))));
    V5 out;
#pragma unroll
    for (int c = 0; c < 5; ++c) out.v[c] = f_side.v[c] + s_side * (us.v[c] - u_side.v[c]);
    return out;
}

// ------------------------------------------------------ characteristic fields
// kernels.py:233-269 roe_average (no floors), as a conserved state
__device__ __forceinline__ V5 roe5(const V5 &l, const V5 &r, double gamma) {
    double rho_l = l.v[0], u_l = l.v[1] / rho_l, v_l = l.v[2] / rho_l, w_l = l.v[3] / rho_l, e_l = l.v[4];
    double p_l = (gamma - 1.0) * (e_l - 0.5 * rho_l * (u_l * u_l + v_l * v_l + w_l * w_l));
    double h_l = (e_l + p_l) / rho_l;
    double rho_r = r.v[0], u_r = r.v[1] / rho_r, v_r = r.v[2] / rho_r, w_r = r.v[3] / rho_r, e_r = r.v[4];
    double p_r = (gamma - 1.0) * (e_r - 0.5 * rho_r * (u_r * u_r + v_r * v_r + w_r * w_r));
    double h_r = (e_r + p_r) / rho_r;
    double sl = sqrt(rho_l), sr = sqrt(rho_r), inv = 1.0 / (sl + sr);
    double u = (sl * u_l + sr * u_r) * inv, v = (sl * v_l + sr * v_r) * inv, w = (sl * w_l + sr * w_r) * inv;
    double h = (sl * h_l + sr * h_r) * inv, rho = sl * sr, q2 = u * u + v * v + w * w;
    double e = (rho * h + 0.5 * (gamma - 1.0) * rho * q2) / gamma;
    V5 o;
    o.v[0] = rho; o.v[1] = rho * u; o.v[2] = rho * v; o.v[3] = rho * w; o.v[4] = e;
    return o;
}

// kernels.py:272-498: the eigenvector matrices of the DIR Jacobian at a
// conserved state, kept as the handful of scalars they are built from
struct Eig5 {
    double nn, t1, t2, vtot, c, c2, h, gm1, scale;
};
template <int DIR>
__device__ __forceinline__ Eig5 eigen5(const V5 &s, double gamma) {
    Eig5 E;
    double rho = s.v[0], u = s.v[1] / rho, v = s.v[2] / rho, w = s.v[3] / rho, energy = s.v[4];
    E.gm1 = gamma - 1.0;
    E.vtot = u * u + v * v + w * w;
    double pressure = E.gm1 * (energy - 0.5 * rho * E.vtot);
    E.c = sqrt(gamma * pressure / rho);
    E.c2 = E.c * E.c;
    E.h = 0.5 * E.vtot + E.c2 / E.gm1;
    E.nn = DIR == 1 ? u : (DIR == 2 ? v : w);
    E.t1 = DIR == 1 ? v : u;
    E.t2 = DIR == 3 ? v : w;
    E.scale = E.gm1 / (2.0 * E.c2);
    return E;
}
// component slots of the normal and the two tangential momenta
template <int DIR> struct Slots { static constexpr int IN = DIR, I1 = DIR == 1 ? 2 : 1, I2 = DIR == 3 ? 2 : 3; };

// row k of the scaled left matrix L (5 entries in conserved-component order)
template <int DIR>
__device__ __forceinline__ void left_row5(int k, const Eig5 &E, double l[5]) {
    constexpr int IN = Slots<DIR>::IN, I1 = Slots<DIR>::I1, I2 = Slots<DIR>::I2;
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (k == 0) {
        r[0] = E.h + E.c * (E.nn - E.c) / E.gm1; r[IN] = -(E.nn + E.c / E.gm1); r[I1] = -E.t1; r[I2] = -E.t2;
        r[4] = 1.0;
    } else if (k == 1) {
        r[0] = -2.0 * E.h + 4.0 * E.c2 / E.gm1; r[IN] = 2.0 * E.nn; r[I1] = 2.0 * E.t1; r[I2] = 2.0 * E.t2;
        r[4] = -2.0;
    } else if (k == 2) {
        r[0] = -2.0 * E.t1 * E.c2 / E.gm1; r[I1] = 2.0 * E.c2 / E.gm1;
    } else if (k == 3) {
        r[0] = -2.0 * E.t2 * E.c2 / E.gm1; r[I2] = 2.0 * E.c2 / E.gm1;
    } else {
        r[0] = E.h - E.c * (E.nn + E.c) / E.gm1; r[IN] = -(E.nn - E.c / E.gm1); r[I1] = -E.t1; r[I2] = -E.t2;
        r[4] = 1.0;
    }
#pragma unroll
    for (int b = 0; b < 5; ++b) l[b] = r[b] * E.scale;
}
// column k of the right matrix R
template <int DIR>
__device__ __forceinline__ void right_col5(int k, const Eig5 &E, double r[5]) {
    constexpr int IN = Slots<DIR>::IN, I1 = Slots<DIR>::I1, I2 = Slots<DIR>::I2;
    if (k == 0) { r[0] = 1.0; r[IN] = E.nn - E.c; r[I1] = E.t1; r[I2] = E.t2; r[4] = E.h - E.nn * E.c; }
    else if (k == 1) { r[0] = 1.0; r[IN] = E.nn; r[I1] = E.t1; r[I2] = E.t2; r[4] = 0.5 * E.vtot; }
    else if (k == 2) { r[0] = 0.0; r[IN] = 0.0; r[I1] = 1.0; r[I2] = 0.0; r[4] = E.t1; }
    else if (k == 3) { r[0] = 0.0; r[IN] = 0.0; r[I1] = 0.0; r[I2] = 1.0; r[4] = E.t2; }
    else { r[0] = 1.0; r[IN] = E.nn + E.c; r[I1] = E.t1; r[I2] = E.t2; r[4] = E.h + E.nn * E.c; }
}

__device__ __forceinline__ double dot5(const double a[5], const V5 &x) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 5; ++b) t = t + a[b] * x.v[b];
    return t;
}

// ------------------------------------------------------------------ kernels
template <int SCHEME>
struct PassArgs {
    Grid3 g;
    typename MlpSweep<SCHEME>::type mlp;
    int eno;
    const Ctl *ctl;
    unsigned long long *counters;   // nullptr unless closure evaluations are counted
};

// kernels.py:802-852 boundary_{x,y,z}_kernel: transmissive copy, in place
__global__ __launch_bounds__(BLOCK) void k3_boundary(double *__restrict__ q, Grid3 g, const Ctl *ctl) {
    if (ctl && ctl->done) return;
    long long idx = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= g.np) return;
    int i = (int)(idx % g.nxt);
    long long r = idx / g.nxt;
    int j = (int)(r % g.nyt), k = (int)(r / g.nyt);
    int si = i < GC ? GC : (i >= g.nx + GC ? g.nx + GC - 1 : i);
    int sj = j < GC ? GC : (j >= g.ny + GC ? g.ny + GC - 1 : j);
    int sk = k < GC ? GC : (k >= g.nz + GC ? g.nz + GC - 1 : k);
    if (si == i && sj == j && sk == k) return;
    const double *s = q + pidx(g, sk, sj, si) * 5;
    double *d = q + idx * 5;
#pragma unroll
    for (int c = 0; c < 5; ++c) d[c] = s[c];
}

// characteristic normal reconstruction (kernels.py:570-588, 870-942);
// blockIdx.y = 0: LR = 2 (face left of the stencil centre) -> NL,
// blockIdx.y = 1: LR = 1 on R (L q) (face right of the centre) -> NR
template <int DIR, int SCHEME>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WavesNormal<SCHEME>::value))) void k3_normal(const double *__restrict__ q, double *__restrict__ NL,
                                                   double *__restrict__ NR, PassArgs<SCHEME> A) {
    if (A.ctl && A.ctl->done) return;
    const Grid3 &g = A.g;
    const int ex = DIR == 1 ? g.nx + 2 : g.nx + 4, ey = DIR == 2 ? g.ny + 2 : g.ny + 4;
    const int ez = DIR == 3 ? g.nz + 2 : g.nz + 4;
    const int ox = DIR == 1 ? 0 : 1, oy = DIR == 2 ? 0 : 1, oz = DIR == 3 ? 0 : 1;
    int i, j, k;
    bool active;                            // idle lanes recompute a clamped position, store nothing
    tile_item<TileAxis<SCHEME, DIR>::value, BLOCK, TileWidth<SCHEME, 0>::value>((int)threadIdx.x, ex, ey, ez, i, j, k, active);
    i += ox; j += oy; k += oz;
    const long long base = pidx(g, k, j, i), st = axis_stride<DIR>(g);
    const bool lr1 = blockIdx.y == 1;
    LaneCounts cnt;

    V5 s[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const double *p = q + (base + m * st) * 5;
#pragma unroll
        for (int c = 0; c < 5; ++c) s[m].v[c] = p[c];
    }
    Eig5 E = eigen5<DIR>(roe5(s[1], s[2], g.gamma), g.gamma);
    if (lr1) {
        // the reference's LR = 2 pass leaves R (L q) in the stencil it reuses
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            double ch[5];
#pragma unroll
            for (int kk = 0; kk < 5; ++kk) {
                double l[5];
                left_row5<DIR>(kk, E, l);
                ch[kk] = dot5(l, s[m]);
            }
            V5 back;
#pragma unroll
            for (int a = 0; a < 5; ++a) back.v[a] = 0.0;
#pragma unroll
            for (int kk = 0; kk < 5; ++kk) {
                double rc[5];
                right_col5<DIR>(kk, E, rc);
#pragma unroll
                for (int a = 0; a < 5; ++a) back.v[a] = back.v[a] + rc[a] * ch[kk];
            }
            s[m] = back;
        }
        E = eigen5<DIR>(roe5(s[2], s[3], g.gamma), g.gamma);
    }
    double h = DIR == 1 ? g.dx : (DIR == 2 ? g.dy : g.dz);
    V5 out;
#pragma unroll
    for (int a = 0; a < 5; ++a) out.v[a] = 0.0;
#pragma unroll 1
    for (int kk = 0; kk < 5; ++kk) {
        double l[5], w[5], rc[5];
        left_row5<DIR>(kk, E, l);
#pragma unroll
        for (int m = 0; m < 5; ++m) w[m] = dot5(l, s[m]);
        bool ev;
        double f = recon_normal<SCHEME>(w, lr1, A.mlp, A.eno, h, ev);
        if (A.counters) cnt.add(0, 4, active, ev);
        right_col5<DIR>(kk, E, rc);
#pragma unroll
        for (int a = 0; a < 5; ++a) out.v[a] = out.v[a] + rc[a] * f;
    }
    double *dst = lr1 ? NR : NL;
    if (active) {
#pragma unroll
        for (int c = 0; c < 5; ++c) dst[c * g.np + base] = out.v[c];
    }
    if (A.counters) cnt.flush_block(A.counters);
}

// first transverse Gauss pass (kernels.py:941-1014): lane -> (side, comp, position)
// S[(first_lr - 1) * 2 + side], side 0 from NL, 1 from NR
template <int DIR, int SCHEME>
__global__ __launch_bounds__(BLOCK) void k3_trans1(const double *__restrict__ NL, const double *__restrict__ NR,
                                                   double *__restrict__ S, PassArgs<SCHEME> A) {
    if (A.ctl && A.ctl->done) return;
    const Grid3 &g = A.g;
    constexpr int T1 = Trans<DIR>::T1;
    // box: normal [0, n+2), t1 targets [0, n_t1), t2 [1, n_t2 + 5)
    const int ex = DIR == 1 ? g.nx + 2 : (T1 == 1 ? g.nx : g.nx + 4);
    const int ey = DIR == 2 ? g.ny + 2 : g.ny;     // y is t1 for x and z sweeps
    const int ez = DIR == 3 ? g.nz + 2 : g.nz + 4; // z is t2 for x and y sweeps
    const int ox = DIR == 3 ? 1 : 0, oy = 0, oz = DIR == 3 ? 0 : 1;
    const int sc = blockIdx.y;                 // side * 5 + comp
    const int side = sc / 5, comp = sc - side * 5;
    int i, j, k;
    bool active;
    tile_item<TileAxis<SCHEME, T1>::value, BLOCK, TileWidth<SCHEME, 1>::value>((int)threadIdx.x, ex, ey, ez, i, j, k, active);
    i += ox; j += oy; k += oz;
    const long long pos = pidx(g, k, j, i), st = axis_stride<T1>(g);
    const double *src = (side == 0 ? NL : NR) + comp * g.np + pos;
    double w[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) w[m] = src[(m + 1) * st];
    double h = T1 == 1 ? g.dx : g.dy;
    double g1, g2;
    bool ev;
    recon_gauss<SCHEME>(w, A.mlp, A.eno, h, g1, g2, ev);
    const long long plane = 5 * g.np;
    if (active) {
        S[(0 * 2 + side) * plane + comp * g.np + pos] = g1;
        S[(1 * 2 + side) * plane + comp * g.np + pos] = g2;
    }
    if (A.counters) {
        LaneCounts cnt;
        cnt.add(2, 5, active, ev);
        cnt.flush_block(A.counters);
    }
}

// second transverse pass + Riemann solves + flag sum (solver.py:176-203,
// kernels.py:1016-1090).  With the SR closures four lanes own a face:
// q = lane & 3 = 2 (first_lr - 1) + side (0: low side, from NR at the face's
// index; 1: high side, from NL one cell up the normal axis); a lane reconstructs
// its side's five components at both second_lr points, swaps one point with its
// partner (lane ^ 1) and solves one Riemann problem (side 0: second_lr = 1,
// side 1: second_lr = 2).  WENO-JS, whose reconstructions are cheap, uses two
// lanes per face (first_lr), each reconstructing both sides and solving two.
template <int SCHEME> struct FluxLanes { static constexpr int value = SCHEME == SCHEME_JS ? 2 : 4; };

template <int DIR, int SCHEME, int RIEMANN>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WavesFlux<SCHEME>::value))) void k3_flux(
    const double *__restrict__ S, double *__restrict__ F, PassArgs<SCHEME> A, double dt_host) {
    if (A.ctl && A.ctl->done) return;
    constexpr int LPF = FluxLanes<SCHEME>::value;
    const Grid3 &g = A.g;
    constexpr int T2 = Trans<DIR>::T2;
    const int ex = DIR == 1 ? g.nx + 1 : g.nx, ey = DIR == 2 ? g.ny + 1 : g.ny, ez = DIR == 3 ? g.nz + 1 : g.nz;
    const int q = (int)(threadIdx.x % LPF);
    const int a = LPF == 4 ? q >> 1 : q;       // first_lr - 1
    int i, j, k;
    bool active;                               // idle lanes still take part in the exchanges
    tile_item<FluxAxis<SCHEME, T2>::value, BLOCK / LPF, TileWidth<SCHEME, 2>::value>((int)(threadIdx.x / LPF), ex, ey, ez, i, j, k, active);
    const long long pos = pidx(g, k, j, i), sn = axis_stride<DIR>(g), st = axis_stride<T2>(g);
    const double h = DIR == 1 ? g.dx : (DIR == 2 ? g.dy : g.dz);
    const double ht = T2 == 1 ? g.dx : g.dz;
    const double dt = A.ctl ? A.ctl->dt : dt_host;
    const double dt_h = dt / h;
    const long long plane = 5 * g.np;
    const double *lo_src = S + (a * 2 + 1) * plane + pos;
    const double *hi_src = S + (a * 2 + 0) * plane + pos + sn;
    LaneCounts cnt;
    double tot[5];
    const int base = threadIdx.x & ~(LPF - 1);
    if constexpr (LPF == 4) {
        const int side = q & 1;
        const double *src = side == 0 ? lo_src : hi_src;
        V5 v1 = {}, v2 = {};   // this side at second_lr = 1, 2
#pragma unroll 1
        for (int c = 0; c < 5; ++c) {
            double w[5], g1, g2;
            bool ev;
#pragma unroll
            for (int m = 0; m < 5; ++m) w[m] = src[c * g.np + (m + 1) * st];
            recon_gauss<SCHEME>(w, A.mlp, A.eno, ht, g1, g2, ev);
            if (A.counters) cnt.add(2, 5, active, ev);
#pragma unroll
            for (int cc = 0; cc < 5; ++cc) {
                if (cc == c) { v1.v[cc] = g1; v2.v[cc] = g2; }
            }
        }
        // side 0 keeps point 1 and receives the high side's point 1; side 1 keeps
        // point 2 and receives the low side's point 2
        V5 lo, hi;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const double send = side == 0 ? v2.v[c] : v1.v[c];
            const double got = __shfl_xor(send, 1);
            lo.v[c] = side == 0 ? v1.v[c] : got;
            hi.v[c] = side == 0 ? got : v2.v[c];
        }
        V5 fb;
        if constexpr (RIEMANN == WSR_RIEMANN_EVILIN) fb = evilin_flux5<DIR>(lo, hi, dt_h, g.gamma);
        else fb = hllc_flux5<DIR>(lo, hi, g.gamma);
        // flags 1..4 = (first_lr, second_lr) (1,1), (2,1), (1,2), (2,2) live on q = 0, 2, 1, 3
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const double sv = fb.v[c] * dt_h;
            const double f1 = __shfl(sv, base + 0), f2 = __shfl(sv, base + 2);
            const double f3 = __shfl(sv, base + 1), f4 = __shfl(sv, base + 3);
            tot[c] = ((f1 + f2) + f3) + f4;
        }
    } else {
        V5 lo1 = {}, lo2 = {}, hi1 = {}, hi2 = {};
#pragma unroll 1
        for (int c = 0; c < 5; ++c) {
            double w[5], g1, g2, e1, e2;
            bool ev;
#pragma unroll
            for (int m = 0; m < 5; ++m) w[m] = lo_src[c * g.np + (m + 1) * st];
            recon_gauss<SCHEME>(w, A.mlp, A.eno, ht, g1, g2, ev);
            if (A.counters) cnt.add(2, 5, active, ev);
#pragma unroll
            for (int m = 0; m < 5; ++m) w[m] = hi_src[c * g.np + (m + 1) * st];
            recon_gauss<SCHEME>(w, A.mlp, A.eno, ht, e1, e2, ev);
            if (A.counters) cnt.add(2, 5, active, ev);
#pragma unroll
            for (int cc = 0; cc < 5; ++cc) {
                if (cc == c) { lo1.v[cc] = g1; lo2.v[cc] = g2; hi1.v[cc] = e1; hi2.v[cc] = e2; }
            }
        }
        V5 fb1, fb2;   // second_lr = 1, 2
        if constexpr (RIEMANN == WSR_RIEMANN_EVILIN) {
            fb1 = evilin_flux5<DIR>(lo1, hi1, dt_h, g.gamma);
            fb2 = evilin_flux5<DIR>(lo2, hi2, dt_h, g.gamma);
        } else {
            fb1 = hllc_flux5<DIR>(lo1, hi1, g.gamma);
            fb2 = hllc_flux5<DIR>(lo2, hi2, g.gamma);
        }
        // flags 1..4: (a=0, 1), (a=1, 1), (a=0, 2), (a=1, 2)
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const double s1 = fb1.v[c] * dt_h, s2 = fb2.v[c] * dt_h;      // this lane's (a, 1), (a, 2)
            const double o1 = __shfl_xor(s1, 1), o2 = __shfl_xor(s2, 1);   // the partner's
            tot[c] = ((s1 + o1) + s2) + o2;                                 // used on the a = 0 lane
        }
    }
    if (active && q == 0) {
#pragma unroll
        for (int c = 0; c < 5; ++c) F[c * g.np + pos] = tot[c];
    }
    if (A.counters) cnt.flush_block(A.counters);
}

// kernels.py:1092-1163 rk_stage{1,2,3}: q <- {adv, 3/4 q0 + 1/4 adv, 1/3 q0 + 2/3 adv},
// adv = q - (Fx_r - Fx_l)/4 - (Fy_r - Fy_l)/4 - (Fz_r - Fz_l)/4 (stage 1 saves q0)
__global__ __launch_bounds__(BLOCK) void k3_update(double *__restrict__ q, double *__restrict__ q0,
                                                   const double *__restrict__ Fx, const double *__restrict__ Fy,
                                                   const double *__restrict__ Fz, Grid3 g, int rk, Ctl *ctl,
                                                   int fold_speed, double *dts) {
    if (ctl && ctl->done) return;
    long long idx = (long long)blockIdx.x * BLOCK + threadIdx.x;
    const long long ncells = (long long)g.nx * g.ny * g.nz;
    const bool active = idx < ncells;
    double sp = 0.0;
    if (active) {
        const int i = (int)(idx % g.nx);
        long long r = idx / g.nx;
        const int j = (int)(r % g.ny), k = (int)(r / g.ny);
        const long long f = pidx(g, k, j, i), e = pidx(g, k + GC, j + GC, i + GC);
        const long long sy = g.nxt, sz = (long long)g.nxt * g.nyt;
        V5 o;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const long long p = c * g.np + f;
            double old = q[e * 5 + c];
            double adv = old - 0.25 * (Fx[p + 1] - Fx[p]) - 0.25 * (Fy[p + sy] - Fy[p]) - 0.25 * (Fz[p + sz] - Fz[p]);
            double v;
            if (rk == 1) { q0[e * 5 + c] = old; v = adv; }
            else if (rk == 2) v = (3.0 / 4.0) * q0[e * 5 + c] + (1.0 / 4.0) * adv;
            else v = (1.0 / 3.0) * q0[e * 5 + c] + (2.0 / 3.0) * adv;
            o.v[c] = v;
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) q[e * 5 + c] = o.v[c];
        if (fold_speed) {
            V5 p = con_to_pri5(o, g.gamma);
            double a = sqrt(g.gamma * p.v[4] / p.v[0]);
            sp = dmax(fabs(p.v[1]) + a, dmax(fabs(p.v[2]) + a, fabs(p.v[3]) + a));
        }
    }
    if (fold_speed) {
        reduce_speed<BLOCK>(sp, active, ctl);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            // solver.py:303-329: time += dt; step += 1; dt_trace
            double dt = ctl->dt;
            dts[ctl->step] = dt;
            ctl->t = ctl->t + dt;
            ctl->dt_last = dt;
            ctl->step = ctl->step + 1;
        }
    }
}

// kernels.py:855-867 max_speed_kernel (no floors)
__global__ __launch_bounds__(BLOCK) void k3_max_speed(const double *__restrict__ q, Grid3 g, Ctl *ctl) {
    long long idx = (long long)blockIdx.x * BLOCK + threadIdx.x;
    const long long ncells = (long long)g.nx * g.ny * g.nz;
    const bool active = idx < ncells;
    double sp = 0.0;
    if (active) {
        const int i = (int)(idx % g.nx);
        long long r = idx / g.nx;
        const int j = (int)(r % g.ny), k = (int)(r / g.ny);
        const double *c = q + pidx(g, k + GC, j + GC, i + GC) * 5;
        V5 s;
#pragma unroll
        for (int m = 0; m < 5; ++m) s.v[m] = c[m];
        V5 p = con_to_pri5(s, g.gamma);
        double a = sqrt(g.gamma * p.v[4] / p.v[0]);
        sp = dmax(fabs(p.v[1]) + a, dmax(fabs(p.v[2]) + a, fabs(p.v[3]) + a));
    }
    reduce_speed<BLOCK>(sp, active, ctl);
}

__global__ void k3_reset_speed(Ctl *ctl) {
    ctl->speed_bits = 0ull;
    ctl->nan_flag = 0;
}

__global__ void k3_loop_init(Ctl *ctl, double t_end, double cfl_h, int max_steps) {
    ctl->t = 0.0; ctl->dt = 0.0; ctl->t_end = t_end; ctl->cfl_h = cfl_h; ctl->dt_last = 0.0;
    ctl->speed_bits = 0ull; ctl->step = 0; ctl->done = 0; ctl->nan_flag = 0; ctl->max_steps = max_steps;
    for (int i = 0; i < NCOUNT; ++i) ctl->counters[i] = 0ull;
}

// solver.py:303-312: stop at time == t_end; dt = cfl min(h) / max (1e10 below
// 1e-15), clipped so that time + dt does not pass t_end.  A NaN speed stops
// the loop (the reference would go on with a NaN dt).
__global__ void k3_dt(Ctl *ctl) {
    if (ctl->done) return;
    double t = ctl->t;
    if (t == ctl->t_end || ctl->step >= ctl->max_steps) { ctl->done = 1; return; }
    double dt;
    if (ctl->nan_flag) {
        dt = __longlong_as_double(0x7ff8000000000000ll);
    } else {
        double m = __longlong_as_double((long long)ctl->speed_bits);
        dt = m < 1.0e-15 ? 1.0e10 : ctl->cfl_h / m;
    }
    if (t + dt > ctl->t_end) dt = ctl->t_end - t;
    if (!isfinite(dt) || dt <= 0.0) { ctl->done = 1; return; }
    ctl->dt = dt;
    ctl->speed_bits = 0ull;
    ctl->nan_flag = 0;
}

// ----------------------------------------------------------- host side
struct Workspace3 {
    Ctl *ctl;
    void *mlp;
    double *q0;            // stage-1 copy (AoS, padded)
    double *NL, *NR;       // normal face values (SoA, padded)
    double *S;             // four first-pass arrays (SoA, padded)
    double *Fx, *Fy, *Fz;  // face flux sums (SoA, padded strides)
};

static size_t ws3_layout(const wsr_euler3_config *cfg, Workspace3 *w, void *base) {
    const size_t np = (size_t)(cfg->nx + 6) * (cfg->ny + 6) * (cfg->nz + 6);
    const size_t arr = np * 5 * sizeof(double);
    size_t off = 0;
    char *b = (char *)base;
    auto take = [&](size_t bytes) { char *p = b + off; off += align256(bytes); return (void *)p; };
    void *ctl = take(sizeof(Ctl));
    void *mlp = take(MLP5_COUNT * sizeof(double));
    void *q0 = take(arr);
    void *nl = take(arr);
    void *nr = take(arr);
    void *s = take(4 * arr);
    void *fx = take(arr);
    void *fy = take(arr);
    void *fz = take(arr);
    if (w) {
        w->ctl = (Ctl *)ctl; w->mlp = mlp; w->q0 = (double *)q0; w->NL = (double *)nl; w->NR = (double *)nr;
        w->S = (double *)s; w->Fx = (double *)fx; w->Fy = (double *)fy; w->Fz = (double *)fz;
    }
    return off;
}

static int check3(const wsr_euler3_config *cfg) {
    if (!cfg) return WSR_E_ARG;
    if (cfg->ghost != 3 || cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return WSR_E_ARG;
    if ((long long)(cfg->nx + 6) * (cfg->ny + 6) * (cfg->nz + 6) > (1ll << 30)) return WSR_E_ARG;
    if (cfg->scheme != WSR_SCHEME_JS && cfg->scheme != WSR_SCHEME_SR64 && cfg->scheme != WSR_SCHEME_SR32 &&
        cfg->scheme != WSR_SCHEME_SRBF16)
        return WSR_E_ARG;
    if (cfg->riemann != WSR_RIEMANN_HLLC && cfg->riemann != WSR_RIEMANN_EVILIN) return WSR_E_ARG;
    if (cfg->boundary != WSR_BC_TRANSMISSIVE) return WSR_E_ARG;
    if (!(cfg->dx > 0.0) || !(cfg->dy > 0.0) || !(cfg->dz > 0.0) || !(cfg->gamma > 1.0)) return WSR_E_ARG;
    bool sr = cfg->scheme != WSR_SCHEME_JS;
    if (sr && !cfg->mlp) return WSR_E_MLP;
    return WSR_OK;
}

static Grid3 make_grid3(const wsr_euler3_config *cfg) {
    Grid3 g;
    g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz;
    g.nxt = cfg->nx + 6; g.nyt = cfg->ny + 6; g.nzt = cfg->nz + 6;
    g.np = (long long)g.nxt * g.nyt * g.nzt;
    g.dx = cfg->dx; g.dy = cfg->dy; g.dz = cfg->dz; g.gamma = cfg->gamma;
    return g;
}

struct ParamWords3 { uint32_t w[MLP5_COUNT * 2]; };

__global__ void k3_put_params(ParamWords3 src, uint32_t *dst, int nwords) {
    for (int i = threadIdx.x; i < nwords; i += blockDim.x) dst[i] = src.w[i];
}

static void enqueue_params3(const wsr_euler3_config *cfg, const Workspace3 &w, hipStream_t st) {
    const int scheme = cfg->scheme;
    if (scheme == WSR_SCHEME_JS) return;
    ParamWords3 pw;
    memset(&pw, 0, sizeof pw);
    int nwords;
    if (scheme == WSR_SCHEME_SR64) {
        double dev[MLP5_COUNT];
        pack_neuron_major((const double *)cfg->mlp, dev);
        memcpy(pw.w, dev, sizeof dev);
        nwords = MLP5_COUNT * 2;
    } else if (scheme == WSR_SCHEME_SRBF16) {
        float ref[MLP5_COUNT];
        memcpy(ref, cfg->mlp, sizeof ref);
        pack_bf16_closure(ref, pw.w);   // BF16 weight pairs and FP32 biases for the dot-product closure
        nwords = BF_WORDS;
    } else {
        float ref[MLP5_COUNT], dev[MLP5_COUNT];
        memcpy(ref, cfg->mlp, sizeof ref);
        pack_neuron_major(ref, dev);
        fold_fp32_closure(dev);   // the FP32 closure evaluates folded parameters
        memcpy(pw.w, dev, sizeof dev);
        nwords = MLP5_COUNT;
    }
    hipLaunchKernelGGL(k3_put_params, dim3(1), dim3(256), 0, st, pw, (uint32_t *)w.mlp, nwords);
}

static inline unsigned blocks_for(long long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

// optional HIP-event pairs around the normal / first-pass / flux launches
constexpr int TIMER3_MAX_STEPS = 32;
constexpr int TIMER3_PAIRS = TIMER3_MAX_STEPS * 3 * 3 * 3;
struct Timer3 {
    hipEvent_t ev[2 * TIMER3_PAIRS];
    int kind[TIMER3_PAIRS];
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

template <int DIR, int SCHEME, int RIEMANN>
static void launch_direction(double *q, const Workspace3 &w, double *F, const Grid3 &g, int eno, const Ctl *ctl,
                             double dt, unsigned long long *counters, hipStream_t st, Timer3 *tm) {
    PassArgs<SCHEME> A;
    memset(&A, 0, sizeof A);
    A.g = g; A.eno = eno; A.ctl = ctl; A.counters = counters;
    if constexpr (SCHEME == SCHEME_SR64) A.mlp.p = (const double *)w.mlp;
    else if constexpr (SCHEME != SCHEME_JS) A.mlp.p = (const float *)w.mlp;
    const unsigned nbn = tile_blocks<TileAxis<SCHEME, DIR>::value, BLOCK, TileWidth<SCHEME, 0>::value>(DIR == 1 ? g.nx + 2 : g.nx + 4, DIR == 2 ? g.ny + 2 : g.ny + 4,
                                                 DIR == 3 ? g.nz + 2 : g.nz + 4);
    if (tm) tm->begin(st);
    hipLaunchKernelGGL((k3_normal<DIR, SCHEME>), dim3(nbn, 2), dim3(BLOCK), 0, st, (const double *)q, w.NL, w.NR, A);
    if (tm) tm->end(0, st);
    constexpr int T1 = Trans<DIR>::T1;
    const unsigned nbt = tile_blocks<TileAxis<SCHEME, T1>::value, BLOCK, TileWidth<SCHEME, 1>::value>(DIR == 1 ? g.nx + 2 : (T1 == 1 ? g.nx : g.nx + 4),
                                                DIR == 2 ? g.ny + 2 : g.ny, DIR == 3 ? g.nz + 2 : g.nz + 4);
    if (tm) tm->begin(st);
    hipLaunchKernelGGL((k3_trans1<DIR, SCHEME>), dim3(nbt, 10), dim3(BLOCK), 0, st, (const double *)w.NL,
                       (const double *)w.NR, w.S, A);
    if (tm) tm->end(1, st);
    const unsigned nbf = tile_blocks<FluxAxis<SCHEME, Trans<DIR>::T2>::value, BLOCK / FluxLanes<SCHEME>::value, TileWidth<SCHEME, 2>::value>(
        DIR == 1 ? g.nx + 1 : g.nx, DIR == 2 ? g.ny + 1 : g.ny, DIR == 3 ? g.nz + 1 : g.nz);
    if (tm) tm->begin(st);
    hipLaunchKernelGGL((k3_flux<DIR, SCHEME, RIEMANN>), dim3(nbf), dim3(BLOCK), 0, st, (const double *)w.S, F, A, dt);
    if (tm) tm->end(2, st);
}

template <int SCHEME, int RIEMANN>
static void enqueue_fluxes_t(double *q, const Workspace3 &w, const Grid3 &g, int eno, const Ctl *ctl, double dt,
                             unsigned long long *counters, hipStream_t st, Timer3 *tm) {
    launch_direction<1, SCHEME, RIEMANN>(q, w, w.Fx, g, eno, ctl, dt, counters, st, tm);
    launch_direction<2, SCHEME, RIEMANN>(q, w, w.Fy, g, eno, ctl, dt, counters, st, tm);
    launch_direction<3, SCHEME, RIEMANN>(q, w, w.Fz, g, eno, ctl, dt, counters, st, tm);
}

template <int SCHEME>
static void enqueue_fluxes_s(int riemann, double *q, const Workspace3 &w, const Grid3 &g, int eno, const Ctl *ctl,
                             double dt, unsigned long long *counters, hipStream_t st, Timer3 *tm) {
    if (riemann == WSR_RIEMANN_EVILIN)
        enqueue_fluxes_t<SCHEME, WSR_RIEMANN_EVILIN>(q, w, g, eno, ctl, dt, counters, st, tm);
    else
        enqueue_fluxes_t<SCHEME, WSR_RIEMANN_HLLC>(q, w, g, eno, ctl, dt, counters, st, tm);
}

// one stage: boundary fill, x / y / z fluxes, update (solver.py:285-301)
static void enqueue_stage3(const wsr_euler3_config *cfg, const Workspace3 &w, double *q, int rk, double dt, Ctl *ctl,
                           int fold_speed, double *dts, hipStream_t st, Timer3 *tm = nullptr) {
    Grid3 g = make_grid3(cfg);
    hipLaunchKernelGGL(k3_boundary, dim3(blocks_for(g.np)), dim3(BLOCK), 0, st, q, g, (const Ctl *)ctl);
    unsigned long long *cnt = (cfg->flags & WSR_FLAG_COUNT_CLOSURE) ? w.ctl->counters : nullptr;
    const int eno = cfg->eno_cutoff, rs = cfg->riemann;
    switch (cfg->scheme) {
    case WSR_SCHEME_JS: enqueue_fluxes_s<SCHEME_JS>(rs, q, w, g, eno, ctl, dt, cnt, st, tm); break;
    case WSR_SCHEME_SR64: enqueue_fluxes_s<SCHEME_SR64>(rs, q, w, g, eno, ctl, dt, cnt, st, tm); break;
    case WSR_SCHEME_SR32: enqueue_fluxes_s<SCHEME_SR32>(rs, q, w, g, eno, ctl, dt, cnt, st, tm); break;
    default: enqueue_fluxes_s<SCHEME_SRBF16>(rs, q, w, g, eno, ctl, dt, cnt, st, tm); break;
    }
    const long long ncells = (long long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k3_update, dim3(blocks_for(ncells)), dim3(BLOCK), 0, st, q, w.q0, (const double *)w.Fx,
                       (const double *)w.Fy, (const double *)w.Fz, g, rk, ctl, fold_speed, dts);
}

}  // namespace e3
}  // namespace wsr

using namespace wsr;
using namespace wsr::e3;

extern "C" {

size_t wsr_euler3_workspace_bytes(const wsr_euler3_config *cfg) {
    if (check3(cfg) == WSR_E_ARG) return 0;
    return ws3_layout(cfg, nullptr, nullptr);
}

int wsr_euler3_step(const wsr_euler3_config *cfg, double *q, double dt, void *workspace, size_t workspace_bytes,
                    void *stream) {
    int rc = check3(cfg);
    if (rc) return rc;
    if (!q || !workspace || !(dt > 0.0) || !isfinite(dt)) return WSR_E_ARG;
    if (workspace_bytes < ws3_layout(cfg, nullptr, nullptr)) return WSR_E_WORKSPACE;
    Workspace3 w;
    ws3_layout(cfg, &w, workspace);
    hipStream_t st = (hipStream_t)stream;
    enqueue_params3(cfg, w, st);
    for (int rk = 1; rk <= 3; ++rk) enqueue_stage3(cfg, w, q, rk, dt, nullptr, 0, nullptr, st);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_euler3_max_speed(const wsr_euler3_config *cfg, const double *q, void *workspace, size_t workspace_bytes,
                         double *max_speed_out, void *stream) {
    int rc = check3(cfg);
    if (rc == WSR_E_ARG) return rc;
    if (!q || !workspace || !max_speed_out) return WSR_E_ARG;
    if (workspace_bytes < ws3_layout(cfg, nullptr, nullptr)) return WSR_E_WORKSPACE;
    Workspace3 w;
    ws3_layout(cfg, &w, workspace);
    hipStream_t st = (hipStream_t)stream;
    Grid3 g = make_grid3(cfg);
    hipLaunchKernelGGL(k3_reset_speed, dim3(1), dim3(1), 0, st, w.ctl);
    hipLaunchKernelGGL(k3_max_speed, dim3(blocks_for((long long)g.nx * g.ny * g.nz)), dim3(BLOCK), 0, st, q, g,
                       w.ctl);
    Ctl host;
    if (hipMemcpyAsync(&host, w.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st) != hipSuccess) return WSR_E_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return WSR_E_HIP;
    double m;
    memcpy(&m, &host.speed_bits, sizeof m);
    *max_speed_out = host.nan_flag ? NAN : m;
    return WSR_OK;
}

int wsr_euler3_loop_init(const wsr_euler3_config *cfg, const double *q, int32_t max_steps, void *workspace,
                         size_t workspace_bytes, void *stream) {
    int rc = check3(cfg);
    if (rc) return rc;
    if (!q || !workspace || max_steps < 0) return WSR_E_ARG;
    if (workspace_bytes < ws3_layout(cfg, nullptr, nullptr)) return WSR_E_WORKSPACE;
    Workspace3 w;
    ws3_layout(cfg, &w, workspace);
    hipStream_t st = (hipStream_t)stream;
    Grid3 g = make_grid3(cfg);
    double h = fmin(cfg->dx, fmin(cfg->dy, cfg->dz));   // python min(dx, min(dy, dz))
    hipLaunchKernelGGL(k3_loop_init, dim3(1), dim3(1), 0, st, w.ctl, cfg->t_end, cfg->cfl * h, max_steps);
    hipLaunchKernelGGL(k3_max_speed, dim3(blocks_for((long long)g.nx * g.ny * g.nz)), dim3(BLOCK), 0, st, q, g,
                       w.ctl);
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_euler3_loop_steps(const wsr_euler3_config *cfg, double *q, double *dts, int32_t n_steps, void *workspace,
                          size_t workspace_bytes, void *stream) {
    int rc = check3(cfg);
    if (rc) return rc;
    if (!q || !dts || !workspace || n_steps < 0) return WSR_E_ARG;
    if (workspace_bytes < ws3_layout(cfg, nullptr, nullptr)) return WSR_E_WORKSPACE;
    Workspace3 w;
    ws3_layout(cfg, &w, workspace);
    hipStream_t st = (hipStream_t)stream;
    if (n_steps > 0) enqueue_params3(cfg, w, st);
    for (int n = 0; n < n_steps; ++n) {
        hipLaunchKernelGGL(k3_dt, dim3(1), dim3(1), 0, st, w.ctl);
        enqueue_stage3(cfg, w, q, 1, 0.0, w.ctl, 0, dts, st);
        enqueue_stage3(cfg, w, q, 2, 0.0, w.ctl, 0, dts, st);
        enqueue_stage3(cfg, w, q, 3, 0.0, w.ctl, 1, dts, st);
    }
    return hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
}

int wsr_euler3_loop_steps_timed(const wsr_euler3_config *cfg, double *q, double *dts, int32_t n_steps,
                                void *workspace, size_t workspace_bytes, void *stream, double *kernel_ms,
                                int32_t *launches) {
    int rc = check3(cfg);
    if (rc) return rc;
    if (!q || !dts || !workspace || n_steps < 0 || n_steps > TIMER3_MAX_STEPS || !kernel_ms || !launches)
        return WSR_E_ARG;
    if (workspace_bytes < ws3_layout(cfg, nullptr, nullptr)) return WSR_E_WORKSPACE;
    Workspace3 w;
    ws3_layout(cfg, &w, workspace);
    hipStream_t st = (hipStream_t)stream;
    static_assert(sizeof(Timer3) < (1u << 20), "timer record size");
    Timer3 *tm = new Timer3();
    tm->cap = 27 * n_steps;
    for (int i = 0; i < 2 * tm->cap; ++i)
        if (hipEventCreate(&tm->ev[i]) != hipSuccess) {
            while (i-- > 0) (void)hipEventDestroy(tm->ev[i]);
            delete tm;
            return WSR_E_HIP;
        }
    int32_t step0 = 0, step1 = 0;
    if (!loop_step_count(w.ctl, st, &step0)) rc = WSR_E_HIP;
    if (n_steps > 0) enqueue_params3(cfg, w, st);
    for (int n = 0; n < n_steps; ++n) {
        hipLaunchKernelGGL(k3_dt, dim3(1), dim3(1), 0, st, w.ctl);
        enqueue_stage3(cfg, w, q, 1, 0.0, w.ctl, 0, dts, st, tm);
        enqueue_stage3(cfg, w, q, 2, 0.0, w.ctl, 0, dts, st, tm);
        enqueue_stage3(cfg, w, q, 3, 0.0, w.ctl, 1, dts, st, tm);
    }
    if (rc == WSR_OK) rc = hipGetLastError() == hipSuccess ? WSR_OK : WSR_E_HIP;
    if (!loop_step_count(w.ctl, st, &step1)) rc = WSR_E_HIP;
    // only the launches of the steps that ran: 27 timed launches per step, in order
    int ran = step1 - step0;
    ran = ran < 0 ? 0 : (ran > n_steps ? n_steps : ran);
    for (int k = 0; k < 3; ++k) { kernel_ms[k] = 0.0; launches[k] = 0; }
    for (int i = 0; i < tm->used && i < 27 * ran; ++i) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, tm->ev[2 * i], tm->ev[2 * i + 1]) == hipSuccess) {
            kernel_ms[tm->kind[i]] += ms;
            launches[tm->kind[i]] += 1;
        }
    }
    for (int i = 0; i < 2 * tm->cap; ++i) (void)hipEventDestroy(tm->ev[i]);
    delete tm;
    return rc;
}

}  // extern "C"
