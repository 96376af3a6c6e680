// wsr_loop.hpp -- state shared by the 2-D step kernels: the device-resident
// time-loop control block, the grid record, the CFL speed reduction, closure
// counters and the double-Mach boundary state.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "wsr_math.hpp"

namespace wsr {

// counters 0-5: closure evaluations (WSR_FLAG_COUNT_CLOSURE); 6: WENO7 values that fell
// back to WENO7-Z (WSR_FLAG_Z_FALLBACK)
constexpr int NCOUNT = 7;
constexpr int SPEED_SLOTS = 16, SPEED_SLOT_STRIDE = 16;

struct Ctl {
    double t, dt, t_end, cfl_h, dt_last;
    unsigned long long speed_bits;
    int32_t step, done, nan_flag, max_steps;
    // closure evaluations (WSR_FLAG_COUNT_CLOSURE): [0] normal reconstructions,
    // [1] of which evaluated the closure (not a plateau), [2] Gauss-pair
    // reconstructions, [3] of which evaluated the closure, [4] / [5] lanes of
    // the waves that ran the closure code (normal / Gauss; 64 per wave pass);
    // [6] WENO7 reconstructed values that fell back to WENO7-Z (WSR_FLAG_Z_FALLBACK)
    unsigned long long counters[NCOUNT];
    // per-block maxima of the fused speed reduction, spread over SPEED_SLOTS
    // addresses on separate 128-byte lines (one address would serialise every
    // block's atomic in one L2 channel); the dt kernel folds them into speed_bits
    unsigned long long speed_slot[SPEED_SLOTS * SPEED_SLOT_STRIDE];
};

struct Grid {
    int nx, ny, nxt, nyt;
    double dx, dy, gamma;
};

// ---------------------------------------------------------------- boundary
__device__ __forceinline__ void dm_exact(double x, double y, double t, double gamma, double q[4]) {
    // weno5_hllc_refsym.py:25-54 double_mach_exact_conserved
    const double root3 = 1.7320508075688772;
    double x0 = 1.0 / 6.0;
    double shock_x = x0 + y / root3 + 20.0 * t / root3;
    double rho, u, v, p;
    if (x < shock_x) { rho = 8.0; u = 8.25 * root3 * 0.5; v = -8.25 * 0.5; p = 116.5; }
    else { rho = 1.4; u = 0.0; v = 0.0; p = 1.0; }
    double e = 0.5 * rho * (u * u + v * v) + p / (gamma - 1.0);
    q[0] = rho; q[1] = rho * u; q[2] = rho * v; q[3] = e;
}

// per-lane closure counts, summed over the wave and added once per wave at the end
struct LaneCounts {
    unsigned c[NCOUNT] = {};
    __device__ __forceinline__ void add(int slot, int wave_slot, bool counted, bool evaluated) {
        c[slot] += counted ? 1u : 0u;
        c[slot + 1] += (counted && evaluated) ? 1u : 0u;
        c[wave_slot] += __ballot(evaluated) != 0ull ? 1u : 0u;
    }
    __device__ __forceinline__ void flush(unsigned long long *dst) {
#pragma unroll
        for (int i = 0; i < NCOUNT; ++i) {
            unsigned v = c[i];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
            if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst + i, (unsigned long long)v);
        }
    }
    // the same sums gathered per block in LDS first: one global atomic per block
    // and counter (every thread of the block must call it)
    __device__ __forceinline__ void flush_block(unsigned long long *dst) {
        __shared__ unsigned blk[NCOUNT];
        if (threadIdx.x < NCOUNT) blk[threadIdx.x] = 0u;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NCOUNT; ++i) {
            unsigned v = c[i];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
            if ((threadIdx.x & 63) == 0 && v) atomicAdd(blk + i, v);
        }
        __syncthreads();
        if (threadIdx.x < NCOUNT && blk[threadIdx.x]) atomicAdd(dst + threadIdx.x, (unsigned long long)blk[threadIdx.x]);
    }
};

__device__ __forceinline__ double cell_speed(const double q[4], double gamma) {
    // warp_weno5_helpers.py:876-897 compute_max_speed_kernel
    double rho = dmax(q[0], 1.0e-15);
    double vx = q[1] / rho, vy = q[2] / rho;
    double p = (gamma - 1.0) * (q[3] - 0.5 * rho * (vx * vx + vy * vy));
    p = dmax(p, 1.0e-15);
    double c = sqrt(gamma * p / rho);
    return dmax(fabs(vx) + c, fabs(vy) + c);
}

// block max of non-negative speeds (NaN reported separately) -> atomics in ctl
template <int NTHREADS, bool SLOTS = false>
__device__ __forceinline__ void reduce_speed(double sp, bool active, Ctl *ctl) {
    __shared__ unsigned long long red[NTHREADS / 64];
    __shared__ int nan_seen;
    if (threadIdx.x == 0) nan_seen = 0;
    __syncthreads();
    unsigned long long bits = 0ull;
    if (active) {
        if (sp != sp) nan_seen = 1;
        else bits = (unsigned long long)__double_as_longlong(sp);
    }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_down(bits, off);
        bits = o > bits ? o : bits;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = bits;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = red[0];
        for (int w = 1; w < NTHREADS / 64; ++w) m = red[w] > m ? red[w] : m;
        if constexpr (SLOTS) atomicMax(&ctl->speed_slot[(blockIdx.x % SPEED_SLOTS) * SPEED_SLOT_STRIDE], m);
        else atomicMax(&ctl->speed_bits, m);
        if (nan_seen) atomicOr(&ctl->nan_flag, 1);
    }
}

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// State, buffer and workspace pointers handed to the C ABI must be 16-byte aligned:
// the kernels move the 32-byte cell and face records as 16-byte halves.  Null
// pointers pass (each entry point checks the ones it needs separately).
inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }
template <class... P> inline bool all_aligned16(const P *...p) { return (aligned16(p) && ...); }

// Steps the device time loop has completed (synchronises the stream first); the
// timed loop entry points use it to count only the launches of steps that ran --
// the launches enqueued after the loop stopped return at once.
inline bool loop_step_count(const Ctl *ctl, hipStream_t st, int32_t *step) {
    Ctl h;
    if (hipStreamSynchronize(st) != hipSuccess) return false;
    if (hipMemcpy(&h, ctl, sizeof(Ctl), hipMemcpyDeviceToHost) != hipSuccess) return false;
    *step = h.step;
    return true;
}

}  // namespace wsr
