/*
 * wsr.h -- C ABI of the MI355X WENO5-SR finite-volume library (libwsr_hip.so).
 *
 * All state pointers are DEVICE pointers (owned by the caller, e.g. torch
 * tensors) in the reference layout: FP64, C-order (ny + 2*ghost, nx + 2*ghost, 4),
 * conserved variables (rho, rho*u, rho*v, E), ghost = 3.  Every call is
 * stream-ordered on `stream` (a hipStream_t; NULL = default stream) and returns
 * 0 on success or a negative WSR_E* code; nothing throws across the ABI.  The
 * library keeps no global mutable state: per-run state lives in the caller's
 * workspace buffer.
 *
 * Reference paths are relative to the reference checkout.
 */
#ifndef WSR_H
#define WSR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reconstruction closures */
enum {
    WSR_SCHEME_JS = 0,     /* WENO5-JS (code/snapshot/warp_weno5_helpers.py:511-570) */
    WSR_SCHEME_SR64 = 1,   /* WENO5-SR, FP64 MLP (teacherfree_lab_weno5_v20_distance_balanced/weno5_rk3_diff_v20_deploy.py:155-281) */
    WSR_SCHEME_SR32 = 2,   /* WENO5-SR, FP32 MLP (teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/warp_mlp_f32/weno5_rk3_diff_mlp_f32.py:152-283) */
    WSR_SCHEME_SRBF16 = 3, /* WENO5-SR, BF16 MLP parameters/activations (no reference counterpart) */
    WSR_SCHEME_Z = 4       /* WENO5-Z p=2 (code/snapshot/weno_z_borges_p2_results/weno5_helpers_z.py:511-570) */
};
/* Riemann fluxes */
enum {
    WSR_RIEMANN_HLLC = 0,  /* code/snapshot/for_paper_results/solvers/weno5_hllc.py:23-115 */
    WSR_RIEMANN_FORCE = 1, /* code/snapshot/warp_weno5_helpers.py:401-435 */
    WSR_RIEMANN_EVILIN = 2 /* WENO7 2-D: code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/
                              warp_weno7_ader4_helpers_classical_only.py:256-342,429-433;
                              3-D WENO5: code/snapshot/warp_weno5_3d_rk3/kernels.py:122-219 */
};
/* boundary conditions */
enum {
    WSR_BC_TRANSMISSIVE = 0, /* weno5_rk3_diff_v20_deploy.py:391-409 */
    WSR_BC_PERIODIC = 1,     /* weno5_rk3_diff_v20_deploy.py:371-388 */
    WSR_BC_DOUBLE_MACH = 2   /* teacherfree_lab_weno5_v20_distance_balanced/weno5_hllc_refsym.py:25-124 */
};
/* error codes */
enum {
    WSR_OK = 0,
    WSR_E_ARG = -1,        /* bad size / enum / null pointer; a 2-D or WENO7 state, buffer or workspace
                              pointer that is not 16-byte aligned (cells and faces move as 16-byte halves) */
    WSR_E_MLP = -2,        /* SR scheme without MLP parameters */
    WSR_E_WORKSPACE = -3,  /* workspace too small */
    WSR_E_HIP = -4         /* a HIP runtime call failed */
};

/* instrumentation flags (wsr_euler2d_config.flags) */
enum {
    WSR_FLAG_COUNT_CLOSURE = 1, /* count reconstructions and closure (MLP) evaluations into the workspace */
    WSR_FLAG_Z_FALLBACK = 2,    /* WENO7-SR: per reconstructed value, the closure output is rejected and the
                                   value reconstructed with WENO7-Z weights when a ratio is not finite or outside
                                   [0, 1], or when the candidate the closure weights most (argmax d_k/(beta_k +
                                   1e-12)^2) has a WENO7-Z indicator above 50 x the smallest one + 240 h^3 (it
                                   crosses a discontinuity).  Counted in counter 6 (wsr_euler2d_counters).  No
                                   reference counterpart. */
    WSR_FLAG_Z7_UNIT_GAUSS_H = 4 /* WENO7-Z: transverse Gauss-point weights with h = 1 (epsilon = 240) instead of
                                   the transverse spacing.  The reference's source passes the spacing
                                   (weno_z_borges_p2_results/weno7_point_rk4_z.py:473-500); its recorded 2-D
                                   WENO7-Z results (experiments/07-09 line cuts) follow h = 1. */
};

/* Grid and method description (a plain C struct; mirrors the fields of
 * Params, code/snapshot/warp_weno5_helpers.py:66-87, plus the method switches). */
typedef struct wsr_euler2d_config {
    int32_t nx, ny, ghost;          /* ghost must be 3 */
    int32_t scheme, riemann, boundary;
    int32_t eno_cutoff;             /* 0/1, SR only (weno5_rk3_diff_v20_deploy.py:258-279) */
    int32_t flags;                  /* WSR_FLAG_* */
    double dx, dy, gamma;
    double cfl, t_end;              /* used by the device-resident time loop only */
    /* host pointer to the 189 closure parameters w1,b1,w2,b2,w3,b3,w4,b4
     * (row-major, no singleton axis): double for SR64, float for SR32/SRBF16;
     * NULL for JS and Z.  Copied into the workspace by each stage / loop call. */
    const void *mlp;
} wsr_euler2d_config;

/* Bytes of device workspace the calls below need (fluxes, boundary-filled
 * stage copy, device time-loop control block). */
size_t wsr_euler2d_workspace_bytes(int32_t nx, int32_t ny, int32_t ghost);

/* One SSPRK3 stage: boundary fill, x and y sweeps (characteristic normal
 * reconstruction, transverse two-point Gauss reconstruction, Riemann flux) and
 * the stage combination u_out = {rhs, 3/4 u0 + 1/4 rhs, 1/3 u0 + 2/3 rhs}.
 * Replaces launch_stage of for_paper_results/solvers/weno5_hllc_mixed.py:121-168,
 * teacherfree_lab_weno5_v20_distance_balanced/weno5_hllc_refsym.py:298-435 and
 * for_paper_results/solvers/weno5_hllc.py:390-454.  Only the interior of u_out
 * is written.  u_stage == u0 is allowed (stage 1); u_out may alias u0. */
int wsr_euler2d_stage(const wsr_euler2d_config *cfg, const double *u_stage, const double *u0,
                      double *u_out, double dt, int32_t rk, double boundary_time,
                      void *workspace, size_t workspace_bytes, void *stream);

/* Allocation-free variant of wsr_euler2d_stage.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and results
 * as wsr_euler2d_stage (the same kernels), without the workspace: the stage works in a
 * file-scope device array of WSR_NOALLOC_SCRATCH_BYTES (32 MiB; each GPU the library
 * runs on holds it until the process ends, and its device address is looked up once
 * per device).  Grids whose workspace (wsr_euler2d_workspace_bytes) exceeds it return
 * WSR_E_WORKSPACE (400 x 400 needs 15.5 MB; up to about 570 x 570 fits);
 * WSR_FLAG_COUNT_CLOSURE is refused (WSR_E_ARG).  The call allocates nothing.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race. */
#define WSR_NOALLOC_SCRATCH_BYTES (32u << 20)
int wsr_euler2d_stage_noalloc(const wsr_euler2d_config *cfg, const double *u_stage, const double *u0,
                              double *u_out, double dt, int32_t rk, double boundary_time, void *stream);

/* Maximum characteristic speed max(|u|+c, |v|+c) over the interior, reduced on
 * the device and returned to the host (synchronises the stream).  Replaces
 * compute_max_speed_kernel + np.max of warp_weno5_helpers.py:189-207,876-897;
 * NaN anywhere gives NaN like np.max. */
int wsr_euler2d_max_speed(const wsr_euler2d_config *cfg, const double *u, void *workspace,
                          size_t workspace_bytes, double *max_speed_out, void *stream);

/* Device-resident time loop (the `while t < t_end` loop of run_to_time,
 * weno5_hllc_mixed.py:185-212 / weno5_hllc_refsym.py:462-544 /
 * weno5_hllc.py:482-506) without a host round trip per step.
 *   loop_init : t = 0, step = 0, first max speed of u (enqueued, no sync).
 *   loop_steps: enqueue n steps; each step computes dt = min(cfl*min(dx,dy)/max_speed,
 *               t_end - t) on the device, stops the loop (later steps become no-ops)
 *               when t >= t_end - 1e-14, dt is non-finite or <= 0, or max_steps is
 *               reached, runs the three stages into u1, u2 and back into u, records
 *               dts[step] (device array of max_steps doubles) and folds the next
 *               max speed into the third stage.
 *   loop_status: synchronise and read step count, time, last dt and the done flag. */
int wsr_euler2d_loop_init(const wsr_euler2d_config *cfg, const double *u, int32_t max_steps,
                          void *workspace, size_t workspace_bytes, void *stream);
int wsr_euler2d_loop_steps(const wsr_euler2d_config *cfg, double *u, double *u1, double *u2,
                           double *dts, int32_t n_steps, void *workspace, size_t workspace_bytes,
                           void *stream);
int wsr_euler2d_loop_status(void *workspace, int32_t *steps, double *t, double *dt_last,
                            int32_t *done, void *stream);

/* Instrumented loop_steps (no reference counterpart; used by the benchmark):
 * the same n_steps (<= 64) steps with a HIP event pair around every sweep-kernel
 * launch (one launch sweeps x and y of a stage) on `stream`; synchronises and
 * returns the summed sweep-kernel time and number of the launches of the steps
 * the loop ran (read from the control block before and after; launches enqueued
 * after the loop stopped are not counted). */
int wsr_euler2d_loop_steps_timed(const wsr_euler2d_config *cfg, double *u, double *u1, double *u2,
                                 double *dts, int32_t n_steps, void *workspace, size_t workspace_bytes,
                                 void *stream, double *sweep_ms, int32_t *sweep_launches);

/* Closure counters accumulated since loop_init when WSR_FLAG_COUNT_CLOSURE is set;
 * copies the first n (1..7) of {normal reconstructions, of which closure
 * evaluated, Gauss-pair reconstructions, of which closure evaluated, lanes of
 * the waves that ran the normal closure, lanes of the waves that ran the Gauss
 * closure, WENO7 values that fell back to WENO7-Z} to out (synchronises the
 * stream).  The last one is collected whenever WSR_FLAG_Z_FALLBACK is set: each
 * used normal face value and each Gauss-point value on a grid face counts once.
 * wsr_euler7_rhs / wsr_euler7_step zero the counters at their start.  No
 * counterpart in the reference: an instrument for the closure fraction that the
 * roofline figures use and for the fallback rate. */
int wsr_euler2d_counters(void *workspace, uint64_t *out, int32_t n, void *stream);

/* ------------------------------------------------------------------------ */
/* 1-D periodic scalar advection (the training / GSTE path)                  */
/* ------------------------------------------------------------------------ */
typedef struct wsr_adv1d_config {
    int32_t batch, n;               /* trajectories x cells (5 <= n <= 2048) */
    int32_t scheme;                 /* WSR_SCHEME_JS / SR64 / SR32 / SRBF16 */
    int32_t reserved;
    double dxinv;                   /* 1 / dx */
    const void *mlp;                /* host pointer to the 189 parameters (as in wsr_euler2d_config) */
} wsr_adv1d_config;

/* Advance u (device, batch x n, FP64, in place) by n_steps signed SSPRK3 steps of
 * q_t + a q_x = 0 with velocities[b] in {-1, +1} (device) and the step sizes
 * dts[n_steps] (device).  Replaces ssprk3_step_signed / rollout_distance of
 * teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/rk3_advection.py:82-186
 * and ssprk3 of teacherfree_lab_weno5/apost_advect.py:41-45 (GSTE).  If
 * checkpoints is not NULL it receives the state at the start of every step and
 * the final state: (batch, n_steps + 1, n) doubles. */
int wsr_adv1d_rollout(const wsr_adv1d_config *cfg, double *u, const double *velocities,
                      const double *dts, int32_t n_steps, double *checkpoints, void *stream);

/* Training (no counterpart in the reference's C/CUDA: the reference differentiates
 * with torch autograd through checkpointed rollouts,
 * teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/fast_losses.py:412-648).
 * Parameters are read from device memory (params_device: 189 doubles for SR64,
 * floats for SR32) so an optimiser can update them in place.
 *   train_forward: rollout as wsr_adv1d_rollout, checkpoints required
 *                  ((batch, n_steps + 1, n) doubles).
 *   train_adjoint: reverse pass of the same rollout from the adjoint lam_final
 *                  (batch x n) of the final state: grad (batch x 189 doubles,
 *                  per-trajectory parameter gradients) and, if lam_initial is not
 *                  NULL, the adjoint of the initial state.  n <= 1024 (SR32) / 400 (SR64). */
int wsr_adv1d_train_forward(const wsr_adv1d_config *cfg, double *u, const double *velocities,
                            const double *dts, int32_t n_steps, double *checkpoints,
                            const void *params_device, void *stream);
int wsr_adv1d_train_adjoint(const wsr_adv1d_config *cfg, const double *checkpoints,
                            const double *lam_final, const double *velocities, const double *dts,
                            int32_t n_steps, const void *params_device, double *grad,
                            double *lam_initial, void *stream);

/* Library version / build identification. */
/* ------------------------------------------------------------------------ */
/* 2-D WENO7 with Shu's downwind fourth-order TVD Runge-Kutta (ghost 4)      */
/* ------------------------------------------------------------------------ */
/* time-loop rules */
enum {
    WSR_LOOP7_POINT = 0,       /* weno7_point_rk4_shu/point_rk4.py:557-575 run_from_initial */
    WSR_LOOP7_DOUBLE_MACH = 1  /* for_paper_results/solvers/weno7_double_mach.py:434-470 run_to_time */
};

/* Mirrors point_rk4.py:26-55 Params (nx, ny, ghost = 4, gamma, cfl, t_end,
 * dx = x_length / nx) plus the method switches.  scheme: WSR_SCHEME_JS
 * (WENO7-JS, warp helpers' expanded indicators), WSR_SCHEME_SR64 / _SR32
 * (WENO7-SR, 6-24-16-16-4 closure in FP64 / FP32), WSR_SCHEME_Z (WENO7-Z p=3);
 * mlp = the 908 parameters w1, b1, ..., b4 in the checkpoint order (FP64 for
 * SR64, FP32 for SR32). */
typedef struct wsr_euler7_config {
    int32_t nx, ny, ghost;
    int32_t scheme, riemann, boundary;
    int32_t eno_cutoff;
    int32_t flags;
    int32_t loop_mode;
    int32_t reserved;
    double dx, dy, gamma;
    double cfl, t_end;
    const void *mlp;
} wsr_euler7_config;

size_t wsr_euler7_workspace_bytes(const wsr_euler7_config *cfg);

/* L(u) into rhs and, when rhs_reverse != NULL, the downwind operator L~(u)
 * into rhs_reverse (padded arrays, ghost cells set to 0); replaces
 * point_rk4.py:410-468 compute_rhs (reverse_upwind 0 / 1) and
 * point_rk4_mlp.py:232-304 compute_rhs_mlp.  dt enters FORCE / EVILIN only;
 * time is the double-Mach boundary time. */
int wsr_euler7_rhs(const wsr_euler7_config *cfg, const double *u, double *rhs, double *rhs_reverse, double dt,
                   double time, void *workspace, size_t workspace_bytes, void *stream);

/* One Shu RK4 step of u in place (point_rk4.py:471-493 launch_shu_rk4_step,
 * point_rk4_mlp.py:307-327, weno7_double_mach.py:354-431). */
int wsr_euler7_step(const wsr_euler7_config *cfg, double *u, double dt, double time, void *workspace,
                    size_t workspace_bytes, void *stream);

/* Device-resident time loop (replaces the host loops of point_rk4.py:537-575
 * and weno7_double_mach.py:434-470): loop_init resets the clock and reduces
 * the max speed of u; loop_steps enqueues n_steps steps (dt on the device,
 * dts[step] = dt); wsr_euler2d_loop_status / wsr_euler2d_counters read the
 * same control block. */
int wsr_euler7_loop_init(const wsr_euler7_config *cfg, const double *u, int32_t max_steps, void *workspace,
                         size_t workspace_bytes, void *stream);
int wsr_euler7_loop_steps(const wsr_euler7_config *cfg, double *u, double *dts, int32_t n_steps, void *workspace,
                          size_t workspace_bytes, void *stream);

/* Instrumented loop_steps (no reference counterpart; used by the benchmark):
 * n_steps (<= 32) steps with a HIP event pair around every sweep launch (one
 * launch sweeps x and y of a right-hand side) on `stream`; synchronises and
 * returns the summed sweep time (ms) and launch count of the sweeps of the steps
 * the loop ran, [0] for the sweeps that serve L and L~, [1] for the L-only sweeps. */
int wsr_euler7_loop_steps_timed(const wsr_euler7_config *cfg, double *u, double *dts, int32_t n_steps,
                                void *workspace, size_t workspace_bytes, void *stream, double *sweep_ms,
                                int32_t *launches);

/* Ghost cells of u in place for cfg->boundary (4 ghost layers; double Mach at
 * `time`): apply_boundary_kernel / apply_periodic_boundary_kernel of
 * teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/warp_weno7_ader4_helpers_classical_only.py:2160-2200
 * as point_rk4.py:453-455 launches them. */
int wsr_euler7_boundary(const wsr_euler7_config *cfg, double *u, double time, void *stream);

/* Max characteristic speed over the interior of u, returned to the host (synchronises
 * the stream; NaN if any cell is NaN): the kernel and np.max of point_rk4.py:136-147
 * compute_dt_from_warp_array.  The workspace needs WSR_SPEED_WORKSPACE_BYTES only (a
 * wsr_euler7_workspace_bytes workspace serves as well). */
#define WSR_SPEED_WORKSPACE_BYTES 4096u
int wsr_euler7_max_speed(const wsr_euler7_config *cfg, const double *u, void *workspace, size_t workspace_bytes,
                         double *max_speed_out, void *stream);

/* ---- beta-provider protocol of the WENO7-SR point step (unfused path) ----
 * The reference drives WENO7-SR through a beta provider: an object whose methods
 * fill arrays of smoothness indicators (beta = 4 r of the reflection-averaged
 * 6-24-16-16-4 closure, ones on a plateau), which separate kernels consume
 * (teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/point_rk4_mlp.py:22-56,
 * 194-273).  These entry points restore that interface; the fused wsr_euler7_* step
 * is the production path and gives the same results for this library's closure.
 * Arrays (FP64, C order, the reference's allocations):
 *   beta_x (ny+8, nx+2, 32) / beta_y (ny+2, nx+8, 32): entry side*16 + comp*4 + k
 *     (side 0: lr 1, the right face of the window's centre cell; side 1: lr 2);
 *   beta_cross_x (ny, nx+2, 256) / beta_cross_y (ny+2, nx, 256): entry
 *     group*128 + side*16 + comp*4 + k (group 0 from x_r / y_r, 1 from x_l / y_l;
 *     side 0 serves Gauss location 1, side 1 location 2);
 *   face values x_l, x_r, y_l, y_r and rhs arrays: padded (ny+8, nx+8, 4);
 *   fluxes fx (ny, nx+1, 4), fy (ny+1, nx, 4).
 * dir 1 = x, 2 = y.  The fill calls need cfg->scheme SR64 or SR32 and cfg->mlp, and
 * a workspace of wsr_beta7_workspace_bytes() for the parameters. */
size_t wsr_beta7_workspace_bytes(void);

/* TorchWeno7PointBeta.fill_normal_x_for / fill_normal_y_for (point_rk4_mlp.py:24-37;
 * for_paper_results/solvers/weno7_warp_beta.py:318-426, 616-648): the normal beta of
 * every window of u (characteristic components, Roe state of each side's face). */
int wsr_beta7_fill_normal(const wsr_euler7_config *cfg, const double *u, double *beta, int32_t dir, void *workspace,
                          size_t workspace_bytes, void *stream);

/* fill_cross_x_point / fill_cross_y_point (point_rk4_mlp.py:38-56;
 * weno7_warp_beta.py:445-560, 650-682): the cross beta of the transverse windows of
 * the face-value arrays (conservative components). */
int wsr_beta7_fill_cross(const wsr_euler7_config *cfg, const double *right, const double *left, double *beta,
                         int32_t dir, void *workspace, size_t workspace_bytes, void *stream);

/* compute_x_stage_point_mlp_kernel / compute_y_stage_point_mlp_kernel
 * (point_rk4_mlp.py:61-109 with warp_weno7_external_overwrite_fullall.py:84-211):
 * both normal face values of every window from external beta (lr 1 -> right,
 * lr 2 -> left). */
int wsr_euler7_stage_ext(const wsr_euler7_config *cfg, const double *u, const double *beta, double *left,
                         double *right, int32_t dir, void *stream);

/* compute_x_flux_point_mlp_kernel / compute_y_flux_point_mlp_kernel
 * (point_rk4_mlp.py:111-186): the cfg->riemann flux at Gauss location loca (1 or 2)
 * of every face from external cross beta; reverse_upwind swaps the states; dt enters
 * FORCE / EVILIN only. */
int wsr_euler7_flux_ext(const wsr_euler7_config *cfg, const double *left, const double *right,
                        const double *beta_cross, double *flux, int32_t dir, int32_t loca, double dt,
                        int32_t reverse_upwind, void *stream);

/* rhs_from_flux_kernel (point_rk4.py:290-312): interior cells of rhs from the four
 * flux arrays (ghost entries are not written). */
int wsr_euler7_rhs_from_flux(const wsr_euler7_config *cfg, double *rhs, const double *fx1, const double *fx2,
                             const double *fy1, const double *fy2, void *stream);

/* shu_stage1/2/3/final kernels (point_rk4.py:315-403) over every padded entry:
 * stage 1 inputs (u, rhs0); 2 (u, rhs_t0, u1, rhs1); 3 (u, rhs_t0, u1, rhs_t1, u2,
 * rhs2); 4 (u, rhs0, u1, rhs1, u2, u3, rhs3); out may alias an input only for stage 4
 * and only the first (the reference writes the new state over u). */
int wsr_euler7_shu_combine(const wsr_euler7_config *cfg, int32_t stage, double dt, const double *const *inputs,
                           double *out, void *stream);

/* ------------------------------------------------------------------ 3-D WENO5
 * The 3-D characteristic WENO5 step of code/snapshot/warp_weno5_3d_rk3:
 * state (nz+6, ny+6, nx+6, 5) FP64 conserved (rho, rho u, rho v, rho w, E),
 * ghost 3; per direction a characteristic normal reconstruction (5 x 5
 * eigenvectors, kernels.py:272-498, 870-942), two transverse Gauss passes
 * giving the four face points (kernels.py:941-1014, solver.py:165-283), the
 * flux at each point (EVILIN, kernels.py:122-219, 1016-1090, or the
 * five-variable HLLC, which the reference does not have in 3-D) and the
 * SSPRK3 update (kernels.py:1092-1163).  Schemes: JS, SR64, SR32, SRBF16
 * (no WENO-Z in the reference's 3-D path).  Boundary: transmissive. */
typedef struct wsr_euler3_config {
    int32_t nx, ny, nz, ghost;
    int32_t scheme, riemann, boundary;
    int32_t eno_cutoff;
    int32_t flags;
    int32_t reserved;
    double dx, dy, dz, gamma;
    double cfl, t_end;
    const void *mlp;
} wsr_euler3_config;

size_t wsr_euler3_workspace_bytes(const wsr_euler3_config *cfg);

/* One SSPRK3 step of q in place with a given dt (solver.py:285-301 _rk_stage
 * x 3 as called by advance, solver.py:303-329). */
int wsr_euler3_step(const wsr_euler3_config *cfg, double *q, double dt, void *workspace, size_t workspace_bytes,
                    void *stream);

/* Max over the interior of max(|u|+a, |v|+a, |w|+a) (solver.py:159-163,
 * kernels.py:855-867), NaN if any cell's speed is NaN. */
int wsr_euler3_max_speed(const wsr_euler3_config *cfg, const double *q, void *workspace, size_t workspace_bytes,
                         double *max_speed_out, void *stream);

/* Device-resident time loop replacing solver.py:303-329 advance and the
 * run_solver loop (solver.py:348-...): dt = cfl min(dx,dy,dz) / max speed
 * (1e10 below 1e-15), clipped to t_end; stops at t == t_end or max_steps.
 * wsr_euler2d_loop_status reads the same control block. */
int wsr_euler3_loop_init(const wsr_euler3_config *cfg, const double *q, int32_t max_steps, void *workspace,
                         size_t workspace_bytes, void *stream);
int wsr_euler3_loop_steps(const wsr_euler3_config *cfg, double *q, double *dts, int32_t n_steps, void *workspace,
                          size_t workspace_bytes, void *stream);

/* Instrumented loop_steps (no reference counterpart; used by the benchmark):
 * the same n_steps (<= 32) steps with a HIP event pair around every normal,
 * first-transverse-pass and flux launch on `stream`; synchronises and returns
 * the summed time (ms) and the number of launches of the steps the loop ran
 * for each of the three kernel kinds (kernel_ms[3], launches[3]).  With
 * WSR_FLAG_COUNT_CLOSURE in cfg->flags the closure counters of
 * wsr_euler2d_counters accumulate over the 3-D passes as well. */
int wsr_euler3_loop_steps_timed(const wsr_euler3_config *cfg, double *q, double *dts, int32_t n_steps,
                                void *workspace, size_t workspace_bytes, void *stream, double *kernel_ms,
                                int32_t *launches);

const char *wsr_version(void);

/* ------------------------------------------------------------------------ */
/* Initial states built on the device (cell averages by tensor-product       */
/* Gauss-Legendre quadrature; nodes / weights are host arrays of nq <= 16     */
/* values).  Each equals the host builder of weno_sr_publishing_amd/initial.py */
/* bit for bit (the vortex up to the last bits of exp and pow).               */
/* ------------------------------------------------------------------------ */

/* Four-quadrant Riemann problem: states[16] = (rho, u, v, p) of the upper-right,
 * upper-left, lower-left and lower-right quadrants (split at x = y = 0.5).  wrap = 0:
 * every padded cell holds its own average (for_paper_results/run_quadrant.py:87-106
 * make_quadrant_state); wrap = 1: the interior averages with the ghosts filled by
 * periodic wrap (the WENO5 re-pad of run_weno5_timing.py:86-104 / run_vortex.py:34-35). */
int wsr_init_quadrant(double *u, int32_t nx, int32_t ny, int32_t ghost, int32_t wrap, const double *states,
                      double x_min, double x_max, double y_min, double y_max, double gamma, const double *nodes,
                      const double *weights, int32_t nq, void *stream);

/* Sod strip centred on 0 along axis 1 (x) or 2 (y), rho/p = 1/1 left, 0.125/0.1 right.
 * Replaces teacherfree_lab_weno5/warp_sod_validation.py:176-210. */
int wsr_init_sod(double *u, int32_t nx, int32_t ny, int32_t ghost, int32_t axis, double x_length, double y_length,
                 double gamma, const double *nodes, const double *weights, int32_t nq, void *stream);

/* Double-Mach reflection at t = 0 (Mach-10 shock through x = 1/6 at 60 degrees).
 * Replaces run_double_mach_compare.py:69-88 make_initial_state. */
int wsr_init_double_mach(double *u, int32_t nx, int32_t ny, int32_t ghost, double x_length, double y_length,
                         double gamma, const double *nodes, const double *weights, int32_t nq, void *stream);

/* Periodic isentropic vortex (strength 5, advected with (1, 1), period 20) at time t.
 * Replaces weno7_point_rk4_shu/point_rk4.py:102-117 cell_average_state. */
int wsr_init_vortex(double *u, int32_t nx, int32_t ny, int32_t ghost, double t, double x_min, double x_max,
                    double y_min, double y_max, double gamma, const double *nodes, const double *weights, int32_t nq,
                    void *stream);

/* 3-D shock-bubble, state (nz+2g, ny+2g, nx+2g, 5).  kind 0: the unit-spacing
 * benchmark case (shock at x < 4, bubble of radius ny/4 centred at (ny/2 + 8, ny/2,
 * nz/2) with density density_ratio; lengths = (nx, ny, nz)); kind 1: Ma = 3 in SI
 * units (warp_weno5_3d_rk3/kernels.py:730-799; density_ratio = helium density).
 * shock[3] = post-shock (rho, u, p). */
int wsr_init_shockbubble3d(double *q, int32_t nx, int32_t ny, int32_t nz, int32_t ghost, int32_t kind,
                           const double *lengths, const double *shock, double density_ratio, double gamma,
                           const double *nodes, const double *weights, int32_t nq, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* WSR_H */
