"""Paper runners with the reference's module names, on the HIP solvers.

``config``                    <- code/snapshot/for_paper_results/config.py (methods, checkpoints, output roots)
``run_sod``                   <- code/snapshot/for_paper_results/run_sod.py
``run_shock_tube_benchmark``  <- code/snapshot/for_paper_results/run_shock_tube_benchmark.py (Lax, blast half)
``run_riemann_1d``            <- code/snapshot/weno_z_borges_p2_results/run_riemann_1d.py (WENO-Z rows)
``riemann_tables``            <- the error tables of weno_z_borges_p2_results/plot_riemann_1d_zoom.py:95-140
                                 and plot_lax_paper.py:22-151 (no plotting)
``run_quadrant``              <- code/snapshot/for_paper_results/run_quadrant.py (C.3-C.6 four-quadrant problems)
``riemann_linecuts``          <- the line-cut tables of weno_z_borges_p2_results/plot_riemann.py:44-57,180-235
``run_vortex``                <- code/snapshot/for_paper_results/run_vortex.py (isentropic vortex order table)
``run_gste``                  <- code/snapshot/for_paper_results/run_gste.py (1-D advection GSTE rollouts)
``run_weno5_timing``          <- code/snapshot/for_paper_results/run_weno5_timing.py
``run_double_mach``           <- code/snapshot/for_paper_results/run_double_mach.py (+ the WENO-Z fields and
                                 the line cuts of weno_z_borges_p2_results/plot_double_mach.py)

Output files (``.npz`` with ``state`` and ``metadata_json``, ``.json`` rows,
``metrics.csv``) keep the reference's names and fields; they go under
``config.RAW`` (``$WSR_PAPER_OUT/raw``, default ``./paper_results/raw``).
"""
