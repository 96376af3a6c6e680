"""Paper runners with the reference's module names, on the HIP solvers.

``config``                    <- code/snapshot/for_paper_results/config.py (methods, checkpoints, output roots)
``run_sod``                   <- code/snapshot/for_paper_results/run_sod.py
``run_shock_tube_benchmark``  <- code/snapshot/for_paper_results/run_shock_tube_benchmark.py (Lax, blast half)
``run_riemann_1d``            <- code/snapshot/weno_z_borges_p2_results/run_riemann_1d.py (WENO-Z rows)
``riemann_tables``            <- the error tables of weno_z_borges_p2_results/plot_riemann_1d_zoom.py:95-140
                                 and plot_lax_paper.py:22-151 (no plotting)

Output files (``.npz`` with ``state`` and ``metadata_json``, ``.json`` rows,
``metrics.csv``) keep the reference's names and fields; they go under
``config.RAW`` (``$WSR_PAPER_OUT/raw``, default ``./paper_results/raw``).
"""
