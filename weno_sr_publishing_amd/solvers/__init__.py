"""Drop-in adapters with the reference's solver entry points.

``weno5_hllc``        <- code/snapshot/for_paper_results/solvers/weno5_hllc.py (WENO5-JS, or inline FP64 SR)
``weno5_hllc_refsym`` <- code/snapshot/teacherfree_lab_weno5_v20_distance_balanced/weno5_hllc_refsym.py (WENO5-SR FP64)
``weno5_hllc_mixed``  <- code/snapshot/for_paper_results/solvers/weno5_hllc_mixed.py (WENO5-SR FP32 MLP)
``euler_methods``     <- code/snapshot/for_paper_results/solvers/euler_methods.py
``point_rk4``         <- code/snapshot/weno7_point_rk4_shu/point_rk4.py (WENO7-JS + Shu RK4)
``point_rk4_mlp``     <- code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/point_rk4_mlp.py (WENO7-SR)
``weno7_double_mach`` <- code/snapshot/for_paper_results/solvers/weno7_double_mach.py
``shockbubble3d``     <- code/snapshot/warp_weno5_3d_rk3/solver.py run_solver (3-D WENO5)
"""
