"""Drop-in adapters with the reference's solver entry points.

``weno5_hllc``        <- code/snapshot/for_paper_results/solvers/weno5_hllc.py (WENO5-JS, or inline FP64 SR)
``weno5_hllc_refsym`` <- code/snapshot/teacherfree_lab_weno5_v20_distance_balanced/weno5_hllc_refsym.py (WENO5-SR FP64)
``weno5_hllc_mixed``  <- code/snapshot/for_paper_results/solvers/weno5_hllc_mixed.py (WENO5-SR FP32 MLP)
``euler_methods``     <- code/snapshot/for_paper_results/solvers/euler_methods.py
"""
