"""CPU oracle for the WENO5 finite-volume step -- test infrastructure only.

Only the test-suite, the smoke entry point and the benchmark's CPU-baseline leg
import this package.  The product path never does.
"""
