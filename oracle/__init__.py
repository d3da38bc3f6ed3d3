"""CPU oracle for the D3Q19 alpha/beta A-A step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this package, and only as the checker. The product package
``cavityflow_cuda_amd`` never imports it.

``oracle.cpu`` binds ``liblbm_oracle.so``, a plain-C restatement of the
reference CPU solver (CLbmInitCPU / CLbmAlphaCPU / CLbmBetaCPU / CLbmSolver).
``oracle.ref`` binds ``_ref/libref_lbm.so``, the reference's own CPU kernels
compiled from /root/reference where that source tree exists.

Parity pins (tests/test_oracle.py): golden vectors written from the compiled
reference by tests/golden/make_golden.py, plus closed forms (rest equilibrium
equals the lattice weights, tau = 1 gives equilibrium, all-GHOST beta is a pure
neighbour shift, FLUID collision conserves mass and momentum).
"""
