"""FP64 NumPy oracle for the QMTools grid hot path.

TEST INFRASTRUCTURE ONLY.  This package restates, in FP64 NumPy, what the
reference CUDA kernels of fullmetalfelix/QMTools-pycuda compute (each function
cites the reference file:line it follows).  It is imported only by `tests/`,
by `__graft_entry__.smoke()` and by the `cpu_baseline` leg of `bench.py`, and
there only as the checker or as the timed CPU baseline; the product path
(`qmtools`) never imports it.

The reference is CUDA-only (pycuda JIT + cuFFT) and cannot be built or run in
this image, so there is no `oracle/_ref`.  Each oracle function is pinned by
closed forms, invariants or brute force in `tests/test_oracle_*.py`.

Modules:
  basisfile  independent parser of the binary basis file (qmtools.py:33-108)
  system     molecule/basis arrays in the reference's ALMO order (qmtools.py:114-171)
  density    basis functions on points, density grid, normalisation (kernel_density.cu:14-155,
             qmtools.py:2381-2430)
  reduce     total / rescale / abs-diff / rel-diff (kernel_utils.cu:54-168)
  seed       nuclear seed charge (kernel_utils.cu:5-50)
  jacobi     Jacobi Poisson sweeps, zero and multipole Dirichlet boundaries
             (kernel_hartree.cu:85-171, qmtools.py:2497-2573)
  fft        total-charge RHS, periodic and free-space FFT Poisson
             (kernel_fft.cu:23-106, qmtools.py:2578-2639)
  gto        analytic GTO-pair + nuclear potential, McMurchie-Davidson / Boys
             (kernel_hartree.cu:288-869, qmtools.py:2436-2489)
  vne        classical V_ne averaged over voxel sub-points (kernel_vne_classic.cu:13-84);
             diffusion pseudo-V_ne relaxation (kernel_vne.cu:13-128, qmtools.py:2677-2745)
"""
