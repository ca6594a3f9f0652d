"""MI355X-native AFQMC walker step (SAFIRE's Cholesky HamiltonianOperations and
Slater-determinant operations as gfx950 HIP kernels behind a C ABI).

Submodules:
  afq        ctypes host mirror of include/afq.h (HamiltonianOperations, det_ops, Propagator)
  synthetic  the synthetic workloads (h1, Cholesky vectors, RHF trial)
"""
