"""CPU oracle for the DMRG two-site H_eff hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy/scipy, the reference's algorithm for the path
the HIP library implements (clover199/DMRG):

  * symmetry algebra and sector enumeration   -- src/useful.cpp
  * block-sparse tensor ``qtensor``            -- src/qtensor.cpp, src/tensor.cpp
  * MPO builders                               -- src/operators.cpp, include/mpo.h
  * matvec ``av``, step functions, sweeps     -- src/core.cpp, src/functions.cpp
  * ARPACK / LAPACK semantics                  -- src/dsaupd.cpp, src/znaupd.cpp,
                                                  src/dgesvd.cpp, src/zgesvd.cpp

Independent physics checks (closed forms, exact diagonalisation) live in
``oracle.exact``.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or as the timed
CPU baseline.  The product path (``dmrg_amd``) never calls into it.

Pinning status of each part (what independent check fixes it):

  * get_map / generate_dim_sym / generate_map  -- invariants + brute-force
    materialisation of H_eff against a dense Kronecker build (tests/test_oracle.py)
  * av / task GEMM                              -- dense H_eff from remove_symmetry
  * svd truncation                              -- U S V reconstruction, orthonormality,
                                                   selection rule as a pure function of S
  * DMRG energies                               -- Kitaev BdG closed form, XY free fermions,
                                                   sparse ED for XYZ and Z3 Potts
  * reference-sign-rule energies for bosonic models (the reference's get_map alpha
    applied to XYZ / Potts)                     -- parity unpinned: no closed form or
    ED describes that operator; only self-consistency (oracle vs HIP) checks it.
"""
