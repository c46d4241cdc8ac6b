"""CPU oracle for the mr_recon NUFFT hot path -- TEST INFRASTRUCTURE ONLY.

A restatement, in plain torch-on-CPU and numpy, of the reference algorithm
(danielabrahamgit/mr_recon) for the path the HIP kernels replace. Each function
cites the reference file:line it follows. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and only as the checker (or, for ``cpu_baseline``, as the timed reference
CPU path) -- never as part of the product path.

How each part is pinned (tests/test_oracle_*.py):
  kb.py          A&S I0 polynomial vs scipy.special.i0; derivative vs central
                 differences.
  bins.py        scaling / wrap / lower-corner arithmetic vs torch's own fp32
                 mul/add/remainder/ceil (the library routines the reference
                 calls) and vs brute force over every grid point.
  interp_ref.py  ungrid/grid vs a pure-Python brute-force sum over all grid
                 points on tiny inputs; grid is the exact adjoint of ungrid.
  nufft_ref.py   forward/adjoint vs the FP64 direct NDFT (ndft.py) within the
                 KB approximation error for each (width, oversampling); adjoint
                 identity; Toeplitz normal == adjoint(forward) and, on a
                 sub-sampled Cartesian grid, == the NDFT normal matrix.
  sense_ref.py   SENSE forward/adjoint identities vs explicit per-coil loops;
                 CG vs torch.linalg.solve on a small dense SPD system.
  ndft.py        closed form; checked on integer k against numpy's FFT.
  dcf_ref.py     Pipe-Menon iteration vs the same iteration on a dense G^H G
                 built from the kernel definition; Cartesian closed form with
                 scipy's I0; invariances (duplicates, shifts, permutations).
  sigpy_ref.py   sigpy's KB interpolate / gridding vs the brute-force sum
                 (widths 1..4); adjointness; == interp_ref for widths >= 2.
  type3_ref.py   type-3 NUFFT vs the reference's exact-sum class
                 (type3_nufft_naive) within the KB error; adjointness;
                 Toeplitz normal == adjoint(forward).
"""
