"""tflows_amd -- MI355X-native drop-in for T-FlowS's preconditioned Krylov solves.

The package directory is ``t-flows_amd/``; it is imported under the name
``tflows_amd`` (see ``load_package`` in the repository's entry points).

Modules:
  solvers  -- Cg / Bicg / Cgs with the reference call signature, on the GPU
              through libtfs_solver.so (HIP, gfx950); plus SpMV, the FP32-inner
              refinement solver, the batched solver and the serving loop.
  mesh     -- synthetic hex-mesh cases and the CRS topology builder.
  cases    -- the measured workloads (sizes, seeds).
  build    -- in-tree build of the native libraries.
"""
