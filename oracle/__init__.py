"""FP64 CPU oracle for the FETI-DP Nedelec path -- TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU in float64 with numpy/scipy, the algorithms
of the reference (Poofee/emagddm) that the MI355X solver replaces, with the
reference's Nedelec defects corrected (see each module).  It is the checker:
only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import it.  The product (``emagddm_amd``) never imports, links or
calls anything here.

Modules and what pins them:

* ``nedelec``   -- Whitney functions, reference curls, element K_e/M_e/F_e.
  Pinned by the reference curl known answers (test/test_shape_function.cpp:
  306-328), a correct 4-point rule and an independent Duffy-collapsed
  Gauss rule, the gradient kernel K_e G_e = 0, symmetry / PSD / rank.
* ``topology``  -- global edge ids (first appearance) and free-dof numbering
  (PEC -> -1).  Pinned by a pure-Python loop restatement of the reference's
  hash-map algorithm and closed-form Kuhn edge counts.
* ``assembly``  -- global free-dof CSR and RHS with orientation signs.
  Pinned by symmetry, SPD, and the manufactured-solution convergence test.
* ``solvers``   -- Jacobi-PCG with CGSolver semantics and a sparse direct
  solve.  Pinned against the reference's own Eigen LLT / CGSolver built into
  ``oracle/_ref`` (golden vectors under tests/golden).
* ``fetidp``    -- dense CPU FETI-DP with the same primal/dual split, scaling
  and PCPG as the GPU solver.  Pinned by equality with the global direct
  solution and by brute-force dense F on micro meshes.
"""
