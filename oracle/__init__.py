"""CPU oracle of the per-frequency solve -- TEST INFRASTRUCTURE ONLY.

A restatement, in numpy / scipy and plain C, of the reference algorithm of
mom25d's per-frequency solve, used as the checker of the GPU path.  Only
`tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py`
may import, call or link anything under `oracle/`; the product package
`mom25d_amd` never does (its solve fails loudly without its HIP library).

Functions and what pins them (tests named in parentheses):
  greens_ref.SlabGreens        Engine/mom25d/greens.py:53-228       air slab vs exact images < 1e-6,
                                                                   lift independence, golden tables
                                                                   from the reference (test_oracle_greens)
  greens_ref.LayeredGreensRef  Engine/mom25d/layered.py:183-415     one layer == slab 1e-12, air stack vs
                                                                   images (test_oracle_greens)
  integrals_ref                Engine/mom25d/integrals.py:5-67      brute-force midpoint rule, polynomial
                                                                   exactness of the rules (test_oracle_integrals)
  assemble_ref.Geometry/assemble Engine/mom25d/solver.py:256-566   golden Z of the reference on the tee
                                                                   mesh at 2 frequencies, symmetry,
                                                                   numpy path == C path (test_oracle_assemble)
  fill.c (oracle_fill_block)   Engine/mom25d/_fast.py:82-158        == numpy path to 1e-12 (test_oracle_assemble)
  deembed_ref                  Engine/mom25d/solver.py:570-670, 1337-1427  wave fit with the reference's
                                                                   optimisers, reference impedances, port
                                                                   samples (through solve_ref.sweep)
  solve_ref.sweep              Engine/mom25d/solver.py:11-31, 569-1427  golden S / gamma / z0_line of the
                                                                   reference on tiny, tee and the via
                                                                   transition (test_oracle_sweep)
Third-party arithmetic used as in the reference: scipy.special.jv / j0
(AMOS / Cephes), scipy.interpolate.CubicSpline, scipy.linalg.lu_factor,
scipy.optimize.minimize(_scalar) -- the reference's own dependencies
(numpy >= 1.22, scipy >= 1.8, no lock file; here numpy 2.2.6, scipy 1.15.3).
"""
