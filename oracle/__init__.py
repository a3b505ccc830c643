"""CPU oracle for the batched propagation path -- TEST INFRASTRUCTURE ONLY.

Restates the reference algorithms (TemGymCore, src/temgym_core/*.py) in
numpy / plain C so that the HIP kernels can be checked against them.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import or run anything here, and only as the checker or as the timed
CPU baseline -- never as part of the product path.

Modules:
  ray_oracle      ray branch: propagation, components, pixel transform,
                  binning (numpy; f32 "exact-order" and f64 variants)
  stem_oracle     4D-STEM shadow-image forward and descan-fit loss/gradient
  beamlet_oracle  Gaussian-beamlet fields (complex128)
  c/              OpenMP C restatements used as the timed CPU baseline

Each function cites the reference file:line it follows.  How each one is
pinned (golden values, closed forms, brute force) is stated in its docstring;
a function with no such check says "parity unpinned".
"""
