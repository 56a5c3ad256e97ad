"""CPU oracle for the xgrid stencil path - TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker / the timed host
baseline.  The product path (``xgrid_amd``) never imports it.

* ``cpu_path``  - restatement of the reference's CPU path: the C99/OpenMP code
  that ``xgrid/lang/generator.py`` emits (one full-grid loop nest per stencil
  statement gated by the boundary mask, the two-phase nest with a malloc'd
  temporary for implicit statements, the ``_at`` accessor with its overstep
  modes), compiled with gcc, called through ctypes with the reference's
  by-value grid structs, on numpy grids whose time levels rotate exactly like
  ``xgrid/xgrid/__init__.py``'s.
* ``restate``   - independent numpy restatements of the workloads (diffusion,
  convection, lid-driven cavity, 3-D heat) written from their definitions, used
  to pin ``cpu_path`` bit-for-bit.
"""
