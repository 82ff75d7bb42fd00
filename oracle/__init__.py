"""Oracle for the voxel-grid render-and-optimise path (TEST INFRASTRUCTURE ONLY).

Two independent CPU restatements of the reference algorithm live here:

* ``oracle.plain`` -- an op-for-op plain-PyTorch restatement of the reference chain
  (src/ray_sampling.py, src/grid_functions.py, scripts/train.py:130-157), dtype- and
  device-generic, with autograd backward.  This is also "the plain method" the
  benchmark times beside the fused HIP path.
* ``oracle/render_ref.c`` -- a scalar per-ray loop in FP64 (front-to-back compositing and an
  analytic reverse-recurrence backward), loaded through ``oracle.cref``.  It is a second
  formulation of the same mathematics used to pin ``oracle.plain`` and to produce FP64
  gradients and central finite differences.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s baseline legs may import
this package, and only as the checker or the measured baseline.  The product package
``plenoxels_amd`` never imports it; it runs on the HIP library or raises.

Parity status of each function is recorded in its docstring ("pinned by: ...").
"""
