"""CPU oracle for the per-shot RTM hot path -- TEST INFRASTRUCTURE ONLY.

This package is a NumPy restatement of the reference package's propagation,
checkpoint/replay and imaging arithmetic, written to reproduce its float32
operation order exactly (and its float64 mode) so that the GPU engine can be
checked against it.  It is the checker, never the thing measured or shipped:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it;
* the product package (``challenge-cup_amd/ertm``) never imports it and has no
  CPU fallback.

Modules and the reference code each follows:

=================  =========================================================
``numerics``       ertm/numerics.py:35-137 (stencil), 143-227 (CPML axis),
                   314-326 (CFL), 398-442 (laplacian, low-pass),
                   466-511 (Ricker, bilinear sampling)
``engines``        ertm/propagation.py:165-983 (media staging, the three
                   engines, sampling, injection, image fields, state)
``gsls``           ertm/viscoelastic.py:47-205 (GSLS constant-Q fit)
``rtm``            ertm/propagation.py:1004-1093 (wavelet, forward shot),
                   ertm/imaging.py:93-502 (checkpoints, replay, migrate)
=================  =========================================================

How each part is pinned (tests/test_oracle_*.py):

* golden vectors written by ``tests/golden/make_golden.py`` from the reference
  itself -- stencils, CPML profiles, GSLS fits, forward records and final
  states for all three engines (cpml/free/none, all source mechanisms),
  sampling/injection with duplicate receivers, replay blocks, and migrated
  images for every imaging condition and wave mode -- compared bit-for-bit;
* closed forms: coefficient identities, derivative of a sine, CFL formula,
  brute-force loop derivatives on tiny inputs, single-mechanism GSLS Q(f_ref);
* the reference's own behavioural thresholds (CPML absorption, moveout,
  explosive/shear energy ratios, layer depth).

Parity status: every function here is pinned by at least one golden vector.
"""

from . import engines, gsls, numerics, rtm  # noqa: F401
