"""CPU oracle for the D2Q9 BGK stream / bounce-back / collide path.

TEST INFRASTRUCTURE ONLY. Nothing in ``latticeboltzmann_amd`` imports this
package. Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may use it, and only as the checker or the timed CPU
baseline, never as the product path.

Contents
--------
``d2q9_numpy``
    numpy restatement of the reference's opt1 arithmetic
    (``simulators/parallel_lid_drive_cavity/cavity_opt1.py:88-245``,
    ``simulators/serial_shear_wave/Python/shear_wave_opt1.py:79-174``) with the
    exact operation order, plus the opt0 ``np.dot`` forms used by
    ``tests/02-CollideTest.py:43-89`` and the shear-wave analysis of
    ``simulators/simple_flows/shear_wave_decay.py:136-167``.
``d2q9_ref.c`` (built by ``oracle/Makefile`` into ``liboracle_d2q9.so``)
    single-thread C restatement of the same step (opt1 order) and of the
    ``_lbkernels`` collide formula (``c/d2q9.h:59-131``), used to cross-check
    the numpy restatement and as the CPU baseline.
``ctypes_ref``
    loader for the C restatement.
``lowprec``
    the reduced-precision (scaled deviation) codec and its conserving rounding.
``surrogate_ref``
    FP64 numpy forward/backward of the warm-start network (optionally with
    BF16 rounding where the GPU rounds), Adam, and the warm-start
    initializer's reconstructions (no reference counterpart).
``cpu_baseline``
    the timed CPU leg of ``bench.py``.

Pinning (what checks each function; the tests live in ``tests/test_oracle_*.py``)
--------------------------------------------------------------------------------
* ``stream``: the reference's streaming known-answer tests
  (``tests/simpleFlowsTest.py:176-196, 198-230, 260-297``), permutation
  invariants, and an index-map brute force.
* ``equilibrium``: closed-form moments (sum f_eq = rho, sum c f_eq = rho u,
  sum c c f_eq = rho (I/3 + u u)), f_eq(1, 0, 0) = w, and agreement with the
  opt0 ``np.dot`` form at 1e-7 (``tests/02-CollideTest.py:94-102``).
* ``collide``: agreement with ``d2q9_collide_ref`` at 1e-7 for omega in
  {0.5, 1.7} (``tests/02-CollideTest.py:104-111``), per-node mass and
  momentum conservation, omega = 1 gives f_eq.
* ``stream_and_bounce_back``: an independent per-node, per-channel brute force
  written from the wall physics (source node outside the box -> reflected
  population, lid momentum term), the u0 = 0 permutation property and the
  per-step lid mass-change identity.
* full steps: shear-wave decay against A0 exp(-nu k^2 t) with
  nu = (1/omega - 1/2)/3 (``simulators/simple_flows/shear_wave_decay.py:162``).
* the C restatement: bitwise equality with the numpy restatement.
"""
