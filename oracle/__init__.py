"""FP64 CPU oracle for the batched FDTD hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy float64 (and in plain C for large sizes, see
``fdtd_ref.c``), the algorithm of the reference's ``BatchedFDTD2D.run()`` /
``BatchedFDTD3D.run()`` and the ADE / DFT pieces the configurations compose with
it.  Every function cites the reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import it, and only as the checker (or as the timed CPU baseline).  The
product path (``ceep_amd``) never imports anything from here.

Pinning (what checks each function against something other than itself):

* ``waveforms.reference_waveform``   -- golden traces (sample 0 is the closed form)
                                       and ``tests/golden/ceep_results_traces.json``.
* ``fdtd2d.run_literal``             -- ``ceep_results.json`` (3 scenarios x 100
                                       samples, produced by the reference itself).
* ``fdtd2d.run_cavity``              -- bitwise equal to ``run_literal`` for
                                       interior sources (sub-box theorem), and the
                                       discrete PEC-cavity dispersion relation.
* ``fdtd3d.run_literal``/``run_cavity`` -- equal to each other, batch
                                       independence, discrete 3D cavity modes.
* ``ade``                            -- homogeneous-medium discrete ADE
                                       permittivity (closed form), Debye -> plain
                                       reduction when delta_eps = 0.
* ``dft.running_dft``                -- DFT of analytic sequences (closed form),
                                       numpy.fft at bin frequencies.
"""
