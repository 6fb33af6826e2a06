"""CPU FP64 oracle for the s2fft spin spherical-harmonic and Wigner transforms.

TEST INFRASTRUCTURE ONLY.  Nothing under ``oracle/`` is part of the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker.  The product path
(``s2fft_amd``) never imports it and fails loudly when its HIP library is
missing.

The oracle is a NumPy restatement of the reference's own algorithm, written
afresh.  Each function cites the reference file:line it follows (paths are
relative to the reference repository root).  Modules:

``geometry``   sampling geometry, quadrature weights, HEALPix phases
               (s2fft/sampling/s2_samples.py, so3_samples.py,
               s2fft/utils/quadrature.py, s2fft/utils/healpix_ffts.py:429-509)
``ffts``       HEALPix ragged ring FFTs and MW/MWSS resampling
               (s2fft/utils/healpix_ffts.py:19-117,187-226,335-371,
               s2fft/utils/resampling.py)
``pm``         Price-McEwen recursion precomputes and the latitudinal steps
               (s2fft/recursions/price_mcewen.py:118-255,
               s2fft/transforms/otf_recursions.py)
``transforms`` forward / inverse spin transforms, refinement, Wigner
               (s2fft/transforms/spherical.py, wigner.py,
               s2fft/utils/iterative_refinement.py)
``signals``    seeded coefficient generators with the reference's draw order
               (s2fft/utils/signal_generator.py)
``pins``       independent mathematics that pins the above: explicit Wigner-d
               sums, scipy spherical harmonics, HEALPix pix2ang closed form,
               brute-force DFTs.  No reference-held golden vectors exist for
               this path (the reference's cached test data is not shipped), so
               every function above is pinned by these checks in
               ``tests/test_oracle_pins.py``.
"""
