"""CPU oracle for the particle-filter hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy (uint32 / float32 / float64), the algorithms of
the reference particle-filter path so that the HIP implementation can be checked
against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
the timed CPU baseline -- never as a product code path.

Modules and what pins them (see ``tests/test_oracle_*.py``):

* ``threefry``  -- Threefry-2x32-20 hash and the JAX partitionable key tree
  (split / random_bits / uniform / normal).  Pinned by the Random123 known-answer
  vectors and, end to end, by the reference's regression CSVs.
* ``lg``        -- linear-Gaussian model components (rinit/rproc/dmeas/rmeas).
* ``filters``   -- simulate, bootstrap pfilter, MOP(alpha), perturbed filter
  (IF2), mif, panel mif.  Pinned by the reference's regression CSVs
  (tests/test_regression/*.csv in the reference) and by the exact Kalman
  log-likelihood.
* ``resample``  -- the build's systematic resampler (fixed-point prefix sums);
  pinned by brute-force exact rational arithmetic on small inputs.
* ``kalman``    -- FP64 Kalman filter log-likelihood (closed form oracle).
* ``seir``, ``binom``, ``sv`` -- models of the larger workloads; see their headers
  for what pins each function.
"""
