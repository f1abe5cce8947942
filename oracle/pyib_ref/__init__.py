"""pyib_ref -- CPU oracle for the PyIB time-step hot path.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy/SciPy restatement of the
reference solver (PyIB, ``src/pyib``), written to be read side by side with it:
every function cites the reference ``file:line`` it follows.  It is the checker
for the HIP product path and the timed host-CPU baseline.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product package never does.

Pinning (what makes this oracle trustworthy), by module:

* ``mesh``      -- sha256 layout fingerprints of the reference D32/D40 meshes
                   (checks/pressure-validation-20260925.zip runs/*/metadata.json),
                   closed-form face/edge/cell counts, golden fixtures.
* ``core``      -- closed-form uniform geometry (Af=h^2, vol_ci=1/h^3, AoL=h),
                   grad-recon exactness, curl/rot adjointness, golden interface
                   geometry written by the reference itself (tests/golden).
* ``sparse``    -- A@s == flow_mat(s), symmetry of A, dense brute force.
* ``driver``    -- CG == dense solve on tiny meshes; golden short trajectories.
* ``immersed``  -- partition of unity, linear reproduction, force conservation,
                   spread/interp duality.
* ``kinematics``-- analytic velocity == central finite difference.
* ``interior``  -- closed-form sphere indicator volume; golden trajectories.
* ``case``      -- golden history rows; checkpoint/config fingerprints; the harmonic fit
                   against the reference's published D32/D40 fits (expected.json).
* ``postprocess``-- the reference's own manufactured-pressure and disconnected-graph tests
                   (tests/test_ib_postprocess.py:76-85, 113-116), restated; dense solve of
                   the pinned graph system on small graphs.
"""
