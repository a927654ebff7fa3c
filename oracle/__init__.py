"""CPU oracles for the direct SHT hot path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product (``directsht_amd``) imports this package.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker (never as the thing measured
or shipped).

Modules
-------
``refalg``  FP64 numpy restatement of the reference algorithm
            (sht/sht.py:112-283, sht/legendre_py.py:4-118,
            sht/shared_interp_funcs.py:88-105, sht/interp_funcs_py.py:16-60,
            sht/shared_utils.py:3-99), including the polar rotation that the
            reference delegates to healpy (sht/sht.py:128-139).
``exact``   FP64 direct sum a_lm = sum_i w_i Y*_lm(theta_i, phi_i), with a
            per-point normalized associated-Legendre recurrence.  This is the
            "exact oracle": it has no spline error.
``wigner``  Wigner small-d matrices at beta = pi/2 and the alm rotation that
            replaces healpy's ``Rotator.rotate_alm``.
``cref``    ctypes loader for ``oracle/_ref/sht_helper.so``: the reference's
            own deprecated C implementation (csht/sht_helper.c) compiled from
            its sources by ``oracle/Makefile``.

Parity status of each oracle function (what pins it) is recorded in the
module docstrings and checked by ``tests/test_oracle_*.py``.
"""
