"""CPU oracle for the monodomain time step — TEST INFRASTRUCTURE ONLY.

This package is a plain numpy/scipy restatement, in float64, of the
reference algorithm (sanpinnawala/torchcor, ``torchcor/core``,
``torchcor/ionic``, ``torchcor/simulator/monodomain.py``).  It exists so the
tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` can check and time the HIP implementation.  It is never the thing
measured and never a fallback: the product package ``torchcor_amd`` does not
import it, and fails loudly when its HIP library is missing.

Modules
-------
``fem``        P1 tetrahedron / surface-triangle local matrices, conductivity
               tensors, COO->CSR assembly (assemble.py:99-248,
               conductivity.py:4-69).
``pcg``        Jacobi preconditioner and the reference PCG with its stopping
               rule and iteration-count convention (preconditioner.py:9-43,
               solver.py:23-76).
``ionic``      ModifiedMS2v, TenTusscherPanfilov, CourtemancheRamirezNattel
               including their lookup tables (ionic/*.py).
``monodomain`` Stimulus schedule and the theta-scheme time loop with LAT/RT
               tracking (stimulation.py:17-63, monodomain.py:83-226).

How each part is pinned (checked against something other than itself) is
listed in ``tests/test_oracle.py``; every function here has such a check.
"""
