"""CPU oracle for the TUPS vertex hot path — TEST INFRASTRUCTURE ONLY.

This package is a plain numpy restatement of the reference Fortran routines
(annakauch/TUPS, ``SRC/*.f90``) that the GPU library replaces.  It exists to
check the HIP path, never to stand in for it:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it;
* the product package ``tups_amd`` never imports it and fails loudly when its
  HIP library is missing.

Every function cites the reference file:line it restates.  Array layouts
follow the reference exactly: vertex arrays are Fortran-ordered numpy arrays
of shape ``(Nz, Nz, Nq)`` (``order='F'``) so that ``X[i, j, q]`` is the
reference's ``X(i+1, j+1, q+1)`` and the memory image is identical.

How each module is pinned (closed forms, invariants, brute force) is listed
in its own header; functions without such a check say "parity unpinned".
"""
