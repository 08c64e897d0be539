"""CPU oracle for the stealthy collective-coordinate hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ``xla_cco`` package
under ``xla-cco_amd/`` and its HIP library) may import, call, link or execute
anything in this directory.  It is used by ``tests/`` as the checker, by
``__graft_entry__.smoke()`` to check one small GPU invocation, and by
``bench.py`` for its ``cpu_baseline`` leg only.

Contents
--------
``oracle.py``
    NumPy restatement of the reference algorithms (k-set scan, trig energy and
    forces, adaptive descent, the optax-0.2.8 L-BFGS semantics, relax batching,
    Dormand-Prince gradient flow, velocity Verlet), each function citing the
    reference file:line it follows.
``cco_oracle.c`` + ``Makefile``
    Plain-C restatement of the trig energy/forces and of both minimizers, with
    OpenMP over configurations: the timed CPU baseline and a second oracle at
    sizes where NumPy is slow.

Parity pins (which check fixes each function) are listed in the header of
``oracle.py``.
"""
