"""CPU oracle for the wavelet potential-recovery path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It wraps ``liboracle.so`` (an FP64 C
restatement of the reference algorithm, built from ``whhd_oracle.c`` by the
Makefile next to it) with numpy-facing helpers.
"""
from .oracle import *  # noqa: F401,F403
