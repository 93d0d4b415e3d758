"""FP64 CPU oracle for the per-particle SPH step (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker.  The product path
(``sph_2_amd``) never imports it.

``oracle.core`` wraps the C restatement (``sph_oracle.c``, built into
``liboracle.so`` by the Makefile here); ``oracle.npref`` is an independent
pure-numpy restatement used to cross-check it at small N.
"""
from .core import *  # noqa: F401,F403
