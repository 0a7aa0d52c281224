"""FP64 CPU oracle for the Widom insertion path -- TEST INFRASTRUCTURE ONLY.

A restatement of PureAdsorb.jl's CPU algorithm (``oracle.c``, built by ``oracle/Makefile``)
plus an independent numpy restatement (``np_oracle.py``). Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import this package,
and only as the checker: the product library never loads it.
"""
from .oracle import *  # noqa: F401,F403
