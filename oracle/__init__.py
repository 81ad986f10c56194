"""TEST INFRASTRUCTURE ONLY -- the oracle.

CPU restatement of the reference FoF (+ unbinding) path. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package, and only as the checker. The product package
``py_fof_amd`` never imports it.

The C sources are built by ``oracle/Makefile`` (``build()`` in
``__graft_entry__`` runs it); this module loads the result with ctypes.
"""
from .oracle import (  # noqa: F401
    oracle_available,
    oracle_fof,
    oracle_candidates,
    oracle_geometry,
    load_oracle,
    oracle_fof_bound,
    oracle_green,
    oracle_select,
    oracle_depth,
    oracle_tl_info,
    oracle_set_force_full,
    oracle_c2_check,
    oracle_tl_pp_check,
    oracle_direct_check,
    oracle_set_link_threads,
)
