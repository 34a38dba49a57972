"""FP64 CPU oracle for the OpenTissue::fem step.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker / CPU timing
baseline.  The product library never imports, links or calls anything here.
"""
from .oracle import Oracle, OracleParams, OracleStats, reference_params, config_params, load_library, build

__all__ = ["Oracle", "OracleParams", "OracleStats", "reference_params", "config_params",
           "load_library", "build"]
