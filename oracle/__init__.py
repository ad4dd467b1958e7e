"""CPU oracle for the IMEX_SfloW2D hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package.  The product library (imex_sflow2d_v2_amd) never imports it.
"""
from .pyoracle import Oracle, build_oracle, load_oracle  # noqa: F401
