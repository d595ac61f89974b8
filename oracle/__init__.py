"""CPU oracle for the Batch-OMP / K-SVD path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  The product library never calls it.
"""
from .oracle import Oracle, load, build  # noqa: F401
