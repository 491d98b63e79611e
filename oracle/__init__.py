"""CPU oracle of the render path (test infrastructure: the checker and the CPU
baseline).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this package; the product library never does."""
from .pyoracle import *  # noqa: F401,F403
