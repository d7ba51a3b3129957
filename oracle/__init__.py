"""CPU oracle of the BD step: test infrastructure, never the product path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it.
"""
