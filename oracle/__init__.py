"""The oracle: a CPU restatement of Elemental's sequential dense factor-and-solve path.

TEST INFRASTRUCTURE ONLY -- used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline; never by the product.
"""
