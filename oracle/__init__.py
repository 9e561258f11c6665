"""The oracle: CPU restatements of the reference CKM1D solver path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this package; the product
(coupled-bubble-control_amd/) never does.
"""
