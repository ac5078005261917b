"""CPU oracle of the rasterise-and-shade path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the reference they compare with; the product never does.
"""
