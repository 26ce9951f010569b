"""CPU oracle of the reference KinectFusion CPU path -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  It is the checker, never the product.
"""
