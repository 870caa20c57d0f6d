"""FP64 CPU oracle -- test infrastructure only (see sim_oracle.py header).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import anything under oracle/, and only as the checker.
"""
