"""The oracle: a CPU (numpy / float64) restatement of gw-gpu-search's filtering hot path.

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package,
and only as the checker (or, for bench.py, as the timed CPU baseline). The product path in gw-gpu-search_amd/ never
imports, calls or falls back to it.

Each function cites the reference file:line it restates. Where a function has no golden vector, closed form or
brute-force check in tests/, its docstring says "parity unpinned".
"""
