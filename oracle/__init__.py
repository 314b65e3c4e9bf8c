"""CPU restatement of libumbreon's per-ray path -- test infrastructure only.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker; the umbreon_amd product never imports it.
"""
