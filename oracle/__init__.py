"""Test oracle package (CPU restatement of the reference's PDLP path).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the mipx_amd package.
"""
