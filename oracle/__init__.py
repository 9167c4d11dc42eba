"""The oracle: a CPU FP64 restatement of the reference sfa/sim step.

TEST INFRASTRUCTURE ONLY. Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker. The product package
(scp_amd) never imports it.
"""
