"""Checker for the stim_amd detection-event sampler: CPU restatements of the reference's algorithm.

Test infrastructure only. Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg; never by the product package (stim_amd), which fails loudly without its HIP library.
"""
