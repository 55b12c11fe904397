"""CPU oracle for the transform path -- test infrastructure only.

Imported by tests/, the smoke check and bench.py's cpu_baseline leg, never by
the product package fft_amd.
"""
