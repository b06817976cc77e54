"""Measurement drivers beside bench.py: plain one-instance solves against the
oracle (`plain`), the iterate-precision study (`precision`) and the device
copy that calibrates the HBM counters (`pmc_copy`).  They import the oracle
as the CPU baseline and checker; the package `reformulation_amd` never
does."""
