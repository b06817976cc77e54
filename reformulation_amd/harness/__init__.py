"""Harness backends: the reference harness's "one more backend" slot.

`backends.praq_hip.solve` and `backends.pria_hip.solve` take the reference
harness's `Instance`, `Tolerances` and `time_limit` and return its `Result`
record (qubo/harness/results.py:24-80, jrp/harness/results.py:25-70).  When
the reference harness is importable its own `Result` class is used, so the
rows are its rows; otherwise `results.py` here supplies a record with the
same fields.  Registration: see REGISTRY.
"""

#: entries a maintainer adds to the reference harness's BACKENDS
#: (qubo/harness/registry.py:17-24, jrp/harness/registry.py:17-23)
REGISTRY = {
    "praq-hip": "reformulation_amd.harness.backends.praq_hip",
    "pria-hip": "reformulation_amd.harness.backends.pria_hip",
}
