"""PRAQ (QUBO roof-dual LP) and PRIA (joint-replenishment LP) on MI355X.

Restarted reflected Halpern PDHG as hand-written gfx950 HIP kernels behind a
C ABI (include/praq_hip.h, include/pria_hip.h, include/reform_batch.h), with
Python host code mirroring the praq-jax / pria-jax entry points:

    from reformulation_amd import praq, pria
    x, y, iters = praq.solve_qubo_problem(problem, praq.Parameters.with_auto_step(None, 1e-6))
    x, y, z, iters, lower, upper = pria.solve_jrp_problem(instance, params)

Many instances of uneven size solve together with `reformulation_amd.batch`,
and `reformulation_amd.serve` admits a stream of requests under an HBM
budget (continuous batching).
"""
from . import _native  # noqa: F401

__all__ = ["praq", "pria", "batch", "serve"]
