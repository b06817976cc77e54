"""One module per solver, each exposing
``solve(instance, *, tolerances, time_limit) -> Result``."""
