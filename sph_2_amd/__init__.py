"""sph_2_amd — MI355X-native per-particle SPH step (gfx950 HIP kernels behind a C ABI).

Drop-in functions with the reference's call signatures live in
``sph_2_amd.libsx``; the resident whole-step run is ``sph_2_amd.plan.SphPlan``;
initial conditions and cooling tables are in ``sph_2_amd.ic``.
"""
from ._lib import FP64, MIXED, SphxError, default_params, load  # noqa: F401

__all__ = ["FP64", "MIXED", "SphxError", "default_params", "load"]
