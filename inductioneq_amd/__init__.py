"""MI355X (gfx950) implementation of the InductionEq right-hand side, RK stages and
CG divergence cleaning, behind a C ABI that keeps the reference kernels' names,
argument order and compile-time settings.

Modules: ``settings`` (reference settings vocabulary), ``schedules`` (RK kernel lists),
``program`` (the C-ABI library over torch device memory), ``harness`` (the reference
driver flow), ``problems`` (benchmark inputs), ``build`` (hipcc build).
"""
from .settings import ProgramConfig, generate_settings  # noqa: F401
from .schedules import INTEGRATORS, STAGES_PER_STEP, rk_schedule  # noqa: F401

__all__ = ["ProgramConfig", "generate_settings", "INTEGRATORS", "STAGES_PER_STEP", "rk_schedule"]
