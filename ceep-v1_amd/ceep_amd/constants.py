"""Physical constants and CFL time steps (src/ceep/core/constants.py:21-27, 105-175)."""
from __future__ import annotations

import math

C_0 = 299_792_458.0
EPS_0 = 8.854187817e-12
MU_0 = 4.0 * math.pi * 1e-7
ETA_0 = math.sqrt(MU_0 / EPS_0)


def cfl_timestep_2d(dx: float, dy: float, courant: float = 0.5) -> float:
    """constants.py:105-140."""
    return courant * (1.0 / (C_0 * math.sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy))))


def cfl_timestep_3d(dx: float, dy: float, dz: float, courant: float = 0.5) -> float:
    """constants.py:143-175."""
    return courant * (1.0 / (C_0 * math.sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz))))
