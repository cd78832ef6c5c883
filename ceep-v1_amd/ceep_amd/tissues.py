"""Cole-Cole tissue models for phantom generation (host side, NumPy FP64).

``TissueProperties.permittivity`` is the 4-term Cole-Cole model with the
static-conductivity term (src/ceep/materials/tissue_database.py:92-118):

    eps*(w) = eps_inf + sum_n delta_eps_n / (1 + (j w tau_n)^(1 - alpha_n)) + sigma_s / (j w eps0)

summed in that order, terms with delta_eps_n = 0 skipped.  ``GABRIEL`` holds
the parameter sets the reference's ``TissueDatabase`` gives the six tissues
its ``BrainPhantom2D`` uses (tissue_database.py:178-384, after Gabriel, Lau
and Gabriel, Phys. Med. Biol. 41 (1996) 2271); the values are data and are
the reference's.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Tuple

import numpy as np

from .constants import EPS_0


@dataclass
class TissueProperties:
    name: str
    eps_inf: float
    delta_eps: Tuple[float, float, float, float]
    tau: Tuple[float, float, float, float]
    alpha: Tuple[float, float, float, float]
    sigma_s: float

    def permittivity(self, frequencies) -> np.ndarray:
        """Complex relative permittivity eps' - j eps'' at ``frequencies`` (Hz)."""
        w = 2 * np.pi * np.asarray(frequencies)
        eps = np.ones_like(w, dtype=complex) * self.eps_inf
        for de, ta, al in zip(self.delta_eps, self.tau, self.alpha):
            if de > 0:
                eps += de / (1.0 + (1j * w * ta) ** (1 - al))
        eps += self.sigma_s / (1j * w * EPS_0)
        return eps

    def conductivity(self, frequencies) -> np.ndarray:
        """Effective conductivity w eps0 |eps''| (S/m) (tissue_database.py:120-137)."""
        w = 2 * np.pi * np.asarray(frequencies)
        return w * EPS_0 * np.abs(self.permittivity(frequencies).imag)


# key: (name, eps_inf, delta_eps, tau [s], alpha, sigma_s [S/m])
GABRIEL = {
    "skin_wet": ("Skin (Wet)", 4.0, (39.0, 280.0, 3.0e4, 1.2e7),
                 (7.23e-12, 32.5e-9, 159e-6, 15.9e-3), (0.0, 0.10, 0.20, 0.20), 0.0004),
    "skull_cortical_bone": ("Bone (Cortical)", 2.5, (10.0, 180.0, 5.0e3, 1.0e7),
                            (13.3e-12, 79.6e-9, 159e-6, 15.9e-3), (0.20, 0.20, 0.20, 0.0), 0.02),
    "cerebrospinal_fluid": ("Cerebrospinal Fluid (CSF)", 4.0, (65.0, 40.0, 0.0, 0.0),
                            (7.96e-12, 1.59e-9, 0.0, 0.0), (0.10, 0.0, 0.0, 0.0), 2.0),
    "brain_gray_matter": ("Brain (Gray Matter)", 4.0, (50.0, 500.0, 5.0e4, 3.5e7),
                          (7.96e-12, 15.9e-9, 106e-6, 5.30e-3), (0.10, 0.15, 0.20, 0.0), 0.02),
    "brain_white_matter": ("Brain (White Matter)", 4.0, (32.0, 100.0, 4.0e4, 3.0e7),
                           (7.96e-12, 7.96e-9, 53.0e-6, 7.96e-3), (0.10, 0.10, 0.30, 0.02), 0.02),
    "hemorrhage": ("Hemorrhage (Blood Clot)", 4.0, (58.0, 5.5e3, 5.5e3, 1.5e7),
                   (8.38e-12, 132e-9, 159e-6, 7.96e-3), (0.10, 0.10, 0.05, 0.00), 0.85),
}


def tissue(key: str) -> TissueProperties:
    """The ``TissueDatabase.get`` of these tissues (KeyError for others)."""
    if key not in GABRIEL:
        raise KeyError(f"Tissue '{key}' not found. Available: {', '.join(sorted(GABRIEL))}")
    return TissueProperties(*GABRIEL[key])
