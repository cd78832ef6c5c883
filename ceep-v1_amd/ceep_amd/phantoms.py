"""Seeded phantom generators that feed ``set_phantom`` (host side, NumPy FP64).

The reference consumes phantoms through duck typing
(src/ceep/solvers/fdtd_2d_batched.py:112-136): ``get_eps_map(f)`` returning
(eps_r, sigma) maps, as its ``BrainPhantom2D`` does with concentric layers
(src/ceep/phantoms/head_models.py:546-631).  The Debye heads here expose
``get_debye_maps()`` -- per-cell (eps_inf, delta_eps, tau, sigma) of a
single-pole Debye medium (materials/dispersive.py:32-46) -- which the batched
solvers turn into the ADE.

* ``DiskPhantom2D``      -- one lossy dielectric disk in vacuum.
* ``DebyeHeadPhantom2D`` -- five concentric jittered ellipses (scalp, skull,
  CSF, grey, white matter) plus a hemorrhage disk inside the white matter.
* ``DebyeHeadPhantom3D`` -- the same layers as ellipsoids, a spherical
  hemorrhage and an optional Lorentz pole in every tissue cell.
* ``BrainPhantom2D``     -- the reference's concentric Cole-Cole head
  (``get_eps_map`` with the Gabriel tissues of ``ceep_amd.tissues``).
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

# (eps_inf, delta_eps, tau [s], sigma [S/m]) per layer, outermost first
HEAD_LAYERS = (
    ("scalp", (4.0, 35.0, 8e-12, 0.7)),
    ("skull", (3.0, 9.0, 20e-12, 0.1)),
    ("csf", (4.0, 65.0, 7e-12, 2.0)),
    ("grey", (4.0, 48.0, 8e-12, 0.8)),
    ("white", (4.0, 34.0, 8e-12, 0.5)),
)
HEMORRHAGE = (4.0, 56.0, 8e-12, 1.3)
HEAD_SEMI_AXES = (160.0, 150.0, 140.0, 135.0, 120.0)   # cells, for a 600-cell grid


def ring_positions(center: Tuple[float, float], radius: float, n: int) -> List[Tuple[int, int]]:
    """n antennas on a circle, ``int()`` truncation as the reference README does."""
    angles = np.linspace(0, 2 * np.pi, n, endpoint=False)
    return [(int(center[0] + radius * np.cos(a)), int(center[1] + radius * np.sin(a)))
            for a in angles]


class DiskPhantom2D:
    """Background eps_r = 1, sigma = 0 and one disk of radius U[r_lo, r_hi] cells
    at a random offset <= max_offset cells from the grid centre with
    eps_r ~ U[eps_lo, eps_hi] and sigma ~ U[sig_lo, sig_hi]."""

    def __init__(self, nx: int, ny: int, seed: int, r=(10.0, 30.0), max_offset=40.0,
                 eps=(10.0, 60.0), sigma=(0.2, 1.5)):
        rng = np.random.default_rng(seed)
        self.nx, self.ny = nx, ny
        self.radius = rng.uniform(*r)
        ang = rng.uniform(0.0, 2.0 * np.pi)
        off = rng.uniform(0.0, max_offset)
        self.center = (nx / 2 + off * np.cos(ang), ny / 2 + off * np.sin(ang))
        self.eps_r_value = rng.uniform(*eps)
        self.sigma_value = rng.uniform(*sigma)

    def get_eps_map(self, frequency: float):
        x, y = np.meshgrid(np.arange(self.nx), np.arange(self.ny), indexing="ij")
        mask = (x - self.center[0]) ** 2 + (y - self.center[1]) ** 2 <= self.radius ** 2
        eps = np.ones((self.nx, self.ny))
        sig = np.zeros((self.nx, self.ny))
        eps[mask] = self.eps_r_value
        sig[mask] = self.sigma_value
        return eps, sig


class DebyeHeadPhantom2D:
    """Seeded Debye head on an (nx, ny) grid.

    Semi-axes are HEAD_SEMI_AXES scaled by min(nx, ny)/600, each jittered by
    U[1-jitter, 1+jitter] per axis and layer; the hemorrhage disk has radius
    U[10, 25] cells (scaled likewise) and a centre drawn uniformly inside the
    white-matter ellipse (shrunk by the radius so it stays inside).
    """

    def __init__(self, nx: int, ny: int, seed: int, jitter: float = 0.05,
                 hemorrhage: bool = True):
        rng = np.random.default_rng(seed)
        self.nx, self.ny = nx, ny
        scale = min(nx, ny) / 600.0
        self.center = (nx / 2.0, ny / 2.0)
        self.axes = []
        for a in HEAD_SEMI_AXES:
            self.axes.append((a * scale * rng.uniform(1 - jitter, 1 + jitter),
                              a * scale * rng.uniform(1 - jitter, 1 + jitter)))
        self.hem = None
        if hemorrhage:
            r = rng.uniform(10.0, 25.0) * scale
            ax, ay = self.axes[-1]
            while True:
                u, v = rng.uniform(-1, 1, size=2)
                if u * u + v * v <= 1.0:
                    break
            cx = self.center[0] + u * max(ax - r, 0.0)
            cy = self.center[1] + v * max(ay - r, 0.0)
            self.hem = (cx, cy, r)

    def _grid(self):
        return np.meshgrid(np.arange(self.nx, dtype=np.float64),
                           np.arange(self.ny, dtype=np.float64), indexing="ij")

    def get_debye_maps(self) -> dict:
        x, y = self._grid()
        shape = (self.nx, self.ny)
        out = {"eps_inf": np.ones(shape), "delta_eps": np.zeros(shape),
               "tau": np.zeros(shape), "sigma": np.zeros(shape)}
        for (ax, ay), (_, (ei, de, tau, sig)) in zip(self.axes, HEAD_LAYERS):
            m = ((x - self.center[0]) / ax) ** 2 + ((y - self.center[1]) / ay) ** 2 <= 1.0
            out["eps_inf"][m], out["delta_eps"][m], out["tau"][m], out["sigma"][m] = ei, de, tau, sig
        if self.hem is not None:
            cx, cy, r = self.hem
            m = (x - cx) ** 2 + (y - cy) ** 2 <= r * r
            ei, de, tau, sig = HEMORRHAGE
            out["eps_inf"][m], out["delta_eps"][m], out["tau"][m], out["sigma"][m] = ei, de, tau, sig
        return out

    def get_eps_map(self, frequency: float):
        """Static snapshot at one frequency: eps' and effective sigma of the
        Debye medium, for callers that want a non-dispersive run."""
        m = self.get_debye_maps()
        w = 2 * math.pi * frequency
        chi = m["delta_eps"] / (1 + 1j * w * m["tau"])
        eps_c = m["eps_inf"] + chi
        sigma = m["sigma"] - eps_c.imag * w * 8.854187817e-12
        return eps_c.real, sigma


class DebyeHeadPhantom3D:
    """Seeded 3D Debye head: ellipsoid layers (semi-axes HEAD_SEMI_AXES scaled
    by min(shape)/600 and jittered per axis), a spherical hemorrhage of radius
    U[8, 15] cells inside the white matter, and optionally a Lorentz pole
    (delta_eps, omega_0, delta) in every tissue cell."""

    def __init__(self, shape, seed: int, jitter: float = 0.05,
                 lorentz=(5.0, 2 * math.pi * 5e9, 1e9), hem_radius=(8.0, 15.0)):
        rng = np.random.default_rng(seed)
        self.shape = tuple(shape)
        scale = min(shape) / 600.0
        self.center = tuple(s / 2.0 for s in shape)
        self.axes = [tuple(a * scale * rng.uniform(1 - jitter, 1 + jitter) for _ in range(3))
                     for a in HEAD_SEMI_AXES]
        r = rng.uniform(*hem_radius) * min(1.0, min(shape) / 200.0)
        ax = self.axes[-1]
        while True:
            u = rng.uniform(-1, 1, size=3)
            if float(u @ u) <= 1.0:
                break
        self.hem = tuple(self.center[d] + u[d] * max(ax[d] - r, 0.0) for d in range(3)) + (r,)
        self.lorentz = lorentz

    def get_debye_maps(self) -> dict:
        g = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in self.shape], indexing="ij")
        shape = self.shape
        out = {"eps_inf": np.ones(shape), "delta_eps": np.zeros(shape),
               "tau": np.zeros(shape), "sigma": np.zeros(shape)}
        tissue = np.zeros(shape, dtype=bool)
        for axes, (_, (ei, de, tau, sig)) in zip(self.axes, HEAD_LAYERS):
            m = sum(((g[d] - self.center[d]) / axes[d]) ** 2 for d in range(3)) <= 1.0
            out["eps_inf"][m], out["delta_eps"][m], out["tau"][m], out["sigma"][m] = ei, de, tau, sig
            tissue |= m
        cx, cy, cz, r = self.hem
        m = (g[0] - cx) ** 2 + (g[1] - cy) ** 2 + (g[2] - cz) ** 2 <= r * r
        ei, de, tau, sig = HEMORRHAGE
        out["eps_inf"][m], out["delta_eps"][m], out["tau"][m], out["sigma"][m] = ei, de, tau, sig
        if self.lorentz is not None:
            de, w0, dl = self.lorentz
            out["lorentz_delta_eps"] = np.where(tissue, de, 0.0)
            out["lorentz_omega_0"] = np.where(tissue, w0, 0.0)
            out["lorentz_delta"] = np.where(tissue, dl, 0.0)
        return out


class BrainPhantom2D:
    """Concentric circular head cross-section with an optional hemorrhage disk
    (src/ceep/phantoms/head_models.py:497-631): skin, skull, CSF, grey and
    white matter of 3, 7, 2 and 10 mm thickness inside a ``head_radius_cm``
    circle centred at (nx // 2, ny // 2).

    ``get_eps_map(f)`` returns (eps_r, sigma) maps.  With the Gabriel tissues,
    eps_r = Re eps*(f) and sigma = -Im eps*(f) 2 pi f 8.854e-12 -- the
    reference's rounded vacuum permittivity at this one place (:581), kept so
    the maps agree bit for bit; otherwise fixed (eps_r, sigma) per layer.
    """

    FIXED = {"skin": (40.0, 1.0), "skull": (12.0, 0.2), "csf": (68.0, 2.5),
             "gray": (52.0, 1.8), "white": (38.0, 1.2), "blood": (61.0, 2.0)}
    TISSUES = {"skin": "skin_wet", "skull": "skull_cortical_bone", "csf": "cerebrospinal_fluid",
               "gray": "brain_gray_matter", "white": "brain_white_matter", "blood": "hemorrhage"}

    def __init__(self, nx: int = 600, ny: int = 600, dx: float = 0.5e-3,
                 hemorrhage_location=None, hemorrhage_radius: float = 1.0,
                 head_radius_cm: float = 9.0, use_gabriel_database: bool = True):
        self.nx, self.ny, self.dx = nx, ny, dx
        self.head_radius = head_radius_cm * 1e-2
        self.hemorrhage_location = hemorrhage_location
        self.hemorrhage_radius = hemorrhage_radius * 1e-2
        self.use_gabriel = use_gabriel_database
        self.skin_thickness = 3e-3
        self.skull_thickness = 7e-3
        self.csf_thickness = 2e-3
        self.gray_thickness = 10e-3

    def _layer_values(self, frequency: float):
        if not self.use_gabriel:
            return dict(self.FIXED)
        from .tissues import tissue
        f = np.array([frequency])
        out = {}
        for layer, key in self.TISSUES.items():
            e = tissue(key).permittivity(f)[0]
            out[layer] = (e.real, -e.imag * 2 * np.pi * frequency * 8.854e-12)
        return out

    def get_eps_map(self, frequency: float):
        vals = self._layer_values(frequency)
        eps_r = np.ones((self.nx, self.ny), dtype=np.float64)
        sigma = np.zeros((self.nx, self.ny), dtype=np.float64)
        cx, cy = self.nx // 2, self.ny // 2
        gx, gy = np.meshgrid(np.arange(self.nx), np.arange(self.ny), indexing="ij")
        r = np.sqrt(((gx - cx) * self.dx) ** 2 + ((gy - cy) * self.dx) ** 2)
        # layer boundaries, outermost first
        edges = [self.head_radius]
        for t in (self.skin_thickness, self.skull_thickness, self.csf_thickness,
                  self.gray_thickness):
            edges.append(edges[-1] - t)
        for k, layer in enumerate(("skin", "skull", "csf", "gray")):
            m = (r <= edges[k]) & (r > edges[k + 1])
            eps_r[m], sigma[m] = vals[layer]
        m = r <= edges[4]
        eps_r[m], sigma[m] = vals["white"]
        if self.hemorrhage_location is not None:
            hx, hy = self.hemorrhage_location
            rh = np.sqrt(((gx - cx) * self.dx - hx * 1e-2) ** 2 +
                         ((gy - cy) * self.dx - hy * 1e-2) ** 2)
            m = rh <= self.hemorrhage_radius
            eps_r[m], sigma[m] = vals["blood"]
        return eps_r, sigma
