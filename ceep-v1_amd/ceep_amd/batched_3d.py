"""``BatchedFDTD3D`` -- drop-in for src/ceep/solvers/fdtd_3d_batched.py:48-1011.

Same constructor, material setters, ``run(verbose)``, ``compute_energy()`` and
``{tx: {rx: trace}}`` output; the time loop runs in HIP kernels
(libceep_fdtd.so) in FP32.  Keyword-only extensions as in ``BatchedFDTD2D``
(``courant``, ``waveform``, ``cpml_R``) plus ``set_debye`` / ``set_lorentz``
for dispersive media (dispersive.py:240-426 composed per E component in the
order of fdtd_2d.py:163-178).
"""
from __future__ import annotations

import time
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _native
from .batched_2d import _check_positions
from .constants import EPS_0, MU_0, cfl_timestep_3d
from .sparams import MultistaticSParameters
from .waveforms import gaussian_derivative


class BatchedFDTD3D:
    """Batched 3D Yee FDTD: one z-directed soft source per batch element."""

    def __init__(self, nx: int, ny: int, nz: int, dx: float, total_steps: int,
                 cpml_thickness: int, source_positions: List[Tuple[int, int, int]],
                 probe_positions: List[Tuple[int, int, int]], frequency: float,
                 delay_factor: float = 5.0, *, courant: Optional[float] = None,
                 waveform: Optional[np.ndarray] = None, cpml_R: float = 1e-8,
                 boundary: str = "pec", cpml_order: int = 4, cpml_sigma_factor: float = 2.0,
                 cpml_alpha_max: float = 0.05, cpml_kappa_max: float = 1.0):
        """``boundary="pec"`` (default) is the reference's batched solver: its
        CPML is inert, so the result is the PEC cavity it encloses.
        ``boundary="cpml"`` is an opt-in physical absorbing boundary: the full
        grid stepped in the order of the reference's single-run ``FDTD3D``
        (solvers/fdtd_3d.py, fdtd_base.py:161-185) with its
        ``CPML(thickness=cpml_thickness, order, sigma_factor, alpha_max,
        kappa_max)`` 3D recursions (boundaries/absorbing.py:567-943); sources
        and probes may lie anywhere in the grid."""
        if boundary not in ("pec", "cpml"):
            raise ValueError("boundary must be 'pec' or 'cpml'")
        self.boundary = boundary
        self.cpml_params = dict(order=cpml_order, sigma_factor=cpml_sigma_factor,
                                alpha_max=cpml_alpha_max, kappa_max=cpml_kappa_max)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx = dx
        self.dy = dx
        self.dz = dx
        self.total_steps = total_steps
        self.cpml_thickness = cpml_thickness
        self.source_positions = source_positions
        self.probe_positions = probe_positions
        self.frequency = frequency
        self.delay_factor = delay_factor
        self.batch = len(source_positions)
        self.courant = courant
        self.cpml_R = cpml_R
        if courant is None:
            self.dt = cfl_timestep_3d(dx, dx, dx, courant=0.5) * 0.99   # :133
        else:
            self.dt = cfl_timestep_3d(dx, dx, dx, courant=courant)
        self._user_waveform = None if waveform is None else np.asarray(waveform, dtype=np.float64)
        shape = (nx, ny, nz)
        self._eps_r = np.ones(shape, dtype=np.float64)
        self._sigma_e = np.zeros(shape, dtype=np.float64)
        self._mu_r = np.ones(shape, dtype=np.float64)
        self._sigma_m = np.zeros(shape, dtype=np.float64)
        self._debye = None
        self._lorentz = None
        self.tuning = None       # launch settings (see autotune.tuning_for); None = defaults
        self._built = False
        self._handle = None
        self.xp = np
        self.probe_data = None

    @property
    def _shape(self):
        return (self.nx, self.ny, self.nz)

    def set_material_region(self, x_start, x_end, y_start, y_end, z_start, z_end,
                            eps_r=1.0, sigma_e=0.0, mu_r=1.0, sigma_m=0.0) -> None:
        """fdtd_3d_batched.py:143-183."""
        r = (slice(x_start, x_end), slice(y_start, y_end), slice(z_start, z_end))
        self._eps_r[r] = eps_r
        self._sigma_e[r] = sigma_e
        self._mu_r[r] = mu_r
        self._sigma_m[r] = sigma_m

    def set_material_sphere(self, center_x, center_y, center_z, radius, eps_r=1.0,
                            sigma_e=0.0, mu_r=1.0, sigma_m=0.0) -> None:
        """fdtd_3d_batched.py:185-225 (mask as written there, meshgrid(z, y, x, 'ij'))."""
        z_grid, y_grid, x_grid = np.meshgrid(np.arange(self.nz), np.arange(self.ny),
                                             np.arange(self.nx), indexing="ij")
        mask = ((x_grid - center_x) ** 2 + (y_grid - center_y) ** 2
                + (z_grid - center_z) ** 2 <= radius ** 2)
        self._eps_r[mask] = eps_r
        self._sigma_e[mask] = sigma_e
        self._mu_r[mask] = mu_r
        self._sigma_m[mask] = sigma_m

    def set_phantom(self, phantom) -> None:
        """fdtd_3d_batched.py:227-264, plus Debye phantoms (``get_debye_maps``)."""
        if hasattr(phantom, "get_debye_maps"):
            m = phantom.get_debye_maps()
            self.set_debye(m["eps_inf"], m["delta_eps"], m["tau"], m["sigma"])
            if "lorentz_delta_eps" in m:
                self.set_lorentz(m["lorentz_delta_eps"], m["lorentz_omega_0"], m["lorentz_delta"])
        elif hasattr(phantom, "get_eps_map_3d"):
            eps_r, sigma_e = phantom.get_eps_map_3d(self.frequency)
            self._eps_r[:] = eps_r
            self._sigma_e[:] = sigma_e
        elif hasattr(phantom, "get_eps_map"):
            eps_r_2d, sigma_e_2d = phantom.get_eps_map(self.frequency)
            for k in range(self.nz):
                self._eps_r[:, :, k] = eps_r_2d
                self._sigma_e[:, :, k] = sigma_e_2d
        elif hasattr(phantom, "get_permittivity_map"):
            eps_real_2d, eps_imag_2d = phantom.get_permittivity_map(self.frequency)
            omega = 2 * np.pi * self.frequency
            for k in range(self.nz):
                self._eps_r[:, :, k] = eps_real_2d
                self._sigma_e[:, :, k] = eps_imag_2d * omega * EPS_0
        elif hasattr(phantom, "eps_r"):
            self._eps_r[:] = phantom.eps_r
            if hasattr(phantom, "sigma_e"):
                self._sigma_e[:] = phantom.sigma_e
        else:
            raise TypeError("Phantom must have get_eps_map_3d(), get_eps_map(), "
                            "get_permittivity_map(), or eps_r attribute")

    def set_debye(self, eps_inf, delta_eps, tau, sigma_e=None) -> None:
        self._eps_r[:] = np.broadcast_to(eps_inf, self._shape)
        if sigma_e is not None:
            self._sigma_e[:] = np.broadcast_to(sigma_e, self._shape)
        self._debye = (np.array(np.broadcast_to(delta_eps, self._shape), dtype=np.float64),
                       np.array(np.broadcast_to(tau, self._shape), dtype=np.float64))

    def set_lorentz(self, delta_eps, omega_0, delta) -> None:
        self._lorentz = tuple(np.array(np.broadcast_to(v, self._shape), dtype=np.float64)
                              for v in (delta_eps, omega_0, delta))

    def _waveform(self):
        if self._user_waveform is not None:
            return self._user_waveform
        return gaussian_derivative(self.total_steps, self.dt, self.frequency, self.delay_factor)

    def _build(self) -> None:
        """fdtd_3d_batched.py:266-326."""
        src = _check_positions(self.source_positions, self._shape, "source")
        prb = _check_positions(self.probe_positions, self._shape, "probe")
        self.waveform = self._waveform()
        if self._handle is not None:
            self._handle.close()
        mode = _native.BOX_CPML if self.boundary == "cpml" else _native.BOX_REFERENCE
        h = _native.Handle(3, self._shape, self.cpml_thickness, mode,
                           self.batch, len(self.probe_positions), self.total_steps,
                           self.dx, self.dt)
        if self.boundary == "cpml":
            h.set_cpml(**self.cpml_params)
        if self._debye is not None:
            h.set_debye(*self._debye)
        if self._lorentz is not None:
            h.set_lorentz(*self._lorentz)
        magnetic = bool(np.any(self._mu_r != 1.0) or np.any(self._sigma_m != 0.0))
        h.set_materials(self._eps_r, self._sigma_e,
                        self._mu_r if magnetic else None, self._sigma_m if magnetic else None)
        h.set_sources(src, self.waveform)
        h.set_probes(prb)
        if self.tuning is not None:
            h.apply_tuning(self.tuning)
        h.reset()
        self._handle = h
        self.inv_dx = 1.0 / self.dx
        self.inv_dy = 1.0 / self.dy
        self.inv_dz = 1.0 / self.dz
        self.probe_data = np.zeros((self.batch, len(self.probe_positions), self.total_steps))
        self._built = True

    def update_materials(self, eps_r, sigma_e) -> None:
        """Replace the eps_r / sigma maps of a built solver in place (device
        buffers, pole maps and launch setup are kept; fields are not reset)."""
        self._eps_r[:] = eps_r
        self._sigma_e[:] = sigma_e
        if not self._built:
            self._build()
            return
        magnetic = bool(np.any(self._mu_r != 1.0) or np.any(self._sigma_m != 0.0))
        self._handle.set_materials(self._eps_r, self._sigma_e,
                                   self._mu_r if magnetic else None,
                                   self._sigma_m if magnetic else None)

    def run(self, verbose: bool = False) -> Dict[int, Dict[int, np.ndarray]]:
        """fdtd_3d_batched.py:886-985 on the GPU."""
        if not self._built:
            self._build()
        t0 = time.time()
        self.probe_data = self._handle.run()
        if verbose:
            elapsed = time.time() - t0
            cells = self.batch * self.nx * self.ny * self.nz * self.total_steps
            print(f"\nCompleted {self.total_steps} steps in {elapsed:.2f}s")
            print(f"Performance: {cells / elapsed / 1e9:.2f} G cell-updates/sec")
        tr = self.probe_data
        return {tx: {rx: tr[tx, rx, :] for rx in range(tr.shape[1])} for tx in range(tr.shape[0])}

    def s_parameters(self, frequencies) -> MultistaticSParameters:
        if self._handle is None or self.probe_data is None:
            raise RuntimeError("run() first")
        f = np.atleast_1d(np.asarray(frequencies, dtype=np.float64))
        return MultistaticSParameters(f, self._handle.dft(f),
                                      [tuple(p) for p in self.probe_positions])

    def compute_energy(self) -> float:
        """fdtd_3d_batched.py:987-1011 (reduction on the GPU)."""
        if not self._built:
            raise RuntimeError("_build() first")
        return self._handle.energy()

    def _field(self, comp):
        if not self._built:
            raise AttributeError("fields exist after _build()")
        return self._handle.field(comp)

    ex = property(lambda self: self._field(0))
    ey = property(lambda self: self._field(1))
    ez = property(lambda self: self._field(2))
    hx = property(lambda self: self._field(3))
    hy = property(lambda self: self._field(4))
    hz = property(lambda self: self._field(5))

    def close(self) -> None:
        if self._handle is not None:
            self._handle.close()
            self._handle = None
            self._built = False
