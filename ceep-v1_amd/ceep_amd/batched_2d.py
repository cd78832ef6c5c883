"""``BatchedFDTD2D`` -- drop-in for src/ceep/solvers/fdtd_2d_batched.py:41-611.

Same constructor, methods, attributes and ``{tx: {rx: trace}}`` output as the
reference; the time loop runs in hand-written HIP kernels (libceep_fdtd.so)
in FP32, and traces come back as float64 views of one ``(B, P, T)`` array.

Keyword-only extensions (defaults reproduce the reference):
  ``courant``   -- Courant number; None keeps the reference dt = dx/(c*sqrt2)*0.99 (:92)
  ``waveform``  -- (T,) or (B, T) source samples; None keeps the built-in
                   Gaussian-derivative pulse (:169-182)
  ``cpml_R``    -- CPML reflection target (:210); the profile it shapes is inert
Dispersive media: ``set_debye`` / ``set_lorentz`` (or a phantom exposing
``get_debye_maps``) add the ADE of src/ceep/materials/dispersive.py:240-426 to
the E update in the order of src/ceep/solvers/fdtd_2d.py:163-178.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _native
from .constants import C_0, EPS_0, MU_0, cfl_timestep_2d
from .sparams import MultistaticSParameters
from .waveforms import gaussian_derivative


def _check_positions(positions, shape, what):
    out = []
    for i, p in enumerate(positions):
        p = tuple(int(v) for v in p)
        if len(p) != len(shape):
            raise ValueError(f"{what} {i} must have {len(shape)} coordinates, got {p}")
        for d, (v, n) in enumerate(zip(p, shape)):
            if not 0 <= v < n:
                raise IndexError(f"{what} {i} index {v} is out of bounds for axis {d} with size {n}")
        out.append(p)
    return out


class BatchedFDTD2D:
    """Batched 2D TMz FDTD: one TX per batch element, shared materials and probes."""

    def __init__(self, nx: int, ny: int, dx: float, total_steps: int, cpml_thickness: int,
                 source_positions: List[Tuple[int, int]],
                 probe_positions: List[Tuple[int, int]], frequency: float,
                 delay_factor: float = 5.0, *, courant: Optional[float] = None,
                 waveform: Optional[np.ndarray] = None, cpml_R: float = 1e-8,
                 boundary: str = "pec", cpml_order: int = 4, cpml_sigma_factor: float = 2.0,
                 cpml_alpha_max: float = 0.05, cpml_kappa_max: float = 1.0):
        """``boundary="pec"`` (default) is the reference's batched solver: its
        CPML is inert, so the result is the PEC cavity it encloses.
        ``boundary="cpml"`` is an opt-in physical absorbing boundary: the full
        grid stepped in the order of the reference's single-run ``FDTD2D``
        with its ``CPML(thickness=cpml_thickness, order, sigma_factor,
        alpha_max, kappa_max)`` (boundaries/absorbing.py:174-561); sources and
        probes may lie anywhere in the grid."""
        if boundary not in ("pec", "cpml"):
            raise ValueError("boundary must be 'pec' or 'cpml'")
        self.boundary = boundary
        self.cpml_params = dict(order=cpml_order, sigma_factor=cpml_sigma_factor,
                                alpha_max=cpml_alpha_max, kappa_max=cpml_kappa_max)
        self.nx = nx
        self.ny = ny
        self.dx = dx
        self.dy = dx
        self.total_steps = total_steps
        self.cpml_n = cpml_thickness
        self.source_positions = source_positions
        self.probe_positions = probe_positions
        self.frequency = frequency
        self.delay_factor = delay_factor
        self.batch = len(source_positions)
        self.courant = courant
        self.cpml_R = cpml_R
        if courant is None:
            self.dt = dx / (C_0 * np.sqrt(2.0)) * 0.99
        else:
            self.dt = cfl_timestep_2d(dx, dx, courant)
        self._user_waveform = None if waveform is None else np.asarray(waveform, dtype=np.float64)
        self._eps_r = np.ones((nx, ny), dtype=np.float64)
        self._sigma_e = np.zeros((nx, ny), dtype=np.float64)
        self._debye = None       # (delta_eps, tau) maps
        self._lorentz = None     # (delta_eps, omega_0, delta) maps
        self._built = False
        self._handle = None
        self.xp = np
        self.probe_data = None
        # kernel launch parameters (see autotune.py); None picks the tuned or
        # built-in defaults for this shape
        self.tuning = None

    # -- materials ------------------------------------------------------------------
    def set_material_circle(self, center_x: int, center_y: int, radius: int,
                            eps_r: float, sigma_e: float = 0.0) -> None:
        """fdtd_2d_batched.py:100-110."""
        y_grid, x_grid = np.meshgrid(np.arange(self.ny), np.arange(self.nx))
        mask = (x_grid - center_x) ** 2 + (y_grid - center_y) ** 2 <= radius ** 2
        self._eps_r[mask] = eps_r
        self._sigma_e[mask] = sigma_e

    def set_phantom(self, phantom) -> None:
        """fdtd_2d_batched.py:112-136, plus Debye phantoms (``get_debye_maps``)."""
        if hasattr(phantom, "get_debye_maps"):
            m = phantom.get_debye_maps()
            self.set_debye(m["eps_inf"], m["delta_eps"], m["tau"], m["sigma"])
            if "lorentz_delta_eps" in m:
                self.set_lorentz(m["lorentz_delta_eps"], m["lorentz_omega_0"], m["lorentz_delta"])
        elif hasattr(phantom, "get_eps_map"):
            eps_r, sigma_e = phantom.get_eps_map(self.frequency)
            self._eps_r[:] = eps_r
            self._sigma_e[:] = sigma_e
        elif hasattr(phantom, "get_permittivity_map"):
            eps_real, eps_imag = phantom.get_permittivity_map(self.frequency)
            self._eps_r[:] = eps_real
            omega = 2 * np.pi * self.frequency
            self._sigma_e[:] = eps_imag * omega * EPS_0
        elif hasattr(phantom, "eps_r"):
            self._eps_r[:] = phantom.eps_r
            if hasattr(phantom, "sigma_e"):
                self._sigma_e[:] = phantom.sigma_e
        else:
            raise TypeError("Phantom must have get_eps_map(), get_permittivity_map(), "
                            "or eps_r attribute")

    def set_debye(self, eps_inf, delta_eps, tau, sigma_e=None) -> None:
        """Single-pole Debye medium per cell (dispersive.py:32-46, 253-261).

        ``eps_inf`` replaces ``_eps_r`` (core/grid.py:178); cells with tau <= 0
        carry no pole.
        """
        shape = (self.nx, self.ny)
        self._eps_r[:] = np.broadcast_to(eps_inf, shape)
        if sigma_e is not None:
            self._sigma_e[:] = np.broadcast_to(sigma_e, shape)
        self._debye = (np.array(np.broadcast_to(delta_eps, shape), dtype=np.float64),
                       np.array(np.broadcast_to(tau, shape), dtype=np.float64))

    def set_lorentz(self, delta_eps, omega_0, delta) -> None:
        """Single Lorentz pole per cell (dispersive.py:68-87, 278-299)."""
        shape = (self.nx, self.ny)
        self._lorentz = tuple(np.array(np.broadcast_to(v, shape), dtype=np.float64)
                              for v in (delta_eps, omega_0, delta))

    # -- build / run ----------------------------------------------------------------
    def _waveform(self) -> np.ndarray:
        if self._user_waveform is not None:
            return self._user_waveform
        return gaussian_derivative(self.total_steps, self.dt, self.frequency, self.delay_factor)

    def _coefficients(self):
        """The (1, nx, ny) ca/cb/da/db maps of _build (:152-165); ADE eps_eff added."""
        dt = self.dt
        eps = self._eps_r * EPS_0
        add = np.zeros_like(eps)
        if self._debye is not None:
            de, tau = self._debye
            v = tau > 0
            add[v] += (2.0 * EPS_0 * de[v] * dt) / (2.0 * tau[v] + dt) / 2.0
        if self._lorentz is not None:
            de, w0, dl = self._lorentz
            v = w0 > 0
            dt2 = dt * dt
            add[v] += 2.0 * EPS_0 * de[v] * w0[v] ** 2 * dt2 / (
                4.0 + 2.0 * dl[v] * dt + w0[v] ** 2 * dt2) / 2.0
        if self._debye is not None or self._lorentz is not None:
            eps = eps + add
        s = self._sigma_e * dt / (2.0 * eps)
        ca = (1.0 - s) / (1.0 + s)
        cb = (dt / eps) / (1.0 + s)
        return ca, cb

    def _make_handle(self, box_mode):
        h = _native.Handle(2, (self.nx, self.ny), self.cpml_n, box_mode, self.batch,
                           len(self.probe_positions), self.total_steps, self.dx, self.dt)
        if self._debye is not None:
            h.set_debye(*self._debye)
        if self._lorentz is not None:
            h.set_lorentz(*self._lorentz)
        h.set_materials(self._eps_r, self._sigma_e)
        from .autotune import tuning_for
        h.apply_tuning(self.tuning if self.tuning is not None else
                       tuning_for(2, (self.nx, self.ny), self.cpml_n, self.batch,
                                  self._debye is not None, self._lorentz is not None))
        return h

    def _build(self) -> None:
        """fdtd_2d_batched.py:138-193: coefficients, waveform, zeroed device state."""
        src = _check_positions(self.source_positions, (self.nx, self.ny), "source")
        prb = _check_positions(self.probe_positions, (self.nx, self.ny), "probe")
        self.waveform = self._waveform()
        if self._handle is not None:
            self._handle.close()
        if self.boundary == "cpml":
            h = self._make_handle(_native.BOX_CPML)
            h.set_cpml(**self.cpml_params)
        else:
            h = self._make_handle(_native.BOX_REFERENCE)
        h.set_sources(src, self.waveform)
        h.set_probes(prb)
        h.reset()
        self._handle = h
        self.da = np.ones((1, self.nx, self.ny))
        self.db = np.full((1, self.nx, self.ny), self.dt / MU_0)
        self.inv_dx = 1.0 / self.dx
        self.inv_dy = 1.0 / self.dy
        n = self.cpml_n
        eta_0 = np.sqrt(MU_0 / EPS_0)
        sigma_max = -(3 + 1) * np.log(self.cpml_R) / (2.0 * eta_0 * n * self.dx)
        d = np.arange(n, dtype=np.float64)
        self.cpml_b_x = np.exp(-sigma_max * ((n - 1 - d) / (n - 1)) ** 3 * self.dt / EPS_0)
        self.cpml_c_x = self.cpml_b_x - 1.0
        self.cpml_b_y, self.cpml_c_y = self.cpml_b_x, self.cpml_c_x
        self.probe_data = np.zeros((self.batch, len(self.probe_positions), self.total_steps))
        self._target_maps = None          # reference medium of the difference form
        self._residual_run = False        # device traces hold trace - reference trace
        self._built = True

    def update_materials(self, eps_r, sigma_e) -> None:
        """Replace the eps_r / sigma maps of a built solver in place (the device
        state, buffers and launch setup are kept; fields are not reset)."""
        self._eps_r[:] = eps_r
        self._sigma_e[:] = sigma_e
        if not self._built:
            self._build()
            return
        self._handle.set_materials(self._eps_r, self._sigma_e)

    @staticmethod
    def _as_dict(traces: np.ndarray) -> Dict[int, Dict[int, np.ndarray]]:
        return {tx: {rx: traces[tx, rx, :] for rx in range(traces.shape[1])}
                for tx in range(traces.shape[0])}

    def run(self) -> Dict[int, Dict[int, np.ndarray]]:
        """fdtd_2d_batched.py:408-538 on the GPU; returns ``{tx: {rx: (T,) float64}}``."""
        if not self._built:
            self._build()
        self.probe_data = self._handle.run()
        self._residual_run = False
        return self._as_dict(self.probe_data)

    def run_cpu(self) -> Dict[int, Dict[int, np.ndarray]]:
        """fdtd_2d_batched.py:540-611: the full-grid PEC box without CPML.

        Runs on the GPU like ``run`` (the name is the reference's); each call
        starts from zero fields, as the reference's does.
        """
        src = _check_positions(self.source_positions, (self.nx, self.ny), "source")
        prb = _check_positions(self.probe_positions, (self.nx, self.ny), "probe")
        h = self._make_handle(_native.BOX_FULL)
        try:
            h.set_sources(src, self._waveform())
            h.set_probes(prb)
            traces = h.run()
        finally:
            h.close()
        return self._as_dict(traces)

    # -- outputs ----------------------------------------------------------------------
    # (B, P, T) float64 traces of the last run.  A gradient call that keeps its
    # traces on the device leaves a fetch function instead, run on first access.
    @property
    def probe_data(self):
        fetch = self.__dict__.get("_probe_fetch")
        if fetch is not None:
            self.__dict__["_probe_data"] = fetch()
            self.__dict__["_probe_fetch"] = None
        return self.__dict__.get("_probe_data")

    @probe_data.setter
    def probe_data(self, value):
        self.__dict__["_probe_fetch"] = None
        self.__dict__["_probe_data"] = value

    def _defer_probe_data(self, fetch) -> None:
        self.__dict__["_probe_fetch"] = fetch
        self.__dict__["_probe_data"] = None

    def s_parameters(self, frequencies) -> MultistaticSParameters:
        """Running DFT (solvers/dft.py:55-90) of the last run's traces as an
        ``(F, RX, TX)`` ``MultistaticSParameters`` (s_params.py:70-99)."""
        if self._handle is None or (self.__dict__.get("_probe_fetch") is None and
                                    self.probe_data is None):
            raise RuntimeError("run() first")
        f = np.atleast_1d(np.asarray(frequencies, dtype=np.float64))
        if getattr(self, "_residual_run", False):
            # after a difference-form run the device holds the residual
            # traces; probe_data (reference + residual) is transformed here
            ph = np.exp(-2j * np.pi * f[:, None] * (np.arange(self.total_steps) * self.dt)[None, :])
            S = np.einsum("ft,bpt->fpb", ph, self.probe_data) * self.dt
        else:
            S = self._handle.dft(f)
        return MultistaticSParameters(f, S, [tuple(p) for p in self.probe_positions])

    # the (1, nx, ny) coefficient maps of _build (:162-165), formed on access
    ca = property(lambda self: self._coefficients()[0][None])
    cb = property(lambda self: self._coefficients()[1][None])

    def _field(self, comp):
        if not self._built:
            raise AttributeError("fields exist after _build()")
        return self._handle.field(comp)

    ez = property(lambda self: self._field(0))
    hx = property(lambda self: self._field(1))
    hy = property(lambda self: self._field(2))

    def compute_energy(self) -> float:
        """(eps0/2 |Ez|^2 + mu0/2 |H|^2) dx dy summed over batch and grid."""
        if not self._built:
            raise RuntimeError("_build() first")
        return self._handle.energy()

    def close(self) -> None:
        if self._handle is not None:
            if self.__dict__.get("_probe_fetch") is not None:
                _ = self.probe_data             # traces left on the device by a gradient call
            self._handle.close()
            self._handle = None
            self._built = False
