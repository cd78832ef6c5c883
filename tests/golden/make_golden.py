"""Write the golden vectors under tests/golden/ from the reference itself.

Run once, in the build container only (the reference is not present on GPU
machines):  ``python tests/golden/make_golden.py``.  It needs the read-only
reference tree at /root/reference and NumPy; nothing of the reference is
copied -- only inputs and the arrays it returns are stored.

Vectors written (all float64, ``numpy.savez_compressed``, no pickles):

* ``ceep_results_traces.json`` -- the three 100-sample traces the reference
  ships in ``ceep_results.json`` (its CuPy run of ``BatchedFDTD2D.run()``).
* ``bf2d_*.npz`` -- ``BatchedFDTD2D.run()`` under the NumPy backend
  (src/ceep/solvers/fdtd_2d_batched.py:408-538), including sources inside the
  CPML so the CPML arithmetic itself is pinned.
* ``bf3d_*.npz`` -- ``BatchedFDTD3D.run()`` (fdtd_3d_batched.py:886-985).
* ``ade2d_*.npz`` -- ``FDTD2D`` with one Debye / one Lorentz pole
  (fdtd_2d.py:142-178, dispersive.py:240-426) on a grid the pulse never
  crosses, pinning the ADE operation order.
* ``das_small.npz`` -- ``DelayAndSumBeamformer.reconstruct`` DAS / DMAS images
  (imaging/beamforming.py:86-313) of a seeded S-matrix.
* ``cpml2d_*.npz`` -- ``FDTD2D`` with a ``CPML`` boundary
  (boundaries/absorbing.py:174-561, fdtd_base.py:161-185): the physical
  absorbing-boundary mode, traces and final fields.
* ``cpml3d_*.npz`` -- ``FDTD3D`` with a ``CPML`` boundary
  (boundaries/absorbing.py:567-943) and mu_r / sigma_m: the 3D physical
  absorbing-boundary mode, traces and final fields.
* ``runcpu_*.npz`` -- ``BatchedFDTD2D.run_cpu()`` (fdtd_2d_batched.py:540-611):
  interior sources with a lossy disk, and sources / probes on and next to the
  PEC walls (including corners and duplicated probes).
* ``brain2d.npz`` -- ``BrainPhantom2D.get_eps_map`` (phantoms/head_models.py:
  497-631) with and without the Gabriel tissues and a hemorrhage, and
  ``TissueProperties.permittivity`` / ``conductivity``
  (materials/tissue_database.py:92-137) of its six tissues.
* ``sparams.npz`` -- ``MultistaticSParameters`` accessors, ``to_db``,
  ``subtract_background``, ``MultistaticDataCollector`` and
  ``extract_s_parameters`` (solvers/s_params.py:29-353) on seeded data.
* ``waveforms.npz`` -- ``ModulatedGaussianSource.waveform`` and ``value_at``
  (sources/waveforms.py:97-132) at the configurations' parameters, plus the
  CFL time steps and constants of core/constants.py:21-27,105-175.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

sys.dont_write_bytecode = True
sys.path.insert(0, os.path.join(REF, "src"))


def _traces_2d(solver):
    try:
        solver.run()
    except AttributeError:
        # fdtd_2d_batched.py:529 calls ``np.asnumpy`` under NumPy; the probe
        # buffer is complete at that point.
        pass
    return np.array(solver.probe_data, dtype=np.float64)


def golden_ceep_results():
    with open(os.path.join(REF, "ceep_results.json")) as f:
        d = json.load(f)
    out = {"parameters": d["parameters"],
           "traces": {k: v["signal_real"] for k, v in d["scenarios"].items()}}
    with open(os.path.join(HERE, "ceep_results_traces.json"), "w") as f:
        json.dump(out, f)


def golden_bf2d():
    from ceep.core.backend import set_backend
    from ceep.solvers.fdtd_2d_batched import BatchedFDTD2D
    set_backend("numpy")

    # (a) the ceep_results.json configuration, empty scenario
    s = BatchedFDTD2D(nx=64, ny=64, dx=0.5e-3, total_steps=100, cpml_thickness=10,
                      source_positions=[(32, 32)], probe_positions=[(32, 32)],
                      frequency=2e9)
    tr = _traces_2d(s)
    np.savez_compressed(os.path.join(HERE, "bf2d_empty64.npz"), traces=tr, dt=s.dt,
             waveform=np.asarray(s.waveform))

    # (b) sources inside the CPML and on interior cells, probes everywhere
    # (x-lo, interior, x-hi, y-lo and corner sources; probes on the rows and
    # columns their fields can reach, since the CPML couples only along its
    # own axis)
    nx, ny, n, T = 40, 36, 5, 60
    src = [(3, 10), (20, 18), (37, 30), (20, 2), (2, 33)]
    prb = [(2, 2), (20, 18), (10, 30), (38, 33), (6, 6), (34, 29), (1, 10), (4, 10),
           (36, 30), (38, 30), (20, 3), (20, 1), (2, 34), (3, 33)]
    s = BatchedFDTD2D(nx=nx, ny=ny, dx=1e-3, total_steps=T, cpml_thickness=n,
                      source_positions=src, probe_positions=prb, frequency=3e9,
                      delay_factor=2.0)
    s.set_material_circle(22, 16, 6, eps_r=4.0, sigma_e=0.3)
    tr = _traces_2d(s)
    np.savez_compressed(os.path.join(HERE, "bf2d_pml_sources.npz"), traces=tr, dt=s.dt,
             waveform=np.asarray(s.waveform), eps_r=s._eps_r, sigma_e=s._sigma_e,
             sources=np.array(src), probes=np.array(prb), nx=nx, ny=ny, n=n, dx=1e-3)

    # (c) interior sources, several probes, a lossy disk and a long run so the
    # cavity walls matter
    nx, ny, n, T = 48, 44, 6, 150
    src = [(20, 20), (30, 25)]
    prb = [(20, 20), (30, 25), (12, 37), (7, 7), (40, 30)]
    s = BatchedFDTD2D(nx=nx, ny=ny, dx=0.5e-3, total_steps=T, cpml_thickness=n,
                      source_positions=src, probe_positions=prb, frequency=4e9)
    s.set_material_circle(26, 22, 7, eps_r=9.0, sigma_e=0.5)
    tr = _traces_2d(s)
    np.savez_compressed(os.path.join(HERE, "bf2d_interior.npz"), traces=tr, dt=s.dt,
             waveform=np.asarray(s.waveform), eps_r=s._eps_r, sigma_e=s._sigma_e,
             sources=np.array(src), probes=np.array(prb), nx=nx, ny=ny, n=n, dx=0.5e-3)


def golden_bf3d():
    from ceep.core.backend import set_backend
    from ceep.solvers.fdtd_3d_batched import BatchedFDTD3D
    set_backend("numpy")

    # (a) interior sources, dielectric sphere
    nx = ny = nz = 24
    n, T = 3, 40
    src = [(12, 12, 12), (14, 10, 11)]
    prb = [(12, 12, 12), (14, 10, 11), (8, 15, 13)]
    s = BatchedFDTD3D(nx=nx, ny=ny, nz=nz, dx=1e-3, total_steps=T, cpml_thickness=n,
                      source_positions=src, probe_positions=prb, frequency=5e9,
                      delay_factor=2.0)
    s.set_material_sphere(13, 12, 12, 4, eps_r=6.0, sigma_e=0.4)
    s.run()
    np.savez_compressed(os.path.join(HERE, "bf3d_sphere.npz"), traces=np.array(s.probe_data),
             dt=s.dt, waveform=np.asarray(s.waveform), eps_r=s._eps_r, sigma_e=s._sigma_e,
             mu_r=s._mu_r, sigma_m=s._sigma_m, sources=np.array(src),
             probes=np.array(prb), shape=np.array([nx, ny, nz]), n=n, dx=1e-3)

    # (b) one source inside the CPML, magnetic material, anisotropic box
    nx, ny, nz = 20, 18, 16
    n, T = 3, 30
    src = [(1, 9, 8), (10, 8, 7)]
    prb = [(1, 9, 8), (10, 8, 7), (2, 2, 2), (17, 15, 13)]
    s = BatchedFDTD3D(nx=nx, ny=ny, nz=nz, dx=1e-3, total_steps=T, cpml_thickness=n,
                      source_positions=src, probe_positions=prb, frequency=5e9,
                      delay_factor=2.0)
    s.set_material_region(6, 14, 5, 12, 4, 12, eps_r=3.0, sigma_e=0.2, mu_r=2.0,
                          sigma_m=50.0)
    s.run()
    np.savez_compressed(os.path.join(HERE, "bf3d_pml_magnetic.npz"), traces=np.array(s.probe_data),
             dt=s.dt, waveform=np.asarray(s.waveform), eps_r=s._eps_r, sigma_e=s._sigma_e,
             mu_r=s._mu_r, sigma_m=s._sigma_m, sources=np.array(src),
             probes=np.array(prb), shape=np.array([nx, ny, nz]), n=n, dx=1e-3,
             energy=s.compute_energy())


def golden_ade2d():
    from ceep.core.backend import set_backend
    from ceep.core.config import GridConfig, SimulationConfig
    from ceep.materials.dispersive import DebyePole, LorentzPole
    from ceep.solvers.fdtd_2d import FDTD2D
    from ceep.sources.waveforms import ModulatedGaussianSource
    set_backend("numpy")

    nx = ny = 48
    T = 22
    dx = 1e-3
    for name, kw in (
        ("debye", dict(eps_inf=4.0, debye_poles=[DebyePole(delta_eps=35.0, tau=8e-12)])),
        ("lorentz", dict(eps_inf=3.0, debye_poles=[LorentzPole(delta_eps=5.0,
                                                               omega_0=2 * np.pi * 5e9,
                                                               delta=1e9)])),
        ("debye_lorentz", dict(eps_inf=4.0, debye_poles=[
            DebyePole(delta_eps=35.0, tau=8e-12),
            LorentzPole(delta_eps=5.0, omega_0=2 * np.pi * 5e9, delta=1e9)])),
    ):
        cfg = SimulationConfig(grid=GridConfig(nx=nx, ny=ny, dx=dx, dy=dx), courant=0.5,
                               total_steps=T)
        src = ModulatedGaussianSource(x=24, y=24, frequency=20e9, bandwidth=150e9,
                                      amplitude=1.0, delay_factor=2.0)
        sim = FDTD2D(config=cfg, sources=[src], boundaries=[],
                     probe_points=[(24, 24), (27, 22), (20, 30)])
        sim.grid.set_material_circle(25, 23, 9, eps_r=kw["eps_inf"], sigma_e=0.7, **kw)
        sim.run(T)
        tr = np.array([sim.probe_data[p] for p in sim.probe_points], dtype=np.float64)
        np.savez_compressed(os.path.join(HERE, f"ade2d_{name}.npz"), traces=tr, dt=cfg.dt,
                 waveform=src.waveform(T, cfg.dt), eps_inf=np.asarray(sim.grid.eps_inf),
                 sigma_e=np.asarray(sim.grid.sigma_e),
                 delta_eps=np.asarray(sim.grid.dispersive.delta_eps[:2]),
                 tau=np.asarray(sim.grid.dispersive.tau[:2]),
                 omega_0=np.asarray(sim.grid.dispersive.omega_0[:2]),
                 damping=np.asarray(sim.grid.dispersive.damping[:2]),
                 pole_types=np.array(sim.grid.dispersive.pole_types[:2]),
                 probes=np.array(sim.probe_points), source=np.array([24, 24]),
                 nx=nx, ny=ny, dx=dx)


def golden_cpml2d():
    """``FDTD2D`` (TMz) with a ``CPML`` boundary (boundaries/absorbing.py:174-561):
    one soft source, probes in the interior, the PML slabs and a corner, a
    lossy disk; default CPML parameters and a second set (thickness 8, order 3,
    sigma_factor 1.5, alpha_max 0.02, kappa_max 3), plus one run with a Debye
    pole in the disk."""
    from ceep.boundaries.absorbing import CPML
    from ceep.core.backend import set_backend
    from ceep.core.config import GridConfig, SimulationConfig
    from ceep.materials.dispersive import DebyePole
    from ceep.solvers.fdtd_2d import FDTD2D
    from ceep.sources.waveforms import ModulatedGaussianSource
    set_backend("numpy")

    nx, ny, T, dx = 60, 52, 240, 1e-3
    probes = [(25, 22), (40, 30), (3, 20), (57, 45), (2, 2), (30, 50)]
    for name, ckw, debye in (
        ("default", dict(thickness=10), None),
        ("custom", dict(thickness=8, order=3, sigma_factor=1.5, alpha_max=0.02, kappa_max=3.0), None),
        ("debye", dict(thickness=10), DebyePole(delta_eps=30.0, tau=8e-12)),
    ):
        cfg = SimulationConfig(grid=GridConfig(nx=nx, ny=ny, dx=dx, dy=dx), courant=0.5,
                               total_steps=T)
        src = ModulatedGaussianSource(x=25, y=22, frequency=20e9, bandwidth=30e9,
                                      amplitude=1.0, delay_factor=3.0)
        cpml = CPML(**ckw)
        sim = FDTD2D(config=cfg, sources=[src], boundaries=[cpml], probe_points=probes)
        kw = {"debye_poles": [debye]} if debye is not None else {}
        sim.grid.set_material_circle(33, 27, 7, eps_r=6.0, sigma_e=0.4, **kw)
        sim.run(T)
        tr = np.array([sim.probe_data[p] for p in probes], dtype=np.float64)
        extra = {}
        if debye is not None:
            extra = dict(delta_eps=np.asarray(sim.grid.dispersive.delta_eps[0]),
                         tau=np.asarray(sim.grid.dispersive.tau[0]))
        np.savez_compressed(os.path.join(HERE, f"cpml2d_{name}.npz"), traces=tr, dt=cfg.dt,
                            waveform=src.waveform(T, cfg.dt),
                            eps_r=np.asarray(sim.grid.eps_inf), sigma_e=np.asarray(sim.grid.sigma_e),
                            ez=np.asarray(sim.grid.ez), hx=np.asarray(sim.grid.hx),
                            hy=np.asarray(sim.grid.hy), probes=np.array(probes),
                            source=np.array([25, 22]), nx=nx, ny=ny, dx=dx,
                            thickness=cpml.thickness, order=cpml.order,
                            sigma_factor=cpml.sigma_factor, alpha_max=cpml.alpha_max,
                            kappa_max=cpml.kappa_max, **extra)


def golden_cpml3d():
    """3D convolutional PML (boundaries/absorbing.py:567-943) around the
    single-run 3D update (solvers/fdtd_3d.py:97-171): one Ez soft source,
    probes in the interior, in the PML slabs and in a corner, a lossy block
    with mu_r / sigma_m; default CPML parameters and a second set.

    ``FDTD3D`` itself cannot be instantiated (it leaves ``get_field`` and
    ``initialize`` of its base class abstract), so the step is driven here in
    the order of ``FDTDBase.step`` (fdtd_base.py:161-185) by calling the
    reference's own pieces on its ``Grid3D``: ``FDTD3D._update_h_fields``,
    ``CPML.apply_h_3d``, ``FDTD3D._update_e_fields``, the source's
    ``value_at`` added at the source cell (``_inject_sources``, fdtd_3d.py:
    173-195), ``CPML.apply_e_3d``, then Ez at the probes."""
    from types import SimpleNamespace
    from ceep.boundaries.absorbing import CPML
    from ceep.core.backend import set_backend
    from ceep.core.config import GridConfig, SimulationConfig
    from ceep.core.grid_3d import Grid3D
    from ceep.solvers.fdtd_3d import FDTD3D
    from ceep.sources.waveforms import ModulatedGaussianSource
    set_backend("numpy")

    nx, ny, nz, T, dx = 30, 26, 24, 90, 1e-3
    probes = [(14, 12, 11), (20, 16, 13), (2, 12, 11), (27, 20, 9), (1, 1, 1), (15, 24, 22)]
    sx, sy, sz = 13, 12, 11
    for name, ckw in (("default", dict(thickness=6)),
                      ("custom", dict(thickness=5, order=3, sigma_factor=1.5, alpha_max=0.02,
                                      kappa_max=3.0))):
        cfg = SimulationConfig(grid=GridConfig(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, dz=dx),
                               courant=0.5, total_steps=T)
        grid = Grid3D(cfg.grid, cfg.dt)
        grid.set_material_region(16, 22, 9, 17, 8, 15, eps_r=5.0, mu_r=1.5, sigma_e=0.3,
                                 sigma_m=800.0)
        src = ModulatedGaussianSource(x=sx, y=sy, z=sz, frequency=20e9, bandwidth=30e9,
                                      amplitude=1.0, delay_factor=3.0)
        cpml = CPML(**ckw)
        cpml.initialize_3d(grid, cfg.dt)
        solver = SimpleNamespace(grid=grid, config=cfg, _use_fused_kernels=False)
        tr = np.zeros((len(probes), T))
        wave = np.zeros(T)
        for step in range(T):
            FDTD3D._update_h_fields(solver)
            cpml.apply_h_3d(grid)
            FDTD3D._update_e_fields(solver)
            val = src.value_at(step, cfg.dt)
            wave[step] = val
            if val != 0.0:
                grid.ez[sx, sy, sz] += val
            cpml.apply_e_3d(grid)
            for q, p in enumerate(probes):
                tr[q, step] = grid.ez[p]
        np.savez_compressed(os.path.join(HERE, f"cpml3d_{name}.npz"), traces=tr, dt=cfg.dt,
                            waveform=wave, eps_r=np.asarray(grid.eps_r),
                            sigma_e=np.asarray(grid.sigma_e), mu_r=np.asarray(grid.mu_r),
                            sigma_m=np.asarray(grid.sigma_m),
                            ez=np.asarray(grid.ez), hx=np.asarray(grid.hx),
                            hz=np.asarray(grid.hz), probes=np.array(probes),
                            source=np.array([sx, sy, sz]), shape=np.array([nx, ny, nz]), dx=dx,
                            thickness=cpml.thickness, order=cpml.order,
                            sigma_factor=cpml.sigma_factor, alpha_max=cpml.alpha_max,
                            kappa_max=cpml.kappa_max)


def golden_run_cpu():
    """``BatchedFDTD2D.run_cpu()`` (fdtd_2d_batched.py:540-611), the full-grid
    PEC box: (a) interior sources, a lossy disk, a duplicated probe; (b)
    sources on each wall, in a corner and next to the walls, probes on the
    walls and in the corners."""
    from ceep.core.backend import set_backend
    from ceep.solvers.fdtd_2d_batched import BatchedFDTD2D
    set_backend("numpy")

    cases = {
        "interior": dict(nx=40, ny=34, dx=0.5e-3, T=120, f=4e9, df=5.0,
                         src=[(20, 17), (5, 30), (33, 9)],
                         prb=[(20, 17), (5, 30), (33, 8), (20, 17), (10, 25)],
                         disk=(18, 15, 6, 9.0, 0.8)),
        "walls": dict(nx=30, ny=28, dx=1e-3, T=90, f=5e9, df=3.0,
                      src=[(0, 10), (29, 14), (12, 0), (1, 1), (28, 26), (15, 27)],
                      prb=[(0, 10), (29, 14), (12, 0), (1, 1), (28, 26), (15, 27), (14, 14),
                           (0, 0), (29, 27), (1, 10)],
                      disk=(14, 13, 5, 4.0, 0.3)),
    }
    for name, c in cases.items():
        s = BatchedFDTD2D(nx=c["nx"], ny=c["ny"], dx=c["dx"], total_steps=c["T"],
                          cpml_thickness=4, source_positions=c["src"],
                          probe_positions=c["prb"], frequency=c["f"], delay_factor=c["df"])
        cx, cy, r, e, sg = c["disk"]
        s.set_material_circle(cx, cy, r, eps_r=e, sigma_e=sg)
        out = s.run_cpu()
        tr = np.array([[out[b][p] for p in range(len(c["prb"]))] for b in range(len(c["src"]))],
                      dtype=np.float64)
        np.savez_compressed(os.path.join(HERE, f"runcpu_{name}.npz"), traces=tr, dt=s.dt,
                            eps_r=s._eps_r, sigma_e=s._sigma_e, sources=np.array(c["src"]),
                            probes=np.array(c["prb"]), nx=c["nx"], ny=c["ny"], dx=c["dx"],
                            frequency=c["f"], delay_factor=c["df"])


def golden_waveforms():
    """``ModulatedGaussianSource.waveform`` / ``value_at`` (sources/waveforms.py:
    97-132) at the configurations' pulse (2 GHz, 1.5 GHz bandwidth, default
    amplitude and delay factor) on the 2D (dx 0.5 mm, Courant 0.5) and 3D
    (dx 1 mm, 0.5 * 0.99) time steps, long enough to pass the envelope's peak,
    and at non-default amplitude / delay factor; the CFL steps and constants."""
    from ceep.core import constants as K
    from ceep.sources.waveforms import ModulatedGaussianSource

    dt2 = K.cfl_timestep_2d(0.5e-3, 0.5e-3, 0.5)
    dt3 = K.cfl_timestep_3d(1e-3, 1e-3, 1e-3, 0.5) * 0.99
    out = dict(dt2=dt2, dt3=dt3, c0=K.C_0, eps0=K.EPS_0, mu0=K.MU_0,
               dt2_aniso=K.cfl_timestep_2d(0.5e-3, 0.7e-3, 0.45),
               dt3_aniso=K.cfl_timestep_3d(1e-3, 0.8e-3, 1.2e-3, 0.4))
    src = ModulatedGaussianSource(x=0, y=0, frequency=2e9, bandwidth=1.5e9)
    out["cfg2d"] = src.waveform(4400, dt2)
    out["cfg3d"] = src.waveform(2800, dt3)
    src = ModulatedGaussianSource(x=0, y=0, frequency=18e9, bandwidth=20e9, amplitude=2.5,
                                  delay_factor=3.0)
    out["custom"] = src.waveform(300, dt2)
    out["custom_value_at"] = np.array([src.value_at(k, dt2) for k in range(300)])
    np.savez_compressed(os.path.join(HERE, "waveforms.npz"), **out)


def golden_brain2d():
    """``BrainPhantom2D.get_eps_map`` (head_models.py:497-631): Gabriel tissues
    at 2 GHz on a 200 x 180 grid with a hemorrhage, the fixed-value layers,
    and the 600^2 default head at 1 GHz; the six tissues' Cole-Cole
    permittivity and conductivity over 1 MHz - 10 GHz."""
    from ceep.materials.tissue_database import TissueDatabase
    from ceep.phantoms.head_models import BrainPhantom2D

    out = {}
    ph = BrainPhantom2D(nx=200, ny=180, dx=0.5e-3, hemorrhage_location=(2.0, -1.5),
                        hemorrhage_radius=1.2, head_radius_cm=4.5)
    out["a_eps"], out["a_sig"] = ph.get_eps_map(2e9)
    ph = BrainPhantom2D(nx=200, ny=180, dx=0.5e-3, hemorrhage_location=(-1.0, 0.5),
                        hemorrhage_radius=0.8, head_radius_cm=4.5, use_gabriel_database=False)
    out["b_eps"], out["b_sig"] = ph.get_eps_map(2e9)
    ph = BrainPhantom2D()
    out["c_eps"], out["c_sig"] = ph.get_eps_map(1e9)
    db = TissueDatabase()
    freqs = np.logspace(6, 10, 25)
    out["freqs"] = freqs
    for key in ("skin_wet", "skull_cortical_bone", "cerebrospinal_fluid", "brain_gray_matter",
                "brain_white_matter", "hemorrhage"):
        t = db.get(key)
        out[f"perm_{key}"] = t.permittivity(freqs)
        out[f"cond_{key}"] = t.conductivity(freqs)
    np.savez_compressed(os.path.join(HERE, "brain2d.npz"), **out)


def golden_sparams():
    """S-matrix accessors and exports (s_params.py:70-231), the collector
    (:234-353; one receiver position missing from a record) and two-port
    extraction (:29-66; incident values at and below 1e-15 included)."""
    from ceep.solvers.s_params import (MultistaticDataCollector, MultistaticSParameters,
                                       extract_s_parameters)

    rng = np.random.default_rng(31)
    F, N = 7, 5
    freqs = np.linspace(1e9, 3e9, F)
    S = rng.standard_normal((F, N, N)) + 1j * rng.standard_normal((F, N, N))
    Sb = rng.standard_normal((F, N, N)) + 1j * rng.standard_normal((F, N, N))
    S[2, 1, 3] = 0.0
    ants = [(10 + 3 * k, 20 - k) for k in range(N)]
    sp = MultistaticSParameters(frequencies=freqs, s_matrix=S, antenna_positions=ants)
    bg = MultistaticSParameters(frequencies=freqs, s_matrix=Sb, antenna_positions=ants)
    out = dict(freqs=freqs, s=S, s_bg=Sb, ants=np.array(ants))
    out["db"] = sp.to_db().s_matrix
    out["diff"] = sp.subtract_background(bg).s_matrix
    out["mono"] = sp.get_monostatic()
    out["mono3"] = sp.get_monostatic(3)
    out["elem"] = sp.get_element(2, 4)
    bi = sp.get_bistatic()
    out["bi_txrx"] = np.array([(t, r) for t, r, _ in bi])
    out["bi_s"] = np.array([v for _, _, v in bi])
    bi_all = sp.get_bistatic(exclude_monostatic=False)
    out["bi_all_txrx"] = np.array([(t, r) for t, r, _ in bi_all])

    class Array:
        def get_antenna_positions(self):
            return ants

    col = MultistaticDataCollector(Array(), freqs, dx=1e-3)
    rec = rng.standard_normal((N, N, F)) + 1j * rng.standard_normal((N, N, F))
    for tx in range(N):
        d = {ants[rx]: rec[tx, rx] for rx in range(N) if (tx, rx) != (1, 2)}
        col.record_transmission(tx, d)
    out["rec"] = rec
    out["collected"] = col.get_s_parameters().s_matrix
    inc = rng.standard_normal(9) + 1j * rng.standard_normal(9)
    inc[2] = 0.0
    inc[5] = 1e-16
    tot = rng.standard_normal(9) + 1j * rng.standard_normal(9)
    trn = rng.standard_normal(9) + 1j * rng.standard_normal(9)
    out["inc"], out["tot"], out["trn"] = inc, tot, trn
    out["s11"], out["s21"] = extract_s_parameters(inc, tot, trn)
    np.savez_compressed(os.path.join(HERE, "sparams.npz"), **out)


def golden_das():
    """``DelayAndSumBeamformer.reconstruct`` (imaging/beamforming.py:86-313),
    DAS and DMAS, raw and normalised, on a seeded S-matrix."""
    from ceep.imaging.beamforming import DelayAndSumBeamformer, ImagingRegion
    from ceep.solvers.s_params import MultistaticSParameters

    rng = np.random.default_rng(12)
    n_ant, F = 6, 13
    freqs = np.linspace(0.5e9, 3.5e9, F)
    S = rng.standard_normal((F, n_ant, n_ant)) + 1j * rng.standard_normal((F, n_ant, n_ant))
    ants = [(20 + int(14 * np.cos(2 * np.pi * k / n_ant)), 18 + int(13 * np.sin(2 * np.pi * k / n_ant)))
            for k in range(n_ant)]
    sp = MultistaticSParameters(frequencies=freqs, s_matrix=S, antenna_positions=ants)
    region = ImagingRegion(x_range=(9, 31), y_range=(7, 27), dx=1e-3, background_permittivity=4.0)
    out = {}
    for method in ("das", "dmas"):
        bf = DelayAndSumBeamformer(sp, ants, region, frequency_range=(1e9, 3e9))
        out[f"{method}_raw"] = bf.reconstruct(method=method, normalize=False)
        out[f"{method}_norm"] = bf.reconstruct(method=method, normalize=True)
    np.savez_compressed(os.path.join(HERE, "das_small.npz"), freqs=freqs, s_matrix=S,
                        antennas=np.array(ants), x_range=np.array(region.x_range),
                        y_range=np.array(region.y_range), dx=region.dx,
                        background_permittivity=region.background_permittivity,
                        frequency_range=np.array([1e9, 3e9]), **out)


if __name__ == "__main__":
    golden_ceep_results()
    golden_bf2d()
    golden_bf3d()
    golden_ade2d()
    golden_cpml2d()
    golden_cpml3d()
    golden_das()
    golden_run_cpu()
    golden_waveforms()
    golden_brain2d()
    golden_sparams()
    print("golden vectors written to", HERE)
